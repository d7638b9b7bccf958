"""Writes tests/golden/stb_bmp.npz: BMP files written by the stb_image_write v1.14 that rtcamp8 vendors.

Runs oracle/_ref/stb_bmp_probe (built from oracle/ref/stb_bmp_probe.c against ext/stb_image_write.h;
only where the reference tree is mounted) for a few image sizes, odd widths included so that every
row padding occurs, and keeps each RGBA8 input (all 256 alpha values appear) with stb's output bytes.
"""
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
SIZES = [(1, 1), (2, 3), (13, 7), (16, 16), (33, 9)]


def main() -> int:
    exe = ROOT / "oracle" / "_ref" / "stb_bmp_probe"
    if not exe.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle" / "ref")], check=True)
    arrays = {}
    with tempfile.TemporaryDirectory() as td:
        for k, (w, h) in enumerate(SIZES):
            rgba, bmp = Path(td) / "px.bin", Path(td) / "o.bmp"
            subprocess.run([str(exe), str(w), str(h), str(1000 + k), str(rgba), str(bmp)], check=True)
            arrays[f"rgba_{w}x{h}"] = np.fromfile(rgba, dtype="<u4").reshape(h, w)
            arrays[f"bmp_{w}x{h}"] = np.frombuffer(bmp.read_bytes(), dtype=np.uint8)
    np.savez_compressed(ROOT / "tests" / "golden" / "stb_bmp.npz", **arrays)
    print(f"wrote stb_bmp.npz ({len(SIZES)} images)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
