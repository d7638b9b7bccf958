"""Writes tests/golden/nvdb_codec.npz: NanoVDB's own .nvdb files and lookups, and its verdict on files
this library writes.

Runs oracle/_ref/nvdb_probe (built from oracle/ref/nvdb_probe.cpp against the NanoVDB 32.3.3 headers
rtcamp8 vendors, with its ZIP codec; only where the reference tree is mounted):

1. ``nvdb_probe write`` -- two FloatGrids and a DoubleGrid written by nanovdb::io::writeGrids with
   Codec::NONE and Codec::ZIP.  Kept: both file images, and per float grid its world / index bbox as the
   reference keeps them, the reference's majorant, the background, the node counts, getValue at integer
   coordinates and the trilinear sample at world points.
2. For two grids built by the CPU oracle (dense 32^3 and sparse 64^3 fBm), the library's encoder writes a
   .nvdb file and ``nvdb_probe check`` reads it back with NanoVDB: GridValidator (detailed), the Partial
   checksum (GridBuilder's default), the bytes nanovdb::gridStats(StatsMode::All) would change (must be
   0), the reference's majorant and NanoVDB's lookups.  Kept: the SHA-256 of the file NanoVDB accepted, the query points and
   NanoVDB's answers.
"""
import ctypes as C
import hashlib
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

N_IJK, N_WORLD = 4000, 2000
ENCODED = {  # name -> FbmConfig keyword arguments
    "dense32": dict(resolution=32, seed=11),
    "sparse64": dict(resolution=64, seed=12, sparse=True, active_leaf_fraction=0.3, contrast=2.0),
}


def parse_grid_bin(raw: bytes) -> dict:
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    out = {"world_bbox": take("<f4", 6).copy(), "index_bbox": take("<i4", 6).copy(), "majorant": take("<f4", 1).copy(),
           "background": take("<f4", 1).copy(), "node_count": take("<u4", 3).copy()}
    (n,) = take("<u4", 1)
    pts = np.frombuffer(raw, dtype=np.dtype([("ijk", "<i4", 3), ("v", "<f4")]), count=n, offset=off)
    off += pts.nbytes
    (m,) = take("<u4", 1)
    wp = np.frombuffer(raw, dtype=np.dtype([("w", "<f4", 3), ("v", "<f4")]), count=m, offset=off)
    out.update(ijk=pts["ijk"][:N_IJK].copy(), ijk_value=pts["v"][:N_IJK].copy(), world=wp["w"][:N_WORLD].copy(),
               world_value=wp["v"][:N_WORLD].copy())
    return out


def main() -> int:
    from rtcamp8_amd import nvdb
    from rtcamp8_amd.scene import FbmConfig
    from oracle_lib import oracle_build_grid

    exe = ROOT / "oracle" / "_ref" / "nvdb_probe"
    if not exe.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle" / "ref")], check=True)
    arrays = {}
    with tempfile.TemporaryDirectory() as td:
        d = Path(td)
        subprocess.run([str(exe), "write", str(d)], check=True)
        arrays["file_none"] = np.frombuffer((d / "grids_none.nvdb").read_bytes(), dtype=np.uint8)
        arrays["file_zip"] = np.frombuffer((d / "grids_zip.nvdb").read_bytes(), dtype=np.uint8)
        for k in range(2):
            for key, val in parse_grid_bin((d / f"grid{k}.bin").read_bytes()).items():
                arrays[f"g{k}_{key}"] = val

        rng = np.random.default_rng(20221203)
        for name, kw in ENCODED.items():
            fb = FbmConfig(**kw)
            hg = oracle_build_grid(fb)
            data = nvdb.write_grids([nvdb.encode_grid(hg.grid, "fbm")], "none")
            (d / "ours.nvdb").write_bytes(data)
            ijk = rng.integers(-4, fb.resolution + 4, (N_IJK, 3)).astype(np.int32)
            world = (rng.random((N_WORLD, 3)) * 1.2 - 0.1).astype(np.float32)
            with open(d / "in.bin", "wb") as f:
                f.write(np.uint32(N_IJK).tobytes() + ijk.tobytes() + np.uint32(N_WORLD).tobytes() + world.tobytes())
            subprocess.run([str(exe), "check", str(d / "ours.nvdb"), str(d / "in.bin"), str(d / "out.bin")], check=True)
            raw = (d / "out.bin").read_bytes()
            valid, checksum_ok, changed, _first = np.frombuffer(raw, dtype="<i4", count=4)
            assert valid == 1 and checksum_ok == 1 and changed == 0, (name, valid, checksum_ok, changed)
            arrays[f"{name}_sha256"] = np.frombuffer(hashlib.sha256(data).digest(), dtype=np.uint8)
            arrays[f"{name}_size"] = np.int64(len(data))
            arrays[f"{name}_majorant"] = np.frombuffer(raw, dtype="<f4", count=1, offset=16).copy()
            arrays[f"{name}_ijk"] = ijk
            arrays[f"{name}_ijk_value"] = np.frombuffer(raw, dtype="<f4", count=N_IJK, offset=20).copy()
            arrays[f"{name}_world"] = world
            arrays[f"{name}_world_value"] = np.frombuffer(raw, dtype="<f4", count=N_WORLD, offset=20 + 4 * N_IJK).copy()
    np.savez_compressed(ROOT / "tests" / "golden" / "nvdb_codec.npz", **arrays)
    print(f"wrote nvdb_codec.npz ({len(arrays)} arrays)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
