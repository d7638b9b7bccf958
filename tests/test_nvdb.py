"""The .nvdb codec (host code of the library, no GPU): NanoVDB-written files decode to the values NanoVDB
itself looks up, files the library writes are the bytes NanoVDB validated, and malformed input is
rejected with an error instead of being read out of bounds.

Golden data: tests/golden/nvdb_codec.npz (tests/golden/make_nvdb_fixture.py, from the NanoVDB 32.3.3
that rtcamp8 vendors).  Lookups on the decoded tables go through the oracle's NanoVDB restatement, whose
indexing is pinned separately by tests/golden/nanovdb_probe.npz.
"""
import ctypes as C
import hashlib
import struct
from pathlib import Path

import numpy as np
import pytest

from rtcamp8_amd import _abi

GOLDEN = Path(__file__).resolve().parent / "golden" / "nvdb_codec.npz"


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(GOLDEN))


def lookups(grid: _abi.Grid, ijk: np.ndarray, world: np.ndarray):
    from oracle_lib import oracle

    o = oracle()
    offs = (C.c_uint32 * 3)()
    vals = np.array([o.oracle_grid_get_value(C.byref(grid), int(a), int(b), int(c), offs) for a, b, c in ijk],
                    dtype=np.float32)
    n = world.shape[0]
    w = np.ascontiguousarray(world, dtype=np.float32)
    out_ijk, out_off, tri = np.zeros((n, 3), np.int32), np.zeros((n, 3), np.uint32), np.zeros(n, np.float32)
    o.oracle_grid_probe(C.byref(grid), n, w.ctypes.data, out_ijk.ctypes.data, out_off.ctypes.data, tri.ctypes.data)
    return vals, tri


def bits(a):
    return np.asarray(a, dtype=np.float32).view(np.uint32)


@pytest.mark.parametrize("codec", ["none", "zip"])
def test_lists_nanovdb_files(fx, codec):
    from rtcamp8_amd import nvdb

    infos = nvdb.list_grids(fx[f"file_{codec}"].tobytes())
    assert [i.name for i in infos] == ["density", "shifted", "doubles"]
    assert [i.grid_type for i in infos] == [1, 1, 2]
    assert all(i.codec == nvdb.CODECS[codec] and i.version[0] == 32 for i in infos)
    for k in range(2):
        assert tuple(infos[k].node_count[:3]) == tuple(int(v) for v in fx[f"g{k}_node_count"])
        assert tuple(infos[k].index_bbox) == tuple(int(v) for v in fx[f"g{k}_index_bbox"])


@pytest.mark.parametrize("codec", ["none", "zip"])
@pytest.mark.parametrize("k", [0, 1])
def test_decodes_nanovdb_files_to_nanovdb_lookups(fx, codec, k):
    """Every getValue and trilinear sample equals NanoVDB's, bit for bit; the bboxes, background and node
    counts are NanoVDB's; the majorant is the reference's half-open loop (grid 1 holds its largest value
    on the bbox's upper face, which that loop skips)."""
    from rtcamp8_amd import nvdb

    with nvdb.read_grid(fx[f"file_{codec}"].tobytes(), k) as h:
        g = h.grid
        assert np.array_equal(bits(list(g.world_bbox_min) + list(g.world_bbox_max)), bits(fx[f"g{k}_world_bbox"]))
        assert list(g.index_bbox_min) + list(g.index_bbox_max) == fx[f"g{k}_index_bbox"].tolist()
        assert bits(g.background) == bits(fx[f"g{k}_background"][0])
        assert bits(g.max_value) == bits(fx[f"g{k}_majorant"][0])
        assert [g.num_leaves, g.num_lower, g.num_upper] == fx[f"g{k}_node_count"].tolist()
        vals, tri = lookups(g, fx[f"g{k}_ijk"], fx[f"g{k}_world"])
        assert np.array_equal(bits(vals), bits(fx[f"g{k}_ijk_value"]))
        assert np.array_equal(bits(tri), bits(fx[f"g{k}_world_value"]))
    if k == 1:
        assert float(fx["g1_majorant"][0]) < 7.5  # the upper-face voxel is not in the reference's loop


def test_non_float_grid_is_rejected(fx):
    from rtcamp8_amd import nvdb
    from rtcamp8_amd._lib import Rtc8xError

    with pytest.raises(Rtc8xError) as e:
        nvdb.read_grid(fx["file_none"].tobytes(), 2)
    assert e.value.status == _abi.RTC8X_ERR_NOT_SUPPORTED and "double" in str(e.value)
    with pytest.raises(Rtc8xError):
        nvdb.read_grid(fx["file_none"].tobytes(), 3)  # beyond the grid count


@pytest.mark.parametrize("name", ["dense32", "sparse64"])
def test_encoder_writes_the_bytes_nanovdb_validated(fx, name):
    """The library's encoder reproduces the file NanoVDB accepted (GridValidator, Partial checksum,
    gridStats leaving every byte unchanged); decoding it gives NanoVDB's lookups and the reference's
    majorant."""
    from rtcamp8_amd import nvdb
    from rtcamp8_amd.scene import FbmConfig
    from oracle_lib import oracle_build_grid

    import make_nvdb_fixture as mk

    hg = oracle_build_grid(FbmConfig(**mk.ENCODED[name]))
    data = nvdb.write_grids([nvdb.encode_grid(hg.grid, "fbm")], "none")
    assert len(data) == int(fx[f"{name}_size"])
    assert hashlib.sha256(data).digest() == fx[f"{name}_sha256"].tobytes()
    with nvdb.read_grid(data) as h:
        assert bits(h.grid.max_value) == bits(fx[f"{name}_majorant"][0])
        vals, tri = lookups(h.grid, fx[f"{name}_ijk"], fx[f"{name}_world"])
        assert np.array_equal(bits(vals), bits(fx[f"{name}_ijk_value"]))
        assert np.array_equal(bits(tri), bits(fx[f"{name}_world_value"]))
    # the oracle's own tables give the same answers
    vals, tri = lookups(hg.grid, fx[f"{name}_ijk"], fx[f"{name}_world"])
    assert np.array_equal(bits(vals), bits(fx[f"{name}_ijk_value"]))


def grid_buffers(file: bytes):
    """The raw grid buffers of a codec-NONE file, located from the segment header and records."""
    from rtcamp8_amd import nvdb

    infos = nvdb.list_grids(file)
    off = 16 + sum(176 + len(i.name.encode()) + 1 for i in infos)
    out = []
    for i in infos:
        out.append(file[off: off + i.grid_size])
        off += i.file_size
    return infos, out


@pytest.mark.parametrize("k", [0, 1])
def test_reencoding_a_nanovdb_grid_gives_nanovdb_bytes(fx, k):
    """A grid NanoVDB's GridBuilder wrote re-encodes to exactly the bytes NanoVDB wrote (activity, stats,
    map, world bbox, checksum), and decode -> encode -> decode keeps every lookup."""
    from rtcamp8_amd import nvdb

    infos, buffers = grid_buffers(fx["file_none"].tobytes())
    with nvdb.decode(buffers[k]) as h:
        assert nvdb.encode_grid(h.grid, infos[k].name) == buffers[k]
    with nvdb.read_grid(fx["file_none"].tobytes(), k) as h:
        enc = nvdb.encode_grid(h.grid, "again")
        g0 = h.grid
        with nvdb.decode(enc) as h2:
            g1 = h2.grid
            for f in ("world_bbox_min", "world_bbox_max", "inv_mat", "vec"):
                assert np.array_equal(bits(list(getattr(g0, f))), bits(list(getattr(g1, f)))), f
            assert list(g0.index_bbox_min) == list(g1.index_bbox_min)
            assert bits(g0.max_value) == bits(g1.max_value) and bits(g0.background) == bits(g1.background)
            vals, tri = lookups(g1, fx[f"g{k}_ijk"], fx[f"g{k}_world"])
            assert np.array_equal(bits(vals), bits(fx[f"g{k}_ijk_value"]))
            assert np.array_equal(bits(tri), bits(fx[f"g{k}_world_value"]))
            assert nvdb.encode_grid(g1, "again") == enc


def test_zip_and_multi_segment_files(fx):
    from rtcamp8_amd import nvdb

    with nvdb.read_grid(fx["file_none"].tobytes(), 0) as h:
        g0 = nvdb.encode_grid(h.grid, "first")
    with nvdb.read_grid(fx["file_none"].tobytes(), 1) as h:
        g1 = nvdb.encode_grid(h.grid, "second")
    zipped = nvdb.write_grids([g0, g1], "zip")
    plain = nvdb.write_grids([g0, g1], "none")
    assert len(zipped) < len(plain) // 4
    two_segments = plain + zipped  # readGrid counts grids across segments
    infos = nvdb.list_grids(two_segments)
    assert [i.name for i in infos] == ["first", "second", "first", "second"]
    assert [i.segment for i in infos] == [0, 0, 1, 1]
    for n in range(4):
        with nvdb.read_grid(two_segments, n) as h:
            assert nvdb.encode_grid(h.grid, infos[n].name) == (g0, g1)[n % 2]


def _corrupt(data: bytes, offset: int, payload: bytes) -> bytes:
    b = bytearray(data)
    b[offset:offset + len(payload)] = payload
    return bytes(b)


def test_malformed_files_are_rejected(fx):
    from rtcamp8_amd import nvdb
    from rtcamp8_amd._lib import Rtc8xError

    good = fx["file_none"].tobytes()
    info = nvdb.list_grids(good)[0]
    grid_at = 16 + sum(176 + len(i.name) + 1 for i in nvdb.list_grids(good))
    assert grid_at == 16 + 3 * 176 + len("density") + len("shifted") + len("doubles") + 3
    cases = {
        "truncated header": good[:10],
        "bad magic": b"X" + good[1:],
        "reversed endianness": struct.pack(">Q", 0x304244566F6E614E) + good[8:],
        "old major version": _corrupt(good, 8, struct.pack("<I", (31 << 21) | (1 << 10))),
        "truncated grid": good[: grid_at + info.grid_size // 2],
        "BLOSC codec": _corrupt(good, 14, struct.pack("<H", 2)),
        # root offset of grid 0 pointed past its end
        "root offset": _corrupt(good, grid_at + 672 + 24, struct.pack("<Q", info.grid_size)),
        # first root tile's child link pointed into the grid header
        "child link": _corrupt(good, grid_at + 672 + 64 + 64 + 8, struct.pack("<q", -64)),
        # misaligned child link
        "misaligned link": _corrupt(good, grid_at + 672 + 64 + 64 + 8,
                                    struct.pack("<q", struct.unpack_from("<q", good, grid_at + 800 + 8)[0] + 4)),
        # node counts that disagree with the links
        "node count": _corrupt(good, grid_at + 672 + 32, struct.pack("<I", 1)),
    }
    for what, data in cases.items():
        with pytest.raises(Rtc8xError) as e:
            with nvdb.read_grid(data, 0):
                pass
        assert e.value.status in (_abi.RTC8X_ERR_INVALID_ARGUMENT, _abi.RTC8X_ERR_NOT_SUPPORTED), what
    # a raw grid buffer with the wrong magic
    with pytest.raises(Rtc8xError):
        nvdb.decode(b"\0" * 1024)


def test_zip_header_claiming_a_huge_grid_is_an_error(fx):
    """A ZIP grid's record claims its inflated size; a claim beyond deflate's 1032:1 bound (2^62 bytes here)
    is rejected before any buffer is sized from it, and so is a claim that disagrees with the stream."""
    from rtcamp8_amd import nvdb
    from rtcamp8_amd._lib import Rtc8xError

    with nvdb.read_grid(fx["file_none"].tobytes(), 0) as h:
        g0 = nvdb.encode_grid(h.grid, "first")
    zipped = nvdb.write_grids([g0], "zip")
    assert nvdb.list_grids(zipped)[0].grid_size == len(g0)
    for claim, status in ((1 << 62, _abi.RTC8X_ERR_INVALID_ARGUMENT), (len(g0) + 4096, _abi.RTC8X_ERR_INVALID_ARGUMENT),
                          (100, _abi.RTC8X_ERR_INVALID_ARGUMENT)):
        bad = _corrupt(zipped, 16, struct.pack("<Q", claim))  # the first record's grid size
        for reader in (lambda d: nvdb.read_grid(d, 0).close(), lambda d: nvdb.load_grid(d, 0)):
            with pytest.raises(Rtc8xError) as e:
                reader(bad)
            assert e.value.status == status, claim


def test_free_host_rejects_tables_it_did_not_allocate():
    from rtcamp8_amd._lib import Rtc8xError, call
    from rtcamp8_amd.scene import FbmConfig
    from oracle_lib import oracle_build_grid

    hg = oracle_build_grid(FbmConfig(resolution=16, seed=3))
    with pytest.raises(Rtc8xError):
        call("rtc8x_grid_free_host", C.byref(hg.grid))
