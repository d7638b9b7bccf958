"""GPU side of the .nvdb codec: volumes loaded from NanoVDB files give NanoVDB's lookups on the device,
render as the oracle renders them, and the benchmark's 512^3 volume survives a write -> read round trip
unchanged (size-independent checks at full size).

Golden data: tests/golden/nvdb_codec.npz (tests/golden/make_nvdb_fixture.py)."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "nvdb_codec.npz"


def device_probe(grid, world: np.ndarray) -> np.ndarray:
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer

    n = world.shape[0]
    dpts = DeviceBuffer.from_array(np.ascontiguousarray(world, dtype=np.float32))
    dijk, doff, dval = DeviceBuffer(12 * n), DeviceBuffer(12 * n), DeviceBuffer(4 * n)
    try:
        call("rtc8x_grid_probe", C.byref(grid), n, C.c_void_p(dpts.ptr), C.c_void_p(dijk.ptr), C.c_void_p(doff.ptr),
             C.c_void_p(dval.ptr), None)
        return dval.download(np.float32, n)
    finally:
        for b in (dpts, dijk, doff, dval):
            b.free()


@pytest.mark.parametrize("codec", ["none", "zip"])
@pytest.mark.parametrize("k", [0, 1])
def test_device_grid_from_nanovdb_file_matches_nanovdb(codec, k):
    from rtcamp8_amd import nvdb
    from rtcamp8_amd._lib import call

    fx = np.load(GOLDEN)
    print(f"[nvdb] load grid {k} ({codec})", flush=True)
    g = nvdb.load_grid(fx[f"file_{codec}"].tobytes(), k)
    try:
        assert g.max_value == float(fx[f"g{k}_majorant"][0])
        vals = device_probe(g, fx[f"g{k}_world"])
        assert np.array_equal(vals.view(np.uint32), fx[f"g{k}_world_value"].view(np.uint32))
    finally:
        call("rtc8x_grid_destroy", C.byref(g))


def test_upload_of_a_raw_grid_buffer():
    """rtc8x_grid_upload_nanovdb takes what a GridHandle holds (no file framing)."""
    from rtcamp8_amd import nvdb
    from rtcamp8_amd._lib import call
    from test_nvdb import grid_buffers

    fx = np.load(GOLDEN)
    _, buffers = grid_buffers(fx["file_none"].tobytes())
    g = nvdb._abi.Grid()
    call("rtc8x_grid_upload_nanovdb", buffers[1], len(buffers[1]), C.byref(g))
    try:
        assert g.max_value == float(fx["g1_majorant"][0])
        vals = device_probe(g, fx["g1_world"])
        assert np.array_equal(vals.view(np.uint32), fx["g1_world_value"].view(np.uint32))
    finally:
        call("rtc8x_grid_destroy", C.byref(g))


def test_render_of_a_nanovdb_volume_matches_the_oracle(tmp_path):
    """PATH_TRACE over a volume read from a NanoVDB-written file (non-zero background, translated map,
    two root tiles) is bit-identical to the FP32 oracle on the same decoded tables."""
    from oracle_lib import trace_parity_case

    fx = np.load(GOLDEN)
    path = tmp_path / "shifted.nvdb"
    path.write_bytes(fx["file_none"].tobytes())
    print("[nvdb] parity trace of grid 1", flush=True)
    res = trace_parity_case("dense128", width=24, height=24, spp=2, tracking=0, volume_file=str(path), volume_grid=1,
                            camera_position=(0.3, 0.5, 3.2), camera_target=(0.3, 0.5, 0.7), density_coeff=3.0)
    assert res["rng_equal"] and res["vertices_equal"]
    assert np.array_equal(res["gpu_accum"].view(np.uint64), res["cpu_accum"].view(np.uint64))
    assert res["lookups"] > 0


def test_device_fbm_download_encodes_like_the_oracle_grid():
    """The device-built fBm volume, downloaded and encoded, is byte-identical to the oracle's build
    encoded the same way (tables, values, leaf order, statistics)."""
    from rtcamp8_amd import nvdb
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.scene import FbmConfig
    from oracle_lib import oracle_build_grid

    fb = FbmConfig(resolution=64, seed=12, sparse=True, active_leaf_fraction=0.3, contrast=2.0)
    g = nvdb._abi.Grid()
    call("rtc8x_grid_build_fbm", C.byref(fb.desc()), C.byref(g), None)
    try:
        with nvdb.download(g) as h:
            ours = nvdb.encode_grid(h.grid, "fbm")
    finally:
        call("rtc8x_grid_destroy", C.byref(g))
    assert ours == nvdb.encode_grid(oracle_build_grid(fb).grid, "fbm")


def test_benchmark_volume_round_trip(tmp_path):
    """sparse512 (65,536 leaves, 143 MB as NanoVDB) -> .nvdb (ZIP) -> device: identical lookups at 2^20
    random points, identical majorant, and an identical rendered frame."""
    from rtcamp8_amd import nvdb
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.tracer import PathTracer

    cfg = sc.CONFIGS["sparse512"].with_(width=96, height=64, spp=2)
    ds = sc.DeviceScene(cfg)
    path = tmp_path / "sparse512.nvdb"
    print("[nvdb] writing sparse512", flush=True)
    size = nvdb.save_device_grid(ds.grid, path, "sparse512", codec="zip")
    assert size < 200 << 20
    info = nvdb.list_grids(path)[0]
    assert info.node_count[0] == ds.grid.num_leaves and info.grid_type == 1
    loaded = nvdb.load_grid(path)
    try:
        assert loaded.max_value == ds.grid.max_value
        assert loaded.num_leaves == ds.grid.num_leaves
        for f in ("world_bbox_min", "world_bbox_max", "inv_mat", "vec"):
            assert list(getattr(loaded, f)) == list(getattr(ds.grid, f)), f
        rng = np.random.default_rng(1)
        pts = rng.random((1 << 20, 3), dtype=np.float32) * np.float32(1.1) - np.float32(0.05)
        a, b = device_probe(ds.grid, pts), device_probe(loaded, pts)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))

        images = []
        for grid in (ds.grid, loaded):
            print("[nvdb] frame", flush=True)
            saved = ds.grid
            ds.grid = grid
            pt = PathTracer(ds)
            pt.reset(7)
            pt.trace(2)
            images.append(pt.image())
            pt.close()
            ds.grid = saved
        assert np.array_equal(images[0].view(np.uint64), images[1].view(np.uint64))
    finally:
        call("rtc8x_grid_destroy", C.byref(loaded))
        ds.close()
