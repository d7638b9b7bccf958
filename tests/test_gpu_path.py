"""GPU parity of the plain delta tracker: gfx950 library vs the CPU oracle on identical inputs.

The bar is bit-exact for the RNG streams, voxel/node indexing and path-vertex counts, and for the
radiance too (the FP32 arithmetic is IEEE-identical on both sides; accumulation is FP64 in the same
order).  A relative tolerance of 1e-5 is the documented floor; these tests assert exact equality.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import download_grid, oracle, oracle_build_grid, ptr, trace_parity_case

pytestmark = pytest.mark.gpu


def _fbm(res, sparse, seed=0x5EED0001, frac=0.25):
    from rtcamp8_amd.scene import FbmConfig

    return FbmConfig(resolution=res, seed=seed, sparse=sparse, active_leaf_fraction=frac)


@pytest.mark.parametrize("res,sparse", [(64, False), (64, True), (128, True)])
def test_grid_build_matches_oracle(res, sparse):
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import call

    fbm = _fbm(res, sparse)
    dg = _abi.Grid()
    call("rtc8x_grid_build_fbm", C.byref(fbm.desc()), C.byref(dg), None)
    try:
        got = download_grid(dg)
        ref = oracle_build_grid(fbm)
        for f in ("num_lower", "num_leaves", "max_value"):
            assert getattr(got.grid, f) == getattr(ref.grid, f), f
        for f in ("world_bbox_min", "world_bbox_max", "index_bbox_min", "index_bbox_max", "cell_origin", "cell_dims",
                  "region_dims"):
            assert list(getattr(got.grid, f)) == list(getattr(ref.grid, f)), f
        for f in ("region_cells", "region_unit"):
            assert getattr(got.grid, f) == getattr(ref.grid, f), f
        for k in ("upper", "lower", "values", "origins", "cells", "regions"):
            assert np.array_equal(got.arrays[k], ref.arrays[k]), k
        if sparse:
            frac = got.grid.num_leaves / (res // 8) ** 3
            assert abs(frac - 0.25) < 0.01
    finally:
        call("rtc8x_grid_destroy", C.byref(dg))


def test_probe_indices_and_values_match_oracle():
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer

    fbm = _fbm(128, True)
    dg = _abi.Grid()
    call("rtc8x_grid_build_fbm", C.byref(fbm.desc()), C.byref(dg), None)
    try:
        rng = np.random.default_rng(7)
        n = 20000
        pts = rng.uniform(-0.05, 1.05, size=(n, 3)).astype(np.float32)
        pts[:2000] = (np.floor(pts[:2000] * 128) / 128).astype(np.float32)  # exact voxel corners
        dpts, dijk, doff, dval = DeviceBuffer.from_array(pts), DeviceBuffer(12 * n), DeviceBuffer(12 * n), DeviceBuffer(4 * n)
        call("rtc8x_grid_probe", C.byref(dg), n, C.c_void_p(dpts.ptr), C.c_void_p(dijk.ptr), C.c_void_p(doff.ptr),
             C.c_void_p(dval.ptr), None)
        hg = download_grid(dg)
        ijk = np.empty(3 * n, np.int32)
        off = np.empty(3 * n, np.uint32)
        val = np.empty(n, np.float32)
        oracle().oracle_grid_probe(C.byref(hg.grid), n, ptr(pts), ptr(ijk), ptr(off), ptr(val))
        assert np.array_equal(dijk.download(np.int32, 3 * n), ijk)
        assert np.array_equal(doff.download(np.uint32, 3 * n), off)
        assert np.array_equal(dval.download(np.float32, n).view(np.uint32), val.view(np.uint32))
    finally:
        call("rtc8x_grid_destroy", C.byref(dg))


@pytest.mark.parametrize("tracking", [0, 1])
def test_plain_tracer_dense_bit_exact(tracking):
    res = trace_parity_case("dense128", width=48, height=40, spp=3, tracking=tracking, fbm_resolution=64,
                            samples_per_launch=1)
    assert res["rng_equal"]
    assert res["vertices_equal"]
    assert np.array_equal(res["gpu_accum"].view(np.uint64), res["cpu_accum"].view(np.uint64))
    assert res["paths"] == 48 * 40 * 3
    assert res["vertices"] == int(res["cpu_stats"][1])
    assert res["lookups"] == int(res["cpu_stats"][2])


@pytest.mark.parametrize("tracking", [0, 1])
def test_plain_tracer_sparse_constant_sky_bit_exact(tracking):
    from rtcamp8_amd import _abi

    res = trace_parity_case("sparse512", width=40, height=24, spp=2, tracking=tracking, fbm_resolution=128,
                            env_kind="constant", sun_enabled=True, max_path_length=64)
    assert res["rng_equal"]
    assert res["vertices_equal"]
    assert np.array_equal(res["gpu_accum"].view(np.uint64), res["cpu_accum"].view(np.uint64))
    assert res["vertices"] == int(res["cpu_stats"][_abi.STAT_VERTICES])


@pytest.mark.parametrize("tracking", [0, 1])
def test_plain_tracer_ragged_image_bit_exact(tracking):
    """An image that is not a multiple of the 8x8 work tiles, several samples per launch."""
    res = trace_parity_case("dense128", width=37, height=29, spp=4, tracking=tracking, fbm_resolution=64,
                            samples_per_launch=2)
    assert res["rng_equal"] and res["vertices_equal"]
    assert np.array_equal(res["gpu_accum"].view(np.uint64), res["cpu_accum"].view(np.uint64))
    assert res["paths"] == 37 * 29 * 4
    assert res["lookups"] == int(res["cpu_stats"][2])


def test_schlick_reference_mode_bit_exact():
    """Reference parity mode: Schlick phase, RR from the first vertex, unbounded paths, env only."""
    from rtcamp8_amd import _abi

    res = trace_parity_case("dense128", width=32, height=32, spp=2, tracking=0, fbm_resolution=64,
                            phase=_abi.PHASE_SCHLICK, forwardness=0.3, max_path_length=0, sun_enabled=False)
    assert res["rng_equal"] and res["vertices_equal"]
    assert np.array_equal(res["gpu_accum"].view(np.uint64), res["cpu_accum"].view(np.uint64))


def test_uploaded_grid_matches_nanovdb_fixture():
    """rtc8x_grid_upload + rtc8x_grid_probe reproduce NanoVDB's own lookups (several root tiles)."""
    from rtcamp8_amd import _abi
    from rtcamp8_amd._lib import call
    from rtcamp8_amd.device import DeviceBuffer

    from oracle_lib import nanovdb_fixture, tables_from_leaves

    fx = nanovdb_fixture()
    hg = tables_from_leaves(fx["leaf_origins"], fx["leaf_values"], float(fx["voxel_size"]), fx["world_bbox"])
    dg = _abi.Grid()
    call("rtc8x_grid_upload", C.byref(hg.grid), C.byref(dg))
    try:
        assert dg.max_value == float(np.max(fx["leaf_values"]))
        n = fx["p_world"].shape[0]
        dpts = DeviceBuffer.from_array(np.ascontiguousarray(fx["p_world"], dtype=np.float32))
        dijk, doff, dval = DeviceBuffer(12 * n), DeviceBuffer(12 * n), DeviceBuffer(4 * n)
        call("rtc8x_grid_probe", C.byref(dg), n, C.c_void_p(dpts.ptr), C.c_void_p(dijk.ptr), C.c_void_p(doff.ptr),
             C.c_void_p(dval.ptr), None)
        assert np.array_equal(dijk.download(np.int32, 3 * n).reshape(n, 3), fx["p_ijk"])
        assert np.array_equal(doff.download(np.uint32, 3 * n).reshape(n, 3), fx["p_offsets"])
        assert np.array_equal(dval.download(np.float32, n).view(np.uint32), fx["p_value"].view(np.uint32))
    finally:
        call("rtc8x_grid_destroy", C.byref(dg))


def test_order_by_tile_cost_groups_tiles_costliest_first():
    """The tile-grouped work order equals its restatement: 8 x 8 tiles (ragged at the right and bottom
    edges) sorted stably by quarter-octave class of their summed cost (saturating at 2^32 - 1) below the
    costliest tile's, each tile's in-image pixels row-major and contiguous."""
    from rtcamp8_amd._lib import call, lib
    from rtcamp8_amd.device import DeviceBuffer

    rng = np.random.default_rng(6)
    for w, h in ((1, 1), (9, 17), (100, 37), (64, 64), (1920, 1080)):
        n = w * h
        print(f"[order] tiles of a {w}x{h} image", flush=True)
        cost = rng.integers(0, 1 << rng.integers(1, 31, n), dtype=np.int64).astype(np.uint32)
        if n >= 64:
            cost[:8] = 0xFFFFFFFF  # the first tile's sum saturates
        dc, do = DeviceBuffer.from_array(cost), DeviceBuffer(4 * n)
        ds = DeviceBuffer(int(lib().rtc8x_order_tiles_scratch_bytes(w, h)))
        call("rtc8x_order_by_tile_cost", None, C.c_void_p(dc.ptr), w, h, C.c_void_p(do.ptr), C.c_void_p(ds.ptr))
        got = do.download(np.uint32, n).astype(np.int64)
        tx, ty = (w + 7) // 8, (h + 7) // 8
        c2 = cost.reshape(h, w).astype(np.uint64)
        def q(c):  # 4 floor(log2 c) + the two bits below the leading one
            e = c.bit_length() - 1
            return 4 * e + ((c >> (e - 2)) & 3 if e >= 2 else (c << (2 - e)) & 3)

        sums = []
        for t in range(tx * ty):
            x0, y0 = (t % tx) * 8, (t // tx) * 8
            sums.append(min(int(c2[y0:y0 + 8, x0:x0 + 8].sum()), 0xFFFFFFFF))
        top = max(sums)
        tiles = []
        for t, s in enumerate(sums):
            x0, y0 = (t % tx) * 8, (t // tx) * 8
            cls = 31 if s == 0 else min(31, q(top) - q(s))
            ys, xs = np.mgrid[y0:min(y0 + 8, h), x0:min(x0 + 8, w)]
            tiles.append((cls, t, (ys * w + xs).ravel()))
        want = np.concatenate([px for _, _, px in sorted(tiles, key=lambda z: (z[0], z[1]))])
        assert np.array_equal(got, want), (w, h)
        for b_ in (dc, do, ds):
            b_.free()


def test_order_by_cost_is_a_stable_costliest_first_permutation():
    from rtcamp8_amd._lib import call, lib
    from rtcamp8_amd.device import DeviceBuffer

    rng = np.random.default_rng(5)
    for n in (1, 1000, 1024, 300001):
        cost = (rng.integers(0, 1 << rng.integers(1, 30, n), dtype=np.int64) % (1 << 31)).astype(np.uint32)
        cost[: min(n, 7)] = [0, 1, 2, 3, 0xFFFFFFFF, 0xFFFFFFFE, 5][: min(n, 7)]
        dc, do = DeviceBuffer.from_array(cost), DeviceBuffer(4 * n)
        ds = DeviceBuffer(max(int(lib().rtc8x_order_scratch_bytes(n)), 4))
        call("rtc8x_order_by_cost", None, C.c_void_p(dc.ptr), n, C.c_void_p(do.ptr), C.c_void_p(ds.ptr))
        order = do.download(np.uint32, n).astype(np.int64)
        assert np.array_equal(np.sort(order), np.arange(n))
        cls = 31 - np.floor(np.log2(np.minimum(cost.astype(np.float64), 0xFFFFFFFE) + 1)).astype(np.int64)
        key = cls[order]
        assert np.all(np.diff(key) >= 0)  # costliest class first
        same = np.diff(key) == 0
        assert np.all(np.diff(order)[same] > 0)  # row-major inside a class
        for b_ in (dc, do, ds):
            b_.free()
