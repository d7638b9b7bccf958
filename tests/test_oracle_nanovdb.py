"""The oracle's grid indexing and ray clipping pinned to NanoVDB itself.

tests/golden/nanovdb_probe.npz holds what the NanoVDB that rtcamp8 vendors computes (built from
ext/nanovdb by tests/golden/make_nanovdb_fixture.py): Floor coordinates, upper/lower/leaf offsets,
uncached trilinear values and Ray<float>::clip results.  The oracle must reproduce them exactly.
"""
import ctypes as C

import numpy as np

from oracle_lib import nanovdb_fixture, oracle, ptr, tables_from_leaves


def _probe_grid(fx):
    return tables_from_leaves(fx["leaf_origins"], fx["leaf_values"], float(fx["voxel_size"]), fx["world_bbox"])


def test_oracle_trilinear_and_offsets_match_nanovdb():
    fx = nanovdb_fixture()
    hg = _probe_grid(fx)
    n = fx["p_world"].shape[0]
    pts = np.ascontiguousarray(fx["p_world"], dtype=np.float32)
    ijk = np.empty(3 * n, np.int32)
    off = np.empty(3 * n, np.uint32)
    val = np.empty(n, np.float32)
    oracle().oracle_grid_probe(C.byref(hg.grid), n, ptr(pts), ptr(ijk), ptr(off), ptr(val))
    assert np.array_equal(ijk.reshape(n, 3), fx["p_ijk"])
    assert np.array_equal(off.reshape(n, 3), fx["p_offsets"])
    assert np.array_equal(val.view(np.uint32), fx["p_value"].view(np.uint32))
    # the fixture exercises several root tiles and both empty and occupied space
    assert hg.grid.num_root_tiles >= 4
    assert (fx["p_value"] == 0).sum() > 100 and (fx["p_value"] > 0).sum() > 1000


def test_oracle_ray_clip_matches_nanovdb():
    fx = nanovdb_fixture()
    bmin = np.ascontiguousarray(fx["world_bbox"][:3], dtype=np.float32)
    bmax = np.ascontiguousarray(fx["world_bbox"][3:], dtype=np.float32)
    t0 = np.zeros(1, np.float32)
    t1 = np.zeros(1, np.float32)
    hits = 0
    for i in range(fx["r_eye"].shape[0]):
        eye = np.ascontiguousarray(fx["r_eye"][i], dtype=np.float32)
        d = np.ascontiguousarray(fx["r_dir"][i], dtype=np.float32)
        hit = oracle().oracle_ray_clip(ptr(eye), ptr(d), C.c_float(fx["r_tmax"][i]), ptr(bmin), ptr(bmax), ptr(t0), ptr(t1))
        assert hit == fx["r_hit"][i], i
        if hit:
            hits += 1
            assert t0.view(np.uint32)[0] == fx["r_t0"][i:i + 1].view(np.uint32)[0], i
            assert t1.view(np.uint32)[0] == fx["r_t1"][i:i + 1].view(np.uint32)[0], i
    assert hits > 200
