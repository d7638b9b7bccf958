"""Test-side bindings of the CPU oracle (oracle/liboracle.so) and GPU-vs-oracle harness helpers.

The oracle is the checker: it receives the same parameter block as the GPU launch, with host copies
of every buffer, and its results are compared with what the gfx950 library produced.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from rtcamp8_amd import _abi

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"

_P = C.c_void_p
_orc = None


def oracle() -> C.CDLL:
    global _orc
    if _orc is not None:
        return _orc
    if not ORACLE_SO.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    _orc = load_oracle(ORACLE_SO)
    return _orc


def load_oracle(path) -> C.CDLL:
    """An oracle library at `path` with its prototypes set (the built one, or a copy built elsewhere)."""
    lib = C.CDLL(str(path))
    protos = {
        "oracle_pcg_next": (C.c_uint32, [C.POINTER(C.c_uint64)]),
        "oracle_pcg_float": (C.c_float, [C.POINTER(C.c_uint64)]),
        "oracle_pcg_fill": (None, [C.c_uint64, C.c_uint32, _P, _P, _P]),
        "oracle_mt19937_64": (None, [C.c_uint64, C.c_uint32, _P]),
        "oracle_seed_rng": (None, [_P, C.c_uint32, C.c_uint32, C.c_uint64]),
        "oracle_lcg_next": (C.c_uint32, [C.POINTER(C.c_uint32)]),
        "oracle_lcg_seeds": (None, [C.c_uint32, C.c_uint32, _P]),
        "oracle_grid_build_fbm": (C.c_int, [C.POINTER(_abi.FbmDesc), C.POINTER(_abi.Grid)]),
        "oracle_grid_free": (None, [C.POINTER(_abi.Grid)]),
        "oracle_fbm_value": (C.c_float, [C.POINTER(_abi.FbmDesc), C.c_int32, C.c_int32, C.c_int32]),
        "oracle_grid_get_value": (C.c_float, [C.POINTER(_abi.Grid), C.c_int32, C.c_int32, C.c_int32, _P]),
        "oracle_grid_probe": (None, [C.POINTER(_abi.Grid), C.c_uint32, _P, _P, _P, _P]),
        "oracle_compute_cell_majorants": (None, [C.POINTER(_abi.Grid), C.c_uint32]),
        "oracle_compute_region_majorants": (None, [C.POINTER(_abi.Grid), C.c_uint32]),
        "oracle_ray_clip": (C.c_int, [_P, _P, C.c_float, _P, _P, _P, _P]),
        "oracle_logf": (C.c_float, [C.c_float]),
        "oracle_sincosf": (None, [C.c_float, _P, _P]),
        "oracle_atan2f": (C.c_float, [C.c_float, C.c_float]),
        "oracle_acosf": (C.c_float, [C.c_float]),
        "oracle_trace_f32": (C.c_int, [C.POINTER(_abi.FrameParams), C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.c_uint32]),
        "oracle_trace_f64": (C.c_int, [C.POINTER(_abi.FrameParams), C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.c_uint32]),
        "oracle_trace_window_f64": (C.c_int, [C.POINTER(_abi.FrameParams), C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.c_uint32, C.c_uint32, C.c_uint32]),
        "oracle_env_sample_f64": (None, [C.POINTER(_abi.EnvLight), C.c_double, C.c_double, _P, _P, _P]),
        "oracle_env_evaluate_f64": (None, [C.POINTER(_abi.EnvLight), _P, _P, _P]),
        "oracle_nrc_prepare": (None, [C.POINTER(_abi.FrameParams), C.c_int32]),
        "oracle_nrc_propagate": (None, [C.POINTER(_abi.FrameParams)]),
        "oracle_nrc_shuffle": (None, [C.POINTER(_abi.FrameParams)]),
        "oracle_nrc_accumulate": (None, [C.POINTER(_abi.FrameParams)]),
        "oracle_nrc_num_params": (C.c_uint32, [C.c_uint32]),
        "oracle_nrc_encode": (None, [_P, C.c_uint32, _P]),
        "oracle_nrc_forward": (None, [_P, C.c_uint32, _P, C.c_uint32, _P, C.c_int]),
        "oracle_nrc_loss": (C.c_double, [_P, C.c_uint32, _P, _P, C.c_uint32, C.c_int]),
        "oracle_nrc_gradient": (C.c_double, [_P, C.c_uint32, _P, _P, C.c_uint32, _P, C.c_int]),
        "oracle_nrc_num_params_ex": (C.c_uint32, [C.c_uint32, C.c_int32]),
        "oracle_nrc_encode_ex": (None, [_P, C.c_uint32, C.c_int32, _P, C.c_uint32, _P]),
        "oracle_nrc_forward_ex": (None, [_P, C.c_uint32, C.c_int32, _P, C.c_uint32, _P, C.c_int]),
        "oracle_nrc_gradient_ex": (C.c_double, [_P, C.c_uint32, C.c_int32, _P, _P, C.c_uint32, _P, C.c_int]),
        "oracle_nrc_adam": (None, [_P, _P, _P, _P, _P, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_float, C.c_uint64]),
        "oracle_round_fp16": (C.c_float, [C.c_float]),
        "oracle_expf": (C.c_float, [C.c_float]),
        "oracle_powf": (C.c_float, [C.c_float, C.c_float]),
        "oracle_tone_map": (None, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                   C.POINTER(_abi.SdrConfig), _P, _P, C.c_int32]),
        "oracle_sdr_convert": (None, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_abi.SdrConfig), _P,
                                      C.c_int32]),
        "oracle_round_bf16": (C.c_float, [C.c_float]),
        "oracle_track_batch_f32": (None, [C.POINTER(_abi.FrameParams), C.c_int32, C.c_uint32, _P, _P, _P, C.c_float,
                                          _P, _P]),
        "oracle_track_batch_f64": (None, [C.POINTER(_abi.FrameParams), C.c_int32, C.c_uint32, _P, _P, _P, C.c_float,
                                          _P, _P]),
        "oracle_phase_sample_batch_f32": (None, [C.c_int32, C.c_float, C.c_uint32, _P, _P, _P]),
        "oracle_phase_sample_batch_f64": (None, [C.c_int32, C.c_float, C.c_uint32, _P, _P, _P]),
        "oracle_phase_value_f32": (C.c_double, [C.c_int32, C.c_float, C.c_double]),
        "oracle_phase_value_f64": (C.c_double, [C.c_int32, C.c_float, C.c_double]),
        "oracle_env_build_mips": (C.c_uint32, [_P, C.c_uint32, C.c_uint32, _P, _P, _P, _P]),
    }
    for name, (res, args) in protos.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def ptr(a: np.ndarray) -> int:
    return int(a.ctypes.data)


def nthreads() -> int:
    return max(1, min(8, os.cpu_count() or 1))


class HostGrid:
    """A grid whose tables live in numpy arrays (host pointers in `.grid`)."""

    def __init__(self, grid: _abi.Grid, arrays: dict):
        self.grid = grid
        self.arrays = arrays


def oracle_build_grid(fbm, lib=None) -> HostGrid:
    """The oracle's own fBm grid build (independent of the library's device build); `lib`: another build of
    the oracle to use."""
    lib = lib or oracle()
    g = _abi.Grid()
    rc = lib.oracle_grid_build_fbm(C.byref(fbm.desc()), C.byref(g))
    assert rc == 0, f"oracle grid build failed: {rc}"
    arrays = {
        "root": np.ctypeslib.as_array((_abi.RootTile * g.num_root_tiles).from_address(g.root_tiles)).copy(),
        "upper": np.ctypeslib.as_array((C.c_uint32 * (g.num_upper * _abi.UPPER_TABLE_SIZE)).from_address(g.upper_tables)).copy(),
        "lower": np.ctypeslib.as_array((C.c_uint32 * (g.num_lower * _abi.LOWER_TABLE_SIZE)).from_address(g.lower_tables)).copy(),
        "values": np.ctypeslib.as_array((C.c_float * (g.num_leaves * _abi.LEAF_SIZE)).from_address(g.leaf_values)).copy(),
        "tiles": np.ctypeslib.as_array((C.c_float * 1).from_address(g.tile_values)).copy(),
        "origins": np.ctypeslib.as_array((C.c_int32 * (3 * g.num_leaves)).from_address(g.leaf_origins)).copy(),
        "cells": np.ctypeslib.as_array((C.c_float * int(np.prod(g.cell_dims[:]))).from_address(g.cell_majorants)).copy(),
        "regions": np.ctypeslib.as_array((C.c_uint8 * int(np.prod(g.region_dims[:]))).from_address(g.region_majorants)).copy(),
    }
    template = _abi.Grid()
    C.memmove(C.byref(template), C.byref(g), C.sizeof(_abi.Grid))
    lib.oracle_grid_free(C.byref(g))
    return host_grid_from_arrays(template, arrays)


def host_grid_from_arrays(template: _abi.Grid, arrays: dict) -> HostGrid:
    g = _abi.Grid()
    C.memmove(C.byref(g), C.byref(template), C.sizeof(_abi.Grid))
    g.root_tiles = arrays["root"].ctypes.data
    g.upper_tables = arrays["upper"].ctypes.data
    g.lower_tables = arrays["lower"].ctypes.data
    g.leaf_values = arrays["values"].ctypes.data
    g.tile_values = arrays["tiles"].ctypes.data
    g.leaf_origins = arrays["origins"].ctypes.data if arrays.get("origins") is not None else None
    g.cell_majorants = arrays["cells"].ctypes.data if arrays.get("cells") is not None else None
    g.region_majorants = arrays["regions"].ctypes.data if arrays.get("regions") is not None else None
    return HostGrid(g, arrays)


def download_grid(dg: _abi.Grid) -> HostGrid:
    """Host copy of a device grid built by the library."""
    from rtcamp8_amd.device import download_ptr

    root = np.empty(dg.num_root_tiles, dtype=np.dtype([("key", "<u8"), ("entry", "<u4"), ("state", "<u4")]))
    raw = download_ptr(dg.root_tiles, np.uint8, root.nbytes)
    root = raw.view(root.dtype).copy()
    arrays = {
        "root": root,
        "upper": download_ptr(dg.upper_tables, np.uint32, dg.num_upper * _abi.UPPER_TABLE_SIZE),
        "lower": download_ptr(dg.lower_tables, np.uint32, dg.num_lower * _abi.LOWER_TABLE_SIZE),
        "values": download_ptr(dg.leaf_values, np.float32, dg.num_leaves * _abi.LEAF_SIZE),
        "tiles": download_ptr(dg.tile_values, np.float32, 1),
        "origins": download_ptr(dg.leaf_origins, np.int32, 3 * dg.num_leaves) if dg.leaf_origins else None,
        "cells": download_ptr(dg.cell_majorants, np.float32, int(np.prod(dg.cell_dims[:]))) if dg.cell_majorants else None,
        "regions": download_ptr(dg.region_majorants, np.uint8, int(np.prod(dg.region_dims[:]))) if dg.region_majorants else None,
    }
    return host_grid_from_arrays(dg, arrays)


def oracle_env_mips(texels: np.ndarray, w: int, h: int):
    """The oracle's own hierarchical importance map of a texture: (mips, dim_x, dim_y, levels, integral)."""
    dx, dy = C.c_uint32(), C.c_uint32()
    levels = oracle().oracle_env_build_mips(ptr(texels), w, h, None, C.byref(dx), C.byref(dy), None)
    assert levels > 0, "map must be 1:1 or 2:1"
    total = sum(max(dx.value >> l, 1) * max(dy.value >> l, 1) for l in range(levels))
    mips, integral = np.zeros(total, np.float32), np.zeros(1, np.float32)
    oracle().oracle_env_build_mips(ptr(texels), w, h, ptr(mips), None, None, ptr(integral))
    return mips, dx.value, dy.value, levels, float(integral[0])


def download_env(de: _abi.EnvLight):
    """Host copy of a device environment light (texels, alias tables).  The hierarchical importance map is
    the oracle's own, built from the texels, and must equal the library's bit for bit."""
    from rtcamp8_amd.device import download_ptr

    e = _abi.EnvLight()
    C.memmove(C.byref(e), C.byref(de), C.sizeof(_abi.EnvLight))
    keep = {}
    if de.kind == _abi.ENV_LATLONG:
        n = de.width * de.height
        keep["texels"] = download_ptr(de.texels, np.float32, 4 * n)
        e.texels = keep["texels"].ctypes.data
        mips, dx, dy, levels, integral = oracle_env_mips(keep["texels"], de.width, de.height)
        assert (dx, dy, levels) == (de.map_dim_x, de.map_dim_y, de.num_mip_levels)
        gpu_mips = download_ptr(de.mips, np.float32, mips.size)
        assert np.array_equal(gpu_mips.view(np.uint32), mips.view(np.uint32)), "importance map differs from the oracle's"
        assert np.float32(integral).view(np.uint32) == np.float32(de.map_integral).view(np.uint32)
        keep["mips"] = mips
        e.mips = keep["mips"].ctypes.data
        keep["rows"] = download_ptr(de.alias_rows, np.uint32, 2 * de.height)
        keep["cols"] = download_ptr(de.alias_cols, np.uint32, 2 * n)
        keep["weights"] = download_ptr(de.alias_weights, np.float32, n)
        e.alias_rows = keep["rows"].ctypes.data
        e.alias_cols = keep["cols"].ctypes.data
        e.alias_weights = keep["weights"].ctypes.data
    return e, keep


def oracle_seed(width: int, height: int, seed: int) -> np.ndarray:
    s = np.empty(width * height, dtype=np.uint64)
    oracle().oracle_seed_rng(ptr(s), width, height, C.c_uint64(seed))
    return s


def trace_parity_case(config: str, width: int, height: int, spp: int, tracking: int, fbm_resolution: int | None = None,
                      samples_per_launch: int | None = None, **overrides) -> dict:
    """Runs PATH_TRACE on the GPU and the FP32 oracle on identical inputs; returns both results."""
    from rtcamp8_amd import scene as sc
    from rtcamp8_amd.tracer import PathTracer

    cfg = sc.CONFIGS[config].with_(width=width, height=height, spp=spp, tracking=tracking, **overrides)
    if fbm_resolution is not None:
        cfg = cfg.with_(fbm=sc.FbmConfig(**{**cfg.fbm.__dict__, "resolution": fbm_resolution}))
    ds = sc.DeviceScene(cfg)
    pt = PathTracer(ds, track_vertices=True)
    per = samples_per_launch or spp
    for _ in range(spp // per):
        pt.trace(per)
    gpu_accum = pt.accum.download(np.float64, 3 * width * height)
    gpu_rng = pt.rng.download(np.uint64, width * height)
    gpu_vert = pt.vertices.download(np.uint32, width * height)
    counters = pt.counters()

    hg = download_grid(ds.grid)
    he, keep = download_env(ds.env)
    p = sc.fill_params(cfg, hg.grid, he, ds.camera)
    rng = oracle_seed(width, height, cfg.rng_seed)
    accum = np.zeros(3 * width * height, dtype=np.float64)
    verts = np.zeros(width * height, dtype=np.uint32)
    stats = np.zeros(_abi.STAT_COUNT, dtype=np.uint64)
    p.rng_states, p.accum, p.path_vertices, p.stats = ptr(rng), ptr(accum), ptr(verts), ptr(stats)
    for _ in range(spp // per):
        rc = oracle().oracle_trace_f32(C.byref(p), _abi.ENTRY_PATH_TRACE, per, 0, height, nthreads())
        assert rc == 0
    pt.close()
    ds.close()
    return {
        "gpu_accum": gpu_accum, "cpu_accum": accum,
        "rng_equal": bool(np.array_equal(gpu_rng, rng)),
        "vertices_equal": bool(np.array_equal(gpu_vert, verts)),
        "gpu_vertices": gpu_vert, "cpu_vertices": verts,
        "paths": counters["paths"], "vertices": counters["vertices"], "lookups": counters["lookups"],
        "cpu_stats": stats,
    }


def nanovdb_fixture():
    """The NanoVDB ground truth (tests/golden/nanovdb_probe.npz) with the probe grid's voxel values
    regenerated from the probe's hash (value = 0.05 + (mix(l*512+v ^ 0xA5A5A5A5) >> 8) / 2^24)."""
    fx = dict(np.load(ROOT / "tests" / "golden" / "nanovdb_probe.npz"))
    nl = fx["leaf_origins"].shape[0]
    x = (np.arange(nl * 512, dtype=np.uint64) ^ np.uint64(0xA5A5A5A5)).astype(np.uint32)
    x ^= x >> np.uint32(16)
    x = (x.astype(np.uint64) * 0x7FEB352D & 0xFFFFFFFF).astype(np.uint32)
    x ^= x >> np.uint32(15)
    x = (x.astype(np.uint64) * 0x846CA68B & 0xFFFFFFFF).astype(np.uint32)
    x ^= x >> np.uint32(16)
    vals = np.float32(0.05) + (x >> np.uint32(8)).astype(np.float32) * np.float32(1.0 / 16777216.0)
    fx["leaf_values"] = vals.reshape(nl, 512)
    assert np.array_equal(fx["leaf_values"][0], fx["leaf0_values"]), "probe value formula drifted"
    return fx


def tables_from_leaves(origins: np.ndarray, values: np.ndarray, voxel_size: float, world_bbox) -> HostGrid:
    """Builds the library's NanoVDB-layout tables for arbitrary leaves (several root tiles allowed)."""
    origins = np.asarray(origins, dtype=np.int32).reshape(-1, 3)
    u = origins.astype(np.int64) & 0xFFFFFFFF
    keys = (u[:, 2] >> 12) | ((u[:, 1] >> 12) << 21) | ((u[:, 0] >> 12) << 42)
    uniq_keys = sorted(set(int(k) for k in keys))
    tile_of = {k: i for i, k in enumerate(uniq_keys)}
    uoff = (((u[:, 0] & 4095) >> 7) << 10) | (((u[:, 1] & 4095) >> 7) << 5) | ((u[:, 2] & 4095) >> 7)
    loff = (((u[:, 0] & 127) >> 3) << 8) | (((u[:, 1] & 127) >> 3) << 4) | ((u[:, 2] & 127) >> 3)
    upper = np.full(len(uniq_keys) * UPPER, TILE, dtype=np.uint32)
    lower_of = {}
    for l in range(len(origins)):
        t = tile_of[int(keys[l])]
        slot = t * UPPER + int(uoff[l])
        if slot not in lower_of:
            lower_of[slot] = len(lower_of)
            upper[slot] = lower_of[slot]
    lower = np.full(max(len(lower_of), 1) * LOWER, TILE, dtype=np.uint32)
    for l in range(len(origins)):
        t = tile_of[int(keys[l])]
        li = lower_of[t * UPPER + int(uoff[l])]
        lower[li * LOWER + int(loff[l])] = l
    root = np.zeros(len(uniq_keys), dtype=np.dtype([("key", "<u8"), ("entry", "<u4"), ("state", "<u4")]))
    root["key"] = np.array(uniq_keys, dtype=np.uint64)
    root["entry"] = np.arange(len(uniq_keys), dtype=np.uint32)
    g = _abi.Grid()
    inv = np.float32(1.0 / voxel_size)
    g.inv_mat[:] = [inv, 0, 0, 0, inv, 0, 0, 0, inv]
    g.vec[:] = [0.0, 0.0, 0.0]
    g.world_bbox_min[:] = list(np.asarray(world_bbox[:3], dtype=np.float32))
    g.world_bbox_max[:] = list(np.asarray(world_bbox[3:], dtype=np.float32))
    g.index_bbox_min[:] = origins.min(axis=0).tolist()
    g.index_bbox_max[:] = (origins.max(axis=0) + 7).tolist()
    g.max_value = float(np.max(values))
    g.num_root_tiles, g.num_upper, g.num_lower, g.num_leaves = len(uniq_keys), len(uniq_keys), len(lower_of), len(origins)
    arrays = {"root": root, "upper": upper, "lower": lower,
              "values": np.ascontiguousarray(values, dtype=np.float32).reshape(-1),
              "tiles": np.zeros(1, np.float32), "origins": origins.reshape(-1).copy(), "cells": None}
    return host_grid_from_arrays(g, arrays)


UPPER = _abi.UPPER_TABLE_SIZE
LOWER = _abi.LOWER_TABLE_SIZE
TILE = _abi.TILE_BIT


class HostNrc:
    """Host copies of an NrcBuffers set (for running the oracle's NRC data path on the same inputs)."""

    FIELDS = {  # name -> (dtype, elements per unit, units fn)
        "num_training_data": (np.uint32, 2), "train_image_size": (np.int32, 4), "target_min_max": (np.float32, 12),
        "target_avg": (np.float32, 6),
    }

    def __init__(self, dev):
        from rtcamp8_amd.nrc import round_up

        n = dev.width * dev.height
        k = getattr(dev, "samples", 1)  # image slices of the render buffers
        ninf = round_up(n, 128)
        self.width, self.height = dev.width, dev.height
        self.a = {
            "num_training_data": dev.num_training_data.download(np.uint32, 2),
            "train_image_size": dev.train_image_size.download(np.int32, 4),
            "target_min_max": dev.target_min_max.download(np.float32, 12),
            "target_avg": dev.target_avg.download(np.float32, 6),
            "inference_queries": dev.inference_queries.download(np.float32, 14 * ninf * k),
            "inference_terminals": dev.inference_terminals.download(np.uint32, 4 * n * k),
            "train_suffix_terminals": dev.train_suffix_terminals.download(np.uint32, 4 * n),
            "inferred_radiance": dev.inferred_radiance.download(np.float32, 3 * ninf * k),
            "per_frame_contribution": dev.per_frame_contribution.download(np.float32, 3 * n * k),
            "train_queries0": dev.train_queries[0].download(np.float32, 14 * _abi.TRAIN_BUFFER_SIZE),
            "train_queries1": dev.train_queries[1].download(np.float32, 14 * _abi.TRAIN_BUFFER_SIZE),
            "train_targets0": dev.train_targets[0].download(np.float32, 3 * _abi.TRAIN_BUFFER_SIZE),
            "train_targets1": dev.train_targets[1].download(np.float32, 3 * _abi.TRAIN_BUFFER_SIZE),
            "train_vertex_info": dev.train_vertex_info.download(np.uint32, 6 * _abi.TRAIN_BUFFER_SIZE),
            "data_shufflers": dev.data_shufflers.download(np.uint32, _abi.NUM_TRAINING_DATA_PER_FRAME),
        }

    def fill(self, nrc):
        a = self.a
        nrc.max_num_training_suffixes = self.width * self.height
        nrc.num_training_data = ptr(a["num_training_data"])
        nrc.train_image_size = ptr(a["train_image_size"])
        nrc.target_min_max = ptr(a["target_min_max"])
        nrc.target_avg = ptr(a["target_avg"])
        nrc.inference_queries = ptr(a["inference_queries"])
        nrc.inference_terminals = ptr(a["inference_terminals"])
        nrc.train_suffix_terminals = ptr(a["train_suffix_terminals"])
        nrc.inferred_radiance = ptr(a["inferred_radiance"])
        nrc.per_frame_contribution = ptr(a["per_frame_contribution"])
        nrc.train_queries[0], nrc.train_queries[1] = ptr(a["train_queries0"]), ptr(a["train_queries1"])
        nrc.train_targets[0], nrc.train_targets[1] = ptr(a["train_targets0"]), ptr(a["train_targets1"])
        nrc.train_vertex_info = ptr(a["train_vertex_info"])
        nrc.data_shufflers = ptr(a["data_shufflers"])
        nrc.scratch = None
