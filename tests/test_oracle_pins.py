"""Pins of the oracle's tracker, phase sampler, NRC data path and importance map against what the
mathematics fixes, independent of the oracle's own code (CPU only).

* Free-flight sampling (optix_pathtracing_kernels.cu:98-126) and ratio tracking
  (renderer_kernel_common.h:360-383) in a homogeneous medium whose majorant is four times its density:
  collision distances are Exp(sigma) with escape probability e^(-sigma d), the number of tentative
  collisions is the majorant's Poisson process, and the ratio-tracking estimator T = (1 - sigma/mu)^K has
  E[T] = e^(-sigma d) and E[T^2] = e^(mu d ((1 - sigma/mu)^2 - 1)).
* The same trackers in a heterogeneous fBm medium against the optical depth integrated by brute force
  (a 20,000-point midpoint rule over the trilinear density along the ray).
* Phase sampling (common_renderer_types.h:2063-2097): histograms of the sampled cosine against the
  closed-form distributions of Henyey-Greenstein and Schlick, uniform azimuth, and the returned value equal
  to the closed-form density.
* propagateRadianceValues (nrc_set_up_kernels.cu:46-86) against the explicit sum over each path's
  vertices, L_j = sum_i (prod_{l<i} T_l) E_i + (prod T) L_terminal, built from explicit vertex lists.
* accumulateInferredRadianceValues (nrc_set_up_kernels.cu:176-219) against direct + T x NRC x reflectance
  per pixel and sample.
* The hierarchical importance map (compute_light_probs.cu:161-211): mip 0 = luminance x (pi/2) sin(theta)
  at texel centres, every level the mean of its 2^l x 2^l block of mip 0, the integral the sum of mip 0.

Every check is a function of the oracle library, so `test_planted_mistakes_fail_their_pins` can run it on
copies of the oracle built with one plausible mistake each (a flipped comparison, a dropped term, a wrong
index) and show that the pin fails.
"""
from __future__ import annotations

import ctypes as C
import math
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np
import pytest

from oracle_lib import ROOT, load_oracle, oracle, oracle_build_grid, ptr, tables_from_leaves
from rtcamp8_amd import _abi
from rtcamp8_amd.scene import FbmConfig

Z = 4.5  # standard errors allowed for a statistical check (two-sided p ~ 7e-6 each)


# ------------------------------------------------------------------ media
def with_majorants(hg, region_cells: int = 0):
    """The host grid with the oracle's own cell and region majorants (local-majorant tracking); region_cells
    0: the default region size."""
    oracle().oracle_compute_cell_majorants(C.byref(hg.grid), region_cells)
    g = hg.grid
    hg.arrays["cells"] = np.ctypeslib.as_array((C.c_float * int(np.prod(g.cell_dims[:]))).from_address(g.cell_majorants)).copy()
    hg.arrays["regions"] = np.ctypeslib.as_array((C.c_uint8 * int(np.prod(g.region_dims[:]))).from_address(g.region_majorants)).copy()
    g.cell_majorants = hg.arrays["cells"].ctypes.data
    g.region_majorants = hg.arrays["regions"].ctypes.data
    return hg


def homogeneous_medium():
    """Index space = world space (voxel size 1); 3 x 3 x 3 leaves of density 1 cover [0, 24)^3, except the
    voxel (23, 23, 23) = 4, which sets the grid maximum and with it the majorant to 4 x the density.  The
    world box [4, 20]^3 is the part of the medium where every trilinear tap reads 1.  With one cell per
    region, the local majorants are 4 in the 8^3 cell (2, 2, 2), whose rim-extended span [15, 25]^3 holds
    that voxel, and 1 elsewhere (cell 1 spans [7, 17] per axis), quantised up to q / 127 of 4."""
    origins = np.array([(x, y, z) for x in (0, 8, 16) for y in (0, 8, 16) for z in (0, 8, 16)], np.int32)
    values = np.ones((27, 512), np.float32)
    values[-1, 511] = 4.0  # leaf (16, 16, 16), voxel (7, 7, 7)
    return with_majorants(tables_from_leaves(origins, values, 1.0, (4.0, 4.0, 4.0, 20.0, 20.0, 20.0)), 1)


def frame(grid: _abi.Grid, sigma: float, tracking: int = _abi.TRACKING_GLOBAL_MAJORANT) -> _abi.FrameParams:
    p = _abi.FrameParams()
    p.grid = grid
    p.density_coeff = sigma
    p.tracking = tracking
    return p


def homogeneous_majorants(o, local: bool):
    """Per ray (along +x from x = 4 to 20): the majorant on [4, 8), [8, 16), [16, 20] (the three cells)."""
    if not local:
        return np.full((o.shape[0], 3), MU)
    unit = np.float32(4.0) / np.float32(127.0)
    while np.float32(127.0) * unit < np.float32(4.0):
        unit = np.nextafter(unit, np.float32(np.inf))
    q1 = 1
    while np.float32(q1) * unit < np.float32(1.0):
        q1 += 1
    lo, hi = float(np.float32(q1) * unit), float(np.float32(127) * unit)
    inner = (o[:, 1] >= 16) & (o[:, 2] >= 16)  # cells y = z = 2; the x cell 2 then sees the 4
    m = np.full((o.shape[0], 3), lo)
    m[inner, 2] = hi
    return SIGMA * m


def track(lib, p, transmittance: bool, o, d, f64: bool, seed: int, tmax: float = 1e30):
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    n = o.shape[0]
    rng = np.random.default_rng(seed).integers(0, 2 ** 63, n, dtype=np.uint64)
    out = np.zeros(n, np.float64)
    lookups = np.zeros(1, np.uint64)
    fn = lib.oracle_track_batch_f64 if f64 else lib.oracle_track_batch_f32
    fn(C.byref(p), 1 if transmittance else 0, n, ptr(rng), ptr(o), ptr(d), tmax, ptr(out), ptr(lookups))
    return out, int(lookups[0])


def axis_rays(n: int, seed: int):
    """Rays along +x through the homogeneous box at random (y, z): they enter at x = 4 and leave at 20."""
    r = np.random.default_rng(seed)
    o = np.zeros((n, 3), np.float32)
    o[:, 1:] = r.uniform(5.0, 19.0, (n, 2))
    d = np.tile(np.array([1.0, 0.0, 0.0], np.float32), (n, 1))
    return o, d


SIGMA, CHORD, MU = 0.1, 16.0, 0.4


def check_free_flight_homogeneous(lib, f64: bool, tracking: int = _abi.TRACKING_GLOBAL_MAJORANT):
    hg = homogeneous_medium()
    p = frame(hg.grid, SIGMA, tracking)
    n = 100_000
    o, d = axis_rays(n, 11)
    t, lookups = track(lib, p, False, o, d, f64, seed=12)
    esc = t < 0
    p_esc = math.exp(-SIGMA * CHORD)
    se = math.sqrt(p_esc * (1 - p_esc) / n)
    assert abs(esc.mean() - p_esc) < Z * se, (esc.mean(), p_esc)
    # collision distances from the entry point: Exp(sigma) truncated at the chord (chi-square, 16 bins)
    s = t[~esc] - 4.0
    assert s.min() >= 0.0 and s.max() <= CHORD
    edges = np.linspace(0.0, CHORD, 17)
    counts = np.histogram(s, edges)[0]
    expect = n * np.diff(1.0 - np.exp(-SIGMA * edges))
    chi2 = float(((counts - expect) ** 2 / expect).sum())
    assert chi2 < 55.0, chi2  # 16 degrees of freedom; P(chi2 > 55) ~ 5e-6
    # tentative collisions form the majorant's Poisson process: E[lookups] = int_0^d mu(s) e^(-sigma s) ds
    # (global: mu E[min(X, d)], X ~ Exp(sigma); local: piecewise over the three cells)
    mus = homogeneous_majorants(o, tracking == _abi.TRACKING_LOCAL_MAJORANT)
    seg = np.array([[0.0, 4.0], [4.0, 12.0], [12.0, 16.0]])
    mean_lookups = float((mus * (np.exp(-SIGMA * seg[:, 0]) - np.exp(-SIGMA * seg[:, 1]))[None] / SIGMA).sum(1).mean())
    assert abs(lookups / n / mean_lookups - 1.0) < 0.01, (lookups / n, mean_lookups)


def check_ratio_tracking_homogeneous(lib, f64: bool, tracking: int = _abi.TRACKING_GLOBAL_MAJORANT):
    hg = homogeneous_medium()
    p = frame(hg.grid, SIGMA, tracking)
    n = 100_000
    o, d = axis_rays(n, 21)
    T, lookups = track(lib, p, True, o, d, f64, seed=22)
    # T = prod over the Poisson(mu) tentative collisions of (1 - sigma / mu): E[T] = e^(-sigma d) and
    # E[T^2] = exp(int mu ((1 - sigma / mu)^2 - 1)), per cell for local majorants
    mus = homogeneous_majorants(o, tracking == _abi.TRACKING_LOCAL_MAJORANT)
    length = np.array([4.0, 8.0, 4.0])
    m1 = math.exp(-SIGMA * CHORD)
    m2 = float(np.exp((mus * ((1 - SIGMA / mus) ** 2 - 1) * length[None]).sum(1)).mean())
    assert abs(T.mean() - m1) < Z * T.std() / math.sqrt(n), (T.mean(), m1)
    assert abs((T * T).mean() - m2) < Z * (T * T).std() / math.sqrt(n), ((T * T).mean(), m2)
    want = float((mus * length[None]).sum(1).mean())  # K ~ Poisson(int mu)
    assert abs(lookups / n / want - 1.0) < 0.01, (lookups / n, want)


def brute_force_optical_depth(hg, sigma: float, o, d, t0: float, t1: float, m: int = 20_000) -> float:
    """sigma x the integral of the trilinear density over [t0, t1], midpoint rule."""
    ts = t0 + (np.arange(m) + 0.5) * ((t1 - t0) / m)
    pts = np.ascontiguousarray(o[None, :] + ts[:, None] * d[None, :], np.float32)
    ijk, offs, val = np.zeros((m, 3), np.int32), np.zeros((m, 3), np.uint32), np.zeros(m, np.float32)
    oracle().oracle_grid_probe(C.byref(hg.grid), m, ptr(pts), ptr(ijk), ptr(offs), ptr(val))
    return sigma * float(val.astype(np.float64).sum()) * (t1 - t0) / m


def clip_unit_cube(o, d):
    inv = 1.0 / d
    a, b = (0.0 - o) * inv, (1.0 - o) * inv
    return max(0.0, float(np.minimum(a, b).max())), float(np.maximum(a, b).min())


def check_tracking_against_brute_force_optical_depth(lib, f64: bool, tracking: int = _abi.TRACKING_GLOBAL_MAJORANT,
                                                     sparse: bool = False):
    """sparse: a 128^3 cloud with 5 % of its leaves, whose empty regions the local tracker skips."""
    hg = oracle_build_grid(FbmConfig(resolution=128, seed=0x0123, sparse=True, active_leaf_fraction=0.05) if sparse else
                           FbmConfig(resolution=16, seed=0x0123), lib)
    # majorant 8 per unit length; for the sparse cloud 40, with the rays aimed at its densest leaves so
    # that they cross empty space and dense regions alike
    sigma = (40.0 if sparse else 8.0) / max(hg.grid.max_value, 1e-6)
    r = np.random.default_rng(31)
    n = 20_000
    if sparse:
        org = hg.arrays["origins"].reshape(-1, 3)
        dense_leaves = org[np.argsort(hg.arrays["values"].reshape(-1, 512).max(axis=1))[-max(1, len(org) // 4):]]
    for ray in range(6):
        target = (dense_leaves[r.integers(len(dense_leaves))] + 4.0) / 128.0 if sparse else r.uniform(0.3, 0.7, 3)
        v = r.normal(size=3)
        o = target + 2.0 * v / np.linalg.norm(v)
        d = (target - o) / np.linalg.norm(target - o)
        o, d = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
        t0, t1 = clip_unit_cube(o, d)
        tau = brute_force_optical_depth(hg, sigma, o, d, t0, t1)
        tmid = 0.5 * (t0 + t1)
        tau_mid = brute_force_optical_depth(hg, sigma, o, d, t0, tmid)
        O, D = np.tile(o, (n, 1)), np.tile(d, (n, 1))
        T, _ = track(lib, frame(hg.grid, sigma, tracking), True, O, D, f64, seed=100 + ray)
        assert abs(T.mean() - math.exp(-tau)) < Z * T.std() / math.sqrt(n) + 1e-4, (ray, T.mean(), math.exp(-tau))
        t, _ = track(lib, frame(hg.grid, sigma, tracking), False, O, D, f64, seed=200 + ray)
        for frac, want in (((t < 0).mean(), math.exp(-tau)), (((t >= 0) & (t < tmid)).mean(), 1 - math.exp(-tau_mid))):
            se = math.sqrt(max(want * (1 - want), 1e-12) / n)
            assert abs(frac - want) < Z * se + 1e-4, (ray, frac, want, tau)


def check_local_tracking_across_empty_space(lib, f64: bool):
    """Local tracking that skips empty regions: a 3 x 3 x 3 block of leaves of density 1 at the centre of a
    128^3 index box (two single leaves at opposite corners set the box), regions of one 8^3 cell, so that
    rays aimed at the block
    cross long stretches of empty regions (skipped in jumps of up to several regions) before they reach it.
    E[T] = e^(-tau) and P(escape) = e^(-tau) with tau from the brute-force integral."""
    origins = [(0, 0, 0), (120, 120, 120)] + [(x, y, z) for x in (56, 64, 72) for y in (56, 64, 72) for z in (56, 64, 72)]
    values = np.ones((len(origins), 512), np.float32)
    hg = with_majorants(tables_from_leaves(np.array(origins, np.int32), values, 1.0 / 128, (0, 0, 0, 1, 1, 1)), 1)
    codes = hg.arrays["regions"]
    assert (codes > 130).any()  # empty regions with at least two empty layers around them
    sigma = 8.0
    r = np.random.default_rng(41)
    n = 20_000
    for ray in range(6):
        target = (64.0 + r.uniform(-6, 6, 3)) / 128.0
        v = r.normal(size=3)
        o = target + 2.0 * v / np.linalg.norm(v)
        d = (target - o) / np.linalg.norm(target - o)
        o, d = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
        t0, t1 = clip_unit_cube(o, d)
        tau = brute_force_optical_depth(hg, sigma, o, d, t0, t1)
        assert 0.3 < tau < 4.0, tau
        O, D = np.tile(o, (n, 1)), np.tile(d, (n, 1))
        T, _ = track(lib, frame(hg.grid, sigma, LOCAL), True, O, D, f64, seed=300 + ray)
        assert abs(T.mean() - math.exp(-tau)) < Z * T.std() / math.sqrt(n) + 1e-4, (ray, T.mean(), math.exp(-tau))
        t, _ = track(lib, frame(hg.grid, sigma, LOCAL), False, O, D, f64, seed=400 + ray)
        want = math.exp(-tau)
        assert abs((t < 0).mean() - want) < Z * math.sqrt(want * (1 - want) / n) + 1e-4, (ray, (t < 0).mean(), want)


# ------------------------------------------------------------------ phase functions
def cos_cdf(kind: int, k: float, mu: np.ndarray) -> np.ndarray:
    """P(cos theta <= mu) of the normalised phase function (integrated over the azimuth)."""
    if kind == _abi.PHASE_SCHLICK:
        if k == 0.0:
            return (mu + 1.0) / 2.0
        return (1 - k * k) / (2 * k) * (1.0 / (1.0 - k * mu) - 1.0 / (1.0 + k))
    if k == 0.0:
        return (mu + 1.0) / 2.0
    return (1 - k * k) / (2 * k) * ((1 + k * k - 2 * k * mu) ** -0.5 - 1.0 / (1.0 + k))


def density(kind: int, k: float, mu):
    if kind == _abi.PHASE_SCHLICK:
        return (1 - k * k) / (4 * math.pi * (1 - k * mu) ** 2)
    return (1 - k * k) / (4 * math.pi * (1 + k * k - 2 * k * mu) ** 1.5)


def check_phase_sampling(lib, f64: bool, kind: int, k: float):
    k = float(np.float32(k))  # the parameter is a float in the frame parameters
    n = 200_000
    u = np.random.default_rng(41 + kind).random((n, 2))
    dirs, vals = np.zeros((n, 3)), np.zeros(n)
    fn = lib.oracle_phase_sample_batch_f64 if f64 else lib.oracle_phase_sample_batch_f32
    fn(kind, k, n, ptr(u), ptr(dirs), ptr(vals))
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=2e-6)
    # local frame: z is the reversed propagation direction, so cos(theta) = -dir.z
    mu = np.clip(-dirs[:, 2], -1.0, 1.0)
    # 40 bins of equal probability under the closed form
    qs = np.linspace(0.0, 1.0, 41)
    grid = np.linspace(-1.0, 1.0, 200_001)
    edges = np.interp(qs, cos_cdf(kind, k, grid), grid)
    edges[0], edges[-1] = -1.0 - 1e-9, 1.0 + 1e-9
    counts = np.histogram(mu, edges)[0]
    expect = n / 40.0
    chi2 = float(((counts - expect) ** 2 / expect).sum())
    assert chi2 < 90.0, chi2  # 39 degrees of freedom; P(chi2 > 90) ~ 6e-6
    phi = np.arctan2(dirs[:, 1], dirs[:, 0])
    pc = np.histogram(phi, np.linspace(-math.pi, math.pi, 33))[0]
    assert float(((pc - n / 32) ** 2 / (n / 32)).sum()) < 80.0  # uniform azimuth, 31 dof
    tol = 2e-5 if not f64 else 1e-7  # both builds take pi as the float constant
    want = density(kind, k, mu)
    assert np.allclose(vals, want, rtol=tol, atol=0), np.abs(vals / want - 1).max()
    value_fn = lib.oracle_phase_value_f64 if f64 else lib.oracle_phase_value_f32
    for m in (-1.0, -0.3, 0.0, 0.5, 0.99):
        assert abs(value_fn(kind, k, m) / density(kind, k, m) - 1.0) < tol


# ------------------------------------------------------------------ NRC data path
INVALID = _abi.INVALID_VERTEX_INDEX


class NrcHost:
    def __init__(self, w: int, h: int, samples: int = 1):
        n = w * h
        q_stride = (n + 127) // 128 * 128
        self.w, self.h, self.n, self.q_stride = w, h, n, q_stride
        self.num_training_data = np.zeros(2, np.uint32)
        self.train_image_size = np.array([w, h, w, h], np.int32)
        self.suffix = np.zeros((n, 4), np.uint32)
        self.suffix[:, 0] = INVALID
        self.inferred = np.zeros((samples * q_stride, 3), np.float32)
        self.inf_queries = np.zeros((samples * q_stride, 14), np.float32)
        self.terminals = np.zeros((samples * n, 4), np.uint32)
        self.contrib = np.zeros((samples * n, 3), np.float32)
        self.vinfo = np.zeros((_abi.TRAIN_BUFFER_SIZE, 6), np.uint32)
        self.targets = np.zeros((_abi.TRAIN_BUFFER_SIZE, 3), np.float32)
        self.queries = np.zeros((_abi.TRAIN_BUFFER_SIZE, 14), np.float32)
        self.accum = np.zeros(3 * n, np.float64)
        self.samples = samples

    def params(self, radiance_scale: float) -> _abi.FrameParams:
        p = _abi.FrameParams()
        p.image_width, p.image_height = self.w, self.h
        p.accum = ptr(self.accum)
        nb = p.nrc
        nb.nrc_buffer_index = 0
        nb.num_samples = self.samples
        nb.radiance_scale = radiance_scale
        nb.max_num_training_suffixes = self.n
        nb.num_training_data = ptr(self.num_training_data)
        nb.train_image_size = ptr(self.train_image_size)
        nb.train_suffix_terminals = ptr(self.suffix)
        nb.inferred_radiance = ptr(self.inferred)
        nb.inference_queries = ptr(self.inf_queries)
        nb.inference_terminals = ptr(self.terminals)
        nb.per_frame_contribution = ptr(self.contrib)
        nb.train_vertex_info = ptr(self.vinfo)
        nb.train_targets[0] = ptr(self.targets)
        nb.train_queries[0] = ptr(self.queries)
        return p


def check_propagate_is_the_per_path_sum(lib, radiance_scale: float):
    r = np.random.default_rng(51)
    h = NrcHost(8, 6)
    npaths = 40
    slots = r.permutation(_abi.TRAIN_BUFFER_SIZE)[: npaths * 8]  # vertices scattered over the buffer
    used = 0
    paths = []
    for i in range(npaths):
        m = int(r.integers(1, 9))
        idx = [int(v) for v in slots[used:used + m]]
        used += m
        T = r.uniform(0.0, 1.2, (m, 3)).astype(np.float32)
        E = r.uniform(0.0, 2.0, (m, 3)).astype(np.float32)  # the NEE-initialised targets
        rho = r.uniform(0.2, 1.0, (m, 3)).astype(np.float32)
        rho[r.random((m, 3)) < 0.1] = 0.0
        for j, v in enumerate(idx):
            h.vinfo[v, :3] = T[j].view(np.uint32)
            h.vinfo[v, 3] = idx[j - 1] if j > 0 else INVALID
            h.targets[v] = E[j]
            h.queries[v, 8:11] = rho[j] * 0.75
            h.queries[v, 11:14] = rho[j] - rho[j] * np.float32(0.75)
        suffix = i  # suffix terminal i ends path i
        has_query = int(r.random() < 0.7)
        h.suffix[suffix] = (idx[-1], has_query, m, 0)
        nrc = r.uniform(-0.5, 3.0, 3).astype(np.float32)
        rho_t = r.uniform(0.1, 1.0, 3).astype(np.float32)
        h.inferred[suffix] = nrc
        h.inf_queries[suffix, 8:11] = rho_t
        paths.append((idx, T.astype(np.float64), E.astype(np.float64),
                      (h.queries[idx, 8:11] + h.queries[idx, 11:14]).astype(np.float64), has_query,
                      nrc.astype(np.float64), rho_t.astype(np.float64)))
    p = h.params(radiance_scale)
    lib.oracle_nrc_propagate(C.byref(p))
    for idx, T, E, rho, has_query, nrc, rho_t in paths:
        m = len(idx)
        term = np.zeros(3)
        if has_query:
            term = np.maximum(nrc, 0.0) * (1.0 / radiance_scale if radiance_scale > 0 else 1.0) * rho_t
        for j in range(m):
            L = np.zeros(3)
            prod = np.ones(3)
            for i in range(j, m):
                L += prod * E[i]
                prod = prod * T[i]
            L += prod * term
            want = np.where(rho[j] != 0.0, L / np.where(rho[j] != 0.0, rho[j], 1.0), 0.0)
            got = h.targets[idx[j]].astype(np.float64)
            assert np.allclose(got, want, rtol=2e-5, atol=1e-6), (j, m, got, want)


def check_accumulate_is_direct_plus_cached_radiance(lib, radiance_scale: float):
    r = np.random.default_rng(61)
    w, hh, S = 9, 7, 3
    h = NrcHost(w, hh, S)
    n = w * hh
    h.contrib[:] = r.uniform(0.0, 1.0, h.contrib.shape)
    h.terminals[:, :3] = r.uniform(0.0, 1.5, (S * n, 3)).astype(np.float32).view(np.uint32)
    h.terminals[:, 3] = (r.random(S * n) < 0.6).astype(np.uint32) | (r.integers(0, 9, S * n).astype(np.uint32) << 8)
    h.inferred[:] = r.uniform(-0.5, 3.0, h.inferred.shape)
    h.inf_queries[:, 8:11] = r.uniform(0.1, 1.0, (h.inf_queries.shape[0], 3))
    h.inf_queries[:, 11:14] = r.uniform(0.0, 0.2, (h.inf_queries.shape[0], 3))
    h.accum[:] = r.uniform(0.0, 5.0, h.accum.shape)
    start = h.accum.copy()
    lib.oracle_nrc_accumulate(C.byref(h.params(radiance_scale)))
    scale = 1.0 / radiance_scale if radiance_scale > 0 else 1.0
    want = start.reshape(n, 3).copy()
    for s in range(S):
        for px in range(n):
            i, qi = s * n + px, s * h.q_stride + px
            T = h.terminals[i, :3].view(np.float32).astype(np.float64)
            cached = np.zeros(3)
            if h.terminals[i, 3] & 1:
                rho = h.inf_queries[qi, 8:11].astype(np.float64) + h.inf_queries[qi, 11:14].astype(np.float64)
                cached = np.maximum(h.inferred[qi].astype(np.float64), 0.0) * scale * rho
            want[px] += h.contrib[i].astype(np.float64) + T * cached
    assert np.allclose(h.accum.reshape(n, 3), want, rtol=1e-6, atol=1e-6)


# ------------------------------------------------------------------ importance map
LUM = np.array([0.2126729, 0.7151522, 0.0721750])  # Rec. 709 / sRGB luminance (CIE Y row of the XYZ matrix)


def check_importance_map_definition(lib, w: int, h: int):
    r = np.random.default_rng(71 + w)
    tex = np.zeros((h, w, 4), np.float32)
    tex[..., :3] = r.uniform(0.0, 2.0, (h, w, 3)) * (r.random((h, w, 1)) < 0.9)
    dx, dy = C.c_uint32(), C.c_uint32()
    levels = lib.oracle_env_build_mips(ptr(tex), w, h, None, C.byref(dx), C.byref(dy), None)
    dx, dy = dx.value, dy.value
    assert dx == 1 << max(0, (w - 1).bit_length()) and dy == max(1 << max(0, (h - 1).bit_length()), dx // 2)
    assert levels == dx.bit_length()
    total = sum(max(dx >> l, 1) * max(dy >> l, 1) for l in range(levels))
    mips, integral = np.zeros(total, np.float32), np.zeros(1, np.float32)
    lib.oracle_env_build_mips(ptr(tex), w, h, ptr(mips), None, None, ptr(integral))
    y = np.arange(h, dtype=np.float64)
    mip0 = np.zeros((dy, dx))
    mip0[:h, :w] = (tex[..., :3].astype(np.float64) @ LUM) * (math.pi / 2) * np.sin(math.pi * (y + 0.5) / h)[:, None]
    scale = mip0.max()
    off = 0
    for l in range(levels):
        lx, ly = max(dx >> l, 1), max(dy >> l, 1)
        got = mips[off:off + lx * ly].astype(np.float64).reshape(ly, lx)
        off += lx * ly
        # the mean of the level's 2^l x 2^l block of mip 0 (rows past the map count as zero)
        b = 1 << l
        pad = np.zeros((ly * b, lx * b))
        pad[:min(dy, ly * b), :] = mip0[:min(dy, ly * b), :]
        want = pad.reshape(ly, b, lx, b).sum(axis=(1, 3)) / (b * b)
        assert np.allclose(got, want, rtol=0, atol=3e-6 * scale), (l, np.abs(got - want).max())
    assert abs(integral[0] / mip0.sum() - 1.0) < 1e-5


def check_cell_majorants(hg):
    """Cell majorants of rtc8x.h's definition, from a dense copy of the voxels: cell c's bound is the largest
    voxel (at least 0; background 0) over index [8c - 1, 8c + 9] per axis."""
    g = hg.grid
    org = hg.arrays["origins"].reshape(-1, 3)
    vals = hg.arrays["values"].reshape(-1, 8, 8, 8)
    co, cd = np.array(g.cell_origin[:]), np.array(g.cell_dims[:])
    lo = co - 8  # a block of background on every side
    dense = np.zeros(tuple(8 * cd + 16), np.float32)
    for n, (x, y, z) in enumerate(org - lo):
        dense[x:x + 8, y:y + 8, z:z + 8] = vals[n]
    cells = hg.arrays["cells"].reshape(cd)
    for x in range(cd[0]):
        for y in range(cd[1]):
            for z in range(cd[2]):
                b = dense[8 * x + 7:8 * x + 18, 8 * y + 7:8 * y + 18, 8 * z + 7:8 * z + 18]
                assert cells[x, y, z] == max(np.float32(0), b.max()), (x, y, z, cells[x, y, z], b.max())


def check_region_majorants(lib, res: int, sparse: bool, region_cells: int = 0):
    """The region table of rtc8x.h's definition: regions of region_cells^3 cells (0: 4, doubled until there
    are at most 32768 of them); for a region with density the least q in 1..127 with q x unit >= its largest
    cell majorant; for an empty region 128 + (its chessboard distance to the nearest region with density) - 1,
    the distance from scipy's chamfer transform.  The cell majorants follow their own definition."""
    from scipy.ndimage import distance_transform_cdt

    hg = oracle_build_grid(FbmConfig(resolution=res, seed=0x0456, sparse=sparse), lib)
    if region_cells:
        hg.grid.cell_majorants = hg.grid.region_majorants = None  # numpy's memory: the oracle must not free it
        lib.oracle_compute_cell_majorants(C.byref(hg.grid), region_cells)
        g = hg.grid
        hg.arrays["cells"] = np.ctypeslib.as_array((C.c_float * int(np.prod(g.cell_dims[:]))).from_address(g.cell_majorants)).copy()
        hg.arrays["regions"] = np.ctypeslib.as_array((C.c_uint8 * int(np.prod(g.region_dims[:]))).from_address(g.region_majorants)).copy()
        g.cell_majorants = hg.arrays["cells"].ctypes.data
        g.region_majorants = hg.arrays["regions"].ctypes.data
    check_cell_majorants(hg)
    g = hg.grid
    cd = np.array(g.cell_dims[:])
    cells = hg.arrays["cells"].reshape(cd)
    rc = region_cells or 4
    while not region_cells and np.prod(-(-cd // rc)) > _abi.MAX_REGIONS:
        rc *= 2
    assert g.region_cells == rc and list(g.region_dims) == list(-(-cd // rc))
    rd = -(-cd // rc)
    pad = np.zeros(rd * rc, np.float32)
    pad[:cd[0], :cd[1], :cd[2]] = cells
    rmax = pad.reshape(rd[0], rc, rd[1], rc, rd[2], rc).max(axis=(1, 3, 5))
    unit = np.float32(g.region_unit)
    assert np.float32(127) * unit >= np.float32(cells.max()) and unit <= np.float32(cells.max()) / 127 * (1 + 1e-6)
    code = hg.arrays["regions"].reshape(rd).astype(np.int64)
    dense = rmax > 0
    q = np.where(dense, code, 0)
    assert ((q >= 1) & (q <= 127))[dense].all()
    assert (q.astype(np.float32) * unit >= rmax)[dense].all()
    assert ((q - 1).astype(np.float32) * unit < rmax)[dense].all()  # q is the least such value
    dist = distance_transform_cdt(~dense, metric="chessboard")
    want = 128 + np.minimum(dist - 1, 127)
    assert np.array_equal(code[~dense], want[~dense])
    assert dense.any() and (not sparse or (~dense).any())


# ------------------------------------------------------------------ the tests
LOCAL = _abi.TRACKING_LOCAL_MAJORANT
TRACKINGS = [_abi.TRACKING_GLOBAL_MAJORANT, LOCAL]


@pytest.mark.parametrize("tracking", TRACKINGS)
@pytest.mark.parametrize("f64", [False, True])
def test_free_flight_in_a_homogeneous_medium(f64, tracking):
    check_free_flight_homogeneous(oracle(), f64, tracking)


@pytest.mark.parametrize("tracking", TRACKINGS)
@pytest.mark.parametrize("f64", [False, True])
def test_ratio_tracking_in_a_homogeneous_medium(f64, tracking):
    check_ratio_tracking_homogeneous(oracle(), f64, tracking)


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("tracking", TRACKINGS)
@pytest.mark.parametrize("f64", [False, True])
def test_tracking_against_the_brute_force_optical_depth(f64, tracking, sparse):
    check_tracking_against_brute_force_optical_depth(oracle(), f64, tracking, sparse)


@pytest.mark.parametrize("f64", [False, True])
def test_local_tracking_across_empty_space(f64):
    check_local_tracking_across_empty_space(oracle(), f64)


@pytest.mark.parametrize("res,sparse,region_cells", [(64, False, 0), (128, True, 0), (128, True, 1), (256, True, 2)])
def test_region_majorants_match_their_definition(res, sparse, region_cells):
    check_region_majorants(oracle(), res, sparse, region_cells)


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("kind,k", [(_abi.PHASE_HENYEY_GREENSTEIN, 0.85), (_abi.PHASE_HENYEY_GREENSTEIN, -0.4),
                                    (_abi.PHASE_HENYEY_GREENSTEIN, 0.0), (_abi.PHASE_SCHLICK, 0.7),
                                    (_abi.PHASE_SCHLICK, -0.5)])
def test_phase_sampling_follows_the_closed_form(f64, kind, k):
    check_phase_sampling(oracle(), f64, kind, k)


@pytest.mark.parametrize("radiance_scale", [0.0, 10.0])
def test_propagate_is_the_per_path_sum(radiance_scale):
    check_propagate_is_the_per_path_sum(oracle(), radiance_scale)


@pytest.mark.parametrize("radiance_scale", [0.0, 10.0])
def test_accumulate_is_direct_plus_cached_radiance(radiance_scale):
    check_accumulate_is_direct_plus_cached_radiance(oracle(), radiance_scale)


@pytest.mark.parametrize("w,h", [(40, 20), (33, 17), (16, 16), (1, 1), (64, 20)])
def test_importance_map_matches_its_definition(w, h):
    check_importance_map_definition(oracle(), w, h)


# (file, text, mistake, check) -- each mistake is one a restatement could plausibly make
MUTATIONS = [
    ("rtc8_oracle_tracer.inc", "if (UNI(rng) < sigma / m->majorant) { *thit = t; return 1; }",
     "if (UNI(rng) > sigma / m->majorant) { *thit = t; return 1; }",
     lambda lib: check_free_flight_homogeneous(lib, False)),
    ("rtc8_oracle_tracer.inc", "m->majorant = (R)p->density_coeff * (R)p->grid.max_value;",
     "m->majorant = (R)p->density_coeff;",
     lambda lib: check_free_flight_homogeneous(lib, False)),
    ("rtc8_oracle_tracer.inc", "vis *= (R)1 - sigma / m->majorant;", "vis *= (R)1 - sigma / (m->majorant + m->coeff);",
     lambda lib: check_ratio_tracking_homogeneous(lib, False)),
    ("rtc8_oracle_tracer.inc", "R vy = v000 + wz * (v001 - v000);", "R vy = v001 + wz * (v000 - v001);",
     lambda lib: check_tracking_against_brute_force_optical_depth(lib, False)),
    ("rtc8_oracle_tracer.inc", "R need = mu * (tend - t);", "R need = mu * (tend - t) * (R)0.5;",
     lambda lib: check_free_flight_homogeneous(lib, False, LOCAL)),
    ("rtc8_oracle_tracer.inc", "s->tnext[a] = s->step[a] != 0 ? SUF(dda_boundary)(s, a) : (R)INFINITY;",
     "s->tnext[a] = s->step[a] != 0 ? SUF(dda_boundary)(s, a) + (R)1 : (R)INFINITY;",
     lambda lib: check_local_tracking_across_empty_space(lib, False)),
    # the cube's full width taken for its half-width (one layer too many would jump into the block itself)
    ("rtc8_oracle_tracer.inc", "int32_t c = s->c[a] + s->step[a] * k;", "int32_t c = s->c[a] + s->step[a] * (2 * k + 1);",
     lambda lib: check_local_tracking_across_empty_space(lib, False)),
    ("rtc8_oracle_core.c", "size_t r = ((size_t)(x / (int32_t)rc) * rd[1] + y / (int32_t)rc) * rd[2] + z / (int32_t)rc;",
     "size_t r = ((size_t)(x / (int32_t)rc) * rd[1] + z / (int32_t)rc) * rd[2] + y / (int32_t)rc;",
     lambda lib: check_region_majorants(lib, 128, True)),
    # the rim forgotten on the far side: the taps at 8c + 8 and the rounding margin beyond
    ("rtc8_oracle_core.c", "static const int32_t first[3] = {7, 0, 0}, last[3] = {7, 7, 1};",
     "static const int32_t first[3] = {7, 0, 0}, last[3] = {7, 7, 0};",
     lambda lib: check_region_majorants(lib, 128, True, 1)),
    ("rtc8_oracle_tracer.inc", "*dir_local = SUF(v3)(cp * sin_t, sp * sin_t, -cos_t);",
     "*dir_local = SUF(v3)(cp * sin_t, sp * sin_t, cos_t);",
     lambda lib: check_phase_sampling(lib, False, _abi.PHASE_HENYEY_GREENSTEIN, 0.85)),
    ("rtc8_oracle_tracer.inc", "cos_t = ((R)2 * u1 + k - (R)1) / ((R)2 * k * u1 - k + (R)1);",
     "cos_t = ((R)2 * u1 + k - (R)1) / ((R)2 * k * u1 + k - (R)1);",
     lambda lib: check_phase_sampling(lib, False, _abi.PHASE_SCHLICK, 0.7)),
    ("rtc8_oracle_tracer.inc", "R dterm = (R)1 - ph->k * cos_t;", "R dterm = (R)1 + ph->k * cos_t;",
     lambda lib: check_phase_sampling(lib, False, _abi.PHASE_SCHLICK, 0.7)),
    ("rtc8_oracle_nrc.c", "float indirect = v->local_throughput[k] * c[k];", "float indirect = c[k];",
     lambda lib: check_propagate_is_the_per_path_sum(lib, 10.0)),
    ("rtc8_oracle_nrc.c", "tgt[k] = ref != 0.0f ? c[k] / ref : 0.0f;", "tgt[k] = c[k];",
     lambda lib: check_propagate_is_the_per_path_sum(lib, 10.0)),
    ("rtc8_oracle_nrc.c", "for (int k = 0; k < 3; ++k) rad[k] = rad[k] * (q->diffuse_reflectance[k] + q->specular_reflectance[k]);",
     "for (int k = 0; k < 3; ++k) rad[k] = rad[k] * q->diffuse_reflectance[k];",
     lambda lib: check_accumulate_is_direct_plus_cached_radiance(lib, 10.0)),
    ("rtc8_oracle_tracer.c", "rtc8x_sinf(RTC8X_PI_F * ((float)y + 0.5f) / (float)h)", "rtc8x_sinf(RTC8X_PI_F * (float)y / (float)h)",
     lambda lib: check_importance_map_definition(lib, 40, 20)),
    ("rtc8_oracle_tracer.c", "src[(size_t)(uy + 1) * sdx + ux + 1] : 0.0f;", "src[(size_t)(uy + 1) * sdx + ux] : 0.0f;",
     lambda lib: check_importance_map_definition(lib, 40, 20)),
]


def build_mutant(tmp: Path, i: int) -> Path:
    fname, text, mistake, _ = MUTATIONS[i]
    src = tmp / f"m{i}" / "oracle"
    shutil.copytree(ROOT / "oracle", src, ignore=shutil.ignore_patterns("*.so", "_ref", "ref"))
    shutil.copytree(ROOT / "include", src.parent / "include")
    f = src / fname
    body = f.read_text()
    assert body.count(text) >= 1, (fname, text)
    f.write_text(body.replace(text, mistake, 1))  # the first occurrence (free flight before ratio tracking)
    subprocess.run(["make", "-s", "-C", str(src)], check=True, capture_output=True)
    return src / "liboracle.so"


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_planted_mistakes_fail_their_pins(tmp_path):
    """Each pin above fails on an oracle built with one planted mistake; the unmodified oracle passes them."""
    with ThreadPoolExecutor(max_workers=4) as ex:
        libs = list(ex.map(lambda i: build_mutant(tmp_path, i), range(len(MUTATIONS))))
    for i, so in enumerate(libs):
        check = MUTATIONS[i][3]
        check(oracle())  # the real oracle passes
        with pytest.raises(AssertionError):
            check(load_oracle(so))
