"""CPU checks that pin the oracle's building blocks with independent evidence:
the C++ standard's mt19937_64 check value, LCG jump-ahead closed forms, the PCG32 state recurrence,
deterministic math against libm, and phase-function normalisation."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle_lib import oracle, ptr


def test_mt19937_64_standard_check_value():
    # [rand.predef]: the 10000th output of a default-constructed mt19937_64 (seed 5489) is 9981545732273789042
    out = np.empty(10000, dtype=np.uint64)
    oracle().oracle_mt19937_64(C.c_uint64(5489), 10000, ptr(out))
    assert int(out[-1]) == 9981545732273789042


def _affine_pow(a, c, n, mod):
    """(A, C) such that n steps of s -> a s + c equal s -> A s + C (Brown's jump-ahead)."""
    acc_a, acc_c = 1, 0
    cur_a, cur_c = a, c
    while n:
        if n & 1:
            acc_a, acc_c = (acc_a * cur_a) % mod, (acc_c * cur_a + cur_c) % mod
        cur_a, cur_c = (cur_a * cur_a) % mod, (cur_c * cur_a + cur_c) % mod
        n >>= 1
    return acc_a, acc_c


def test_pcg32_state_recurrence_and_output_range():
    seed = 591842031321323413
    n = 12345
    outs = np.empty(n, dtype=np.uint32)
    end = np.zeros(1, dtype=np.uint64)
    oracle().oracle_pcg_fill(C.c_uint64(seed), n, ptr(outs), None, ptr(end))
    a, c = _affine_pow(6364136223846793005, 1, n, 2 ** 64)
    assert int(end[0]) == (a * seed + c) % 2 ** 64
    # uniform floats: 23-bit grid in [0, 1), mean 1/2
    f = np.empty(200000, dtype=np.float32)
    oracle().oracle_pcg_fill(C.c_uint64(7), f.size, None, ptr(f), None)
    assert f.min() >= 0.0 and f.max() < 1.0
    assert np.all((f * np.float32(2 ** 23)) == np.floor(f * np.float32(2 ** 23)))
    assert abs(f.mean() - 0.5) < 0.005
    # XSH-RR of the first state, computed from its definition
    s = seed
    x = ((s >> 18) ^ s) >> 27 & 0xFFFFFFFF
    rot = s >> 59
    assert int(outs[0]) == ((x >> rot) | (x << ((-rot) & 31))) & 0xFFFFFFFF


def test_lcg_matches_jump_ahead():
    seeds = np.empty(65536, dtype=np.uint32)
    oracle().oracle_lcg_seeds(471313181, 65536, ptr(seeds))
    # with the 32-bit wrap before the mod 2^31, the sequence lives in Z/2^31 (2^31 divides 2^32)
    for k in (1, 2, 1000, 65536):
        a, c = _affine_pow(1103515245, 12345, k, 2 ** 31)
        assert int(seeds[k - 1]) == (a * 471313181 + c) % 2 ** 31


def test_deterministic_math_against_libm():
    orc = oracle()
    rng = np.random.default_rng(3)
    xs = np.concatenate([1.0 - rng.integers(0, 2 ** 23, 20000) * 2.0 ** -23, [1e-40, 1e-30, 1.0, 3e38]])
    for x in xs.astype(np.float32):
        got = orc.oracle_logf(C.c_float(x))
        ref = math.log(float(x))
        assert abs(got - ref) <= 2.5 * abs(np.spacing(np.float32(ref))) + 1e-45, x
    s, c = np.zeros(1, np.float32), np.zeros(1, np.float32)
    for x in (rng.random(20000) * 2 * math.pi).astype(np.float32):
        orc.oracle_sincosf(C.c_float(x), ptr(s), ptr(c))
        assert abs(s[0] - math.sin(float(x))) < 2e-7 and abs(c[0] - math.cos(float(x))) < 2e-7
    for y, x in rng.uniform(-1, 1, size=(20000, 2)).astype(np.float32):
        assert abs(orc.oracle_atan2f(C.c_float(y), C.c_float(x)) - math.atan2(float(y), float(x))) < 5e-7
        assert abs(orc.oracle_acosf(C.c_float(y)) - math.acos(float(y))) < 5e-7


@pytest.mark.parametrize("kind,k", [(0, 0.0), (0, 0.7), (0, -0.4), (1, 0.0), (1, 0.85), (1, -0.5)])
def test_phase_functions_normalise_and_mean_cosine(kind, k):
    """Integral of p over the sphere is 1; HG's mean cosine is g (closed form); Schlick's is
    (2k - (1-k^2) ln((1+k)/(1-k))) / (2k^2) ... evaluated numerically from the same density."""
    mu = np.linspace(-1, 1, 200001)
    if kind == 0:
        p = (1 - k * k) / (4 * math.pi * (1 - k * mu) ** 2)
    else:
        p = (1 - k * k) / (4 * math.pi * (1 + k * k - 2 * k * mu) ** 1.5)
    integral = np.trapezoid(p, mu) * 2 * math.pi
    assert abs(integral - 1.0) < 1e-6
    if kind == 1:
        assert abs(np.trapezoid(p * mu, mu) * 2 * math.pi - k) < 1e-6
