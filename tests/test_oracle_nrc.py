"""CPU checks of the NRC oracle: the network gradient against central finite differences, the FP16 /
BF16 rounding emulation against numpy, the Adam step against its closed form, the shuffle's LCG
permutation, and the 32-bit Mersenne Twister used for the per-frame offsets."""
import ctypes as C

import numpy as np
import pytest

from oracle_lib import oracle, ptr
from rtcamp8_amd import _abi


def random_queries(n, seed=1):
    rng = np.random.default_rng(seed)
    q = np.zeros((n, 14), dtype=np.float32)
    q[:, 0:3] = rng.random((n, 3))
    q[:, 3] = rng.random(n) * np.pi
    q[:, 4] = rng.random(n) * 2 * np.pi
    q[:, 5] = q[:, 3]
    q[:, 6] = q[:, 4]
    q[:, 7] = 0.0
    q[:, 8:11] = 0.99
    return q


def random_weights(L, seed=2, scale=0.3, encoding=0, table_scale=0.1):
    """MLP weights (output rows beyond RGB zero) followed, with the hash grid, by random tables."""
    P = oracle().oracle_nrc_num_params(L)
    P_all = oracle().oracle_nrc_num_params_ex(L, encoding)
    rng = np.random.default_rng(seed)
    w = np.zeros(P_all, dtype=np.float32)
    w[:P] = (rng.standard_normal(P) * scale / 8).astype(np.float32)
    w[L * 4096 + 3 * 64:P] = 0.0  # padded output rows
    w[P:] = (rng.uniform(-1, 1, P_all - P) * table_scale).astype(np.float32)
    return w


def test_gradient_matches_central_differences():
    L, n = 2, 24
    q = random_queries(n)
    t = np.random.default_rng(3).random((n, 3)).astype(np.float32)
    w = random_weights(L)
    P = w.size
    grad = np.zeros(P, dtype=np.float64)
    oracle().oracle_nrc_gradient(ptr(w), L, ptr(q), ptr(t), n, ptr(grad), 0)
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(4096, 20, replace=False), 4096 + rng.choice(4096, 20, replace=False),
                          2 * 4096 + rng.choice(192, 20, replace=False)])
    # the normaliser lum(p)^2 + 0.01 is held constant by the loss definition, so differentiate the loss
    # with the normaliser frozen at the current prediction
    base = np.zeros(3 * n)
    oracle().oracle_nrc_forward(ptr(w), L, ptr(q), n, ptr(base), 0)
    p0 = base.reshape(n, 3)
    denom = (p0 @ np.array([0.2126729, 0.7151522, 0.0721750])) ** 2 + 0.01

    def frozen_loss(wv):
        out = np.zeros(3 * n)
        oracle().oracle_nrc_forward(ptr(wv), L, ptr(q), n, ptr(out), 0)
        return float((((out.reshape(n, 3) - t) ** 2).sum(axis=1) / denom).sum())

    for i in idx:
        h = 1e-4  # small enough not to cross ReLU kinks of the 24 samples
        wp, wm = w.astype(np.float64).copy(), w.astype(np.float64).copy()
        wp[i] += h
        wm[i] -= h
        fd = (frozen_loss(wp.astype(np.float32)) - frozen_loss(wm.astype(np.float32))) / (
            float(np.float32(wp[i])) - float(np.float32(wm[i])))
        assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(grad[i])), (i, fd, grad[i])


def test_fp16_and_bf16_rounding_emulation():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-8, 5, 20000),
                        [65504.0, 65519.0, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 0.0, -0.0]]).astype(np.float32)
    got = np.array([oracle().oracle_round_fp16(C.c_float(v)) for v in x], dtype=np.float32)
    assert np.array_equal(got, x.astype(np.float16).astype(np.float32))
    got_b = np.array([oracle().oracle_round_bf16(C.c_float(v)) for v in x[:2000]], dtype=np.float32)
    # bf16 keeps the top 16 bits: the result is representable and within half an ulp (2^-8 relative)
    assert np.all((got_b.view(np.uint32) & 0xFFFF) == 0)
    rel = np.abs(got_b - x[:2000]) / np.maximum(np.abs(x[:2000]), 1e-30)
    assert np.all(rel <= 2.0 ** -8 + 1e-12)


def test_adam_first_step_closed_form():
    """With zero moments, step 1 moves each weight by lr * g / (|g| + eps') (bias-corrected Adam)."""
    P = 1000
    rng = np.random.default_rng(6)
    master = rng.standard_normal(P).astype(np.float32)
    m, v, ema = np.zeros(P, np.float32), np.zeros(P, np.float32), np.zeros(P, np.float32)
    g = rng.standard_normal(P).astype(np.float32)
    w0 = master.copy()
    oracle().oracle_nrc_adam(ptr(master), ptr(m), ptr(v), ptr(ema), ptr(g), P, 1e-2, 0.9, 0.99, 1e-8, 0.0, 0.99, 1)
    gt = g.astype(np.float64)
    expect = w0 - 1e-2 * gt / (np.abs(gt) + 1e-8)
    assert np.allclose(master, expect, rtol=0, atol=2e-6)
    assert np.allclose(ema, 0.01 * master, rtol=1e-6)


def test_shuffle_lcg_is_a_permutation():
    """renderer_main.cpp:220-229 seeds 65536 shufflers with consecutive LCG states; the low 16 bits of a
    full-period LCG mod 2^31 (c odd, a = 1 mod 4) have period 2^16, so next() % 65536 is a bijection in
    every frame."""
    seeds = np.empty(65536, dtype=np.uint32)
    oracle().oracle_lcg_seeds(471313181, 65536, ptr(seeds))
    s = seeds.astype(np.uint64)
    for frame in range(3):
        s = ((s * 1103515245 + 12345) & 0xFFFFFFFF) % (1 << 31)
        dst = s % 65536
        assert np.unique(dst).size == 65536


def test_mt19937_standard_check_value():
    from rtcamp8_amd.nrc import Mt19937

    g = Mt19937(5489)
    for _ in range(9999):
        g()
    assert g() == 4123659995  # [rand.predef]: 10000th output of default-constructed mt19937


def test_encoding_ranges():
    q = random_queries(50)
    e = np.zeros(50 * 64)
    oracle().oracle_nrc_encode(ptr(q), 50, ptr(e))
    e = e.reshape(50, 64)
    assert np.all(e[:, :56] >= 0) and np.all(e[:, :56] <= 1)
    assert np.all(e[:, 62] == 1) and np.all(e[:, 63] == 0)
    assert np.allclose(e[:, 56:59], 0.99, atol=1e-6)


def test_hash_grid_geometry():
    """Level 0 (scale 15) is dense with (15 + 2)^3 corners rounded up to 8; the 15 finer levels are hashed."""
    L = 2
    total = oracle().oracle_nrc_num_params_ex(L, 1) - oracle().oracle_nrc_num_params(L)
    assert total == 2 * (4920 + 15 * 32768)


def test_hash_encoding_is_continuous_and_interpolates():
    """Trilinear interpolation is continuous across cell faces and exact at the dense level's corners."""
    L = 2
    w = random_weights(L, seed=7, encoding=1)
    P = oracle().oracle_nrc_num_params(L)
    q = random_queries(2, seed=8)
    q[1] = q[0]
    q[0, 0] = 5.0 / 15.0 - 0.5 / 15.0 - 1e-7  # just below a level-0 cell face (p * 15 + 0.5 = 5)
    q[1, 0] = 5.0 / 15.0 - 0.5 / 15.0 + 1e-7
    e = np.zeros(2 * 64)
    oracle().oracle_nrc_encode_ex(ptr(w), L, 1, ptr(q), 2, ptr(e))
    e = e.reshape(2, 64)
    # across the level-0 face the coarse features match (finer levels move by their scale x 2e-7)
    assert np.max(np.abs(e[0, :8] - e[1, :8])) < 1e-5
    # at a level-0 corner the features are the table entry there
    q2 = random_queries(1, seed=9)
    q2[0, 0:3] = [(3 - 0.5) / 15.0, (4 - 0.5) / 15.0, (6 - 0.5) / 15.0]  # corner (3, 4, 6)
    e2 = np.zeros(64)
    oracle().oracle_nrc_encode_ex(ptr(w), L, 1, ptr(q2), 1, ptr(e2))
    idx = 3 + 4 * 17 + 6 * 17 * 17
    assert np.allclose(e2[0:2], w[P + 2 * idx:P + 2 * idx + 2], rtol=0, atol=1e-6)
    assert np.all(e2[32:52] >= 0) and e2[58] == 1 and np.all(e2[59:] == 0)


def test_hash_gradient_matches_central_differences():
    L, n = 2, 16
    q = random_queries(n, seed=10)
    t = np.random.default_rng(11).random((n, 3)).astype(np.float32)
    w = random_weights(L, seed=12, encoding=1)
    P = oracle().oracle_nrc_num_params(L)
    grad = np.zeros(w.size)
    oracle().oracle_nrc_gradient_ex(ptr(w), L, 1, ptr(q), ptr(t), n, ptr(grad), 0)
    base = np.zeros(3 * n)
    oracle().oracle_nrc_forward_ex(ptr(w), L, 1, ptr(q), n, ptr(base), 0)
    denom = (base.reshape(n, 3) @ np.array([0.2126729, 0.7151522, 0.0721750])) ** 2 + 0.01

    def frozen_loss(wv):
        out = np.zeros(3 * n)
        oracle().oracle_nrc_forward_ex(ptr(wv), L, 1, ptr(q), n, ptr(out), 0)
        return float((((out.reshape(n, 3) - t) ** 2).sum(axis=1) / denom).sum())

    touched = np.flatnonzero(grad[P:]) + P
    assert touched.size > 100
    rng = np.random.default_rng(13)
    idx = np.concatenate([rng.choice(touched, 25, replace=False), rng.choice(4096, 5, replace=False)])
    for i in idx:
        h = 1e-4
        wp, wm = w.copy(), w.copy()
        wp[i] += np.float32(h)
        wm[i] -= np.float32(h)
        fd = (frozen_loss(wp) - frozen_loss(wm)) / (float(wp[i]) - float(wm[i]))
        assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(grad[i])), (i, fd, grad[i])
