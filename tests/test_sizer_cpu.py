"""The learned-sizing oracle pinned on the CPU: MLP gradients against finite
differences, the HLL estimator against known cardinalities, and the
feature layout on hand-built sketches."""

import math
import os

import numpy as np
import pytest

from oracle import sizer_ref as sr

DIN, H, DOUT = 72, 64, 2


def _params(seed, din=DIN, h=H, dout=DOUT):
    rng = np.random.default_rng(seed)
    n = din * h + h + h * h + h + dout * h + dout
    return rng.normal(0, 0.3, n)


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    din, h, dout = 9, 7, 2
    p = _params(4, din, h, dout)
    X = rng.normal(0, 1, (5, din))
    T = rng.normal(0, 1, (5, dout))
    g = sr.gradient(p, X, T, din, h, dout)
    eps = 1e-6
    fd = np.zeros_like(p)
    for i in range(p.size):
        q = p.copy()
        q[i] += eps
        lp = sr.loss(q, X, T, din, h, dout)
        q[i] -= 2 * eps
        lm = sr.loss(q, X, T, din, h, dout)
        fd[i] = (lp - lm) / (2 * eps)
    assert np.max(np.abs(g - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_oracle_adam_first_step_moves_by_lr():
    # with bias correction the first step is lr * sign(g) (up to eps)
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    q, m, v = sr.adam(p, g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(p - q, 0.01 * np.sign(g), rtol=1e-4)


@pytest.mark.parametrize("n", [100, 3000, 20000, 300000])
def test_hll_estimate_tracks_cardinality(n):
    rng = np.random.default_rng(n)
    vals = rng.integers(0, 1 << 62, n, dtype=np.int64)
    regs = np.zeros(sr.HLL_M, np.uint8)
    for v in vals.tolist():
        h = sr.mix64(v)
        idx = h >> (64 - sr.HLL_BITS)
        rest = ((h << sr.HLL_BITS) & sr.M64) | (1 << (sr.HLL_BITS - 1))
        regs[idx] = max(regs[idx], 64 - rest.bit_length() + 1)
    E = sr.hll_estimate(regs)
    assert abs(E / n - 1) < 5 * 1.04 / 64


def test_features_of_hand_built_sketch():
    regs = np.zeros(sr.HLL_M, np.uint8)
    regs[:1024] = 3
    regs[1024:2048] = 5
    f = sr.features(regs, occ_s=3000.0, valid_s=25.0, frac=0.01, k=31, P=8)
    # mean register 2.0 -> centre 2: value v lands in bin v + 30
    assert f[30] == pytest.approx(0.5) and f[33] == pytest.approx(0.25) and f[35] == pytest.approx(0.25)
    # 2048 zero registers: linear counting E = m ln 2
    E = sr.HLL_M * math.log(2)
    assert f[64] == pytest.approx(E / 3000.0, rel=1e-6)
    assert f[65] == pytest.approx(math.log1p(2 * (1 - E / 3000.0) / 0.01) / 5, rel=1e-6)
    assert f[66] == pytest.approx(31 / 64) and f[67] == pytest.approx((120 + 30) / 256)
    assert f[68] == pytest.approx(0.1) and f[69] == pytest.approx(3 / 16) and f[71] == 1.0


def test_register_rank_and_index_layout():
    # the hash with only the lowest bit set has its first set bit (after the
    # 12 index bits) at position 52 -> rank 52; all-zero rest -> rank 53
    # through the sentinel bit; index = top 12 bits
    def rank_of(h):
        rest = ((h << sr.HLL_BITS) & sr.M64) | (1 << (sr.HLL_BITS - 1))
        return 64 - rest.bit_length() + 1
    assert rank_of(1) == 52
    assert rank_of(0) == 53
    assert rank_of(1 << 51) == 1
    assert (0xABC << 52) >> (64 - sr.HLL_BITS) == 0xABC


def test_committed_weights_shape():
    from genomix_amd import sizer
    if not os.path.exists(sizer.WEIGHTS):
        pytest.skip("sizing weights not trained yet")
    p = np.load(sizer.WEIGHTS, allow_pickle=False)
    assert p.dtype == np.float32 and p.size == DIN * H + H + H * H + H + DOUT * H + DOUT
    assert np.all(np.isfinite(p))
