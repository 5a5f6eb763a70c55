"""Learned table sizing on the GPU against the CPU oracle (oracle/sizer_ref.py).

* The MLP kernels: forward, gradient (backward) and Adam against the float64
  restatement; training reduces the loss.
* The HLL sketch of the build: registers bit-identical to the oracle's,
  features equal to float32 rounding.
* Sizing never changes the graph: identical bytes with and without the
  network, and with a network that predicts far too few k-mers (the
  grow-and-rebuild fallback runs and reports itself).  Sizing applies to
  the hash-table aggregation method.
* What the trained network gains: on read sets outside its training seeds,
  a table several times smaller than the occurrence-count sizing, without
  rebuilds.
"""

import os

import numpy as np
import pytest

from oracle import oracle
from oracle import sizer_ref as sr

pytestmark = pytest.mark.gpu

DIN, H, DOUT = 72, 64, 2


@pytest.fixture(scope="module")
def gx():
    import genomix_amd.gx as g
    return g


@pytest.fixture(scope="module")
def sizer():
    from genomix_amd import sizer as s
    return s


def _batch(seed, B=37):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (B, DIN)).astype(np.float32)
    T = rng.normal(0, 1, (B, DOUT)).astype(np.float32)
    return X, T


def test_mlp_forward_matches_oracle(sizer):
    p = sizer.init_params(seed=5)
    p[DIN * H:DIN * H + H] = 0.1  # non-zero biases
    m = sizer.Mlp(params=p)
    X, _ = _batch(1)
    y = m.forward(X)
    ref, _ = sr.forward(p, X, DIN, H, DOUT)
    assert np.allclose(y, ref, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("B", [1, 37, 256])
def test_mlp_gradient_matches_oracle(sizer, B):
    p = sizer.init_params(seed=B)
    m = sizer.Mlp(params=p)
    X, T = _batch(B, B)
    g, loss = m.gradient(X, T)
    gref = sr.gradient(p, X, T, DIN, H, DOUT)
    assert abs(loss - sr.loss(p, X, T, DIN, H, DOUT)) < 1e-5 * max(1.0, abs(loss))
    scale = np.max(np.abs(gref))
    assert np.max(np.abs(g - gref)) < 1e-4 * scale


def test_adam_step_matches_oracle(sizer):
    p = sizer.init_params(seed=9)
    m = sizer.Mlp(params=p)
    X, T = _batch(2, 64)
    q = p.astype(np.float64)
    mo = np.zeros_like(q)
    vo = np.zeros_like(q)
    for t in range(1, 4):
        g = sr.gradient(q, X, T, DIN, H, DOUT)
        q, mo, vo = sr.adam(q, g, mo, vo, t, 1e-3)
        m.train_step(X, T, 1e-3)
    assert np.max(np.abs(m.params() - q)) < 2e-5


def test_training_fits_a_smooth_target(sizer):
    rng = np.random.default_rng(4)
    X = rng.uniform(0, 1, (2048, DIN)).astype(np.float32)
    T = np.stack([np.sin(3 * X[:, 64]) - X[:, 66], X[:, 65] * X[:, 67]], 1).astype(np.float32)
    m = sizer.Mlp(params=sizer.init_params(seed=2))
    hist = sizer.train(m, X, T, epochs=60, batch=128, lr=3e-3)
    assert hist[-1] < 0.05 * hist[0]


def _text(seed, G=30000, cov=20, paired=True, k=None):
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=G, read_len=100, coverage=cov, error_rate=0.01, paired=paired,
                           insert_mean=300, insert_sd=25, seed=seed)
    return sim.simulate(spec)


@pytest.mark.parametrize("k", [21, 31, 55])
def test_hll_registers_and_features_match_oracle(gx, k):
    text = _text(600 + k, G=8000, cov=8)
    with gx.GraphBuilder(k, nparts=4) as b:
        b.add_reads(text)
        b.finish()
        f, regs, pred = b.features()
    rref, occ, valid, frac = sr.hll_registers(text, k)
    assert np.array_equal(regs, rref)
    fref = sr.features(rref, float(occ), float(valid), frac, k, 4)
    assert np.allclose(f, fref, rtol=1e-6, atol=1e-7)
    assert pred == [0.0, 0.0]


def test_sizing_never_changes_the_graph(gx, sizer):
    text = _text(71, G=60000, cov=25)
    base = oracle.build_graph(text, 31)[0]
    # a network that predicts ~e^-20 of the occurrences: far too small a table
    p = np.zeros(sizer.Mlp().n, np.float32)
    p[-2] = -20.0
    tiny = sizer.Mlp(params=p)
    with gx.GraphBuilder(31, sizer=tiny, method="hash") as b:
        b.add_reads(text)
        b.finish()
        assert b.part_bytes(0) == base
        st = b.stats()
        assert st["table_rebuilds"] >= 1 and st["predicted_distinct"] >= 1
        u1 = b.unitigs()
    with gx.GraphBuilder(31, method="hash") as b:
        b.add_reads(text)
        b.finish()
        assert b.stats()["table_rebuilds"] == 0
        assert b.unitigs() == u1


def test_trained_network_sizes_the_table(gx, sizer):
    if not os.path.exists(sizer.WEIGHTS):
        pytest.skip("sizing weights not trained yet")
    net = sizer.load()
    from genomix_amd import sim
    # read sets from seeds outside the training range, sizes inside it and above
    cases = [sim.ReadSetSpec(genome_len=150_000, read_len=150, coverage=40, error_rate=0.01, paired=True,
                             insert_mean=450, insert_sd=40, repeat_frac=0.1, repeat_len=(1000, 5000),
                             repeat_div=0.02, seed=990001),
             sim.ReadSetSpec(genome_len=1_000_000, read_len=100, coverage=30, error_rate=0.015, seed=990002)]
    for spec, k in zip(cases, (31, 21)):
        text = sim.simulate(spec)
        with gx.GraphBuilder(k, sizer=net, method="hash") as b:
            b.add_reads(text)
            b.finish()
            got = b.part_bytes(0)
            st = b.stats()
        with gx.GraphBuilder(k, method="hash") as b:
            b.add_reads(text)
            b.finish()
            assert b.part_bytes(0) == got
            plain = b.stats()
        D = st["distinct"]
        assert abs(st["predicted_distinct"] / D - 1) < 0.3
        assert st["table_rebuilds"] == 0
        assert st["table_capacity"] * 2 <= plain["table_capacity"]
