"""The coverage cutoff (include/genomix_stats.h: gx_coverage_cutoff and
gx_fit_mixture) against the oracle's FittingMixture restatement over the
expanded coverage data.  The fit is floating point: the cutoff must be equal
and the fitted parameters within a relative 1e-9 (the library sums over
coverage bins weighted by their counts, the reference over one entry per
node, so the two round differently)."""

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

REL = 1e-9


@pytest.fixture(scope="module")
def gx():
    import genomix_amd.gx as g
    return g


def _same(got, want):
    assert got["cutoff"] == want["cutoff"]
    for key in ("exp_mean", "normal_mean", "normal_std", "prob_exp", "prob_normal"):
        assert got[key] == pytest.approx(want[key], rel=REL), key


def _expand(bins):
    return np.repeat(np.array([float(v) for v in bins], dtype=np.float64),
                     np.array(list(bins.values()), dtype=np.int64))


def _coverage_like(seed, n_err, n_gen, mean):
    rng = np.random.default_rng(seed)
    d = np.concatenate([rng.exponential(1.5, n_err), rng.normal(mean, mean / 4, n_gen),
                        rng.integers(3 * mean, 20 * mean, n_gen // 50)])
    return np.maximum(np.round(d), 1).astype(np.int64)


@pytest.mark.parametrize("seed,n_err,n_gen,mean,iters", [(1, 50000, 40000, 30, 10), (2, 200000, 20000, 12, 10),
                                                         (3, 5000, 100000, 60, 25), (4, 100, 100, 8, 0)])
def test_fit_over_bins_matches_expanded_oracle(gx, seed, n_err, n_gen, mean, iters):
    d = _coverage_like(seed, n_err, n_gen, mean)
    vals, cnts = np.unique(d, return_counts=True)
    got = gx.fit_mixture(vals.tolist(), cnts.tolist(), iterations=iters)
    _same(got, oracle.fitting_mixture(d.astype(np.float64), float(d.max()), iters))


def test_fit_edge_cases(gx):
    # one value: both means equal it, the normal's sd takes the 1e-9f floor
    got = gx.fit_mixture([16], [100])
    _same(got, oracle.fitting_mixture(np.full(100, 16.0), 16.0))
    assert got["normal_std"] == float(np.float32(1e-9)) and got["cutoff"] == 0.0
    # an explicit max below the crossing: no cutoff
    d = _coverage_like(5, 20000, 20000, 30)
    vals, cnts = np.unique(d, return_counts=True)
    _same(gx.fit_mixture(vals.tolist(), cnts.tolist(), max_value=10.0),
          oracle.fitting_mixture(d.astype(np.float64), 10.0))
    for args in (([], []), ([5], [0]), ([0], [3])):
        with pytest.raises(gx.GxError) as e:
            gx.fit_mixture(*args)
        assert e.value.code == gx.GX_EINVAL
    with pytest.raises(gx.GxError) as e:
        gx.fit_mixture([3], [1], iterations=-1)
    assert e.value.code == gx.GX_EINVAL


def _reads(seed, G, L, cov, err):
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=G, read_len=L, coverage=cov, error_rate=err, paired=True,
                           insert_mean=3 * L, insert_sd=L // 4, repeat_frac=0.05, repeat_len=(300, 900),
                           repeat_div=0.01, seed=seed)
    return sim.simulate(spec)


@pytest.mark.parametrize("k,cov", [(21, 40), (31, 60), (55, 80)])
def test_driver_flow_build_cutoff_clean(gx, k, cov):
    """BUILD_HYRACKS, the cutoff over the node graph's coverage counters,
    then REMOVE_BAD_COVERAGE with it and MERGE
    (GenomixDriver.java:192-197,119-136), against the oracle step by step."""
    text = _reads(4400 + k, 30000, 100, cov, 0.01)
    recs = oracle.build_graph(text, k)[0]
    want_bins = oracle.graph_stats(recs, k)["coverage-bins"]
    want = oracle.fitting_mixture(_expand(want_bins), max(float(v) for v in want_bins), 10)
    with gx.GraphBuilder(k) as b:
        b.add_reads(text)
        b.finish()
        got = b.coverage_cutoff()
        _same(got, want)
        # erroneous k-mers at coverage 1-2 sit below the cutoff, the genome's
        # peak above it
        assert 2 <= got["cutoff"] < got["normal_mean"]
        b.run_unitigs()
        cleaned = b.clean_bad_coverage(got["cutoff"])
    assert cleaned == oracle.clean_bad_coverage(recs, k, want["cutoff"])


def test_cutoff_needs_statistics(gx):
    import ctypes
    with gx.GraphBuilder(21) as b:
        b.add_reads(_reads(9, 5000, 100, 10, 0.01))
        b.finish()
        out = (ctypes.c_double * 6)()
        assert gx.lib().gx_coverage_cutoff(b._h, 10, out) == gx.GX_ESTATE
        b.run_unitigs()
        # statistics of the merged graph serve as well
        got = b.coverage_cutoff(graph="unitigs")
        bins = b.graph_stats("unitigs")["coverage-bins"]
    _same(got, oracle.fitting_mixture(_expand(bins), max(float(v) for v in bins), 10))


def test_statistics_do_not_outlive_their_graph(gx):
    """A new build, a reset, a re-configuration or a graph step drops the
    previous statistics: the cutoff then needs gx_graph_stats again."""
    import ctypes
    out = (ctypes.c_double * 6)()
    text = _reads(9, 5000, 100, 10, 0.01)
    with gx.GraphBuilder(21) as b:
        b.add_reads(text)
        b.finish()
        b.coverage_cutoff()
        assert gx.lib().gx_coverage_cutoff(b._h, 10, out) == gx.GX_OK
        b.reset()
        assert gx.lib().gx_coverage_cutoff(b._h, 10, out) == gx.GX_ESTATE
        b.add_reads(text)
        b.finish()
        assert gx.lib().gx_coverage_cutoff(b._h, 10, out) == gx.GX_ESTATE
        b.coverage_cutoff()
        b.run_unitigs()
        assert gx.lib().gx_coverage_cutoff(b._h, 10, out) == gx.GX_ESTATE
        b.coverage_cutoff(graph="unitigs")
        b.clean_bad_coverage(2.0, emit=False)
        assert gx.lib().gx_coverage_cutoff(b._h, 10, out) == gx.GX_ESTATE
