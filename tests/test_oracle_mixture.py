"""The oracle's FittingMixture restatement (FittingMixture.java:89-219) pinned
without the reference (commons-math and a JVM are absent; the reference has
no test of it): its densities and initial state against scipy.stats, its
fixed point against the mixture likelihood's stationarity, its estimates
against the generating parameters of a separated sample, its cutoff rule
against scipy densities, and the closed form of a single-valued data set."""

import numpy as np
import pytest
from scipy import stats

from oracle import oracle

FLOOR_SD = float(np.float32(1e-9))  # the float literal 0.000000001f, widened


def _sample(seed, n_exp=20000, exp_mean=2.0, n_norm=20000, mu=40.0, sd=4.0):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.exponential(exp_mean, n_exp), rng.normal(mu, sd, n_norm)])


def _loglik(d, em, mu, sd, pe):
    f = pe * stats.expon.pdf(d, scale=em) + (1 - pe) * stats.norm.pdf(d, mu, sd)
    return np.log(f).sum()


@pytest.mark.parametrize("outliers", [False, True])
def test_initial_expectation_step(outliers):
    """No M step: the initial distributions (exponential mean 5, normal 20/5)
    and the priors after one E step, recomputed with scipy's densities.
    Outliers (repeat k-mers of coverage 500 and 5000) make the normal's and
    then also the exponential's density underflow to 0, where the 1e-9 floors
    decide the memberships."""
    d = _sample(1, 3000, 3.0, 3000, 30.0, 6.0)
    if outliers:
        d = np.concatenate([d, np.full(300, 500.0), np.full(200, 5000.0)])
    got = oracle.fitting_mixture(d, d.max(), iterations=0)
    assert (got["exp_mean"], got["normal_mean"], got["normal_std"]) == (5.0, 20.0, 5.0)
    a = 0.5 * stats.expon.pdf(d, scale=5.0)
    c = 0.5 * stats.norm.pdf(d, 20.0, 5.0)
    a[a == 0] = 1e-9
    c[c == 0] = 1e-9
    pe = (a / (a + c)).sum() / len(d)
    assert got["prob_exp"] == pytest.approx(pe, rel=1e-12)
    assert got["prob_exp"] + got["prob_normal"] == pytest.approx(1.0, rel=1e-15)


@pytest.mark.parametrize("seed,params", [(2, (2.0, 40.0, 4.0)), (3, (1.5, 25.0, 5.0)), (4, (3.0, 30.0, 6.0))])
def test_converges_to_generating_parameters(seed, params):
    em, mu, sd = params
    d = _sample(seed, exp_mean=em, mu=mu, sd=sd)
    got = oracle.fitting_mixture(d, d.max(), iterations=300)
    assert got["exp_mean"] == pytest.approx(em, rel=0.05)
    assert got["normal_mean"] == pytest.approx(mu, rel=0.02)
    assert got["normal_std"] == pytest.approx(sd, rel=0.05)
    assert got["prob_exp"] == pytest.approx(0.5, abs=0.02)


@pytest.mark.parametrize("seed,params", [(2, (2.0, 40.0, 4.0)), (9, (4.0, 12.0, 3.0)), (11, (6.0, 20.0, 8.0)),
                                         (12, (2.0, 25.0, 10.0))])
def test_fixed_point_is_stationary(seed, params):
    """EM's fixed point is a stationary point of the mixture likelihood
    (where the 1e-9 membership floors never bind, as here, overlapping
    components included): every partial derivative, by central differences
    of the log-likelihood with scipy densities, vanishes."""
    em, mu, sd = params
    d = _sample(seed, exp_mean=em, mu=mu, sd=sd)
    got = oracle.fitting_mixture(d, d.max(), iterations=500)
    theta = np.array([got["exp_mean"], got["normal_mean"], got["normal_std"], got["prob_exp"]])
    for j in range(4):
        h = 1e-5 * theta[j]
        up, dn = theta.copy(), theta.copy()
        up[j] += h
        dn[j] -= h
        grad = (_loglik(d, *up) - _loglik(d, *dn)) / (2 * h)
        assert abs(grad * theta[j]) / len(d) < 1e-6, j


def test_likelihood_rises_over_the_reference_iteration_count():
    """Ten rounds, as GenomixDriver runs, from the fixed initial guess: the
    likelihood after ten is above that after one."""
    d = _sample(5, 8000, 2.5, 12000, 35.0, 6.0)
    r1 = oracle.fitting_mixture(d, d.max(), iterations=1)
    r10 = oracle.fitting_mixture(d, d.max(), iterations=10)
    l1 = _loglik(d, r1["exp_mean"], r1["normal_mean"], r1["normal_std"], r1["prob_exp"])
    l10 = _loglik(d, r10["exp_mean"], r10["normal_mean"], r10["normal_std"], r10["prob_exp"])
    assert l10 > l1


@pytest.mark.parametrize("max_value", [None, 10.0, 1.0, 23.0, 24.0])
def test_cutoff_rule(max_value):
    """The cutoff is the first integer c in [1, max) where the weighted
    normal density exceeds the weighted exponential one (scipy densities at
    the fitted parameters), else 0."""
    d = _sample(6)
    mx = d.max() if max_value is None else max_value
    got = oracle.fitting_mixture(d, mx, iterations=10)
    want = 0.0
    c = 1.0
    while c < mx:
        if got["prob_exp"] * stats.expon.pdf(c, scale=got["exp_mean"]) < \
                got["prob_normal"] * stats.norm.pdf(c, got["normal_mean"], got["normal_std"]):
            want = c
            break
        c += 1
    assert got["cutoff"] == want
    if max_value is None:
        assert 10 < got["cutoff"] < 35


@pytest.mark.parametrize("c", [8.0, 32.0])
def test_single_value_closed_form(c):
    """All data equal c: both means are c, the normal's sd is 0 and takes the
    floor 1e-9f, so its density vanishes off c and there is no cutoff below
    max = c.  (c a power of two: the weighted sums are then exactly c times
    the membership sums, so the closed form holds without rounding.)"""
    d = np.full(500, c)
    got = oracle.fitting_mixture(d, c, iterations=3)
    assert got["exp_mean"] == c and got["normal_mean"] == c
    assert got["normal_std"] == FLOOR_SD
    assert got["cutoff"] == 0.0
    assert got["prob_normal"] > 1 - 1e-6


def test_order_of_the_data_changes_rounding_only():
    """The reference walks its data in the Hadoop counter group's order; any
    order gives the same fit up to rounding and the same cutoff."""
    d = np.round(_sample(7, 5000, 2.0, 5000, 30.0, 5.0)).clip(1)
    a = oracle.fitting_mixture(np.sort(d), d.max())
    b = oracle.fitting_mixture(np.random.default_rng(0).permutation(d), d.max())
    assert a["cutoff"] == b["cutoff"]
    for key in ("exp_mean", "normal_mean", "normal_std", "prob_exp"):
        assert a[key] == pytest.approx(b[key], rel=1e-12)


def test_no_coverage_is_an_error():
    with pytest.raises(oracle.OracleError):
        oracle.fitting_mixture(np.zeros(0), 5.0)
    with pytest.raises(oracle.OracleError):
        oracle.fitting_mixture(np.ones(3), 0.0)
