"""Diagnostics against closed forms and library routines."""
import numpy as np

from achr import stats as st


def ar1(rng, chains, draws, k, rho):
    x = np.empty((chains, draws, k))
    x[:, 0] = rng.standard_normal((chains, k)) / np.sqrt(1 - rho * rho)
    e = rng.standard_normal((chains, draws, k))
    for t in range(1, draws):
        x[:, t] = rho * x[:, t - 1] + e[:, t]
    return x


def test_ess_iid_is_about_n():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((8, 500, 5))
    e = st.ess(x)
    assert np.all(np.abs(e / 4000 - 1) < 0.15)


def test_ess_ar1_closed_form():
    """AR(1) with coefficient rho has ESS = N (1 - rho) / (1 + rho)."""
    rng = np.random.default_rng(1)
    for rho in (0.5, 0.9):
        x = ar1(rng, 8, 4000, 4, rho)
        expect = 8 * 4000 * (1 - rho) / (1 + rho)
        e = st.ess(x)
        assert np.all(np.abs(e / expect - 1) < 0.2), (rho, e, expect)
        b = st.bulk_ess(x)
        assert np.all(np.abs(b / expect - 1) < 0.25), (rho, b, expect)


def test_autocov_brute_force():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 17, 3))
    ac = st.autocov(x)
    xc = x - x.mean(axis=1, keepdims=True)
    for lag in range(17):
        brute = (xc[:, : 17 - lag] * xc[:, lag:]).sum(axis=1) / 17
        np.testing.assert_allclose(ac[:, lag], brute, atol=1e-12)


def test_rhat_detects_unmixed_chains():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 400, 2))
    assert np.all(st.rhat(x) < 1.02)
    x[0] += 3.0  # one chain stuck elsewhere
    assert np.all(st.rhat(x) > 1.2)


def test_ks_matches_scipy():
    from scipy.stats import ks_2samp
    rng = np.random.default_rng(4)
    a = rng.standard_normal((300, 3))
    b = rng.standard_normal((250, 3)) + np.array([0.0, 0.3, 1.0])
    ours = st.ks_2samp(a, b)
    ref = [ks_2samp(a[:, j], b[:, j]).statistic for j in range(3)]
    np.testing.assert_allclose(ours, ref, atol=1e-12)


def test_moment_error_fraction_of_range():
    lb, ub = np.array([0.0, -1.0, 2.0]), np.array([1.0, 1.0, 2.0])
    samples = np.array([[0.4, 0.6], [0.0, 0.2], [2.0, 2.0]])
    em, es = st.moment_error(samples, np.array([0.5, 0.0]).tolist() + [2.0], np.zeros(3), lb, ub)
    assert abs(em - 0.05) < 1e-12  # reaction 1: |0.1 - 0| / 2; fixed reaction 2 ignored
