"""The exact finish of warmup LPs (achr_lp_polish: a primal active-set method
in the coordinates of an orthonormal null-space basis, host code) against
scipy HiGHS goldens (tests/golden/make_fva_golden.py), on the CPU: every
flux-variability LP of the 72 x 95 model and 120 of the reference's P. putida
model (VFWarmup/lib/P_Putida.mps), each started from the origin (inside the
bounds of both) or, for the small model, from an infeasible point that the
method first pulls into the bounds. The reference solves these LPs with
GLPK's simplex (VFWarmup/createWarmupPtsGLPK.c:100-175)."""
import numpy as np
import scipy.linalg

import achr
from conftest import golden, load_c1, parse_mps_py


def highs(name):
    return np.loadtxt(golden(name), delimiter=",")[:, 1:]


def run_fva(S, lb, ub, cols, start, xfeas):
    N = scipy.linalg.null_space(S)
    n = S.shape[1]
    out = np.empty((len(cols), 2))
    for r, i in enumerate(cols):
        for k, sgn in enumerate((1.0, -1.0)):
            c = np.zeros(n)
            c[i] = sgn
            rc, x = achr.lp_polish(N, lb, ub, c, xfeas, start)
            assert rc == 0, (i, sgn, rc)
            # on S x = 0 to the bounds' feasibility tolerance (1e-9 of each bound's size)
            assert np.abs(S @ x).max() < 5e-9 * max(1.0, np.abs(x).max())
            assert (x >= lb).all() and (x <= ub).all()
            out[r, k] = x[i]
    return out


def test_polish_matches_highs_on_c1():
    S, lb, ub, _, We = load_c1()
    g = highs("c1_fva_highs.csv")
    xfeas = We.mean(axis=1)  # the exact warmup's centre: inside the bounds, on S x = 0
    cols = np.arange(S.shape[1])
    for start in (xfeas, xfeas + 5.0 * np.sin(np.arange(S.shape[1]))):  # the second violates bounds
        got = run_fva(S, lb, ub, cols, start, xfeas)
        err = np.abs(got - g) / (1 + np.abs(g))
        assert err.max() < 1e-7, (err.max(), np.argwhere(err > 1e-7)[:5])


def test_polish_matches_highs_on_p_putida_sample():
    S, lb, ub = parse_mps_py(golden("P_Putida.mps"))
    assert (lb <= 0).all() and (ub >= 0).all()  # the origin is feasible
    g = highs("p_putida_fva_highs.csv")
    cols = np.random.default_rng(3).choice(S.shape[1], 60, replace=False)
    z = np.zeros(S.shape[1])
    got = run_fva(S, lb, ub, cols, z, z)
    err = np.abs(got - g[cols]) / (1 + np.abs(g[cols]))
    print(f"P. putida sample: worst relative error {err.max():.3g}")
    assert err.max() < 1e-7
