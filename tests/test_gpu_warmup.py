"""Warmup generation on the GPU (flux variability + random-objective LPs,
batched PDHG) against scipy HiGHS goldens (tests/golden/make_fva_golden.py):
every LP reported converged has the HiGHS optimum to 1e-6 relative, the
points keep the reference's layout and move toward their mean, and the
reference's CI shape check (P. putida: 1,060 x 2,120) holds."""
import numpy as np
import pytest
import scipy.sparse as sp

import achr
from conftest import golden

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def highs(name):
    return np.loadtxt(golden(name), delimiter=",")[:, 1:]


def matrix(model):
    rp, ci, va = model.csr()
    return sp.csr_matrix((va, ci, rp), shape=(model.m, model.n))


def test_c1_flux_variability_all_converge_to_highs():
    model = achr.Model.read_mps(golden("c1.mps"))
    n = model.n
    W, fva, conv, rep = achr.warmup_points(model, 2 * n + 10, seed=3)
    say(f"c1 report {rep}")
    assert conv.all() and rep["converged"] == 2 * n + 10
    g = highs("c1_fva_highs.csv")
    np.testing.assert_allclose(fva, g, rtol=0, atol=1e-7 * (1 + np.abs(g)).max())
    lb, ub = model.bounds()
    S = matrix(model)
    # the move toward the mean is an affine map of the LP solutions: undo it
    # and every solution lies in the bounds and (to the solver's tolerance) on S x = 0
    c = W.mean(axis=1)
    X = (W - 0.67 * c[:, None]) / 0.33
    assert np.abs(S @ X).max() < 1e-6
    assert (X >= lb[:, None] - 1e-9).all() and (X <= ub[:, None] + 1e-9).all()
    # columns 2i / 2i+1 are the minimum / maximum of reaction i
    np.testing.assert_allclose(X[np.arange(n), 2 * np.arange(n)], fva[:, 0], atol=1e-9)
    np.testing.assert_allclose(X[np.arange(n), 2 * np.arange(n) + 1], fva[:, 1], atol=1e-9)


def test_p_putida_all_lps_match_highs():
    """Every one of the 2,120 flux-variability LPs of the reference's P.
    putida model equals the HiGHS optimum to 1e-6 relative: the batched
    iteration converges most of them, the exact host finish the rest (the
    reference's GLPK simplex solves all, createWarmupPtsGLPK.c:100-175)."""
    model = achr.Model.read_mps(golden("P_Putida.mps"))
    n = model.n
    W, fva, conv, rep = achr.warmup_points(model, 2 * n, max_iters=40000)
    say(f"P. putida report {rep}")
    assert W.shape == (1060, 2120)  # the reference's CI shape check (vfwarmup.yml)
    g = highs("p_putida_fva_highs.csv")
    err = np.abs(fva - g) / (1 + np.abs(g))
    say(f"converged {int(conv.sum())} of {2 * n} ({rep['polished']} by the exact finish); worst error {err.max():.3g}")
    assert conv.all() and rep["converged"] == 2 * n and rep["polished"] > 0
    assert err.max() < 1e-6
    lb, ub = model.bounds()
    c = W.mean(axis=1)
    X = (W - 0.67 * c[:, None]) / 0.33
    assert (X >= lb[:, None] - 1e-6).all() and (X <= ub[:, None] + 1e-6).all()
    # S x = 0 to the bounds' feasibility tolerance (1e-9 of each bound's size),
    # below the 1e-6 resolution of the "%f" file these points are written to
    assert np.abs(matrix(model) @ X).max() < 5e-6
    # without the finish, the iteration alone leaves some LPs approximate
    _, fva0, conv0, rep0 = achr.warmup_points(model, 2 * n, max_iters=40000, polish=False)
    assert rep0["polished"] == 0 and rep0["converged"] == int(conv0.sum()) < 2 * n


def test_generator_cli_then_sampler_cli(tmp_path):
    """The reference's two programs end to end: createWarmupPts writes
    <model><nPts>warmup.csv and <model>_fva_bounds.csv beside the model, and
    ACHRCuda samples from that warmup file."""
    import os
    import shutil
    import subprocess
    shutil.copy(golden("c1.mps"), tmp_path / "m.mps")
    env = dict(os.environ, ACHR_SEED="5")
    r = subprocess.run([achr.WARMUP_CLI_PATH, "m.mps", "150", "-1"], cwd=tmp_path, env=env, capture_output=True,
                       text=True, timeout=120)
    say(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "setting nPts=190" in r.stdout and "190 of 190 LPs converged" in r.stdout
    lines = (tmp_path / "m190warmup.csv").read_text().splitlines()
    assert len(lines) == 95 and all(len(l.split(",")) == 190 for l in lines)
    fva = (tmp_path / "m_fva_bounds.csv").read_text().splitlines()
    assert fva[0] == "reaction,min,max" and len(fva) == 96
    g = highs("c1_fva_highs.csv")
    vals = np.array([[float(v) for v in l.split(",")[1:]] for l in fva[1:]])
    np.testing.assert_allclose(vals, g, atol=1.5e-6)  # "%f"
    W = np.array([[float(v) for v in l.split(",")] for l in lines])
    np.testing.assert_allclose(W, achr.read_warmup_csv(str(tmp_path / "m190warmup.csv"), 95), atol=0)
    r = subprocess.run([achr.CLI_PATH, "m.mps", "m190warmup.csv", "1", "20", "30"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pts = np.array([[float(v) for v in l.split(",")] for l in (tmp_path / "File0").read_text().splitlines()])
    model = achr.Model.read_mps(golden("c1.mps"))
    lb, ub = model.bounds()
    assert pts.shape == (95, 20)
    assert (pts >= lb[:, None] - 1e-6).all() and (pts <= ub[:, None] + 1e-6).all()
