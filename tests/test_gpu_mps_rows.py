"""Sampling a polytope given by MPS rows as written (E / L / G rows,
right-hand sides, RANGES; achr_model_read_mps_rows) on the GPU:
- closed forms: uniform on the 2-simplex x + y + z = 1 (a right-hand side),
  the triangle x + y <= 1 (an L row) and the band 1 <= x + y <= 1.5 in the
  unit square (a ranged G row) have means 1/3, 1/3 and 11/18 and, for the
  first two, variances 1/18; exact hit-and-run over a direction bank must
  hit them to a few standard errors;
- every sampled point, plain ACHR or bank, meets the file's rows;
- the sampler over such a model is the same path: GPU and oracle agree bit
  for bit;
- flux variability of the file columns equals scipy HiGHS on the file's own
  constraint system;
- createWarmupPts then ACHRCuda (ACHR_MPS_ROWS=1) end to end."""
import os
import subprocess

import numpy as np
import pytest

import achr
import oracle as O
from achr import synth
from test_mps_rows import file_lp_bounds, mixed_case, sides, write_mps

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def rows_ok(A, rows, lb, ub, X, tol):
    act = A @ X
    ok = (X >= lb[:, None] - tol).all() and (X <= ub[:, None] + tol).all()
    for i, (t, b, R) in enumerate(rows):
        lo, hi = sides(t, b, R)
        ok = ok and (act[i] >= lo - tol).all() and (act[i] <= hi + tol).all()
    return bool(ok)


SHAPES = {
    # name: A, rows, column bounds, exact means, exact variances (None: not checked)
    "simplex": (np.array([[1.0, 1.0, 1.0]]), [("E", 1.0, None)], (0.0, 1.0), [1 / 3] * 3, [1 / 18] * 3),
    "triangle": (np.array([[1.0, 1.0]]), [("L", 1.0, None)], (0.0, 1.0), [1 / 3] * 2, [1 / 18] * 2),
    "band": (np.array([[1.0, 1.0]]), [("G", 1.0, 0.5)], (0.0, 1.0), [11 / 18] * 2, None),
}


@pytest.mark.parametrize("name", list(SHAPES))
def test_uniform_moments_on_row_polytopes(tmp_path, name):
    say(f"shape {name}")
    A, rows, (lo, hi), mean, var = SHAPES[name]
    nf = A.shape[1]
    lb, ub = np.full(nf, lo), np.full(nf, hi)
    p = tmp_path / "m.mps"
    write_mps(p, A, rows, lb, ub, extra_free_row=False)
    model = achr.Model.read_mps(str(p), rows=True)
    model.compute_nullspace()
    W, fva, conv, rep = achr.warmup_points(model, 2 * nf + 20, seed=4)
    assert conv.all(), rep
    Lm, Um = model.bounds()
    rng = np.random.default_rng(3)
    free = synth.free_directions(model.nullspace(), Lm, Um)
    bank = free @ rng.standard_normal((free.shape[1], 64))
    bank /= np.linalg.norm(bank, axis=0)
    P, S = 20000, 300
    smp = achr.Sampler(model, W, achr.Options(reproject_mode=2))
    smp.set_directions(bank)
    smp.use_directions(True)
    smp.run(P, S, seed=11)
    X = smp.points()
    assert rows_ok(A, rows, lb, ub, X[:nf], 1e-9)
    m_hat, v_hat = X[:nf].mean(axis=1), X[:nf].var(axis=1)
    sem = np.sqrt(X[:nf].var(axis=1) / P)
    say(f"{name}: means {m_hat} (exact {mean}), variances {v_hat}, sem {sem}")
    assert (np.abs(m_hat - mean) < 5 * sem).all()
    if var is not None:
        assert (np.abs(v_hat - var) < 0.05 * np.asarray(var)).all()
    # plain ACHR over the same model: every point meets the rows
    plain = achr.Sampler(model, W, achr.Options(reproject_mode=2))
    plain.run(2000, 300, seed=12)
    Xp = plain.points()
    assert rows_ok(A, rows, lb, ub, Xp[:nf], 1e-7)  # the LP warmup meets the rows to the solver's tolerance
    say(f"{name}: plain ACHR means {Xp[:nf].mean(axis=1)}")


def test_mixed_rows_fva_parity_and_feasibility(tmp_path):
    A, rows, lb, ub = mixed_case(2)
    nf = A.shape[1]
    p = tmp_path / "m.mps"
    write_mps(p, A, rows, lb, ub)
    model = achr.Model.read_mps(str(p), rows=True)
    model.compute_nullspace()
    W, fva, conv, rep = achr.warmup_points(model, 2 * nf + 30, seed=6)
    say(f"mixed report {rep}")
    assert conv.all()
    g = np.array([file_lp_bounds(A, rows, lb, ub, j) for j in range(nf)])
    np.testing.assert_allclose(fva[:nf], g, rtol=0, atol=1e-7 * (1 + np.abs(g)).max())
    assert np.isnan(fva[nf:]).all()
    # the sampler over the rows model is the plain path: bit-equal to the oracle
    Lm, Um = model.bounds()
    rp, ci, va = model.csr()
    P, S, seed = 40, 60, 7
    smp = achr.Sampler(model, W, achr.Options(lanes=32, reproject_mode=1))
    smp.run(P, S, seed=seed)
    o, ost, _ = O.sample(warm=W, lb=Lm, ub=Um, rowptr=rp, colidx=ci, vals=va, N=model.nullspace(), m=model.m, P=P,
                         S=S, seed=seed, lanes=32, threads=8)
    np.testing.assert_array_equal(smp.stats(), ost)
    np.testing.assert_array_equal(smp.points(), o)
    assert rows_ok(A, rows, lb, ub, o[:nf], 1e-7)


def test_generator_cli_then_sampler_cli_with_rows(tmp_path):
    A, rows, lb, ub = mixed_case(3)
    nf = A.shape[1]
    write_mps(tmp_path / "m.mps", A, rows, lb, ub)
    env = dict(os.environ, ACHR_SEED="2")
    r = subprocess.run([achr.WARMUP_CLI_PATH, "m.mps", "30"], cwd=tmp_path, env=env, capture_output=True, text=True,
                       timeout=120)
    say(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "30 of 30 LPs converged" in r.stdout and "slack / right-hand-side columns" in r.stdout
    lines = (tmp_path / "m30warmup.csv").read_text().splitlines()
    assert len(lines) == nf
    Wf = np.array([[float(v) for v in l.split(",")] for l in lines])
    assert rows_ok(A, rows, lb, ub, Wf, 2e-5)  # "%f"
    fva = (tmp_path / "m_fva_bounds.csv").read_text().splitlines()
    assert len(fva) == nf + 1
    vals = np.array([[float(v) for v in l.split(",")[1:]] for l in fva[1:]])
    g = np.array([file_lp_bounds(A, rows, lb, ub, j) for j in range(nf)])
    np.testing.assert_allclose(vals, g, atol=1.5e-6)
    r = subprocess.run([achr.CLI_PATH, "m.mps", "m30warmup.csv", "2", "50", "40"], cwd=tmp_path,
                       env=dict(env, ACHR_MPS_ROWS="1"), capture_output=True, text=True, timeout=120)
    say(r.stdout)
    assert r.returncode == 0, r.stderr
    for k in range(2):
        pts = np.array([[float(v) for v in l.split(",")] for l in (tmp_path / f"File{k}").read_text().splitlines()])
        assert pts.shape == (nf, 50)
        # directions from the "%f" warmup leave the rows by ~1e-6 per unit step
        # between the residual checks
        assert rows_ok(A, rows, lb, ub, pts, 1e-4)
    # without the switch ACHRCuda takes the reference's view of the same file
    # (every row an equality with right-hand side 0), which this warmup
    # does not lie on: it still runs, writing n_file lines
    r = subprocess.run([achr.CLI_PATH, "m.mps", "m30warmup.csv", "1", "10", "5"], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert len((tmp_path / "File0").read_text().splitlines()) == nf
