"""MPS rows as written (achr_model_read_mps_rows): E / L / G rows with right-hand
sides and RANGES become {x' : S' x' = 0, lb' <= x' <= ub'} with slack columns
and one column fixed at 1. Host code only. Pinned by scipy HiGHS: the minimum
and maximum of every file column over the file's own constraint system (the
MPS range rules: E with R > 0 is [b, b + R], with R < 0 [b + R, b]; L is
[b - |R|, b]; G is [b, b + |R|]) equal those over the returned set; and a
point's completion meets the returned set exactly when the point meets the
file's rows."""
import numpy as np
import pytest
from scipy.optimize import linprog

import achr

BIG = np.finfo(float).max


def sides(t, b, R):
    if t == "E":
        return (b, b) if R is None else ((b, b + R) if R > 0 else (b + R, b))
    if t == "L":
        return (-np.inf if R is None else b - abs(R), b)
    return (b, np.inf if R is None else b + abs(R))


def write_mps(path, A, rows, lb, ub, extra_free_row=True):
    m, n = A.shape
    out = ["NAME rows", "ROWS", " N  obj"]
    out += [f" {t}  r{i}" for i, (t, _, _) in enumerate(rows)]
    if extra_free_row:
        out.append(" N  free")
    out.append("COLUMNS")
    for j in range(n):
        out.append(f"    c{j}  obj  {j + 1}")
        out += [f"    c{j}  r{i}  {float(A[i, j])!r}" for i in range(m) if A[i, j] != 0]
        if extra_free_row:
            out.append(f"    c{j}  free  1")
    out.append("RHS")
    out += [f"    rhs  r{i}  {float(b)!r}" for i, (_, b, _) in enumerate(rows) if b != 0]
    out.append("    rhs  obj  7")  # the objective's constant: ignored
    out.append("RANGES")
    out += [f"    rng  r{i}  {float(R)!r}" for i, (_, _, R) in enumerate(rows) if R is not None]
    out.append("BOUNDS")
    for j in range(n):
        out.append(f" LO BND c{j} {float(lb[j])!r}")
        out.append(f" UP BND c{j} {float(ub[j])!r}")
    out.append("ENDATA")
    path.write_text("\n".join(out) + "\n")


def mixed_case(seed):
    rng = np.random.default_rng(seed)
    n = 9
    types = ["E", "E", "E", "E", "L", "L", "G", "G", "E"]
    ranges = [None, 0.8, -0.6, None, None, 1.5, None, 0.9, None]
    A = np.round(rng.uniform(-2, 2, (len(types), n)), 3)
    A[rng.random(A.shape) < 0.4] = 0.0
    lb = -rng.uniform(1, 3, n)
    ub = rng.uniform(1, 3, n)
    x0 = rng.uniform(-0.3, 0.3, n)
    act = A @ x0
    rows = []
    for t, R, a in zip(types, ranges, act):
        if t == "E":
            b = a if R is None else (a - 0.3 * R)  # x0 inside the range
        elif t == "L":
            b = a + 0.4
        else:
            b = a - 0.4
        rows.append((t, float(np.round(b, 6)) if R is not None or t != "E" else float(a), R))
    rows[-1] = ("E", 0.0, None)  # an equality with right-hand side 0
    return A, rows, lb, ub


def file_lp_bounds(A, rows, lb, ub, j):
    A_eq, b_eq, A_ub, b_ub = [], [], [], []
    for a, (t, b, R) in zip(A, rows):
        lo, hi = sides(t, b, R)
        if lo == hi:
            A_eq.append(a)
            b_eq.append(lo)
            continue
        if np.isfinite(hi):
            A_ub.append(a)
            b_ub.append(hi)
        if np.isfinite(lo):
            A_ub.append(-a)
            b_ub.append(-lo)
    res = []
    for sgn in (1, -1):
        c = np.zeros(A.shape[1])
        c[j] = sgn
        r = linprog(c, A_ub=np.array(A_ub) if A_ub else None, b_ub=b_ub or None, A_eq=np.array(A_eq) if A_eq else None,
                    b_eq=b_eq or None, bounds=list(zip(lb, ub)), method="highs")
        assert r.status == 0
        res.append(sgn * r.fun)
    return res


def model_lp_bounds(model, j):
    S = model.dense()
    lb, ub = model.bounds()
    res = []
    for sgn in (1, -1):
        c = np.zeros(model.n)
        c[j] = sgn
        r = linprog(c, A_eq=S, b_eq=np.zeros(S.shape[0]), bounds=list(zip(lb, ub)), method="highs")
        assert r.status == 0
        res.append(sgn * r.fun)
    return res


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rows_view_has_the_files_polytope(tmp_path, seed):
    A, rows, lb, ub = mixed_case(seed)
    p = tmp_path / "m.mps"
    write_mps(p, A, rows, lb, ub)
    model = achr.Model.read_mps(str(p), rows=True)
    n_file = A.shape[1]
    n_slack = sum(sides(t, b, R)[0] != sides(t, b, R)[1] for t, b, R in rows)
    has_rhs = any(sides(t, b, R)[0] == sides(t, b, R)[1] != 0 for t, b, R in rows)
    assert model.file_columns == n_file
    assert model.n == n_file + n_slack + has_rhs
    assert model.m == len(rows)  # the free row is dropped
    Lm, Um = model.bounds()
    np.testing.assert_array_equal(Lm[:n_file], lb)
    np.testing.assert_array_equal(Um[:n_file], ub)
    if has_rhs:
        assert Lm[-1] == Um[-1] == 1.0
    for j in range(n_file):
        np.testing.assert_allclose(model_lp_bounds(model, j), file_lp_bounds(A, rows, lb, ub, j), rtol=0, atol=1e-7)
    # the reference's view of the same file: every row (the free one too) an
    # equality with right-hand side 0, no extra columns
    ref = achr.Model.read_mps(str(p))
    assert (ref.n, ref.m, ref.file_columns) == (n_file, len(rows) + 1, n_file)


def test_completion_meets_the_set_exactly_when_the_point_meets_the_rows(tmp_path):
    rng = np.random.default_rng(9)
    n = 6
    rows = [("L", 1.5, None), ("L", 0.8, 2.5), ("G", -1.5, None), ("G", -0.5, 1.8), ("E", -0.4, 1.6), ("E", 0.6, -1.7)]
    A = np.round(rng.uniform(-1, 1, (len(rows), n)), 2)
    lb, ub = -np.ones(n), np.ones(n)
    p = tmp_path / "m.mps"
    write_mps(p, A, rows, lb, ub, extra_free_row=False)
    model = achr.Model.read_mps(str(p), rows=True)
    assert model.n == n + len(rows)  # every row has two sides: slacks only, no right-hand-side column
    X = rng.uniform(-1, 1, (n, 4000))
    Xc = model.complete_points(X)
    np.testing.assert_array_equal(Xc[:n], X)
    S = model.dense()
    assert np.abs(S @ Xc).max() < 1e-14
    Lm, Um = model.bounds()
    in_model = ((Xc >= Lm[:, None]) & (Xc <= Um[:, None])).all(axis=0)
    act = A @ X
    in_file = np.ones(X.shape[1], bool)
    for i, (t, b, R) in enumerate(rows):
        lo, hi = sides(t, b, R)
        in_file &= (act[i] >= lo) & (act[i] <= hi)
    assert 200 < in_file.sum() < 3800
    np.testing.assert_array_equal(in_model, in_file)


def test_rhs_column_and_completion(tmp_path):
    A = np.array([[1.0, 1.0, 1.0]])
    p = tmp_path / "simplex.mps"
    write_mps(p, A, [("E", 1.0, None)], np.zeros(3), np.ones(3))
    model = achr.Model.read_mps(str(p), rows=True)
    assert (model.n, model.m, model.file_columns) == (4, 1, 3)
    np.testing.assert_array_equal(model.dense(), [[1, 1, 1, -1]])
    Xc = model.complete_points(np.array([[0.2], [0.3], [0.5]]))
    np.testing.assert_array_equal(Xc[:, 0], [0.2, 0.3, 0.5, 1.0])


@pytest.mark.parametrize("rows,lb,ub,why", [
    ([("L", 1.0, None)], [-BIG, 0.0], [1.0, 1.0], "unbounded below"),  # x0 + x1 <= 1 with x0 free below
    ([("G", 5.0, None)], [0.0, 0.0], [1.0, 1.0], "cannot be met"),     # x0 + x1 >= 5 in the unit square
    ([("E", 3.0, None)], [0.0, 0.0], [1.0, 1.0], "cannot be met"),     # x0 + x1 = 3 in the unit square
])
def test_rows_view_errors(tmp_path, rows, lb, ub, why):
    p = tmp_path / "bad.mps"
    A = np.array([[1.0, 1.0]])
    lines = ["NAME bad", "ROWS", " N  obj", f" {rows[0][0]}  r0", "COLUMNS", "    c0  r0  1", "    c1  r0  1", "RHS",
             f"    rhs  r0  {rows[0][1]}", "BOUNDS"]
    for j in range(2):
        lines.append(f" MI BND c{j}" if lb[j] == -BIG else f" LO BND c{j} {lb[j]}")
        lines.append(f" UP BND c{j} {ub[j]}")
    p.write_text("\n".join(lines + ["ENDATA"]) + "\n")
    assert A.shape == (1, 2)
    with pytest.raises(achr.AchrError) as e:
        achr.Model.read_mps(str(p), rows=True)
    assert e.value.code == -3 and why in str(e.value)
    achr.Model.read_mps(str(p))  # the reference's view reads it
