"""Pins of the CPU oracle: published vectors, an independent Philox, library
routines and closed forms. No GPU needed."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import golden

# Random123 known-answer vectors for philox4x32-10 (Salmon et al., SC'11, kat_vectors).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_kat(ctr, key, expect):
    assert O.philox(ctr, key) == expect


def test_philox_matches_rocrand_engine():
    rows = [l.split() for l in open(golden("philox_rocrand.txt")) if not l.startswith("#")]
    assert len(rows) == 64
    for r in rows:
        v = [int(t, 16) for t in r]
        assert O.philox(tuple(v[0:4]), tuple(v[4:6])) == tuple(v[6:10])


def test_stream_uniforms_range_and_moments():
    u = np.array([O.stream_u01(123, i, 7, 0, 0, 1) for i in range(20000)])
    assert (u >= 0).all() and (u < 1).all()
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1 / 12) < 0.005
    # tags, chains and files select different streams
    assert O.stream_u01(1, 0, 0, 0, 0, 0) != O.stream_u01(1, 0, 0, 0, 0, 1)
    assert O.stream_u01(1, 0, 0, 0, 0, 0) != O.stream_u01(1, 0, 1, 0, 0, 0)
    assert O.stream_u01(1, 0, 0, 0, 0, 0) != O.stream_u01(1, 0, 0, 1, 0, 0)


def test_center_is_mean(c1):
    S, lb, ub, Wf, We = c1
    c = O.center(We)
    np.testing.assert_allclose(c, We.mean(axis=1), rtol=0, atol=1e-12)


def test_nullspace_matches_scipy_projector(c1):
    import scipy.linalg
    S = c1[0]
    N = O.nullspace(S)
    N2 = scipy.linalg.null_space(S)
    assert N.shape == N2.shape == (95, 23)
    assert np.abs(S @ N).max() < 1e-12
    assert np.abs(N.T @ N - np.eye(N.shape[1])).max() < 1e-12
    assert np.abs(N @ N.T - N2 @ N2.T).max() < 1e-12
    diag = np.loadtxt(golden("c1_nullproj_diag.txt"))
    np.testing.assert_allclose(np.diag(N @ N.T), diag, atol=1e-12)


def test_nullspace_rank_deficient_and_trivial():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((4, 9))
    S = np.vstack([A, A[0] + A[1], 2 * A[2]])  # rank 4, 6 rows
    N = O.nullspace(S)
    assert N.shape == (9, 5)
    assert np.abs(S @ N).max() < 1e-12
    # a full-column-rank S has a trivial null space
    assert O.nullspace(rng.standard_normal((5, 3))).shape == (3, 0)


def _box_model(lb, ub, W):
    n = len(lb)
    return dict(warm=W, lb=np.asarray(lb, float), ub=np.asarray(ub, float), rowptr=np.zeros(1, np.int32),
                colidx=np.zeros(1, np.int32), vals=np.zeros(1), N=np.eye(n), m=0)


def test_one_step_chord_on_a_box():
    """A box (m = 0): one step from the centre lands at c + t u with t from the
    textbook chord [max_i (bound_i - c_i)/u_i over the entering side, min_i over
    the leaving side] and t = r (max - min) + min."""
    rng = np.random.default_rng(3)
    n, W = 5, 12
    lb = -rng.uniform(1, 3, n)
    ub = rng.uniform(1, 3, n)
    Wm = lb[:, None] + (ub - lb)[:, None] * rng.uniform(0.1, 0.9, (n, W))
    c = Wm.mean(axis=1)
    seed, P = 99, 6
    pts, st, _ = O.sample(**_box_model(lb, ub, Wm), P=P, S=1, seed=seed, lanes=16)
    for b in range(P):
        up = O.stream_u01(seed, 0, b, 0, 0, 0)
        w = Wm[:, min(int(W * up), W - 1)]
        u = (w - c) / np.linalg.norm(w - c)
        lo = np.where(u > 0, (lb - c) / u, (ub - c) / u)
        hi = np.where(u > 0, (ub - c) / u, (lb - c) / u)
        tmin, tmax = lo.max(), hi.min()
        r = O.stream_u01(seed, 0, b, 0, 0, 1)
        x = c + (r * (tmax - tmin) + tmin) * u
        np.testing.assert_allclose(pts[:, b], x, rtol=1e-12, atol=1e-12)
    assert (st[:, 0] == 0).all()


def test_box_uniform_moments_and_ks():
    """Hit-and-run on a box converges to the uniform law: mean (lb+ub)/2,
    std (ub-lb)/sqrt(12), and a small KS distance to U(lb, ub)."""
    lb = np.array([0.0, -1.0, 2.0])
    ub = np.array([1.0, 3.0, 2.5])
    rng = np.random.default_rng(5)
    Wm = lb[:, None] + (ub - lb)[:, None] * rng.uniform(0, 1, (3, 20))
    pts, st, _ = O.sample(**_box_model(lb, ub, Wm), P=4000, S=60, seed=11, lanes=16, threads=8)
    mean = pts.mean(axis=1)
    std = pts.std(axis=1)
    np.testing.assert_allclose(mean, (lb + ub) / 2, atol=0.02 * (ub - lb).max())
    np.testing.assert_allclose(std, (ub - lb) / np.sqrt(12), rtol=0.03)
    for i in range(3):
        z = np.sort((pts[i] - lb[i]) / (ub[i] - lb[i]))
        ks = np.max(np.abs(np.arange(1, len(z) + 1) / len(z) - z))
        assert ks < 0.03


def test_samples_feasible_and_clipped(c1):
    S, lb, ub, Wf, We = c1
    rp, ci, va = O.csr_of(S)
    N = O.nullspace(S)
    pts, st, _ = O.sample(warm=Wf, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=64, S=200,
                          seed=3, lanes=32)
    assert (pts >= lb[:, None]).all() and (pts <= ub[:, None]).all()
    # %f warmup: every residual check finds |S x| > 1e-9 and reprojects
    assert (st[:, 1] == st[:, 2]).all() and st[0, 2] == 20
    # exact warmup: drift stays far below the trigger and no reprojection happens
    pts2, st2, _ = O.sample(warm=We, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=64, S=200,
                            seed=3, lanes=32)
    assert (st2[:, 1] == 0).all()
    assert np.abs(S @ pts2).max() < 1e-9


def test_bank_clip_orders_signed_zeros():
    """Bank mode clips with the IEEE 754-2019 minimumNumber / maximumNumber,
    which order -0 below +0 (the kernel's v_min_f64 / v_max_f64): a blocked
    reaction at +0.0 under the upper bound -0.0 ends at -0.0, where a
    compare-and-select clip (+0 > -0 is false) would keep +0.0; a reaction at
    +0.0 on the lower bound +0.0 stays +0.0. Every point is inside its
    bounds, and the moving reactions are unaffected by the zeros."""
    from conftest import signed_zero_bank_case
    S, lb, ub, N, W, c0, bank = signed_zero_bank_case()
    rp, ci, va = O.csr_of(S)
    pts, st, _ = O.sample(warm=W, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=2, P=8, S=60, seed=9,
                          lanes=32, bank=bank, c0=c0)
    assert (st[:, 0] < 60).all() and (st[:, 3] == 0).all()
    bits = pts.view(np.uint64)
    assert (bits[0] == np.uint64(0x8000000000000000)).all()
    assert (bits[1] == 0).all()
    assert (pts >= lb[:, None]).all() and (pts <= ub[:, None]).all()
    assert np.abs(pts[2:] - c0[2:, None]).min() > 0  # the free reactions moved
    # zero steps: the start point as it is (no clip without an attempt)
    pts0, _, _ = O.sample(warm=W, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=2, P=2, S=0, seed=9,
                          lanes=32, bank=bank, c0=c0)
    assert (pts0.view(np.uint64)[0] == 0).all()


def test_reduction_lanes_change_only_rounding(c1):
    """Different lane trees reorder one sum; trajectories agree to rounding."""
    S, lb, ub, Wf, We = c1
    rp, ci, va = O.csr_of(S)
    N = O.nullspace(S)
    kw = dict(warm=We, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=16, S=50, seed=1)
    a = O.sample(lanes=16, **kw)[0]
    b = O.sample(lanes=256, **kw)[0]
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9)


def test_centre_telescopes_through_trace(c1):
    """c_T = (W c0 + sum_{s<T} x_s) / (W + T): the recorded trajectory of a chain
    rebuilds the centre the next step uses; checked through a one-step replay."""
    S, lb, ub, Wf, We = c1
    rp, ci, va = O.csr_of(S)
    N = O.nullspace(S)
    W = We.shape[1]
    T = 12
    kw = dict(warm=We, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=1, seed=4, lanes=16)
    _, _, tr = O.sample(S=T + 1, trace_chains=1, trace_thin=1, **kw)
    c0 = We.mean(axis=1)
    cT = (W * c0 + tr[0, :T].sum(axis=0)) / (W + T)
    # step T+1 from x_T along (w - c_T): recompute with numpy and compare with the trace
    up = O.stream_u01(4, T, 0, 0, 0, 0)  # no misses on this model, attempt index == step index
    w = We[:, min(int(W * up), W - 1)]
    u = (w - cT) / np.linalg.norm(w - cT)
    x = tr[0, T - 1]
    du, dl = ub - x, x - lb
    keep = (du > 1e-14) & (dl > 1e-14)
    pos, neg = keep & (u > 1e-9), keep & (u < -1e-9)
    tmin = max(np.max(-dl[pos] / u[pos], initial=-np.inf), np.max(du[neg] / u[neg], initial=-np.inf))
    tmax = min(np.min(du[pos] / u[pos], initial=np.inf), np.min(-dl[neg] / u[neg], initial=np.inf))
    r = O.stream_u01(4, T, 0, 0, 0, 1)
    xn = np.clip(x + (r * (tmax - tmin) + tmin) * u, lb, ub)
    np.testing.assert_allclose(tr[0, T], xn, rtol=1e-9, atol=1e-9)


def test_miss_cap_flags_degenerate_polytope():
    """Every warmup point equal to the centre: |w - c| = 0 is a miss each time;
    the chain stops after max_misses consecutive misses and is flagged."""
    n = 3
    lb, ub = -np.ones(n), np.ones(n)
    Wm = np.zeros((n, 4))
    pts, st, _ = O.sample(**_box_model(lb, ub, Wm), P=2, S=5, seed=0, lanes=16, max_misses=50)
    assert (st[:, 3] == 1).all() and (st[:, 0] == 51).all()
    np.testing.assert_array_equal(pts, 0.0)


def _chord(x, u, lb, ub):
    """Chord [tmin, tmax] of fillrandPoint (ACHRCuda.cu:266-288) in numpy."""
    du, dl = ub - x, x - lb
    keep = (du > 1e-14) & (dl > 1e-14)
    pos, neg = keep & (u > 1e-9), keep & (u < -1e-9)
    tmin = max(np.max(-dl[pos] / u[pos], initial=-np.inf), np.max(du[neg] / u[neg], initial=-np.inf))
    tmax = min(np.min(du[pos] / u[pos], initial=np.inf), np.min(-dl[neg] / u[neg], initial=np.inf))
    return tmin, tmax


@pytest.mark.parametrize("step", [0, 10, 20])
def test_reprojection_value_replayed_with_scipy_projector(c1, step):
    """The value of the reprojection x <- N (N^T x) (reprojectPoint /
    reprojectPoint2, ACHRCuda.cu:208-230, triggered at :325-343) on the
    "%f" warmup, where every check reprojects: step `step` (a check step) is
    replayed in numpy from the recorded trajectory as
    clip(P (x + t u)) with P the projector of scipy's null_space basis, not
    the oracle's N. A transposed N, a dropped null-space column or a skipped
    projection moves the point by far more than the tolerance (asserted
    below for each)."""
    import scipy.linalg

    S, lb, ub, Wf, _ = c1
    rp, ci, va = O.csr_of(S)
    N = O.nullspace(S)
    W = Wf.shape[1]
    seed = 8
    kw = dict(warm=Wf, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=1, seed=seed, lanes=32)
    _, st, tr = O.sample(S=step + 1, trace_chains=1, trace_thin=1, **kw)
    assert st[0, 0] == 0  # no misses: attempt index == step index
    assert st[0, 1] == st[0, 2] == step // 10 + 1  # every check reprojected
    c0 = Wf.mean(axis=1)
    cT = (W * c0 + tr[0, :step].sum(axis=0)) / (W + step)
    x = c0 if step == 0 else tr[0, step - 1]
    up = O.stream_u01(seed, step, 0, 0, 0, 0)
    w = Wf[:, min(int(W * up), W - 1)]
    u = (w - cT) / np.linalg.norm(w - cT)
    tmin, tmax = _chord(x, u, lb, ub)
    r = O.stream_u01(seed, step, 0, 0, 0, 1)
    y = x + (r * (tmax - tmin) + tmin) * u
    assert np.abs(S @ y).max() > 1e-9  # the check fires on this point
    Ns = scipy.linalg.null_space(S)
    proj = Ns @ (Ns.T @ y)
    assert np.abs(S @ proj).max() < 1e-12  # the projected point is on S x = 0 before the clip
    expect = np.clip(proj, lb, ub)
    tol = 1e-12 * (1 + np.abs(expect).max())
    np.testing.assert_allclose(tr[0, step], expect, rtol=0, atol=tol)
    # each plausible mistake lands 100 times that tolerance or more away
    assert np.abs(np.clip(y, lb, ub) - expect).max() > 100 * tol  # projection skipped
    for j in (0, Ns.shape[1] // 2, Ns.shape[1] - 1):  # one null-space direction dropped
        Nd = np.delete(N, j, axis=1)
        assert np.abs(np.clip(Nd @ (Nd.T @ y), lb, ub) - expect).max() > 100 * tol
    Nt = N.ravel(order="F").reshape(N.shape, order="C")  # column-major buffer read row-major
    assert np.abs(np.clip(Nt @ (Nt.T @ y), lb, ub) - expect).max() > 100 * tol


def test_reprojection_count_and_value_at_check_period_3(c1):
    """Checks every 3 steps: steps 0, 3, 6 reproject, the steps between do not
    (their points are x + t u clipped, off S x = 0 by the step alone)."""
    S, lb, ub, Wf, _ = c1
    rp, ci, va = O.csr_of(S)
    N = O.nullspace(S)
    kw = dict(warm=Wf, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=1, seed=2, lanes=16,
              check_every=3)
    _, st, tr = O.sample(S=8, trace_chains=1, trace_thin=1, **kw)
    assert st[0, 2] == 3 and st[0, 1] == 3
    res = np.abs(S @ tr[0].T).max(axis=0)
    # a clipped projected point keeps a residual only from its clipped coordinates
    for k in range(8):
        on_bound = np.isclose(tr[0, k], lb) | np.isclose(tr[0, k], ub)
        if k % 3 == 0 and not on_bound.any():
            assert res[k] < 1e-12
