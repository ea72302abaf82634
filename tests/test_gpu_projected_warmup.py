"""Plain ACHR from a warmup set moved onto the sampled set {S x = 0, fixed
reactions at their bound} (achr_sampler_project_warmup). The reference
samples from its "%f" warmup file as read (ACHRCuda.cu:676-689), so every
check reprojects (:325-343); from projected points only the clip to the
bounds (:193-206) can leave S x = 0, and with batched FP64 projection rounds
a chain whose check does not flag runs on instead of stopping. The moved
points are checked against a least-squares projection in numpy, and the
chains against the oracle bit for bit, in-kernel and batched."""
import numpy as np
import pytest

import achr
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_projected_warmup_matches_numpy_and_samples_bitwise(c1):
    S, lb, ub, Wf, _ = c1
    N = O.nullspace(S)
    rp, ci, va = O.csr_of(S)
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    fixed = lb == ub
    A = np.vstack([S, np.eye(S.shape[1])[fixed]])
    b = np.concatenate([np.zeros(S.shape[0]), lb[fixed]])
    want = Wf - np.linalg.lstsq(A, A @ Wf - b[:, None], rcond=None)[0]
    runs = {}
    for mode in (1, 2):
        smp = achr.Sampler(model, Wf, achr.Options(lanes=32, reproject_mode=mode))
        smp.project_warmup()
        Wp = smp.warmup()
        assert Wp.shape == Wf.shape
        assert np.abs(Wp - want).max() < 1e-9 * (1 + np.abs(want).max())
        assert np.abs(S @ Wp).max() < 1e-11
        np.testing.assert_array_equal(Wp[fixed], np.repeat(lb[fixed][:, None], Wf.shape[1], axis=1))
        np.testing.assert_allclose(smp.center(), Wp.mean(axis=1), rtol=0, atol=1e-13)
        say(f"mode {mode}: sampling from the projected warmup")
        pts = smp.sample(64, 300, seed=14)
        st = smp.stats()
        o, ost, _ = O.sample(warm=Wp, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=64, S=300,
                             seed=14, lanes=32, c0=smp.center(), threads=8)
        np.testing.assert_array_equal(st, ost)
        np.testing.assert_array_equal(pts, o)
        runs[mode] = (pts, st, smp.timing())
    np.testing.assert_array_equal(runs[1][0], runs[2][0])
    st = runs[2][1]
    say(f"reprojections {int(st[:, 1].sum())} of {int(st[:, 2].sum())} checks; timing {runs[2][2]}")
    assert st[:, 1].sum() < st[:, 2].sum()  # the "%f" warmup reprojects at every check


def test_batched_rounds_run_unflagged_chains_on(c1):
    """Batched FP64 rounds where only some checks flag: a chain stops only at a
    check that flags it and resumes after the round; the points equal the
    in-kernel projection's and the oracle's, on the "%f" warmup (every check
    flags) and on the exact one (none does) in one sampler each."""
    S, lb, ub, Wf, We = c1
    N = O.nullspace(S)
    rp, ci, va = O.csr_of(S)
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    for name, W in (("%f", Wf), ("exact", We)):
        g = {}
        for mode in (1, 2):
            smp = achr.Sampler(model, W, achr.Options(lanes=64, reproject_mode=mode, trace_chains=4, trace_thin=9))
            g[mode] = (smp.sample(50, 95, seed=3), smp.stats(), smp.trace())
        o, ost, otr = O.sample(warm=W, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=50, S=95,
                               seed=3, lanes=64, trace_chains=4, trace_thin=9, threads=8)
        for mode in (1, 2):
            say(f"{name} warmup, mode {mode}")
            np.testing.assert_array_equal(g[mode][1], ost)
            np.testing.assert_array_equal(g[mode][0], o)
            np.testing.assert_array_equal(g[mode][2], otr[:4])
