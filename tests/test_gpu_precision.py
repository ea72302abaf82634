"""Matrix-core contractions and the reduced-precision projection rounds.

The GEMM is checked against numpy in FP64 for every layout the library uses
(either operand transposed), with tolerances stated per precision. The
projection modes are checked against the FP64 mode on the same chains:
they are the same operator in exact arithmetic, so drift |S x| and the
chains' moments must agree."""
import numpy as np
import pytest

import achr
import oracle as O

pytestmark = pytest.mark.gpu

# Relative Frobenius error bounds: FP32 MFMA ~ K * 2^-24; split-BF16x3 drops the
# lo*lo term (~2^-16 * 2^-16 relative) plus FP32 accumulation; BF16 ~ 2^-8.
TOL = {0: 2e-6, 1: 5e-6, 2: 1.5e-2}


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("shape", [(130, 200, 70), (256, 512, 512), (37, 64, 33), (300, 129, 1000)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_matches_numpy(prec, shape, ta, tb):
    M, N, K = shape
    print(f"gemm prec={prec} M={M} N={N} K={K} ta={ta} tb={tb}", flush=True)
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    got = achr.gemm_f32(A.T.copy() if ta else A, B.T.copy() if tb else B, prec=prec, alpha=0.5, C=C0, beta=-2.0,
                        transA=ta, transB=tb)
    ref = 0.5 * (A.astype(np.float64) @ B.astype(np.float64)) - 2.0 * C0
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < TOL[prec], err


# Bounds for the residual rounds' kernel. Elementwise, guaranteed: FP32
# accumulation of K products (3K for the three split passes) costs at most
# about (terms + 2) 2^-24 of sum_k |a_k b_k|; the split a = hi + bf16(a - hi)
# represents each operand to 2^-16 of itself and the dropped lo * lo term is
# below 2^-16 |a b|, so the split adds 3 2^-16 of the same sum. A wrong
# pairing of k, a dropped pass or a transposed operand misses by O(1).
# Frobenius, statistical (Gaussian data): about 5e-6 for the split (its
# representation error does not average out over k), ~1e-7 for FP32.
SUB_TOL = {0: 2e-6, 1: 1e-5}


def check_sub(A, B, upd, prec):
    """upd (the FP64 change of the rows) against A @ B in FP64."""
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = A64 @ B64
    K = A.shape[1]
    mag = np.abs(A64) @ np.abs(B64)
    terms = K if prec == 0 else 3 * K
    bound = ((terms + 2) * 2.0 ** -24 + (0 if prec == 0 else 3 * 2.0 ** -16)) * mag + 1e-300
    worst = float((np.abs(upd - ref) / bound).max())
    err = np.linalg.norm(upd - ref) / np.linalg.norm(ref)
    print(f"relative error {err:.3g}, worst elementwise {worst:.3g} of its bound", flush=True)
    assert worst <= 1.0 and err < SUB_TOL[prec], (worst, err)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("shape", [(300, 300, 37, 1), (300, 257, 600, 100), (1000, 700, 513, 33),
                                   (40, 3, 256, 16), (513, 512, 1, 2000)])
def test_residual_round_kernel_matches_numpy(prec, shape):
    """The residual rounds' kernel C[r] -= A[r] B on packed B: ragged tiles
    in every dimension (K from 1, N from 1, partial row tiles), rows gathered
    in random order when M < rows, the rest left bit for bit."""
    rows, M, N, K = shape
    print(f"gemm_sub prec={prec} rows={rows} M={M} N={N} K={K}", flush=True)
    rng = np.random.default_rng(rows + M + N + K)
    A = rng.standard_normal((rows, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    C0 = rng.standard_normal((rows, N))
    idx = None if M == rows else rng.permutation(rows)[:M].astype(np.int32)
    got = achr.gemm_sub_f64(A, B, C0, idx=idx, prec=prec)
    sel = np.arange(rows) if idx is None else idx
    check_sub(A[sel], B, C0[sel] - got[sel], prec)
    rest = np.setdiff1d(np.arange(rows), sel)
    assert np.array_equal(got[rest], C0[rest])


@pytest.fixture(scope="module")
def c1_models(c1):
    S, lb, ub, Wf, We = c1
    N = O.nullspace(S)
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    return S, lb, ub, Wf, model


def _run(model, W, prec, P=2000, S=101, fp64_every=1000, seed=3):
    print(f"projection prec={prec} P={P} S={S}", flush=True)
    smp = achr.Sampler(model, W, achr.Options(reproject_mode=2, proj_prec=prec, fp64_every=fp64_every, lanes=32))
    pts = smp.sample(P, S, seed=seed)
    return pts, smp.stats()


def test_residual_projection_matches_fp64_on_c1(c1_models):
    """The check at s = 100 is the last step, so the final points are fresh
    from a projection (then clipped): drift of the residual forms must be as
    small as FP64's, and the chains' moments must agree."""
    S, lb, ub, Wf, model = c1_models
    ref, st_ref = _run(model, Wf, achr.PROJ_FP64)
    drift_ref = np.abs(S @ ref).max()
    for prec in (achr.PROJ_RES_BF16X3, achr.PROJ_RES_F32):
        pts, st = _run(model, Wf, prec)
        assert (st[:, 1] == st_ref[:, 1]).all()  # same checks, all flagged
        drift = np.abs(S @ pts).max()
        assert drift < max(10 * drift_ref, 1e-9), (prec, drift, drift_ref)
        rng = ub - lb
        live = rng > 0
        assert np.max(np.abs(pts.mean(1) - ref.mean(1))[live] / rng[live]) < 0.01
        assert np.max(np.abs(pts.std(1) - ref.std(1))[live] / rng[live]) < 0.01


def test_null_form_low_precision_reports_its_drift(c1_models):
    """N (N^T x) with x rounded to FP32 loses |x| * 2^-24 per coordinate; the
    residual form does not. Both run; the drift of each is what the bench reports."""
    S, lb, ub, Wf, model = c1_models
    pts_null, _ = _run(model, Wf, achr.PROJ_NULL_F32)
    pts_res, _ = _run(model, Wf, achr.PROJ_RES_F32)
    d_null = np.abs(S @ pts_null).max()
    d_res = np.abs(S @ pts_res).max()
    assert np.isfinite(d_null) and d_null < 1e-3
    assert d_res <= d_null * 10


def test_periodic_fp64_round_on_medium_model():
    """A 330 x 600 model, residual BF16x3 rounds with an FP64 round every 5th."""
    from achr import synth
    import scipy.linalg
    m = synth.make_model(330, 600, 17, irrev_frac=0.3, n_fixed=6, width=(5.0, 50.0))
    N = scipy.linalg.null_space(m.dense())
    W = synth.round_like_csv(synth.chord_warmup(N, m.x0, m.lb, m.ub, 400, seed=3))
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(N)
    Sd = m.dense()
    print("medium model fp64 vs bf16x3", flush=True)
    a = achr.Sampler(model, W, achr.Options(reproject_mode=2, lanes=64)).sample(512, 151, seed=8)
    b = achr.Sampler(model, W, achr.Options(reproject_mode=2, lanes=64, proj_prec=achr.PROJ_RES_BF16X3,
                                            fp64_every=5)).sample(512, 151, seed=8)
    assert np.abs(Sd @ b).max() < max(10 * np.abs(Sd @ a).max(), 1e-9)
    rng = m.ub - m.lb
    live = rng > 0
    assert np.max(np.abs(a.mean(1) - b.mean(1))[live] / rng[live]) < 0.02


def test_pseudo_inverse_routes_agree():
    """S+ comes from a Cholesky factorisation of S^T S + sigma N N^T when N
    spans null(S), else from the eigen-decomposition of S S^T. A basis with
    one column zeroed leaves the first route singular along that direction,
    so the sampler falls back to the second. Residual-form FP32 rounds from
    the two (no FP64 rounds, so S+ is the only difference) give the same
    chains to the rounding of S+, and both keep |S x| at FP64's level."""
    from achr import synth
    import scipy.linalg
    m = synth.make_model(330, 600, 21, irrev_frac=0.3, n_fixed=6, width=(5.0, 50.0))
    Sd = m.dense()
    N = scipy.linalg.null_space(Sd)
    W = synth.round_like_csv(synth.chord_warmup(N, m.x0, m.lb, m.ub, 400, seed=5))
    Nz = N.copy()
    Nz[:, 0] = 0.0
    pts = []
    for name, basis in (("cholesky route", N), ("eigen route", Nz)):
        print(f"330 x 600, res_f32 rounds, {name}", flush=True)
        model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
        model.set_nullspace(basis)
        smp = achr.Sampler(model, W, achr.Options(reproject_mode=2, lanes=64, proj_prec=achr.PROJ_RES_F32,
                                                  fp64_every=1 << 30))
        pts.append(smp.sample(256, 41, seed=9))
        assert (smp.stats()[:, 1] > 0).all()
        del smp
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(N)
    ref = achr.Sampler(model, W, achr.Options(reproject_mode=2, lanes=64)).sample(256, 41, seed=9)
    rng = m.ub - m.lb
    live = rng > 0
    d = (np.abs(pts[0] - pts[1]).max(axis=1)[live] / rng[live]).max()
    print(f"routes differ by {d:.3g} of range", flush=True)
    assert d < 1e-8
    drift_ref = np.abs(Sd @ ref).max()
    for p in pts:
        assert np.abs(Sd @ p).max() < max(10 * drift_ref, 1e-9)
