"""The batching, reduced-precision and learned layers at the shapes of the
workloads they serve, on the GPU, through the C ABI:

* a batch whose models hold 2,300 to 12,800 reactions, so that the
  1,024-lane persistent kernels with 3 to 13 register slots run (plain,
  with in-graph FP64 projection rounds, and along a direction bank), each
  model bit-equal to the oracle (ACHRcu/ACHRCuda.cu:361-385 is one model per
  process; the oracle restates it per model);
* the matrix-core contractions of the projection at the genome model's
  sizes (K = 5,800 metabolites, 10,600 reactions, null dimension ~4,800)
  against numpy in FP64;
* FP32 and split-BF16x3 projection rounds on the 5,800 x 10,600 model
  against FP64 rounds on the same chains (the reference reprojects with
  FP64 GEMVs, ACHRCuda.cu:208-230, :338-343);
* the learned pipeline with width 512 and rank 512 on that model: unit bank
  columns in the free null space, a sign-symmetric law, and bank-mode chains
  bit-equal to the oracle.
"""
import time

import numpy as np
import pytest
import scipy.sparse as sp

import achr
import oracle as O
from achr import synth
from test_gpu_precision import check_sub

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def gpu_nullspace(rows, cols, vals, m, n):
    mdl = achr.Model.from_coo(n, m, rows, cols, vals, -np.ones(n), np.ones(n))
    mdl.compute_nullspace()
    return mdl.nullspace()


# ---------------------------------------------------------------- batches

def big_case(n, seed, *, csv=False, W=1000):
    t = time.time()
    sm, Nf = synth.batch_model(n, seed, gpu_nullspace)
    warm = synth.chord_warmup(None, sm.x0, sm.lb, sm.ub, W, seed=seed + 1, free=Nf, mix=8)
    if csv:
        warm = synth.round_like_csv(warm)
    model = achr.Model.from_coo(sm.n, sm.m, sm.rows, sm.cols, sm.vals, sm.lb, sm.ub)
    model.set_nullspace(sm.N)
    rp, ci, va = synth.csr(sm)
    say(f"model {sm.m} x {sm.n} (d={sm.N.shape[1]}, free {Nf.shape[1]}) in {time.time() - t:.1f}s")
    return dict(sm=sm, Nf=Nf, lb=sm.lb, ub=sm.ub, N=sm.N, rp=rp, ci=ci, va=va, m=sm.m, model=model, warm=warm)


@pytest.fixture(scope="module")
def big_cases(c1):
    S, lb, ub, Wf, _ = c1
    N = O.nullspace(S)
    rp, ci, va = O.csr_of(S)
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    small = dict(lb=lb, ub=ub, N=N, rp=rp, ci=ci, va=va, m=S.shape[0], model=model, warm=Wf)
    # 3, 7 and 13 slots of 1,024 lanes; the 6,500-reaction model's "%f"
    # warmup makes every check reproject inside the graph
    return [small, big_case(2300, 501), big_case(6500, 502, csv=True), big_case(12800, 503)]


def test_batch_of_genome_sized_models_matches_oracle(big_cases):
    cases = big_cases
    rng = np.random.default_rng(5)
    # the largest model samples along a direction bank of free null-space vectors
    Nf = cases[3]["Nf"]
    bank = Nf @ rng.standard_normal((Nf.shape[1], 400))
    bank /= np.linalg.norm(bank, axis=0)
    smps = []
    for k, c in enumerate(cases):
        smp = achr.Sampler(c["model"], c["warm"], achr.Options(model_id=k + 1))
        if k == 3:
            smp.set_directions(bank)
            smp.use_directions(True)
        smps.append(smp)
    batch = achr.Batch(smps)
    lanes = batch.lanes().tolist()
    say(f"lanes {lanes}, modes {[s.schedule()[1] for s in smps]}")
    assert lanes[1:] == [1024, 1024, 1024]
    P, S, seed, file = [19, 5, 3, 4], 40, 31, 1
    batch.run(P, S, seed=seed, file=file)
    say(f"timing {batch.timing()}")
    got = [(s.points(), s.stats()) for s in smps]
    for k, c in enumerate(cases):
        say(f"oracle model {k}: n={c['lb'].size} P={P[k]} S={S} lanes={lanes[k]}")
        o_pts, o_st, _ = O.sample(warm=c["warm"], lb=c["lb"], ub=c["ub"], rowptr=c["rp"], colidx=c["ci"],
                                  vals=c["va"], N=c["N"], m=c["m"], P=P[k], S=S, seed=seed, file=file, model=k + 1,
                                  lanes=lanes[k], threads=8, bank=bank if k == 3 else None,
                                  c0=smps[k].start_point())
        np.testing.assert_array_equal(got[k][1], o_st)
        np.testing.assert_array_equal(got[k][0], o_pts)
    assert (got[2][1][:, 1] == got[2][1][:, 2]).all() and got[2][1][0, 1] == 4  # every check reprojected
    assert got[1][1][:, 1].sum() == 0 and got[3][1][:, 1].sum() == 0  # exact warmups / bank: none did
    # the same models each on their own: the same samples
    for k, s in enumerate(smps):
        s.run(P[k], S, seed=seed, file=file)
        np.testing.assert_array_equal(s.points(), got[k][0])


# ------------------------------------------------------ genome-scale model

@pytest.fixture(scope="module")
def genome():
    """The 5,800 x 10,600 model (3.2 nonzeros per column, 40% irreversible,
    212 fixed), its free null space, and 2,000 chord warmup points stored
    with "%f" precision."""
    t = time.time()
    sm = synth.config2(gpu_nullspace, seed=2)
    fixed = np.nonzero(sm.lb == sm.ub)[0]
    r = np.concatenate([sm.rows, sm.m + np.arange(len(fixed))]).astype(np.int32)
    c = np.concatenate([sm.cols, fixed]).astype(np.int32)
    v = np.concatenate([sm.vals, np.ones(len(fixed))])
    Nf = gpu_nullspace(r, c, v, sm.m + len(fixed), sm.n)
    W = synth.round_like_csv(synth.chord_warmup(None, sm.x0, sm.lb, sm.ub, 2000, seed=102, free=Nf, mix=8))
    model = achr.Model.from_coo(sm.n, sm.m, sm.rows, sm.cols, sm.vals, sm.lb, sm.ub)
    model.set_nullspace(sm.N)
    Ssp = sp.csr_matrix((sm.vals, (sm.rows, sm.cols)), shape=(sm.m, sm.n))
    say(f"genome model {sm.m} x {sm.n}, d={sm.N.shape[1]}, free {Nf.shape[1]}, in {time.time() - t:.1f}s")
    return dict(sm=sm, Nf=Nf, W=W, model=model, Ssp=Ssp)


# Relative Frobenius error bounds, as in the small-shape GEMM test: FP32 MFMA
# and split-BF16x3 ~ sqrt(K) 2^-24 with FP32 accumulation over K.
GEMM_TOL = {0: 2e-6, 1: 5e-6}


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("form", ["residual", "null"])
def test_projection_gemm_at_genome_sizes(prec, form):
    """The two contractions of a projection round at the genome model's
    sizes, in the operand layouts the rounds use: residual rows (P x 5,800)
    times S+ (10,600 reactions contiguous per metabolite), and chains
    (P x 10,600) times N (column major, K contiguous)."""
    rng = np.random.default_rng(7 + prec)
    M = 256
    if form == "residual":
        K, N = 5800, 10600
        A = rng.standard_normal((M, K)).astype(np.float32)
        B = (rng.standard_normal((K, N)) * 1e-2).astype(np.float32)
        say(f"gemm {form} prec={prec} {M}x{N}x{K}")
        got = achr.gemm_f32(A, B, prec=prec)
        ref = A.astype(np.float64) @ B.astype(np.float64)
    else:
        K, N = 10600, 4817
        A = rng.standard_normal((M, K)).astype(np.float32)
        Bt = (rng.standard_normal((N, K)) * 1e-2).astype(np.float32)  # N^T rows: reactions contiguous
        say(f"gemm {form} prec={prec} {M}x{N}x{K}")
        got = achr.gemm_f32(A, Bt, prec=prec, transB=True)
        ref = A.astype(np.float64) @ Bt.T.astype(np.float64)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    say(f"relative error {err:.3g}")
    assert err < GEMM_TOL[prec]


@pytest.mark.parametrize("prec", [0, 1])
def test_residual_round_kernel_at_genome_sizes(prec):
    """The kernel of the residual rounds (packed S+, 256 x 256 tiles) at the
    genome model's sizes: 700 of 1,000 residual rows, gathered in random
    order, each updated as C[r] -= A[r] B with K = 5,800 and N = 10,600
    (neither a multiple of the tile), against the FP64 product of the same
    FP32 operands; the other rows are left bit for bit."""
    rng = np.random.default_rng(17 + prec)
    rows, M, K, N = 1000, 700, 5800, 10600
    A = rng.standard_normal((rows, K)).astype(np.float32)
    B = (rng.standard_normal((K, N)) * 1e-2).astype(np.float32)
    C0 = rng.standard_normal((rows, N))
    idx = rng.permutation(rows)[:M].astype(np.int32)
    say(f"gemm_sub prec={prec} {M} of {rows} rows, {N}x{K}")
    got = achr.gemm_sub_f64(A, B, C0, idx=idx, prec=prec)
    check_sub(A[idx], B, C0[idx] - got[idx], prec)
    rest = np.setdiff1d(np.arange(rows), idx)
    assert np.array_equal(got[rest], C0[rest])


def test_reduced_precision_rounds_on_genome_model(genome):
    """FP32 and split-BF16x3 residual rounds (FP64 every 5th round) against
    FP64 rounds on the same 512 chains and seed, 101 steps, every check
    reprojecting ("%f" warmup): the last check is the last step, so the final
    points are fresh from a projection; their drift stays within 10x FP64's
    and the chains agree with the FP64 run to a small fraction of the range."""
    sm, W, model, Ssp = genome["sm"], genome["W"], genome["model"], genome["Ssp"]
    rng = sm.ub - sm.lb
    live = rng > 0
    P, S = 512, 101
    res = {}
    for name, prec in (("fp64", achr.PROJ_FP64), ("res_f32", achr.PROJ_RES_F32),
                       ("res_bf16x3", achr.PROJ_RES_BF16X3)):
        say(f"{name}: sampler + {P} x {S}")
        smp = achr.Sampler(model, W, achr.Options(reproject_mode=2, proj_prec=prec, fp64_every=5))
        pts = smp.sample(P, S, seed=4)
        st = smp.stats()
        res[name] = (pts, st, float(np.abs(Ssp @ pts).max()))
        say(f"{name}: drift {res[name][2]:.3g}, reprojections {int(st[:, 1].sum())}, timing {smp.timing()}")
        del smp
    ref, st_ref, drift_ref = res["fp64"]
    assert (st_ref[:, 1] == st_ref[:, 2]).all() and st_ref[0, 2] == 11
    for name in ("res_f32", "res_bf16x3"):
        pts, st, drift = res[name]
        assert (st[:, 2] == st_ref[:, 2]).all()
        assert drift < max(10 * drift_ref, 1e-9), (name, drift, drift_ref)
        dpt = (np.abs(pts - ref).max(axis=1)[live] / rng[live]).max()
        dm = (np.abs(pts.mean(1) - ref.mean(1))[live] / rng[live]).max()
        ds = (np.abs(pts.std(1) - ref.std(1))[live] / rng[live]).max()
        say(f"{name} vs fp64: point {dpt:.3g}, mean {dm:.3g}, std {ds:.3g} of range")
        assert dm < 1e-3 and ds < 1e-3 and dpt < 1e-2


def test_learned_pipeline_width_512_on_genome_model(genome):
    """learn_directions with the config's network (width 512) and a rank-512
    principal subspace, trained on a traced plain run of the genome model:
    the loss falls, every bank column has unit norm, lies in null(S) and is
    zero on the fixed reactions, the bank's law is sign-symmetric (its mean
    direction is as small as that of K independent symmetric unit vectors),
    and bank-mode chains are bit-equal to the oracle."""
    sm, W, model, Ssp = genome["sm"], genome["W"], genome["model"], genome["Ssp"]
    P = 1024
    smp = achr.Sampler(model, W, achr.Options(trace_chains=P, trace_thin=100, reproject_mode=2,
                                              proj_prec=achr.PROJ_RES_BF16X3))
    say("training run")
    smp.run(P, 600, seed=99)
    smp.timing()
    K = 4096
    say("learn")
    rep = smp.learn_directions(rank=512, width=512, iters=200, batch=2048, lr=1e-3, bank_size=K, seed=7)
    say(f"learn report {rep}")
    assert rep["samples"] >= 512 and rep["loss_end"] < 0.5 * rep["loss_start"]
    bank = smp.directions()
    assert bank.shape == (sm.n, K)
    assert np.abs(np.linalg.norm(bank, axis=0) - 1.0).max() < 1e-12
    assert np.abs(Ssp @ bank).max() < 1e-11
    fixed = sm.lb == sm.ub
    assert np.abs(bank[fixed]).max() < 1e-12
    mean_dir = np.linalg.norm(bank.mean(axis=1))
    say(f"|mean direction| {mean_dir:.4g} (1/sqrt(K) = {1 / np.sqrt(K):.4g})")
    assert mean_dir < 3.0 / np.sqrt(K)
    # bank-mode chains (1,024 lanes, 11 slots) against the oracle
    smp.set_run_options(0, 0, 100)
    Pb, Sb = 6, 150
    got = smp.sample(Pb, Sb, seed=12)
    st = smp.stats()
    rp, ci, va = synth.csr(sm)
    say("oracle bank run")
    o_pts, o_st, _ = O.sample(warm=W, lb=sm.lb, ub=sm.ub, rowptr=rp, colidx=ci, vals=va, N=sm.N, m=sm.m, P=Pb,
                              S=Sb, seed=12, lanes=smp.schedule()[0], check_every=100, threads=6, bank=bank,
                              c0=smp.start_point())
    np.testing.assert_array_equal(st, o_st)
    np.testing.assert_array_equal(got, o_pts)
    assert st[:, 1].sum() == 0 and np.abs(Ssp @ got).max() < 1e-9
