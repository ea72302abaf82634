"""GPU parity: the HIP sampler, called through the C ABI, against the CPU oracle
on the same inputs and Philox stream. Integer/step structure is compared
exactly and the FP64 trajectories bit for bit (the kernels and the oracle
share one arithmetic contract: the same fused multiply-adds and the same
reduction tree for the chain's lane count)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import achr
import oracle as O
from conftest import golden

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c1_setup(c1):
    S, lb, ub, Wf, We = c1
    N = O.nullspace(S)
    rp, ci, va = O.csr_of(S)
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    return dict(S=S, lb=lb, ub=ub, Wf=Wf, We=We, N=N, rp=rp, ci=ci, va=va, model=model)


def run_both(st, warm, *, lanes, mode, P, S, seed, file=0, chain0=0, trace=0, thin=0, threads=8):
    say(f"case lanes={lanes} mode={mode} P={P} S={S} seed={seed} file={file} chain0={chain0}")
    opts = achr.Options(lanes=lanes, reproject_mode=mode, trace_chains=trace, trace_thin=thin)
    smp = achr.Sampler(st["model"], warm, opts)
    assert smp.schedule() == (lanes, mode)
    smp.run(P, S, seed=seed, file=file, chain0=chain0)
    g_pts, g_st = smp.points(), smp.stats()
    g_tr = smp.trace() if trace else None
    o_pts, o_st, o_tr = O.sample(warm=warm, lb=st["lb"], ub=st["ub"], rowptr=st["rp"], colidx=st["ci"],
                                 vals=st["va"], N=st["N"], m=st["S"].shape[0], P=P, S=S, seed=seed, file=file,
                                 chain0=chain0, lanes=lanes, threads=threads, trace_chains=trace, trace_thin=thin)
    return g_pts, g_st, g_tr, o_pts, o_st, o_tr


@pytest.mark.parametrize("lanes", [16, 32, 64, 128, 256, 1024])
@pytest.mark.parametrize("which", ["csv6", "exact"])
def test_in_kernel_projection_bitwise(c1_setup, lanes, which):
    warm = c1_setup["Wf"] if which == "csv6" else c1_setup["We"]
    g, gs, _, o, os_, _ = run_both(c1_setup, warm, lanes=lanes, mode=1, P=130, S=250, seed=2024)
    np.testing.assert_array_equal(gs, os_)
    np.testing.assert_array_equal(g, o)
    if which == "csv6":
        assert (gs[:, 1] == 25).all()  # every check reprojects
    else:
        assert (gs[:, 1] == 0).all()


@pytest.mark.parametrize("lanes", [32, 256])
def test_batched_projection_bitwise(c1_setup, lanes):
    g, gs, _, o, os_, _ = run_both(c1_setup, c1_setup["Wf"], lanes=lanes, mode=2, P=200, S=123, seed=5)
    np.testing.assert_array_equal(gs, os_)
    np.testing.assert_array_equal(g, o)


@pytest.mark.parametrize("lanes,mode", [(32, 1), (1024, 1), (256, 2)])
def test_edge_sizes_one_chain_ragged_and_zero_steps(c1_setup, lanes, mode):
    """One chain, chain counts that fill no whole block, and zero steps. The
    reference starts each chain at the warmup centre (ACHRCuda.cu:374-377) and
    writes its point only inside the step loop (:322, :350) into a buffer it
    clears before every file (:739), so with stepsPerPoint 0 its File<k> holds
    zeros, which the ACHRCuda command line here reproduces
    (test_cli_zero_steps_writes_zeros). The library's achr_sample returns a
    chain's start point, the centre, after zero steps. A run of no chains is
    refused."""
    for P, S in [(1, 150), (3, 57), (5, 0)]:
        g, gs, _, o, os_, _ = run_both(c1_setup, c1_setup["Wf"], lanes=lanes, mode=mode, P=P, S=S, seed=31)
        np.testing.assert_array_equal(gs, os_)
        np.testing.assert_array_equal(g, o)
        assert g.shape == (95, P)
    smp = achr.Sampler(c1_setup["model"], c1_setup["Wf"], achr.Options(lanes=lanes, reproject_mode=mode))
    np.testing.assert_array_equal(g, np.repeat(smp.center()[:, None], 5, axis=1))
    assert (gs == 0).all()
    for P, S in [(0, 10), (-1, 10), (4, -1)]:
        with pytest.raises(achr.AchrError) as e:
            smp.run(P, S, seed=1)
        assert e.value.code == -1


def test_trace_file_and_chain_offset(c1_setup):
    g, gs, gt, o, os_, ot = run_both(c1_setup, c1_setup["Wf"], lanes=64, mode=1, P=40, S=100, seed=9, file=3,
                                     chain0=1000, trace=5, thin=10)
    np.testing.assert_array_equal(g, o)
    np.testing.assert_array_equal(gt, ot[:5, :10])
    assert np.all(g >= c1_setup["lb"][:, None]) and np.all(g <= c1_setup["ub"][:, None])


def test_medium_model_batched_gemm_bitwise():
    """A 330 x 600 model (d ~ 270): the batched FP64 GEMM path with chains
    flagged every round, against the oracle's sequential dot products."""
    from achr import synth
    import scipy.linalg
    m = synth.make_model(330, 600, 17, irrev_frac=0.3, n_fixed=6, width=(5.0, 50.0))
    N = scipy.linalg.null_space(m.dense())
    W = synth.round_like_csv(synth.chord_warmup(N, m.x0, m.lb, m.ub, 400, seed=3))
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(N)
    rp, ci, va = O.csr_of(m.dense())
    st = dict(S=m.dense(), lb=m.lb, ub=m.ub, N=N, rp=rp, ci=ci, va=va, model=model)
    for lanes in (64, 256):
        g, gs, _, o, os_, _ = run_both(st, W, lanes=lanes, mode=2, P=96, S=60, seed=31)
        np.testing.assert_array_equal(gs, os_)
        assert gs[:, 1].sum() > 0
        np.testing.assert_array_equal(g, o)


@pytest.mark.parametrize("P,exact", [(600, False), (1300, False), (1300, True)])
def test_projection_round_widths_bitwise(P, exact):
    """Batched FP64 rounds pick their tile width on the device from the
    round's count of flagged chains: 128 x 32 tiles up to 1,024 chains,
    128 x 128 above (runs of more than 256 chains). Every check flags with the
    "%f" warmup (600 and 1,300 chains per round), few do with exact points;
    each output keeps the sequential FMA order, so both widths equal the
    oracle bit for bit."""
    from achr import synth
    import scipy.linalg
    m = synth.make_model(330, 600, 17, irrev_frac=0.3, n_fixed=6, width=(5.0, 50.0))
    N = scipy.linalg.null_space(m.dense())
    W = synth.chord_warmup(N, m.x0, m.lb, m.ub, 400, seed=3)
    if not exact:
        W = synth.round_like_csv(W)
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(N)
    rp, ci, va = O.csr_of(m.dense())
    st = dict(S=m.dense(), lb=m.lb, ub=m.ub, N=N, rp=rp, ci=ci, va=va, model=model)
    g, gs, _, o, os_, _ = run_both(st, W, lanes=64, mode=2, P=P, S=31, seed=7, threads=8)
    np.testing.assert_array_equal(gs, os_)
    np.testing.assert_array_equal(g, o)
    flagged = int(gs[:, 1].sum())
    print(f"P {P} exact {exact}: {flagged} reprojections of {int(gs[:, 2].sum())} checks", flush=True)
    assert (flagged < gs[:, 2].sum()) if exact else (flagged == gs[:, 2].sum())


def test_gpu_nullspace_matches_scipy(c1_setup):
    import scipy.linalg
    m = achr.Model.read_mps(golden("c1.mps"))
    d = m.compute_nullspace()
    N = m.nullspace()
    N2 = scipy.linalg.null_space(c1_setup["S"])
    assert d == N2.shape[1] == 23
    assert np.abs(c1_setup["S"] @ N).max() < 1e-12
    assert np.abs(N.T @ N - np.eye(d)).max() < 1e-12
    assert np.abs(N @ N.T - N2 @ N2.T).max() < 1e-12


def test_gpu_nullspace_rank_deficient():
    rng = np.random.default_rng(0)
    A = rng.integers(-2, 3, size=(20, 60)).astype(float)
    S = np.vstack([A, A[0] + A[1], A[2] - A[3]])
    m = achr.Model.from_dense(S, -np.ones(60), np.ones(60))
    d = m.compute_nullspace()
    N = m.nullspace()
    assert d == 60 - np.linalg.matrix_rank(S) == 40
    assert np.abs(S @ N).max() < 1e-10


def test_box_uniform_moments_gpu():
    lb = np.array([0.0, -1.0, 2.0, -5.0])
    ub = np.array([1.0, 3.0, 2.5, 5.0])
    rng = np.random.default_rng(5)
    # 400 warmup points keep the chain's own points from dominating the
    # adaptive centre (ACHR's bias on few warmup points).
    Wm = lb[:, None] + (ub - lb)[:, None] * rng.uniform(0, 1, (4, 400))
    model = achr.Model.from_coo(4, 0, [], [], [], lb, ub)
    model.set_nullspace(np.eye(4))
    pts = achr.Sampler(model, Wm, achr.Options(lanes=16)).sample(20000, 200, seed=11)
    assert np.all(np.abs(pts.mean(axis=1) - (lb + ub) / 2) < 0.01 * (ub - lb))
    np.testing.assert_allclose(pts.std(axis=1), (ub - lb) / np.sqrt(12), rtol=0.02)


def test_miss_cap_forced_from_host():
    """Warmup points all equal to their centre: every direction is zero; each
    chain stops after max_misses and is flagged (the reference would spin)."""
    lb, ub = -np.ones(3), np.ones(3)
    model = achr.Model.from_coo(3, 0, [], [], [], lb, ub)
    model.set_nullspace(np.eye(3))
    smp = achr.Sampler(model, np.zeros((3, 4)), achr.Options(lanes=16, max_misses=50))
    pts = smp.sample(8, 5, seed=0)
    st = smp.stats()
    assert (st[:, 3] == 1).all() and (st[:, 0] == 51).all()
    np.testing.assert_array_equal(pts, 0.0)


def test_cli_drop_in(tmp_path, c1_setup):
    """ACHRCuda model.mps warmup.csv nFiles nPoints nSteps writes File0..: n
    lines x P comma-separated %f values, equal to the oracle's points; with
    --lossless also File<k>.bin, bit-equal to them."""
    npath = tmp_path / "N.bin"
    np.asfortranarray(c1_setup["N"]).ravel(order="F").tofile(npath)
    env = dict(os.environ, ACHR_SEED="77", ACHR_NULLSPACE=str(npath))
    say("cli run")
    r = subprocess.run([achr.CLI_PATH, golden("c1.mps"), golden("c1_warmup.csv"), "2", "50", "80", "--lossless"],
                       cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for f in range(2):
        lines = (tmp_path / f"File{f}").read_text().splitlines()
        assert len(lines) == 95 and all(len(l.split(",")) == 50 for l in lines)
        o, _, _ = O.sample(warm=c1_setup["Wf"], lb=c1_setup["lb"], ub=c1_setup["ub"], rowptr=c1_setup["rp"],
                           colidx=c1_setup["ci"], vals=c1_setup["va"], N=c1_setup["N"], m=72, P=50, S=80, seed=77,
                           file=f, lanes=32)
        expect = [",".join("%f" % v for v in o[i]) for i in range(95)]
        assert lines == expect
        np.testing.assert_array_equal(achr.read_points_bin(tmp_path / f"File{f}.bin"), o)
    r = subprocess.run([achr.CLI_PATH, golden("c1.mps"), golden("c1_warmup.csv"), "1", "5", "5", "--lossy"],
                       cwd=tmp_path, env=env, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Usage" in r.stderr
    r = subprocess.run([achr.CLI_PATH, "missing.mps", "w.csv", "1", "1", "1"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0 and "MPS" in r.stderr


def test_cli_zero_steps_writes_zeros(tmp_path, c1_setup):
    """ACHRCuda m.mps w.csv 2 7 0: the reference clears its points buffer
    before each file (ACHRCuda.cu:739) and its step loop never runs, so each
    File<k> is n lines of P zeros in "%f"."""
    npath = tmp_path / "N.bin"
    np.asfortranarray(c1_setup["N"]).ravel(order="F").tofile(npath)
    env = dict(os.environ, ACHR_NULLSPACE=str(npath))
    say("cli run, zero steps")
    r = subprocess.run([achr.CLI_PATH, golden("c1.mps"), golden("c1_warmup.csv"), "2", "7", "0", "--lossless"],
                       cwd=tmp_path, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for f in range(2):
        assert (tmp_path / f"File{f}").read_text() == ("0.000000," * 6 + "0.000000\n") * 95
        np.testing.assert_array_equal(achr.read_points_bin(tmp_path / f"File{f}.bin"), np.zeros((95, 7)))


@pytest.mark.parametrize("mets,rxns,lanes", [(4900, 9000, 1024), (3800, 7000, 512)])
def test_genome_scale_schedule_bitwise(mets, rxns, lanes):
    """Genome-scale sizes: 9 slots per thread at 1,024 lanes (every slot
    count has a kernel there) or 14 padded to 16 at 512 lanes, with the
    centre in LDS (72 KiB / 64 KiB of dynamic LDS), batched FP64 projection
    on a null space from the GPU solver."""
    from achr import synth

    def ns(rows, cols, vals, m, n):
        mdl = achr.Model.from_coo(n, m, rows, cols, vals, -np.ones(n), np.ones(n))
        mdl.compute_nullspace()
        return mdl.nullspace()

    m = synth.make_model(mets, rxns, 23, irrev_frac=0.4, n_fixed=rxns // 100, width=(100.0, 1000.0), x0_scale=100.0,
                         nullspace_fn=ns)
    W = synth.round_like_csv(synth.chord_warmup(m.N, m.x0, m.lb, m.ub, 300, seed=5))
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(m.N)
    import scipy.sparse as sp
    Sd = sp.csr_matrix((m.vals, (m.rows, m.cols)), shape=(m.m, m.n))
    Sd.sort_indices()
    st = dict(S=np.zeros((m.m, 1)), lb=m.lb, ub=m.ub, N=m.N, rp=Sd.indptr.astype(np.int32),
              ci=Sd.indices.astype(np.int32), va=Sd.data, model=model)
    g, gs, _, o, os_, _ = run_both(st, W, lanes=lanes, mode=2, P=8, S=25, seed=3)
    np.testing.assert_array_equal(gs, os_)
    np.testing.assert_array_equal(g, o)
    assert np.abs(Sd @ g).max() < 1e-3
