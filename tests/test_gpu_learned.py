"""Learned direction bank: directions stay in the free null space (no
drift), the sampler targets the same uniform law as isotropic hit-and-run,
and the trained rounding map mixes faster than isotropic directions (the
method without the network) on a polytope with long internal-cycle
directions."""
import numpy as np
import pytest

import achr
from achr import stats as st
from achr import synth

pytestmark = pytest.mark.gpu


def cycle_model(seed=11, cycles=10, m_rest=40, n_rest=100):
    """A network with `cycles` internal two-reaction cycles (A_i -> B_i,
    B_i -> A_i, bounds +-1000: long directions of the polytope) beside a random
    block whose reactions are bounded within +-(1..2) of an interior point
    (short directions): variance ratios of about 1e6, as reversible internal
    loops produce in genome-scale models. Warmup: FVA with the reference's
    0.33 / 0.67 move, stored with "%f" precision."""
    import scipy.linalg
    rng = np.random.default_rng(seed)
    m, n = 2 * cycles + m_rest, 2 * cycles + n_rest
    S = np.zeros((m, n))
    for i in range(cycles):
        S[2 * i, 2 * i], S[2 * i + 1, 2 * i] = -1.0, 1.0
        S[2 * i, 2 * i + 1], S[2 * i + 1, 2 * i + 1] = 1.0, -1.0
    rr, cc, vv = synth.random_stoich(m_rest, n_rest, rng)
    S[2 * cycles + rr, 2 * cycles + cc] = vv
    N = scipy.linalg.null_space(S)
    x0 = N @ rng.standard_normal(N.shape[1])
    x0[2 * cycles:] /= np.abs(x0[2 * cycles:]).max()
    x0[:2 * cycles] = 0.0
    x0 = N @ (N.T @ x0)
    width = np.concatenate([np.full(2 * cycles, 1000.0), rng.uniform(1.0, 2.0, n_rest)])
    lb, ub = x0 - width, x0 + width
    r, c = np.nonzero(S)
    sm = synth.SynthModel(S, r.astype(np.int32), c.astype(np.int32), S[r, c], m, n, lb, ub, x0, N)
    W = synth.round_like_csv(synth.fva_warmup(sm))
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    return S, lb, ub, model, W


def _cycle_error(smp, P, steps, seed):
    """Runs the sampler; returns (moment error of the cycle fluxes against
    their closed form U(-1000, 1000), final points, traces)."""
    smp.run(P, steps, seed=seed)
    tr = smp.trace()
    a = tr[:, tr.shape[1] // 2:, :20].reshape(-1, 20)
    err = max(np.abs(a.mean(0)).max() / 2000.0, (np.abs(a.std(0) - 2000 / np.sqrt(12)) / (2000 / np.sqrt(12))).max())
    return err, tr


def test_learned_bank_no_drift_same_law_and_faster_mixing():
    """Chains start at the FVA warmup centre, where every cycle flux sits at
    -983: the sampler has to travel the long cycle directions to reach the
    uniform law. Isotropic hit-and-run (the method without the network)
    barely moves along them; the learned rounding map does."""
    S, lb, ub, model, W = cycle_model()
    P, steps = 256, 6000
    opts = achr.Options(lanes=64, trace_chains=P, trace_thin=20, check_every=100)
    errs = {}
    for name, kw in (("learned", dict(rank=16, width=128, iters=300, batch=1024, lr=2e-3, seed=5)),
                     ("untrained", dict(rank=16, width=128, iters=0, seed=5)),
                     ("isotropic", dict(isotropic=True, seed=5))):
        print(f"{name}: ACHR training run + learn", flush=True)
        smp = achr.Sampler(model, W, opts)
        smp.run(P, 2000, seed=1)
        rep = smp.learn_directions(**kw)
        if name == "learned":
            assert rep["loss_end"] < 0.2 * rep["loss_start"]
        err, tr = _cycle_error(smp, P, steps, 2)
        errs[name] = err
        print(name, rep, f"cycle moment error {err:.4f}", flush=True)
        # bank directions keep S x = 0 (the first check removes the warmup
        # centre's "%f" residual)
        assert np.abs(S @ tr[:, -1, :].T).max() < 1e-9
    assert errs["learned"] < 0.15
    assert errs["isotropic"] > 0.5
    assert errs["learned"] < 0.5 * errs["untrained"]


def test_bank_mode_bitwise_against_oracle(c1):
    """The bank sampler (stored unit directions, no centre) on the 72 x 95
    model: GPU and oracle agree bit for bit, in-kernel and batched checks."""
    import oracle as O
    S, lb, ub, Wf, We = c1
    N = O.nullspace(S)
    rng = np.random.default_rng(3)
    free = synth.free_directions(N, lb, ub)
    bank = free @ rng.standard_normal((free.shape[1], 150))
    bank /= np.linalg.norm(bank, axis=0)
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    rp, ci, va = O.csr_of(S)
    for lanes, mode in ((32, 1), (64, 2), (256, 1), (1024, 2)):
        print(f"bank lanes={lanes} mode={mode}", flush=True)
        smp = achr.Sampler(model, Wf, achr.Options(lanes=lanes, reproject_mode=mode, trace_chains=5, trace_thin=7))
        smp.set_directions(bank)
        smp.use_directions(True)
        g = smp.sample(96, 150, seed=4)
        o, ost, otr = O.sample(warm=Wf, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=96,
                               S=150, seed=4, lanes=lanes, bank=bank, trace_chains=5, trace_thin=7,
                               c0=smp.start_point())
        np.testing.assert_array_equal(smp.stats(), ost)
        np.testing.assert_array_equal(g.view(np.uint64), o.view(np.uint64))
        np.testing.assert_array_equal(smp.trace().view(np.uint64), otr.view(np.uint64))  # the clipped points


def test_bank_clip_signed_zeros_bitwise():
    """The bank clip is v_max_f64 / v_min_f64 (-0 below +0) in the kernel and
    the same IEEE order in the oracle: a blocked reaction at +0.0 under the
    upper bound -0.0 ends at -0.0 on both, compared as bit patterns (an
    equality test of values would not see the sign of a zero), for both
    bound placements (LDS at 1,024 lanes, memory at 32) and both projection
    modes."""
    import oracle as O
    from conftest import signed_zero_bank_case
    S, lb, ub, N, W, c0, bank = signed_zero_bank_case()
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    rp, ci, va = O.csr_of(S)
    for lanes, mode in ((32, 1), (1024, 2)):
        print(f"signed zeros lanes={lanes} mode={mode}", flush=True)
        smp = achr.Sampler(model, W, achr.Options(lanes=lanes, reproject_mode=mode, trace_chains=3, trace_thin=7))
        smp.set_directions(bank)
        smp.use_directions(True)
        s0 = smp.start_point()
        assert (s0.view(np.uint64)[:2] == 0).all()
        g = smp.sample(40, 60, seed=9)
        o, ost, otr = O.sample(warm=W, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=2, P=40, S=60, seed=9,
                               lanes=lanes, bank=bank, trace_chains=3, trace_thin=7, c0=s0)
        np.testing.assert_array_equal(smp.stats(), ost)
        np.testing.assert_array_equal(g.view(np.uint64), o.view(np.uint64))
        np.testing.assert_array_equal(smp.trace().view(np.uint64), otr.view(np.uint64))
        assert (g.view(np.uint64)[0] == np.uint64(0x8000000000000000)).all()
        assert (g.view(np.uint64)[1] == 0).all()


def test_bank_mode_edge_sizes(c1):
    """Bank sampler with one chain, chain counts that fill no whole block, and
    zero steps (each chain returns its start point as it is), bit-equal to the
    oracle."""
    import oracle as O
    S, lb, ub, Wf, We = c1
    N = O.nullspace(S)
    rng = np.random.default_rng(8)
    free = synth.free_directions(N, lb, ub)
    bank = free @ rng.standard_normal((free.shape[1], 40))
    bank /= np.linalg.norm(bank, axis=0)
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    rp, ci, va = O.csr_of(S)
    for lanes, mode in ((32, 1), (1024, 2)):
        smp = achr.Sampler(model, Wf, achr.Options(lanes=lanes, reproject_mode=mode))
        smp.set_directions(bank)
        smp.use_directions(True)
        s0 = smp.start_point()
        for P, Ss in ((1, 150), (3, 57), (5, 0)):
            print(f"bank edge lanes={lanes} mode={mode} P={P} S={Ss}", flush=True)
            g = smp.sample(P, Ss, seed=21)
            o, ost, _ = O.sample(warm=Wf, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=P,
                                 S=Ss, seed=21, lanes=lanes, bank=bank, c0=s0)
            np.testing.assert_array_equal(smp.stats(), ost)
            np.testing.assert_array_equal(g, o)
        np.testing.assert_array_equal(g, np.repeat(s0[:, None], 5, axis=1))


def test_bank_chains_start_on_the_flux_set_and_never_reproject():
    """The centre of a "%f" warmup is off S x = 0 by the file's rounding. Bank
    chains start from the nearest point with S x = 0 and every fixed reaction
    at its bound (checked against a dense least-squares solution); bank
    directions keep both, so the residual checks of a bank run never flag and
    the points stay on that set. The fixed rows of N are linearly dependent
    here (80 rows of rank 79), the case a plain conjugate-gradient solve of the
    fixed-reaction correction gets wrong."""
    import scipy.linalg
    m = synth.make_model(330, 600, 43, irrev_frac=0.3, n_fixed=80, width=(5.0, 50.0))
    S = m.dense()
    N = scipy.linalg.null_space(S)
    W = synth.round_like_csv(synth.chord_warmup(N, m.x0, m.lb, m.ub, 800, seed=44))
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(N)
    fixed = m.lb == m.ub
    assert fixed.sum() == 80 and np.linalg.matrix_rank(N[fixed], tol=1e-10) == 79
    smp = achr.Sampler(model, W, achr.Options(reproject_mode=2))
    c = smp.center()
    assert np.abs(S @ c).max() > 1e-8
    rng = np.random.default_rng(2)
    free = synth.free_directions(N, m.lb, m.ub)
    bank = free @ rng.standard_normal((free.shape[1], 500))
    smp.set_directions(bank / np.linalg.norm(bank, axis=0))
    np.testing.assert_array_equal(smp.start_point(), c)  # ACHR directions start at the centre
    smp.use_directions(True)
    x0 = smp.start_point()
    A = np.vstack([S, np.eye(m.n)[fixed]])
    b = np.concatenate([np.zeros(m.m), m.lb[fixed]])
    want = c - np.linalg.lstsq(A, A @ c - b, rcond=None)[0]
    assert np.abs(x0 - want).max() < 1e-9
    assert np.abs(S @ x0).max() < 1e-11
    np.testing.assert_array_equal(x0[fixed], m.lb[fixed])
    smp.run(200, 300, seed=6)
    stt = smp.stats()
    assert stt[:, 2].min() >= 29 and stt[:, 1].sum() == 0  # 30 checks per chain, none reprojects
    pts = smp.points()
    assert np.abs(S @ pts).max() < 1e-10
    np.testing.assert_array_equal(pts[fixed], np.repeat(m.lb[fixed][:, None], 200, axis=1))


def test_learned_and_isotropic_directions_agree_at_convergence():
    """On a 165 x 300 model, long enough to mix (3e6 steps), the learned
    rounding map and isotropic free-null-space directions (exact hit-and-run)
    both reach split-R-hat < 1.1 and their 1,000 chain endpoints agree within
    sampling error per reaction: the learned law changes the speed, not the
    target. (Plain ACHR started from these chord warmups is still unmixed at
    this length; its adaptive centre is not a Markov chain.)"""
    import scipy.linalg
    m = synth.make_model(165, 300, 51, irrev_frac=0.3, n_fixed=6, width=(5.0, 50.0))
    N = scipy.linalg.null_space(m.dense())
    W = synth.chord_warmup(N, m.x0, m.lb, m.ub, 600, seed=52)
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(N)
    P, S = 1000, 3_000_000
    out = {}
    for name, kw in (("learned", dict(rank=64, width=256, iters=300, batch=1024, lr=2e-3, seed=3)),
                     ("isotropic", dict(isotropic=True, seed=4))):
        print(f"{name}: training run + bank", flush=True)
        smp = achr.Sampler(model, W, achr.Options(model_id=2, trace_chains=P, trace_thin=50))
        smp.run(P, 2000, seed=9)
        smp.learn_directions(**kw)
        smp.set_run_options(64, S // 100, 100)
        print(f"{name}: {S} steps", flush=True)
        smp.run(P, S, seed=6)
        out[name] = (smp.points(), smp.trace())
    A, B = out["learned"][0], out["isotropic"][0]
    rng = m.ub - m.lb
    live = (rng > 0) & (np.maximum(A.std(1), B.std(1)) > 1e-9 * rng)
    for name in out:
        rh = np.nanmax(st.rhat(out[name][1][:, :, live]))
        print(f"{name}: max split-R-hat {rh:.3f}", flush=True)
        assert rh < 1.1
    z = np.abs(A.mean(1) - B.mean(1))[live] / np.sqrt((A.var(1) + B.var(1))[live] / P)
    zs = np.abs(A.std(1) - B.std(1))[live] / np.sqrt((A.var(1) + B.var(1))[live] / (2 * P))
    print(f"max |z| of means {z.max():.2f}, of stds {zs.max():.2f} over {live.sum()} reactions", flush=True)
    assert z.max() < 5.0 and zs.max() < 5.0
