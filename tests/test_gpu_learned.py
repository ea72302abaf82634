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


def _ess_per_step(smp, P, S, seed):
    smp.run(P, S, seed=seed)
    tr = smp.trace()
    live = (smp_bounds[1] - smp_bounds[0]) > 0
    return st.bulk_ess(tr[:, :, live]), tr


smp_bounds = None


def test_learned_bank_no_drift_same_law_and_faster_mixing():
    global smp_bounds
    S, lb, ub, model, W = cycle_model()
    smp_bounds = (lb, ub)
    live = (ub - lb) > 0
    P, steps = 256, 2000
    opts = achr.Options(lanes=64, trace_chains=P, trace_thin=20, check_every=100)
    print("plain ACHR training run", flush=True)
    plain = achr.Sampler(model, W, opts)
    plain.run(P, steps, seed=1)
    print("learn", flush=True)
    rep = plain.learn_directions(rank=16, width=128, iters=300, batch=1024, lr=2e-3, seed=5)
    assert rep["loss_end"] < rep["loss_start"]
    print(rep, flush=True)
    learned_ess, tr_l = _ess_per_step(plain, P, steps, 2)
    pts_l = tr_l[:, -1, :].T
    iso = achr.Sampler(model, W, opts)
    iso.run(P, 10, seed=0)
    iso.learn_directions(isotropic=True, seed=5)
    iso_ess, tr_i = _ess_per_step(iso, P, steps, 2)
    # 1. bank directions keep S x = 0 (after the first FP64 check removes the
    #    warmup centre's "%f" residual)
    assert np.abs(S @ pts_l).max() < 1e-9
    # 2. stationary law: the polytope is the product of the cycle cube
    #    [-1000, 1000]^10 and the rest, so each cycle flux is U(-1000, 1000):
    #    mean 0, std 2000/sqrt(12), small KS distance to the uniform law.
    a = tr_l[:, tr_l.shape[1] // 2:, :20].reshape(-1, 20)
    assert np.all(np.abs(a.mean(0)) < 0.05 * 2000)
    np.testing.assert_allclose(a.std(0), 2000 / np.sqrt(12), rtol=0.08)
    for j in range(0, 20, 2):
        z = np.sort((a[:, j] + 1000) / 2000)
        assert np.max(np.abs(np.arange(1, len(z) + 1) / len(z) - z)) < 0.08
    # 3. the trained map mixes faster than isotropic directions
    gain = np.nanmin(learned_ess) / np.nanmin(iso_ess)
    print(f"min-ESS learned {np.nanmin(learned_ess):.1f} isotropic {np.nanmin(iso_ess):.1f} gain {gain:.2f}",
          flush=True)
    assert gain > 1.5


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
    for lanes, mode in ((32, 1), (64, 2), (256, 1)):
        print(f"bank lanes={lanes} mode={mode}", flush=True)
        smp = achr.Sampler(model, Wf, achr.Options(lanes=lanes, reproject_mode=mode))
        smp.set_directions(bank)
        smp.use_directions(True)
        g = smp.sample(96, 150, seed=4)
        o, ost, _ = O.sample(warm=Wf, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=96, S=150,
                             seed=4, lanes=lanes, bank=bank)
        np.testing.assert_array_equal(smp.stats(), ost)
        np.testing.assert_array_equal(g, o)
