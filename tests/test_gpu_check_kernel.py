"""The residual check of batched rounds in its own kernel (check_kernel 2).

Every chain stops at its check with x stored; one launch sums the rows of S
for all of them (x staged in LDS, S from L2) and sets the flags, the
reprojection counters and, for the reduced-precision rounds, the FP32
residual rows. The rows are summed in CSR order from +0.0 as the in-kernel
check and the oracle sum them (findMaxAbs, ACHRcu/ACHRCuda.cu:232-240,
:325-338), so a run gives the same bits with the check in either place:
points, counters, traces, and the oracle's."""
import numpy as np
import pytest

import achr
import oracle as O
from achr import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def genome_case(mets, rxns, seed, exact=False):
    def ns(rows, cols, vals, m, n):
        mdl = achr.Model.from_coo(n, m, rows, cols, vals, -np.ones(n), np.ones(n))
        mdl.compute_nullspace()
        return mdl.nullspace()

    m = synth.make_model(mets, rxns, seed, irrev_frac=0.4, n_fixed=rxns // 100, width=(100.0, 1000.0),
                         x0_scale=100.0, nullspace_fn=ns)
    W = synth.chord_warmup(m.N, m.x0, m.lb, m.ub, 300, seed=seed + 1)
    if not exact:
        W = synth.round_like_csv(W)
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(m.N)
    import scipy.sparse as sp
    Sd = sp.csr_matrix((m.vals, (m.rows, m.cols)), shape=(m.m, m.n))
    Sd.sort_indices()
    return m, W, model, Sd


def run(model, W, P, S, seed, **opts):
    smp = achr.Sampler(model, W, achr.Options(reproject_mode=2, **opts))
    smp.run(P, S, seed=seed)
    out = (smp.points(), smp.stats(), smp.trace() if opts.get("trace_chains") else None, smp.timing())
    del smp
    return out


@pytest.mark.parametrize("lanes,form", [(32, 2), (256, 2), (1024, 2), (256, 3)])
def test_check_kernel_on_c1_bitwise(c1, lanes, form):
    """Every check of the "%f" warmup reprojects; ragged chain counts. Form
    2 holds S in registers (the 72 x 95 table fits), form 3 reads it from
    L2."""
    S_, lb, ub, Wf, _ = c1
    N = O.nullspace(S_)
    model = achr.Model.from_dense(S_, lb, ub)
    model.set_nullspace(N)
    rp, ci, va = O.csr_of(S_)
    for P, S in [(1, 31), (77, 123)]:
        say(f"c1 lanes {lanes} form {form} P {P} S {S}")
        g, gs, _, tm = run(model, Wf, P, S, 11, lanes=lanes, check_kernel=form)
        o, os_, _ = O.sample(warm=Wf, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S_.shape[0], P=P, S=S,
                             seed=11, lanes=lanes)
        np.testing.assert_array_equal(gs, os_)
        np.testing.assert_array_equal(g, o)
        assert (gs[:, 1] == gs[:, 2]).all()
        assert tm["check_launches"] == tm["step_launches"] - 1 > 0


@pytest.mark.parametrize("exact", [False, True])
def test_check_kernel_genome_schedule_equals_in_kernel_and_oracle(exact):
    """1,024-lane chains at 4,400 x 8,000: with the "%f" warmup every check
    flags; with exact warmup points only the checks after clipped steps do,
    so most chains stop at a check without a projection and resume."""
    say(f"genome case 4400 x 8000, exact warmup {exact}")
    m, W, model, Sd = genome_case(4400, 8000, 41, exact=exact)
    P, S, seed = 12, 47, 6
    a = run(model, W, P, S, seed, lanes=1024, check_kernel=2, trace_chains=3, trace_thin=9)
    b = run(model, W, P, S, seed, lanes=1024, check_kernel=1, trace_chains=3, trace_thin=9)
    c = run(model, W, P, S, seed, lanes=1024, check_kernel=3, trace_chains=3, trace_thin=9)
    say(f"timing with the check kernel {a[3]}, inside the step kernel {b[3]}, S from L2 {c[3]}")
    for other in (b, c):
        np.testing.assert_array_equal(a[1], other[1])
        np.testing.assert_array_equal(a[0], other[0])
        np.testing.assert_array_equal(a[2], other[2])
    assert a[3]["check_launches"] > 0 and c[3]["check_launches"] > 0 and b[3]["check_launches"] == 0
    o, ost, otr = O.sample(warm=W, lb=m.lb, ub=m.ub, rowptr=Sd.indptr.astype(np.int32),
                           colidx=Sd.indices.astype(np.int32), vals=Sd.data, N=m.N, m=m.m, P=P, S=S, seed=seed,
                           lanes=1024, threads=8, trace_chains=3, trace_thin=9)
    np.testing.assert_array_equal(a[1], ost)
    np.testing.assert_array_equal(a[0], o)
    np.testing.assert_array_equal(a[2], otr)
    flagged = int(a[1][:, 1].sum())
    checks = int(a[1][:, 2].sum())
    say(f"reprojections {flagged} of {checks} checks")
    if exact:
        assert flagged < checks
    else:
        assert flagged == checks


@pytest.mark.parametrize("prec", [achr.PROJ_RES_F32, achr.PROJ_RES_BF16X3])
def test_check_kernel_residual_rows_of_reduced_precision_rounds(prec):
    """The residual-form rounds contract the check's FP32 row sums with S+:
    the check kernel writes the same rows (both forms), so the runs are
    identical."""
    say(f"reduced precision {prec}, 2200 x 4000, 512 lanes")
    m, W, model, Sd = genome_case(2200, 4000, 43)
    P, S, seed = 40, 61, 8
    a = run(model, W, P, S, seed, lanes=512, check_kernel=2, proj_prec=prec, fp64_every=3)
    b = run(model, W, P, S, seed, lanes=512, check_kernel=1, proj_prec=prec, fp64_every=3)
    c = run(model, W, P, S, seed, lanes=512, check_kernel=3, proj_prec=prec, fp64_every=3)
    for other in (b, c):
        np.testing.assert_array_equal(a[1], other[1])
        np.testing.assert_array_equal(a[0], other[0])
    assert a[3]["check_launches"] > 0
    assert np.abs(Sd @ a[0]).max() < 1e-3


def test_check_kernel_on_p_putida_hub_rows():
    """P. putida (911 x 1,060): hub metabolites give slices wider than a
    wavefront holds, so the kernel reads S from L2; the run equals the
    in-kernel check's and the oracle's."""
    from conftest import golden
    model = achr.Model.read_mps(golden("P_Putida.mps"))
    model.compute_nullspace()
    N = model.nullspace()
    lb, ub = model.bounds()
    # a short warmup set from the LP generator: its points need not be
    # vertices for a comparison of two placements of the same check
    W, _, _, _ = achr.warmup_points(model, 2 * model.n, max_iters=2000, polish=False)
    P, S, seed = 24, 35, 2
    a = run(model, W, P, S, seed, lanes=512, check_kernel=2)
    b = run(model, W, P, S, seed, lanes=512, check_kernel=1)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[0], b[0])
    assert a[3]["check_launches"] > 0
    rp, ci, va = model.csr()
    o, ost, _ = O.sample(warm=W, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=model.m, P=P, S=S, seed=seed,
                         lanes=512, threads=8)
    np.testing.assert_array_equal(a[1], ost)
    np.testing.assert_array_equal(a[0], o)


def test_check_kernel_without_rows_and_with_empty_rows():
    """A box (no rows of S: the residual is 0 and nothing is projected) and a
    model whose S has empty rows (slices of width 0, left out of the
    resident plan) at 1,024 lanes: the check kernel's runs equal the
    in-kernel check's and the oracle's."""
    rng = np.random.default_rng(9)
    n = 2500
    lb = -rng.uniform(1.0, 5.0, n)
    ub = rng.uniform(1.0, 5.0, n)
    W = lb[:, None] + (ub - lb)[:, None] * rng.uniform(0.2, 0.8, (n, 64))
    box = achr.Model.from_coo(n, 0, [], [], [], lb, ub)
    box.set_nullspace(np.eye(n))
    # rows 0..9 empty, then 40 rows coupling random pairs of reactions
    m = 50
    rows = np.repeat(np.arange(10, m), 2)
    cols = rng.choice(n, size=2 * (m - 10), replace=False)
    vals = np.tile([1.0, -1.0], m - 10)
    Sd = np.zeros((m, n))
    Sd[rows, cols] = vals
    N = O.nullspace(Sd)
    x0 = N @ rng.standard_normal(N.shape[1])
    x0 *= 0.5 / np.abs(x0).max()
    Wc = x0[:, None] + N @ (0.1 * rng.standard_normal((N.shape[1], 64)))
    coupled = achr.Model.from_coo(n, m, rows, cols, vals, lb, ub)
    coupled.set_nullspace(N)
    rp, ci, va = O.csr_of(Sd)
    for name, model, Wm, Nm, mm, (r_, c_, v_) in (
            ("box", box, W, np.eye(n), 0, O.csr_of(np.zeros((0, n)))),
            ("empty rows", coupled, Wc, N, m, (rp, ci, va))):
        say(f"{name}: n {n}, m {mm}")
        P, S, seed = 9, 33, 4
        a = run(model, Wm, P, S, seed, lanes=1024, check_kernel=2)
        b = run(model, Wm, P, S, seed, lanes=1024, check_kernel=1)
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_array_equal(a[0], b[0])
        assert a[3]["check_launches"] > 0
        o, ost, _ = O.sample(warm=Wm, lb=lb, ub=ub, rowptr=r_, colidx=c_, vals=v_, N=Nm, m=mm, P=P, S=S,
                             seed=seed, lanes=1024, threads=8)
        np.testing.assert_array_equal(a[1], ost)
        np.testing.assert_array_equal(a[0], o)
