"""Genome-scale chains of 1,024 lanes keep the per-reaction bounds on chip:
the first 9 slot rows in LDS as (lb, ub) pairs, up to 2 more rows in
registers, any further row read per step. Bank mode at 10, 11 and 12 slots
(so every kind of storage is exercised), plain ACHR at 8 slots (all in
LDS) and at 9 and 11 slots (centre in LDS, next warmup point prefetched):
GPU and oracle agree bit for bit, points, counters and traces. The bank
mode's reciprocal (hardware estimate + three Newton steps) must equal IEEE
division on the whole range the chord uses."""
import numpy as np
import pytest

import achr
import oracle as O
from achr import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def genome_case(mets, rxns, seed):
    def ns(rows, cols, vals, m, n):
        mdl = achr.Model.from_coo(n, m, rows, cols, vals, -np.ones(n), np.ones(n))
        mdl.compute_nullspace()
        return mdl.nullspace()

    m = synth.make_model(mets, rxns, seed, irrev_frac=0.4, n_fixed=rxns // 100, width=(100.0, 1000.0),
                         x0_scale=100.0, nullspace_fn=ns)
    W = synth.round_like_csv(synth.chord_warmup(m.N, m.x0, m.lb, m.ub, 300, seed=seed + 1))
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(m.N)
    import scipy.sparse as sp
    Sd = sp.csr_matrix((m.vals, (m.rows, m.cols)), shape=(m.m, m.n))
    Sd.sort_indices()
    return m, W, model, Sd


def oracle_run(m, W, Sd, *, bank, lanes, P, S, seed, trace, thin, c0=None):
    return O.sample(warm=W, lb=m.lb, ub=m.ub, rowptr=Sd.indptr.astype(np.int32), colidx=Sd.indices.astype(np.int32),
                    vals=Sd.data, N=m.N, m=m.m, P=P, S=S, seed=seed, lanes=lanes, threads=8, bank=bank,
                    trace_chains=trace, trace_thin=thin, c0=c0)


@pytest.mark.parametrize("mets,rxns,slots", [(5200, 9500, 10), (5800, 10600, 11), (6600, 12000, 12)])
def test_bank_chain_with_resident_bounds_bitwise(mets, rxns, slots):
    say(f"bank case {mets} x {rxns}")
    m, W, model, Sd = genome_case(mets, rxns, 31)
    rng = np.random.default_rng(5)
    free = synth.free_directions(m.N, m.lb, m.ub)
    bank = free @ rng.standard_normal((free.shape[1], 400))
    bank /= np.linalg.norm(bank, axis=0)
    smp = achr.Sampler(model, W, achr.Options(lanes=1024, reproject_mode=2, trace_chains=3, trace_thin=7))
    smp.set_directions(bank)
    smp.use_directions(True)
    assert smp.schedule() == (1024, 2)
    P, S, seed = 6, 30, 9
    smp.run(P, S, seed=seed)
    o, ost, otr = oracle_run(m, W, Sd, bank=bank, lanes=1024, P=P, S=S, seed=seed, trace=3, thin=7,
                             c0=smp.start_point())
    np.testing.assert_array_equal(smp.stats(), ost)
    np.testing.assert_array_equal(smp.points(), o)
    np.testing.assert_array_equal(smp.trace(), otr)
    assert ost[:, 2].min() >= 2  # the residual checks ran inside the run


@pytest.mark.parametrize("mets,rxns,lanes", [(4400, 8000, 1024), (2200, 4000, 512)])
def test_plain_chain_with_bounds_in_lds_bitwise(mets, rxns, lanes):
    say(f"plain case {mets} x {rxns} lanes {lanes}")
    m, W, model, Sd = genome_case(mets, rxns, 37)
    smp = achr.Sampler(model, W, achr.Options(lanes=lanes, reproject_mode=2))
    assert smp.schedule() == (lanes, 2)
    P, S, seed = 6, 25, 4
    smp.run(P, S, seed=seed)
    o, ost, _ = oracle_run(m, W, Sd, bank=None, lanes=lanes, P=P, S=S, seed=seed, trace=0, thin=0)
    np.testing.assert_array_equal(smp.stats(), ost)
    np.testing.assert_array_equal(smp.points(), o)


@pytest.mark.parametrize("mets,rxns", [(5000, 9200), (5800, 10600)])
def test_plain_chain_with_prefetch_and_centre_in_lds_bitwise(mets, rxns):
    """Plain ACHR at 9 and 11 slots of 1,024 lanes: the next warmup point is
    prefetched into registers while the running centre lives in LDS."""
    say(f"plain prefetch case {mets} x {rxns}")
    m, W, model, Sd = genome_case(mets, rxns, 41)
    smp = achr.Sampler(model, W, achr.Options(lanes=1024, reproject_mode=2, trace_chains=2, trace_thin=5))
    assert smp.schedule() == (1024, 2)
    P, S, seed = 6, 25, 6
    smp.run(P, S, seed=seed)
    o, ost, otr = oracle_run(m, W, Sd, bank=None, lanes=1024, P=P, S=S, seed=seed, trace=2, thin=5)
    np.testing.assert_array_equal(smp.stats(), ost)
    np.testing.assert_array_equal(smp.points(), o)
    np.testing.assert_array_equal(smp.trace(), otr)


def test_bank_reciprocal_is_ieee_division():
    rng = np.random.default_rng(11)
    count = 1 << 22
    sig = rng.integers(0, 1 << 52, count, dtype=np.uint64)
    # significands where rounding is delicate: all ones, all zeros (powers of
    # two), one bit set at either end, and those next to them
    special = np.array([(1 << 52) - 1, 0, 1, 1 << 51, (1 << 52) - 2, (1 << 51) - 1, (1 << 51) + 1, 2, 3],
                       dtype=np.uint64)
    sig[:special.size * 61] = np.repeat(special, 61)
    exp = rng.integers(1023 - 30, 1023 + 30, count, dtype=np.uint64)  # |u| in [2^-30, 2^30)
    exp[:special.size * 61] = np.tile(np.arange(1023 - 30, 1023 + 31, dtype=np.uint64), special.size)
    sign = rng.integers(0, 2, count, dtype=np.uint64)
    u = ((sign << np.uint64(63)) | (exp << np.uint64(52)) | sig).view(np.float64)
    r = achr.bank_reciprocal(u)
    bad = np.flatnonzero(r.view(np.uint64) != (1.0 / u).view(np.uint64))
    assert bad.size == 0, f"{bad.size} misrounded, first u = {u[bad[:4]]!r}"


def test_chord_quotient_is_ieee_division():
    """The plain sampler's division replacement (a correctly rounded
    reciprocal and one corrected product) against IEEE division, bit for bit
    with signed zeros: 4M quotients with |b| in [2^-30, 2^30] (the norm and
    every |u_i| the chord uses) and |a| in [2^-45, 2^15] or zero, delicate
    significands of b included."""
    rng = np.random.default_rng(12)
    count = 1 << 22

    def f64(sign, exp, sig):
        return ((sign << np.uint64(63)) | (exp << np.uint64(52)) | sig).view(np.float64)

    sig = rng.integers(0, 1 << 52, count, dtype=np.uint64)
    special = np.array([(1 << 52) - 1, 0, 1, 1 << 51, (1 << 52) - 2, (1 << 51) - 1, (1 << 51) + 1, 2, 3],
                       dtype=np.uint64)
    sig[:special.size * 61] = np.repeat(special, 61)
    exp = rng.integers(1023 - 30, 1023 + 30, count, dtype=np.uint64)
    exp[:special.size * 61] = np.tile(np.arange(1023 - 30, 1023 + 31, dtype=np.uint64), special.size)
    b = f64(rng.integers(0, 2, count, dtype=np.uint64), exp, sig)
    asig = rng.integers(0, 1 << 52, count, dtype=np.uint64)
    asig[count // 2: count // 2 + special.size * 61] = np.repeat(special, 61)
    a = f64(rng.integers(0, 2, count, dtype=np.uint64), rng.integers(1023 - 45, 1023 + 15, count, dtype=np.uint64),
            asig)
    a[:1000] = 0.0
    a[1000:2000] = -0.0
    q = achr.quotient(a, b)
    want = a / b
    bad = np.flatnonzero(q.view(np.uint64) != want.view(np.uint64))
    assert bad.size == 0, f"{bad.size} misrounded, first a = {a[bad[:3]]!r}, b = {b[bad[:3]]!r}"
