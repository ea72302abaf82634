"""GPU batches: several models of uneven size in one job (one persistent
launch per schedule bucket and segment, the segments replayed from a HIP
graph), called through the C ABI. Every model's points and counters must equal
the oracle's for that model's Philox stream (model_id) and lane count bit for
bit, and equal a run of the same sampler on its own."""
import numpy as np
import pytest

import achr
import oracle as O
from achr import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_case(mets, rxns, seed, *, csv=False, W=None):
    import scipy.linalg
    m = synth.make_model(mets, rxns, seed, irrev_frac=0.3, n_fixed=max(1, rxns // 60), width=(5.0, 50.0))
    N = scipy.linalg.null_space(m.dense())
    warm = synth.chord_warmup(N, m.x0, m.lb, m.ub, W or 2 * rxns, seed=seed + 1)
    if csv:
        warm = synth.round_like_csv(warm)
    model = achr.Model.from_coo(m.n, m.m, m.rows, m.cols, m.vals, m.lb, m.ub)
    model.set_nullspace(N)
    rp, ci, va = O.csr_of(m.dense())
    return dict(S=m.dense(), lb=m.lb, ub=m.ub, N=N, rp=rp, ci=ci, va=va, model=model, warm=warm)


@pytest.fixture(scope="module")
def cases(c1):
    S, lb, ub, Wf, _ = c1
    N = O.nullspace(S)
    rp, ci, va = O.csr_of(S)
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    out = [dict(S=S, lb=lb, ub=ub, N=N, rp=rp, ci=ci, va=va, model=model, warm=Wf)]
    # n = 40 (16-32 lanes), 200 (64), 600 (batched projection rounds, "%f"
    # warmup so that checks reproject), 1300 (512 lanes, block per chain)
    for mets, rxns, seed, csv in ((22, 40, 41, False), (110, 200, 42, False), (330, 600, 43, True),
                                  (700, 1300, 44, False)):
        say(f"building model {mets} x {rxns}")
        out.append(make_case(mets, rxns, seed, csv=csv, W=min(2 * rxns, 800)))
    return out


def oracle_of(c, *, P, S, seed, file, model_id, lanes, bank=None, c0=None):
    say(f"oracle n={c['S'].shape[1]} P={P} S={S} lanes={lanes} model={model_id}")
    pts, st, _ = O.sample(warm=c["warm"], lb=c["lb"], ub=c["ub"], rowptr=c["rp"], colidx=c["ci"], vals=c["va"],
                          N=c["N"], m=c["S"].shape[0], P=P, S=S, seed=seed, file=file, model=model_id, lanes=lanes,
                          threads=8, bank=bank, c0=c0)
    return pts, st


def samplers_for(cases, lanes=None):
    out = []
    for k, c in enumerate(cases):
        opts = achr.Options(model_id=k + 1, lanes=0 if lanes is None else lanes[k])
        out.append(achr.Sampler(c["model"], c["warm"], opts))
    return out


def test_batch_matches_oracle_and_solo_runs(cases):
    P = [37, 130, 64, 50, 21]  # ragged, not multiples of the chains per workgroup
    S, seed, file = 60, 77, 2
    smps = samplers_for(cases)
    batch = achr.Batch(smps)
    lanes = batch.lanes()
    say(f"lanes {lanes.tolist()}, modes {[s.schedule()[1] for s in smps]}")
    assert len(set(lanes.tolist())) >= 3
    batch.run(P, S, seed=seed, file=file)
    t = batch.timing()
    say(f"timing {t}")
    assert t["graphs"] >= 1 and t["buckets"] >= 3 and t["segments"] == (S + 9) // 10 + 1
    got = [(s.points(), s.stats()) for s in smps]
    for k, c in enumerate(cases):
        o_pts, o_st = oracle_of(c, P=P[k], S=S, seed=seed, file=file, model_id=k + 1, lanes=int(lanes[k]))
        np.testing.assert_array_equal(got[k][1], o_st)
        np.testing.assert_array_equal(got[k][0], o_pts)
    assert got[3][1][:, 1].sum() > 0  # the "%f" model reprojected inside the graph
    # the same work launched eagerly, and each model on its own
    batch.run(P, S, seed=seed, file=file, graph=False)
    for k, s in enumerate(smps):
        np.testing.assert_array_equal(s.points(), got[k][0])
    for k, s in enumerate(smps):
        say(f"solo run of model {k}")
        s.run(P[k], S, seed=seed, file=file)
        np.testing.assert_array_equal(s.points(), got[k][0])
        np.testing.assert_array_equal(s.stats(), got[k][1])


def test_batch_edge_sizes(cases):
    """A batch of one to three chains per model, then zero steps (every chain
    returns the warmup centre), in the graph: each model bit-equal to the
    oracle. A model with no chains is refused."""
    smps = samplers_for(cases)
    batch = achr.Batch(smps)
    lanes = batch.lanes()
    P = [1, 3, 1, 2, 1]
    for S in (57, 0):
        say(f"batch edge P={P} S={S}")
        batch.run(P, S, seed=13, file=1)
        for k, c in enumerate(cases):
            o_pts, o_st = oracle_of(c, P=P[k], S=S, seed=13, file=1, model_id=k + 1, lanes=int(lanes[k]))
            np.testing.assert_array_equal(smps[k].stats(), o_st)
            np.testing.assert_array_equal(smps[k].points(), o_pts)
    for k, s in enumerate(smps):
        np.testing.assert_array_equal(s.points(), np.repeat(s.center()[:, None], P[k], axis=1))
    with pytest.raises(achr.AchrError) as e:
        batch.run([1, 0, 1, 1, 1], 10, seed=1)
    assert e.value.code == -1


def test_autotune_then_pinned_lanes_reproduce(cases):
    P = [64, 96, 64, 40, 24]
    smps = samplers_for(cases)
    batch = achr.Batch(smps)
    ms = batch.autotune(P, probe_steps=12, seed=1)
    lanes = batch.lanes()
    say(f"autotuned lanes {lanes.tolist()}")
    for k in range(len(cases)):
        cand = ms[k] > 0
        assert cand.any()
        j = achr.Batch.TPC.index(int(lanes[k]))
        assert cand[j] and ms[k, j] == ms[k][cand].min()
    batch.run(P, 40, seed=5)
    first = [s.points() for s in smps]
    for k, c in enumerate(cases):
        o_pts, _ = oracle_of(c, P=P[k], S=40, seed=5, file=0, model_id=k + 1, lanes=int(lanes[k]))
        np.testing.assert_array_equal(first[k], o_pts)
    # pinned lanes in a fresh batch give the same samples
    b2 = achr.Batch(samplers_for(cases))
    b2.set_lanes(lanes)
    b2.run(P, 40, seed=5)
    for k, s in enumerate(b2.samplers):
        np.testing.assert_array_equal(s.points(), first[k])
    with pytest.raises(achr.AchrError):
        b2.set_lanes([16] * len(cases))  # 16 lanes cannot hold 1300 reactions


def test_batch_with_direction_bank_and_rerun(cases):
    """A bank-mode sampler (no centre, 256 lanes: the min/max reduction slots
    alternate between attempts) next to plain ones; then a second run of the
    same batch with other chain counts."""
    c = cases[3]
    rng = np.random.default_rng(8)
    free = synth.free_directions(c["N"], c["lb"], c["ub"])
    bank = free @ rng.standard_normal((free.shape[1], 300))
    bank /= np.linalg.norm(bank, axis=0)
    smps = samplers_for(cases, lanes=[0, 0, 0, 256, 0])
    smps[3].set_directions(bank)
    smps[3].use_directions(True)
    batch = achr.Batch(smps)
    for P, seed in (([16, 64, 32, 80, 8], 11), ([40, 8, 70, 33, 30], 12)):
        batch.run(P, 50, seed=seed)
        lanes = batch.lanes()
        for k, cc in enumerate(cases):
            o_pts, o_st = oracle_of(cc, P=P[k], S=50, seed=seed, file=0, model_id=k + 1, lanes=int(lanes[k]),
                                    bank=bank if k == 3 else None, c0=smps[k].start_point())
            np.testing.assert_array_equal(smps[k].stats(), o_st)
            np.testing.assert_array_equal(smps[k].points(), o_pts)
