"""Continuing a run, and checkpoint / resume: S1 steps and then S2 more (in the
same sampler, or after saving the chains' state and loading it into a new
sampler of the same model) equal one run of S1 + S2 steps bit for bit:
points, counters and traces, for plain ACHR and the direction bank, with the
checks inside the kernel and as batched projection rounds. The oracle pins
the long run itself."""
import numpy as np
import pytest

import achr
import oracle as O
from achr import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def setup(c1):
    S, lb, ub, Wf, _ = c1
    N = O.nullspace(S)
    rp, ci, va = O.csr_of(S)
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(N)
    rng = np.random.default_rng(12)
    free = synth.free_directions(N, lb, ub)
    bank = free @ rng.standard_normal((free.shape[1], 120))
    bank /= np.linalg.norm(bank, axis=0)
    return dict(S=S, lb=lb, ub=ub, W=Wf, N=N, rp=rp, ci=ci, va=va, model=model, bank=bank)


def make(st, lanes, mode, bank):
    smp = achr.Sampler(st["model"], st["W"], achr.Options(lanes=lanes, reproject_mode=mode, trace_chains=3,
                                                           trace_thin=7))
    if bank:
        smp.set_directions(st["bank"])
        smp.use_directions(True)
    return smp


@pytest.mark.parametrize("lanes,mode,bank,S1,S2", [(32, 1, False, 37, 58), (1024, 2, False, 40, 25),
                                                  (64, 2, True, 33, 41), (256, 1, True, 20, 31)])
def test_continue_equals_one_long_run(setup, tmp_path, lanes, mode, bank, S1, S2):
    say(f"case lanes={lanes} mode={mode} bank={bank} S1={S1} S2={S2}")
    P, seed, file = 40, 5, 3
    long = make(setup, lanes, mode, bank)
    long.run(P, S1 + S2, seed=seed, file=file)
    pts, sts, tr = long.points(), long.stats(), long.trace()
    o, ost, otr = O.sample(warm=setup["W"], lb=setup["lb"], ub=setup["ub"], rowptr=setup["rp"], colidx=setup["ci"],
                           vals=setup["va"], N=setup["N"], m=setup["S"].shape[0], P=P, S=S1 + S2, seed=seed,
                           file=file, lanes=lanes, bank=setup["bank"] if bank else None, trace_chains=3,
                           trace_thin=7, c0=long.start_point(), threads=8)
    np.testing.assert_array_equal(sts, ost)
    np.testing.assert_array_equal(pts, o)
    np.testing.assert_array_equal(tr, otr)
    assert sts[:, 1].sum() > 0 or bank  # plain chains on the "%f" warmup reproject

    smp = make(setup, lanes, mode, bank)
    smp.run(P, S1, seed=seed, file=file)
    smp.save_state(str(tmp_path / "state.bin"))
    smp.continue_run(S2)
    np.testing.assert_array_equal(smp.stats(), sts)
    np.testing.assert_array_equal(smp.points(), pts)
    np.testing.assert_array_equal(smp.trace(), tr)

    fresh = make(setup, lanes, mode, bank)
    fresh.load_state(str(tmp_path / "state.bin"))
    assert (fresh.P, fresh.S) == (P, S1)
    fresh.continue_run(S2)
    np.testing.assert_array_equal(fresh.stats(), sts)
    np.testing.assert_array_equal(fresh.points(), pts)


def test_several_continues_and_refusals(setup, tmp_path):
    P = 24
    long = make(setup, 64, 2, False)
    long.run(P, 35, seed=8)
    smp = make(setup, 64, 2, False)
    smp.run(P, 10, seed=8)
    before = smp.points()
    smp.continue_run(0)
    np.testing.assert_array_equal(smp.points(), before)
    smp.continue_run(10)
    smp.continue_run(15)
    np.testing.assert_array_equal(smp.points(), long.points())
    np.testing.assert_array_equal(smp.stats(), long.stats())
    np.testing.assert_array_equal(smp.trace(), long.trace())
    with pytest.raises(achr.AchrError):
        make(setup, 64, 2, False).continue_run(5)  # nothing to continue
    smp.set_run_options(3, 7, 20)
    with pytest.raises(achr.AchrError):
        smp.continue_run(5)  # other check period than the run's
    smp.set_run_options(3, 7, 10)
    smp.save_state(str(tmp_path / "s.bin"))
    # a sampler whose own schedule has other lanes resumes the saved run with
    # the run's 64 lanes (they fix the norm's reduction tree)
    other = make(setup, 128, 2, False)
    other.load_state(str(tmp_path / "s.bin"))
    other.continue_run(10)
    longer = make(setup, 64, 2, False)
    longer.run(P, 45, seed=8)
    np.testing.assert_array_equal(other.points(), longer.points())
    np.testing.assert_array_equal(other.stats(), longer.stats())
    raw = bytearray((tmp_path / "s.bin").read_bytes())
    raw[20:24] = (48).to_bytes(4, "little")  # header field `lanes`: no schedule has 48 threads per chain
    (tmp_path / "s48.bin").write_bytes(bytes(raw))
    with pytest.raises(achr.AchrError):
        make(setup, 64, 2, False).load_state(str(tmp_path / "s48.bin"))
    with pytest.raises(achr.AchrError):
        make(setup, 64, 2, True).load_state(str(tmp_path / "s.bin"))  # bank sampler, ACHR state
    (tmp_path / "bad.bin").write_bytes(b"not a state file at all, nothing here to read")
    with pytest.raises(achr.AchrError):
        make(setup, 64, 2, False).load_state(str(tmp_path / "bad.bin"))


def test_batch_run_continues_and_resumes_with_its_own_lanes(setup, tmp_path):
    """A batch may run a sampler's chains with other lanes than the sampler's
    own schedule (pinned or autotuned); the lanes fix the norm's reduction
    tree, so continuing such a run, or saving it and resuming it in a fresh
    sampler, must keep them: S1 steps in the batch then S2 more equal one
    batch run of S1 + S2 steps, bit for bit, and the oracle at those lanes."""
    P, seed, file, S1, S2 = 24, 8, 1, 30, 27
    lanes = [64, 256]

    def pair():
        out = []
        for k in range(2):
            smp = achr.Sampler(setup["model"], setup["W"], achr.Options(model_id=k + 1, reproject_mode=2))
            out.append(smp)
        assert all(s.schedule()[0] == 32 for s in out)  # their own schedule: 32 lanes
        return out

    longs = pair()
    b = achr.Batch(longs)
    b.set_lanes(lanes)
    b.run([P, P], S1 + S2, seed=seed, file=file)
    want = [(s.points(), s.stats()) for s in longs]
    for k in range(2):
        o, ost, _ = O.sample(warm=setup["W"], lb=setup["lb"], ub=setup["ub"], rowptr=setup["rp"],
                             colidx=setup["ci"], vals=setup["va"], N=setup["N"], m=setup["S"].shape[0], P=P,
                             S=S1 + S2, seed=seed, file=file, model=k + 1, lanes=lanes[k], threads=8)
        np.testing.assert_array_equal(want[k][1], ost)
        np.testing.assert_array_equal(want[k][0], o)
    shorts = pair()
    b2 = achr.Batch(shorts)
    b2.set_lanes(lanes)
    b2.run([P, P], S1, seed=seed, file=file)
    path = tmp_path / "batch1.state"
    shorts[1].save_state(str(path))
    say("continue model 0 in place")
    shorts[0].continue_run(S2)
    np.testing.assert_array_equal(shorts[0].stats(), want[0][1])
    np.testing.assert_array_equal(shorts[0].points(), want[0][0])
    say("resume model 1 in a fresh sampler")
    fresh = achr.Sampler(setup["model"], setup["W"], achr.Options(model_id=2, reproject_mode=2))
    fresh.load_state(str(path))
    fresh.continue_run(S2)
    np.testing.assert_array_equal(fresh.stats(), want[1][1])
    np.testing.assert_array_equal(fresh.points(), want[1][0])
