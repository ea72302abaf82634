"""achr_sample_noalloc (the entry point for callers that cannot allocate
device memory) against achr_sample on the same sampler, through the C ABI:
two calls in a row, the second with fewer chains and steps, each bit-equal
(points and counters) to achr_sample's run, in-kernel and batched projection,
on the 72 x 95 model's "%f" warmup (every check reprojects) and on the
genome-sized schedule (1,024 lanes, 11 slots). A run larger than the static
scratch is refused with ACHR_E_ARG before anything is queued."""
import ctypes as C

import numpy as np
import pytest

import achr
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


class DeviceBuffer:
    """Device memory of the caller (the outputs), through the HIP runtime."""

    def __init__(self, nbytes):
        achr.lib()
        self.hip = C.CDLL("libamdhip64.so")
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(max(nbytes, 8))) == 0
        self.nbytes = nbytes

    def fetch(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def __del__(self):
        if getattr(self, "p", None) and self.p.value:
            self.hip.hipFree(self.p)


def noalloc(smp, P, S, seed, file=0, chain0=0):
    n = smp.n
    out = DeviceBuffer(8 * n * P)
    st = DeviceBuffer(16 * P)
    rc = achr.lib().achr_sample_noalloc(smp._h, P, S, C.c_uint64(seed), C.c_uint32(file), chain0, out.p, st.p)
    assert rc == 0, achr.lib().achr_last_error()
    return out.fetch((n, P), np.float64), st.fetch((P, 4), np.int32)


@pytest.mark.parametrize("lanes,mode", [(32, 1), (32, 2), (256, 2)])
def test_noalloc_equals_sample_twice(c1, lanes, mode):
    S, lb, ub, Wf, _ = c1
    model = achr.Model.from_dense(S, lb, ub)
    model.set_nullspace(O.nullspace(S))
    smp = achr.Sampler(model, Wf, achr.Options(lanes=lanes, reproject_mode=mode))
    for P, Ss, seed, file in ((300, 120, 5, 2), (37, 41, 6, 0)):
        say(f"lanes={lanes} mode={mode} P={P} S={Ss}")
        g, gs = noalloc(smp, P, Ss, seed, file)
        want = smp.sample(P, Ss, seed=seed, file=file)
        np.testing.assert_array_equal(gs, smp.stats())
        np.testing.assert_array_equal(g, want)
        assert gs[:, 1].sum() > 0  # the "%f" warmup reprojects


def test_noalloc_genome_schedule_and_refusal():
    """A 4,900 x 9,000 model (1,024 lanes, 9 slots, batched FP64 rounds): 200
    then 64 chains, bit-equal to achr_sample; 400 chains exceed the scratch."""
    from achr import synth

    def ns(rows, cols, vals, m, n):
        mdl = achr.Model.from_coo(n, m, rows, cols, vals, -np.ones(n), np.ones(n))
        mdl.compute_nullspace()
        return mdl.nullspace()

    sm = synth.make_model(4900, 9000, 21, irrev_frac=0.3, n_fixed=90, width=(5.0, 50.0), nullspace_fn=ns)
    W = synth.round_like_csv(synth.chord_warmup(sm.N, sm.x0, sm.lb, sm.ub, 600, seed=22, mix=8))
    model = achr.Model.from_coo(sm.n, sm.m, sm.rows, sm.cols, sm.vals, sm.lb, sm.ub)
    model.set_nullspace(sm.N)
    smp = achr.Sampler(model, W, achr.Options(reproject_mode=2))
    assert smp.schedule()[0] == 1024
    for P, Ss in ((200, 35), (64, 22)):
        say(f"genome schedule P={P} S={Ss}")
        g, gs = noalloc(smp, P, Ss, 9)
        want = smp.sample(P, Ss, seed=9)
        np.testing.assert_array_equal(gs, smp.stats())
        np.testing.assert_array_equal(g, want)
    out = DeviceBuffer(8)
    rc = achr.lib().achr_sample_noalloc(smp._h, 400, 10, C.c_uint64(1), C.c_uint32(0), 0, out.p, None)
    assert rc == -1 and b"scratch" in achr.lib().achr_last_error()
