"""Python view of the C ABI: models, null spaces, samplers.

Mirrors the reference's host flow (ACHRcu/ACHRCuda.cu:504-802): read the
model, compute the null space, read the warmup points and their centre, then
run files of P chains x S steps. Arrays cross the boundary as numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import ctypes as _ctypes
import os
import struct
from dataclasses import dataclass

import numpy as np

from ._lib import AchrLearnOptions, AchrOptions, AchrWarmupOptions, check, lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Model:
    """Stoichiometric model {x : S x = 0, lb <= x <= ub} held by libachr."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def from_coo(cls, n: int, m: int, rows, cols, vals, lb, ub) -> "Model":
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        lb = np.ascontiguousarray(lb, dtype=np.float64)
        ub = np.ascontiguousarray(ub, dtype=np.float64)
        if lb.shape != (n,) or ub.shape != (n,):
            raise ValueError("bounds must have n entries")
        h = C.c_void_p()
        check(lib().achr_model_create(n, m, len(vals), _ip(rows), _ip(cols), _dp(vals), _dp(lb), _dp(ub),
                                      C.byref(h)), "achr_model_create")
        return cls(h)

    @classmethod
    def from_dense(cls, S: np.ndarray, lb, ub) -> "Model":
        S = np.asarray(S, dtype=np.float64)
        r, c = np.nonzero(S)
        return cls.from_coo(S.shape[1], S.shape[0], r, c, S[r, c], lb, ub)

    @classmethod
    def read_mps(cls, path: str, rows: bool = False) -> "Model":
        """rows=False: the reference sampler's view (every row an equality with
        right-hand side 0). rows=True: row types, right-hand sides and ranges
        as written, carried by auxiliary columns after the file's own."""
        h = C.c_void_p()
        fn = "achr_model_read_mps_rows" if rows else "achr_model_read_mps"
        check(getattr(lib(), fn)(str(path).encode(), C.byref(h)), fn)
        return cls(h)

    @property
    def file_columns(self) -> int:
        """Leading columns that come from the file (n unless read with rows=True)."""
        nf = C.c_int()
        check(lib().achr_model_file_columns(self._h, C.byref(nf)), "achr_model_file_columns")
        return nf.value

    def complete_points(self, X_file: np.ndarray) -> np.ndarray:
        """(n_file x P) points of the file's columns -> (n x P) with the
        auxiliary columns filled: slacks take their row's activity, the
        right-hand-side column 1."""
        X_file = np.asarray(X_file, dtype=np.float64)
        nf, n = self.file_columns, self.n
        if X_file.ndim != 2 or X_file.shape[0] != nf:
            raise ValueError(f"expected {nf} rows")
        P = X_file.shape[1]
        xi = np.ascontiguousarray(X_file.T)
        xo = np.empty((P, n))
        check(lib().achr_model_complete_points(self._h, _dp(xi), P, _dp(xo)), "achr_model_complete_points")
        return np.ascontiguousarray(xo.T)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().achr_model_destroy(h)
            self._h = None

    @property
    def dims(self):
        n, m, nnz, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().achr_model_dims(self._h, C.byref(n), C.byref(m), C.byref(nnz), C.byref(d)), "dims")
        return n.value, m.value, nnz.value, d.value

    @property
    def n(self) -> int:
        return self.dims[0]

    @property
    def m(self) -> int:
        return self.dims[1]

    def bounds(self):
        n = self.n
        lb = np.empty(n)
        ub = np.empty(n)
        check(lib().achr_model_bounds(self._h, _dp(lb), _dp(ub)), "bounds")
        return lb, ub

    def csr(self):
        n, m, nnz, _ = self.dims
        rp = np.empty(m + 1, dtype=np.int32)
        ci = np.empty(max(nnz, 1), dtype=np.int32)
        va = np.empty(max(nnz, 1))
        check(lib().achr_model_csr(self._h, _ip(rp), _ip(ci), _dp(va)), "csr")
        return rp, ci[:nnz], va[:nnz]

    def dense(self) -> np.ndarray:
        n, m, _, _ = self.dims
        rp, ci, va = self.csr()
        S = np.zeros((m, n))
        for r in range(m):
            S[r, ci[rp[r]:rp[r + 1]]] = va[rp[r]:rp[r + 1]]
        return S

    def compute_nullspace(self) -> int:
        d = C.c_int()
        check(lib().achr_model_compute_nullspace(self._h, C.byref(d)), "achr_model_compute_nullspace")
        return d.value

    def set_nullspace(self, N: np.ndarray) -> None:
        N = np.asfortranarray(N, dtype=np.float64)
        n, d = N.shape
        if n != self.n:
            raise ValueError("N must have n rows")
        flat = np.ascontiguousarray(N.ravel(order="F"))
        check(lib().achr_model_set_nullspace(self._h, _dp(flat), d), "achr_model_set_nullspace")

    def nullspace(self) -> np.ndarray:
        n, _, _, d = self.dims
        flat = np.empty(max(n * d, 1))
        check(lib().achr_model_get_nullspace(self._h, _dp(flat)), "achr_model_get_nullspace")
        return flat[: n * d].reshape((n, d), order="F")


@dataclass
class Options:
    """achr_options with the reference defaults (ACHRCuda.cu:505-512, :325, :338)."""

    dTol: float = 1e-14
    uTol: float = 1e-9
    maxMinTol: float = 1e-9
    reprojTol: float = 1e-9
    check_every: int = 10
    max_misses: int = 100000
    lanes: int = 0
    reproject_mode: int = 0
    trace_chains: int = 0
    trace_thin: int = 0
    proj_prec: int = 0
    fp64_every: int = 10
    model_id: int = 0
    check_kernel: int = 0

    def to_c(self) -> AchrOptions:
        return AchrOptions(self.dTol, self.uTol, self.maxMinTol, self.reprojTol, self.check_every,
                           self.max_misses, self.lanes, self.reproject_mode, self.trace_chains, self.trace_thin,
                           self.proj_prec, self.fp64_every, self.model_id, self.check_kernel)


PROJ_FP64, PROJ_RES_BF16X3, PROJ_RES_F32, PROJ_NULL_F32, PROJ_NULL_BF16X3 = range(5)


class Sampler:
    """Device-resident sampler over a model, its null space and warmup set."""

    def __init__(self, model: Model, warmup: np.ndarray | None = None, opts: Options | None = None,
                 warmup_csv: str | None = None):
        self.model = model
        self.opts = opts or Options()
        co = self.opts.to_c()
        h = C.c_void_p()
        if warmup_csv is not None:
            check(lib().achr_sampler_create_csv(model._h, str(warmup_csv).encode(), C.byref(co), C.byref(h)),
                  "achr_sampler_create_csv")
        else:
            w = np.asfortranarray(warmup, dtype=np.float64)
            if w.shape[0] != model.n:
                raise ValueError("warmup must be n x W")
            flat = np.ascontiguousarray(w.ravel(order="F"))
            check(lib().achr_sampler_create(model._h, _dp(flat), w.shape[1], C.byref(co), C.byref(h)),
                  "achr_sampler_create")
        self._h = h
        self.n = model.n
        self.P = 0
        self.S = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().achr_sampler_destroy(h)
            self._h = None

    def center(self) -> np.ndarray:
        c = np.empty(self.n)
        check(lib().achr_sampler_center(self._h, _dp(c)), "center")
        return c

    def project_warmup(self) -> None:
        """Moves the warmup points onto {S x = 0, fixed reactions at their bound}
        and recomputes the centre (achr_sampler_project_warmup)."""
        check(lib().achr_sampler_project_warmup(self._h), "achr_sampler_project_warmup")

    def warmup(self) -> np.ndarray:
        """The sampler's warmup set, n x W."""
        W = C.c_int()
        check(lib().achr_sampler_get_warmup(self._h, None, C.byref(W)), "achr_sampler_get_warmup")
        out = np.empty((self.n, W.value), order="F")
        check(lib().achr_sampler_get_warmup(self._h, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(W)),
              "achr_sampler_get_warmup")
        return out

    def schedule(self):
        lanes, mode = C.c_int(), C.c_int()
        check(lib().achr_sampler_schedule(self._h, C.byref(lanes), C.byref(mode)), "schedule")
        return lanes.value, mode.value

    def set_stream(self, stream_ptr: int) -> None:
        check(lib().achr_sampler_set_stream(self._h, C.c_void_p(stream_ptr)), "set_stream")

    def run(self, P: int, S: int, seed: int = 0, file: int = 0, chain0: int = 0) -> None:
        """Queues one file (P chains x S accepted steps); results stay on device."""
        check(lib().achr_sample(self._h, P, S, C.c_uint64(seed), C.c_uint32(file), chain0), "achr_sample")
        self.P, self.S = P, S

    def continue_run(self, S_more: int) -> None:
        """Runs the chains of the last run (or loaded state) S_more steps further."""
        check(lib().achr_sample_continue(self._h, S_more), "achr_sample_continue")
        self.S += S_more

    def save_state(self, path: str) -> None:
        check(lib().achr_sampler_save_state(self._h, os.fsencode(path)), "achr_sampler_save_state")

    def load_state(self, path: str) -> None:
        """Loads a saved chain state (the header's P and S become this sampler's run size)."""
        check(lib().achr_sampler_load_state(self._h, os.fsencode(path)), "achr_sampler_load_state")
        with open(path, "rb") as f:
            head = f.read(20)
        _n, self.P, self.S = struct.unpack("<3i", head[8:20])

    def points(self) -> np.ndarray:
        """n x P, reaction major (the reference's File<k> layout)."""
        out = np.empty((self.n, self.P))
        check(lib().achr_sample_points(self._h, _dp(out)), "achr_sample_points")
        return out

    def stats(self) -> np.ndarray:
        st = np.empty((self.P, 4), dtype=np.int32)
        check(lib().achr_sample_stats(self._h, st.ctypes.data_as(C.POINTER(C.c_int32))), "achr_sample_stats")
        return st

    def trace(self) -> np.ndarray:
        tc = min(self.opts.trace_chains, self.P)
        L = self.S // self.opts.trace_thin if self.opts.trace_thin > 0 else 0
        tr = np.empty((tc, L, self.n))
        check(lib().achr_sample_trace(self._h, _dp(tr)), "achr_sample_trace")
        return tr

    def timing(self):
        tot, stp, prj = C.c_double(), C.c_double(), C.c_double()
        ns, np_ = C.c_int(), C.c_int()
        check(lib().achr_sample_timing(self._h, C.byref(tot), C.byref(stp), C.byref(prj), C.byref(ns),
                                       C.byref(np_)), "achr_sample_timing")
        chk, nc = C.c_double(), C.c_int()
        check(lib().achr_sample_timing_check(self._h, C.byref(chk), C.byref(nc)), "achr_sample_timing_check")
        return {"ms_total": tot.value, "ms_step": stp.value, "ms_proj": prj.value,
                "step_launches": ns.value, "proj_launches": np_.value,
                "ms_check": chk.value, "check_launches": nc.value}

    def sample(self, P: int, S: int, seed: int = 0, file: int = 0, chain0: int = 0) -> np.ndarray:
        self.run(P, S, seed, file, chain0)
        return self.points()

    def learn_directions(self, rank: int = 512, width: int = 512, iters: int = 400, batch: int = 2048,
                         lr: float = 1e-3, bank_size: int = 0, isotropic: bool = False, seed: int = 7) -> dict:
        """Learned direction bank from the last traced run; switches the sampler to it."""
        o = AchrLearnOptions(rank, width, iters, batch, lr, bank_size, 1 if isotropic else 0, seed)
        rep = np.zeros(8)
        check(lib().achr_sampler_learn_directions(self._h, C.byref(o), _dp(rep)), "achr_sampler_learn_directions")
        keys = ["free_dim", "samples", "loss_start", "loss_end", "sigma_perp2", "captured", "ms_total", "ms_train"]
        return dict(zip(keys, rep.tolist()))

    def set_directions(self, bank: np.ndarray) -> None:
        b = np.asfortranarray(bank, dtype=np.float64)
        flat = np.ascontiguousarray(b.ravel(order="F"))
        check(lib().achr_sampler_set_directions(self._h, _dp(flat), b.shape[1]), "achr_sampler_set_directions")

    def directions(self) -> np.ndarray:
        """The sampler's direction bank (n x K), empty if it has none."""
        K = C.c_int()
        check(lib().achr_sampler_get_directions(self._h, None, C.byref(K)), "achr_sampler_get_directions")
        out = np.empty((self.n, K.value), order="F")
        if K.value:
            check(lib().achr_sampler_get_directions(self._h, _dp(out), C.byref(K)), "achr_sampler_get_directions")
        return out

    def set_run_options(self, trace_chains: int, trace_thin: int, check_every: int) -> None:
        check(lib().achr_sampler_set_run_options(self._h, trace_chains, trace_thin, check_every),
              "achr_sampler_set_run_options")
        self.opts.trace_chains, self.opts.trace_thin, self.opts.check_every = trace_chains, trace_thin, check_every

    def start_point(self) -> np.ndarray:
        """Where the chains of the next run start (the centre, or the bank-mode start)."""
        out = np.empty(self.n)
        check(lib().achr_sampler_start_point(self._h, _dp(out)), "achr_sampler_start_point")
        return out

    def use_directions(self, on: bool) -> None:
        check(lib().achr_sampler_use_directions(self._h, 1 if on else 0), "achr_sampler_use_directions")


class Batch:
    """Several samplers run as one job (achr_batch_*): one persistent launch
    per schedule bucket and segment, each segment replayed from a HIP graph."""

    TPC = (16, 32, 64, 128, 256, 512, 1024)

    def __init__(self, samplers):
        self.samplers = list(samplers)
        arr = (C.c_void_p * len(self.samplers))(*[s._h.value for s in self.samplers])
        h = C.c_void_p()
        check(lib().achr_batch_create(arr, len(self.samplers), C.byref(h)), "achr_batch_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().achr_batch_destroy(h)
            self._h = None

    def _ints(self, v):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.int32), (len(self.samplers),)))
        return a, a.ctypes.data_as(C.POINTER(C.c_int))

    def autotune(self, P, probe_steps: int = 20, seed: int = 0) -> np.ndarray:
        """Times every candidate threads-per-chain per size class; returns the
        probe times (models x 7, -1 where not a candidate)."""
        _, pp = self._ints(P)
        ms = np.full((len(self.samplers), len(self.TPC)), -1.0)
        check(lib().achr_batch_autotune(self._h, pp, probe_steps, C.c_uint64(seed), _dp(ms)), "achr_batch_autotune")
        return ms

    def lanes(self) -> np.ndarray:
        out = np.zeros(len(self.samplers), dtype=np.int32)
        check(lib().achr_batch_lanes(self._h, out.ctypes.data_as(C.POINTER(C.c_int))), "achr_batch_lanes")
        return out

    def set_lanes(self, lanes) -> None:
        _, pp = self._ints(lanes)
        check(lib().achr_batch_set_lanes(self._h, pp), "achr_batch_set_lanes")

    def run(self, P, S: int, seed: int = 0, file: int = 0, graph: bool = True) -> None:
        a, pp = self._ints(P)
        check(lib().achr_batch_sample(self._h, pp, S, C.c_uint64(seed), C.c_uint32(file), 1 if graph else 0),
              "achr_batch_sample")
        for smp, p in zip(self.samplers, a.tolist()):
            smp.P, smp.S = p, S

    def timing(self) -> dict:
        ms = C.c_double()
        seg, bk, gr, nodes = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().achr_batch_timing(self._h, C.byref(ms), C.byref(seg), C.byref(bk), C.byref(gr), C.byref(nodes)),
              "achr_batch_timing")
        return {"ms_total": ms.value, "segments": seg.value, "buckets": bk.value, "graphs": gr.value,
                "graph_nodes": nodes.value}


def philox4x32_10(ctr, key) -> tuple:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().achr_philox4x32_10(c, k, o)
    return tuple(o)


def write_points_csv(path: str, points: np.ndarray) -> None:
    p = np.ascontiguousarray(points, dtype=np.float64)
    check(lib().achr_write_points_csv(str(path).encode(), _dp(p), p.shape[0], p.shape[1]), "write_points_csv")


def warmup_points(model: "Model", nPts: int, max_iters: int | None = None, tol: float | None = None,
                  seed: int | None = None, check_every: int | None = None, polish: bool | None = None):
    """Warmup set by flux variability + random-objective LPs on the GPU
    (achr_warmup_generate). Returns (W: n x nPts, one point per column;
    fva: n x 2 minimum / maximum per reaction; converged: nPts bools, one per
    LP; report dict)."""
    o = AchrWarmupOptions()
    lib().achr_warmup_options_default(C.byref(o))
    if max_iters is not None:
        o.max_iters = max_iters
    if tol is not None:
        o.tol = tol
    if seed is not None:
        o.seed = seed
    if check_every is not None:
        o.check_every = check_every
    if polish is not None:
        o.polish = 1 if polish else 0
    n = model.n
    W = np.empty((n, nPts), order="F")
    fva = np.empty((2, n))
    rep = np.empty(6)
    conv = np.empty(nPts, dtype=np.int32)
    check(lib().achr_warmup_generate(model._h, nPts, C.byref(o), W.ctypes.data_as(C.POINTER(C.c_double)), _dp(fva),
                                     conv.ctypes.data_as(C.POINTER(C.c_int)), _dp(rep)), "achr_warmup_generate")
    return W, fva.T.copy(), conv.astype(bool), {"iterations": int(rep[0]), "converged": int(rep[1]),
                                                "ms": float(rep[2]), "scaled_norm": float(rep[3]),
                                                "polished": int(rep[4]), "polish_ms": float(rep[5])}


def write_points_bin(path: str, points: np.ndarray) -> None:
    """Lossless point file: b"ACHRPT01", int32 n, int32 P, n x P float64 (reaction major)."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    check(lib().achr_write_points_bin(str(path).encode(), _dp(p), p.shape[0], p.shape[1]), "write_points_bin")


def read_points_bin(path: str) -> np.ndarray:
    ptr = C.POINTER(C.c_double)()
    n, P = C.c_int(), C.c_int()
    check(lib().achr_read_points_bin(str(path).encode(), C.byref(ptr), C.byref(n), C.byref(P)), "read_points_bin")
    try:
        return np.ctypeslib.as_array(ptr, shape=(n.value * P.value,)).reshape(n.value, P.value).copy()
    finally:
        lib().achr_free(ptr)


def read_warmup_csv(path: str, n_expected: int = 0) -> np.ndarray:
    ptr = C.POINTER(C.c_double)()
    W = C.c_int()
    check(lib().achr_read_warmup_csv(str(path).encode(), n_expected, C.byref(ptr), C.byref(W)),
          "achr_read_warmup_csv")
    try:
        # column major n x W -> numpy (n, W)
        flat = np.ctypeslib.as_array(ptr, shape=(W.value * max(n_expected, 1),)) if n_expected else None
        if flat is None:
            raise ValueError("n_expected is required")
        return flat.reshape((n_expected, W.value), order="F").copy()
    finally:
        lib().achr_free(ptr)


def gemm_f32(A, B, prec: int = 0, alpha: float = 1.0, C=None, beta: float = 0.0, transA=False, transB=False):
    """C = alpha * op(A) op(B) + beta * C on the matrix cores (prec 0 FP32, 1 BF16x3, 2 BF16).

    op(X) = X.T when trans is set; the transpose is a stride swap, not a copy."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    if transA:
        K, M = A.shape
        sa_m, sa_k = 1, M
    else:
        M, K = A.shape
        sa_m, sa_k = K, 1
    if transB:
        N, K2 = B.shape
        sb_k, sb_n = 1, K2
    else:
        K2, N = B.shape
        sb_k, sb_n = N, 1
    if K2 != K:
        raise ValueError("inner dimensions differ")
    out = np.zeros((M, N), dtype=np.float32) if C is None else np.array(C, dtype=np.float32, order="C")
    P = _ctypes.POINTER(_ctypes.c_float)
    check(lib().achr_gemm_f32(M, N, K, A.ctypes.data_as(P), sa_m, sa_k, B.ctypes.data_as(P), sb_k, sb_n,
                              out.ctypes.data_as(P), N, 1, alpha, beta, prec), "achr_gemm_f32")
    return out


def gemm_sub_f64(A, B, C, idx=None, prec: int = 0, timing: bool = False):
    """The residual projection round's kernel: a copy of C (FP64, rows x N)
    with rows idx (all when None) reduced by A[idx] @ B (A rows x K, B K x N,
    FP32 operands, prec 0 FP32 MFMA, 1 split-BF16x3). With timing, returns
    (result, kernel milliseconds)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    out = np.array(C, dtype=np.float64, order="C")
    rows, K = A.shape
    K2, N = B.shape
    if K2 != K or out.shape != (rows, N):
        raise ValueError("shapes differ")
    iv = None if idx is None else np.ascontiguousarray(idx, dtype=np.int32)
    M = rows if iv is None else int(iv.size)
    F = _ctypes.POINTER(_ctypes.c_float)
    ms = _ctypes.c_float(0.0)
    check(lib().achr_gemm_sub_f64(rows, M, N, K, A.ctypes.data_as(F), B.ctypes.data_as(F),
                                  out.ctypes.data_as(_ctypes.POINTER(_ctypes.c_double)),
                                  None if iv is None else iv.ctypes.data_as(_ctypes.POINTER(_ctypes.c_int)), prec,
                                  _ctypes.byref(ms) if timing else None),
          "achr_gemm_sub_f64")
    return (out, float(ms.value)) if timing else out


def bank_reciprocal(u) -> np.ndarray:
    """The device reciprocal of the direction-bank chord, elementwise."""
    u = np.ascontiguousarray(u, dtype=np.float64).ravel()
    r = np.empty_like(u)
    check(lib().achr_bank_reciprocal(_dp(u), _dp(r), u.size), "achr_bank_reciprocal")
    return r


def quotient(a, b) -> np.ndarray:
    """The device's division replacement of the plain chord, elementwise a / b."""
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    b = np.ascontiguousarray(b, dtype=np.float64).ravel()
    if a.shape != b.shape:
        raise ValueError("a and b differ in size")
    q = np.empty_like(a)
    check(lib().achr_quotient(_dp(a), _dp(b), _dp(q), a.size), "achr_quotient")
    return q


def lp_polish(N, lb, ub, c, xfeas, x0, max_iters: int = 0):
    """achr_lp_polish: min c^T x over {x = N z, lb <= x <= ub}; returns (status, x)."""
    N = np.asfortranarray(N, dtype=np.float64)
    n, d = N.shape
    flat = np.ascontiguousarray(N.ravel(order="F")) if d else np.zeros(1)
    x = np.array(x0, dtype=np.float64)
    args = [np.ascontiguousarray(v, dtype=np.float64) for v in (lb, ub, c, xfeas)]
    rc = lib().achr_lp_polish(n, d, _dp(flat), *[_dp(v) for v in args], _dp(x), max_iters or 50 * d + 1000)
    return rc, x


def _fp(a):
    return a.ctypes.data_as(_ctypes.POINTER(_ctypes.c_float))


class DirNet:
    """The direction network (include/achr.h achr_dirnet_*)."""

    def __init__(self, latent: int, width: int, outdim: int, seed: int = 0):
        h = C.c_void_p()
        check(lib().achr_dirnet_create(latent, width, outdim, C.c_uint64(seed), C.byref(h)), "achr_dirnet_create")
        self._h = h
        self.L, self.H, self.O = latent, width, outdim

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().achr_dirnet_destroy(h)
            self._h = None

    def param_count(self) -> int:
        c = C.c_long()
        check(lib().achr_dirnet_param_count(self._h, C.byref(c)), "param_count")
        return c.value

    def params(self) -> np.ndarray:
        p = np.empty(self.param_count(), dtype=np.float32)
        check(lib().achr_dirnet_get_params(self._h, _fp(p)), "get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().achr_dirnet_set_params(self._h, _fp(p)), "set_params")

    def set_target(self, Ct: np.ndarray) -> None:
        Ct = np.ascontiguousarray(Ct, dtype=np.float32)
        check(lib().achr_dirnet_set_target(self._h, _fp(Ct)), "set_target")

    def loss_grad(self, Z: np.ndarray):
        Z = np.ascontiguousarray(Z, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(self.param_count(), dtype=np.float32)
        check(lib().achr_dirnet_loss_grad(self._h, _fp(Z), Z.shape[0], _fp(loss), _fp(g)), "loss_grad")
        return float(loss[0]), g

    def train(self, iters: int, batch: int, lr: float, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(iters, 1), dtype=np.float32)
        check(lib().achr_dirnet_train(self._h, iters, batch, lr, C.c_uint64(seed), _fp(losses)), "train")
        return losses[:iters]

    def sample(self, Z: np.ndarray, prec: int = 0) -> np.ndarray:
        Z = np.ascontiguousarray(Z, dtype=np.float32)
        V = np.empty((Z.shape[0], self.O), dtype=np.float32)
        check(lib().achr_dirnet_sample(self._h, _fp(Z), Z.shape[0], prec, _fp(V)), "sample")
        return V
