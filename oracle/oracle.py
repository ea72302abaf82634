"""ctypes binding of the CPU oracle (oracle/liboracle.so).

ORACLE — test infrastructure only: imported by tests/, by the smoke check and
by the CPU-baseline leg of bench.py, as the checker. The shipped library never
uses it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"oracle not built: make -C {_HERE}")
        h = C.CDLL(LIB_PATH)
        PD = C.POINTER(C.c_double)
        PI = C.POINTER(C.c_int)
        h.oracle_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        h.oracle_philox4x32_10.restype = None
        h.oracle_stream_u01.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        h.oracle_stream_u01.restype = C.c_double
        h.oracle_center.argtypes = [PD, C.c_int, C.c_int, PD]
        h.oracle_center.restype = None
        h.oracle_nullspace.argtypes = [PD, C.c_int, C.c_int, PD]
        h.oracle_nullspace.restype = C.c_int
        h.oracle_sample.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, PD, PD, PD, PD, PI, PI, PD, PD,
                                    C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32,
                                    C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_int, PD, C.POINTER(C.c_int32), PD, C.c_int, C.c_int, C.c_int,
                                    PD, C.c_int]
        h.oracle_sample.restype = C.c_int
        _lib = h
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return tuple(o)


def stream_u01(seed, index, chain, file=0, model=0, tag=0):
    return lib().oracle_stream_u01(seed, index, chain, file, model, tag)


def center(W: np.ndarray) -> np.ndarray:
    W = np.asfortranarray(W, dtype=np.float64)
    n, k = W.shape
    flat = np.ascontiguousarray(W.ravel(order="F"))
    c = np.empty(n)
    lib().oracle_center(_dp(flat), n, k, _dp(c))
    return c


def nullspace(S: np.ndarray) -> np.ndarray:
    S = np.asarray(S, dtype=np.float64)
    m, n = S.shape
    flat = np.ascontiguousarray(S.ravel(order="F"))
    N = np.empty(n * n)
    d = lib().oracle_nullspace(_dp(flat), m, n, _dp(N))
    return N[: n * d].reshape((n, d), order="F")


def csr_of(S: np.ndarray):
    m, n = S.shape
    rp = [0]
    ci, va = [], []
    for r in range(m):
        nz = np.nonzero(S[r])[0]
        ci.extend(nz.tolist())
        va.extend(S[r, nz].tolist())
        rp.append(len(ci))
    return (np.array(rp, dtype=np.int32), np.array(ci if ci else [0], dtype=np.int32),
            np.array(va if va else [0.0]))


def sample(*, warm, lb, ub, rowptr, colidx, vals, N, m, P, S, seed=0, file=0, model=0, chain0=0, lanes=64,
           check_every=10, max_misses=100000, dTol=1e-14, uTol=1e-9, maxMinTol=1e-9, reprojTol=1e-9,
           threads=1, b_lo=0, b_hi=None, trace_chains=0, trace_thin=0, c0=None, bank=None):
    """Runs chains [b_lo, b_hi) of a P-chain file; returns (points n x P, stats P x 4, trace)."""
    warm = np.asfortranarray(warm, dtype=np.float64)
    n, W = warm.shape
    wflat = np.ascontiguousarray(warm.ravel(order="F"))
    if c0 is None:
        c0 = center(warm)
    N = np.zeros((n, 0)) if N is None else np.asfortranarray(N, dtype=np.float64)
    d = N.shape[1]
    Nflat = np.ascontiguousarray(N.ravel(order="F")) if d else np.zeros(1)
    lb = np.ascontiguousarray(lb, dtype=np.float64)
    ub = np.ascontiguousarray(ub, dtype=np.float64)
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    colidx = np.ascontiguousarray(colidx, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    if b_hi is None:
        b_hi = P
    out = np.zeros((n, P))
    stats = np.zeros((P, 4), dtype=np.int32)
    L = S // trace_thin if trace_thin > 0 else 0
    trace = np.zeros((max(trace_chains, 1), max(L, 1), n)) if trace_chains and L else None
    if bank is not None:
        bank = np.asfortranarray(bank, dtype=np.float64)
        bflat = np.ascontiguousarray(bank.ravel(order="F"))
        K = bank.shape[1]
    else:
        bflat, K = None, 0
    rc = lib().oracle_sample(n, m, W, d, _dp(wflat), _dp(lb), _dp(ub), _dp(np.ascontiguousarray(c0)),
                             rowptr.ctypes.data_as(C.POINTER(C.c_int)), colidx.ctypes.data_as(C.POINTER(C.c_int)),
                             _dp(vals), _dp(Nflat), P, S, b_lo, b_hi, seed, file, model, chain0, lanes,
                             check_every, max_misses, dTol, uTol, maxMinTol, reprojTol, threads, _dp(out),
                             stats.ctypes.data_as(C.POINTER(C.c_int32)),
                             _dp(trace) if trace is not None else None, trace_chains, trace_thin, L,
                             _dp(bflat) if bflat is not None else None, K)
    if rc != 0:
        raise ValueError(f"oracle_sample rejected its arguments (lanes={lanes})")
    return out, stats, trace
