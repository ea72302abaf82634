"""The host code that reads untrusted input, under AddressSanitizer and
UndefinedBehaviorSanitizer: the MPS reader (both views of the rows), the
warmup CSV reader, the point files and the model set-up of the C ABI (the
host-only sources, achr_io.cpp), and the CPU oracle. The reference only
exits on errors (ACHRcu/Utilities.cu:28-37) and reads its MPS file through
GLPK/CPLEX (ACHRCuda.cu:563, :601) and its CSV with fixed buffers
(:387-412); here every malformed input must come back as an error code, with
no sanitizer report.

Both libraries are built with -fsanitize=address,undefined (Makefile targets
`sanitize`) and loaded into a child Python process with the sanitizer
runtime preloaded; the child runs the fixtures, a set of hand-made malformed
files and 300 seeded mutations (byte flips, truncations, spliced lines) of
the 72 x 95 model and its warmup file."""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IO_ASAN = os.path.join(ROOT, "achr.cu_amd", "build", "asan", "libachr_io_asan.so")
ORACLE_ASAN = os.path.join(ROOT, "oracle", "liboracle_asan.so")


def _runtime():
    p = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def asan_env():
    rt = _runtime()
    if rt is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    for d in (os.path.join(ROOT, "achr.cu_amd"), os.path.join(ROOT, "oracle")):
        subprocess.run(["make", "-s", "-C", d, "sanitize"], check=True, capture_output=True)
    env = dict(os.environ, LD_PRELOAD=rt, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", PYTHONPATH=os.path.join(ROOT, "oracle"))
    return env


DRIVER = textwrap.dedent(r'''
    import ctypes as C, os, random, sys
    import numpy as np
    io_path, golden, tmp = sys.argv[1], sys.argv[2], sys.argv[3]
    L = C.CDLL(io_path)
    P, PD, PI = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.achr_model_read_mps.argtypes = [C.c_char_p, C.POINTER(P)]
    L.achr_model_read_mps_rows.argtypes = [C.c_char_p, C.POINTER(P)]
    L.achr_model_dims.argtypes = [P, PI, PI, PI, PI]
    L.achr_model_destroy.argtypes = [P]
    L.achr_model_destroy.restype = None
    L.achr_model_file_columns.argtypes = [P, PI]
    L.achr_model_complete_points.argtypes = [P, PD, C.c_int, PD]
    L.achr_model_create.argtypes = [C.c_int, C.c_int, C.c_int, PI, PI, PD, PD, PD, C.POINTER(P)]
    L.achr_read_warmup_csv.argtypes = [C.c_char_p, C.c_int, C.POINTER(PD), PI]
    L.achr_free.argtypes = [P]
    L.achr_free.restype = None
    L.achr_read_points_bin.argtypes = [C.c_char_p, C.POINTER(PD), PI, PI]
    L.achr_write_points_bin.argtypes = [C.c_char_p, PD, C.c_int, C.c_int]
    L.achr_write_points_csv.argtypes = [C.c_char_p, PD, C.c_int, C.c_int]
    L.achr_write_fva_csv.argtypes = [C.c_char_p, P, PD]
    L.achr_model_set_nullspace.argtypes = [P, PD, C.c_int]
    L.achr_model_get_nullspace.argtypes = [P, PD]
    L.achr_last_error.restype = C.c_char_p
    counts = {}

    def note(kind, rc):
        counts.setdefault(kind, {}).setdefault(rc, 0)
        counts[kind][rc] += 1

    def model_ops(h):
        n, m, nnz, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        assert L.achr_model_dims(h, C.byref(n), C.byref(m), C.byref(nnz), C.byref(d)) == 0
        nf = C.c_int()
        L.achr_model_file_columns(h, C.byref(nf))
        x = np.ones(max(nf.value, 1) * 2)
        y = np.empty(n.value * 2)
        L.achr_model_complete_points(h, x.ctypes.data_as(PD), 2, y.ctypes.data_as(PD))
        fva = np.zeros(2 * n.value)
        L.achr_write_fva_csv(os.path.join(tmp, "fva.csv").encode(), h, fva.ctypes.data_as(PD))
        N = np.eye(n.value)[:, : min(3, n.value)].ravel(order="F").copy()
        assert L.achr_model_set_nullspace(h, N.ctypes.data_as(PD), min(3, n.value)) == 0
        out = np.empty(N.size)
        assert L.achr_model_get_nullspace(h, out.ctypes.data_as(PD)) == 0
        return n.value, m.value

    def read_mps(path, kind):
        for fn in (L.achr_model_read_mps, L.achr_model_read_mps_rows):
            h = P()
            rc = fn(path.encode(), C.byref(h))
            note(kind, rc)
            if rc == 0:
                dims = model_ops(h)
                L.achr_model_destroy(h)
            else:
                assert not h.value and L.achr_last_error()
        return rc

    def read_csv(path, kind, n_expected=0):
        w, W = PD(), C.c_int()
        rc = L.achr_read_warmup_csv(path.encode(), n_expected, C.byref(w), C.byref(W))
        note(kind, rc)
        if rc == 0:
            s = sum(w[i] for i in range(0, W.value * max(n_expected, 1), 97))
            L.achr_free(w)
        return rc

    def write(name, data):
        p = os.path.join(tmp, name)
        with open(p, "wb") as f:
            f.write(data)
        return p

    # fixtures
    for f in ("P_Putida.mps", "c1.mps"):
        assert read_mps(os.path.join(golden, f), "fixture") == 0
    assert read_csv(os.path.join(golden, "c1_warmup.csv"), "fixture", 95) == 0
    mps = open(os.path.join(golden, "c1.mps"), "rb").read()
    csv = open(os.path.join(golden, "c1_warmup.csv"), "rb").read()

    # hand-made malformed models
    bad_mps = [b"", b"\0\0\0", b"NAME x\nROWS\n N obj\n", b"ROWS\n Q r\nENDATA\n", b"FOO\nENDATA\n",
               b"ROWS\n E r\nCOLUMNS\n x r 1 2\nENDATA\n", b"ROWS\n E r\nCOLUMNS\n x q 1\nENDATA\n",
               b"ROWS\n E r\nCOLUMNS\n x r 1e99999\nENDATA\n", b"ROWS\n E r\nCOLUMNS\n x r 1\nBOUNDS\n UP B y 1\nENDATA\n",
               b"ROWS\n E r\nCOLUMNS\n x r 1\nBOUNDS\n UP B x\nENDATA\n", b"ROWS\n E r\nCOLUMNS\n x r 1\nBOUNDS\n ZZ B x 1\nENDATA\n",
               b"ROWS\n E r\n L s\n G t\nCOLUMNS\n x r 1 s 1\n x t 1\nRHS\n R r 1 s 2\n t 3\nRANGES\n R r -1 s 2\n R t 1\n"
               b"BOUNDS\n UP B x 5\n MI B x\nENDATA\n",
               b"ROWS\n L s\nCOLUMNS\n x s 1\nBOUNDS\n FR B x\nENDATA\n",
               b"ROWS\n E r\nCOLUMNS\n x r 1\n" + b" " * 100000 + b"y r 2\nENDATA\n",
               b"ROWS\n E r\nCOLUMNS\n" + b"".join(b" x%d r 1\n" % i for i in range(3000)) + b"RHS\n R r 5\nENDATA\n",
               bytes(random.Random(1).getrandbits(8) for _ in range(20000))]
    for k, data in enumerate(bad_mps):
        read_mps(write(f"bad{k}.mps", data), "bad_mps")

    bad_csv = [b"", b"\n\n", b"1,2\n3\n", b"1,x\n", b"1,,2\n", b"1,2,\n", b"1\0,2\n", b"1.5\r\n2.5\r\n",
               b",".join(b"1.0" for _ in range(400000)) + b"\n" + b",".join(b"2.0" for _ in range(400000)) + b"\n",
               b"nan,inf,-inf\n1e308,1e-320,-0\n", bytes(random.Random(2).getrandbits(8) for _ in range(20000))]
    for k, data in enumerate(bad_csv):
        read_csv(write(f"bad{k}.csv", data), "bad_csv")
        read_csv(write(f"bad{k}.csv", data), "bad_csv", 2)

    # seeded mutations of the fixtures
    rng = random.Random(7)
    for k in range(150):
        for src, fn in ((mps, read_mps), (csv, read_csv)):
            b = bytearray(src)
            op = k % 3
            if op == 0:
                for _ in range(1 + rng.randrange(8)):
                    b[rng.randrange(len(b))] = rng.randrange(256)
            elif op == 1:
                b = b[: rng.randrange(len(b))]
            else:
                lines = bytes(b).split(b"\n")
                i, j = rng.randrange(len(lines)), rng.randrange(len(lines))
                lines.insert(i, lines[j])
                b = bytearray(b"\n".join(lines))
            fn(write(f"mut{k}.{'mps' if src is mps else 'csv'}", bytes(b)), "mutated")

    # point files
    pts = np.arange(12.0).reshape(3, 4)
    p = os.path.join(tmp, "p.bin")
    assert L.achr_write_points_bin(p.encode(), pts.ctypes.data_as(PD), 3, 4) == 0
    assert L.achr_write_points_csv(os.path.join(tmp, "p.csv").encode(), pts.ctypes.data_as(PD), 3, 4) == 0
    raw = open(p, "rb").read()
    for k, data in enumerate([raw, raw[:20], raw[:-1], b"ACHRPT01" + np.array([2**31 - 1, 2**31 - 1], np.int32).tobytes(),
                              b"ACHRPT01" + np.array([-1, 4], np.int32).tobytes(), b"XXXXXXXX" + raw[8:]]):
        out, n, Pp = PD(), C.c_int(), C.c_int()
        rc = L.achr_read_points_bin(write(f"pt{k}.bin", data).encode(), C.byref(out), C.byref(n), C.byref(Pp))
        note("points", rc)
        if rc == 0:
            assert [out[i] for i in range(12)] == list(range(12))
            L.achr_free(out)

    # models from triplets, out-of-range entries and sizes
    r = np.array([0, 1, 5], np.int32); c = np.array([0, 1, 1], np.int32); v = np.ones(3); lb = -np.ones(2); ub = np.ones(2)
    for (n, m, nnz) in ((2, 2, 2), (2, 2, 3), (0, 2, 0), (2, -1, 0), (2, 2, -1)):
        h = P()
        rc = L.achr_model_create(n, m, nnz, r.ctypes.data_as(PI), c.ctypes.data_as(PI), v.ctypes.data_as(PD),
                                 lb.ctypes.data_as(PD), ub.ctypes.data_as(PD), C.byref(h))
        note("triplets", rc)
        if rc == 0:
            model_ops(h)
            L.achr_model_destroy(h)

    # the oracle: plain and bank-mode chains, null space, centre, Philox
    import oracle as O
    O.LIB_PATH = sys.argv[4]
    sys.path.insert(0, os.path.dirname(golden))
    from conftest import load_c1
    S, lb, ub, Wf, We = load_c1()
    N = O.nullspace(S)
    rp, ci, va = O.csr_of(S)
    for lanes, W in ((32, Wf), (1024, We)):
        pts, st, tr = O.sample(warm=W, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=6, S=60,
                               seed=3, lanes=lanes, trace_chains=2, trace_thin=7)
        assert np.isfinite(pts).all()
    bank = N[:, :5] / np.linalg.norm(N[:, :5], axis=0)
    O.sample(warm=Wf, lb=lb, ub=ub, rowptr=rp, colidx=ci, vals=va, N=N, m=S.shape[0], P=3, S=40, seed=1, lanes=16,
             bank=bank, c0=O.center(We))
    O.philox((1, 2, 3, 4), (5, 6))
    # the warmup generator's exact LP finish (host code), on the 72 x 95 model
    L.achr_lp_polish.argtypes = [C.c_int, C.c_int, PD, PD, PD, PD, PD, PD, C.c_int]
    Nf = np.ascontiguousarray(N.ravel(order="F"))
    xf = O.center(We)
    for i in range(0, S.shape[1], 7):
        for sgn in (1.0, -1.0):
            c = np.zeros(S.shape[1])
            c[i] = sgn
            x = xf + 3.0 * np.cos(np.arange(S.shape[1]))
            rc = L.achr_lp_polish(S.shape[1], N.shape[1], Nf.ctypes.data_as(PD), lb.ctypes.data_as(PD),
                                  ub.ctypes.data_as(PD), c.ctypes.data_as(PD), xf.ctypes.data_as(PD),
                                  x.ctypes.data_as(PD), 5000)
            note("lp_polish", rc)
    print("SANITIZED OK", counts, flush=True)
''')


def test_host_readers_and_oracle_under_asan_ubsan(asan_env, tmp_path):
    drv = tmp_path / "drive.py"
    drv.write_text(DRIVER)
    r = subprocess.run([sys.executable, str(drv), IO_ASAN, os.path.join(ROOT, "tests", "golden"), str(tmp_path),
                        ORACLE_ASAN], env=asan_env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-6000:]
    assert r.returncode == 0 and "SANITIZED OK" in r.stdout, out[-6000:]
    print(r.stdout.strip().splitlines()[-1])
