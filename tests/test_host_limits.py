"""Entry points whose host buffers are sized by their inputs report a failed
allocation as ACHR_E_NOMEM instead of letting the C++ exception escape the C
ABI (which would end the calling process). Each call runs in a child process
whose address space is capped (RLIMIT_AS) a few MiB above what it already
maps: the MPS reader, the warmup CSV reader, a model from triplets with 50
million rows, and the null space of a 30,000-reaction model (its dense
S^T S alone is 7.2 GB; the allocation fails before any device call). The
reference exits or asserts instead (ACHRcu/Utilities.cu:28-37, :76-85)."""
import os
import subprocess
import sys
import textwrap

import achr

CHILD = textwrap.dedent(r'''
    import ctypes as C, resource, sys
    sys.path.insert(0, sys.argv[1])
    import achr
    L = achr.lib()
    which, path = sys.argv[2], sys.argv[3]

    def vm_bytes():
        for line in open("/proc/self/status"):
            if line.startswith("VmSize:"):
                return int(line.split()[1]) * 1024

    h = C.c_void_p()
    args = None
    if which == "nullspace":
        r = (C.c_int * 1)(0); c = (C.c_int * 1)(0); v = (C.c_double * 1)(1.0)
        n = 30000
        lb = (C.c_double * n)(*([-1.0] * n)); ub = (C.c_double * n)(*([1.0] * n))
        assert L.achr_model_create(n, 1, 1, r, c, v, lb, ub, C.byref(h)) == 0
    w, W = C.POINTER(C.c_double)(), C.c_int()
    cap = vm_bytes() + (16 << 20)
    soft, hard = resource.getrlimit(resource.RLIMIT_AS)
    resource.setrlimit(resource.RLIMIT_AS, (cap, hard))
    if which == "mps":
        rc = L.achr_model_read_mps(path.encode(), C.byref(h))
    elif which == "csv":
        rc = L.achr_read_warmup_csv(path.encode(), 0, C.byref(w), C.byref(W))
    elif which == "triplets":
        z = C.c_double(0.0)
        rc = L.achr_model_create(1, 50000000, 0, None, None, None, C.byref(z), C.byref(z), C.byref(h))
    else:
        d = C.c_int()
        rc = L.achr_model_compute_nullspace(h, C.byref(d))
    resource.setrlimit(resource.RLIMIT_AS, (soft, hard))
    print("RC", rc, L.achr_last_error().decode(), flush=True)
''')


def _run(which, path, tmp_path):
    drv = tmp_path / "child.py"
    drv.write_text(CHILD)
    pkg = os.path.dirname(os.path.dirname(achr.__file__))
    r = subprocess.run([sys.executable, str(drv), pkg, which, str(path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("RC ")][-1]
    return int(line.split()[1]), line


def test_mps_reader_reports_nomem(tmp_path):
    p = tmp_path / "big.mps"
    with open(p, "w") as f:
        f.write("NAME big\nROWS\n N obj\n")
        f.writelines(f" E R{i}\n" for i in range(50000))
        f.write("COLUMNS\n")
        f.writelines(f"    C{j} R{j % 50000} 1 R{(j * 7) % 50000} -1\n" for j in range(600000))
        f.write("RHS\nBOUNDS\nENDATA\n")
    rc, line = _run("mps", p, tmp_path)
    assert rc == -7, line


def test_csv_reader_reports_nomem(tmp_path):
    p = tmp_path / "big.csv"
    row = ",".join(["1.5"] * 20000) + "\n"
    with open(p, "w") as f:
        f.writelines(row for _ in range(400))
    rc, line = _run("csv", p, tmp_path)
    assert rc == -7, line


def test_model_from_triplets_reports_nomem(tmp_path):
    rc, line = _run("triplets", "-", tmp_path)
    assert rc == -7, line


def test_nullspace_reports_nomem_before_touching_a_device(tmp_path):
    rc, line = _run("nullspace", "-", tmp_path)
    assert rc == -7, line
