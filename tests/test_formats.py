"""Host-side boundary: the C ABI loads and exports what include/achr.h declares,
and the MPS / CSV readers and writers (host code, no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import achr
from achr import _lib
from conftest import ROOT, golden, parse_mps_py


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "achr.h")).read()
    declared = set(re.findall(r"\b(achr_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    L = achr.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert b"gfx950" in L.achr_version()


def test_options_defaults_are_the_reference_tolerances():
    o = _lib.AchrOptions()
    achr.lib().achr_options_default(ctypes.byref(o))
    assert (o.dTol, o.uTol, o.maxMinTol, o.reprojTol, o.check_every) == (1e-14, 1e-9, 1e-9, 1e-9, 10)


def test_library_philox_matches_kat():
    assert achr.philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)
    rows = [l.split() for l in open(golden("philox_rocrand.txt")) if not l.startswith("#")]
    for r in rows[:16]:
        v = [int(t, 16) for t in r]
        assert achr.philox4x32_10(tuple(v[0:4]), tuple(v[4:6])) == tuple(v[6:10])


def test_read_mps_p_putida_counts():
    """The model the reference ships (VFWarmup/lib/P_Putida.mps): 911 E rows +
    the objective, 1,060 columns; 1,040 upper bounds at 1000 and 20 at 0; lower
    bounds 347 at -1000, 11 at -100, one at -20, one at -6.3, the rest 0."""
    m = achr.Model.read_mps(golden("P_Putida.mps"))
    n, mm, nnz, d = m.dims
    assert (n, mm, d) == (1060, 911, -1)
    lb, ub = m.bounds()
    assert (ub == 1000).sum() == 1040 and (ub == 0).sum() == 20
    assert (lb == -1000).sum() == 347 and (lb == -100).sum() == 11
    assert (lb == -20).sum() == 1 and (lb == -6.3).sum() == 1 and (lb == 0).sum() == 700
    S, L, U = parse_mps_py(golden("P_Putida.mps"))
    np.testing.assert_array_equal(m.dense(), S)
    np.testing.assert_array_equal(lb, L)
    np.testing.assert_array_equal(ub, U)
    assert nnz == np.count_nonzero(S)


def test_read_mps_c1_matches_python_parse(c1):
    S, lb, ub, _, _ = c1
    m = achr.Model.read_mps(golden("c1.mps"))
    np.testing.assert_array_equal(m.dense(), S)
    L, U = m.bounds()
    np.testing.assert_array_equal(L, lb)
    np.testing.assert_array_equal(U, ub)
    assert (lb == ub).sum() == 5


def test_mps_bound_types_and_defaults(tmp_path):
    p = tmp_path / "b.mps"
    p.write_text("NAME x\nROWS\n N  obj\n E  r1\n N  free2\nCOLUMNS\n    a  r1  1  obj  3\n    b  r1  -2\n"
                 "    c  free2  1\n    d  r1  1\nRHS\n    rhs  r1  5\nBOUNDS\n UP BND a -4\n FR BND b\n MI BND c\n"
                 " FX BND d 2.5\nENDATA\n")
    m = achr.Model.read_mps(str(p))
    lb, ub = m.bounds()
    big = np.finfo(float).max
    np.testing.assert_array_equal(lb, [-big, -big, -big, 2.5])
    np.testing.assert_array_equal(ub, [-4, big, big, 2.5])
    # first N row is the objective; a later N row stays a (free) row of S; RHS is ignored
    np.testing.assert_array_equal(m.dense(), [[1, -2, 0, 1], [0, 0, 1, 0]])


@pytest.mark.parametrize("text,code", [
    ("ROWS\n E r\nCOLUMNS\n    a  q  1\nENDATA\n", -3),       # unknown row
    ("ROWS\n E r\nCOLUMNS\n    a  r  x1\nENDATA\n", -3),      # bad number
    ("ROWS\n E r\nCOLUMNS\n    a  r  1\n", -3),               # no ENDATA
])
def test_mps_errors(tmp_path, text, code):
    p = tmp_path / "bad.mps"
    p.write_text(text)
    with pytest.raises(achr.AchrError) as e:
        achr.Model.read_mps(str(p))
    assert e.value.code == code
    with pytest.raises(achr.AchrError) as e:
        achr.Model.read_mps(str(tmp_path / "missing.mps"))
    assert e.value.code == -2


def test_warmup_csv_reader(c1):
    _, _, _, Wf, We = c1
    W = achr.read_warmup_csv(golden("c1_warmup_exact.csv"), 95)
    np.testing.assert_array_equal(W, We)
    W = achr.read_warmup_csv(golden("c1_warmup.csv"), 95)
    np.testing.assert_array_equal(W, Wf)


def test_warmup_csv_errors(tmp_path):
    p = tmp_path / "w.csv"
    p.write_text("1,2,3\n4,5\n")
    with pytest.raises(achr.AchrError) as e:
        achr.read_warmup_csv(str(p), 2)
    assert e.value.code == -3
    p.write_text("1,2\n3,4\n")
    with pytest.raises(achr.AchrError):
        achr.read_warmup_csv(str(p), 3)  # reaction count mismatch


def test_warmup_csv_reader_in_parallel_and_its_first_error(tmp_path):
    """Files of many lines are parsed by several threads: the values equal
    numpy's parse bit for bit (empty lines, CR LF and blanks around values
    as the reference's reader takes them), and of several bad lines spread
    over the threads' blocks the first in file order is reported."""
    rng = np.random.default_rng(4)
    X = rng.standard_normal((700, 37)) * 10.0 ** rng.integers(-6, 6, (700, 37))
    lines = [",".join(repr(float(v)) for v in r) for r in X]
    lines[5] = " " + lines[5].replace(",", " ,\t", 3)
    p = tmp_path / "w.csv"
    p.write_bytes(("\r\n".join(lines[:300]) + "\n\n" + "\n".join(lines[300:]) + "\n").encode())
    W = achr.read_warmup_csv(str(p), 700)
    assert W.view(np.uint64).tolist() == X.view(np.uint64).tolist()
    bad = list(lines)
    bad[650] = bad[650] + ",oops"
    bad[420] = bad[420].rsplit(",", 1)[0]
    bad[444] = "x" + bad[444]
    p.write_text("\n".join(bad) + "\n")
    with pytest.raises(achr.AchrError) as e:
        achr.read_warmup_csv(str(p), 700)
    assert "line 421 has 36 values, expected 37" in str(e.value)


def test_points_csv_writer_layout(tmp_path):
    pts = np.array([[1.0, -2.5, 3.1234567], [0.0, 1e-7, -0.0000004]])
    p = tmp_path / "File0"
    achr.write_points_csv(str(p), pts)
    lines = p.read_text().splitlines()
    assert lines == ["1.000000,-2.500000,3.123457", "0.000000,0.000000,-0.000000"]


@pytest.mark.parametrize("n,P", [(3, 5), (37, 701), (1021, 64)])
def test_points_csv_writer_bytes_equal_sequential_printf(tmp_path, n, P):
    """The writer formats rows on several threads in batches; the file must
    be byte for byte the sequential "%f" text (the reference's File<k>,
    ACHRCuda.cu:744-756), whatever the row count against the thread count,
    with signed zeros, subnormals, huge values and infinities."""
    rng = np.random.default_rng(n)
    pts = rng.standard_normal((n, P)) * 10.0 ** rng.integers(-8, 8, (n, P))
    pts[0, :3] = [-0.0, 5e-324, -1e308][:min(3, P)]
    pts[-1, -2:] = [np.inf, -np.inf]
    p = tmp_path / "File0"
    achr.write_points_csv(str(p), pts)
    want = "".join(",".join("%f" % v for v in row) + "\n" for row in pts)
    assert p.read_bytes() == want.encode()


def test_model_create_validates():
    with pytest.raises(achr.AchrError):
        achr.Model.from_coo(3, 2, [0, 5], [0, 1], [1.0, 1.0], np.zeros(3), np.ones(3))
    m = achr.Model.from_coo(3, 2, [0, 0, 1, 0], [0, 1, 2, 0], [1.0, 2.0, 3.0, 1.0], np.zeros(3), np.ones(3))
    np.testing.assert_array_equal(m.dense(), [[2, 2, 0], [0, 0, 3]])  # duplicates summed


def test_batch_rejects_bad_arguments_without_touching_the_device():
    L = achr.lib()
    h = ctypes.c_void_p()
    assert L.achr_batch_create(None, 0, ctypes.byref(h)) == -1
    two_null = (ctypes.c_void_p * 2)(None, None)
    assert L.achr_batch_create(two_null, 2, ctypes.byref(h)) == -1
    assert b"NULL" in L.achr_last_error()
    assert L.achr_batch_create(two_null, 2, None) == -1
    P = (ctypes.c_int * 2)(10, 10)
    assert L.achr_batch_sample(None, P, 10, 0, 0, 1) == -1
    assert L.achr_batch_autotune(None, P, 10, 0, None) == -1
    assert L.achr_batch_lanes(None, P) == -1
    assert L.achr_batch_set_lanes(None, P) == -1
    assert L.achr_batch_set_stream(None, None) == -1
    assert L.achr_batch_timing(None, None, None, None, None, None) == -1
    L.achr_batch_destroy(None)


def test_options_model_id_default_and_range():
    o = _lib.AchrOptions()
    achr.lib().achr_options_default(ctypes.byref(o))
    assert o.model_id == 0 and o.fp64_every == 10


def test_points_bin_round_trip_is_lossless(tmp_path):
    """The lossless point file keeps every bit (signed zeros, subnormals,
    infinities, NaN payloads) and its layout is the documented one."""
    rng = np.random.default_rng(4)
    pts = rng.standard_normal((7, 5)) * 10.0 ** rng.integers(-300, 300, (7, 5))
    pts[0, :4] = [-0.0, 5e-324, np.inf, -np.inf]
    pts[1, 0] = np.frombuffer(np.uint64(0x7FF8000000000123).tobytes(), dtype=np.float64)[0]
    achr.write_points_bin(tmp_path / "p.bin", pts)
    back = achr.read_points_bin(tmp_path / "p.bin")
    np.testing.assert_array_equal(back.view(np.uint64), pts.view(np.uint64))
    raw = (tmp_path / "p.bin").read_bytes()
    assert raw[:8] == b"ACHRPT01" and np.frombuffer(raw[8:16], "<i4").tolist() == [7, 5]
    np.testing.assert_array_equal(np.frombuffer(raw[16:], "<f8").reshape(7, 5).view(np.uint64), pts.view(np.uint64))
    (tmp_path / "short.bin").write_bytes(raw[:-8])
    with pytest.raises(achr.AchrError):
        achr.read_points_bin(tmp_path / "short.bin")
    with pytest.raises(achr.AchrError):
        achr.read_points_bin(tmp_path / "missing.bin")
    (tmp_path / "huge.bin").write_bytes(b"ACHRPT01" + np.array([2**31 - 1, 2**31 - 1], "<i4").tobytes())
    with pytest.raises(achr.AchrError):
        achr.read_points_bin(tmp_path / "huge.bin")  # refused before allocating


def test_fva_csv_names_and_layout(tmp_path):
    """The raw FVA file of the warmup generator: header, then one
    "name,%f,%f" line per reaction with the MPS column names."""
    model = achr.Model.read_mps(golden("c1.mps"))
    n = model.n
    fva = np.stack([-np.arange(n) / 3.0, np.arange(n) * 2.5]).astype(np.float64)  # rows: min, max
    p = tmp_path / "m_fva_bounds.csv"
    rc = achr.lib().achr_write_fva_csv(str(p).encode(), model._h, fva.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    lines = p.read_text().splitlines()
    assert lines[0] == "reaction,min,max" and len(lines) == n + 1
    names = [l.split(",")[0] for l in lines[1:]]
    assert len(set(names)) == n and "?" not in names
    assert lines[1 + 7].split(",")[1:] == ["%f" % (-7 / 3.0), "%f" % 17.5]
