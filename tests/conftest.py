import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "achr.cu_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden(name: str) -> str:
    return os.path.join(GOLDEN, name)


def load_c1():
    """Config-1 fixture: (S dense, lb, ub, warmup %f, warmup exact) read with numpy only."""
    S, lb, ub = parse_mps_py(golden("c1.mps"))
    Wf = np.loadtxt(golden("c1_warmup.csv"), delimiter=",")
    We = np.loadtxt(golden("c1_warmup_exact.csv"), delimiter=",")
    return S, lb, ub, Wf, We


def parse_mps_py(path: str):
    """Independent whitespace MPS parser (N/E rows, COLUMNS, BOUNDS UP/LO/FX) for tests."""
    rows, cols = {}, {}
    ent = []
    lb, ub = {}, {}
    sec = None
    obj = None
    with open(path) as f:
        for line in f:
            if not line.strip():
                continue
            if not line[0].isspace():
                sec = line.split()[0]
                continue
            t = line.split()
            if sec == "ROWS":
                if t[0] == "N" and obj is None:
                    obj = t[1]
                else:
                    rows[t[1]] = len(rows)
            elif sec == "COLUMNS":
                if t[0] not in cols:
                    cols[t[0]] = len(cols)
                for k in range(1, len(t) - 1, 2):
                    if t[k] != obj:
                        ent.append((rows[t[k]], cols[t[0]], float(t[k + 1])))
            elif sec == "BOUNDS":
                j = cols[t[2]]
                v = float(t[3])
                if t[0] == "UP":
                    ub[j] = v
                elif t[0] == "LO":
                    lb[j] = v
                elif t[0] == "FX":
                    lb[j] = ub[j] = v
    n, m = len(cols), len(rows)
    S = np.zeros((m, n))
    for r, c, v in ent:
        S[r, c] += v
    L = np.array([lb.get(j, 0.0) for j in range(n)])
    U = np.array([ub.get(j, np.finfo(float).max) for j in range(n)])
    return S, L, U


@pytest.fixture(scope="session")
def c1():
    return load_c1()


def has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
