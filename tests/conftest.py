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


def signed_zero_bank_case(seed=5):
    """Six reactions, two of them blocked by S (x_0 = x_1 = 0): the blocked
    coordinates start at +0.0 and no direction moves them. Reaction 0 has
    the upper bound -0.0, reaction 1 the lower bound +0.0. Returns (S, lb, ub,
    N, warmup, start point, bank)."""
    n = 6
    S = np.zeros((2, n))
    S[0, 0] = S[1, 1] = 1.0
    lb = np.array([-1.0, 0.0, -1.0, -2.0, -1.0, -1.0])
    ub = np.array([-0.0, 1.0, 1.0, 2.0, 1.0, 1.0])
    N = np.eye(n)[:, 2:].copy()
    rng = np.random.default_rng(seed)
    W = np.zeros((n, 12))
    W[2:] = lb[2:, None] + (ub - lb)[2:, None] * rng.uniform(0.2, 0.8, (n - 2, 12))
    c0 = W.mean(axis=1)
    bank = N @ rng.standard_normal((N.shape[1], 20))
    bank /= np.linalg.norm(bank, axis=0)
    assert not np.signbit(c0[:2]).any() and not np.signbit(bank[:2]).any()
    return S, lb, ub, N, W, c0, bank


def has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
