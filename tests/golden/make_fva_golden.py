"""Writes flux-variability goldens with scipy's HiGHS (dual simplex): for each
reaction of the model, min and max of its flux subject to S x = 0 and the
column bounds, the LPs of the reference's warmup generator
(VFWarmup/createWarmupPtsGLPK.c:95-175, objective e_i, GLP_MIN / GLP_MAX).

    python tests/golden/make_fva_golden.py tests/golden/c1.mps tests/golden/c1_fva_highs.csv
    python tests/golden/make_fva_golden.py tests/golden/P_Putida.mps tests/golden/p_putida_fva_highs.csv

Output: one line per reaction, "index,min,max" with 17 significant digits.
The MPS file is read with the library's own reader (its counts are pinned
against an independent parser in tests/test_formats.py)."""
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.optimize import linprog

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "achr.cu_amd"))
import achr  # noqa: E402


def main(mps, out):
    model = achr.Model.read_mps(mps)
    n, m = model.n, model.m
    rp, ci, va = model.csr()
    S = sp.csr_matrix((va, ci, rp), shape=(m, n))
    lb, ub = model.bounds()
    bounds = list(zip(lb, ub))
    rows = []
    for i in range(n):
        vals = []
        for sgn in (1.0, -1.0):
            c = np.zeros(n)
            c[i] = sgn
            r = linprog(c, A_eq=S, b_eq=np.zeros(m), bounds=bounds, method="highs-ds",
                        options={"primal_feasibility_tolerance": 1e-10, "dual_feasibility_tolerance": 1e-10})
            if r.status != 0:
                raise RuntimeError(f"reaction {i}: {r.message}")
            vals.append(sgn * r.fun)
        rows.append(f"{i},{vals[0]:.17g},{vals[1]:.17g}")
    with open(out, "w") as f:
        f.write("\n".join(rows) + "\n")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
