"""Writes the config-1 fixtures (E. coli-core-sized synthetic model, seed 1):

  c1.mps                 72 x 95 model, fixed-format MPS
  c1_warmup.csv          190 FVA warmup points, "%f" (what VFWarmup writes)
  c1_warmup_exact.csv    the same points, "%.17g" (lossless)
  c1_nullproj_diag.txt   diagonal of the null-space projector N N^T from scipy.linalg.null_space

Run from the repository root: python tests/golden/make_config1.py
(needs scipy for the HiGHS LPs and the null space).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "achr.cu_amd"))
from achr import synth  # noqa: E402

m = synth.config1(1)
synth.write_mps(m, os.path.join(HERE, "c1.mps"), name="config1")
W = synth.fva_warmup(m)
synth.write_warmup_csv(W, os.path.join(HERE, "c1_warmup.csv"), "%f")
synth.write_warmup_csv(W, os.path.join(HERE, "c1_warmup_exact.csv"), "%.17g")
N = synth.nullspace_dense(m.dense())
np.savetxt(os.path.join(HERE, "c1_nullproj_diag.txt"), np.diag(N @ N.T), fmt="%.17g")
print("config-1 fixtures written:", m.m, "x", m.n, "d =", N.shape[1], "W =", W.shape[1])
