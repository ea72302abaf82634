"""Seeded synthetic flux models and warmup sets.

The reference ships one real model (P. putida, VFWarmup/lib/P_Putida.mps) and
no sampler inputs; its warmup sets come from VFWarmup, an LP (GLPK/CPLEX +
MPI) program. The workloads of this project are synthetic models with the
structure of metabolic networks:

* S is m x n, each column with 2-4 nonzeros from {+-1, +-2} on distinct rows
  (P(2, 3, 4) = (.3, .2, .5): 3.2 per column on average);
* an interior point x0 is drawn in null(S) (|x0|_inf <= x0_scale);
* bounds are x0 -+ U(lo, hi); a fraction of reactions is irreversible
  (lb raised to 0 where x0 > 0) and a few are fixed (lb = ub = x0);
* bounds are rounded to 4 decimals so the MPS file carries them exactly in
  12-character fixed-format fields.

Warmup sets:
* ``fva_warmup``: the reference generator's recipe (createWarmupPtsGLPK.c):
  for reaction i, minimise then maximise x_i (columns 2i and 2i+1, :101-173),
  clip to bounds (movePtsBds, :36-46), then move every point to
  0.33 x + 0.67 c where c is their mean (:48-65, :283-284). Solved with
  scipy's HiGHS here (GLPK is not available); only for small models.
* ``chord_warmup``: for models too large for 2n LPs: endpoints of random
  chords through x0 along null-space directions (alternately the far and near
  end, like the max / min pair of an FVA), then the same 0.33 / 0.67 move.
* ``round_like_csv``: the "%f" text the reference's warmup writer produces
  (createWarmupPtsGLPK.c:286-297) keeps 6 decimals; sampling from such a file
  sees residuals |S w| ~ 1e-6, which is what triggers the reference's
  reprojection every 10 steps.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SynthModel:
    S: np.ndarray | None          # dense m x n (small models) or None
    rows: np.ndarray              # COO of S
    cols: np.ndarray
    vals: np.ndarray
    m: int
    n: int
    lb: np.ndarray
    ub: np.ndarray
    x0: np.ndarray
    N: np.ndarray | None = None   # orthonormal null-space basis used to draw x0

    def dense(self) -> np.ndarray:
        if self.S is not None:
            return self.S
        S = np.zeros((self.m, self.n))
        np.add.at(S, (self.rows, self.cols), self.vals)
        return S


def random_stoich(m: int, n: int, rng: np.random.Generator, probs=(0.3, 0.2, 0.5)):
    """COO triplets of an m x n matrix with 2-4 nonzeros per column in {+-1, +-2}."""
    k = rng.choice([2, 3, 4], size=n, p=probs)
    rows, cols, vals = [], [], []
    for j in range(n):
        r = rng.choice(m, size=int(k[j]), replace=False)
        v = rng.choice([-2.0, -1.0, 1.0, 2.0], size=int(k[j]))
        rows.append(r)
        cols.append(np.full(k[j], j))
        vals.append(v)
    return (np.concatenate(rows).astype(np.int32), np.concatenate(cols).astype(np.int32),
            np.concatenate(vals))


def nullspace_dense(S: np.ndarray) -> np.ndarray:
    """Orthonormal basis of null(S) by SVD (small models)."""
    import scipy.linalg
    return scipy.linalg.null_space(S)


def make_model(m: int, n: int, seed: int, *, irrev_frac: float, n_fixed: int, width=(5.0, 50.0),
               x0_scale: float = 10.0, N: np.ndarray | None = None, nullspace_fn=None) -> SynthModel:
    """Synthetic model; N (or nullspace_fn(rows, cols, vals, m, n)) supplies null(S) for large n."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = random_stoich(m, n, rng)
    S = None
    if N is None:
        if nullspace_fn is not None:
            N = nullspace_fn(rows, cols, vals, m, n)
        else:
            S = np.zeros((m, n))
            np.add.at(S, (rows, cols), vals)
            N = nullspace_dense(S)
    d = N.shape[1]
    if d == 0:
        raise ValueError("S has a trivial null space")
    z = rng.standard_normal(d)
    x0 = N @ z
    x0 *= x0_scale * rng.uniform(0.5, 1.0) / np.max(np.abs(x0))
    fixed = rng.choice(n, size=n_fixed, replace=False) if n_fixed else np.array([], dtype=int)
    if len(fixed):
        # Move x0 inside null(S) so that the fixed reactions sit exactly on
        # their 4-decimal values.
        target = np.round(x0[fixed], 4)
        NF = N[fixed, :]
        delta, *_ = np.linalg.lstsq(NF, target - x0[fixed], rcond=None)
        x0 = x0 + N @ delta
    lo, hi = width
    lb = np.round(x0 - rng.uniform(lo, hi, size=n), 4)
    ub = np.round(x0 + rng.uniform(lo, hi, size=n), 4)
    irrev = rng.random(n) < irrev_frac
    lb = np.where(irrev & (x0 > 0), np.maximum(lb, 0.0), lb)
    if len(fixed):
        # 4-decimal fixed values when x0 reaches them exactly (they then fit
        # a fixed-format MPS field); otherwise x0's own values, since fixed
        # reactions with dependent null-space rows cannot take arbitrary
        # rounded values and stay on S x = 0.
        v = np.round(x0[fixed], 4)
        if np.max(np.abs(x0[fixed] - v)) > 1e-12:
            v = x0[fixed]
        lb[fixed] = v
        ub[fixed] = v
    if S is None and n <= 4000:
        S = np.zeros((m, n))
        np.add.at(S, (rows, cols), vals)
    return SynthModel(S, rows, cols, vals, m, n, lb, ub, x0, N)


def config1(seed: int = 1) -> SynthModel:
    """E. coli-core-sized model: 72 x 95, 30% irreversible, 5 fixed."""
    return make_model(72, 95, seed, irrev_frac=0.30, n_fixed=5, width=(5.0, 50.0), x0_scale=10.0)


def config2(nullspace_fn, seed: int = 2) -> SynthModel:
    """Genome-scale model: 5,800 x 10,600, 40% irreversible, 2% fixed, bounds +-(100..1000)."""
    return make_model(5800, 10600, seed, irrev_frac=0.40, n_fixed=212, width=(100.0, 1000.0), x0_scale=100.0,
                      nullspace_fn=nullspace_fn)


def move_to_center(W: np.ndarray) -> np.ndarray:
    """0.33 x + 0.67 c for every column (createWarmupPtsGLPK.c:59-65)."""
    c = W.sum(axis=1) / W.shape[1]
    return W * 0.33 + 0.67 * c[:, None]


def fva_warmup(model: SynthModel) -> np.ndarray:
    """2n FVA points (min, max of each reaction), clipped and moved to the centre."""
    from scipy.optimize import linprog
    S = model.dense()
    n = model.n
    W = np.empty((n, 2 * n))
    bounds = list(zip(model.lb, model.ub))
    for i in range(n):
        for sense, col in ((1.0, 2 * i), (-1.0, 2 * i + 1)):  # min -> 2i, max -> 2i+1
            c = np.zeros(n)
            c[i] = sense
            res = linprog(c, A_eq=S, b_eq=np.zeros(model.m), bounds=bounds, method="highs")
            if res.status != 0:
                raise RuntimeError(f"LP {i} ({'min' if sense > 0 else 'max'}) failed: {res.message}")
            W[:, col] = np.clip(res.x, model.lb, model.ub)
    return move_to_center(W)


def free_directions(N: np.ndarray, lb: np.ndarray, ub: np.ndarray) -> np.ndarray:
    """Orthonormal basis of {v in span(N) : v_i = 0 for fixed reactions (lb = ub)}."""
    import scipy.linalg
    fixed = np.nonzero(lb == ub)[0]
    if len(fixed) == 0:
        return N
    Z = scipy.linalg.null_space(N[fixed, :])
    return N @ Z


def chord_warmup(N: np.ndarray, x0: np.ndarray, lb: np.ndarray, ub: np.ndarray, W: int, seed: int,
                 shrink: float = 0.999, free: np.ndarray | None = None, mix: int = 0) -> np.ndarray:
    """W points on random chords through x0 along directions that keep S x = 0
    and the fixed reactions fixed, then moved to the centre.

    Directions are Gaussian combinations of the free null-space basis (`free`,
    computed from N when not given): of every basis vector (mix = 0), or of
    `mix` randomly chosen ones (O(n W mix) work for large models)."""
    rng = np.random.default_rng(seed)
    N = free_directions(N, lb, ub) if free is None else free
    if mix > 0:
        N = np.asfortranarray(N)  # column gathers
    n, d = N.shape
    out = np.empty((n, W))
    B = 512
    for k0 in range(0, W, B):
        k1 = min(W, k0 + B)
        if mix > 0:
            idx = rng.integers(0, d, size=(k1 - k0, mix))
            coef = rng.standard_normal((k1 - k0, mix))
            V = np.zeros((n, k1 - k0))
            for j in range(mix):
                V += N[:, idx[:, j]] * coef[:, j]
        else:
            G = rng.standard_normal((d, k1 - k0))
            V = N @ G
        pos = V > 1e-12
        flat = ~(pos | (V < -1e-12))
        a, b = (ub - x0)[:, None], (lb - x0)[:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            up = np.where(pos, a, b) / V
            dn = np.where(pos, b, a) / V
        up[flat] = np.inf
        dn[flat] = -np.inf
        tmax = up.min(axis=0)
        tmin = dn.max(axis=0)
        far = (np.arange(k0, k1) % 2) == 1
        t = np.where(far, tmax, tmin) * shrink
        out[:, k0:k1] = x0[:, None] + V * t[None, :]
    out = np.clip(out, lb[:, None], ub[:, None])
    return move_to_center(out)


def batch_model(n: int, seed: int, nullspace_fn):
    """One model of the 64-model batch workload: m = 0.55 n metabolites, the
    sparsity law above, 30% irreversible, 2% fixed, bounds x0 -+ (5..50).
    Returns it with an orthonormal basis of its free null space
    {v : S v = 0, v_i = 0 for fixed i}, both from nullspace_fn(rows, cols,
    vals, m, n)."""
    mm = int(round(0.55 * n))
    sm = make_model(mm, n, seed, irrev_frac=0.30, n_fixed=max(1, int(round(0.02 * n))), width=(5.0, 50.0),
                    x0_scale=10.0, nullspace_fn=nullspace_fn)
    fixed = np.nonzero(sm.lb == sm.ub)[0]
    # free directions: null([S; e_i for the fixed reactions i])
    r = np.concatenate([sm.rows, mm + np.arange(len(fixed))]).astype(np.int32)
    c = np.concatenate([sm.cols, fixed]).astype(np.int32)
    v = np.concatenate([sm.vals, np.ones(len(fixed))])
    return sm, nullspace_fn(r, c, v, mm + len(fixed), n)


def csr(sm: SynthModel):
    """CSR (rowptr, colidx, vals) of S, columns ascending within a row."""
    order = np.lexsort((sm.cols, sm.rows))
    rows, cols, vals = sm.rows[order], sm.cols[order], sm.vals[order]
    rp = np.zeros(sm.m + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    return np.cumsum(rp).astype(np.int32), cols.astype(np.int32), vals


def round_like_csv(W: np.ndarray) -> np.ndarray:
    """Values as a "%f" warmup CSV stores them (6 decimals)."""
    return np.round(W, 6)


def fmt_g12(v: float) -> str:
    s = f"{v:.12g}"
    if len(s) > 12:
        s = f"{v:.6f}".rstrip("0").rstrip(".")
    return s


def write_mps(model: SynthModel, path: str, name: str = "synthetic") -> None:
    """Fixed-format MPS in the dialect of VFWarmup/test_fva_validation.py:58-86."""
    order = np.lexsort((model.rows, model.cols))
    rows, cols, vals = model.rows[order], model.cols[order], model.vals[order]
    with open(path, "w") as f:
        f.write(f"{'NAME':<14}{name}\n")
        f.write("ROWS\n")
        f.write(" N  COST\n")
        for i in range(model.m):
            f.write(f" E  EQ{i + 1}\n")
        f.write("COLUMNS\n")
        for r, c, v in zip(rows, cols, vals):
            f.write(f"    {'X' + str(c + 1):<10}{'EQ' + str(r + 1):<10}{fmt_g12(v):<12}\n")
        f.write("RHS\n")
        f.write("BOUNDS\n")
        for j in range(model.n):
            cn = f"X{j + 1}"
            lb, ub = model.lb[j], model.ub[j]
            if lb == ub:
                f.write(f" FX BOUND     {cn:<10}{fmt_g12(lb):<12}\n")
            else:
                if lb != 0:
                    f.write(f" LO BOUND     {cn:<10}{fmt_g12(lb):<12}\n")
                f.write(f" UP BOUND     {cn:<10}{fmt_g12(ub):<12}\n")
        f.write("ENDATA\n")


def write_warmup_csv(W: np.ndarray, path: str, fmt: str = "%f") -> None:
    """One line per reaction, one comma-separated column per point (createWarmupPtsGLPK.c:286-297)."""
    with open(path, "w") as f:
        for i in range(W.shape[0]):
            f.write(",".join(fmt % v for v in W[i]) + "\n")
