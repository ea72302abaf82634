/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C, FP64 restatement of the reference ACHR chain loop, used as the
 * checker for the HIP sampler (tests, the smoke check and the CPU baseline
 * leg of the benchmark). Nothing in the shipped library links or calls it.
 *
 * What it restates (reference citations are paths in marouenbg/ACHR.cu):
 *   oracle_philox4x32_10  — the counter-based stream that replaces cuRAND
 *                           (ACHRcu/ACHRCuda.cu:291-295, :309, :371).
 *                           Pinned by the Random123 known-answer vectors and
 *                           by rocRAND's philox4x32_10 engine (tests).
 *   oracle_center         — createCenterPt, c0 = (1/W) * sum_k w_k
 *                           (ACHRCuda.cu:414-434). Pinned against numpy mean.
 *   oracle_nullspace      — computeKernelCuda: eigenvectors of S^T S whose
 *                           eigenvalue (= singular value of S^T S) is below
 *                           n * sigma_0 * 2.2204e-16 (ACHRCuda.cu:436-502,
 *                           :41, :474-483). Cyclic Jacobi. Pinned against the
 *                           projector of scipy.linalg.null_space goldens.
 *   oracle_sample         — stepPointProgress / createPoint / fillrandPoint /
 *                           advNextStep / findMaxAbs / reprojectPoint(2) /
 *                           correctBounds (ACHRCuda.cu:193-385), with the
 *                           tolerances of main (:505, :508, :512) and the
 *                           ordering notes of VFWarmup/ACHRSamplerMOD.m:77-185.
 *                           Pinned by closed forms (box polytope chord bounds,
 *                           uniform moments, telescoping centre identity,
 *                           clip invariant) in tests/test_oracle.py.
 *
 * Arithmetic contract shared with the HIP kernels (so that both produce the
 * same bits from the same inputs):
 *   - built with -ffp-contract=off; fused multiply-add appears only where
 *     written as fma(): the sum of squares, the step t = fma(r, max-min, min),
 *     the advance x = fma(t, u, x), the residual and projection dot products,
 *     and the centre update c = fma(alpha, c, beta*x). These are the places
 *     where nvcc's default contraction fuses the reference's device code.
 *   - the squared norm is reduced over `lanes` partial sums (element i goes to
 *     lane i % lanes, accumulated in increasing i), then lanes are combined by
 *     a pairwise tree within groups of min(lanes, 64) and a second pairwise
 *     tree across groups; the kernels use the same tree for the same lanes.
 *   - dot products of the projection N (N^T x) and of the residual S x are
 *     sequential in the reference's loop order (ACHRCuda.cu:214-216, :226-228).
 *
 * Bank mode (oracle_sample with a bank): the library's learned sampler —
 * the direction of an attempt is a stored unit column of the bank (picked
 * like a warmup column), with no normalisation and no centre, and x is
 * clipped with the IEEE 754-2019 maximum and minimum (-0 below +0, see
 * clip_ieee) at the start of every attempt and after every accepted step
 * (the kernel's v_max_f64 / v_min_f64); everything else is the loop above.
 * No counterpart in the reference.
 *
 * Deliberate differences from the reference (all documented behaviour of the
 * new sampler, not approximations):
 *   - independent uniforms for the warmup pick and the step length (the
 *     reference reuses one cuRAND state for both, ACHRCuda.cu:291, :297);
 *   - a chord with no bounding coordinate, an infinite end or a zero
 *     direction counts as a miss (the reference dereferences an empty range,
 *     ACHRCuda.cu:283-288, or divides by zero, :264-267);
 *   - a chain that misses more than max_misses times in a row stops and is
 *     flagged (the reference retries without bound, :316-319).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_EPSILON 2.2204e-16 /* ACHRCuda.cu:41 */

/* ------------------------------------------------------------------ Philox */

static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint32_t h0, l0, h1, l1;
    mulhilo32(0xD2511F53u, c0, &h0, &l0);
    mulhilo32(0xCD9E8D57u, c2, &h1, &l1);
    uint32_t n0 = h1 ^ c1 ^ k0, n1 = l1, n2 = h0 ^ c3 ^ k1, n3 = l0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double stream_u01(uint64_t seed, uint32_t index, uint32_t chain, uint32_t file,
                                uint32_t model, uint32_t tag) {
  uint32_t ctr[4] = {index, chain, file, (model << 4) | tag};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  oracle_philox4x32_10(ctr, key, o);
  uint64_t bits = (((uint64_t)o[1] << 32) | o[0]) >> 11;
  return (double)bits * 0x1.0p-53;
}

double oracle_stream_u01(uint64_t seed, uint32_t index, uint32_t chain, uint32_t file, uint32_t model,
                         uint32_t tag) {
  return stream_u01(seed, index, chain, file, model, tag);
}

/* ------------------------------------------------------- centre, null space */

void oracle_center(const double *warm, int n, int W, double *c0) {
  const double alpha = 1.0 / (double)W;
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int k = 0; k < W; ++k) s += warm[(size_t)k * n + i];
    c0[i] = alpha * s;
  }
}

/* Cyclic Jacobi eigen-decomposition of the symmetric n x n matrix A (column
 * major, destroyed). Eigenvectors go to V (column major), eigenvalues to ev. */
static void jacobi_eig(double *A, int n, double *V, double *ev) {
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) V[i + (size_t)j * n] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int j = 0; j < n; ++j) {
      diag += A[j + (size_t)j * n] * A[j + (size_t)j * n];
      for (int i = 0; i < j; ++i) off += A[i + (size_t)j * n] * A[i + (size_t)j * n];
    }
    if (off <= 1e-60 * (diag > 0 ? diag : 1.0) || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p + (size_t)q * n];
        if (apq == 0.0) continue;
        double app = A[p + (size_t)p * n], aqq = A[q + (size_t)q * n];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) { /* columns p, q */
          double akp = A[k + (size_t)p * n], akq = A[k + (size_t)q * n];
          A[k + (size_t)p * n] = c * akp - s * akq;
          A[k + (size_t)q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) { /* rows p, q */
          double apk = A[p + (size_t)k * n], aqk = A[q + (size_t)k * n];
          A[p + (size_t)k * n] = c * apk - s * aqk;
          A[q + (size_t)k * n] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + (size_t)p * n], vkq = V[k + (size_t)q * n];
          V[k + (size_t)p * n] = c * vkp - s * vkq;
          V[k + (size_t)q * n] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < n; ++i) ev[i] = A[i + (size_t)i * n];
}

/* S is m x n column major. Writes up to n columns of N (n x d, column major)
 * and returns d, the number of eigenvalues of S^T S below n*max*EPSILON. */
int oracle_nullspace(const double *S, int m, int n, double *N_out) {
  double *A = (double *)calloc((size_t)n * n, sizeof(double));
  double *V = (double *)malloc((size_t)n * n * sizeof(double));
  double *ev = (double *)malloc((size_t)n * sizeof(double));
  int *ord = (int *)malloc((size_t)n * sizeof(int));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) {
      double s = 0.0;
      for (int k = 0; k < m; ++k) s += S[k + (size_t)i * m] * S[k + (size_t)j * m];
      A[i + (size_t)j * n] = s;
      A[j + (size_t)i * n] = s;
    }
  jacobi_eig(A, n, V, ev);
  /* singular values of a symmetric PSD matrix are |eigenvalues|; sort descending */
  for (int i = 0; i < n; ++i) ord[i] = i;
  for (int i = 1; i < n; ++i) {
    int t = ord[i], j = i - 1;
    while (j >= 0 && fabs(ev[ord[j]]) < fabs(ev[t])) { ord[j + 1] = ord[j]; --j; }
    ord[j + 1] = t;
  }
  double tol = (double)n * fabs(ev[ord[0]]) * ORACLE_EPSILON;
  int istart = n;
  for (int i = 0; i < n; ++i)
    if (fabs(ev[ord[i]]) < tol) { istart = i; break; }
  int d = n - istart;
  for (int j = 0; j < d; ++j) {
    int src = ord[istart + j];
    for (int i = 0; i < n; ++i) N_out[i + (size_t)j * n] = V[i + (size_t)src * n];
  }
  free(A); free(V); free(ev); free(ord);
  return d;
}

/* --------------------------------------------------------------- sampling */

/* Canonical reduction of per-lane partial sums (see header comment). */
static double canon_tree(double *lanes, int L) {
  int G = L < 64 ? L : 64;
  int NG = L / G;
  for (int g = 0; g < NG; ++g) {
    double *v = lanes + (size_t)g * G;
    for (int o = G / 2; o >= 1; o /= 2)
      for (int l = 0; l < o; ++l) v[l] = v[l] + v[l + o];
  }
  for (int o = NG / 2; o >= 1; o /= 2)
    for (int g = 0; g < o; ++g) lanes[(size_t)g * G] = lanes[(size_t)g * G] + lanes[(size_t)(g + o) * G];
  return lanes[0];
}

typedef struct {
  int n, m, W, d;
  const double *warm;   /* n x W column major */
  const double *lb, *ub, *c0;
  const int *rowptr, *colidx; /* CSR of S, m rows, columns ascending */
  const double *vals;
  const double *N;      /* n x d column major, may be NULL when d == 0 */
  const double *bank;   /* optional: n x K unit directions (bank mode) */
  int K;
} oracle_model;

typedef struct {
  uint64_t seed;
  uint32_t file, model;
  int chain0;           /* global index of chain 0 in the Philox counter */
  int lanes;            /* reduction lanes of the kernel schedule being mirrored */
  int check_every;      /* ACHRCuda.cu:325 -> 10 */
  int max_misses;
  double dTol, uTol, maxMinTol, reprojTol; /* :512, :508, :505, :338 */
  int trace_chains, trace_thin, trace_len;
} oracle_opts;

/* The IEEE 754-2019 maximum / minimum of two doubles: -0 orders below +0,
 * and a NaN operand gives the other one. clip_ieee(x, lb, ub) =
 * min(max(x, lb), ub) is idempotent. */
static double max_ieee(double a, double b) {
  if (a > b || isnan(b)) return a;
  if (b > a || isnan(a)) return b;
  return signbit(a) ? b : a; /* equal: +0 over -0 */
}
static double min_ieee(double a, double b) {
  if (a < b || isnan(b)) return a;
  if (b < a || isnan(a)) return b;
  return signbit(a) ? a : b; /* equal: -0 over +0 */
}
static double clip_ieee(double x, double lb, double ub) { return min_ieee(max_ieee(x, lb), ub); }

/* Per-chain counters written to stats[4*b + k]:
 *   k=0 misses, k=1 reprojections, k=2 residual checks, k=3 flags (1 = stuck) */
static void run_chain(const oracle_model *M, const oracle_opts *o, int P, int S, int b, double *out,
                      int32_t *stats, double *trace, double *wk) {
  const int n = M->n;
  double *x = wk, *c = wk + n, *u = wk + 2 * n, *r = wk + 3 * n; /* r: m + d */
  double *lanes = (double *)malloc((size_t)o->lanes * sizeof(double));
  for (int i = 0; i < n; ++i) { x[i] = M->c0[i]; c[i] = M->c0[i]; }
  uint32_t attempt = 0;
  int s = 0, consec = 0;
  int32_t misses = 0, nrep = 0, nchk = 0, flags = 0;
  const uint32_t chain = (uint32_t)(o->chain0 + b);
  while (s < S) {
    double up = stream_u01(o->seed, attempt, chain, o->file, o->model, 0u);
    attempt++;
    const int npick = M->bank ? M->K : M->W;
    int id = (int)((double)npick * up);
    if (id >= npick) id = npick - 1;
    double nrm = 1.0;
    if (M->bank) {
      /* bank mode: a stored unit direction, no centre */
      memcpy(u, M->bank + (size_t)id * n, (size_t)n * sizeof(double));
    } else {
      const double *w = M->warm + (size_t)id * n;
      for (int l = 0; l < o->lanes; ++l) lanes[l] = 0.0;
      for (int i = 0; i < n; ++i) {
        u[i] = w[i] - c[i];
        lanes[i % o->lanes] = fma(u[i], u[i], lanes[i % o->lanes]);
      }
      nrm = sqrt(canon_tree(lanes, o->lanes));
    }
    double mn = -INFINITY, mx = INFINITY;
    for (int i = 0; i < n; ++i) {
      const double ui = M->bank ? u[i] : u[i] / nrm;
      u[i] = ui;
      if (M->bank) x[i] = clip_ieee(x[i], M->lb[i], M->ub[i]);
      const double du = M->ub[i] - x[i], dl = x[i] - M->lb[i];
      if (du > o->dTol && dl > o->dTol && (ui > o->uTol || ui < -o->uTol)) {
        /* reference (ACHRCuda.cu:270-279): the two distances divided by u_i;
         * bank mode (no reference counterpart): one reciprocal of u_i, then
         * two products */
        double lo, hi;
        if (M->bank) {
          const double rcp = 1.0 / ui;
          lo = -dl * rcp;
          hi = du * rcp;
        } else {
          lo = -dl / ui;
          hi = du / ui;
        }
        if (ui < 0.0) { const double tmp = lo; lo = hi; hi = tmp; }
        if (lo > mn) mn = lo;
        if (hi < mx) mx = hi;
      }
    }
    int miss = !(nrm > 0.0) || !(mn > -INFINITY) || !(mx < INFINITY) ||
               (fabs(mn) < o->maxMinTol && fabs(mx) < o->maxMinTol) || (mn > mx);
    if (miss) {
      misses++;
      if (++consec > o->max_misses) { flags |= 1; break; }
      continue;
    }
    consec = 0;
    const double rs = stream_u01(o->seed, (uint32_t)s, chain, o->file, o->model, 1u);
    const double t = fma(rs, mx - mn, mn);
    for (int i = 0; i < n; ++i) x[i] = fma(t, u[i], x[i]);
    if (s % o->check_every == 0) {
      nchk++;
      double rho = 0.0;
      for (int j = 0; j < M->m; ++j) {
        double acc = 0.0;
        for (int p = M->rowptr[j]; p < M->rowptr[j + 1]; ++p) acc = fma(M->vals[p], x[M->colidx[p]], acc);
        double a = fabs(acc);
        if (a > rho) rho = a;
      }
      if (rho > o->reprojTol && M->d > 0) {
        nrep++;
        double *tmp = r;
        for (int j = 0; j < M->d; ++j) {
          const double *Nj = M->N + (size_t)j * n;
          double acc = 0.0;
          for (int i = 0; i < n; ++i) acc = fma(Nj[i], x[i], acc);
          tmp[j] = acc;
        }
        for (int i = 0; i < n; ++i) {
          double acc = 0.0;
          for (int j = 0; j < M->d; ++j) acc = fma(M->N[i + (size_t)j * n], tmp[j], acc);
          x[i] = acc;
        }
      }
    }
    const double alpha = (double)(M->W + s) / (double)(M->W + s + 1);
    const double beta = 1.0 / (double)(M->W + s + 1);
    for (int i = 0; i < n; ++i) {
      if (M->bank) {
        x[i] = clip_ieee(x[i], M->lb[i], M->ub[i]);
        continue;
      }
      if (x[i] > M->ub[i]) x[i] = M->ub[i];
      else if (x[i] < M->lb[i]) x[i] = M->lb[i];
      c[i] = fma(alpha, c[i], beta * x[i]);
    }
    s++;
    if (trace && b < o->trace_chains && o->trace_thin > 0 && s % o->trace_thin == 0) {
      int k = s / o->trace_thin - 1;
      if (k < o->trace_len) memcpy(trace + ((size_t)b * o->trace_len + k) * n, x, (size_t)n * sizeof(double));
    }
  }
  for (int i = 0; i < n; ++i) out[(size_t)i * P + b] = x[i];
  if (stats) {
    stats[4 * b + 0] = misses;
    stats[4 * b + 1] = nrep;
    stats[4 * b + 2] = nchk;
    stats[4 * b + 3] = flags;
  }
  free(lanes);
}

/* Runs chains [b_lo, b_hi) of a P-chain file; out is n x P (reaction major,
 * out[i*P + b], the layout of the reference's File<k>, ACHRCuda.cu:744-756).
 * threads <= 1 runs serially; otherwise OpenMP over chains. */
int oracle_sample(int n, int m, int W, int d, const double *warm, const double *lb, const double *ub,
                  const double *c0, const int *rowptr, const int *colidx, const double *vals,
                  const double *N, int P, int S, int b_lo, int b_hi, uint64_t seed, uint32_t file,
                  uint32_t model, int chain0, int lanes, int check_every, int max_misses, double dTol,
                  double uTol, double maxMinTol, double reprojTol, int threads, double *out,
                  int32_t *stats, double *trace, int trace_chains, int trace_thin, int trace_len,
                  const double *bank, int K) {
  if (lanes < 1 || (lanes & (lanes - 1)) || (lanes > 64 && lanes % 64)) return -1;
  if (bank && K <= 0) return -1;
  oracle_model M = {n, m, W, d, warm, lb, ub, c0, rowptr, colidx, vals, N, bank, K};
  oracle_opts o = {seed, file, model, chain0, lanes, check_every, max_misses, dTol, uTol, maxMinTol,
                   reprojTol, trace_chains, trace_thin, trace_len};
  const size_t wk_len = (size_t)4 * n + (size_t)m + (size_t)d + 8;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 1 ? threads : 1) if (threads > 1)
  for (int b = b_lo; b < b_hi; ++b) {
    double *wk = (double *)malloc(wk_len * sizeof(double));
    run_chain(&M, &o, P, S, b, out, stats, trace, wk);
    free(wk);
  }
  return 0;
}
