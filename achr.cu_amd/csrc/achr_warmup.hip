// Warmup points by flux variability analysis on the GPU: every LP of the
// reference's generator (VFWarmup/createWarmupPtsGLPK.c:67-182) -- for each
// reaction min and max of its flux, then random objectives -- solved together
// by one batched primal-dual hybrid gradient method with the restart scheme of
// PDLP (Applegate, Diaz, Hinder, Lubin, Lu, O'Donoghue, Schudy, "Practical
// large-scale linear programming using primal-dual hybrid gradient", NeurIPS
// 2021): diagonal (Ruiz + Pock-Chambolle) rescaling, restarts to the current
// or averaged iterate on KKT-error decay, primal-weight updates at restarts.
// The reference runs GLPK's simplex one LP at a time over MPI ranks; here the
// LPs are the columns of the iterates (X: n x B, Y: m x B, column b contiguous
// along b so every sparse gather is a coalesced row segment).
//
// LP of column b:  min c_b^T x  s.t.  S x = 0,  lb <= x <= ub.
// Column b = 2 i is the minimum, 2 i + 1 the maximum (as -min of -c) of pair
// i: c = e_i for i < n, a random vector in [-0.5, 0.5)^n (Philox) otherwise,
// one per column like the reference's rand() per solve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/achr.h"
#include "achr_host.h"
#include "achr_internal.h"
#include "achr_philox.h"
#include "achr_polish.h"

using achr_host::DevBuf;
using achr_host::fail;

namespace {

constexpr int kChunk = 64;  // rows / columns per partial sum of the KKT reductions
constexpr uint32_t kTagObjective = 2;

struct Lp {
  int n, m, B, nfva;  // columns b < 2 nfva are FVA objectives
  const int* rp;      // CSR of the scaled matrix (m rows)
  const int* ci;
  const double* va;
  const int* cp;      // CSC (n columns)
  const int* ri;
  const double* vc;
  const double* l;    // scaled bounds
  const double* u;
  const double* dc;   // column scaling (objective c~ = dc c)
  const double* Cr;   // random objectives, row major n x (B - 2 nfva), unscaled
};

__device__ __forceinline__ double objective(const Lp& p, int i, int b) {
  const double sgn = (b & 1) ? -1.0 : 1.0;
  if (b < 2 * p.nfva) return (i == (b >> 1)) ? sgn * p.dc[i] : 0.0;
  const int R = p.B - 2 * p.nfva;
  return sgn * p.dc[i] * p.Cr[static_cast<size_t>(i) * R + (b - 2 * p.nfva)];
}

// x <- clamp(x - tau (c - A^T y)); xbar = 2 x_new - x; xsum += x_new.
// Grid: LPs across x, columns across y (strided). The column's CSC entries,
// scaling and bounds depend on the block only, so they are read once per
// wavefront through the scalar unit; the iterate reads are coalesced over b.
__global__ __launch_bounds__(256) void primal_kernel(Lp p, double* __restrict__ X, double* __restrict__ Xbar,
                                                     double* __restrict__ Xsum, const double* __restrict__ Y,
                                                     const double* __restrict__ tau, const int* __restrict__ done) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= p.B || done[b]) return;
  const double tb = tau[b];
  for (int i = blockIdx.y; i < p.n; i += gridDim.y) {
    const size_t idx = static_cast<size_t>(i) * p.B + b;
    double g = 0.0;
    for (int q = p.cp[i]; q < p.cp[i + 1]; ++q) g = fma(p.vc[q], Y[static_cast<size_t>(p.ri[q]) * p.B + b], g);
    const double x = X[idx];
    double xn = x - tb * (objective(p, i, b) - g);
    xn = fmin(fmax(xn, p.l[i]), p.u[i]);
    X[idx] = xn;
    Xbar[idx] = 2.0 * xn - x;
    Xsum[idx] += xn;
  }
}

// y <- y - sigma A xbar; ysum += y_new (grid as in primal_kernel, rows across y)
__global__ __launch_bounds__(256) void dual_kernel(Lp p, double* __restrict__ Y, double* __restrict__ Ysum,
                                                   const double* __restrict__ Xbar, const double* __restrict__ sig,
                                                   const int* __restrict__ done) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= p.B || done[b]) return;
  const double sb = sig[b];
  for (int j = blockIdx.y; j < p.m; j += gridDim.y) {
    const size_t idx = static_cast<size_t>(j) * p.B + b;
    double r = 0.0;
    for (int q = p.rp[j]; q < p.rp[j + 1]; ++q) r = fma(p.va[q], Xbar[static_cast<size_t>(p.ci[q]) * p.B + b], r);
    const double y = Y[idx] - sb * r;
    Y[idx] = y;
    Ysum[idx] += y;
  }
}

// Partial sums over row chunks: |A x|^2 and |y - y_r|^2 for the current and
// the averaged iterate (avg = sum / den). part[(chunk * 4 + k) * B + b].
__global__ __launch_bounds__(256) void kkt_rows_kernel(Lp p, const double* __restrict__ X,
                                                       const double* __restrict__ Xsum, const double* __restrict__ Y,
                                                       const double* __restrict__ Ysum, const double* __restrict__ Yr,
                                                       const double* __restrict__ den, double* __restrict__ part) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  if (b >= p.B) return;
  const double inv = 1.0 / den[b];
  double pc = 0.0, pa = 0.0, dyc = 0.0, dya = 0.0;
  const int j1 = std::min(p.m, (chunk + 1) * kChunk);
  for (int j = chunk * kChunk; j < j1; ++j) {
    double rc = 0.0, ra = 0.0;
    for (int q = p.rp[j]; q < p.rp[j + 1]; ++q) {
      const size_t o = static_cast<size_t>(p.ci[q]) * p.B + b;
      rc = fma(p.va[q], X[o], rc);
      ra = fma(p.va[q], Xsum[o] * inv, ra);
    }
    pc = fma(rc, rc, pc);
    pa = fma(ra, ra, pa);
    const size_t o = static_cast<size_t>(j) * p.B + b;
    const double ec = Y[o] - Yr[o], ea = Ysum[o] * inv - Yr[o];
    dyc = fma(ec, ec, dyc);
    dya = fma(ea, ea, dya);
  }
  double* out = part + static_cast<size_t>(chunk) * 4 * p.B + b;
  out[0] = pc;
  out[p.B] = pa;
  out[2 * p.B] = dyc;
  out[3 * p.B] = dya;
}

// Partial sums over column chunks: c^T x, the dual objective
// sum_i (lambda_i >= 0 ? l_i : u_i) lambda_i with lambda = c - A^T y, and
// |x - x_r|^2, each for the current and the averaged iterate.
__global__ __launch_bounds__(256) void kkt_cols_kernel(Lp p, const double* __restrict__ X,
                                                       const double* __restrict__ Xsum, const double* __restrict__ Xr,
                                                       const double* __restrict__ Y, const double* __restrict__ Ysum,
                                                       const double* __restrict__ den, double* __restrict__ part) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  if (b >= p.B) return;
  const double inv = 1.0 / den[b];
  double poc = 0.0, poa = 0.0, doc = 0.0, doa = 0.0, dxc = 0.0, dxa = 0.0;
  const int i1 = std::min(p.n, (chunk + 1) * kChunk);
  for (int i = chunk * kChunk; i < i1; ++i) {
    double gc = 0.0, ga = 0.0;
    for (int q = p.cp[i]; q < p.cp[i + 1]; ++q) {
      const size_t o = static_cast<size_t>(p.ri[q]) * p.B + b;
      gc = fma(p.vc[q], Y[o], gc);
      ga = fma(p.vc[q], Ysum[o] * inv, ga);
    }
    const double c = objective(p, i, b);
    const size_t o = static_cast<size_t>(i) * p.B + b;
    const double xc = X[o], xa = Xsum[o] * inv;
    poc = fma(c, xc, poc);
    poa = fma(c, xa, poa);
    const double lc = c - gc, la = c - ga;
    doc += (lc >= 0.0 ? p.l[i] : p.u[i]) * lc;
    doa += (la >= 0.0 ? p.l[i] : p.u[i]) * la;
    const double ec = xc - Xr[o], ea = xa - Xr[o];
    dxc = fma(ec, ec, dxc);
    dxa = fma(ea, ea, dxa);
  }
  double* out = part + static_cast<size_t>(chunk) * 6 * p.B + b;
  out[0] = poc;
  out[p.B] = poa;
  out[2 * p.B] = doc;
  out[3 * p.B] = doa;
  out[4 * p.B] = dxc;
  out[5 * p.B] = dxa;
}

struct ColState {
  double* w;      // primal weight
  double* tau;
  double* sig;
  double* kr;     // KKT error at the last restart
  double* kprev;  // KKT error of the last candidate
  double* den;    // iterates in the running average
  int* since;     // iterations since the last restart
  int* done;
  int* action;    // 0 none, 1 restart to the current iterate, 2 to the average
};

// One thread per column: KKT errors of the current and averaged iterates,
// the restart decision and the primal-weight update.
__global__ __launch_bounds__(256) void decide_kernel(int B, int rchunks, int cchunks, const double* __restrict__ prow,
                                                     const double* __restrict__ pcol, ColState cs, double eta,
                                                     int checked, long long total_iters, double tol, int first) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  cs.action[b] = 0;
  if (cs.done[b]) return;
  double s[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int c = 0; c < rchunks; ++c)
    for (int k = 0; k < 4; ++k) s[k] += prow[(static_cast<size_t>(c) * 4 + k) * B + b];
  for (int c = 0; c < cchunks; ++c)
    for (int k = 0; k < 6; ++k) s[4 + k] += pcol[(static_cast<size_t>(c) * 6 + k) * B + b];
  const double prc = sqrt(s[0]), pra = sqrt(s[1]);
  const double gapc = fabs(s[4] - s[6]), gapa = fabs(s[5] - s[7]);
  const double kc = sqrt(prc * prc + gapc * gapc), ka = sqrt(pra * pra + gapa * gapa);
  const bool avg = ka < kc;
  const double kcand = avg ? ka : kc;
  if (first) {
    cs.kr[b] = kcand;
    cs.kprev[b] = kcand;
    return;
  }
  const double pr = avg ? pra : prc, gap = avg ? gapa : gapc;
  const double po = avg ? s[5] : s[4], dob = avg ? s[7] : s[6];
  const bool converged = pr <= tol && gap <= tol * (1.0 + fabs(po) + fabs(dob));
  cs.since[b] += checked;
  const double kr = cs.kr[b];
  const bool restart = converged || kcand <= 0.2 * kr || (kcand <= 0.8 * kr && kcand > cs.kprev[b]) ||
                       static_cast<double>(cs.since[b]) >= 0.36 * static_cast<double>(total_iters);
  cs.kprev[b] = kcand;
  if (!restart) return;
  cs.action[b] = avg ? 2 : 1;
  const double dx = sqrt(avg ? s[9] : s[8]), dy = sqrt(avg ? s[3] : s[2]);
  if (dx > 1e-10 && dy > 1e-10) {
    const double w = exp(0.5 * log(dy / dx) + 0.5 * log(cs.w[b]));
    cs.w[b] = w;
    cs.tau[b] = eta / w;
    cs.sig[b] = eta * w;
  }
  cs.kr[b] = kcand;
  cs.since[b] = 0;
  if (converged) cs.done[b] = 1;
}

// Applies the restarts: the chosen iterate becomes current, restart point and
// the start of a new average.
__global__ __launch_bounds__(256) void restart_kernel(size_t rows, int B, double* __restrict__ Z,
                                                      double* __restrict__ Zsum, double* __restrict__ Zr,
                                                      const int* __restrict__ action, const double* __restrict__ den) {
  const size_t idx = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
  if (idx >= rows * B) return;
  const int b = static_cast<int>(idx % B);
  const int a = action[b];
  if (a == 0) return;
  const double z = a == 2 ? Zsum[idx] / den[b] : Z[idx];
  Z[idx] = z;
  Zsum[idx] = z;
  Zr[idx] = z;
}

// Counts of the running averages: + iters for every live column, or 1 for
// the columns just restarted (reset).
__global__ __launch_bounds__(256) void count_kernel(int B, double* __restrict__ den, const int* __restrict__ action,
                                                    const int* __restrict__ done, int iters, int reset) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (reset) {
    if (action[b] != 0) den[b] = 1.0;
  } else if (!done[b]) {
    den[b] += iters;
  }
}

__global__ __launch_bounds__(256) void count_done_kernel(int B, const int* __restrict__ done, int* __restrict__ out) {
  __shared__ int s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  int c = 0;
  for (int b = threadIdx.x; b < B; b += blockDim.x) c += done[b];
  atomicAdd(&s, c);
  __syncthreads();
  if (threadIdx.x == 0) *out = s;
}

unsigned blocks_for(size_t count) { return static_cast<unsigned>((count + 255) / 256); }

int warmup_generate(const achr_model* M, int nPts, const achr_warmup_options* opts, double* warmup, double* fva,
                    int* converged, double* report);

}  // namespace

extern "C" {

void achr_warmup_options_default(achr_warmup_options* o) {
  if (!o) return;
  o->max_iters = 40000;
  o->check_every = 64;
  o->tol = 1e-9;
  o->seed = 1;
  o->polish = 1;
}

int achr_warmup_generate(const achr_model* M, int nPts, const achr_warmup_options* opts, double* warmup, double* fva,
                         int* converged, double* report) {
  return achr_host::no_throw([&] { return warmup_generate(M, nPts, opts, warmup, fva, converged, report); });
}

}  // extern "C"

namespace {

// The LPs the iteration left unconverged (done[b] == 0), finished exactly
// on the host (achr_polish.cpp) in the coordinates of an orthonormal null-space
// basis of S: from the model's own, or one computed for the purpose when the
// model has none. Feasible reference point: the mean of the converged
// solutions (of all, when none converged). xs: n x B, reaction major.
int polish_unconverged(const achr_model* M, int nf, int R, const std::vector<double>& Cr, std::vector<double>& xs,
                       std::vector<int>& done, int* polished, double* ms) {
  const auto t0 = std::chrono::steady_clock::now();
  const int n = M->n, B = static_cast<int>(done.size());
  std::vector<double> N;
  int d = M->d;
  if (d >= 0) {
    N = M->N;
  } else {
    achr_model tmp = *M;
    const int rc = achr_model_compute_nullspace(&tmp, &d);
    if (rc != ACHR_OK) return rc;
    N.swap(tmp.N);
  }
  *polished = 0;
  if (static_cast<double>(n) * d > 3e7) return ACHR_OK;  // beyond the host method's reach
  std::vector<double> Nr(static_cast<size_t>(n) * std::max(d, 1));
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < n; ++i) Nr[static_cast<size_t>(i) * d + j] = N[static_cast<size_t>(j) * n + i];
  std::vector<int> todo;
  int nconv = 0;
  for (int b = 0; b < B; ++b) {
    if (done[b]) ++nconv;
    else todo.push_back(b);
  }
  std::vector<double> xf(n, 0.0);
  for (int i = 0; i < n; ++i) {
    const double* row = xs.data() + static_cast<size_t>(i) * B;
    double s = 0.0;
    for (int b = 0; b < B; ++b)
      if (done[b] || nconv == 0) s += row[b];
    xf[i] = s / (nconv ? nconv : B);
  }
  std::vector<int> ok(todo.size(), 0);
  const unsigned hw = std::thread::hardware_concurrency();
  achr_host::polish_many(static_cast<int>(todo.size()), static_cast<int>(std::min(hw ? hw : 4u, 16u)), [&](int q) {
    const int b = todo[q];
    std::vector<double> c(n, 0.0), x(n);
    const double sgn = (b & 1) ? -1.0 : 1.0;
    if (b < 2 * nf) c[b >> 1] = sgn;
    else
      for (int i = 0; i < nf; ++i) c[i] = sgn * Cr[static_cast<size_t>(i) * R + (b - 2 * nf)];
    for (int i = 0; i < n; ++i) x[i] = xs[static_cast<size_t>(i) * B + b];
    if (achr_host::polish_lp(n, d, Nr.data(), M->lb.data(), M->ub.data(), c.data(), xf.data(), x.data(),
                             50 * d + 1000) == 0) {
      for (int i = 0; i < n; ++i) xs[static_cast<size_t>(i) * B + b] = x[i];
      ok[q] = 1;
    }
  });
  for (size_t q = 0; q < todo.size(); ++q)
    if (ok[q]) {
      done[todo[q]] = 1;
      ++*polished;
    }
  *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ACHR_OK;
}

int warmup_generate(const achr_model* M, int nPts, const achr_warmup_options* opts, double* warmup, double* fva,
                    int* converged, double* report) {
  if (!M || !warmup) return fail(ACHR_E_ARG, "NULL argument");
  achr_warmup_options o;
  achr_warmup_options_default(&o);
  if (opts) o = *opts;
  // FVA and random objectives range over the file's columns; the auxiliary
  // columns of a model read with its rows honoured only carry its rows.
  const int n = M->n, m = M->m, nf = M->n_file;
  if (nPts < 2 * nf)
    return fail(ACHR_E_ARG, "nPts=%d is below 2 n = %d (one min and one max per reaction)", nPts, 2 * nf);
  if (nPts % 2) return fail(ACHR_E_ARG, "nPts=%d must be even (a min and a max per objective)", nPts);
  if (o.max_iters <= 0 || o.check_every <= 0 || !(o.tol > 0.0)) return fail(ACHR_E_ARG, "bad warmup options");
  const int B = nPts;
  if (static_cast<double>(n) * B >= 9.0e18 / 8) return fail(ACHR_E_ARG, "problem too large");
  const auto t0 = std::chrono::steady_clock::now();

  // Rescaling A~ = Dr S Dc: Ruiz (rows and columns to unit inf-norm), then
  // Pock-Chambolle (alpha = 1). Empty rows or columns keep scale 1.
  const int nnz = M->rowptr[m];
  std::vector<double> va(M->vals.begin(), M->vals.begin() + nnz), dr(m, 1.0), dc(n, 1.0);
  auto scale_pass = [&](bool inf_norm) {
    std::vector<double> rs(m, 0.0), cs(n, 0.0);
    for (int j = 0; j < m; ++j)
      for (int q = M->rowptr[j]; q < M->rowptr[j + 1]; ++q) {
        const double a = std::fabs(va[q]);
        const int i = M->colidx[q];
        if (inf_norm) {
          rs[j] = std::max(rs[j], a);
          cs[i] = std::max(cs[i], a);
        } else {
          rs[j] += a;
          cs[i] += a;
        }
      }
    for (int j = 0; j < m; ++j) rs[j] = rs[j] > 0.0 ? std::sqrt(rs[j]) : 1.0;
    for (int i = 0; i < n; ++i) cs[i] = cs[i] > 0.0 ? std::sqrt(cs[i]) : 1.0;
    for (int j = 0; j < m; ++j)
      for (int q = M->rowptr[j]; q < M->rowptr[j + 1]; ++q) va[q] /= rs[j] * cs[M->colidx[q]];
    for (int j = 0; j < m; ++j) dr[j] /= rs[j];
    for (int i = 0; i < n; ++i) dc[i] /= cs[i];
  };
  for (int k = 0; k < 20; ++k) scale_pass(true);
  scale_pass(false);
  // CSC of the scaled matrix
  std::vector<int> cp(n + 1, 0), ri(nnz);
  std::vector<double> vc(nnz);
  for (int q = 0; q < nnz; ++q) ++cp[M->colidx[q] + 1];
  for (int i = 0; i < n; ++i) cp[i + 1] += cp[i];
  {
    std::vector<int> pos(cp.begin(), cp.end() - 1);
    for (int j = 0; j < m; ++j)
      for (int q = M->rowptr[j]; q < M->rowptr[j + 1]; ++q) {
        const int at = pos[M->colidx[q]]++;
        ri[at] = j;
        vc[at] = va[q];
      }
  }
  // ||A~||_2 by power iteration on A~^T A~
  double nrm = 1.0;
  {
    std::vector<double> v(n), t(m), z(n);
    for (int i = 0; i < n; ++i) v[i] = 1.0 + 0.01 * (i % 7);
    for (int it = 0; it < 80; ++it) {
      for (int j = 0; j < m; ++j) {
        double a = 0.0;
        for (int q = M->rowptr[j]; q < M->rowptr[j + 1]; ++q) a += va[q] * v[M->colidx[q]];
        t[j] = a;
      }
      std::fill(z.begin(), z.end(), 0.0);
      for (int j = 0; j < m; ++j)
        for (int q = M->rowptr[j]; q < M->rowptr[j + 1]; ++q) z[M->colidx[q]] += va[q] * t[j];
      double s = 0.0;
      for (double x : z) s += x * x;
      s = std::sqrt(s);
      if (!(s > 0.0)) break;
      nrm = std::sqrt(s);
      for (int i = 0; i < n; ++i) v[i] = z[i] / s;
    }
  }
  const double eta = 0.95 / nrm;
  std::vector<double> ls(n), us(n);
  for (int i = 0; i < n; ++i) {
    ls[i] = M->lb[i] / dc[i];
    us[i] = M->ub[i] / dc[i];
  }
  // random objectives (one per column beyond the FVA pairs), Philox U(-0.5, 0.5)
  const int R = B - 2 * nf;
  std::vector<double> Cr(static_cast<size_t>(n) * std::max(R, 1), 0.0);
  for (int r = 0; r < R; ++r)
    for (int i = 0; i < nf; ++i)
      Cr[static_cast<size_t>(i) * R + r] =
          achr::stream_u01(o.seed, static_cast<uint32_t>(i), static_cast<uint32_t>(r), 0, 0, kTagObjective) - 0.5;

  DevBuf<int> d_rp, d_ci, d_cp, d_ri, d_done, d_since, d_action, d_cnt;
  DevBuf<double> d_va, d_vc, d_l, d_u, d_dc, d_Cr, X, Xbar, Xsum, Xr, Y, Ysum, Yr, prow, pcol;
  DevBuf<double> w, tau, sig, kr, kprev, den;
  HIP_TRY(d_rp.upload(M->rowptr.data(), m + 1, nullptr));
  HIP_TRY(d_ci.upload(M->colidx.data(), std::max(nnz, 1), nullptr));
  HIP_TRY(d_va.upload(va.data(), std::max(nnz, 1), nullptr));
  HIP_TRY(d_cp.upload(cp.data(), n + 1, nullptr));
  HIP_TRY(d_ri.upload(ri.data(), std::max(nnz, 1), nullptr));
  HIP_TRY(d_vc.upload(vc.data(), std::max(nnz, 1), nullptr));
  HIP_TRY(d_l.upload(ls.data(), n, nullptr));
  HIP_TRY(d_u.upload(us.data(), n, nullptr));
  HIP_TRY(d_dc.upload(dc.data(), n, nullptr));
  HIP_TRY(d_Cr.upload(Cr.data(), Cr.size(), nullptr));
  const size_t nB = static_cast<size_t>(n) * B, mB = static_cast<size_t>(std::max(m, 1)) * B;
  // start: x = clamp(0), y = 0
  {
    std::vector<double> x0(nB);
    for (int i = 0; i < n; ++i) {
      const double v = std::min(std::max(0.0, ls[i]), us[i]);
      std::fill(x0.begin() + static_cast<size_t>(i) * B, x0.begin() + static_cast<size_t>(i + 1) * B, v);
    }
    HIP_TRY(X.upload(x0.data(), nB, nullptr));
    HIP_TRY(Xsum.upload(x0.data(), nB, nullptr));
    HIP_TRY(Xr.upload(x0.data(), nB, nullptr));
  }
  HIP_TRY(Xbar.ensure(nB));
  HIP_TRY(Y.ensure(mB));
  HIP_TRY(Ysum.ensure(mB));
  HIP_TRY(Yr.ensure(mB));
  HIP_TRY(hipMemset(Y.p, 0, sizeof(double) * mB));
  HIP_TRY(hipMemset(Ysum.p, 0, sizeof(double) * mB));
  HIP_TRY(hipMemset(Yr.p, 0, sizeof(double) * mB));
  const int rchunks = std::max(1, (m + kChunk - 1) / kChunk), cchunks = (n + kChunk - 1) / kChunk;
  HIP_TRY(prow.ensure(static_cast<size_t>(rchunks) * 4 * B));
  HIP_TRY(pcol.ensure(static_cast<size_t>(cchunks) * 6 * B));
  {
    std::vector<double> one(B, 1.0), et(B, eta), zero(B, 0.0);
    HIP_TRY(w.upload(one.data(), B, nullptr));
    HIP_TRY(den.upload(one.data(), B, nullptr));
    HIP_TRY(tau.upload(et.data(), B, nullptr));
    HIP_TRY(sig.upload(et.data(), B, nullptr));
    HIP_TRY(kr.upload(zero.data(), B, nullptr));
    HIP_TRY(kprev.upload(zero.data(), B, nullptr));
  }
  HIP_TRY(d_done.ensure(B));
  HIP_TRY(d_since.ensure(B));
  HIP_TRY(d_action.ensure(B));
  HIP_TRY(d_cnt.ensure(1));
  HIP_TRY(hipMemset(d_done.p, 0, sizeof(int) * B));
  HIP_TRY(hipMemset(d_since.p, 0, sizeof(int) * B));
  HIP_TRY(hipMemset(d_action.p, 0, sizeof(int) * B));

  Lp p{n, m, B, nf, d_rp.p, d_ci.p, d_va.p, d_cp.p, d_ri.p, d_vc.p, d_l.p, d_u.p, d_dc.p, d_Cr.p};
  ColState cs{w.p, tau.p, sig.p, kr.p, kprev.p, den.p, d_since.p, d_done.p, d_action.p};
  const dim3 kb(256), rg(blocks_for(B), rchunks), cg(blocks_for(B), cchunks);
  // iteration kernels: LPs across x, columns / rows across y (grid-strided past 65,535)
  const unsigned bx = static_cast<unsigned>((B + 255) / 256);
  const dim3 pg(bx, static_cast<unsigned>(std::min(n, 65535))), dg(bx, static_cast<unsigned>(std::max(std::min(m, 65535), 1)));
  auto kkt = [&]() {
    hipLaunchKernelGGL(kkt_rows_kernel, rg, kb, 0, nullptr, p, X.p, Xsum.p, Y.p, Ysum.p, Yr.p, den.p, prow.p);
    hipLaunchKernelGGL(kkt_cols_kernel, cg, kb, 0, nullptr, p, X.p, Xsum.p, Xr.p, Y.p, Ysum.p, den.p, pcol.p);
  };
  kkt();
  hipLaunchKernelGGL(decide_kernel, dim3(blocks_for(B)), kb, 0, nullptr, B, rchunks, cchunks, prow.p, pcol.p, cs, eta,
                     0, 1LL, o.tol, 1);
  HIP_TRY(hipGetLastError());
  long long it = 0;
  int ndone = 0;
  int checks = 0;
  while (it < o.max_iters) {
    const int k = static_cast<int>(std::min<long long>(o.check_every, o.max_iters - it));
    for (int s = 0; s < k; ++s) {
      hipLaunchKernelGGL(primal_kernel, pg, kb, 0, nullptr, p, X.p, Xbar.p, Xsum.p, Y.p, tau.p, d_done.p);
      if (m > 0) hipLaunchKernelGGL(dual_kernel, dg, kb, 0, nullptr, p, Y.p, Ysum.p, Xbar.p, sig.p, d_done.p);
    }
    it += k;
    // the averages hold den + k iterates now
    hipLaunchKernelGGL(count_kernel, dim3(blocks_for(B)), kb, 0, nullptr, B, den.p, d_action.p, d_done.p, k, 0);
    kkt();
    hipLaunchKernelGGL(decide_kernel, dim3(blocks_for(B)), kb, 0, nullptr, B, rchunks, cchunks, prow.p, pcol.p, cs,
                       eta, k, it, o.tol, 0);
    hipLaunchKernelGGL(restart_kernel, dim3(blocks_for(nB)), kb, 0, nullptr, static_cast<size_t>(n), B, X.p, Xsum.p,
                       Xr.p, d_action.p, den.p);
    if (m > 0)
      hipLaunchKernelGGL(restart_kernel, dim3(blocks_for(mB)), kb, 0, nullptr, static_cast<size_t>(m), B, Y.p, Ysum.p,
                         Yr.p, d_action.p, den.p);
    hipLaunchKernelGGL(count_kernel, dim3(blocks_for(B)), kb, 0, nullptr, B, den.p, d_action.p, d_done.p, 0, 1);
    HIP_TRY(hipGetLastError());
    if (++checks % 8 == 0 || it >= o.max_iters) {
      hipLaunchKernelGGL(count_done_kernel, dim3(1), kb, 0, nullptr, B, d_done.p, d_cnt.p);
      HIP_TRY(hipMemcpy(&ndone, d_cnt.p, sizeof(int), hipMemcpyDeviceToHost));
      if (ndone == B) break;
    }
  }
  HIP_TRY(hipDeviceSynchronize());
  const double solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();

  // points: unscale, clip to the bounds (movePtsBds), FVA columns, then the
  // move toward the centre x <- 0.33 x + 0.67 c (createWarmupPtsGLPK.c:36-65)
  std::vector<double> xs(nB);
  std::vector<int> done(B);
  HIP_TRY(hipMemcpy(xs.data(), X.p, sizeof(double) * nB, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(done.data(), d_done.p, sizeof(int) * B, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i) {
    double* row = xs.data() + static_cast<size_t>(i) * B;
    for (int b = 0; b < B; ++b) {
      double v = row[b] * dc[i];
      if (v < M->lb[i]) v = M->lb[i];
      else if (v > M->ub[i]) v = M->ub[i];
      row[b] = v;
    }
  }
  int polished = 0;
  double polish_ms = 0.0;
  if (o.polish && ndone < B) {
    const int rc = polish_unconverged(M, nf, R, Cr, xs, done, &polished, &polish_ms);
    if (rc != ACHR_OK) return rc;
    ndone += polished;
  }
  if (converged) std::copy(done.begin(), done.end(), converged);
  std::vector<double> cen(n, 0.0);
  for (int i = 0; i < n; ++i) {
    const double* row = xs.data() + static_cast<size_t>(i) * B;
    if (fva) {
      fva[i] = i < nf ? row[2 * i] : NAN;
      fva[static_cast<size_t>(n) + i] = i < nf ? row[2 * i + 1] : NAN;
    }
    double sum = 0.0;
    for (int b = 0; b < B; ++b) sum += row[b];
    cen[i] = sum / B;
  }
  for (int i = 0; i < n; ++i) {
    const double* row = xs.data() + static_cast<size_t>(i) * B;
    for (int b = 0; b < B; ++b) warmup[static_cast<size_t>(b) * n + i] = row[b] * 0.33 + 0.67 * cen[i];
  }
  if (report) {
    report[0] = static_cast<double>(it);
    report[1] = static_cast<double>(ndone);
    report[2] = solve_ms;
    report[3] = nrm;
    report[4] = static_cast<double>(polished);
    report[5] = polish_ms;
  }
  return ACHR_OK;
}

}  // namespace
