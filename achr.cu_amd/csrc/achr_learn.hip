// Learned direction law: network handles and the sample -> rounding map ->
// direction bank pipeline (declared in include/achr.h). Host code; the
// device work is the matrix-core GEMM, the FP64 column GEMM, the network's
// kernels and rocSOLVER for one small symmetric eigenproblem.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>
#include <vector>

#include "achr_dirnet.h"
#include "achr_gemm.h"
#include "achr_host.h"
#include "achr_internal.h"

using achr_host::DevBuf;
using achr_host::fail;
using achr_host::on_threads;

struct achr_dirnet {
  achr::DirNet net;
  bool has_target = false;
  ~achr_dirnet() { net.release(); }
};

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int upload_target(achr_dirnet* d, const float* C) {
  achr::DirNet& n = d->net;
  HIP_TRY(n.ensure_cov());
  HIP_TRY(hipMemcpy(n.target, C, sizeof(float) * n.O * n.O, hipMemcpyHostToDevice));
  double s = 0.0;
  for (size_t i = 0; i < static_cast<size_t>(n.O) * n.O; ++i) s += static_cast<double>(C[i]) * C[i];
  n.target_norm2 = static_cast<float>(s > 0.0 ? s : 1.0);
  d->has_target = true;
  return ACHR_OK;
}

// loss from the error matrix left in net.cov by backward()
int read_loss(achr::DirNet& n, float* loss) {
  std::vector<float> E(static_cast<size_t>(n.O) * n.O);
  HIP_TRY(hipMemcpy(E.data(), n.cov, sizeof(float) * E.size(), hipMemcpyDeviceToHost));
  double s = 0.0;
  for (float e : E) s += static_cast<double>(e) * e;
  *loss = static_cast<float>(s / n.target_norm2);
  return ACHR_OK;
}

}  // namespace

extern "C" {

int achr_dirnet_create(int latent, int width, int outdim, uint64_t seed, achr_dirnet** out) {
  if (!out || latent <= 0 || width <= 0 || outdim <= 0) return fail(ACHR_E_ARG, "bad network shape");
  *out = nullptr;
  auto* d = new (std::nothrow) achr_dirnet();
  if (!d) return fail(ACHR_E_NOMEM, "network allocation");
  hipError_t e = d->net.alloc(latent, width, outdim);
  if (e != hipSuccess) {
    delete d;
    return fail(ACHR_E_DEVICE, "network allocation: %s", hipGetErrorString(e));
  }
  std::vector<float> host;
  d->net.init_params(seed, host);
  e = hipMemcpy(d->net.params, host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    delete d;
    return fail(ACHR_E_DEVICE, "network upload: %s", hipGetErrorString(e));
  }
  *out = d;
  return ACHR_OK;
}

void achr_dirnet_destroy(achr_dirnet* net) { delete net; }

int achr_dirnet_param_count(const achr_dirnet* d, long* count) {
  if (!d || !count) return fail(ACHR_E_ARG, "NULL argument");
  *count = static_cast<long>(d->net.param_count());
  return ACHR_OK;
}

int achr_dirnet_get_params(const achr_dirnet* d, float* params) {
  if (!d || !params) return fail(ACHR_E_ARG, "NULL argument");
  HIP_TRY(hipMemcpy(params, d->net.params, sizeof(float) * d->net.param_count(), hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_dirnet_set_params(achr_dirnet* d, const float* params) {
  if (!d || !params) return fail(ACHR_E_ARG, "NULL argument");
  HIP_TRY(hipMemcpy(d->net.params, params, sizeof(float) * d->net.param_count(), hipMemcpyHostToDevice));
  return ACHR_OK;
}

int achr_dirnet_set_target(achr_dirnet* d, const float* C) {
  if (!d || !C) return fail(ACHR_E_ARG, "NULL argument");
  return upload_target(d, C);
}

int achr_dirnet_loss_grad(achr_dirnet* d, const float* Z, int B, float* loss, float* grads) {
  if (!d || !Z || B <= 0) return fail(ACHR_E_ARG, "bad arguments");
  if (!d->has_target) return fail(ACHR_E_STATE, "no target covariance");
  achr::DirNet& n = d->net;
  HIP_TRY(n.ensure_batch(B));
  HIP_TRY(hipMemcpy(n.z, Z, sizeof(float) * B * n.L, hipMemcpyHostToDevice));
  HIP_TRY(n.forward(B, achr::kPrecF32, nullptr));
  HIP_TRY(n.backward(B, n.target_norm2, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  if (loss) {
    const int rc = read_loss(n, loss);
    if (rc != ACHR_OK) return rc;
  }
  if (grads) HIP_TRY(hipMemcpy(grads, n.grads, sizeof(float) * n.param_count(), hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_dirnet_train(achr_dirnet* d, int iters, int batch, float lr, uint64_t seed, float* losses) {
  if (!d || iters < 0 || batch <= 0 || !(lr > 0.0f)) return fail(ACHR_E_ARG, "bad training arguments");
  if (!d->has_target) return fail(ACHR_E_STATE, "no target covariance");
  achr::DirNet& n = d->net;
  HIP_TRY(n.ensure_batch(batch));
  for (int it = 0; it < iters; ++it) {
    HIP_TRY(n.gaussian_latents(batch, seed, static_cast<uint32_t>(it), nullptr));
    HIP_TRY(n.forward(batch, achr::kPrecF32, nullptr));
    HIP_TRY(n.backward(batch, n.target_norm2, nullptr));
    if (losses) {
      HIP_TRY(hipDeviceSynchronize());
      const int rc = read_loss(n, losses + it);
      if (rc != ACHR_OK) return rc;
    }
    HIP_TRY(n.adam_step(lr, ++n.adam_t, nullptr));
  }
  HIP_TRY(hipDeviceSynchronize());
  return ACHR_OK;
}

int achr_dirnet_sample(achr_dirnet* d, const float* Z, int B, int prec, float* V) {
  if (!d || !Z || !V || B <= 0 || (prec != achr::kPrecF32 && prec != achr::kPrecBF16))
    return fail(ACHR_E_ARG, "bad arguments");
  achr::DirNet& n = d->net;
  HIP_TRY(n.ensure_batch(B));
  HIP_TRY(hipMemcpy(n.z, Z, sizeof(float) * B * n.L, hipMemcpyHostToDevice));
  HIP_TRY(n.forward(B, prec, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(V, n.v, sizeof(float) * B * n.O, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

void achr_learn_options_default(achr_learn_options* o) {
  if (!o) return;
  o->rank = 512;
  o->width = 512;
  o->iters = 400;
  o->batch = 2048;
  o->lr = 1e-3f;
  o->bank_size = 0;  // 0: 2n directions, like the reference's 2n warmup points
  o->isotropic = 0;
  o->seed = 7;
}

int achr_sampler_set_directions(achr_sampler* S, const double* bank, int K) {
  if (!S || !bank || K <= 0) return fail(ACHR_E_ARG, "bad bank");
  if (S->sched.tpc < 64 && 8.0 * S->n * K >= 2147483648.0) return fail(ACHR_E_ARG, "bank too large for the schedule");
  HIP_TRY(S->dirs.upload(bank, static_cast<size_t>(S->n) * K, nullptr));
  S->K = K;
  return ACHR_OK;
}

int achr_sampler_get_directions(const achr_sampler* S, double* bank, int* K) {
  if (!S || !K) return fail(ACHR_E_ARG, "NULL argument");
  *K = S->K;
  if (bank && S->K > 0)
    HIP_TRY(hipMemcpy(bank, S->dirs.p, sizeof(double) * static_cast<size_t>(S->n) * S->K, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_sampler_set_run_options(achr_sampler* S, int trace_chains, int trace_thin, int check_every) {
  if (!S || trace_chains < 0 || (trace_chains > 0 && trace_thin <= 0) || check_every <= 0)
    return fail(ACHR_E_ARG, "bad run options");
  S->opt.trace_chains = trace_chains;
  S->opt.trace_thin = trace_thin;
  S->opt.check_every = check_every;
  return ACHR_OK;
}

// Start of bank-mode chains: the point of {x : S x = 0, x_i = lb_i for every
// fixed reaction (lb_i = ub_i)} nearest the warmup centre c0. Bank directions
// keep both conditions, so chains started there stay on that set up to
// rounding and the residual check never has to reproject. (c0 is off it by
// the warmup file's rounding, and a chain started at c0 would undo every
// reprojection at its next clip of the fixed reactions.) Host FP64, with the
// orthonormal null-space basis N: y = N N^T c0 is the nearest point of
// null(S); the nearest one that also meets x_F = lb_F is y + N da with da the
// least-norm solution of N_F da = lb_F - y_F (N_F: the fixed rows of N,
// possibly dependent). A pivoted, reorthogonalised Gram-Schmidt pass over the
// rows of N_F picks independent rows P and gives N_P = L Q (Q orthonormal, L
// lower triangular), so da = Q^T L^{-1} (lb_P - y_P). Finally x_F = lb_F.
static int compute_bank_start(achr_sampler* S) {
  const achr_model* M = S->model;
  const int n = S->n, d = S->d;
  std::vector<double> x = S->c0;
  if (d > 0 && M->N.size() == static_cast<size_t>(n) * d) {
    const double* N = M->N.data();
    auto col = [&](int j) { return N + static_cast<size_t>(j) * n; };
    std::vector<double> t(d, 0.0);
    for (int j = 0; j < d; ++j) {
      const double* cj = col(j);
      double acc = 0.0;
      for (int i = 0; i < n; ++i) acc += cj[i] * S->c0[i];
      t[j] = acc;
    }
    std::fill(x.begin(), x.end(), 0.0);
    for (int j = 0; j < d; ++j) {
      const double* cj = col(j);
      for (int i = 0; i < n; ++i) x[i] += cj[i] * t[j];
    }
    std::vector<int> F;
    for (int i = 0; i < n; ++i)
      if (M->lb[i] == M->ub[i]) F.push_back(i);
    const int f = static_cast<int>(F.size());
    if (f > 0) {
      auto dot = [d](const double* a, const double* b) {
        double s = 0.0;
        for (int j = 0; j < d; ++j) s += a[j] * b[j];
        return s;
      };
      std::vector<double> NF(static_cast<size_t>(f) * d), R;  // f x d row major
      for (int r = 0; r < f; ++r)
        for (int j = 0; j < d; ++j) NF[static_cast<size_t>(r) * d + j] = col(j)[F[r]];
      R = NF;
      double scale = 0.0;
      for (int r = 0; r < f; ++r) scale = std::max(scale, std::sqrt(dot(&NF[static_cast<size_t>(r) * d], &NF[static_cast<size_t>(r) * d])));
      std::vector<int> piv, rest(f);
      for (int r = 0; r < f; ++r) rest[r] = r;
      std::vector<double> Q;  // rank x d row major
      while (!rest.empty()) {
        int best = -1;
        double bn = 0.0;
        for (int k = 0; k < static_cast<int>(rest.size()); ++k) {
          const double* rr = &R[static_cast<size_t>(rest[k]) * d];
          const double v = std::sqrt(dot(rr, rr));
          if (v > bn) { bn = v; best = k; }
        }
        if (best < 0 || !(bn > 1e-10 * scale)) break;
        const int j = rest[best];
        rest.erase(rest.begin() + best);
        std::vector<double> q(R.begin() + static_cast<size_t>(j) * d, R.begin() + static_cast<size_t>(j + 1) * d);
        for (int k = 0; k < static_cast<int>(piv.size()); ++k) {  // second orthogonalisation
          const double* qk = &Q[static_cast<size_t>(k) * d];
          const double c = dot(qk, q.data());
          for (int jj = 0; jj < d; ++jj) q[jj] -= c * qk[jj];
        }
        const double qn = std::sqrt(dot(q.data(), q.data()));
        for (double& v : q) v /= qn;
        for (int rr : rest) {
          double* row = &R[static_cast<size_t>(rr) * d];
          const double c = dot(row, q.data());
          for (int jj = 0; jj < d; ++jj) row[jj] -= c * q[jj];
        }
        Q.insert(Q.end(), q.begin(), q.end());
        piv.push_back(j);
      }
      const int rk = static_cast<int>(piv.size());
      // L z = lb_P - y_P with L[a][b] = N_F[piv[a]] . Q[b] (b <= a)
      std::vector<double> z(rk, 0.0);
      for (int a = 0; a < rk; ++a) {
        const double* row = &NF[static_cast<size_t>(piv[a]) * d];
        double v = M->lb[F[piv[a]]] - x[F[piv[a]]];
        for (int b = 0; b < a; ++b) v -= dot(row, &Q[static_cast<size_t>(b) * d]) * z[b];
        z[a] = v / dot(row, &Q[static_cast<size_t>(a) * d]);
      }
      std::vector<double> da(d, 0.0);
      for (int a = 0; a < rk; ++a)
        for (int jj = 0; jj < d; ++jj) da[jj] += Q[static_cast<size_t>(a) * d + jj] * z[a];
      for (int j = 0; j < d; ++j) {
        const double* cj = col(j);
        for (int i = 0; i < n; ++i) x[i] += cj[i] * da[j];
      }
      for (int k = 0; k < f; ++k) x[F[k]] = M->lb[F[k]];
    }
  }
  S->s0 = x;
  HIP_TRY(S->s0d.upload(S->s0.data(), static_cast<size_t>(n), nullptr));
  HIP_TRY(hipDeviceSynchronize());
  return ACHR_OK;
}

// Warmup points onto {x : S x = 0, x_F = lb_F}: w <- s0 + N_f N_f^T (w - s0)
// with s0 the bank start (a point of that set) and N_f the orthonormal basis
// of its directions, both FP64 products on the matrix cores in the oracle's
// dot-product order; the fixed coordinates are then set to their bound
// exactly. The centre is recomputed from the moved points as at creation.
static int free_basis(const achr_model* M, const double* dN, DevBuf<double>& Nf, int* df_out);

static int project_warmup(achr_sampler* S) {
  const achr_model* M = S->model;
  const int n = S->n, W = S->W;
  if (S->d <= 0) return fail(ACHR_E_STATE, "the model has no null space to project onto");
  if (S->s0.empty()) {
    const int rc = compute_bank_start(S);
    if (rc != ACHR_OK) return rc;
  }
  DevBuf<double> Nf, D, T;
  int df = 0;
  int rc = free_basis(M, S->Ncm.p, Nf, &df);
  if (rc != ACHR_OK) return rc;
  HIP_TRY(D.ensure(static_cast<size_t>(n) * W));
  HIP_TRY(T.ensure(static_cast<size_t>(df) * W));
  DevBuf<int> wcount;
  HIP_TRY(wcount.upload(&W, 1, nullptr));
  HIP_TRY(achr::launch_sub_cols(S->warm.p, S->s0d.p, D.p, n, W, nullptr));
  // T(:, k) = N_f^T D(:, k): A(r, i) = N_f[i + r*n]
  HIP_TRY(achr::launch_dgemm_cols(df, n, Nf.p, n, 1, D.p, n, T.p, df, wcount.p, W, nullptr));
  // W(:, k) = N_f T(:, k) + s0: A(i, r) = N_f[i + r*n]
  HIP_TRY(achr::launch_dgemm_cols(n, df, Nf.p, 1, n, T.p, df, S->warm.p, n, wcount.p, W, nullptr));
  HIP_TRY(achr::launch_add_cols_fix(S->warm.p, S->s0d.p, S->lb.p, S->ub.p, n, W, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  // the centre of the moved points, summed in point order as at creation
  std::vector<double> sum(n, 0.0), chunk;
  const int per = std::max(1, static_cast<int>((64LL << 20) / (8LL * n)));
  for (int k0 = 0; k0 < W; k0 += per) {
    const int cnt = std::min(per, W - k0);
    chunk.resize(static_cast<size_t>(cnt) * n);
    HIP_TRY(hipMemcpy(chunk.data(), S->warm.p + static_cast<size_t>(k0) * n, sizeof(double) * chunk.size(),
                      hipMemcpyDeviceToHost));
    for (int k = 0; k < cnt; ++k)
      for (int i = 0; i < n; ++i) sum[i] += chunk[static_cast<size_t>(k) * n + i];
  }
  const double alpha = 1.0 / static_cast<double>(W);
  for (int i = 0; i < n; ++i) S->c0[i] = alpha * sum[i];
  HIP_TRY(S->c0d.upload(S->c0.data(), static_cast<size_t>(n), nullptr));
  S->s0.clear();  // the bank start follows the new centre when a bank is switched on
  S->have_run = false;
  return ACHR_OK;
}

int achr_sampler_project_warmup(achr_sampler* S) {
  if (!S) return fail(ACHR_E_ARG, "sampler is NULL");
  return achr_host::no_throw([&] { return project_warmup(S); });
}

int achr_sampler_get_warmup(const achr_sampler* S, double* warmup, int* W) {
  if (!S || !W) return fail(ACHR_E_ARG, "NULL argument");
  *W = S->W;
  if (warmup)
    HIP_TRY(hipMemcpy(warmup, S->warm.p, sizeof(double) * static_cast<size_t>(S->n) * S->W, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_sampler_use_directions(achr_sampler* S, int on) {
  if (!S) return fail(ACHR_E_ARG, "sampler is NULL");
  if (on && S->K == 0) return fail(ACHR_E_STATE, "no direction bank");
  if (on && S->s0.empty()) {
    const int rc = achr_host::no_throw([&] { return compute_bank_start(S); });
    if (rc != ACHR_OK) return rc;
  }
  S->use_dirs = on ? 1 : 0;
  return ACHR_OK;
}

// Orthonormal basis N_f of {v in span(N): v_i = 0 for fixed reactions},
// n x df column major FP64 on the device.
static int free_basis(const achr_model* M, const double* dN, DevBuf<double>& Nf, int* df_out) {
  const int n = M->n, d = M->d;
  std::vector<int> fixed;
  for (int i = 0; i < n; ++i)
    if (M->lb[i] == M->ub[i]) fixed.push_back(i);
  const int nf = static_cast<int>(fixed.size());
  if (nf == 0) {
    HIP_TRY(Nf.ensure(static_cast<size_t>(n) * d));
    HIP_TRY(hipMemcpy(Nf.p, dN, sizeof(double) * n * d, hipMemcpyDeviceToDevice));
    *df_out = d;
    return ACHR_OK;
  }
  // The fixed rows N_F (nf x d) span, within the coordinates of N, the
  // directions that move a fixed reaction; N_f = N Q2 with Q2 an orthonormal
  // basis of their complement. The nf x nf Gram N_F N_F^T has the nonzero
  // eigenvalues of N_F^T N_F: its eigenpairs at or above d * lambda_max *
  // 2.2204e-16 give the range R = N_F^T U / sqrt(lambda) (d x r0), and the
  // last d - r0 columns of the Householder Q of R are Q2 (rocSOLVER geqrf and
  // orgqr), in place of an eigen-decomposition of the d x d N_F^T N_F.
  std::vector<double> NF(static_cast<size_t>(nf) * d);
  for (int r = 0; r < nf; ++r)
    for (int j = 0; j < d; ++j) NF[static_cast<size_t>(r) + static_cast<size_t>(j) * nf] = M->N[static_cast<size_t>(fixed[r]) + static_cast<size_t>(j) * n];
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return fail(ACHR_E_SOLVER, "rocblas_create_handle");
  DevBuf<double> dNF, dA, dD, dE, dU, dQ, dtau;
  DevBuf<int> dinfo;
  int rc = ACHR_OK;
  std::vector<double> ev(nf), Ah(static_cast<size_t>(nf) * nf);
  const double one = 1.0, zero = 0.0;
  do {
    if (dNF.upload(NF.data(), NF.size(), nullptr) != hipSuccess || dA.ensure(static_cast<size_t>(nf) * nf) != hipSuccess ||
        dD.ensure(nf) != hipSuccess || dE.ensure(nf) != hipSuccess || dinfo.ensure(1) != hipSuccess) {
      rc = fail(ACHR_E_NOMEM, "free basis buffers");
      break;
    }
    rocblas_int info = 0;
    if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, nf, nf, d, &one, dNF.p, nf, dNF.p, nf,
                      &zero, dA.p, nf) != rocblas_status_success ||
        rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, nf, dA.p, nf, dD.p, dE.p, dinfo.p) !=
            rocblas_status_success ||
        hipMemcpy(&info, dinfo.p, sizeof info, hipMemcpyDeviceToHost) != hipSuccess || info != 0) {
      rc = fail(ACHR_E_SOLVER, "free basis eigenproblem");
      break;
    }
    if (hipMemcpy(ev.data(), dD.p, sizeof(double) * nf, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(Ah.data(), dA.p, sizeof(double) * Ah.size(), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(ACHR_E_DEVICE, "free basis download");
      break;
    }
    double lmax = 0.0;
    for (double v : ev) lmax = std::max(lmax, std::fabs(v));
    const double tol = std::max(1.0, static_cast<double>(d)) * lmax * 2.2204e-16;
    // U / sqrt(lambda) for the range eigenpairs (nf x r0)
    std::vector<double> Us;
    int r0 = 0;
    for (int k = 0; k < nf; ++k) {
      if (!(std::fabs(ev[k]) >= tol) || !(ev[k] > 0.0)) continue;
      const double sc = 1.0 / std::sqrt(ev[k]);
      for (int i = 0; i < nf; ++i) Us.push_back(Ah[static_cast<size_t>(i) + static_cast<size_t>(k) * nf] * sc);
      ++r0;
    }
    const int df = d - r0;
    if (df <= 0) {
      rc = fail(ACHR_E_STATE, "the fixed reactions leave no free direction");
      break;
    }
    if (r0 == 0) {
      if (Nf.ensure(static_cast<size_t>(n) * d) != hipSuccess ||
          hipMemcpy(Nf.p, dN, sizeof(double) * n * d, hipMemcpyDeviceToDevice) != hipSuccess) {
        rc = fail(ACHR_E_NOMEM, "free basis");
        break;
      }
      *df_out = d;
      break;
    }
    // dQ (d x d): R = N_F^T (U / sqrt(lambda)) in its first r0 columns, then Q
    if (dU.upload(Us.data(), Us.size(), nullptr) != hipSuccess || dQ.ensure(static_cast<size_t>(d) * d) != hipSuccess ||
        dtau.ensure(r0) != hipSuccess || Nf.ensure(static_cast<size_t>(n) * df) != hipSuccess) {
      rc = fail(ACHR_E_NOMEM, "free basis");
      break;
    }
    if (rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, d, r0, nf, &one, dNF.p, nf, dU.p, nf,
                      &zero, dQ.p, d) != rocblas_status_success ||
        rocsolver_dgeqrf(h, d, r0, dQ.p, d, dtau.p) != rocblas_status_success ||
        rocsolver_dorgqr(h, d, d, r0, dQ.p, d, dtau.p) != rocblas_status_success ||
        rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, df, d, &one, dN, n,
                      dQ.p + static_cast<size_t>(r0) * d, d, &zero, Nf.p, n) != rocblas_status_success ||
        hipDeviceSynchronize() != hipSuccess) {
      rc = fail(ACHR_E_SOLVER, "free basis product");
      break;
    }
    *df_out = df;
  } while (false);
  rocblas_destroy_handle(h);
  return rc;
}

// Cholesky-QR (twice) of the df x q FP64 matrix G (column major, host): G <- Q.
// The products run on host threads, each element summed in the sequential
// order (columns of A dealt cyclically, G updated by row blocks).
static bool chol_qr(std::vector<double>& G, int rows, int q) {
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<double> A(static_cast<size_t>(q) * q, 0.0);
    on_threads([&](int t, int T) {
      for (int a = t; a < q; a += T)
        for (int b = a; b < q; ++b) {
          const double* ga = G.data() + static_cast<size_t>(a) * rows;
          const double* gb = G.data() + static_cast<size_t>(b) * rows;
          double s = 0.0;
          for (int i = 0; i < rows; ++i) s += ga[i] * gb[i];
          A[a + static_cast<size_t>(b) * q] = s;
          A[b + static_cast<size_t>(a) * q] = s;
        }
    });
    // A = R^T R (upper R)
    std::vector<double> R(static_cast<size_t>(q) * q, 0.0);
    for (int j = 0; j < q; ++j) {
      double s = A[j + static_cast<size_t>(j) * q];
      for (int k = 0; k < j; ++k) s -= R[k + static_cast<size_t>(j) * q] * R[k + static_cast<size_t>(j) * q];
      if (!(s > 0.0)) return false;
      const double rjj = std::sqrt(s);
      R[j + static_cast<size_t>(j) * q] = rjj;
      for (int i = j + 1; i < q; ++i) {
        double t = A[j + static_cast<size_t>(i) * q];
        for (int k = 0; k < j; ++k) t -= R[k + static_cast<size_t>(j) * q] * R[k + static_cast<size_t>(i) * q];
        R[j + static_cast<size_t>(i) * q] = t / rjj;
      }
    }
    // G <- G R^{-1}, column by column within each thread's block of rows
    on_threads([&](int t, int T) {
      const int i0 = static_cast<int>(static_cast<long long>(rows) * t / T);
      const int i1 = static_cast<int>(static_cast<long long>(rows) * (t + 1) / T);
      for (int j = 0; j < q; ++j) {
        double* gj = G.data() + static_cast<size_t>(j) * rows;
        for (int k = 0; k < j; ++k) {
          const double r = R[k + static_cast<size_t>(j) * q];
          const double* gk = G.data() + static_cast<size_t>(k) * rows;
          for (int i = i0; i < i1; ++i) gj[i] -= r * gk[i];
        }
        const double inv = 1.0 / R[j + static_cast<size_t>(j) * q];
        for (int i = i0; i < i1; ++i) gj[i] *= inv;
      }
    });
  }
  return true;
}

static achr::GemmArgs gemm(int M, int N, int K, const void* A, long sa_m, long sa_k, const float* B, long sb_k,
                           long sb_n, void* C, long sc_m, long sc_n, float alpha = 1.0f, float beta = 0.0f,
                           int epi = achr::kEpiStoreF32) {
  achr::GemmArgs g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = A;
  g.sa_m = sa_m;
  g.sa_k = sa_k;
  g.B = B;
  g.sb_k = sb_k;
  g.sb_n = sb_n;
  g.C = C;
  g.sc_m = sc_m;
  g.sc_n = sc_n;
  g.alpha = alpha;
  g.beta = beta;
  g.epi = epi;
  return g;
}

#define GEMM_TRY(g, prec, rows)                                                               \
  do {                                                                                        \
    hipError_t e_ = achr::launch_gemm((g), (prec), (rows), nullptr);                          \
    if (e_ != hipSuccess) return fail(ACHR_E_DEVICE, "gemm: %s", hipGetErrorString(e_));      \
  } while (0)

int achr_sampler_learn_directions(achr_sampler* S, const achr_learn_options* opts, double* report) {
  if (!S) return fail(ACHR_E_ARG, "sampler is NULL");
  achr_learn_options o;
  achr_learn_options_default(&o);
  if (opts) o = *opts;
  const achr_model* M = S->model;
  const int n = S->n;
  if (S->d <= 0) return fail(ACHR_E_STATE, "empty null space");
  const double t0 = now_ms();
  DevBuf<double> Nf;
  int df = 0;
  int rc = free_basis(M, S->Ncm.p, Nf, &df);
  if (rc != ACHR_OK) return rc;
  const int K = o.bank_size > 0 ? o.bank_size : 2 * n;
  if (S->sched.tpc < 64 && 8.0 * n * K >= 2147483648.0) return fail(ACHR_E_ARG, "bank too large for the schedule");
  const int r = std::min(o.rank, df);
  long Msamp = 0;
  double loss0 = 0.0, loss1 = 0.0, sig2 = 1.0, captured = 0.0, train_ms = 0.0;
  DevBuf<float> Q32;     // df x r column major (Q[i + j*df]), orthonormal principal axes
  DevBuf<float> Qs32;    // the same axes scaled by sqrt(lambda_j)
  DevBuf<float> Wmix;    // K x df row major
  HIP_TRY(Wmix.ensure(static_cast<size_t>(K) * df));
  const uint64_t seed = o.seed;
  if (o.isotropic || r <= 0) {
    HIP_TRY(achr::launch_gauss(Wmix.p, static_cast<size_t>(K) * df, seed, 0xB0u, nullptr));
  } else {
    if (!S->have_run || S->trace_chains_run <= 0 || S->trace_len <= 0)
      return fail(ACHR_E_STATE, "learning needs a traced run (trace_chains, trace_thin)");
    Msamp = static_cast<long>(S->trace_chains_run) * S->trace_len;
    HIP_TRY(hipStreamSynchronize(S->stream));
    // Samples in free null-space coordinates: Y = (X - mean) N_f, M x df FP32.
    DevBuf<float> Nf32, Y, mf;
    DevBuf<double> mean, mean_work;
    HIP_TRY(Nf32.ensure(static_cast<size_t>(n) * df));
    HIP_TRY(achr::launch_to_f32(Nf.p, Nf32.p, static_cast<size_t>(n) * df, nullptr));
    HIP_TRY(Y.ensure(static_cast<size_t>(Msamp) * df));
    HIP_TRY(mean.ensure(n));
    HIP_TRY(mf.ensure(df));
    HIP_TRY(mean_work.ensure(static_cast<size_t>(achr::kColMeanStripes) * n));
    HIP_TRY(achr::launch_col_means(S->trace.p, Msamp, n, mean.p, mean_work.p, nullptr));
    {
      achr::GemmArgs g = gemm(static_cast<int>(Msamp), df, n, S->trace.p, n, 1, Nf32.p, 1, n, Y.p, df, 1);
      g.a_f64 = 1;
      GEMM_TRY(g, achr::kPrecBF16x3, static_cast<int>(Msamp));
      achr::GemmArgs gm = gemm(1, df, n, mean.p, n, 1, Nf32.p, 1, n, mf.p, df, 1);
      gm.a_f64 = 1;
      GEMM_TRY(gm, achr::kPrecBF16x3, 1);
      HIP_TRY(achr::launch_sub_row(Y.p, Msamp, df, mf.p, nullptr));
    }
    // Randomised range finder: G = Y^T (Y Omega), Omega df x q Gaussian.
    const int q = std::min(df, r + 16);
    DevBuf<float> Om, P1, Gd;
    HIP_TRY(Om.ensure(static_cast<size_t>(df) * q));
    HIP_TRY(P1.ensure(static_cast<size_t>(Msamp) * q));
    HIP_TRY(Gd.ensure(static_cast<size_t>(df) * q));
    HIP_TRY(achr::launch_gauss(Om.p, static_cast<size_t>(df) * q, seed, 0xB1u, nullptr));
    GEMM_TRY(gemm(static_cast<int>(Msamp), q, df, Y.p, df, 1, Om.p, q, 1, P1.p, q, 1), achr::kPrecBF16x3,
             static_cast<int>(Msamp));
    // Gd (df x q, row major) = Y^T P1
    GEMM_TRY(gemm(df, q, static_cast<int>(Msamp), Y.p, 1, df, P1.p, q, 1, Gd.p, q, 1), achr::kPrecBF16x3, df);
    std::vector<float> Gh(static_cast<size_t>(df) * q);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(Gh.data(), Gd.p, sizeof(float) * Gh.size(), hipMemcpyDeviceToHost));
    std::vector<double> G(static_cast<size_t>(df) * q);  // column major
    for (int i = 0; i < df; ++i)
      for (int j = 0; j < q; ++j) G[static_cast<size_t>(i) + static_cast<size_t>(j) * df] = Gh[static_cast<size_t>(i) * q + j];
    if (!chol_qr(G, df, q)) return fail(ACHR_E_SOLVER, "range finder basis is rank deficient");
    std::vector<float> Qh(static_cast<size_t>(df) * q);
    for (size_t i = 0; i < Qh.size(); ++i) Qh[i] = static_cast<float>(G[i]);
    DevBuf<float> Qq;
    HIP_TRY(Qq.upload(Qh.data(), Qh.size(), nullptr));
    // Bq = Y Q (M x q), Cq = Bq^T Bq / M (q x q); keep the top-r eigen-directions of Cq.
    DevBuf<float> Bq, Cq;
    HIP_TRY(Bq.ensure(static_cast<size_t>(Msamp) * q));
    HIP_TRY(Cq.ensure(static_cast<size_t>(q) * q));
    GEMM_TRY(gemm(static_cast<int>(Msamp), q, df, Y.p, df, 1, Qq.p, 1, df, Bq.p, q, 1), achr::kPrecBF16x3,
             static_cast<int>(Msamp));
    GEMM_TRY(gemm(q, q, static_cast<int>(Msamp), Bq.p, 1, q, Bq.p, q, 1, Cq.p, q, 1,
                  1.0f / static_cast<float>(Msamp)),
             achr::kPrecBF16x3, q);
    // total variance |Y|_F^2 / M
    std::vector<float> Cqh(static_cast<size_t>(q) * q);
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(Cqh.data(), Cq.p, sizeof(float) * Cqh.size(), hipMemcpyDeviceToHost));
    {
      double tot = 0.0;
      HIP_TRY(achr::sum_squares(Y.p, static_cast<size_t>(Msamp) * df, &tot));
      tot /= static_cast<double>(Msamp);
      // symmetric eigen-decomposition of Cq (q x q) by rocSOLVER dsyevd on the
      // device (a cyclic Jacobi on the host took 0.9 s at q = 528); A holds
      // the eigenvalues on its diagonal afterwards, V the eigenvectors
      std::vector<double> A(static_cast<size_t>(q) * q), V(static_cast<size_t>(q) * q, 0.0);
      for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j)
          V[i + static_cast<size_t>(j) * q] =
              0.5 * (Cqh[static_cast<size_t>(i) * q + j] + Cqh[static_cast<size_t>(j) * q + i]);
      {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) return fail(ACHR_E_SOLVER, "rocblas_create_handle");
        DevBuf<double> dA, dD, dE;
        DevBuf<int> dinfo;
        int info = 0;
        std::vector<double> ev(static_cast<size_t>(q));
        const bool ok = dA.upload(V.data(), V.size(), nullptr) == hipSuccess && dD.ensure(q) == hipSuccess &&
                        dE.ensure(q) == hipSuccess && dinfo.ensure(1) == hipSuccess &&
                        rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, q, dA.p, q, dD.p, dE.p,
                                         dinfo.p) == rocblas_status_success &&
                        hipMemcpy(&info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess && info == 0 &&
                        hipMemcpy(V.data(), dA.p, sizeof(double) * V.size(), hipMemcpyDeviceToHost) == hipSuccess &&
                        hipMemcpy(ev.data(), dD.p, sizeof(double) * ev.size(), hipMemcpyDeviceToHost) == hipSuccess;
        rocblas_destroy_handle(h);
        if (!ok) return fail(ACHR_E_SOLVER, "principal axes eigenproblem");
        for (int i = 0; i < q; ++i) A[i + static_cast<size_t>(i) * q] = ev[static_cast<size_t>(i)];
      }
      std::vector<int> ord(q);
      for (int i = 0; i < q; ++i) ord[i] = i;
      std::sort(ord.begin(), ord.end(), [&](int a, int b) {
        return A[a + static_cast<size_t>(a) * q] > A[b + static_cast<size_t>(b) * q];
      });
      // Q_r = Q V[:, top r] (principal axes), lambda_j their variances.
      std::vector<float> Qr(static_cast<size_t>(df) * r, 0.0f);
      std::vector<double> lam(r);
      captured = 0.0;
      for (int j = 0; j < r; ++j) {
        lam[j] = std::max(0.0, A[ord[j] + static_cast<size_t>(ord[j]) * q]);
        captured += lam[j];
      }
      // Qr[:, j] = G V[:, ord[j]], each element summed over k in order; the
      // k loop outside the i loop reads G by columns (host threads by j)
      on_threads([&](int t, int T) {
        std::vector<double> acc(static_cast<size_t>(df));
        for (int j = t; j < r; j += T) {
          const double* vcol = V.data() + static_cast<size_t>(ord[j]) * q;
          std::fill(acc.begin(), acc.end(), 0.0);
          for (int k = 0; k < q; ++k) {
            const double* gk = G.data() + static_cast<size_t>(k) * df;
            const double vk = vcol[k];
            for (int i = 0; i < df; ++i) acc[static_cast<size_t>(i)] += gk[i] * vk;
          }
          float* qcol = Qr.data() + static_cast<size_t>(j) * df;
          for (int i = 0; i < df; ++i) qcol[i] = static_cast<float>(acc[static_cast<size_t>(i)]);
        }
      });
      sig2 = df > r ? std::max(tot - captured, 0.0) / (df - r) : 0.0;
      captured = tot > 0.0 ? captured / tot : 0.0;
      // The network emits whitened directions (target: identity second
      // moment); the rounding map scales principal axis j by sqrt(lambda_j).
      // Working in whitened axes keeps the fit relative per axis, so the
      // 1e5 : 1 variance ratios of internal cycles do not leak into the
      // narrow directions. Scales are normalised to unit mean variance.
      HIP_TRY(Q32.upload(Qr.data(), Qr.size(), nullptr));
      double mean_lam = 0.0;
      for (int j = 0; j < r; ++j) mean_lam += lam[j];
      mean_lam /= r;
      const double scale = mean_lam > 0.0 ? 1.0 / mean_lam : 1.0;
      std::vector<float> Qs(Qr.size());
      for (int j = 0; j < r; ++j) {
        const float sj = static_cast<float>(std::sqrt(lam[j] * scale));
        for (int i = 0; i < df; ++i) Qs[static_cast<size_t>(i) + static_cast<size_t>(j) * df] = Qr[static_cast<size_t>(i) + static_cast<size_t>(j) * df] * sj;
      }
      HIP_TRY(Qs32.upload(Qs.data(), Qs.size(), nullptr));
      std::vector<float> Cr(static_cast<size_t>(r) * r, 0.0f);
      for (int j = 0; j < r; ++j) Cr[static_cast<size_t>(j) * r + j] = 1.0f;
      sig2 *= scale;
      // Train the network.
      achr_dirnet* net = nullptr;
      rc = achr_dirnet_create(r, o.width, r, seed, &net);
      if (rc != ACHR_OK) return rc;
      rc = achr_dirnet_set_target(net, Cr.data());
      const double tt = now_ms();
      std::vector<float> losses(std::max(o.iters, 1));
      if (rc == ACHR_OK) rc = achr_dirnet_train(net, o.iters, o.batch, o.lr, seed, losses.data());
      train_ms = now_ms() - tt;
      if (rc == ACHR_OK && o.iters > 0) {
        loss0 = losses.front();
        loss1 = losses[o.iters - 1];
      }
      // Bank coordinates: Wmix (K x df) = sigma (G - (G Q) Q^T) + V Q^T, V = v(z) (BF16 inference).
      if (rc == ACHR_OK) {
        achr::DirNet& nn = net->net;
        DevBuf<float> Vb, GQ;
        HIP_TRY(Vb.ensure(static_cast<size_t>(K) * r));
        HIP_TRY(GQ.ensure(static_cast<size_t>(K) * r));
        const int chunk = 4096;
        for (int k0 = 0; k0 < K; k0 += chunk) {
          const int cnt = std::min(chunk, K - k0);
          HIP_TRY(nn.ensure_batch(cnt));
          HIP_TRY(nn.gaussian_latents(cnt, seed ^ 0x5bd1e995u, static_cast<uint32_t>(0x10000 + k0), nullptr));
          HIP_TRY(nn.forward(cnt, achr::kPrecBF16, nullptr));
          HIP_TRY(hipMemcpyAsync(Vb.p + static_cast<size_t>(k0) * r, nn.v, sizeof(float) * cnt * r,
                                 hipMemcpyDeviceToDevice, nullptr));
        }
        const float sig = static_cast<float>(std::sqrt(sig2));
        HIP_TRY(achr::launch_gauss(Wmix.p, static_cast<size_t>(K) * df, seed, 0xB2u, nullptr));
        // GQ = G Q (K x r)
        GEMM_TRY(gemm(K, r, df, Wmix.p, df, 1, Q32.p, 1, df, GQ.p, r, 1), achr::kPrecBF16x3, K);
        // Wmix <- sigma (Wmix - GQ Q^T) + V Qs^T
        GEMM_TRY(gemm(K, df, r, GQ.p, r, 1, Q32.p, df, 1, Wmix.p, df, 1, -sig, sig), achr::kPrecBF16x3, K);
        GEMM_TRY(gemm(K, df, r, Vb.p, r, 1, Qs32.p, df, 1, Wmix.p, df, 1, 1.0f, 1.0f), achr::kPrecBF16x3, K);
      }
      achr_dirnet_destroy(net);
      if (rc != ACHR_OK) return rc;
    }
  }
  // Bank (n x K column major, FP64) = N_f Wmix^T, then unit columns.
  DevBuf<double> W64;
  HIP_TRY(W64.ensure(static_cast<size_t>(K) * df));
  HIP_TRY(achr::launch_to_f64(Wmix.p, W64.p, static_cast<size_t>(K) * df, nullptr));
  HIP_TRY(S->dirs.ensure(static_cast<size_t>(n) * K));
  DevBuf<int> kcount;
  HIP_TRY(kcount.upload(&K, 1, nullptr));
  HIP_TRY(achr::launch_dgemm_cols(n, df, Nf.p, 1, n, W64.p, df, S->dirs.p, n, kcount.p, K, nullptr));
  HIP_TRY(achr::launch_normalize_cols(S->dirs.p, n, K, nullptr));
  HIP_TRY(hipDeviceSynchronize());
  S->K = K;
  const int rc_start = achr_sampler_use_directions(S, 1);
  if (rc_start != ACHR_OK) return rc_start;
  if (report) {
    report[0] = df;
    report[1] = static_cast<double>(Msamp);
    report[2] = loss0;
    report[3] = loss1;
    report[4] = sig2;
    report[5] = captured;
    report[6] = now_ms() - t0;
    report[7] = train_ms;
  }
  return ACHR_OK;
}

}  // extern "C"
