// C ABI of the ACHR sampler: models, null space, warmup centre, sampling runs,
// file formats. Host code; the device work is in achr_step.hip and
// achr_project.hip.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <climits>
#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/achr.h"
#include "achr_gemm.h"
#include "achr_host.h"
#include "achr_internal.h"
#include "achr_philox.h"

using achr_host::DevBuf;
using achr_host::fail;

extern "C" {

void achr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  const achr::u32x4 r = achr::philox4x32_10({ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
  out[0] = r.v0;
  out[1] = r.v1;
  out[2] = r.v2;
  out[3] = r.v3;
}

void achr_options_default(achr_options* o) {
  if (!o) return;
  o->dTol = 1e-14;
  o->uTol = 1e-9;
  o->maxMinTol = 1e-9;
  o->reprojTol = 1e-9;
  o->check_every = 10;
  o->max_misses = 100000;
  o->lanes = 0;
  o->reproject_mode = 0;
  o->trace_chains = 0;
  o->trace_thin = 0;
  o->proj_prec = ACHR_PROJ_FP64;
  o->fp64_every = 10;
  o->model_id = 0;
  o->check_kernel = 0;
}


static int compute_nullspace(achr_model* M, int* null_dim) {
  if (!M) return fail(ACHR_E_ARG, "model is NULL");
  const int n = M->n;
  // S^T S from the CSR rows on the host: (S^T S)_{ab} = sum_r S_ra S_rb.
  std::vector<double> A(static_cast<size_t>(n) * n, 0.0);
  for (int r = 0; r < M->m; ++r) {
    for (int p = M->rowptr[r]; p < M->rowptr[r + 1]; ++p)
      for (int q = M->rowptr[r]; q < M->rowptr[r + 1]; ++q)
        A[static_cast<size_t>(M->colidx[p]) + static_cast<size_t>(M->colidx[q]) * n] += M->vals[p] * M->vals[q];
  }
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return fail(ACHR_E_SOLVER, "rocblas_create_handle");
  double *dA = nullptr, *dD = nullptr, *dE = nullptr;
  rocblas_int* dinfo = nullptr;
  int rc = ACHR_OK;
  std::vector<double> ev(n);
  rocblas_int info = 0;
  do {
    if (hipMalloc(&dA, sizeof(double) * A.size()) != hipSuccess || hipMalloc(&dD, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&dE, sizeof(double) * n) != hipSuccess || hipMalloc(&dinfo, sizeof(rocblas_int)) != hipSuccess) {
      rc = fail(ACHR_E_NOMEM, "device allocation for the %d x %d eigenproblem", n, n);
      break;
    }
    if (hipMemcpy(dA, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice) != hipSuccess) {
      rc = fail(ACHR_E_DEVICE, "upload S^T S");
      break;
    }
    rocblas_status s = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, dA, n, dD, dE, dinfo);
    if (s != rocblas_status_success) {
      rc = fail(ACHR_E_SOLVER, "rocsolver_dsyevd status %d", static_cast<int>(s));
      break;
    }
    if (hipMemcpy(&info, dinfo, sizeof info, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ev.data(), dD, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(A.data(), dA, sizeof(double) * A.size(), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(ACHR_E_DEVICE, "download eigenpairs");
      break;
    }
    if (info != 0) {
      rc = fail(ACHR_E_SOLVER, "rocsolver_dsyevd info %d", static_cast<int>(info));
      break;
    }
  } while (false);
  (void)hipFree(dA);
  (void)hipFree(dD);
  (void)hipFree(dE);
  (void)hipFree(dinfo);
  rocblas_destroy_handle(h);
  if (rc != ACHR_OK) return rc;
  // Eigenvalues ascend; the largest singular value of S^T S is max |lambda|.
  double smax = 0.0;
  for (double v : ev) smax = std::max(smax, std::fabs(v));
  const double tol = static_cast<double>(n) * smax * 2.2204e-16;  // ACHRCuda.cu:41, :474
  std::vector<int> keep;
  for (int i = 0; i < n; ++i)
    if (std::fabs(ev[i]) < tol) keep.push_back(i);
  // Order the basis like the reference (descending singular value): the
  // largest of the small eigenvalues first.
  std::sort(keep.begin(), keep.end(), [&](int a, int b) { return std::fabs(ev[a]) > std::fabs(ev[b]); });
  const int d = static_cast<int>(keep.size());
  M->N.assign(static_cast<size_t>(n) * d, 0.0);
  for (int j = 0; j < d; ++j)
    std::copy(A.begin() + static_cast<size_t>(keep[j]) * n, A.begin() + static_cast<size_t>(keep[j] + 1) * n,
              M->N.begin() + static_cast<size_t>(j) * n);
  M->d = d;
  if (null_dim) *null_dim = d;
  return ACHR_OK;
}

int achr_model_compute_nullspace(achr_model* M, int* null_dim) {
  return achr_host::no_throw([&] { return compute_nullspace(M, null_dim); });
}

// ------------------------------------------------------ pseudo-inverse S+

// build_pinv_f32 stores S+ as FP32, n x m column major: by pinv_chol when it
// succeeds, else by the eigen route, S+ = S^T (S S^T)+ with the
// eigen-decomposition of S S^T (rocSOLVER syevd, eigenvalues below
// m * lambda_max * 2.2204e-16 dropped, the rule the reference applies to
// S^T S), (S S^T)+ by one rocBLAS DGEMM, then S^T times it straight from the
// sparse columns of S (launch_pinv_csc; S is never made dense on this route).
// x - S+ (S x) is the orthogonal projection of x onto null(S), the same
// operator as N N^T x.
// With the null-space basis N (n x d, on the device as dN) S+ needs no
// eigen-decomposition: G = S^T S + sigma N N^T is symmetric positive definite
// when N spans null(S) (eigenvalues: those of S^T S on the row space of S,
// sigma on null(S); sigma, the largest diagonal entry of S^T S, keeps both
// ranges alike), and since N N^T S^T = 0, G^-1 S^T = S+. One Cholesky
// factorisation and two triangular solves against the dense S^T (built on the
// device only, then overwritten by S+). Returns false, and the caller takes
// the eigen route, when the factorisation fails, or when for r = S z (z a
// fixed pseudo-random vector) S+ r, which is z's part in the row space of S,
// misses z - N (N^T z) by more than 1e-7 of max |z|: N must be an orthonormal
// basis of all of null(S) (a direction missing from N leaves G singular
// there, and its roundoff-sized pivot puts garbage along it).
static bool pinv_chol(const achr_model* M, const double* dN, DevBuf<float>& out) {
  const int n = M->n, m = M->m, d = M->d;
  double sigma = 0.0;
  {
    std::vector<double> cn(n, 0.0);
    for (size_t p = 0; p < M->vals.size(); ++p) cn[M->colidx[p]] += M->vals[p] * M->vals[p];
    for (double v : cn) sigma = std::max(sigma, v);
  }
  if (!(sigma > 0.0) || (d > 0 && !dN)) return false;
  // check vector: r = S z, t = S^T r
  std::vector<double> z(n), r(m, 0.0), t(n, 0.0), y(n);
  uint64_t st = 0x9E3779B97F4A7C15ull;
  for (int i = 0; i < n; ++i) {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    z[i] = static_cast<double>(st >> 11) * 0x1.0p-52 - 1.0;
  }
  double rmax = 0.0;
  for (int row = 0; row < m; ++row) {
    double a = 0.0;
    for (int p = M->rowptr[row]; p < M->rowptr[row + 1]; ++p) a += M->vals[p] * z[M->colidx[p]];
    r[row] = a;
    rmax = std::max(rmax, std::fabs(a));
  }
  for (int row = 0; row < m; ++row)
    for (int p = M->rowptr[row]; p < M->rowptr[row + 1]; ++p) t[M->colidx[p]] += M->vals[p] * r[row];
  if (!(rmax > 0.0)) return false;
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return false;
  DevBuf<double> dB, dG, dt, dv, dz, dw;
  DevBuf<int> drp, dci, dinfo;
  bool ok = false;
  do {
    const size_t nm = static_cast<size_t>(n) * m, nn = static_cast<size_t>(n) * n;
    if (dB.ensure(nm) != hipSuccess || dG.ensure(nn) != hipSuccess || dinfo.ensure(1) != hipSuccess ||
        drp.upload(M->rowptr.data(), M->rowptr.size(), nullptr) != hipSuccess ||
        dci.upload(M->colidx.data(), std::max<size_t>(M->colidx.size(), 1), nullptr) != hipSuccess ||
        dv.upload(M->vals.data(), std::max<size_t>(M->vals.size(), 1), nullptr) != hipSuccess ||
        dt.upload(t.data(), t.size(), nullptr) != hipSuccess || dz.upload(z.data(), z.size(), nullptr) != hipSuccess ||
        dw.ensure(std::max(d, 1)) != hipSuccess)
      break;
    if (hipMemset(dB.p, 0, sizeof(double) * nm) != hipSuccess ||
        achr::launch_dense_st(m, n, drp.p, dci.p, dv.p, dB.p, nullptr) != hipSuccess)
      break;
    const double one = 1.0, zero = 0.0;
    rocblas_int info = 1;
    if (rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_none, n, m, &one, dB.p, n, &zero, dG.p, n) !=
            rocblas_status_success ||
        (d > 0 && rocblas_dsyrk(h, rocblas_fill_upper, rocblas_operation_none, n, d, &sigma, dN, n, &one, dG.p, n) !=
                      rocblas_status_success) ||
        rocsolver_dpotrf(h, rocblas_fill_upper, n, dG.p, n, dinfo.p) != rocblas_status_success ||
        hipMemcpy(&info, dinfo.p, sizeof info, hipMemcpyDeviceToHost) != hipSuccess || info != 0)
      break;
    // y = G^-1 S^T r = S+ r; S y must give back r
    if (rocsolver_dpotrs(h, rocblas_fill_upper, n, 1, dG.p, n, dt.p, n) != rocblas_status_success ||
        hipMemcpy(y.data(), dt.p, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
      break;
    double err = 0.0;
    for (int row = 0; row < m; ++row) {
      double a = 0.0;
      for (int p = M->rowptr[row]; p < M->rowptr[row + 1]; ++p) a += M->vals[p] * y[M->colidx[p]];
      err = std::max(err, std::fabs(a - r[row]));
    }
    // dz <- z - N (N^T z), the row-space part of z
    const double minus_one = -1.0;
    if (d > 0 && (rocblas_dgemv(h, rocblas_operation_transpose, n, d, &one, dN, n, dz.p, 1, &zero, dw.p, 1) !=
                      rocblas_status_success ||
                  rocblas_dgemv(h, rocblas_operation_none, n, d, &minus_one, dN, n, dw.p, 1, &one, dz.p, 1) !=
                      rocblas_status_success))
      break;
    std::vector<double> zr(n);
    if (hipMemcpy(zr.data(), dz.p, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) break;
    double zmax = 0.0, dev = 0.0;
    for (int i = 0; i < n; ++i) {
      zmax = std::max(zmax, std::fabs(z[i]));
      dev = std::max(dev, std::fabs(y[i] - zr[i]));
    }
    if (!(err <= 1e-6 * rmax) || !(dev <= 1e-7 * zmax)) break;
    if (rocsolver_dpotrs(h, rocblas_fill_upper, n, m, dG.p, n, dB.p, n) != rocblas_status_success) break;
    if (out.ensure(nm) != hipSuccess || achr::launch_to_f32(dB.p, out.p, nm, nullptr) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
      break;
    ok = true;
  } while (false);
  rocblas_destroy_handle(h);
  if (!ok) (void)hipGetLastError();
  return ok;
}

static int build_pinv_f32(const achr_model* M, const double* dN, DevBuf<float>& out) {
  const int n = M->n, m = M->m;
  if (m == 0) {
    HIP_TRY(out.ensure(1));
    return ACHR_OK;
  }
  if (pinv_chol(M, dN, out)) return ACHR_OK;
  // Columns of S (CSC): rows ascending within a column.
  std::vector<int> colptr(static_cast<size_t>(n) + 1, 0);
  for (int r = 0; r < m; ++r)
    for (int p = M->rowptr[r]; p < M->rowptr[r + 1]; ++p) ++colptr[M->colidx[p] + 1];
  for (int i = 0; i < n; ++i) colptr[i + 1] += colptr[i];
  std::vector<int> rowidx(M->vals.size() > 0 ? M->vals.size() : 1), fillp(colptr.begin(), colptr.end() - 1);
  std::vector<double> cvals(rowidx.size(), 0.0);
  for (int r = 0; r < m; ++r)
    for (int p = M->rowptr[r]; p < M->rowptr[r + 1]; ++p) {
      const int q = fillp[M->colidx[p]]++;
      rowidx[q] = r;
      cvals[q] = M->vals[p];
    }
  // S S^T on the host from the columns of S.
  std::vector<double> G(static_cast<size_t>(m) * m, 0.0);
  for (int j = 0; j < n; ++j)
    for (int a = colptr[j]; a < colptr[j + 1]; ++a)
      for (int b = colptr[j]; b < colptr[j + 1]; ++b)
        G[static_cast<size_t>(rowidx[a]) + static_cast<size_t>(rowidx[b]) * m] += cvals[a] * cvals[b];
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) return fail(ACHR_E_SOLVER, "rocblas_create_handle");
  DevBuf<double> dG, dD, dE, dW, dGinv, dv;
  DevBuf<int> dinfo, dcp, dri;
  int rc = ACHR_OK;
  std::vector<double> ev(m);
  do {
    const size_t mm = static_cast<size_t>(m) * m, mn = static_cast<size_t>(m) * n;
    if (dG.upload(G.data(), mm, nullptr) != hipSuccess || dD.ensure(m) != hipSuccess || dE.ensure(m) != hipSuccess ||
        dinfo.ensure(1) != hipSuccess || dW.ensure(mm) != hipSuccess || dGinv.ensure(mm) != hipSuccess ||
        dcp.upload(colptr.data(), colptr.size(), nullptr) != hipSuccess ||
        dri.upload(rowidx.data(), rowidx.size(), nullptr) != hipSuccess ||
        dv.upload(cvals.data(), cvals.size(), nullptr) != hipSuccess) {
      rc = fail(ACHR_E_NOMEM, "device buffers for the pseudo-inverse");
      break;
    }
    if (rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, m, dG.p, m, dD.p, dE.p, dinfo.p) !=
        rocblas_status_success) {
      rc = fail(ACHR_E_SOLVER, "rocsolver_dsyevd (S S^T)");
      break;
    }
    rocblas_int info = 0;
    if (hipMemcpy(&info, dinfo.p, sizeof info, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ev.data(), dD.p, 8 * static_cast<size_t>(m), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(G.data(), dG.p, 8 * mm, hipMemcpyDeviceToHost) != hipSuccess || info != 0) {
      rc = fail(ACHR_E_SOLVER, "syevd (S S^T) info %d", static_cast<int>(info));
      break;
    }
    double lmax = 0.0;
    for (double v : ev) lmax = std::max(lmax, std::fabs(v));
    const double tol = static_cast<double>(m) * lmax * 2.2204e-16;
    // U diag(1/lambda) for kept eigenpairs, zero columns otherwise
    std::vector<double> Us(G.size());
    for (int k = 0; k < m; ++k) {
      const double sc = ev[k] >= tol ? 1.0 / ev[k] : 0.0;
      for (int i = 0; i < m; ++i) Us[static_cast<size_t>(i) + static_cast<size_t>(k) * m] = G[static_cast<size_t>(i) + static_cast<size_t>(k) * m] * sc;
    }
    if (hipMemcpy(dW.p, Us.data(), 8 * mm, hipMemcpyHostToDevice) != hipSuccess) {
      rc = fail(ACHR_E_DEVICE, "upload U/lambda");
      break;
    }
    const double one = 1.0, zero = 0.0;
    // (S S^T)+ = (U / lambda) U^T; dG holds U
    if (rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, m, m, m, &one, dW.p, m, dG.p, m, &zero,
                      dGinv.p, m) != rocblas_status_success) {
      rc = fail(ACHR_E_SOLVER, "rocblas_dgemm for (S S^T)+");
      break;
    }
    hipError_t e = out.ensure(mn);
    if (e == hipSuccess) e = achr::launch_pinv_csc(n, m, dcp.p, dri.p, dv.p, dGinv.p, out.p, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = fail(ACHR_E_DEVICE, "S+ from the columns of S: %s", hipGetErrorString(e));
  } while (false);
  rocblas_destroy_handle(h);
  return rc;
}

// ---------------------------------------------------------------- sampler

static achr::Schedule pick_schedule(int n, int lanes_req) {
  achr::Schedule s{};
  static const int kTpc[] = {16, 32, 64, 128, 256, 512, 1024};
  if (lanes_req > 0) {
    s.tpc = lanes_req;
    s.e = achr::elems_for(n, lanes_req);
  } else {
    s.tpc = 1024;
    s.e = achr::elems_for(n, 1024);
    for (int t : kTpc) {
      if ((n + t - 1) / t <= 4) {
        s.tpc = t;
        s.e = achr::elems_for(n, t);
        break;
      }
    }
  }
  s.cpb = s.tpc <= 64 ? 256 / s.tpc : 1;
  return s;
}

void achr_sampler_destroy(achr_sampler* S) {
  if (!S) return;
  for (hipEvent_t e : S->ev_pool) (void)hipEventDestroy(e);
  if (S->ev_begin) (void)hipEventDestroy(S->ev_begin);
  if (S->ev_end) (void)hipEventDestroy(S->ev_end);
  delete S;
}

static int sampler_create(const achr_model* M, const double* warmup, int W, const achr_options* opts,
                          achr_sampler** out) {
  if (!M || !warmup || !out) return fail(ACHR_E_ARG, "NULL argument");
  *out = nullptr;
  if (W <= 0) return fail(ACHR_E_ARG, "no warmup points");
  if (M->d < 0) return fail(ACHR_E_STATE, "model has no null space: call achr_model_compute_nullspace or set one");
  achr_options o;
  achr_options_default(&o);
  if (opts) o = *opts;
  if (o.check_every <= 0 || o.max_misses < 0) return fail(ACHR_E_ARG, "bad check_every / max_misses");
  if (o.check_kernel < 0 || o.check_kernel > 3) return fail(ACHR_E_ARG, "bad check_kernel=%d", o.check_kernel);
  if (o.lanes != 0 && !achr::schedule_supported(o.lanes, achr::elems_for(M->n, o.lanes)))
    return fail(ACHR_E_ARG, "lanes=%d cannot hold n=%d reactions", o.lanes, M->n);
  if (o.trace_chains > 0 && o.trace_thin <= 0) return fail(ACHR_E_ARG, "trace_thin must be > 0");
  if (o.proj_prec < ACHR_PROJ_FP64 || o.proj_prec > ACHR_PROJ_NULL_BF16X3 || o.fp64_every < 1)
    return fail(ACHR_E_ARG, "bad proj_prec %d / fp64_every %d", o.proj_prec, o.fp64_every);
  if (o.model_id < 0 || o.model_id >= (1 << 27)) return fail(ACHR_E_ARG, "model_id %d outside 0..2^27-1", o.model_id);
  if (o.proj_prec != ACHR_PROJ_FP64 && o.reproject_mode == 1)
    return fail(ACHR_E_ARG, "reduced-precision projection runs in batched rounds (reproject_mode 2)");
  // owned here until it is returned: every early exit, an allocation failure
  // included, releases what the sampler holds
  std::unique_ptr<achr_sampler, void (*)(achr_sampler*)> S(new (std::nothrow) achr_sampler(), achr_sampler_destroy);
  if (!S) return fail(ACHR_E_NOMEM, "sampler allocation");
  S->model = M;
  S->n = M->n;
  S->m = M->m;
  S->W = W;
  S->d = M->d;
  S->opt = o;
  S->sched = pick_schedule(M->n, o.lanes);
  // Groups smaller than a wavefront address the warmup set with 32-bit byte
  // offsets from its base.
  if (S->sched.tpc < 64 && 8.0 * M->n * W >= 2147483648.0) {
    if (o.lanes != 0) return fail(ACHR_E_ARG, "lanes=%d needs n*W*8 < 2 GiB", o.lanes);
    S->sched = pick_schedule(M->n, 64);
  }
  if (S->sched.e < 0) return fail(ACHR_E_ARG, "n=%d exceeds the largest register schedule (1024 x 16)", M->n);
  const double nd = static_cast<double>(M->n) * static_cast<double>(M->d);
  S->mode = o.reproject_mode != 0 ? o.reproject_mode : (nd <= 65536.0 && o.proj_prec == ACHR_PROJ_FP64 ? 1 : 2);
  // Warmup centre on the host: (1/W) * sum_k w_k, k ascending (createCenterPt).
  // Host threads take ranges of coordinates and add the points one after the
  // other (contiguous reads), so each coordinate is still summed from +0.0 in
  // k order.
  const int n = M->n;
  S->c0.assign(n, 0.0);
  const double alpha = 1.0 / static_cast<double>(W);
  achr_host::on_threads([&](int t, int T) {
    const int i0 = static_cast<int>(static_cast<long>(n) * t / T);
    const int i1 = static_cast<int>(static_cast<long>(n) * (t + 1) / T);
    double* c = S->c0.data();
    for (int k = 0; k < W; ++k) {
      const double* wk = warmup + static_cast<size_t>(k) * n;
      for (int i = i0; i < i1; ++i) c[i] += wk[i];
    }
    for (int i = i0; i < i1; ++i) c[i] = alpha * c[i];
  });
  // N row major (n x d), transposed in 64 x 64 tiles; threads take 64-row bands.
  const int d = M->d;
  std::vector<double> Nrm(static_cast<size_t>(n) * std::max(d, 0));
  achr_host::on_threads([&](int t, int T) {
    for (int ib = 64 * t; ib < n; ib += 64 * T) {
      const int ie = std::min(n, ib + 64);
      for (int jb = 0; jb < d; jb += 64) {
        const int je = std::min(d, jb + 64);
        for (int j = jb; j < je; ++j) {
          const double* src = M->N.data() + static_cast<size_t>(j) * n;
          for (int i = ib; i < ie; ++i) Nrm[static_cast<size_t>(i) * d + j] = src[i];
        }
      }
    }
  });
  // S as a sliced ELL table for the residual check of chains of 64 or more
  // lanes: rows sorted by length (longest first), 64 rows to a slice, each
  // slice as wide as its longest row, entries of a row in CSR order.
  const int m = M->m;
  std::vector<int> order(m);
  for (int r = 0; r < m; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return M->rowptr[x + 1] - M->rowptr[x] > M->rowptr[y + 1] - M->rowptr[y];
  });
  const int nslices = (m + 63) / 64;
  std::vector<int> eptr(static_cast<size_t>(nslices) + 1, 0), erow(static_cast<size_t>(std::max(nslices, 1)) * 64, -1);
  for (int sl = 0; sl < nslices; ++sl) {
    const int r0 = order[static_cast<size_t>(sl) * 64];
    eptr[sl + 1] = eptr[sl] + 64 * (M->rowptr[r0 + 1] - M->rowptr[r0]);
  }
  std::vector<int> ecol(std::max(eptr[nslices], 1), -1);
  std::vector<double> eval(ecol.size(), 0.0);
  for (int sl = 0; sl < nslices; ++sl)
    for (int l = 0; l < 64 && sl * 64 + l < m; ++l) {
      const int r = order[static_cast<size_t>(sl) * 64 + l];
      erow[static_cast<size_t>(sl) * 64 + l] = r;
      for (int q = M->rowptr[r]; q < M->rowptr[r + 1]; ++q) {
        const size_t at = static_cast<size_t>(eptr[sl]) + 64 * static_cast<size_t>(q - M->rowptr[r]) + l;
        ecol[at] = M->colidx[q];
        eval[at] = M->vals[q];
      }
    }
  S->nslices = nslices;
  hipStream_t s0 = nullptr;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = S->ell_ptr.upload(eptr.data(), eptr.size(), s0);
  if (e == hipSuccess) e = S->ell_row.upload(erow.data(), erow.size(), s0);
  if (e == hipSuccess) e = S->ell_col.upload(ecol.data(), ecol.size(), s0);
  if (e == hipSuccess) e = S->ell_val.upload(eval.data(), eval.size(), s0);
  {
    achr_host::CheckPlan plan;
    achr_host::build_check_plan(n, nslices, eptr, erow, ecol, eval, &plan);
    S->chk_ok = plan.ok;
    if (plan.ok) {
      if (e == hipSuccess) e = S->chk_ent.upload(plan.ent.data(), plan.ent.size(), s0);
      if (e == hipSuccess) e = S->chk_row.upload(plan.rowmap.data(), plan.rowmap.size(), s0);
      if (e == hipSuccess) e = S->chk_vtab.upload(plan.vtab.data(), plan.vtab.size(), s0);
      S->chk_nv = static_cast<int>(plan.vtab.size());
      for (int w = 0; w < achr::kCheckWaves; ++w) {
        S->chk_wlen[w] = plan.wlen[w];
        S->chk_woff[w] = plan.woff[w];
        S->chk_roff[w] = plan.roff[w];
        S->chk_wends[w] = plan.wends[w];
      }
    }
  }
  if (e == hipSuccess) e = S->warm.upload(warmup, static_cast<size_t>(n) * W, s0);
  if (e == hipSuccess) e = S->lb.upload(M->lb.data(), n, s0);
  if (e == hipSuccess) e = S->ub.upload(M->ub.data(), n, s0);
  if (e == hipSuccess) e = S->c0d.upload(S->c0.data(), n, s0);
  if (e == hipSuccess) e = S->rowptr.upload(M->rowptr.data(), M->rowptr.size(), s0);
  if (e == hipSuccess) e = S->colidx.upload(M->colidx.data(), M->colidx.size(), s0);
  if (e == hipSuccess) e = S->vals.upload(M->vals.data(), M->vals.size(), s0);
  if (e == hipSuccess) e = S->Ncm.upload(M->N.data(), M->N.size(), s0);
  if (e == hipSuccess) e = S->Nrm.upload(Nrm.data(), Nrm.size(), s0);
  if (e == hipSuccess && (o.proj_prec == ACHR_PROJ_NULL_F32 || o.proj_prec == ACHR_PROJ_NULL_BF16X3)) {
    e = S->N32.ensure(std::max<size_t>(M->N.size(), 1));
    if (e == hipSuccess) e = achr::launch_to_f32(S->Ncm.p, S->N32.p, M->N.size(), nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
  }
  if (e == hipSuccess) e = hipEventCreate(&S->ev_begin);
  if (e == hipSuccess) e = hipEventCreate(&S->ev_end);
  if (e != hipSuccess) return fail(ACHR_E_DEVICE, "sampler upload: %s", hipGetErrorString(e));
  if (o.proj_prec == ACHR_PROJ_RES_BF16X3 || o.proj_prec == ACHR_PROJ_RES_F32) {
    DevBuf<float> Sp32;  // S+, n x m column major (i contiguous)
    const int rc = build_pinv_f32(M, S->Ncm.p, Sp32);
    if (rc != ACHR_OK) return rc;
    // B(k, i) = S+(i, k) for the rounds' X -= R S+^T, packed once
    const int gp = o.proj_prec == ACHR_PROJ_RES_F32 ? achr::kPrecF32 : achr::kPrecBF16x3;
    S->r32_ld = achr::packed_k(std::max(M->m, 1));
    e = S->Bpk.ensure(std::max<size_t>(achr::packed_b_bytes(M->m, n), 1));
    if (e == hipSuccess) e = achr::launch_pack_b(Sp32.p, n, 1, M->m, n, gp, S->Bpk.p, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(ACHR_E_DEVICE, "packing S+: %s", hipGetErrorString(e));
  }
  *out = S.release();
  return ACHR_OK;
}

int achr_sampler_create(const achr_model* M, const double* warmup, int W, const achr_options* opts,
                        achr_sampler** out) {
  return achr_host::no_throw([&] { return sampler_create(M, warmup, W, opts, out); });
}

static int sampler_create_csv(const achr_model* M, const char* path, const achr_options* opts,
                              achr_sampler** out) {
  if (!M) return fail(ACHR_E_ARG, "model is NULL");
  double* w = nullptr;
  int W = 0, n = 0;
  // A model with auxiliary columns also takes a file of its file columns
  // only (what the reference's generator writes); the rest are completed.
  int rc = achr_host::read_warmup_csv(path, M->n_file < M->n ? 0 : M->n, &w, &W, &n);
  if (rc != ACHR_OK) return rc;
  if (n != M->n && n != M->n_file) {
    std::free(w);
    return fail(ACHR_E_FORMAT, "%s: %d lines but the model has %d columns (%d in the file)", path, n, M->n, M->n_file);
  }
  if (n != M->n) {
    std::vector<double> full;
    try {
      full.resize(static_cast<size_t>(M->n) * W);
    } catch (const std::bad_alloc&) {
      std::free(w);
      return fail(ACHR_E_NOMEM, "warmup allocation");
    }
    achr_model_complete_points(M, w, W, full.data());
    rc = achr_sampler_create(M, full.data(), W, opts, out);
  } else {
    rc = achr_sampler_create(M, w, W, opts, out);
  }
  std::free(w);
  return rc;
}

int achr_sampler_create_csv(const achr_model* M, const char* path, const achr_options* opts,
                            achr_sampler** out) {
  return achr_host::no_throw([&] { return sampler_create_csv(M, path, opts, out); });
}

int achr_sampler_set_stream(achr_sampler* S, void* stream) {
  if (!S) return fail(ACHR_E_ARG, "sampler is NULL");
  S->stream = static_cast<hipStream_t>(stream);
  return ACHR_OK;
}

int achr_sampler_center(const achr_sampler* S, double* c0) {
  if (!S || !c0) return fail(ACHR_E_ARG, "NULL argument");
  std::copy(S->c0.begin(), S->c0.end(), c0);
  return ACHR_OK;
}

int achr_sampler_start_point(const achr_sampler* S, double* x0) {
  if (!S || !x0) return fail(ACHR_E_ARG, "NULL argument");
  const std::vector<double>& v = S->use_dirs ? S->s0 : S->c0;
  std::copy(v.begin(), v.end(), x0);
  return ACHR_OK;
}

int achr_sampler_schedule(const achr_sampler* S, int* lanes, int* mode) {
  if (!S) return fail(ACHR_E_ARG, "sampler is NULL");
  if (lanes) *lanes = S->sched.tpc;
  if (mode) *mode = S->mode;
  return ACHR_OK;
}

static int event_at(achr_sampler* S, size_t i, hipEvent_t* ev) {
  while (S->ev_pool.size() <= i) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    S->ev_pool.push_back(e);
  }
  *ev = S->ev_pool[i];
  return ACHR_OK;
}

}  // extern "C"

namespace achr_host {

int run_limits(const achr_sampler* S, const achr::Schedule& sc, int P, int Ssteps) {
  if (S->sched.e < 0 || sc.e < 0) return fail(ACHR_E_ARG, "no register schedule for n=%d", S->n);
  if (sc.tpc < 64) {
    // Groups smaller than a wavefront address per-chain rows and the warmup
    // set with 32-bit byte offsets from the array base.
    const double lim = 2147483648.0;
    const double trace_rows = static_cast<double>(std::min(S->opt.trace_chains, P)) *
                              (S->opt.trace_thin > 0 ? Ssteps / S->opt.trace_thin : 0);
    const int W = S->use_dirs ? S->K : S->W;
    if (8.0 * P * S->n >= lim || 8.0 * P * std::max(S->d, 1) >= lim || 8.0 * trace_rows * S->n >= lim ||
        8.0 * S->n * W >= lim)
      return fail(ACHR_E_ARG, "P=%d too large for a %d-lane schedule (32-bit row offsets)", P, sc.tpc);
  }
  return ACHR_OK;
}

int segments_for(const achr_sampler* S, int Ssteps) {
  const int checks = (Ssteps + S->opt.check_every - 1) / S->opt.check_every;
  return (S->mode == 2) ? checks + 1 : 1;
}

int prepare_run(achr_sampler* S, int P, int Ssteps, uint64_t seed, uint32_t file, int chain0, hipStream_t st,
                achr::StepArgs* args) {
  if (P <= 0 || Ssteps < 0 || chain0 < 0) return fail(ACHR_E_ARG, "bad P=%d S=%d chain0=%d", P, Ssteps, chain0);
  S->have_run = false;  // results of an earlier run are gone once its buffers are resized or reset
  const int n = S->n;
  const size_t Pn = static_cast<size_t>(P) * n;
  const int tr_chains = std::min(S->opt.trace_chains, P);
  const int tr_len = S->opt.trace_thin > 0 ? Ssteps / S->opt.trace_thin : 0;
  HIP_TRY(S->X.ensure(Pn));
  HIP_TRY(S->C.ensure(Pn));
  HIP_TRY(S->T.ensure(static_cast<size_t>(P) * std::max(S->d, 1)));
  HIP_TRY(S->out.ensure(Pn));
  HIP_TRY(S->st.ensure(P));
  HIP_TRY(S->stats.ensure(P));
  HIP_TRY(S->flag.ensure(P));
  HIP_TRY(S->idx.ensure(P));
  HIP_TRY(S->count.ensure(1));
  const int prec = S->opt.proj_prec;
  const bool res_form = prec == ACHR_PROJ_RES_BF16X3 || prec == ACHR_PROJ_RES_F32;
  const bool null_low = prec == ACHR_PROJ_NULL_F32 || prec == ACHR_PROJ_NULL_BF16X3;
  if (res_form) {
    // residual rows, zero in the padding columns the GEMM reads past m
    HIP_TRY(S->R32.ensure(static_cast<size_t>(P) * S->r32_ld));
    HIP_TRY(hipMemsetAsync(S->R32.p, 0, sizeof(float) * P * S->r32_ld, st));
  }
  if (null_low) HIP_TRY(S->T32.ensure(static_cast<size_t>(P) * std::max(S->d, 1)));
  S->rounds_fp64 = 0;
  S->rounds_low = 0;
  if (tr_chains > 0 && tr_len > 0) HIP_TRY(S->trace.ensure(static_cast<size_t>(tr_chains) * tr_len * n));
  S->P = P;
  S->S = Ssteps;
  S->trace_len = (tr_chains > 0) ? tr_len : 0;
  HIP_TRY(hipMemsetAsync(S->st.p, 0, sizeof(int4) * P, st));
  HIP_TRY(hipMemsetAsync(S->stats.p, 0, sizeof(int4) * P, st));
  HIP_TRY(hipMemsetAsync(S->flag.p, 0, sizeof(int) * P, st));
  S->ev_step.clear();
  S->ev_proj.clear();
  S->ev_check.clear();
  S->seed = seed;
  S->file = file;
  S->chain0 = chain0;
  S->trace_chains_run = S->trace_len > 0 ? tr_chains : 0;
  fill_args(S, P, Ssteps, seed, file, chain0, args);
  return ACHR_OK;
}

void fill_args(const achr_sampler* S, int P, int Ssteps, uint64_t seed, uint32_t file, int chain0,
               achr::StepArgs* args) {
  const int n = S->n;
  const int tr_chains = std::min(S->opt.trace_chains, P);
  achr::StepArgs& a = *args;
  a = achr::StepArgs{};
  a.n = n;
  a.m = S->m;
  a.W = S->use_dirs ? S->K : S->W;
  a.d = S->d;
  a.warm = S->use_dirs ? S->dirs.p : S->warm.p;
  a.dir_mode = S->use_dirs;
  a.lb = S->lb.p;
  a.ub = S->ub.p;
  a.c0 = S->use_dirs ? S->s0d.p : S->c0d.p;
  a.rowptr = S->rowptr.p;
  a.colidx = S->colidx.p;
  a.vals = S->vals.p;
  a.nslices = S->nslices;
  a.ell_ptr = S->ell_ptr.p;
  a.ell_col = S->ell_col.p;
  a.ell_val = S->ell_val.p;
  a.ell_row = S->ell_row.p;
  a.Ncm = S->Ncm.p;
  a.Nrm = S->Nrm.p;
  a.P = P;
  a.S = Ssteps;
  a.seed = seed;
  a.file = file;
  a.model = static_cast<uint32_t>(S->opt.model_id);
  a.chain0 = chain0;
  a.dTol = S->opt.dTol;
  a.uTol = S->opt.uTol;
  a.maxMinTol = S->opt.maxMinTol;
  a.reprojTol = S->opt.reprojTol;
  a.check_every = S->opt.check_every;
  a.max_misses = S->opt.max_misses;
  a.mode = S->mode;
  a.stop_flagged_only = S->mode == 2 && S->opt.proj_prec == ACHR_PROJ_FP64;
  a.check_budget = 10;
  a.X = S->X.p;
  a.C = S->C.p;
  a.T = S->T.p;
  a.st = S->st.p;
  a.stats = S->stats.p;
  a.flag = S->flag.p;
  const int prec = S->opt.proj_prec;
  a.R32 = (prec == ACHR_PROJ_RES_BF16X3 || prec == ACHR_PROJ_RES_F32) ? S->R32.p : nullptr;
  a.r32_ld = S->r32_ld;
  a.out = S->out.p;
  a.trace = S->trace_len > 0 ? S->trace.p : nullptr;
  a.trace_chains = tr_chains;
  a.trace_thin = S->opt.trace_thin;
  a.trace_len = S->trace_len;
}

int projection_kind(const achr_sampler* S, int seg) {
  if (S->mode != 2 || S->d <= 0) return kProjNone;
  const int prec = S->opt.proj_prec;
  if (prec == ACHR_PROJ_FP64 || (seg + 1) % S->opt.fp64_every == 0) return kProjFP64;
  return kProjLow;
}

int enqueue_projection(achr_sampler* S, int kind, hipStream_t st) {
  if (kind == kProjNone) return ACHR_OK;
  const int n = S->n;
  const int P = S->P;
  const int prec = S->opt.proj_prec;
  const bool res_form = prec == ACHR_PROJ_RES_BF16X3 || prec == ACHR_PROJ_RES_F32;
  const bool fp64_round = kind == kProjFP64;
  hipError_t le = hipSuccess;
  le = achr::launch_compact(S->flag.p, P, S->idx.p, S->count.p, st);
  if (le == hipSuccess && fp64_round) {
    le = achr::launch_project(S->Ncm.p, n, S->d, S->X.p, S->T.p, S->idx.p, S->count.p, P, st);
  } else if (le == hipSuccess && res_form) {
    // X(q, :) -= sum_k R(q, k) S+(:, k): A = residual rows, B = S+^T packed
    achr::GemmSubArgs g{};
    g.M = P;
    g.N = n;
    g.Kp = S->r32_ld;
    g.Mdev = S->count.p;
    g.A = S->R32.p;
    g.lda = S->r32_ld;
    g.idxA = S->idx.p;
    g.B = S->Bpk.p;
    g.C = S->X.p;
    g.ldc = n;
    g.idxC = S->idx.p;
    le = achr::launch_gemm_sub(g, prec == ACHR_PROJ_RES_F32 ? achr::kPrecF32 : achr::kPrecBF16x3, P, st);
  } else if (le == hipSuccess) {
    const int gp = prec == ACHR_PROJ_NULL_F32 ? achr::kPrecF32 : achr::kPrecBF16x3;
    // T(q, j) = sum_i X(q, i) N(i, j)
    achr::GemmArgs g{};
    g.M = P;
    g.N = S->d;
    g.K = n;
    g.Mdev = S->count.p;
    g.A = S->X.p;
    g.a_f64 = 1;
    g.sa_m = n;
    g.sa_k = 1;
    g.idxA = S->idx.p;
    g.B = S->N32.p;
    g.sb_k = 1;
    g.sb_n = n;
    g.C = S->T32.p;
    g.sc_m = S->d;
    g.sc_n = 1;
    g.alpha = 1.0f;
    g.epi = achr::kEpiStoreF32;
    le = achr::launch_gemm(g, gp, P, st);
    // X(q, i) = sum_j T(q, j) N(i, j)
    achr::GemmArgs h{};
    h.M = P;
    h.N = n;
    h.K = S->d;
    h.Mdev = S->count.p;
    h.A = S->T32.p;
    h.sa_m = S->d;
    h.sa_k = 1;
    h.B = S->N32.p;
    h.sb_k = n;
    h.sb_n = 1;
    h.C = S->X.p;
    h.sc_m = n;
    h.sc_n = 1;
    h.idxC = S->idx.p;
    h.alpha = 1.0f;
    h.epi = achr::kEpiStoreF64;
    if (le == hipSuccess) le = achr::launch_gemm(h, gp, P, st);
  }
  if (le != hipSuccess) return fail(ACHR_E_DEVICE, "projection launch: %s", hipGetErrorString(le));
  return ACHR_OK;
}

bool use_check_kernel(const achr_sampler* S, const achr::Schedule& sc) {
  // plain chains only: a direction bank keeps fixed reactions fixed, its
  // checks rarely flag and run every 100 steps, and it has no kernel without
  // the in-kernel check
  if (S->mode != 2 || S->d <= 0 || S->use_dirs || S->n > achr::kMaxCheckReactions || S->opt.check_kernel == 1)
    return false;
  return S->opt.check_kernel >= 2 || sc.tpc >= 512;
}

achr::CheckArgs check_args(const achr_sampler* S, const achr::StepArgs& a) {
  achr::CheckArgs c{};
  c.n = a.n;
  c.d = a.d;
  c.P = a.P;
  c.nslices = a.nslices;
  c.ell_ptr = a.ell_ptr;
  c.ell_col = a.ell_col;
  c.ell_val = a.ell_val;
  c.ell_row = a.ell_row;
  c.X = a.X;
  c.st = a.st;
  c.stats = a.stats;
  c.flag = a.flag;
  c.R32 = a.R32;
  c.r32_ld = a.r32_ld;
  c.reprojTol = a.reprojTol;
  if (S->chk_ok && S->opt.check_kernel != 3) {
    c.ent = S->chk_ent.p;
    c.rowmap = S->chk_row.p;
    c.vtab = S->chk_vtab.p;
    c.nv = S->chk_nv;
    for (int w = 0; w < achr::kCheckWaves; ++w) {
      c.wlen[w] = S->chk_wlen[w];
      c.woff[w] = S->chk_woff[w];
      c.roff[w] = S->chk_roff[w];
      c.wends[w] = S->chk_wends[w];
    }
  }
  return c;
}

int enqueue_segments(achr_sampler* S, const achr::StepArgs& args, int segments, int checks_done, hipStream_t st) {
  achr::StepArgs a = args;
  a.check_out = use_check_kernel(S, S->run_sched) ? 1 : 0;
  size_t evn = 0;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventRecord(S->ev_begin, st));
  for (int seg = 0; seg < segments; ++seg) {
    int rc = event_at(S, evn, &e0);
    if (rc == ACHR_OK) rc = event_at(S, evn + 1, &e1);
    if (rc != ACHR_OK) return rc;
    HIP_TRY(hipEventRecord(e0, st));
    const hipError_t le = achr::launch_step(a, S->run_sched, st);
    if (le != hipSuccess) return fail(ACHR_E_DEVICE, "step kernel launch: %s", hipGetErrorString(le));
    HIP_TRY(hipEventRecord(e1, st));
    S->ev_step.emplace_back(static_cast<int>(evn), static_cast<int>(evn + 1));
    evn += 2;
    if (a.check_out && seg + 1 < segments) {
      rc = event_at(S, evn, &e0);
      if (rc == ACHR_OK) rc = event_at(S, evn + 1, &e1);
      if (rc != ACHR_OK) return rc;
      HIP_TRY(hipEventRecord(e0, st));
      const hipError_t ce = achr::launch_check(check_args(S, a), st);
      if (ce != hipSuccess) return fail(ACHR_E_DEVICE, "check kernel launch: %s", hipGetErrorString(ce));
      HIP_TRY(hipEventRecord(e1, st));
      S->ev_check.emplace_back(static_cast<int>(evn), static_cast<int>(evn + 1));
      evn += 2;
    }
    const int kind = seg + 1 < segments ? projection_kind(S, checks_done + seg) : kProjNone;
    if (kind != kProjNone) {
      rc = event_at(S, evn, &e0);
      if (rc == ACHR_OK) rc = event_at(S, evn + 1, &e1);
      if (rc != ACHR_OK) return rc;
      HIP_TRY(hipEventRecord(e0, st));
      rc = enqueue_projection(S, kind, st);
      if (rc != ACHR_OK) return rc;
      if (kind == kProjFP64) ++S->rounds_fp64; else ++S->rounds_low;
      HIP_TRY(hipEventRecord(e1, st));
      S->ev_proj.emplace_back(static_cast<int>(evn), static_cast<int>(evn + 1));
      evn += 2;
    }
  }
  HIP_TRY(hipEventRecord(S->ev_end, st));
  note_run(S, S->run_sched);
  return ACHR_OK;
}

void note_run(achr_sampler* S, const achr::Schedule& sc) {
  S->run_use_dirs = S->use_dirs;
  S->run_check_every = S->opt.check_every;
  S->run_trace_thin = S->opt.trace_thin;
  S->run_trace_chains = S->opt.trace_chains;
  S->run_sched = sc;
}

}  // namespace achr_host

extern "C" {

int achr_sample(achr_sampler* S, int P, int Ssteps, uint64_t seed, uint32_t file, int chain0) {
  if (!S) return fail(ACHR_E_ARG, "sampler is NULL");
  int rc = achr_host::run_limits(S, S->sched, P, Ssteps);
  if (rc != ACHR_OK) return rc;
  hipStream_t st = S->stream;
  achr::StepArgs a{};
  rc = achr_host::prepare_run(S, P, Ssteps, seed, file, chain0, st, &a);
  if (rc != ACHR_OK) return rc;
  S->run_sched = S->sched;
  rc = achr_host::enqueue_segments(S, a, achr_host::segments_for(S, Ssteps), 0, st);
  if (rc != ACHR_OK) return rc;
  S->have_run = true;
  return ACHR_OK;
}

int achr_sample_continue(achr_sampler* S, int S_more) {
  if (!S) return fail(ACHR_E_ARG, "sampler is NULL");
  if (S_more < 0) return fail(ACHR_E_ARG, "bad S_more=%d", S_more);
  if (!S->have_run) return fail(ACHR_E_STATE, "no run to continue");
  if (S->run_use_dirs != S->use_dirs || S->run_check_every != S->opt.check_every ||
      S->run_trace_thin != S->opt.trace_thin || S->run_trace_chains != S->opt.trace_chains)
    return fail(ACHR_E_STATE, "directions, check period or trace settings changed since the run");
  if (S_more == 0) return ACHR_OK;
  if (S->S > INT_MAX - S_more) return fail(ACHR_E_ARG, "too many steps");
  const int S0 = S->S, S1 = S0 + S_more;
  int rc = achr_host::run_limits(S, S->run_sched, S->P, S1);
  if (rc != ACHR_OK) return rc;
  hipStream_t st = S->stream;
  if (S->trace_len > 0) {
    // more rows per traced chain, keeping the rows already recorded
    const int old_len = S->trace_len, new_len = S1 / S->opt.trace_thin;
    const size_t row = static_cast<size_t>(S->n) * sizeof(double);
    DevBuf<double> grown;
    HIP_TRY(grown.ensure(static_cast<size_t>(S->trace_chains_run) * new_len * S->n));
    HIP_TRY(hipMemcpy2DAsync(grown.p, row * new_len, S->trace.p, row * old_len, row * old_len, S->trace_chains_run,
                             hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    S->trace = std::move(grown);
    S->trace_len = new_len;
  }
  S->have_run = false;
  S->ev_step.clear();
  S->ev_proj.clear();
  S->ev_check.clear();
  achr::StepArgs a{};
  achr_host::fill_args(S, S->P, S1, S->seed, S->file, S->chain0, &a);
  const int ce = S->opt.check_every;
  const int done = (S0 + ce - 1) / ce, total = (S1 + ce - 1) / ce;
  S->S = S1;
  rc = achr_host::enqueue_segments(S, a, S->mode == 2 ? total - done + 1 : 1, done, st);
  if (rc != ACHR_OK) return rc;
  S->have_run = true;
  return ACHR_OK;
}

namespace {
// Checkpoint file: this header, then X (P x n doubles, chain major), C (the
// same, plain ACHR only), st and stats (P x 4 int32 each), little endian.
struct StateHeader {
  char magic[8];  // "ACHRST01"
  int32_t n, P, S, lanes, use_dirs, warm_count, check_every, mode, model_id, chain0;
  uint32_t file, pad;
  uint64_t seed;
};
constexpr char kStateMagic[8] = {'A', 'C', 'H', 'R', 'S', 'T', '0', '1'};
}  // namespace

static int save_state(const achr_sampler* S, const char* path);
static int load_state(achr_sampler* S, const char* path);

int achr_sampler_save_state(const achr_sampler* S, const char* path) {
  return achr_host::no_throw([&] { return save_state(S, path); });
}

int achr_sampler_load_state(achr_sampler* S, const char* path) {
  return achr_host::no_throw([&] { return load_state(S, path); });
}

}  // extern "C"

static int save_state(const achr_sampler* S, const char* path) {
  if (!S || !path) return fail(ACHR_E_ARG, "NULL argument");
  if (!S->have_run) return fail(ACHR_E_STATE, "no run");
  HIP_TRY(hipEventSynchronize(S->ev_end));
  const size_t Pn = static_cast<size_t>(S->P) * S->n;
  std::vector<double> X(Pn), C(S->use_dirs ? 0 : Pn);
  std::vector<int32_t> st(4 * static_cast<size_t>(S->P)), stats(4 * static_cast<size_t>(S->P));
  HIP_TRY(hipMemcpy(X.data(), S->X.p, sizeof(double) * Pn, hipMemcpyDeviceToHost));
  if (!C.empty()) HIP_TRY(hipMemcpy(C.data(), S->C.p, sizeof(double) * Pn, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(st.data(), S->st.p, sizeof(int32_t) * st.size(), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(stats.data(), S->stats.p, sizeof(int32_t) * stats.size(), hipMemcpyDeviceToHost));
  StateHeader h{};
  std::memcpy(h.magic, kStateMagic, 8);
  h.n = S->n;
  h.P = S->P;
  h.S = S->S;
  h.lanes = S->run_sched.tpc;
  h.use_dirs = S->use_dirs;
  h.warm_count = S->use_dirs ? S->K : S->W;
  h.check_every = S->opt.check_every;
  h.mode = S->mode;
  h.model_id = S->opt.model_id;
  h.chain0 = S->chain0;
  h.file = S->file;
  h.seed = S->seed;
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(ACHR_E_IO, "cannot write %s: %s", path, std::strerror(errno));
  bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(X.data(), sizeof(double), Pn, f) == Pn &&
            (C.empty() || std::fwrite(C.data(), sizeof(double), Pn, f) == Pn) &&
            std::fwrite(st.data(), sizeof(int32_t), st.size(), f) == st.size() &&
            std::fwrite(stats.data(), sizeof(int32_t), stats.size(), f) == stats.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(ACHR_E_IO, "write error on %s", path);
  return ACHR_OK;
}

static int load_state(achr_sampler* S, const char* path) {
  if (!S || !path) return fail(ACHR_E_ARG, "NULL argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(ACHR_E_IO, "cannot read %s: %s", path, std::strerror(errno));
  StateHeader h{};
  if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kStateMagic, 8) != 0) {
    std::fclose(f);
    return fail(ACHR_E_FORMAT, "%s is not a sampler state file", path);
  }
  const int warm_count = S->use_dirs ? S->K : S->W;
  // the saved run's lanes (a batch may have chosen them), if they can hold n
  const achr::Schedule sc{h.lanes, achr::elems_for(S->n, h.lanes), h.lanes <= 64 ? 256 / std::max(h.lanes, 1) : 1};
  if (h.n != S->n || !achr::schedule_supported(sc.tpc, sc.e) || h.use_dirs != S->use_dirs || h.warm_count != warm_count ||
      h.check_every != S->opt.check_every || h.mode != S->mode || h.model_id != S->opt.model_id || h.P <= 0 ||
      h.P > (1 << 28) || h.S < 0 || h.chain0 < 0) {
    std::fclose(f);
    return fail(ACHR_E_STATE, "%s was saved by a sampler with other settings (n, lanes that cannot hold n, "
                "directions, warmup or bank size, check period, reprojection mode or model_id)", path);
  }
  const size_t Pn = static_cast<size_t>(h.P) * h.n;
  std::vector<double> X(Pn), C(h.use_dirs ? 0 : Pn);
  std::vector<int32_t> st(4 * static_cast<size_t>(h.P)), stats(4 * static_cast<size_t>(h.P));
  const bool ok = std::fread(X.data(), sizeof(double), Pn, f) == Pn &&
                  (C.empty() || std::fread(C.data(), sizeof(double), Pn, f) == Pn) &&
                  std::fread(st.data(), sizeof(int32_t), st.size(), f) == st.size() &&
                  std::fread(stats.data(), sizeof(int32_t), stats.size(), f) == stats.size();
  std::fclose(f);
  if (!ok) return fail(ACHR_E_FORMAT, "%s is truncated", path);
  int rc = achr_host::run_limits(S, sc, h.P, h.S);
  if (rc != ACHR_OK) return rc;
  hipStream_t stream = S->stream;
  achr::StepArgs a{};
  rc = achr_host::prepare_run(S, h.P, h.S, h.seed, h.file, h.chain0, stream, &a);
  if (rc != ACHR_OK) return rc;
  S->trace_len = 0;  // traces are not part of the state
  S->trace_chains_run = 0;
  std::vector<double> out(Pn);
  for (int b = 0; b < h.P; ++b)
    for (int i = 0; i < h.n; ++i) out[static_cast<size_t>(i) * h.P + b] = X[static_cast<size_t>(b) * h.n + i];
  HIP_TRY(S->X.upload(X.data(), Pn, stream));
  if (!C.empty()) HIP_TRY(S->C.upload(C.data(), Pn, stream));
  HIP_TRY(S->out.upload(out.data(), Pn, stream));
  HIP_TRY(hipMemcpyAsync(S->st.p, st.data(), sizeof(int32_t) * st.size(), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipMemcpyAsync(S->stats.p, stats.data(), sizeof(int32_t) * stats.size(), hipMemcpyHostToDevice, stream));
  HIP_TRY(hipEventRecord(S->ev_begin, stream));
  HIP_TRY(hipEventRecord(S->ev_end, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  achr_host::note_run(S, sc);
  S->have_run = true;
  return ACHR_OK;
}

extern "C" {

int achr_sample_points(achr_sampler* S, double* out) {
  if (!S || !out) return fail(ACHR_E_ARG, "NULL argument");
  if (!S->have_run) return fail(ACHR_E_STATE, "no run");
  HIP_TRY(hipEventSynchronize(S->ev_end));
  HIP_TRY(hipMemcpy(out, S->out.p, sizeof(double) * static_cast<size_t>(S->P) * S->n, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_sample_points_device(achr_sampler* S, void** dev_ptr) {
  if (!S || !dev_ptr) return fail(ACHR_E_ARG, "NULL argument");
  if (!S->have_run) return fail(ACHR_E_STATE, "no run");
  *dev_ptr = S->out.p;
  return ACHR_OK;
}

int achr_sample_stats(achr_sampler* S, int32_t* stats) {
  if (!S || !stats) return fail(ACHR_E_ARG, "NULL argument");
  if (!S->have_run) return fail(ACHR_E_STATE, "no run");
  HIP_TRY(hipEventSynchronize(S->ev_end));
  HIP_TRY(hipMemcpy(stats, S->stats.p, sizeof(int4) * S->P, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_sample_trace(achr_sampler* S, double* trace) {
  if (!S || !trace) return fail(ACHR_E_ARG, "NULL argument");
  if (!S->have_run) return fail(ACHR_E_STATE, "no run");
  if (S->trace_len == 0) return fail(ACHR_E_STATE, "no trace recorded (trace_chains / trace_thin)");
  HIP_TRY(hipEventSynchronize(S->ev_end));
  const size_t cnt = static_cast<size_t>(S->trace_chains_run) * S->trace_len * S->n;
  HIP_TRY(hipMemcpy(trace, S->trace.p, sizeof(double) * cnt, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_sample_timing(achr_sampler* S, double* ms_total, double* ms_step, double* ms_proj, int* n_step,
                       int* n_proj) {
  if (!S) return fail(ACHR_E_ARG, "sampler is NULL");
  if (!S->have_run) return fail(ACHR_E_STATE, "no run");
  HIP_TRY(hipEventSynchronize(S->ev_end));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, S->ev_begin, S->ev_end));
  double tstep = 0.0, tproj = 0.0;
  for (auto& pr : S->ev_step) {
    HIP_TRY(hipEventElapsedTime(&ms, S->ev_pool[pr.first], S->ev_pool[pr.second]));
    tstep += ms;
  }
  for (auto& pr : S->ev_proj) {
    HIP_TRY(hipEventElapsedTime(&ms, S->ev_pool[pr.first], S->ev_pool[pr.second]));
    tproj += ms;
  }
  HIP_TRY(hipEventElapsedTime(&ms, S->ev_begin, S->ev_end));
  if (ms_total) *ms_total = ms;
  if (ms_step) *ms_step = tstep;
  if (ms_proj) *ms_proj = tproj;
  if (n_step) *n_step = static_cast<int>(S->ev_step.size());
  if (n_proj) *n_proj = static_cast<int>(S->ev_proj.size());
  return ACHR_OK;
}

int achr_sample_timing_check(achr_sampler* S, double* ms_check, int* n_check) {
  if (!S) return fail(ACHR_E_ARG, "sampler is NULL");
  if (!S->have_run) return fail(ACHR_E_STATE, "no run");
  HIP_TRY(hipEventSynchronize(S->ev_end));
  double t = 0.0;
  for (auto& pr : S->ev_check) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, S->ev_pool[pr.first], S->ev_pool[pr.second]));
    t += ms;
  }
  if (ms_check) *ms_check = t;
  if (n_check) *n_check = static_cast<int>(S->ev_check.size());
  return ACHR_OK;
}

int achr_read_bandwidth(long bytes, int reps, double* gb_per_s) {
  if (bytes < (1L << 20) || reps < 1 || !gb_per_s) return fail(ACHR_E_ARG, "bad read-rate arguments");
  const size_t words = static_cast<size_t>(bytes) / 16;  // 16-byte loads
  DevBuf<double> buf, sums;
  HIP_TRY(buf.ensure(words * 2));
  HIP_TRY(hipMemset(buf.p, 0, words * 16));
  int blocks = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(achr::read_stream_blocks(&blocks));
  HIP_TRY(sums.ensure(static_cast<size_t>(blocks)));
  HIP_TRY(hipEventCreate(&e0));
  hipError_t e = hipEventCreate(&e1);
  float best = 1e30f;
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) e = achr::launch_read_stream(buf.p, words, sums.p, blocks, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) return fail(ACHR_E_DEVICE, "read-rate kernel: %s", hipGetErrorString(e));
  *gb_per_s = 16.0 * static_cast<double>(words) / (best * 1e-3) / 1e9;
  return ACHR_OK;
}

int achr_bank_reciprocal(const double* u, double* r, long count) {
  if (!u || !r || count < 0) return fail(ACHR_E_ARG, "bad reciprocal arguments");
  if (count == 0) return ACHR_OK;
  const size_t nc = static_cast<size_t>(count);
  DevBuf<double> du, dr;
  HIP_TRY(du.upload(u, nc, nullptr));
  HIP_TRY(dr.ensure(nc));
  const hipError_t e = achr::launch_bank_reciprocal(du.p, dr.p, nc, nullptr);
  if (e != hipSuccess) return fail(ACHR_E_DEVICE, "reciprocal launch: %s", hipGetErrorString(e));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(r, dr.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_quotient(const double* a, const double* b, double* q, long count) {
  if (!a || !b || !q || count < 0) return fail(ACHR_E_ARG, "bad quotient arguments");
  if (count == 0) return ACHR_OK;
  const size_t nc = static_cast<size_t>(count);
  DevBuf<double> da, db, dq;
  HIP_TRY(da.upload(a, nc, nullptr));
  HIP_TRY(db.upload(b, nc, nullptr));
  HIP_TRY(dq.ensure(nc));
  const hipError_t e = achr::launch_quotient(da.p, db.p, dq.p, nc, nullptr);
  if (e != hipSuccess) return fail(ACHR_E_DEVICE, "quotient launch: %s", hipGetErrorString(e));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(q, dq.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_gemm_f32(int M, int N, int K, const float* A, long sa_m, long sa_k, const float* B, long sb_k, long sb_n,
                  float* C, long sc_m, long sc_n, float alpha, float beta, int prec) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return fail(ACHR_E_ARG, "bad GEMM arguments");
  if ((sa_m != 1 && sa_k != 1) || (sb_k != 1 && sb_n != 1)) return fail(ACHR_E_ARG, "each operand needs a unit stride");
  auto extent = [](long r, long c, long sr, long sc) { return (r - 1) * sr + (c - 1) * sc + 1; };
  const size_t na = static_cast<size_t>(extent(M, K, sa_m, sa_k));
  const size_t nb = static_cast<size_t>(extent(K, N, sb_k, sb_n));
  const size_t nc = static_cast<size_t>(extent(M, N, sc_m, sc_n));
  DevBuf<float> dA, dB, dC;
  HIP_TRY(dA.upload(A, na, nullptr));
  HIP_TRY(dB.upload(B, nb, nullptr));
  HIP_TRY(dC.upload(C, nc, nullptr));
  achr::GemmArgs g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.A = dA.p;
  g.sa_m = sa_m;
  g.sa_k = sa_k;
  g.B = dB.p;
  g.sb_k = sb_k;
  g.sb_n = sb_n;
  g.C = dC.p;
  g.sc_m = sc_m;
  g.sc_n = sc_n;
  g.alpha = alpha;
  g.beta = beta;
  g.epi = achr::kEpiStoreF32;
  const hipError_t e = achr::launch_gemm(g, prec, M, nullptr);
  if (e != hipSuccess) return fail(ACHR_E_ARG, "gemm launch: %s", hipGetErrorString(e));
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(C, dC.p, sizeof(float) * nc, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

int achr_gemm_sub_f64(int rows, int M, int N, int K, const float* A, const float* B, double* C, const int* idx,
                      int prec, float* kernel_ms) {
  if (rows <= 0 || M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return fail(ACHR_E_ARG, "bad GEMM arguments");
  if (prec != 0 && prec != 1) return fail(ACHR_E_ARG, "prec %d: 0 (FP32) or 1 (split BF16x3)", prec);
  if (!idx && M > rows) return fail(ACHR_E_ARG, "M=%d rows of %d", M, rows);
  if (idx)
    for (int q = 0; q < M; ++q)
      if (idx[q] < 0 || idx[q] >= rows) return fail(ACHR_E_ARG, "idx[%d]=%d outside %d rows", q, idx[q], rows);
  const int Kp = achr::packed_k(K);
  std::vector<float> Ap(static_cast<size_t>(rows) * Kp, 0.0f);
  for (int r = 0; r < rows; ++r)
    std::copy(A + static_cast<size_t>(r) * K, A + static_cast<size_t>(r) * K + K, Ap.begin() + static_cast<size_t>(r) * Kp);
  DevBuf<float> dA, dB;
  DevBuf<unsigned char> dP;
  DevBuf<double> dC;
  DevBuf<int> dI, dM;
  HIP_TRY(dA.upload(Ap.data(), Ap.size(), nullptr));
  HIP_TRY(dB.upload(B, static_cast<size_t>(K) * N, nullptr));
  HIP_TRY(dC.upload(C, static_cast<size_t>(rows) * N, nullptr));
  HIP_TRY(dM.upload(&M, 1, nullptr));
  if (idx) HIP_TRY(dI.upload(idx, M, nullptr));
  HIP_TRY(dP.ensure(achr::packed_b_bytes(K, N)));
  const int gp = prec == 0 ? achr::kPrecF32 : achr::kPrecBF16x3;
  hipError_t e = achr::launch_pack_b(dB.p, N, 1, K, N, gp, dP.p, nullptr);
  achr::GemmSubArgs g{};
  g.M = M;
  g.N = N;
  g.Kp = Kp;
  g.Mdev = dM.p;
  g.A = dA.p;
  g.lda = Kp;
  g.idxA = idx ? dI.p : nullptr;
  g.B = dP.p;
  g.C = dC.p;
  g.ldc = N;
  g.idxC = idx ? dI.p : nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess && kernel_ms) e = hipEventCreate(&t0);
  if (e == hipSuccess && kernel_ms) e = hipEventCreate(&t1);
  if (e == hipSuccess && kernel_ms) e = hipEventRecord(t0, nullptr);
  if (e == hipSuccess) e = achr::launch_gemm_sub(g, gp, M, nullptr);
  if (e == hipSuccess && kernel_ms) e = hipEventRecord(t1, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && kernel_ms) e = hipEventElapsedTime(kernel_ms, t0, t1);
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  if (e != hipSuccess) return fail(ACHR_E_DEVICE, "gemm_sub: %s", hipGetErrorString(e));
  HIP_TRY(hipMemcpy(C, dC.p, sizeof(double) * rows * N, hipMemcpyDeviceToHost));
  return ACHR_OK;
}

}  // extern "C"
