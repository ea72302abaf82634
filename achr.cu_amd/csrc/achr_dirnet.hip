// Direction network: a 3-layer MLP that maps a Gaussian latent z (r
// numbers) to a direction in an r-dimensional coordinate system of the free
// null space, trained so that the directions' second moment matches a target
// covariance C (the samples' covariance in that system) — a learned rounding
// map of the polytope.
//
//   F(z) = W3 gelu(W2 gelu(W1 z + b1) + b2) + b3,   v(z) = F(z) - F(-z)
//
// v is odd, so the direction law is symmetric and, being a function of a
// fresh latent only, independent of the chain's state: hit-and-run along such
// directions keeps the uniform law on the polytope stationary.
// Loss: L = |V^T V / B - C|_F^2 / |C|_F^2 over a batch of B latents.
// Training and inference use the library's matrix-core GEMM (FP32 for
// training, BF16 for inference) and the elementwise kernels below; the
// parameters are FP32 and updated with Adam.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

#include "achr_dirnet.h"
#include "achr_gemm.h"
#include "achr_philox.h"

namespace achr {
namespace {

constexpr float kInvSqrt2 = 0.70710678118654752f;
constexpr float kInvSqrt2Pi = 0.39894228040143268f;

__global__ void gelu_fwd_kernel(const float* __restrict__ a, float* __restrict__ h, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const float x = a[i];
    h[i] = 0.5f * x * (1.0f + erff(x * kInvSqrt2));
  }
}

// dA = dH * (Phi(a) + a phi(a)), in place on dH
__global__ void gelu_bwd_kernel(float* __restrict__ dh, const float* __restrict__ a, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const float x = a[i];
    const float cdf = 0.5f * (1.0f + erff(x * kInvSqrt2));
    const float pdf = kInvSqrt2Pi * expf(-0.5f * x * x);
    dh[i] *= cdf + x * pdf;
  }
}

// Y[r, :] = bias for every row (the GEMM then adds with beta = 1).
__global__ void bias_rows_kernel(float* __restrict__ y, const float* __restrict__ bias, int rows, int cols) {
  const size_t n = static_cast<size_t>(rows) * cols;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    y[i] = bias[i % cols];
}

// db[j] = sum_r dY[r, j] in a fixed order: 64 columns per workgroup, lane =
// column, wavefront w summing rows w, w + 16, w + 32, ... (coalesced 256-byte
// rows, sixteen independent chains), then the 16 partials of a column added
// pairwise in a fixed tree. (One thread per column over all rows left all
// but a few CUs idle: 0.94 ms for 4,096 x 512.)
constexpr int kColsumWaves = 16;
__global__ __launch_bounds__(64 * kColsumWaves) void colsum_kernel(const float* __restrict__ dy,
                                                                   float* __restrict__ db, int rows, int cols) {
  __shared__ float part[kColsumWaves][64];
  const int lane = static_cast<int>(threadIdx.x) & 63, w = static_cast<int>(threadIdx.x) >> 6;
  const int j = static_cast<int>(blockIdx.x) * 64 + lane;
  float s = 0.0f;
  if (j < cols) {
#pragma unroll 8
    for (int r = w; r < rows; r += kColsumWaves) s += dy[static_cast<size_t>(r) * cols + j];
  }
  part[w][lane] = s;
  __syncthreads();
  if (w == 0 && j < cols) {
    float p[kColsumWaves];
#pragma unroll
    for (int i = 0; i < kColsumWaves; ++i) p[i] = part[i][lane];
#pragma unroll
    for (int o = kColsumWaves / 2; o >= 1; o >>= 1)
#pragma unroll
      for (int i = 0; i < o; ++i) p[i] = p[i] + p[i + o];
    db[j] = p[0];
  }
}

// C (m x n, row major) = alpha * P^T Q, with P (k x m) and Q (k x n) row major:
// the weight-side products (dW = dY^T X, V^T V) have small outputs and long
// k, 16 of this library's 128 x 128 tiles for a 512 x 512 output, so they go
// to rocBLAS (column major: C^T = Q^T P, i.e. sgemm(N, T) on the stored
// arrays), with atomics off so that every run sums in the same order.
hipError_t sgemm_cm(void*& blas, rocblas_operation ta, rocblas_operation tb, int m, int n, int k, float alpha,
                    const float* A, int lda, const float* B, int ldb, float beta, float* C, int ldc, hipStream_t s) {
  if (!blas) {
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return hipErrorNotInitialized;
    if (rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed) != rocblas_status_success) {
      rocblas_destroy_handle(h);
      return hipErrorNotInitialized;
    }
    blas = h;
  }
  rocblas_handle h = static_cast<rocblas_handle>(blas);
  if (rocblas_set_stream(h, s) != rocblas_status_success) return hipErrorInvalidHandle;
  const rocblas_status st = rocblas_sgemm(h, ta, tb, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc);
  return st == rocblas_status_success ? hipSuccess : hipErrorLaunchFailure;
}

hipError_t gemm_tn(void*& blas, const float* P, const float* Q, float* C, int k, int m, int n, float alpha,
                   hipStream_t s) {
  return sgemm_cm(blas, rocblas_operation_none, rocblas_operation_transpose, n, m, k, alpha, Q, n, P, m, 0.0f, C, n,
                  s);
}

// Z2 = [Z; -Z] from Z (B x L).
__global__ void mirror_kernel(const float* __restrict__ z, float* __restrict__ z2, int B, int L) {
  const size_t n = static_cast<size_t>(B) * L;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    z2[i] = z[i];
    z2[i + n] = -z[i];
  }
}

// V = F[:B] - F[B:]
__global__ void odd_combine_kernel(const float* __restrict__ f, float* __restrict__ v, int B, int O) {
  const size_t n = static_cast<size_t>(B) * O;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    v[i] = f[i] - f[i + n];
}

// dF = [dV; -dV]
__global__ void odd_split_kernel(const float* __restrict__ dv, float* __restrict__ df, int B, int O) {
  const size_t n = static_cast<size_t>(B) * O;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    df[i] = dv[i];
    df[i + n] = -dv[i];
  }
}

// E = A / B - C (in place on A)
__global__ void cov_err_kernel(float* __restrict__ a, const float* __restrict__ c, float invB, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    a[i] = a[i] * invB - c[i];
}

// Standard normal latents from the Philox stream (Box-Muller on two uniforms).
__global__ void gauss_kernel(float* __restrict__ z, size_t n, uint64_t seed, uint32_t iter) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const u32x4 r = stream_block(seed, static_cast<uint32_t>(i), iter, 0xD1u, 0u, kTagDir);
    const double u1 = (static_cast<double>(r.v0) + 1.0) * 0x1.0p-32;  // (0, 1]
    const double u2 = static_cast<double>(r.v1) * 0x1.0p-32;
    z[i] = static_cast<float>(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  }
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

dim3 grid_for(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 16384) b = 16384;
  if (b == 0) b = 1;
  return dim3(static_cast<unsigned>(b));
}

// Y (rows x out) = X (rows x in) W^T (+ bias rows): W is out x in row major.
// FP32 (training) on rocBLAS: at rows = 4,096 and 512 outputs this library's
// 128 x 128 tiles give 128 workgroups for 256 CUs (113 us against rocBLAS's
// smaller tiles); column major, Y^T = W X^T is sgemm(T, N) on the stored
// arrays. The BF16 inference of the bank stays on gemm_kernel.
hipError_t linear_fwd(void*& blas, const float* X, const float* W, const float* bias, float* Y, int rows, int in,
                      int out, int prec, hipStream_t s) {
  hipLaunchKernelGGL(bias_rows_kernel, grid_for(static_cast<size_t>(rows) * out), dim3(256), 0, s, Y, bias, rows,
                     out);
  if (prec == kPrecF32)
    return sgemm_cm(blas, rocblas_operation_transpose, rocblas_operation_none, out, rows, in, 1.0f, W, in, X, in, 1.0f,
                    Y, out, s);
  GemmArgs g{};
  g.M = rows;
  g.N = out;
  g.K = in;
  g.A = X;
  g.sa_m = in;
  g.sa_k = 1;
  g.B = W;
  g.sb_k = 1;
  g.sb_n = in;
  g.C = Y;
  g.sc_m = out;
  g.sc_n = 1;
  g.alpha = 1.0f;
  g.beta = 1.0f;
  g.epi = kEpiStoreF32;
  return launch_gemm(g, prec, rows, s);
}

// dX (rows x in) = dY (rows x out) W ; dW (out x in) = dY^T X ; db = colsum(dY)
hipError_t linear_bwd(void*& blas, const float* X, const float* W, const float* dY, float* dX, float* dW, float* db,
                      int rows, int in, int out, hipStream_t s) {
  hipError_t e = hipSuccess;
  if (dX) {
    // column major: dX^T (in x rows) = W^T dY^T
    e = sgemm_cm(blas, rocblas_operation_none, rocblas_operation_none, in, rows, out, 1.0f, W, in, dY, out, 0.0f, dX,
                 in, s);
    if (e != hipSuccess) return e;
  }
  // dW (out x in) = dY^T X
  e = gemm_tn(blas, dY, X, dW, rows, out, in, 1.0f, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(colsum_kernel, dim3((out + 63) / 64), dim3(64 * kColsumWaves), 0, s, dY, db, rows, out);
  return hipGetLastError();
}

}  // namespace

size_t DirNet::param_count() const {
  return static_cast<size_t>(H) * L + H + static_cast<size_t>(H) * H + H + static_cast<size_t>(O) * H + O;
}

hipError_t DirNet::alloc(int latent, int width, int outdim) {
  L = latent;
  H = width;
  O = outdim;
  const size_t np = param_count();
  hipError_t e = hipMalloc(&params, np * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&grads, np * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&adam_m, np * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&adam_v, np * sizeof(float));
  if (e == hipSuccess) e = hipMemset(adam_m, 0, np * sizeof(float));
  if (e == hipSuccess) e = hipMemset(adam_v, 0, np * sizeof(float));
  return e;
}

void DirNet::release() {
  if (blas) {
    rocblas_destroy_handle(static_cast<rocblas_handle>(blas));
    blas = nullptr;
  }
  for (float** p : {&params, &grads, &adam_m, &adam_v, &z, &z2, &a1, &h1, &a2, &h2, &f, &v, &dv, &df, &cov, &target,
                    &scratch})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  cap_B = 0;
  scratch_cap = 0;
}

hipError_t DirNet::ensure_batch(int B) {
  if (B <= cap_B) return hipSuccess;
  for (float** p : {&z, &z2, &a1, &h1, &a2, &h2, &f, &v, &dv, &df})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  const size_t B2 = 2 * static_cast<size_t>(B);
  hipError_t e = hipMalloc(&z, sizeof(float) * B * L);
  if (e == hipSuccess) e = hipMalloc(&z2, sizeof(float) * B2 * L);
  if (e == hipSuccess) e = hipMalloc(&a1, sizeof(float) * B2 * H);
  if (e == hipSuccess) e = hipMalloc(&h1, sizeof(float) * B2 * H);
  if (e == hipSuccess) e = hipMalloc(&a2, sizeof(float) * B2 * H);
  if (e == hipSuccess) e = hipMalloc(&h2, sizeof(float) * B2 * H);
  if (e == hipSuccess) e = hipMalloc(&f, sizeof(float) * B2 * O);
  if (e == hipSuccess) e = hipMalloc(&v, sizeof(float) * B * O);
  if (e == hipSuccess) e = hipMalloc(&dv, sizeof(float) * B * O);
  if (e == hipSuccess) e = hipMalloc(&df, sizeof(float) * B2 * O);
  if (e == hipSuccess) cap_B = B;
  return e;
}

hipError_t DirNet::ensure_cov() {
  if (cov) return hipSuccess;
  hipError_t e = hipMalloc(&cov, sizeof(float) * O * O);
  if (e == hipSuccess) e = hipMalloc(&target, sizeof(float) * O * O);
  return e;
}

// Forward of the odd map on this->z (B x L) into this->v (B x O).
hipError_t DirNet::forward(int B, int prec, hipStream_t s) {
  float* W1 = params;
  float* b1 = W1 + static_cast<size_t>(H) * L;
  float* W2 = b1 + H;
  float* b2 = W2 + static_cast<size_t>(H) * H;
  float* W3 = b2 + H;
  float* b3 = W3 + static_cast<size_t>(O) * H;
  const int B2 = 2 * B;
  hipLaunchKernelGGL(mirror_kernel, grid_for(static_cast<size_t>(B) * L), dim3(256), 0, s, z, z2, B, L);
  hipError_t e = linear_fwd(blas, z2, W1, b1, a1, B2, L, H, prec, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gelu_fwd_kernel, grid_for(static_cast<size_t>(B2) * H), dim3(256), 0, s, a1, h1,
                     static_cast<size_t>(B2) * H);
  e = linear_fwd(blas, h1, W2, b2, a2, B2, H, H, prec, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gelu_fwd_kernel, grid_for(static_cast<size_t>(B2) * H), dim3(256), 0, s, a2, h2,
                     static_cast<size_t>(B2) * H);
  e = linear_fwd(blas, h2, W3, b3, f, B2, H, O, prec, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(odd_combine_kernel, grid_for(static_cast<size_t>(B) * O), dim3(256), 0, s, f, v, B, O);
  return hipGetLastError();
}

// Loss gradient for the batch in this->z against this->target; grads into this->grads.
hipError_t DirNet::backward(int B, float cnorm2, hipStream_t s) {
  float* W1 = params;
  float* b1 = W1 + static_cast<size_t>(H) * L;
  float* W2 = b1 + H;
  float* b2 = W2 + static_cast<size_t>(H) * H;
  float* W3 = b2 + H;
  (void)b1;
  (void)b2;
  float* gW1 = grads;
  float* gb1 = gW1 + static_cast<size_t>(H) * L;
  float* gW2 = gb1 + H;
  float* gb2 = gW2 + static_cast<size_t>(H) * H;
  float* gW3 = gb2 + H;
  float* gb3 = gW3 + static_cast<size_t>(O) * H;
  const int B2 = 2 * B;
  // cov = V^T V (O x O), then E = cov / B - C
  hipError_t e = gemm_tn(blas, v, v, cov, B, O, O, 1.0f, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(cov_err_kernel, grid_for(static_cast<size_t>(O) * O), dim3(256), 0, s, cov, target,
                     1.0f / static_cast<float>(B), static_cast<size_t>(O) * O);
  // dV = (4 / (B |C|^2)) V E; column major: dV^T (O x B) = E^T V^T, the
  // stored E (row major) read as E^T
  e = sgemm_cm(blas, rocblas_operation_none, rocblas_operation_none, O, B, O, 4.0f / (static_cast<float>(B) * cnorm2),
               cov, O, v, O, 0.0f, dv, O, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(odd_split_kernel, grid_for(static_cast<size_t>(B) * O), dim3(256), 0, s, dv, df, B, O);
  // layer 3: dH2 goes to a scratch buffer; the layer-2 input gradient then
  // overwrites h2, which layer 3 no longer needs.
  if (!scratch || scratch_cap < static_cast<size_t>(B2) * H) {
    if (scratch) (void)hipFree(scratch);
    scratch = nullptr;
    scratch_cap = 0;
    e = hipMalloc(&scratch, sizeof(float) * static_cast<size_t>(B2) * H);
    if (e != hipSuccess) return e;
    scratch_cap = static_cast<size_t>(B2) * H;
  }
  float* dH = scratch;
  e = linear_bwd(blas, h2, W3, df, dH, gW3, gb3, B2, H, O, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gelu_bwd_kernel, grid_for(static_cast<size_t>(B2) * H), dim3(256), 0, s, dH, a2,
                     static_cast<size_t>(B2) * H);
  // layer 2: dH1 overwrites h2 (no longer needed)
  e = linear_bwd(blas, h1, W2, dH, h2, gW2, gb2, B2, H, H, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gelu_bwd_kernel, grid_for(static_cast<size_t>(B2) * H), dim3(256), 0, s, h2, a1,
                     static_cast<size_t>(B2) * H);
  // layer 1 (no input gradient)
  return linear_bwd(blas, z2, W1, h2, nullptr, gW1, gb1, B2, L, H, s);
}

hipError_t DirNet::gaussian_latents(int B, uint64_t seed, uint32_t iter, hipStream_t s) {
  hipLaunchKernelGGL(gauss_kernel, grid_for(static_cast<size_t>(B) * L), dim3(256), 0, s, z,
                     static_cast<size_t>(B) * L, seed, iter);
  return hipGetLastError();
}

hipError_t DirNet::adam_step(float lr, int t, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, static_cast<float>(t));
  const float c2 = 1.0f - std::pow(b2, static_cast<float>(t));
  hipLaunchKernelGGL(adam_kernel, grid_for(param_count()), dim3(256), 0, s, params, grads, adam_m, adam_v,
                     param_count(), lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

void DirNet::init_params(uint64_t seed, std::vector<float>& host) const {
  // Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases, from the Philox stream.
  host.assign(param_count(), 0.0f);
  size_t off = 0;
  auto fill = [&](int rows, int cols, uint32_t tag) {
    const float a = 1.0f / std::sqrt(static_cast<float>(cols));
    for (size_t i = 0; i < static_cast<size_t>(rows) * cols; ++i) {
      const double u = stream_u01(seed, static_cast<uint32_t>(i), tag, 0xD0u, 0u, kTagDir);
      host[off + i] = static_cast<float>((2.0 * u - 1.0) * a);
    }
    off += static_cast<size_t>(rows) * cols;
  };
  fill(H, L, 1);
  off += H;
  fill(H, H, 2);
  off += H;
  fill(O, H, 3);
}

}  // namespace achr
