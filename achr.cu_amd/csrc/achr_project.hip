// Batched FP64 null-space reprojection X <- N (N^T X) over the chains that
// asked for it in the current round.
//
// The reference reprojects one chain at a time with two GEMVs
// (reprojectPoint / reprojectPoint2, ACHRcu/ACHRCuda.cu:208-230), which at
// genome scale reads the whole of N (n x d doubles) per chain per check.
// Here the flagged chains of a round are gathered and both products run as
// tiled FP64 GEMMs on the matrix cores, so N is read once per 128 chains.
// Each output is accumulated in increasing k starting from +0.0 — the loop
// order of the reference's two kernels — so the result is bitwise the
// sequential fused-multiply-add dot product the host oracle computes:
// v_mfma_f64_16x16x4f64 adds its four products to the accumulator as four
// chained fused multiply-adds in k order (checked bit for bit against that
// chain, cancellations and wide exponent ranges included), and consecutive
// instructions continue the chain through the accumulator.
#include <algorithm>

#include "achr_internal.h"
#include "achr_philox.h"

namespace achr {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 16;
constexpr int kThreads = 256;                     // 4 wavefronts, each a 64 x 64 quarter of the tile
constexpr int kPerThread = (BK * BM) / kThreads;  // 8 elements of A and of B per thread and slice
constexpr int kRowStride = BK + 2;  // row-major tiles (k contiguous): 18 doubles per row
constexpr int kColStride = BM + 16; // k-major tile of A: 144 doubles per k
// With these strides both the fragment reads (16 rows x 4 k per wavefront)
// and the tile stores spread over the LDS banks at the 2-way minimum that 64
// lanes of 8 bytes need (3-way for the row-major stores).

// C(r, q) = sum_k A(r, k) * B(k, q) for r < M, k < K, q < *count.
// A(r, k) = A[r*sar + k*sak] (AK: sak == 1, rows contiguous in k); column q
// of B starts at B + colB(q)*ldb and is contiguous in k; column q of C starts
// at Cm + colC(q)*ldc, contiguous in r. colB(q) = idxB ? idxB[q] : q,
// likewise colC.
//
// 128 x 128 output tile per 256-thread workgroup. Wavefront w owns rows
// (w / 2) * 64 + [0, 64) and columns (w % 2) * 64 + [0, 64) as 4 x 4 blocks of
// 16 x 16, each accumulated by v_mfma_f64_16x16x4f64 (lane l supplies A row
// l % 16, k l / 16 and B k l / 16, column l % 16; it holds D rows
// l / 16 + 4 i, column l % 16). Per 4-deep k step a lane reads 4 + 4 doubles
// from LDS for 16 matrix instructions: 0.25 bytes per multiply-add. The k
// loop runs over slices of 16 with two LDS stages: the next slice is loaded
// into registers while the current one is consumed, and stored into the other
// stage, one barrier per slice.
//
// Loads are unconditional: rows past M and columns past the count read row
// M - 1 / column count - 1 (their outputs are never stored, and an output
// depends only on its own row of A and column of B), and k past K reads
// k = K - 1 and is replaced by 0.0, which adds nothing. Operands contiguous
// in k are read 16 lanes per row, whole 128-byte lines per instruction.
//
// A round of few columns (the flagged chains of one check period) gets only
// a few column tiles, too few workgroups for the CUs: TBN = 32 is the same
// kernel with 128 x 32 tiles (wavefront w owns rows 32 w + [0, 32) as 2 x 2
// blocks), four times the workgroups per column. Both widths are enqueued
// and each runs only when the round's count (on the device) is in its range
// (COLS_LE: at most kNarrowCols columns; else more), so the choice needs no
// host synchronisation. Each output is the same k-ordered chain either way.
constexpr int kNarrowCols = 1024;
template <bool AK, int TBN = BN, int COLS_LE = 0>
__global__ __launch_bounds__(kThreads, 2) void dgemm_mfma_kernel(int M, int K, const double* __restrict__ A,
                                                               long sar, long sak, const double* __restrict__ B,
                                                               long ldb, const int* __restrict__ idxB,
                                                               double* __restrict__ Cm, long ldc,
                                                               const int* __restrict__ idxC,
                                                               const int* __restrict__ count) {
  const int ncols = *count;
  if (COLS_LE > 0 && ncols > COLS_LE) return;
  if (COLS_LE < 0 && ncols <= -COLS_LE) return;
  const int q0 = blockIdx.y * TBN;
  if (q0 >= ncols) return;
  const int r0 = blockIdx.x * BM;
  constexpr int kPB = (BK * TBN) / kThreads;  // elements of B per thread and slice
  constexpr int WM = TBN == BN ? 64 : 32, WN = TBN == BN ? 64 : 32;  // rows / columns per wavefront
  constexpr int NI = WM / 16, NJ = WN / 16;
  __shared__ double As[2][AK ? BM * kRowStride : BK * kColStride];
  __shared__ double Bs[2][TBN * kRowStride];
  const int t = threadIdx.x;

  // k-contiguous operands: lane k = t % 16 of rows / columns t / 16 + 16 e.
  // A in general (AK false): 128 consecutive rows of k = t / 128 + 2 e.
  const int lk = t & 15, lr = t >> 4;
  const double* pA[AK ? kPerThread : 1];
  long a_kstride = 0;
  if constexpr (AK) {
#pragma unroll
    for (int e = 0; e < kPerThread; ++e)
      pA[e] = A + static_cast<long>(min(r0 + lr + 16 * e, M - 1)) * sar;
  } else {
    pA[0] = A + static_cast<long>(min(r0 + (t & 127), M - 1)) * sar;
    a_kstride = sak;
  }
  const double* pB[kPB];
#pragma unroll
  for (int e = 0; e < kPB; ++e) {
    const int q = min(q0 + lr + 16 * e, ncols - 1);
    pB[e] = B + (idxB ? static_cast<long>(idxB[q]) : static_cast<long>(q)) * ldb;
  }

  double pa[kPerThread], pb[kPB];
  auto load_slice = [&](int k0) {
    const int kb = k0 + lk;
    const long kc = min(kb, K - 1);
#pragma unroll
    for (int e = 0; e < kPB; ++e) {
      const double v = pB[e][kc];
      pb[e] = kb < K ? v : 0.0;
    }
    if constexpr (AK) {
#pragma unroll
      for (int e = 0; e < kPerThread; ++e) {
        const double v = pA[e][kc];
        pa[e] = kb < K ? v : 0.0;
      }
    } else {
#pragma unroll
      for (int e = 0; e < kPerThread; ++e) {
        const int ka = k0 + (t >> 7) + 2 * e;
        const double v = pA[0][static_cast<long>(min(ka, K - 1)) * a_kstride];
        pa[e] = ka < K ? v : 0.0;
      }
    }
  };
  auto store_slice = [&](int st) {
#pragma unroll
    for (int e = 0; e < kPerThread; ++e) {
      if constexpr (AK)
        As[st][(lr + 16 * e) * kRowStride + lk] = pa[e];
      else
        As[st][((t >> 7) + 2 * e) * kColStride + (t & 127)] = pa[e];
    }
#pragma unroll
    for (int e = 0; e < kPB; ++e) Bs[st][(lr + 16 * e) * kRowStride + lk] = pb[e];
  };

  const int lane = t & 63, w = t >> 6;
  const int wr = TBN == BN ? (w >> 1) * 64 : w * 32, wc = TBN == BN ? (w & 1) * 64 : 0;
  const int fr = lane & 15, fk = lane >> 4;
  f64x4 acc[NI][NJ];
#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  if (K > 0) {  // K == 0 (an empty null space): the outputs are zero
    load_slice(0);
    store_slice(0);
  }
  __syncthreads();
  int st = 0;
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) load_slice(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double av[NI], bv[NJ];
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        if constexpr (AK)
          av[i] = As[st][(wr + 16 * i + fr) * kRowStride + kk + fk];
        else
          av[i] = As[st][(kk + fk) * kColStride + wr + 16 * i + fr];
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) bv[j] = Bs[st][(wc + 16 * j + fr) * kRowStride + kk + fk];
      // raised wave priority while issuing matrix instructions: the other
      // wavefront's loads and LDS traffic fill the gaps (-1.4% per round)
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
      __builtin_amdgcn_s_setprio(0);
    }
    if (more) store_slice(st ^ 1);
    __syncthreads();
    st ^= 1;
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int q = q0 + wc + 16 * j + fr;
    if (q >= ncols) continue;
    double* ccol = Cm + (idxC ? static_cast<long>(idxC[q]) : static_cast<long>(q)) * ldc;
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = r0 + wr + 16 * i + fk + 4 * e;
        if (row < M) ccol[row] = acc[i][j][e];
      }
  }
}

__global__ __launch_bounds__(1024) void compact_kernel(const int* __restrict__ flag, int P, int* __restrict__ idx,
                                                        int* __restrict__ count) {
  __shared__ int wsum[16];
  __shared__ int base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int c0 = 0; c0 < P; c0 += 1024) {
    const int i = c0 + threadIdx.x;
    const int f = (i < P && flag[i]) ? 1 : 0;
    // inclusive scan within the wavefront
    int v = f;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(v, o);
      if (lane >= o) v += y;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    int off = base;
    for (int j = 0; j < w; ++j) off += wsum[j];
    if (f) idx[off + v - 1] = i;
    __syncthreads();
    if (threadIdx.x == 1023) base = off + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = base;
}

__global__ void to_f32_kernel(const double* __restrict__ src, float* __restrict__ dst, size_t count) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[i] = static_cast<float>(src[i]);
}

__global__ __launch_bounds__(256) void normalize_cols_kernel(double* __restrict__ A, int n) {
  double* col = A + static_cast<size_t>(blockIdx.x) * n;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s = fma(col[i], col[i], s);
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  const double nrm = sqrt(red[0]);
  if (nrm > 0.0)
    for (int i = threadIdx.x; i < n; i += 256) col[i] = col[i] / nrm;
}

// Column sums of one stripe of rows (blockIdx.y of gridDim.y equal stripes),
// each in row order; col_means_finish_kernel adds the stripes in stripe order,
// so the mean is the same on every run.
__global__ void col_partial_kernel(const double* __restrict__ X, long rows, int cols, double* __restrict__ part) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  const long per = (rows + gridDim.y - 1) / gridDim.y;
  const long r0 = static_cast<long>(blockIdx.y) * per;
  const long r1 = r0 + per < rows ? r0 + per : rows;
  double s = 0.0;
  for (long r = r0; r < r1; ++r) s += X[r * cols + j];
  part[static_cast<size_t>(blockIdx.y) * cols + j] = s;
}

__global__ void col_means_finish_kernel(const double* __restrict__ part, int stripes, long rows, int cols,
                                        double* __restrict__ mean) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  for (int k = 0; k < stripes; ++k) s += part[static_cast<size_t>(k) * cols + j];
  mean[j] = s / static_cast<double>(rows);
}

__global__ void sub_row_kernel(float* __restrict__ Y, long rows, int cols, const float* __restrict__ v) {
  const size_t n = static_cast<size_t>(rows) * cols;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    Y[i] -= v[i % cols];
}

__global__ void gauss_fill_kernel(float* __restrict__ z, size_t n, uint64_t seed, uint32_t sid) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const u32x4 r = stream_block(seed, static_cast<uint32_t>(i), static_cast<uint32_t>(i >> 32), sid, 0u, kTagDir);
    const double u1 = (static_cast<double>(r.v0) + 1.0) * 0x1.0p-32;
    const double u2 = static_cast<double>(r.v1) * 0x1.0p-32;
    z[i] = static_cast<float>(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  }
}

__global__ void axpby_kernel(float* __restrict__ y, const float* __restrict__ x, float alpha, float beta, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    y[i] = alpha * y[i] + beta * x[i];
}

__global__ void to_f64_kernel(const float* __restrict__ src, double* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[i] = static_cast<double>(src[i]);
}

__global__ __launch_bounds__(256) void sumsq_kernel(const float* __restrict__ x, size_t n, double* __restrict__ part) {
  double s = 0.0;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    s += static_cast<double>(x[i]) * x[i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

unsigned blocks_for(size_t n) {
  size_t b = (n + 255) / 256;
  return static_cast<unsigned>(b > 16384 ? 16384 : (b == 0 ? 1 : b));
}

}  // namespace

hipError_t launch_dgemm_cols(int M, int K, const double* A, long sar, long sak, const double* B, long ldb,
                             double* C, long ldc, const int* count, int max_cols, hipStream_t stream) {
  const unsigned gq = static_cast<unsigned>((max_cols + BN - 1) / BN);
  if (M <= 0 || gq == 0) return hipSuccess;
  if (sak == 1)
    hipLaunchKernelGGL(dgemm_mfma_kernel<true>, dim3((M + BM - 1) / BM, gq), dim3(kThreads), 0, stream, M, K, A, sar,
                       sak, B, ldb, static_cast<const int*>(nullptr), C, ldc, static_cast<const int*>(nullptr), count);
  else
    hipLaunchKernelGGL(dgemm_mfma_kernel<false>, dim3((M + BM - 1) / BM, gq), dim3(kThreads), 0, stream, M, K, A,
                       sar, sak, B, ldb, static_cast<const int*>(nullptr), C, ldc, static_cast<const int*>(nullptr),
                       count);
  return hipGetLastError();
}

hipError_t launch_normalize_cols(double* A, int n, int K, hipStream_t stream) {
  if (K <= 0) return hipSuccess;
  hipLaunchKernelGGL(normalize_cols_kernel, dim3(K), dim3(256), 0, stream, A, n);
  return hipGetLastError();
}

hipError_t launch_col_means(const double* X, long rows, int cols, double* mean, double* work, hipStream_t stream) {
  if (cols <= 0) return hipSuccess;
  const dim3 grid((cols + 255) / 256);
  hipLaunchKernelGGL(col_partial_kernel, dim3(grid.x, kColMeanStripes), dim3(256), 0, stream, X, rows, cols, work);
  hipLaunchKernelGGL(col_means_finish_kernel, grid, dim3(256), 0, stream, work, kColMeanStripes, rows, cols, mean);
  return hipGetLastError();
}

hipError_t launch_sub_row(float* Y, long rows, int cols, const float* v, hipStream_t stream) {
  hipLaunchKernelGGL(sub_row_kernel, dim3(blocks_for(static_cast<size_t>(rows) * cols)), dim3(256), 0, stream, Y,
                     rows, cols, v);
  return hipGetLastError();
}

hipError_t launch_gauss(float* z, size_t count, uint64_t seed, uint32_t sid, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(gauss_fill_kernel, dim3(blocks_for(count)), dim3(256), 0, stream, z, count, seed, sid);
  return hipGetLastError();
}

hipError_t launch_axpby(float* y, const float* x, float alpha, float beta, size_t count, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(axpby_kernel, dim3(blocks_for(count)), dim3(256), 0, stream, y, x, alpha, beta, count);
  return hipGetLastError();
}

hipError_t launch_to_f64(const float* src, double* dst, size_t count, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(to_f64_kernel, dim3(blocks_for(count)), dim3(256), 0, stream, src, dst, count);
  return hipGetLastError();
}

hipError_t sum_squares(const float* x, size_t count, double* out) {
  constexpr unsigned kBlocks = 1024;
  double* part = nullptr;
  hipError_t e = hipMalloc(&part, sizeof(double) * kBlocks);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sumsq_kernel, dim3(kBlocks), dim3(256), 0, nullptr, x, count, part);
  e = hipGetLastError();
  double h[kBlocks];
  if (e == hipSuccess) e = hipMemcpy(h, part, sizeof h, hipMemcpyDeviceToHost);
  (void)hipFree(part);
  if (e != hipSuccess) return e;
  double s = 0.0;
  for (unsigned i = 0; i < kBlocks; ++i) s += h[i];
  *out = s;
  return hipSuccess;
}

hipError_t launch_to_f32(const double* src, float* dst, size_t count, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((count + 255) / 256, 65536);
  hipLaunchKernelGGL(to_f32_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, src, dst, count);
  return hipGetLastError();
}

namespace {
// S+ (n x m column major, FP32) = S^T G with G = (S S^T)+ symmetric (m x m):
// S+[i + k*n] = sum over the entries (r, v) of column i of S of v * G[r + k*m],
// accumulated in FP64 in the column's row order, rounded once. Threads run
// over i, so stores are coalesced; the reads of G gather within one column.
// One thread per row r of S adds the row's entries into column r of S^T.
__global__ __launch_bounds__(256) void dense_st_kernel(int m, int n, const int* __restrict__ rowptr,
                                                       const int* __restrict__ colidx, const double* __restrict__ vals,
                                                       double* __restrict__ B) {
  const int r = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  if (r >= m) return;
  double* col = B + static_cast<size_t>(r) * n;
  for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) col[colidx[p]] += vals[p];
}

__global__ __launch_bounds__(256) void pinv_csc_kernel(int n, int m, const int* __restrict__ colptr,
                                                       const int* __restrict__ rowidx, const double* __restrict__ vals,
                                                       const double* __restrict__ G, float* __restrict__ out) {
  const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
  const int k = static_cast<int>(blockIdx.y);
  if (i >= n) return;
  const double* Gk = G + static_cast<size_t>(k) * m;
  double acc = 0.0;
  for (int p = colptr[i]; p < colptr[i + 1]; ++p) acc = fma(vals[p], Gk[rowidx[p]], acc);
  out[i + static_cast<size_t>(k) * n] = static_cast<float>(acc);
}
}  // namespace

namespace {
// Y[i + k*n] = X[i + k*n] - v[i] over the columns of an n x cols matrix.
__global__ __launch_bounds__(256) void sub_cols_kernel(const double* __restrict__ X, const double* __restrict__ v,
                                                       double* __restrict__ Y, int n, int cols) {
  const size_t total = static_cast<size_t>(n) * cols;
  for (size_t q = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; q < total;
       q += static_cast<size_t>(gridDim.x) * 256) {
    const int i = static_cast<int>(q % static_cast<size_t>(n));
    Y[q] = X[q] - v[i];
  }
}
// X[i + k*n] += v[i], then X[i + k*n] = lb[i] where lb[i] == ub[i].
__global__ __launch_bounds__(256) void add_cols_fix_kernel(double* __restrict__ X, const double* __restrict__ v,
                                                           const double* __restrict__ lb,
                                                           const double* __restrict__ ub, int n, int cols) {
  const size_t total = static_cast<size_t>(n) * cols;
  for (size_t q = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; q < total;
       q += static_cast<size_t>(gridDim.x) * 256) {
    const int i = static_cast<int>(q % static_cast<size_t>(n));
    X[q] = lb[i] == ub[i] ? lb[i] : X[q] + v[i];
  }
}
}  // namespace

hipError_t launch_sub_cols(const double* X, const double* v, double* Y, int n, int cols, hipStream_t stream) {
  const size_t total = static_cast<size_t>(n) * cols;
  if (total == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL(sub_cols_kernel, dim3(blocks), dim3(256), 0, stream, X, v, Y, n, cols);
  return hipGetLastError();
}

hipError_t launch_add_cols_fix(double* X, const double* v, const double* lb, const double* ub, int n, int cols,
                               hipStream_t stream) {
  const size_t total = static_cast<size_t>(n) * cols;
  if (total == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL(add_cols_fix_kernel, dim3(blocks), dim3(256), 0, stream, X, v, lb, ub, n, cols);
  return hipGetLastError();
}

hipError_t launch_dense_st(int m, int n, const int* rowptr, const int* colidx, const double* vals, double* B,
                           hipStream_t stream) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(dense_st_kernel, dim3(static_cast<unsigned>((m + 255) / 256)), dim3(256), 0, stream, m, n, rowptr,
                     colidx, vals, B);
  return hipGetLastError();
}

hipError_t launch_pinv_csc(int n, int m, const int* colptr, const int* rowidx, const double* vals, const double* G,
                           float* out, hipStream_t stream) {
  if (n <= 0 || m <= 0) return hipSuccess;
  hipLaunchKernelGGL(pinv_csc_kernel, dim3(static_cast<unsigned>((n + 255) / 256), static_cast<unsigned>(m)),
                     dim3(256), 0, stream, n, m, colptr, rowidx, vals, G, out);
  return hipGetLastError();
}

hipError_t launch_compact(const int* flag, int P, int* idx, int* count, hipStream_t stream) {
  hipLaunchKernelGGL(compact_kernel, dim3(1), dim3(1024), 0, stream, flag, P, idx, count);
  return hipGetLastError();
}

namespace {
// Four 16-byte loads in flight per thread and iteration, consecutive threads
// on consecutive words: each wavefront reads whole 1 KiB runs.
__global__ __launch_bounds__(256) void read_stream_kernel(const double2* __restrict__ buf, size_t words,
                                                         double* __restrict__ sums) {
  __shared__ double part[4];
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  double acc = 0.0;
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < words; i += 4 * stride) {
    const double2 a = buf[i], b = buf[i + stride], c = buf[i + 2 * stride], d = buf[i + 3 * stride];
    acc += (a.x + a.y) + (b.x + b.y) + (c.x + c.y) + (d.x + d.y);
  }
  for (; i < words; i += stride) acc += buf[i].x + buf[i].y;
  for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sums[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}
}  // namespace

hipError_t read_stream_blocks(int* blocks) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&read_stream_kernel), 256, 0);
  if (e == hipSuccess) *blocks = std::max(cus, 1) * std::max(per_cu, 1);
  return e;
}

hipError_t launch_read_stream(const double* buf, size_t words, double* sums, int blocks, hipStream_t stream) {
  hipLaunchKernelGGL(read_stream_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream,
                     reinterpret_cast<const double2*>(buf), words, sums);
  return hipGetLastError();
}

hipError_t launch_project(const double* Ncm, int n, int d, double* X, double* T, const int* idx,
                          const int* count, int P, hipStream_t stream) {
  const unsigned gq = static_cast<unsigned>((P + BN - 1) / BN);
  if (gq == 0 || n <= 0) return hipSuccess;
  // rounds of at most kNarrowCols flagged chains take 128 x 32 tiles, larger
  // ones 128 x 128; the kernel of the other width returns at once
  const bool narrow = P > kNarrowCols / 4;  // with fewer chains the wide tiles fill the device anyway
  const unsigned gn = static_cast<unsigned>((std::min(P, kNarrowCols) + 31) / 32);
  // T(:, q) = N^T X(:, idx[q]):   A(j, i) = N[i + j*n]  (sar = n, sak = 1); nothing when d == 0
  if (d > 0) {
    if (narrow) {
      hipLaunchKernelGGL((dgemm_mfma_kernel<true, 32, kNarrowCols>), dim3((d + BM - 1) / BM, gn), dim3(kThreads), 0,
                         stream, d, n, Ncm, static_cast<long>(n), 1L, X, static_cast<long>(n), idx, T,
                         static_cast<long>(d), static_cast<const int*>(nullptr), count);
      hipLaunchKernelGGL((dgemm_mfma_kernel<true, BN, -kNarrowCols>), dim3((d + BM - 1) / BM, gq), dim3(kThreads), 0,
                         stream, d, n, Ncm, static_cast<long>(n), 1L, X, static_cast<long>(n), idx, T,
                         static_cast<long>(d), static_cast<const int*>(nullptr), count);
    } else {
      hipLaunchKernelGGL(dgemm_mfma_kernel<true>, dim3((d + BM - 1) / BM, gq), dim3(kThreads), 0, stream, d, n, Ncm,
                         static_cast<long>(n), 1L, X, static_cast<long>(n), idx, T, static_cast<long>(d),
                         static_cast<const int*>(nullptr), count);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  // X(:, idx[q]) = N T(:, q):     A(i, j) = N[i + j*n]  (sar = 1, sak = n)
  if (narrow) {
    hipLaunchKernelGGL((dgemm_mfma_kernel<false, 32, kNarrowCols>), dim3((n + BM - 1) / BM, gn), dim3(kThreads), 0,
                       stream, n, d, Ncm, 1L, static_cast<long>(n), T, static_cast<long>(d),
                       static_cast<const int*>(nullptr), X, static_cast<long>(n), idx, count);
    hipLaunchKernelGGL((dgemm_mfma_kernel<false, BN, -kNarrowCols>), dim3((n + BM - 1) / BM, gq), dim3(kThreads), 0,
                       stream, n, d, Ncm, 1L, static_cast<long>(n), T, static_cast<long>(d),
                       static_cast<const int*>(nullptr), X, static_cast<long>(n), idx, count);
  } else {
    hipLaunchKernelGGL(dgemm_mfma_kernel<false>, dim3((n + BM - 1) / BM, gq), dim3(kThreads), 0, stream, n, d, Ncm,
                       1L, static_cast<long>(n), T, static_cast<long>(d), static_cast<const int*>(nullptr), X,
                       static_cast<long>(n), idx, count);
  }
  return hipGetLastError();
}

}  // namespace achr
