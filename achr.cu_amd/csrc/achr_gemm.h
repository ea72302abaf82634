// Matrix-core contractions shared by the projection rounds and the
// direction network (see achr_gemm.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace achr {

enum GemmPrec : int { kPrecF32 = 0, kPrecBF16x3 = 1, kPrecBF16 = 2 };
enum GemmEpi : int { kEpiStoreF32 = 0, kEpiStoreF64 = 1, kEpiSubF64 = 2 };

// C(i, j) = alpha * sum_k A(i, k) B(k, j) [+ beta * C] (FP32 accumulation).
//   A(i, k) = A[rowA(i) * sa_m + k * sa_k], rowA(i) = idxA ? idxA[i] : i;
//             A is FP32, or FP64 when a_f64 (converted on load);
//   B(k, j) = B[k * sb_k + j * sb_n];
//   C(i, j) at C[rowC(i) * sc_m + j * sc_n], rowC(i) = idxC ? idxC[i] : i;
//   epi: store FP32 (alpha, beta), store FP64, or subtract from FP64.
// One of (sa_m, sa_k) and one of (sb_k, sb_n) must be 1; gathered A rows
// must be contiguous in k. M is *Mdev when Mdev is set (device-side count),
// in which case max_rows bounds the launch.
struct GemmArgs {
  int M, N, K;
  const int* Mdev;
  const void* A;
  int a_f64;
  long sa_m, sa_k;
  const int* idxA;
  const float* B;
  long sb_k, sb_n;
  void* C;
  long sc_m, sc_n;
  const int* idxC;
  float alpha, beta;
  int epi;
};

hipError_t launch_gemm(const GemmArgs& g, int prec, int max_rows, hipStream_t s);

// The projection rounds' residual form on packed operands:
//   C[rowC(i), j] -= sum_k A[rowA(i) * lda + k] B(k, j)  (i < M, j < N)
// with FP32 accumulation (kPrecF32 or kPrecBF16x3). A is FP32, k contiguous,
// lda a multiple of 4 and columns K..Kp-1 zero (Kp = packed_k(K));
// B is packed once by launch_pack_b for the same precision; C is FP64.
// rowA(i) = idxA ? idxA[i] : i, likewise rowC. M is *Mdev when Mdev is set,
// with max_rows bounding the launch.
struct GemmSubArgs {
  int M, N, Kp;
  const int* Mdev;
  const float* A;
  long lda;
  const int* idxA;
  const void* B;
  double* C;
  long ldc;
  const int* idxC;
};
// The reduction length gemm_sub runs over: K rounded up to its slab depth.
int packed_k(int K);
// Bytes of B packed for a K x N operand (either precision).
size_t packed_b_bytes(int K, int N);
// Packs B(k, j) = src[k * sk + j * sj] (K x N, FP32) for gemm_sub: 256-column
// tiles of slabs, [tile][slab][column][k], zero beyond K and N; FP32 values,
// or per slab the BF16 high parts followed by the BF16 low parts.
hipError_t launch_pack_b(const float* src, long sk, long sj, int K, int N, int prec, void* dst, hipStream_t s);
hipError_t launch_gemm_sub(const GemmSubArgs& g, int prec, int max_rows, hipStream_t s);

}  // namespace achr
