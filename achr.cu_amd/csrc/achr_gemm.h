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

}  // namespace achr
