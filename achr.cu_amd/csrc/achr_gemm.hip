// Dense contractions on the gfx950 matrix cores.
//
// C(i, j) = alpha * sum_k A(i, k) B(k, j) (+ beta C) with FP32 accumulation,
// in three precisions:
//   kPrecF32   v_mfma_f32_32x32x2_f32 on FP32 operands;
//   kPrecBF16x3 each FP32 operand split as hi = bf16(x), lo = bf16(x - hi),
//               acc += lo*hi' + hi*lo' + hi*hi' with v_mfma_f32_32x32x16_bf16
//               (about FP32 accuracy at three BF16 passes);
//   kPrecBF16  one BF16 pass (inference of the direction network).
// Operands are strided views (either dimension may be the contiguous one),
// A's rows may be gathered through an index list whose length is read on
// the device, and the epilogue writes FP32, or stores / subtracts into an
// FP64 array (the chains' points), so a projection round needs no host
// round trip. Tiles: 128 x 128 x 32 per 256-thread workgroup, each of the 4
// wavefronts owning a 64 x 64 block of 2 x 2 MFMA 32 x 32 tiles.
#include "achr_gemm.h"

namespace achr {
namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int kThreads = 256;
constexpr int PADF = 1;   // FP32 LDS row padding (floats)
constexpr int PADB = 8;   // BF16 LDS row padding (bf16 elements, keeps 16-byte alignment)

__device__ __forceinline__ unsigned short bf16_bits_rne(float x) {
  const unsigned u = __builtin_bit_cast(unsigned, x);
  const unsigned r = u + 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(r >> 16);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
  return __builtin_bit_cast(float, static_cast<unsigned>(h) << 16);
}

// Loads the 128 x 32 tile of A (rows i0.., k0..) or of B^T (columns j0..,
// k0..) as FP32 values vals[16] for this thread; position (r, k) of element
// e is returned through rr/kk. `kcontig` picks the coalesced layout. All 16
// loads are issued unconditionally, so they are in flight together: a row
// past nrows reads row nrows - 1 and k past K reads K - 1, and those values
// are then multiplied by 0 (operands are finite). A select instead of the
// product lets the compiler skip the load under a branch and wait for each
// load before the next. Needs nrows >= 1 and K >= 1.
template <typename T>
__device__ __forceinline__ void load_tile(const T* base, long s_row, long s_k, bool kcontig, const int* idx,
                                          int row0, int nrows, int k0, int K, float (&vals)[16], int& r_first,
                                          int& k_first, bool& along_k) {
  const int t = threadIdx.x;
  if (kcontig) {
    // thread -> one row, 16 consecutive k
    const int r = t >> 1;
    const int kb = (t & 1) * 16;
    r_first = r;
    k_first = kb;
    along_k = true;
    const int gr = row0 + r;
    const int grc = min(gr, nrows - 1);
    const long roff = static_cast<long>(idx ? idx[grc] : grc) * s_row;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int gk = k0 + kb + e;
      vals[e] = static_cast<float>(base[roff + min(gk, K - 1)]) * ((gr < nrows && gk < K) ? 1.0f : 0.0f);
    }
  } else {
    // thread -> one k, 16 consecutive rows (rows contiguous, no gather)
    const int kk = t >> 3;
    const int rb = (t & 7) * 16;
    r_first = rb;
    k_first = kk;
    along_k = false;
    const int gk = k0 + kk;
    const long koff = static_cast<long>(min(gk, K - 1)) * s_k;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int gr = row0 + rb + e;
      vals[e] = static_cast<float>(base[static_cast<long>(min(gr, nrows - 1)) + koff]) *
                ((gr < nrows && gk < K) ? 1.0f : 0.0f);
    }
  }
}

template <int PREC, typename TA>
__global__ __launch_bounds__(kThreads) void gemm_kernel(GemmArgs g) {
  const int M = g.Mdev ? *g.Mdev : g.M;
  const int i0 = blockIdx.y * BM;
  const int j0 = blockIdx.x * BN;
  if (i0 >= M) return;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int wr = w >> 1, wc = w & 1;

  constexpr bool kBF = PREC != kPrecF32;
  constexpr int LDW = kBF ? (BK + PADB) : (BK + PADF);
  __shared__ __attribute__((aligned(16))) float sf[kBF ? 1 : 2 * BM * (BK + PADF)];
  __shared__ __attribute__((aligned(16))) unsigned short sh[kBF ? 4 * BM * (BK + PADB) : 1];
  float* As = sf;
  float* Bs = sf + BM * LDW;
  unsigned short* Ah = sh;
  unsigned short* Al = sh + BM * LDW;
  unsigned short* Bh = sh + 2 * BM * LDW;
  unsigned short* Bl = sh + 3 * BM * LDW;

  v16f acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;

  const TA* A = static_cast<const TA*>(g.A);
  const bool a_kc = g.sa_k == 1;
  const bool b_kc = g.sb_k == 1;
  // BF16 paths load the next K-slice into registers while the matrix cores
  // work on the current one (issued right after the LDS stores: -4% per
  // residual round); the FP32 path loads each slice at the top of its
  // iteration (the early load made it 4% slower).
  float va[16], vb[16];
  int ra, ka, rb, kb;
  bool ala, alb;
  if (g.K > 0) {
    load_tile<TA>(A, g.sa_m, g.sa_k, a_kc, g.idxA, i0, M, 0, g.K, va, ra, ka, ala);
    load_tile<float>(g.B, g.sb_n, g.sb_k, b_kc, nullptr, j0, g.N, 0, g.K, vb, rb, kb, alb);
  }
  for (int k0 = 0; k0 < g.K; k0 += BK) {
    if constexpr (!kBF) {
      if (k0 > 0) {
        load_tile<TA>(A, g.sa_m, g.sa_k, a_kc, g.idxA, i0, M, k0, g.K, va, ra, ka, ala);
        load_tile<float>(g.B, g.sb_n, g.sb_k, b_kc, nullptr, j0, g.N, k0, g.K, vb, rb, kb, alb);
      }
    }
    __syncthreads();  // previous tile fully consumed
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int ar = ala ? ra : ra + e, ak = ala ? ka + e : ka;
      const int br = alb ? rb : rb + e, bk = alb ? kb + e : kb;
      if constexpr (kBF) {
        const unsigned short ah = bf16_bits_rne(va[e]);
        const unsigned short bh = bf16_bits_rne(vb[e]);
        Ah[ar * LDW + ak] = ah;
        Bh[br * LDW + bk] = bh;
        if constexpr (PREC == kPrecBF16x3) {
          Al[ar * LDW + ak] = bf16_bits_rne(va[e] - bf16_to_f32(ah));
          Bl[br * LDW + bk] = bf16_bits_rne(vb[e] - bf16_to_f32(bh));
        }
      } else {
        As[ar * LDW + ak] = va[e];
        Bs[br * LDW + bk] = vb[e];
      }
    }
    __syncthreads();
    if constexpr (kBF) {
      if (k0 + BK < g.K) {
        load_tile<TA>(A, g.sa_m, g.sa_k, a_kc, g.idxA, i0, M, k0 + BK, g.K, va, ra, ka, ala);
        load_tile<float>(g.B, g.sb_n, g.sb_k, b_kc, nullptr, j0, g.N, k0 + BK, g.K, vb, rb, kb, alb);
      }
    }
    if constexpr (kBF) {
#pragma unroll
      for (int ks = 0; ks < BK; ks += 16) {
        const int kof = ks + 8 * (lane >> 5);
        v8bf ahf[2], alf[2], bhf[2], blf[2];
#pragma unroll
        for (int t2 = 0; t2 < 2; ++t2) {
          const int ar = wr * 64 + t2 * 32 + (lane & 31);
          const int br = wc * 64 + t2 * 32 + (lane & 31);
          ahf[t2] = *reinterpret_cast<const v8bf*>(Ah + ar * LDW + kof);
          bhf[t2] = *reinterpret_cast<const v8bf*>(Bh + br * LDW + kof);
          if constexpr (PREC == kPrecBF16x3) {
            alf[t2] = *reinterpret_cast<const v8bf*>(Al + ar * LDW + kof);
            blf[t2] = *reinterpret_cast<const v8bf*>(Bl + br * LDW + kof);
          }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            if constexpr (PREC == kPrecBF16x3) {
              acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alf[a], bhf[b], acc[a][b], 0, 0, 0);
              acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahf[a], blf[b], acc[a][b], 0, 0, 0);
            }
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahf[a], bhf[b], acc[a][b], 0, 0, 0);
          }
      }
    } else {
#pragma unroll 4
      for (int ks = 0; ks < BK; ks += 2) {
        const int kof = ks + (lane >> 5);
        float af[2], bf[2];
#pragma unroll
        for (int t2 = 0; t2 < 2; ++t2) {
          af[t2] = As[(wr * 64 + t2 * 32 + (lane & 31)) * LDW + kof];
          bf[t2] = Bs[(wc * 64 + t2 * 32 + (lane & 31)) * LDW + kof];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[a], bf[b], acc[a][b], 0, 0, 0);
      }
    }
  }

  // C layout of a 32 x 32 MFMA tile: column lane % 32, rows 8*(r/4) + 4*(lane/32) + r%4.
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = j0 + wc * 64 + b * 32 + (lane & 31);
      if (j >= g.N) continue;
      // Four rows at a time: addresses (rows past M take row M - 1's, never
      // stored), then the reads of C before any write, so they are in flight
      // together without holding a whole column's worth of registers.
#pragma unroll
      for (int r0 = 0; r0 < 16; r0 += 4) {
        long ci[4];
        bool ok[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = r0 + q;
          const int i = i0 + wr * 64 + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          ok[q] = i < M;
          const int ic = min(i, M - 1);
          ci[q] = static_cast<long>(g.idxC ? g.idxC[ic] : ic) * g.sc_m + static_cast<long>(j) * g.sc_n;
        }
        if (g.epi == kEpiStoreF32) {
          float* C = static_cast<float*>(g.C);
          float old[4] = {0.0f, 0.0f, 0.0f, 0.0f};
          if (g.beta != 0.0f) {
#pragma unroll
            for (int q = 0; q < 4; ++q) old[q] = C[ci[q]];
          }
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (ok[q]) C[ci[q]] = g.beta != 0.0f ? g.alpha * acc[a][b][r0 + q] + g.beta * old[q] : g.alpha * acc[a][b][r0 + q];
        } else if (g.epi == kEpiStoreF64) {
          double* C = static_cast<double*>(g.C);
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (ok[q]) C[ci[q]] = static_cast<double>(g.alpha * acc[a][b][r0 + q]);
        } else {
          double* C = static_cast<double*>(g.C);
          double old[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) old[q] = C[ci[q]];
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (ok[q]) C[ci[q]] = old[q] - static_cast<double>(g.alpha * acc[a][b][r0 + q]);
        }
      }
    }
}

template <int PREC>
hipError_t launch_prec(const GemmArgs& g, int max_rows, hipStream_t s) {
  const dim3 grid((g.N + BN - 1) / BN, (max_rows + BM - 1) / BM);
  if (grid.x == 0 || grid.y == 0) return hipSuccess;
  if (g.a_f64)
    hipLaunchKernelGGL((gemm_kernel<PREC, double>), grid, dim3(kThreads), 0, s, g);
  else
    hipLaunchKernelGGL((gemm_kernel<PREC, float>), grid, dim3(kThreads), 0, s, g);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_gemm(const GemmArgs& g, int prec, int max_rows, hipStream_t s) {
  if (g.sa_m != 1 && g.sa_k != 1) return hipErrorInvalidValue;
  if (g.sb_k != 1 && g.sb_n != 1) return hipErrorInvalidValue;
  if (g.idxA && g.sa_k != 1) return hipErrorInvalidValue;  // gathered rows must be k-contiguous
  switch (prec) {
    case kPrecF32: return launch_prec<kPrecF32>(g, max_rows, s);
    case kPrecBF16x3: return launch_prec<kPrecBF16x3>(g, max_rows, s);
    case kPrecBF16: return launch_prec<kPrecBF16>(g, max_rows, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace achr
