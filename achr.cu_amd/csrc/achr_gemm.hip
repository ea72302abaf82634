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

#include <algorithm>

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

// ---------------------------------------------------------------------------
// gemm_sub: the residual form of the projection rounds, C -= A B on packed B.
// 256 x 256 tiles per 512-thread workgroup: 8 wavefronts as 2 x 4, each
// owning a 128 x 64 block of 4 x 2 MFMA 32 x 32 tiles (128 accumulator
// registers). k advances in 16-deep slabs through two LDS buffers: the
// global loads of slab s + 1 are in flight while the matrix cores work on
// slab s, with one barrier per slab. B arrives packed (one contiguous 16 KiB
// block per column tile and slab, already split into BF16 parts for the
// split path); A's gathered rows are read 16 bytes per thread and split on
// the way into LDS. Workgroups are renumbered so that each XCD (dispatch is
// round-robin over eight) runs one contiguous range of tiles, ordered so
// that kGroupM row tiles sweep the column tiles together: the 32 workgroups
// an XCD holds at once read 4 A slabs and 8 B slabs, shared in its L2.
constexpr int GT = 256;  // tile rows and columns
constexpr int GK = 16;   // slab depth
constexpr int kSubThreads = 512;
constexpr int kGroupM = 4;
constexpr int kXcd = 8;
constexpr int PF = GK + 4;  // FP32 LDS row pitch (floats): b128 reads of 8 and 16 lanes hit distinct banks
constexpr int PB = GK + 8;  // BF16 LDS row pitch (bf16 elements), likewise
constexpr int kSlabChunks = GT * GK * 4 / 16;  // 16-byte chunks per packed slab block (1,024)

typedef float v4f __attribute__((ext_vector_type(4)));

template <int PREC>
constexpr size_t sub_lds_bytes() {
  return PREC == kPrecF32 ? size_t(2) * 2 * GT * PF * sizeof(float) : size_t(2) * 4 * GT * PB * sizeof(unsigned short);
}

// hi = bf16(x), lo = bf16(x - hi) for four values, as packed bf16 pairs.
__device__ __forceinline__ void split4(const v4f& x, uint2& hi, uint2& lo) {
  unsigned short h[4], l[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    h[e] = bf16_bits_rne(x[e]);
    l[e] = bf16_bits_rne(x[e] - bf16_to_f32(h[e]));
  }
  hi.x = h[0] | (static_cast<unsigned>(h[1]) << 16);
  hi.y = h[2] | (static_cast<unsigned>(h[3]) << 16);
  lo.x = l[0] | (static_cast<unsigned>(l[1]) << 16);
  lo.y = l[2] | (static_cast<unsigned>(l[3]) << 16);
}

template <int PREC>
__global__ __launch_bounds__(kSubThreads) void gemm_sub_kernel(GemmSubArgs g, int mt, int nt) {
  extern __shared__ __attribute__((aligned(16))) unsigned char glds[];
  const int M = g.Mdev ? *g.Mdev : g.M;
  // XCD-contiguous numbering, then kGroupM row tiles per column sweep
  const int nwg = gridDim.x, id = blockIdx.x;
  const int xcd = id % kXcd, kx = id / kXcd, q = nwg / kXcd, rem = nwg % kXcd;
  const int L = xcd < rem ? xcd * (q + 1) + kx : rem * (q + 1) + (xcd - rem) * q + kx;
  const int grp = L / (kGroupM * nt);
  const int first = grp * kGroupM;
  const int gsz = min(kGroupM, mt - first);
  const int within = L - grp * kGroupM * nt;
  const int tm = first + within % gsz, tn = within / gsz;
  const int i0 = tm * GT, j0 = tn * GT;
  if (i0 >= M) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 2, wn = w & 3;
  const int lr = lane & 31, h = lane >> 5;
  const int nslab = g.Kp / GK;

  // This thread's slab loads: rows (t / 4) and (t / 4) + 128 of A at k
  // 4 (t % 4) .. + 3, and 16-byte chunks t and t + 512 of the B block.
  const int ra = t >> 2, k4 = (t & 3) * 4;
  const float* pa[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int i = min(i0 + ra + u * 128, M - 1);
    pa[u] = g.A + static_cast<long>(g.idxA ? g.idxA[i] : i) * g.lda + k4;
  }
  const v4f* pb = static_cast<const v4f*>(g.B) + static_cast<size_t>(tn) * nslab * kSlabChunks + t;

  v4f va[2], vb[2];
  auto gload = [&](int s) {
    va[0] = *reinterpret_cast<const v4f*>(pa[0] + s * GK);
    va[1] = *reinterpret_cast<const v4f*>(pa[1] + s * GK);
    vb[0] = pb[static_cast<size_t>(s) * kSlabChunks];
    vb[1] = pb[static_cast<size_t>(s) * kSlabChunks + 512];
  };
  auto lstore = [&](int buf) {
    if constexpr (PREC == kPrecF32) {
      float* As = reinterpret_cast<float*>(glds) + buf * 2 * GT * PF;
      float* Bs = As + GT * PF;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        // FP32 B chunk c holds column c / 4, k 4 (c % 4) .. + 3: the same
        // (row, k) positions as this thread's A chunks
        *reinterpret_cast<v4f*>(As + (ra + u * 128) * PF + k4) = va[u];
        *reinterpret_cast<v4f*>(Bs + (ra + u * 128) * PF + k4) = vb[u];
      }
    } else {
      unsigned short* Ah = reinterpret_cast<unsigned short*>(glds) + buf * 4 * GT * PB;
      unsigned short* Al = Ah + GT * PB;
      unsigned short* Bh = Al + GT * PB;
      unsigned short* Bl = Bh + GT * PB;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        uint2 hi, lo;
        split4(va[u], hi, lo);
        *reinterpret_cast<uint2*>(Ah + (ra + u * 128) * PB + k4) = hi;
        *reinterpret_cast<uint2*>(Al + (ra + u * 128) * PB + k4) = lo;
      }
      // BF16 block: chunks 0..511 the high parts, 512..1023 the low parts;
      // chunk c of either holds column c / 2, k 8 (c % 2) .. + 7
      const int bc = t >> 1, bk = (t & 1) * 8;
      *reinterpret_cast<v4f*>(Bh + bc * PB + bk) = vb[0];
      *reinterpret_cast<v4f*>(Bl + bc * PB + bk) = vb[1];
    }
  };

  v16f acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0.0f;

  auto compute = [&](int buf) {
    if constexpr (PREC == kPrecF32) {
      // lane (row lr, half h) of MFMA step st takes k = 8 h + st of the slab;
      // both operands use that assignment, so the 8 steps cover k 0..15
      const float* As = reinterpret_cast<const float*>(glds) + buf * 2 * GT * PF;
      const float* Bs = As + GT * PF;
#pragma unroll
      for (int qq = 0; qq < 2; ++qq) {
        v4f a[4], b[2];
#pragma unroll
        for (int x = 0; x < 4; ++x) a[x] = *reinterpret_cast<const v4f*>(As + (wm * 128 + x * 32 + lr) * PF + h * 8 + qq * 4);
#pragma unroll
        for (int y = 0; y < 2; ++y) b[y] = *reinterpret_cast<const v4f*>(Bs + (wn * 64 + y * 32 + lr) * PF + h * 8 + qq * 4);
#pragma unroll
        for (int st = 0; st < 4; ++st)
#pragma unroll
          for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
              acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[x][st], b[y][st], acc[x][y], 0, 0, 0);
      }
    } else {
      const unsigned short* Ah = reinterpret_cast<const unsigned short*>(glds) + buf * 4 * GT * PB;
      const unsigned short* Al = Ah + GT * PB;
      const unsigned short* Bh = Al + GT * PB;
      const unsigned short* Bl = Bh + GT * PB;
      v8bf ah[4], al[4], bh[2], bl[2];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const int o = (wm * 128 + x * 32 + lr) * PB + h * 8;
        ah[x] = *reinterpret_cast<const v8bf*>(Ah + o);
        al[x] = *reinterpret_cast<const v8bf*>(Al + o);
      }
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        const int o = (wn * 64 + y * 32 + lr) * PB + h * 8;
        bh[y] = *reinterpret_cast<const v8bf*>(Bh + o);
        bl[y] = *reinterpret_cast<const v8bf*>(Bl + o);
      }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
        }
    }
  };

  gload(0);
  lstore(0);
  __syncthreads();
  for (int s = 0; s < nslab; ++s) {
    const int cur = s & 1;
    if (s + 1 < nslab) gload(s + 1);
    compute(cur);
    // the other buffer was last read in iteration s - 1, before its barrier
    if (s + 1 < nslab) lstore(cur ^ 1);
    __syncthreads();
  }

  // C -= acc; MFMA tile layout: column lane % 32, row 8 (r / 4) + 4 (lane / 32) + r % 4
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int j = j0 + wn * 64 + y * 32 + lr;
      if (j >= g.N) continue;
#pragma unroll
      for (int r0 = 0; r0 < 16; r0 += 4) {
        long ci[4];
        bool ok[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = r0 + e;
          const int i = i0 + wm * 128 + x * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
          ok[e] = i < M;
          const int ic = min(i, M - 1);
          ci[e] = static_cast<long>(g.idxC ? g.idxC[ic] : ic) * g.ldc + j;
        }
        double old[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) old[e] = g.C[ci[e]];
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (ok[e]) g.C[ci[e]] = old[e] - static_cast<double>(acc[x][y][r0 + e]);
      }
    }
}

template <int PREC>
__global__ __launch_bounds__(256) void pack_b_kernel(const float* __restrict__ src, long sk, long sj, int K, int N,
                                                    int nslab, size_t total, void* __restrict__ dst) {
  for (size_t e = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * 256) {
    const size_t blk = e >> 12;           // (tile, slab) block of 256 x 16 values
    const int in = static_cast<int>(e & 4095);
    const int col = in >> 4, kk = in & 15;
    const size_t tn = blk / nslab;
    const int s = static_cast<int>(blk - tn * nslab);
    const int k = s * GK + kk;
    const long j = static_cast<long>(tn) * GT + col;
    const float v = (k < K && j < N) ? src[k * sk + j * sj] : 0.0f;
    if constexpr (PREC == kPrecF32) {
      static_cast<float*>(dst)[e] = v;
    } else {
      const unsigned short hi = bf16_bits_rne(v);
      unsigned short* d = static_cast<unsigned short*>(dst) + blk * 8192 + in;
      d[0] = hi;
      d[4096] = bf16_bits_rne(v - bf16_to_f32(hi));
    }
  }
}

template <int PREC>
hipError_t launch_sub(const GemmSubArgs& g, int max_rows, hipStream_t s) {
  const int mt = (max_rows + GT - 1) / GT, nt = (g.N + GT - 1) / GT;
  if (mt <= 0 || nt <= 0 || g.Kp <= 0) return hipSuccess;
  const size_t lds = sub_lds_bytes<PREC>();
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&gemm_sub_kernel<PREC>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((gemm_sub_kernel<PREC>), dim3(mt * nt), dim3(kSubThreads), lds, s, g, mt, nt);
  return hipGetLastError();
}

}  // namespace

int packed_k(int K) { return (K + GK - 1) / GK * GK; }

size_t packed_b_bytes(int K, int N) {
  const size_t nt = (static_cast<size_t>(N) + GT - 1) / GT, nslab = (static_cast<size_t>(K) + GK - 1) / GK;
  return nt * nslab * GT * GK * 4;
}

hipError_t launch_pack_b(const float* src, long sk, long sj, int K, int N, int prec, void* dst, hipStream_t s) {
  if (K <= 0 || N <= 0) return hipSuccess;
  const int nslab = (K + GK - 1) / GK;
  const size_t total = packed_b_bytes(K, N) / 4;
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>((total + 255) / 256, 65536));
  if (prec == kPrecF32)
    hipLaunchKernelGGL((pack_b_kernel<kPrecF32>), dim3(blocks), dim3(256), 0, s, src, sk, sj, K, N, nslab, total, dst);
  else if (prec == kPrecBF16x3)
    hipLaunchKernelGGL((pack_b_kernel<kPrecBF16x3>), dim3(blocks), dim3(256), 0, s, src, sk, sj, K, N, nslab, total,
                       dst);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_gemm_sub(const GemmSubArgs& g, int prec, int max_rows, hipStream_t s) {
  if (g.Kp % GK != 0 || g.lda % 4 != 0 || g.lda < g.Kp || g.N <= 0) return hipErrorInvalidValue;
  if (prec == kPrecF32) return launch_sub<kPrecF32>(g, max_rows, s);
  if (prec == kPrecBF16x3) return launch_sub<kPrecBF16x3>(g, max_rows, s);
  return hipErrorInvalidValue;
}

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
