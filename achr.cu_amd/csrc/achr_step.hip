// FP64 ACHR chain-step kernel for gfx950.
//
// One chain is owned by a group of TPC threads (TPC = 16..64: a slice of one
// wavefront, several chains per workgroup; TPC = 128..1024: one workgroup per
// chain). Element i of the n-vector belongs to thread i % TPC, slot i / TPC,
// and the chain's point x and running centre c stay in registers for the
// whole launch; per accepted step the only HBM traffic is the gathered warmup
// column (n doubles). The loop body restates createPoint/fillrandPoint/
// advNextStep/findMaxAbs/reprojectPoint(2)/correctBounds of the reference
// (ACHRcu/ACHRCuda.cu:193-359) with the reference's tolerances and order:
//   pick w -> u = (w - c)/|w - c| -> chord [minStep, maxStep] from the bound
//   distances of coordinates farther than dTol from both bounds and with
//   |u_i| > uTol -> miss test -> x += t u -> every check_every accepted steps
//   max|S x| > reprojTol => x = N (N^T x) -> clip -> centre update.
// Reductions: the squared norm uses the canonical lane tree (shuffle xor
// within 64 lanes, then a pairwise tree over the wavefront partials); minima
// and maxima are order independent.
//
// Two ways to run the reprojection:
//   mode 1: inside this kernel, per chain (N read through L2; for small n*d);
//   mode 2: the kernel stops after the advance of every check step, flags the
//           chains whose residual exceeds the tolerance, and the host enqueues
//           a batched projection GEMM over the flagged chains before the next
//           segment resumes them (chains stay aligned on accepted steps, so a
//           round of check_every steps is one segment for every chain).
#include <math.h>

#include "achr_internal.h"
#include "achr_philox.h"

namespace achr {
namespace {

constexpr int kRedSlots = 16;  // wavefront partials per reduction (<= 1024 threads)
// Largest slot count that prefetches warmup points: at 11 slots (1,024
// lanes, n = 10,600) the prefetch spills 4 VGPRs and takes 15% off a plain
// step; 12 and more would spill the point. The persistent batch kernel
// (several 16-64-lane chains per workgroup), which holds its work-item loop
// as well, keeps the prefetch to 8 slots.
constexpr int kPrefetchMaxE = 11, kPrefetchMaxEBatch = 8;
constexpr int kBoundLookahead = 2;  // slots between a bound load and its use
// LDS slot arrays of the block-per-chain reductions: the norm, two pairs for
// min/max and the residual maximum. Consecutive reductions must use disjoint
// arrays (a wavefront may write the next partial while a slower one still
// reads the last): with a centre the norm separates two min/max reductions;
// without one (direction bank) the min/max pair alternates with the attempt
// parity.
constexpr int kRedSum = 0, kRedMinMax0 = 1, kRedMax = 3, kRedMinMax1 = 4, kRedArrays = 6;
// A chain that has its workgroup to itself keeps the bounds of its first
// kLdsBoundSlots slots in LDS, as (lb, ub) pairs that each thread writes and
// reads for its own elements only (no barrier); further slots read theirs
// from L2 each step (holding them in registers instead spills). 9 slots of
// pairs fill 144 KiB at 1,024 lanes, and at fewer lanes as much per thread,
// so a CU still holds 1,024 such threads. A plain chain with more than 8
// slots uses that LDS for its centre (8 E bytes per thread) and the bounds
// of as many first slots as fit beside it in a thread's 159 bytes.
constexpr int kLdsBoundSlots = 9;
constexpr int kLdsBytesPerThread = 159;

template <int TPC, int E, int CPB, bool DIR>
struct BoundPlan {
  static constexpr bool centre = !DIR && E > 8;
  static constexpr bool resident = CPB == 1 && (DIR || E <= 8);
  static constexpr int beside = (kLdsBytesPerThread - 8 * E) / 16;
  static constexpr int lds = resident ? (E < kLdsBoundSlots ? E : kLdsBoundSlots)
                                      : (CPB == 1 && centre && beside > 0 ? (beside < E ? beside : E) : 0);
  static constexpr int base = centre ? TPC * E : 0;  // doubles past the reduction arrays
};

// Buffer-resource loads: 32-bit byte offsets (one VGPR per slot, shared by
// every per-reaction array) and hardware range checking — an offset at or
// beyond num_records reads 0.0 without touching memory. Zero is inert in
// every formula of the step (u = 0 adds nothing to the norm and is below
// uTol; a zero x with zero bounds is not farther than dTol from them, stays
// unclipped, and keeps a zero centre), so slots past n need no masks.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer, gfx9-family data format
constexpr int kOob = 0x7FFFFFF8;        // an offset every resource here rejects

__device__ __forceinline__ rsrc_t make_rsrc(const double* p, long long bytes) {
  const int nrec = bytes <= 0 ? 0 : (bytes > 0x7FFFFFF0LL ? 0x7FFFFFF0 : static_cast<int>(bytes));
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), static_cast<short>(0), nrec, kRsrcWord3);
}

__device__ __forceinline__ double bload(rsrc_t r, int off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

template <int TPC>
__device__ __forceinline__ void group_sync() {
  if constexpr (TPC > 64) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
}

// A double moved between lanes by a DPP pattern (VALU only, no LDS trip).
// A lane whose source lies outside its row (row_shl) reads 0 (bound_ctrl):
// the same +0.0 that a zero "old" operand would leave, without the two moves
// that set the destination to that old value first.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const int l = static_cast<int>(b), h = static_cast<int>(b >> 32);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, l, CTRL, 0xF, 0xF, true));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, h, CTRL, 0xF, 0xF, true));
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
// v + (v of lane ^ X) for X = 32 or 16 by the gfx950 lane swaps: after the
// swap every lane holds its own value and its partner's, one in each of the
// two outputs, so their sum is the butterfly stage's (addition commutes).
template <int X>
__device__ __forceinline__ double xor_sum(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned l = static_cast<unsigned>(b), h = static_cast<unsigned>(b >> 32);
  auto swap = [](unsigned u) {
    return X == 32 ? __builtin_amdgcn_permlane32_swap(u, u, false, false)
                   : __builtin_amdgcn_permlane16_swap(u, u, false, false);
  };
  const auto lo = swap(l), hi = swap(h);
  const double a = __builtin_bit_cast(double, (static_cast<unsigned long long>(hi[0]) << 32) | lo[0]);
  const double c = __builtin_bit_cast(double, (static_cast<unsigned long long>(hi[1]) << 32) | lo[1]);
  return a + c;
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b), lane));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane));
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
template <int TPC>
__device__ __forceinline__ double group_sum(double v, double* red) {
  constexpr int G = TPC < 64 ? TPC : 64;
  if constexpr (G == 64) {
    // The butterfly's stages 32 and 16 by lane swaps (each lane's sum of its
    // value and its partner's); stages 8, 4, 2 and 1 by row_shl:o, which
    // hands lane i the value of lane i + o = i ^ o for the lanes below o in
    // each row. Lane 0 depends only on such lanes, so it forms exactly the
    // butterfly's sums (addition commutes); it is the lane whose value is
    // used (the wavefront's partial, or broadcast).
    v = xor_sum<32>(v);
    v = xor_sum<16>(v);
    v = v + dpp_f64<0x108>(v);
    v = v + dpp_f64<0x104>(v);
    v = v + dpp_f64<0x102>(v);
    v = v + dpp_f64<0x101>(v);
    if constexpr (TPC == 64) v = readlane_f64(v, 0);
  } else {
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) v = v + __shfl_xor(v, o);
  }
  if constexpr (TPC > 64) {
    constexpr int NW = TPC / 64;
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double p[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) p[i] = red[i];
#pragma unroll
    for (int o = NW / 2; o >= 1; o >>= 1) {
#pragma unroll
      for (int i = 0; i < o; ++i) p[i] = p[i] + p[i + o];
    }
    v = p[0];
  }
  return v;
}

// v_max_f64 / v_min_f64: the IEEE 754-2019 maximum and minimum of two
// doubles, with -0 ordered below +0 on both operand orders and a NaN operand
// giving the other one (denormals kept; checked on the device). For values
// that are never NaN nor signed zeros of opposite sign (chord ends,
// +-infinity, absolute residuals) that is the same bits as compare-and-
// assign; the bank-mode clip relies on the signed-zero order itself.
// Written as the instructions themselves: fmax / fmin would first quiet
// possible signalling NaNs in every operand that comes from a shuffle or
// from LDS, one more instruction per operand.
__device__ __forceinline__ double vmax(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double vmin(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Maximum (MAX) or minimum over each 16-lane row, left in every lane of the
// row: quad exchanges, then the half-row and row mirrors. The extremum of a
// set does not depend on the pairing, so this equals a butterfly exactly.
template <bool MAX>
__device__ __forceinline__ double row_extremum(double v) {
  auto op = [](double a, double b) { return MAX ? vmax(a, b) : vmin(a, b); };
  v = op(v, dpp_f64<0xB1>(v));   // quad_perm [1,0,3,2]
  v = op(v, dpp_f64<0x4E>(v));   // quad_perm [2,3,0,1]
  v = op(v, dpp_f64<0x141>(v));  // row_half_mirror
  return op(v, dpp_f64<0x140>(v));  // row_mirror
}
// The same over a wavefront's 64 values: the four row results are combined
// from lanes 0, 16, 32 and 48 (uniform over the wavefront).
template <bool MAX>
__device__ __forceinline__ double wave_extremum(double v) {
  auto op = [](double a, double b) { return MAX ? vmax(a, b) : vmin(a, b); };
  v = row_extremum<MAX>(v);
  return op(op(readlane_f64(v, 0), readlane_f64(v, 16)), op(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Maxima and minima of values that are never NaN over the chain's lanes.
template <int TPC>
__device__ __forceinline__ void group_minmax(double& mn, double& mx, double* red_mn, double* red_mx) {
  constexpr int G = TPC < 64 ? TPC : 64;
  if constexpr (G == 64) {
    mn = wave_extremum<true>(mn);
    mx = wave_extremum<false>(mx);
  } else if constexpr (G >= 16) {
    // groups of 16 or 32 lanes: each 16-lane row by DPP, then the two rows
    // of a 32-lane group exchanged once
    mn = row_extremum<true>(mn);
    mx = row_extremum<false>(mx);
    if constexpr (G == 32) {
      mn = vmax(mn, __shfl_xor(mn, 16));
      mx = vmin(mx, __shfl_xor(mx, 16));
    }
  } else {
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) {
      mn = vmax(mn, __shfl_xor(mn, o));
      mx = vmin(mx, __shfl_xor(mx, o));
    }
  }
  if constexpr (TPC > 64) {
    constexpr int NW = TPC / 64;
    if ((threadIdx.x & 63) == 0) {
      red_mn[threadIdx.x >> 6] = mn;
      red_mx[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    // lane i of every wavefront takes wavefront i's partial (lanes past NW
    // repeat partial 0) and the first row reduces them
    const int l = static_cast<int>(threadIdx.x & 63);
    const int src = l < NW ? l : 0;
    mn = readlane_f64(row_extremum<true>(red_mn[src]), 0);
    mx = readlane_f64(row_extremum<false>(red_mx[src]), 0);
  }
}

template <int TPC>
__device__ __forceinline__ double group_max(double v, double* red) {
  constexpr int G = TPC < 64 ? TPC : 64;
  if constexpr (G == 64) {
    v = wave_extremum<true>(v);
  } else {
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) v = vmax(v, __shfl_xor(v, o));
  }
  if constexpr (TPC > 64) {
    constexpr int NW = TPC / 64;
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const int l = static_cast<int>(threadIdx.x & 63);
    v = readlane_f64(row_extremum<true>(red[l < NW ? l : 0]), 0);
  }
  return v;
}

// Lane 0's value on every lane. For a value that already is equal on all
// lanes of the wavefront (a reduced norm, chord end or residual of a chain
// that owns whole wavefronts) this changes nothing, but it tells the compiler
// so: the chain's counters, its branches and the Philox streams derived from
// them then live in scalar registers and run on the scalar unit.
__device__ __forceinline__ double wave_uniform(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(u & 0xFFFFFFFFu)));
  const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(u >> 32)));
  return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// 1/b rounded to nearest for |b| in [2^-30, 2^30] (the components of a
// bank direction that the chord uses: |u_i| > uTol): the hardware estimate
// refined by three Newton steps y <- y + y (1 - b y), each residual 1 - b y
// exact through the FMA. Two steps leave y within one ulp of 1/b (one of its
// two neighbours); from such a y the third step rounds to nearest
// (P. Markstein, IA-64 and Elementary Functions, 2000, ch. 8). Zero or
// denormal components, which the chord never uses, may give a non-finite
// value. Checked against IEEE division by achr_bank_reciprocal's test.
__device__ __forceinline__ double rcp_rn(double b) {
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  return fma(y, e, y);
}

// a / b rounded to nearest, given y = rcp_rn(b): q = a y is within an ulp of
// a / b, the residual a - b q is exact through the FMA, and q + r y rounds to
// the correctly rounded quotient (Markstein's final correction, P.
// Markstein, IA-64 and Elementary Functions, 2000, ch. 8), for |b| in
// [2^-30, 2^30] and quotients and residuals clear of the subnormal range. A
// zero numerator returns a y, which carries IEEE division's signed zero.
// Checked against IEEE division by achr_quotient's test.
__device__ __forceinline__ double div_rn(double a, double b, double y) {
  const double q = a * y;
  const double r = fma(-q, b, a);
  const double qc = fma(r, y, q);
  return a == 0.0 ? q : qc;
}
// The same for b > 0 without a select: the residual is formed with the
// opposite sign, s = q b - a = -(a - q b) exactly, and the correction is
// q - s y. For a = +-0, q = a y carries a's sign, s = +0 and q - s y adds a
// zero of q's sign, so the result is a's signed zero as IEEE division gives;
// otherwise every step equals div_rn's (rounding to nearest is symmetric).
__device__ __forceinline__ double div_rn_pos(double a, double b, double y) {
  const double q = a * y;
  return fma(fma(q, b, -a), -y, q);
}
// The same without the zero case, for numerators that are never zero when
// the quotient is used (chord distances beyond dTol).
__device__ __forceinline__ double div_rn_nz(double a, double b, double y) {
  const double q = a * y;
  return fma(fma(-q, b, a), y, q);
}

__device__ __forceinline__ void bstore(rsrc_t r, int off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned int, v), r,
                                        off, 0, 0);
}

// Addressing of one n-vector per chain (the chain's row of X, T, a trace
// row): with TPC >= 64 the chain index is wave-uniform and the row gets its
// own resource; smaller groups share a wavefront with other chains, so they
// address from the array base with a per-lane offset (the host keeps such
// arrays under 2 GiB) and send slots past the row length out of range.
template <int TPC>
struct Row {
  rsrc_t r;
  int base;  // element offset of the row from the resource base
  int len;
  __device__ __forceinline__ Row(double* arr, long long row, int ld, int len_) : len(len_) {
    if constexpr (TPC >= 64) {
      const long long ru = __builtin_amdgcn_readfirstlane(static_cast<int>(row));
      r = make_rsrc(arr + ru * ld, 8LL * len_);
      base = 0;
    } else {
      r = make_rsrc(arr, 0x7FFFFFF0LL);
      base = static_cast<int>(row) * ld;
    }
  }
  __device__ __forceinline__ int off(int i) const {
    if constexpr (TPC >= 64) {
      return i * 8;
    } else {
      return i < len ? (base + i) * 8 : kOob;
    }
  }
  __device__ __forceinline__ double ld(int i) const { return bload(r, off(i)); }
  __device__ __forceinline__ void st(int i, double v) const { bstore(r, off(i), v); }
};

// One segment of chain b of the model `a` (b is group-uniform; with TPC > 64
// the block is one chain, so every early return is block-uniform). The
// arguments come by value: a reference to the kernel's argument block keeps
// the compiler from scalarising it and costs about 40 more spilled VGPRs.
template <int TPC, int E, int CPB, bool DIR, bool BATCH>
__device__ __forceinline__ void run_chain(const StepArgs a, const int b_in, double* lds) {
  // With 64 or more lanes the chain owns whole wavefronts: its index is
  // wavefront-uniform, and so is everything it branches on.
  constexpr bool kUni = TPC >= 64;
  const int b = kUni ? __builtin_amdgcn_readfirstlane(b_in) : b_in;
  double* red = lds;  // 4 reduction slot arrays; only block-per-chain schedules use them
  const int g = threadIdx.x / TPC;
  const int lane = threadIdx.x % TPC;
  if (b >= a.P) return;
  if (a.mode == 2 && lane == 0) a.flag[b] = 0;
  int4 st = a.st[b];
  int4 sts = a.stats[b];
  if (st.z == kDone) {
    // A finished chain resumes where it stopped when its run is extended
    // past its step count (its state is in X and C), unless the miss cap
    // ended it.
    if (st.y >= a.S || (sts.w & 1)) return;
    st.z = kRunning;
  }
  const int n = a.n;
  const int d = a.d;

  const rsrc_t r_lb = make_rsrc(a.lb, 8LL * n);
  const rsrc_t r_ub = make_rsrc(a.ub, 8LL * n);
  using BP = BoundPlan<TPC, E, CPB, DIR>;
  double2* bl = reinterpret_cast<double2*>(lds + kRedArrays * kRedSlots + BP::base);  // (lb, ub) of the LDS slots
#pragma unroll
  for (int k = 0; k < BP::lds; ++k) {
    const int off = (lane + k * TPC) * 8;
    bl[lane + k * TPC] = make_double2(bload(r_lb, off), bload(r_ub, off));
  }
  // (lb, ub) of slot k, from LDS or memory
  auto bounds = [&](int k) -> double2 {
    if (k < BP::lds) return bl[lane + k * TPC];
    const int off = (lane + k * TPC) * 8;
    return make_double2(bload(r_lb, off), bload(r_ub, off));
  };
  const Row<TPC> xrow(a.X, b, n, n);   // spill of x for the residual / projection
  const Row<TPC> crow(a.C, b, n, n);
  const Row<TPC> trow(a.T, b, d, d);   // N^T x of the in-kernel projection

  // Large slot counts keep the running centre in LDS (its own slots only, so
  // no barrier is needed); x and the direction stay in registers. With a
  // direction bank (DIR) there is no centre.
  constexpr bool kCL = !DIR && E > 8;
  double x[E];
  double c[(DIR || kCL) ? 1 : E];
  double* cl = lds + kRedArrays * kRedSlots + static_cast<size_t>(g) * (TPC * E);  // this chain's slice
  if (st.z == kFresh) {
    const rsrc_t r_c0 = make_rsrc(a.c0, 8LL * n);
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const double v = bload(r_c0, (lane + k * TPC) * 8);
      x[k] = v;
      if constexpr (DIR) {
      } else if constexpr (kCL) {
        cl[lane + k * TPC] = v;
      } else {
        c[k] = v;
      }
    }
    st = make_int4(0, 0, kRunning, 0);
  } else {
#pragma unroll
    for (int k = 0; k < E; ++k) {
      x[k] = xrow.ld(lane + k * TPC);
      if constexpr (!DIR) {
        const double cv = crow.ld(lane + k * TPC);
        if constexpr (kCL) cl[lane + k * TPC] = cv; else c[k] = cv;
      }
    }
  }

  // Small models address the warmup set from its base (the host only picks
  // such schedules when W*n*8 < 2^31).
  const rsrc_t r_wall = make_rsrc(a.warm, 8LL * n * a.W);

  uint32_t attempt = static_cast<uint32_t>(st.x);
  int s = st.y;
  int consec = st.w;
  bool pending = st.z == kPendingCheck;
  bool stopped_at_check = false;
  int checks_here = 0;  // checks passed in this launch
  const uint32_t chain = static_cast<uint32_t>(a.chain0 + b);

  // The warmup point (or bank direction) of attempt a + 1 is loaded while
  // attempt a finishes: the pick is a function of the attempt counter only.
  // Without a bank, schedules of more than kPrefetchMaxE slots load the
  // point at the top of the step instead (the prefetch registers would spill).
  constexpr bool kPre = DIR || E <= (BATCH ? kPrefetchMaxEBatch : kPrefetchMaxE);
  double wn[kPre ? E : 1];
  auto load_dir = [&](uint32_t att) {
    if constexpr (!kPre) return;
    const int idn = pick_index(stream_u01(a.seed, att, chain, a.file, a.model, kTagPick), a.W);
    if constexpr (TPC >= 64) {
      const int idu = __builtin_amdgcn_readfirstlane(idn);
      const rsrc_t r_w = make_rsrc(a.warm + static_cast<size_t>(idu) * n, 8LL * n);
#pragma unroll
      for (int k = 0; k < E; ++k) wn[k] = bload(r_w, (lane + k * TPC) * 8);
    } else {
      const int wbase = idn * n;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const int i = lane + k * TPC;
        wn[k] = bload(r_wall, i < n ? (wbase + i) * 8 : kOob);
      }
    }
  };
  load_dir(attempt);
  // Bank mode clips x to its bounds at the top of every attempt's bounds
  // loop, with the IEEE 754-2019 maximum and minimum (v_max_f64 /
  // v_min_f64: -0 orders below +0, a NaN operand yields the other): two
  // instructions per slot and no condition. The clip is idempotent, so
  // clipping after a miss or twice changes nothing; points that leave the
  // kernel after an accepted step are clipped here first. clip_due is
  // uniform over the chain: set by an accepted step (or a resume after the
  // check of one), cleared by a clip.
  bool clip_due = pending;
  auto clip_all = [&]() {
    if (!clip_due) return;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const double2 bk = bounds(k);
      x[k] = vmin(vmax(x[k], bk.x), bk.y);
    }
    clip_due = false;
  };

  for (;;) {
    if (!pending) {
      if (s >= a.S) break;
      // the step length's uniform, drawn ahead (the scalar unit forms it
      // while the vector unit works on the direction and the chord)
      const double rs = stream_u01(a.seed, static_cast<uint32_t>(s), chain, a.file, a.model, kTagStep);
      double u[E];
      double ss = 0.0;
      if constexpr (kPre) {
#pragma unroll
        for (int k = 0; k < E; ++k) {
          if constexpr (DIR) {
            u[k] = wn[k];
          } else {
            u[k] = wn[k] - (kCL ? cl[lane + k * TPC] : c[k]);
            ss = fma(u[k], u[k], ss);
          }
        }
      } else {
        const int id = pick_index(stream_u01(a.seed, attempt, chain, a.file, a.model, kTagPick), a.W);
        if constexpr (TPC >= 64) {
          // id is uniform over the wavefront: a per-column resource in SGPRs.
          const int idu = __builtin_amdgcn_readfirstlane(id);
          const rsrc_t r_w = make_rsrc(a.warm + static_cast<size_t>(idu) * n, 8LL * n);
#pragma unroll
          for (int k = 0; k < E; ++k) {
            u[k] = bload(r_w, (lane + k * TPC) * 8) - (kCL ? cl[lane + k * TPC] : c[k]);
            ss = fma(u[k], u[k], ss);
          }
        } else {
          const int wbase = id * n;
#pragma unroll
          for (int k = 0; k < E; ++k) {
            const int i = lane + k * TPC;
            u[k] = bload(r_wall, i < n ? (wbase + i) * 8 : kOob) - (kCL ? cl[i] : c[k]);
            ss = fma(u[k], u[k], ss);
          }
        }
      }
      // Bounds read from memory (slots past the LDS-resident rows) are
      // requested kLook slots ahead of their use (the first ones before the
      // norm's reduction), so their L2 latency overlaps the divisions of the
      // slots before. The persistent batch kernel, short of registers, reads
      // them at use.
      constexpr int kLook = BATCH ? 0 : kBoundLookahead;
      double2 bq[E];
#pragma unroll
      for (int k = 0; k < (E < kLook ? E : kLook); ++k)
        if (k >= BP::lds) bq[k] = bounds(k);
      // Bank directions are unit vectors: no norm, no division.
      double nrm = 1.0;
      if constexpr (!DIR) {
        ss = group_sum<TPC>(ss, red + kRedSum * kRedSlots);
        if constexpr (kUni) ss = wave_uniform(ss);
        nrm = sqrt(ss);
        // u / |u| (the reference divides each component, ACHRCuda.cu:264-267):
        // one correctly rounded reciprocal of the norm and a corrected
        // product per component give the same bits as the divisions; norms
        // outside [2^-30, 2^30] divide.
        if (nrm >= 0x1p-30 && nrm <= 0x1p30) {
          const double yn = rcp_rn(nrm);
#pragma unroll
          for (int k = 0; k < E; ++k) u[k] = div_rn_pos(u[k], nrm, yn);
        } else {
#pragma unroll
          for (int k = 0; k < E; ++k) u[k] = u[k] / nrm;
        }
      }
      double mn = -INFINITY, mx = INFINITY;
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const double ui = u[k];
        if (kLook > 0 && k + kLook < E && k + kLook >= BP::lds) bq[k + kLook] = bounds(k + kLook);
        const double2 bk = (kLook > 0 && k >= BP::lds) ? bq[k] : bounds(k);
        const double lbi = bk.x, ubi = bk.y;
        if constexpr (DIR) x[k] = vmin(vmax(x[k], lbi), ubi);  // the attempt's clip (above)
        const double du = ubi - x[k];
        const double dl = x[k] - lbi;
        // Both branches of the reference (ACHRCuda.cu:270-279) divide the same
        // two distances by u_i; divide once and select, so a wavefront whose
        // lanes disagree on the sign does not run both sets of divisions.
        // Here the two quotients share one correctly rounded reciprocal of
        // u_i and are corrected to the IEEE quotients (|u_i| > uTol >= 2^-30
        // and both distances > dTol wherever the chord uses them). A bank
        // direction (no reference counterpart) takes the plain products.
        // A coordinate the chord does not use gets a NaN reciprocal instead:
        // its candidates are NaN, and v_max_f64 / v_min_f64 return the other
        // operand for a NaN, so it leaves mn and mx as they are, the same as
        // skipping it (one 64-bit select per slot instead of two).
        double r_lo, r_hi;
        const bool use = du > a.dTol && dl > a.dTol && fabs(ui) > a.uTol;
        const double rcp = use ? rcp_rn(ui) : __builtin_nan("");
        if constexpr (DIR) {
          r_lo = -dl * rcp;
          r_hi = du * rcp;
        } else {
          r_lo = div_rn_nz(-dl, ui, rcp);
          r_hi = div_rn_nz(du, ui, rcp);
        }
        // A kept coordinate has both distances > dTol > 0, so the two
        // quotients have opposite signs: the lower candidate is their
        // minimum and the upper their maximum whatever the sign of u_i
        // (the reference's two branches pick exactly these). Kept
        // candidates are never NaN, so fmax / fmin equal its compare-and-
        // assign; coordinates that are not kept contribute nothing.
        const double lo = fmin(r_lo, r_hi);
        const double hi = fmax(r_lo, r_hi);
        mn = vmax(mn, lo);
        mx = vmin(mx, hi);
        // Schedule four slots at a time: their bound loads overlap, while
        // hoisting every slot's loads and FP64 divisions ahead would multiply
        // the live registers by E.
        if constexpr (E > 4) {
          if ((k & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
      }

      clip_due = false;
      double* rmm = red + ((DIR && (attempt & 1u)) ? kRedMinMax1 : kRedMinMax0) * kRedSlots;
      group_minmax<TPC>(mn, mx, rmm, rmm + kRedSlots);
      if constexpr (kUni) {
        mn = wave_uniform(mn);
        mx = wave_uniform(mx);
      }
      const bool miss = !(nrm > 0.0) || !(mn > -INFINITY) || !(mx < INFINITY) ||
                        (fabs(mn) < a.maxMinTol && fabs(mx) < a.maxMinTol) || (mn > mx);
      if (miss) {
        ++attempt;
        load_dir(attempt);
        ++sts.x;
        if (++consec > a.max_misses) {
          sts.w |= 1;
          break;
        }
        continue;
      }
      consec = 0;
      const double t = fma(rs, mx - mn, mn);
#pragma unroll
      for (int k = 0; k < E; ++k) x[k] = fma(t, u[k], x[k]);
      clip_due = true;
      // Prefetch the next attempt's point once this direction is consumed
      // (never two of them in registers); it lands during the residual
      // check, the clip, the centre update and the next step's set-up.
      ++attempt;
      load_dir(attempt);

      if (s % a.check_every == 0) {
        ++sts.z;
#pragma unroll
        for (int k = 0; k < E; ++k) xrow.st(lane + k * TPC, x[k]);
        // the residual of this check is left to the check kernel, which
        // runs between this launch and the projection round (plain chains)
        if constexpr (!DIR) {
          if (a.check_out) {
            stopped_at_check = true;
            break;
          }
        }
        group_sync<TPC>();
        double rho = 0.0;
        float* rrow = (a.mode == 2 && a.R32) ? a.R32 + static_cast<size_t>(b) * a.r32_ld : nullptr;
        if constexpr (TPC >= 64) {
          // Sliced ELL: wavefront v of the chain takes slices v, v + TPC/64,
          // ...; lane l sums row l of the slice. The trip count is the
          // slice's width for every lane and the entry loads are one
          // coalesced line per wavefront, with no row pointer to wait for;
          // each row's products are added in CSR order, as above.
          constexpr int NW = TPC / 64;
          const int l = lane & 63;
          for (int sl = __builtin_amdgcn_readfirstlane(lane >> 6); sl < a.nslices; sl += NW) {
            const int e0 = __builtin_amdgcn_readfirstlane(a.ell_ptr[sl]);
            const int width = (__builtin_amdgcn_readfirstlane(a.ell_ptr[sl + 1]) - e0) >> 6;
            double acc = 0.0;
            for (int e = 0; e < width; e += 4) {
              int ci[4];
              double vv[4], xv[4];
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                const int at = e0 + (min(e + q, width - 1) << 6) + l;
                ci[q] = a.ell_col[at];
                vv[q] = a.ell_val[at];
              }
#pragma unroll
              for (int q = 0; q < 4; ++q) xv[q] = xrow.ld(ci[q] < 0 ? 0 : ci[q]);
#pragma unroll
              for (int q = 0; q < 4; ++q)
                if (e + q < width && ci[q] >= 0) acc = fma(vv[q], xv[q], acc);
            }
            const int row = a.ell_row[(sl << 6) + l];
            if (row >= 0) {
              const double av = fabs(acc);
              if (av > rho) rho = av;
              if (rrow) rrow[row] = static_cast<float>(acc);
            }
          }
        } else
        for (int j = lane; j < a.m; j += TPC) {
          double acc = 0.0;
          const int p0 = a.rowptr[j], p1 = a.rowptr[j + 1];
          // four nonzeros at a time: their index, value and x loads are in
          // flight together (positions past the row read its last entry and
          // are skipped), then the products are added in order
          for (int p = p0; p < p1; p += 4) {
            int ci[4];
            double vv[4], xv[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int q = min(p + e, p1 - 1);
              ci[e] = a.colidx[q];
              vv[e] = a.vals[q];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) xv[e] = xrow.ld(ci[e]);
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (p + e < p1) acc = fma(vv[e], xv[e], acc);
          }
          const double av = fabs(acc);
          if (av > rho) rho = av;
          if (rrow) rrow[j] = static_cast<float>(acc);
        }
        rho = group_max<TPC>(rho, red + kRedMax * kRedSlots);
        if constexpr (kUni) rho = wave_uniform(rho);
        const bool reproj = rho > a.reprojTol && d > 0;
        if (reproj) ++sts.y;
        // Mode 2 hands flagged chains to the batched projection round after
        // this launch. A chain that needs no projection either stops too (the
        // rounds then stay aligned with the checks, which reduced-precision
        // runs use to place their FP64 rounds) or, with FP64 rounds, runs on
        // for up to check_budget checks per launch: its projection, whenever
        // it comes, is the same column-independent FP64 product, so the
        // results do not change, and the budget keeps a launch from waiting
        // on the few chains that never stop.
        ++checks_here;
        if (a.mode == 2 && (reproj || !a.stop_flagged_only || checks_here >= a.check_budget)) {
          if (lane == 0) a.flag[b] = reproj ? 1 : 0;
          stopped_at_check = true;
          break;
        }
        if (reproj) {  // mode 1
          for (int j = lane; j < d; j += TPC) {
            double acc = 0.0;
            const double* Nr = a.Nrm + j;
            for (int i = 0; i < n; ++i) acc = fma(Nr[static_cast<size_t>(i) * d], xrow.ld(i), acc);
            trow.st(j, acc);
          }
          group_sync<TPC>();
          // x_i = sum_j N[i, j] tmp_j, j ascending: column j of N is one
          // resource, its slots are this thread's rows.
#pragma unroll
          for (int k = 0; k < E; ++k) x[k] = 0.0;
          for (int j = 0; j < d; ++j) {
            const rsrc_t r_nc = make_rsrc(a.Ncm + static_cast<size_t>(j) * n, 8LL * n);
            const double tj = trow.ld(j);
#pragma unroll
            for (int k = 0; k < E; ++k) x[k] = fma(bload(r_nc, (lane + k * TPC) * 8), tj, x[k]);
          }
          group_sync<TPC>();
        }
      }
    }
    pending = false;
    if constexpr (!DIR) {
      // clip, then the centre update with the clipped point
      const double alpha = static_cast<double>(a.W + s) / static_cast<double>(a.W + s + 1);
      const double beta = 1.0 / static_cast<double>(a.W + s + 1);
#pragma unroll
      for (int k = 0; k < E; ++k) {
        const double2 bk = bounds(k);
        // as selects: written as branches, the compiler loads the lower
        // bound inside the branch, so every slot waits out two dependent
        // memory round trips
        const double xk = x[k];
        const bool above = xk > bk.y;
        const bool below = xk < bk.x;
        x[k] = above ? bk.y : (below ? bk.x : xk);
        if constexpr (kCL) {
          cl[lane + k * TPC] = fma(alpha, cl[lane + k * TPC], beta * x[k]);
        } else {
          c[k] = fma(alpha, c[k], beta * x[k]);
        }
        if constexpr (E > 4) {
          if ((k & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    ++s;
    if (a.trace != nullptr && b < a.trace_chains && s % a.trace_thin == 0) {
      const int kk = s / a.trace_thin - 1;
      if (kk < a.trace_len) {
        if constexpr (DIR) clip_all();  // the recorded point is the clipped one
        const Row<TPC> tr(a.trace, static_cast<long long>(b) * a.trace_len + kk, n, n);
#pragma unroll
        for (int k = 0; k < E; ++k) tr.st(lane + k * TPC, x[k]);
      }
    }
  }

  if (stopped_at_check) {
#pragma unroll
    for (int k = 0; k < E; ++k)
      if constexpr (!DIR) crow.st(lane + k * TPC, kCL ? cl[lane + k * TPC] : c[k]);
    st = make_int4(static_cast<int>(attempt), s, kPendingCheck, consec);
  } else {
    if constexpr (DIR) clip_all();
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int i = lane + k * TPC;
      if (i < n) a.out[static_cast<size_t>(i) * a.P + b] = x[k];
      // the final state too, from which an extended run continues
      xrow.st(i, x[k]);
      if constexpr (!DIR) crow.st(i, kCL ? cl[i] : c[k]);
    }
    st = make_int4(static_cast<int>(attempt), s, kDone, consec);
  }
  if (lane == 0) {
    a.st[b] = st;
    a.stats[b] = sts;
  }
}

template <int TPC, int E, int CPB, bool DIR>
__global__ __launch_bounds__(TPC* CPB) void step_kernel(StepArgs a) {
  extern __shared__ double lds[];
  run_chain<TPC, E, CPB, DIR, false>(a, static_cast<int>(blockIdx.x) * CPB + static_cast<int>(threadIdx.x) / TPC, lds);
}

// Several models in one launch, over an item list of {model, first chain},
// each item naming CPB consecutive chains of one model (the host pads every
// model to whole items). The item index is uniform over the workgroup, so
// the model's arguments load into scalar registers. A chain that owns its
// workgroup (128-1,024 lanes) takes one item per workgroup, and the grid is
// the item list: a loop over items kept the chain's registers live across
// iterations and spilled (31 VGPRs at 11 slots), and the chain then runs
// with the single-model kernel's schedule (warmup point prefetched up to 11
// slots, bounds requested ahead). Groups of several chains per workgroup
// (16-64 lanes) run persistently: the grid is sized to what stays resident
// and workgroup w takes items w, w + grid, w + 2*grid, ...
template <int TPC, int E, int CPB, bool DIR>
__global__ __launch_bounds__(TPC* CPB) void step_batch_kernel(const StepArgs* __restrict__ tab,
                                                              const int2* __restrict__ items, int nitems) {
  extern __shared__ double lds[];
  if constexpr (CPB == 1) {
    const int it = blockIdx.x;
    if (it >= nitems) return;
    const int2 w = items[it];
    run_chain<TPC, E, CPB, DIR, false>(tab[w.x], w.y + static_cast<int>(threadIdx.x) / TPC, lds);
  } else {
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
      const int2 w = items[it];
      run_chain<TPC, E, CPB, DIR, true>(tab[w.x], w.y + static_cast<int>(threadIdx.x) / TPC, lds);
    }
  }
}

template <int TPC, int E, bool DIR, bool BATCH>
const void* kernel_for() {
  constexpr int CPB = TPC <= 64 ? 256 / TPC : 1;
  if constexpr (BATCH)
    return reinterpret_cast<const void*>(&step_batch_kernel<TPC, E, CPB, DIR>);
  else
    return reinterpret_cast<const void*>(&step_kernel<TPC, E, CPB, DIR>);
}

template <int TPC, bool DIR, bool BATCH>
const void* kernel_for_e(int e) {
  if constexpr (TPC == 1024) {
    switch (e) {
      case 3: return kernel_for<TPC, 3, DIR, BATCH>();
      case 5: return kernel_for<TPC, 5, DIR, BATCH>();
      case 6: return kernel_for<TPC, 6, DIR, BATCH>();
      case 7: return kernel_for<TPC, 7, DIR, BATCH>();
      case 9: return kernel_for<TPC, 9, DIR, BATCH>();
      case 10: return kernel_for<TPC, 10, DIR, BATCH>();
      case 11: return kernel_for<TPC, 11, DIR, BATCH>();
      case 13: return kernel_for<TPC, 13, DIR, BATCH>();
      case 14: return kernel_for<TPC, 14, DIR, BATCH>();
      case 15: return kernel_for<TPC, 15, DIR, BATCH>();
      default: break;
    }
  }
  switch (e) {
    case 1: return kernel_for<TPC, 1, DIR, BATCH>();
    case 2: return kernel_for<TPC, 2, DIR, BATCH>();
    case 4: return kernel_for<TPC, 4, DIR, BATCH>();
    case 8: return kernel_for<TPC, 8, DIR, BATCH>();
    case 12: return kernel_for<TPC, 12, DIR, BATCH>();
    case 16: return kernel_for<TPC, 16, DIR, BATCH>();
    default: return nullptr;
  }
}

template <bool DIR, bool BATCH>
const void* kernel_for_tpc(int tpc, int e) {
  switch (tpc) {
    case 16: return kernel_for_e<16, DIR, BATCH>(e);
    case 32: return kernel_for_e<32, DIR, BATCH>(e);
    case 64: return kernel_for_e<64, DIR, BATCH>(e);
    case 128: return kernel_for_e<128, DIR, BATCH>(e);
    case 256: return kernel_for_e<256, DIR, BATCH>(e);
    case 512: return kernel_for_e<512, DIR, BATCH>(e);
    case 1024: return kernel_for_e<1024, DIR, BATCH>(e);
    default: return nullptr;
  }
}

const void* step_function(const Schedule& s, bool dir, bool batch) {
  if (dir) return batch ? kernel_for_tpc<true, true>(s.tpc, s.e) : kernel_for_tpc<true, false>(s.tpc, s.e);
  return batch ? kernel_for_tpc<false, true>(s.tpc, s.e) : kernel_for_tpc<false, false>(s.tpc, s.e);
}

// Dynamic LDS beyond 64 KiB must be opted into per kernel.
hipError_t allow_lds(const void* fn, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
}

// Slot counts with a kernel: every count up to 16 for 1,024-lane chains
// (the schedule of every model above 2,048 reactions, which then pads at
// most one partial slot row), 1, 2, 4, 8, 12 and 16 for smaller groups.
bool elems_instantiated(int tpc, int e) {
  if (tpc == 1024) return e >= 1 && e <= 16;
  return e == 1 || e == 2 || e == 4 || e == 8 || e == 12 || e == 16;
}

}  // namespace

bool schedule_supported(int tpc, int e) {
  bool tpc_ok = tpc == 16 || tpc == 32 || tpc == 64 || tpc == 128 || tpc == 256 || tpc == 512 || tpc == 1024;
  return tpc_ok && elems_instantiated(tpc, e);
}

int elems_for(int n, int tpc) {
  const int need = (n + tpc - 1) / tpc;
  for (int v = need < 1 ? 1 : need; v <= 16; ++v)
    if (elems_instantiated(tpc, v)) return v;
  return -1;
}

namespace {
__global__ __launch_bounds__(256) void bank_reciprocal_kernel(const double* __restrict__ u, double* __restrict__ r,
                                                             size_t count) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    r[i] = rcp_rn(u[i]);
}
}  // namespace

namespace {
__global__ __launch_bounds__(256) void quotient_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                      double* __restrict__ q, size_t count) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    q[i] = b[i] > 0.0 ? div_rn_pos(a[i], b[i], rcp_rn(b[i])) : div_rn(a[i], b[i], rcp_rn(b[i]));
}
}  // namespace

hipError_t launch_quotient(const double* a, const double* b, double* q, size_t count, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((count + 255) / 256, 65536);
  hipLaunchKernelGGL(quotient_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, a, b, q, count);
  return hipGetLastError();
}

hipError_t launch_bank_reciprocal(const double* u, double* r, size_t count, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((count + 255) / 256, 65536);
  hipLaunchKernelGGL(bank_reciprocal_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, u, r, count);
  return hipGetLastError();
}

size_t step_lds_bytes(bool dir, const Schedule& s) {
  size_t bytes = kRedArrays * kRedSlots * sizeof(double);
  const bool centre = s.e > 8 && !dir;
  if (centre) bytes += static_cast<size_t>(s.cpb) * s.tpc * s.e * sizeof(double);  // centres in LDS
  // (lb, ub) pairs of the first slots (BoundPlan)
  int rows = 0;
  if (s.cpb == 1 && (dir || s.e <= 8)) rows = std::min(s.e, kLdsBoundSlots);
  else if (s.cpb == 1 && centre) rows = std::max(0, std::min(s.e, (kLdsBytesPerThread - 8 * s.e) / 16));
  return bytes + static_cast<size_t>(rows) * s.tpc * 2 * sizeof(double);
}

hipError_t launch_step(const StepArgs& a, const Schedule& s, hipStream_t stream) {
  if (!schedule_supported(s.tpc, s.e)) return hipErrorInvalidValue;
  const bool dir = a.dir_mode != 0;
  const void* fn = step_function(s, dir, false);
  const size_t lds = step_lds_bytes(dir, s);
  hipError_t e = allow_lds(fn, lds);
  if (e != hipSuccess) return e;
  const unsigned grid = static_cast<unsigned>((a.P + s.cpb - 1) / s.cpb);
  StepArgs arg = a;
  void* args[] = {&arg};
  return hipLaunchKernel(fn, dim3(grid), dim3(s.tpc * s.cpb), args, lds, stream);
}

hipError_t prepare_step_batch(const Schedule& s, bool dir, int* resident) {
  if (!schedule_supported(s.tpc, s.e)) return hipErrorInvalidValue;
  const void* fn = step_function(s, dir, true);
  const size_t lds = step_lds_bytes(dir, s);
  hipError_t e = allow_lds(fn, lds);
  if (e != hipSuccess) return e;
  int per_cu = 0, dev = 0, cus = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, s.tpc * s.cpb, lds);
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  *resident = (per_cu > 0 ? per_cu : 1) * cus;
  return hipSuccess;
}

hipError_t launch_step_batch(const StepArgs* tab, const int2* items, int nitems, int grid, const Schedule& s,
                             bool dir, hipStream_t stream) {
  if (!schedule_supported(s.tpc, s.e) || grid <= 0) return hipErrorInvalidValue;
  const void* fn = step_function(s, dir, true);
  const size_t lds = step_lds_bytes(dir, s);
  void* args[] = {&tab, &items, &nitems};
  return hipLaunchKernel(fn, dim3(static_cast<unsigned>(grid)), dim3(s.tpc * s.cpb), args, lds, stream);
}

}  // namespace achr
