// Residual check of a whole segment's chains in one launch (batched rounds).
//
// The reference checks each chain every 10th accepted step with a COO SpMV
// r = S x and a serial max |r| (findMaxAbs, ACHRcu/ACHRCuda.cu:232-240,
// :325-338); a residual above 1e-9 triggers the reprojection (:338-343).
// Inside the step kernel that check costs a chain-per-CU schedule a fifth of
// its step: S (0.43 MB at 5,800 x 10,600) is re-read past L2 for every chain
// and every check, the gathers of the other chains having evicted it, and its
// three dependent loads per slice wait out memory latency with nothing else
// to run. Here the step kernel stops every chain at its check with x stored
// in its row of X, and this kernel evaluates the residuals of all of them:
// a workgroup stages one chain's x in LDS (one coalesced read of n doubles),
// its wavefronts sum the rows of S from a sliced ELL table that stays in L2
// across the chains of the launch, and the next chain's x is loaded into
// registers while the current one is summed.
//
// Each row is summed exactly as the step kernel and the oracle sum it: the
// row's nonzeros in CSR order, fma(value, x, acc) from +0.0; the chain's
// residual is the largest absolute row sum (order independent). The flag,
// the reprojection counter and the FP32 residual rows of the
// reduced-precision rounds are the ones the in-kernel check writes, so the
// two placements give identical results.
#include <algorithm>
#include <map>
#include <vector>

#include "achr_host.h"
#include "achr_internal.h"

namespace achr {
namespace {

constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxSlots = kMaxCheckReactions / kThreads;  // x elements per thread

using rsrc_t = __amdgpu_buffer_rsrc_t;

__device__ __forceinline__ rsrc_t row_rsrc(const double* p, int n) {
  // raw buffer: offsets at or past 8 n bytes read 0.0 without a memory access
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(p), static_cast<short>(0), 8 * n, 0x00020000);
}

__device__ __forceinline__ double vmax(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = vmax(v, __shfl_xor(v, o));
  return v;
}

// NT threads per workgroup, XS slots of the next chain's x per thread
// (XS * NT >= n). Models of at most 4,096 reactions take 256-thread
// workgroups, several per CU, so the checks of several chains overlap; the
// entries of a slice are read eight at a time, the next eight requested
// before the current ones are summed (long rows of hub metabolites would
// otherwise wait out an L2 round trip every few entries).
template <int NT, int XS>
__global__ __launch_bounds__(NT) void residual_check_kernel(CheckArgs a) {
  constexpr int NW = NT / 64;
  constexpr int U = 8;
  extern __shared__ double xs[];  // n doubles of the current chain, then NW partial maxima
  double* red = xs + a.n;
  const int tid = static_cast<int>(threadIdx.x);
  const int l = tid & 63;
  const int w = tid >> 6;
  const int n = a.n;
  double pre[XS];
  auto pending = [&](int b) { return b < a.P && a.st[b].z == kPendingCheck; };
  auto fetch = [&](int b) {
    const rsrc_t r = row_rsrc(a.X + static_cast<size_t>(b) * n, n);
#pragma unroll
    for (int k = 0; k < XS; ++k)
      pre[k] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (tid + k * NT) * 8, 0, 0));
  };
  int b = static_cast<int>(blockIdx.x);
  bool have = pending(b);
  if (have) fetch(b);
  while (b < a.P) {
    const int bn = b + static_cast<int>(gridDim.x);
    const bool have_n = pending(bn);
    if (have) {
#pragma unroll
      for (int k = 0; k < XS; ++k)
        if (tid + k * NT < n) xs[tid + k * NT] = pre[k];
      __syncthreads();
      if (have_n) fetch(bn);  // in flight while this chain's rows are summed
      float* rrow = a.R32 ? a.R32 + static_cast<size_t>(b) * a.r32_ld : nullptr;
      double rho = 0.0;
      for (int sl = w; sl < a.nslices; sl += NW) {
        const int e0 = __builtin_amdgcn_readfirstlane(a.ell_ptr[sl]);
        const int width = (__builtin_amdgcn_readfirstlane(a.ell_ptr[sl + 1]) - e0) >> 6;
        double acc = 0.0;
        int ci[U];
        double vv[U];
        auto ld = [&](int e) {
#pragma unroll
          for (int q = 0; q < U; ++q) {
            const int at = e0 + (min(e + q, width - 1) << 6) + l;
            ci[q] = a.ell_col[at];
            vv[q] = a.ell_val[at];
          }
        };
        if (width > 0) ld(0);
        for (int e = 0; e < width; e += U) {
          int cc[U];
          double cv[U];
#pragma unroll
          for (int q = 0; q < U; ++q) {
            cc[q] = ci[q];
            cv[q] = vv[q];
          }
          if (e + U < width) ld(e + U);
#pragma unroll
          for (int q = 0; q < U; ++q) {
            const double xv = xs[cc[q] < 0 ? 0 : cc[q]];
            if (e + q < width && cc[q] >= 0) acc = fma(cv[q], xv, acc);
          }
        }
        const int row = a.ell_row[(sl << 6) + l];
        if (row >= 0) {
          const double av = fabs(acc);
          if (av > rho) rho = av;
          if (rrow) rrow[row] = static_cast<float>(acc);
        }
      }
      rho = wave_max(rho);
      if (l == 0) red[w] = rho;
      __syncthreads();
      if (tid == 0) {
        double r = red[0];
#pragma unroll
        for (int i = 1; i < NW; ++i) r = vmax(r, red[i]);
        const bool reproj = r > a.reprojTol && a.d > 0;
        a.flag[b] = reproj ? 1 : 0;
        if (reproj) ++a.stats[b].y;
      }
      // the partials and xs are rewritten only after every wavefront is done
      // with this chain: the next chain's LDS stores follow this barrier
      __syncthreads();
    } else if (have_n) {
      fetch(bn);
    }
    b = bn;
    have = have_n;
  }
}

// The resident form: a wavefront's entries of S stay in registers across the
// chains of the launch, so per chain the only memory reads are the chain's x
// (staged in LDS as above) and L2 is not involved. Each lane walks its rows of
// the wavefront's slices one after another; the slice ends are uniform over
// the wavefront (ELL), marked in a bitmask, where the row's sum is finished.
template <int R, int XS>
__global__ __launch_bounds__(kThreads) void residual_check_resident_kernel(CheckArgs a) {
  extern __shared__ double xs[];  // n doubles of x, nv values of S, kWaves partial maxima
  double* vt = xs + a.n;
  double* red = vt + a.nv;
  const int tid = static_cast<int>(threadIdx.x);
  const int l = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int n = a.n;
  for (int i = tid; i < a.nv; i += kThreads) vt[i] = a.vtab[i];
  const int J = a.wlen[w];
  const unsigned long long ends = a.wends[w];
  unsigned ent[R];
#pragma unroll
  for (int j = 0; j < R; ++j) ent[j] = j < J ? a.ent[a.woff[w] + 64 * j + l] : kCheckPad;
  const int* rows = a.rowmap + a.roff[w] + l;
  double pre[XS];  // x of the next chain, XS * kThreads >= n
  auto pending = [&](int b) { return b < a.P && a.st[b].z == kPendingCheck; };
  auto fetch = [&](int b) {
    const rsrc_t r = row_rsrc(a.X + static_cast<size_t>(b) * n, n);
#pragma unroll
    for (int k = 0; k < XS; ++k)
      pre[k] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (tid + k * kThreads) * 8, 0, 0));
  };
  int b = static_cast<int>(blockIdx.x);
  bool have = pending(b);
  if (have) fetch(b);
  while (b < a.P) {
    const int bn = b + static_cast<int>(gridDim.x);
    const bool have_n = pending(bn);
    if (have) {
#pragma unroll
      for (int k = 0; k < XS; ++k)
        if (tid + k * kThreads < n) xs[tid + k * kThreads] = pre[k];
      __syncthreads();  // also orders the value table before its first use
      if (have_n) fetch(bn);  // in flight while this chain's rows are summed
      float* rrow = a.R32 ? a.R32 + static_cast<size_t>(b) * a.r32_ld : nullptr;
      double rho = 0.0, acc = 0.0;
      int k = 0;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        if (j < J) {
          unsigned wd = ent[j];
          // opaque to the compiler: otherwise it hoists both LDS addresses
          // and the padding test of every entry out of the chain loop, which
          // triples the registers the entries occupy
          asm volatile("" : "+v"(wd));
          const double xv = xs[wd & 0xFFFFu];
          const double vv = vt[(wd >> 16) & 0x7FFFu];
          const double nx = fma(vv, xv, acc);
          acc = (wd & kCheckPad) ? acc : nx;
          if ((ends >> j) & 1ull) {
            const double av = fabs(acc);
            if (av > rho) rho = av;
            if (rrow) {
              const int row = rows[64 * k];
              if (row >= 0) rrow[row] = static_cast<float>(acc);
            }
            acc = 0.0;
            ++k;
          }
        }
        // four entries at a time: their LDS reads overlap without every
        // entry's operands being hoisted into registers at once
        if ((j & 3) == 3) __builtin_amdgcn_sched_barrier(0);
      }
      rho = wave_max(rho);
      if (l == 0) red[w] = rho;
      __syncthreads();
      if (tid == 0) {
        double r = red[0];
#pragma unroll
        for (int i = 1; i < kWaves; ++i) r = vmax(r, red[i]);
        const bool reproj = r > a.reprojTol && a.d > 0;
        a.flag[b] = reproj ? 1 : 0;
        if (reproj) ++a.stats[b].y;
      }
      __syncthreads();
    } else if (have_n) {
      fetch(bn);
    }
    b = bn;
    have = have_n;
  }
}

}  // namespace

constexpr int kSmallCheckN = 4096;  // models up to this size: 256-thread workgroups
size_t check_lds_bytes(int n) { return (static_cast<size_t>(n) + kWaves) * sizeof(double); }

size_t resident_lds_bytes(int n, int nv) { return (static_cast<size_t>(n) + nv + kWaves) * sizeof(double); }

hipError_t launch_check(const CheckArgs& a, hipStream_t stream) {
  if (a.P <= 0) return hipSuccess;
  if (a.n <= 0 || a.n > kMaxCheckReactions) return hipErrorInvalidValue;
  const int xs = (a.n + kThreads - 1) / kThreads;
  const bool small = !a.ent && a.n <= kSmallCheckN;
  const void* fn = reinterpret_cast<const void*>(&residual_check_kernel<kThreads, kMaxSlots>);
  if (small) fn = reinterpret_cast<const void*>(&residual_check_kernel<256, kSmallCheckN / 256>);
  if (a.ent) {
    if (xs <= 4)
      fn = reinterpret_cast<const void*>(&residual_check_resident_kernel<kCheckResident, 4>);
    else if (xs <= 8)
      fn = reinterpret_cast<const void*>(&residual_check_resident_kernel<kCheckResident, 8>);
    else if (xs <= 12)
      fn = reinterpret_cast<const void*>(&residual_check_resident_kernel<kCheckResident, 12>);
    else
      fn = reinterpret_cast<const void*>(&residual_check_resident_kernel<kCheckResident, 16>);
  }
  const size_t lds = a.ent ? resident_lds_bytes(a.n, a.nv) : check_lds_bytes(a.n);
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  int dev = 0, cus = 0, per_cu = 1;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int threads = small ? 256 : kThreads;
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, lds);
  if (e != hipSuccess) return e;
  // as many workgroups as stay resident (one per CU when the staged x fills
  // most of the LDS), each taking every grid-th chain
  const int grid = std::min(a.P, std::max(cus, 1) * std::max(per_cu, 1));
  CheckArgs arg = a;
  void* args[] = {&arg};
  return hipLaunchKernel(fn, dim3(static_cast<unsigned>(grid)), dim3(threads), args, lds, stream);
}

}  // namespace achr

namespace achr_host {

void build_check_plan(int n, int nslices, const std::vector<int>& eptr, const std::vector<int>& erow,
                      const std::vector<int>& ecol, const std::vector<double>& eval, CheckPlan* plan) {
  using achr::kCheckWaves;
  *plan = CheckPlan{};
  if (n <= 0 || n > achr::kMaxCheckReactions || nslices <= 0) return;
  // values of S as a table (bit patterns, so -0.0 and 0.0 stay apart)
  std::map<unsigned long long, int> index;
  std::vector<double> vtab;
  for (size_t i = 0; i < ecol.size(); ++i) {
    if (ecol[i] < 0) continue;
    const unsigned long long key = __builtin_bit_cast(unsigned long long, eval[i]);
    if (index.emplace(key, static_cast<int>(vtab.size())).second) vtab.push_back(eval[i]);
    if (vtab.size() > 0x7FFF) return;
  }
  if (vtab.empty()) vtab.push_back(0.0);  // padding words read entry 0
  if (achr::resident_lds_bytes(n, static_cast<int>(vtab.size())) > 160 * 1024) return;
  // slices, widest first (their order in the table), each to the wavefront
  // with the fewest entries so far
  std::vector<std::vector<int>> mine(kCheckWaves);
  int load[kCheckWaves] = {};
  for (int sl = 0; sl < nslices; ++sl) {
    const int width = (eptr[sl + 1] - eptr[sl]) / 64;
    if (width == 0) continue;  // rows with no entries: residual 0, nothing to sum
    int best = 0;
    for (int w = 1; w < kCheckWaves; ++w)
      if (load[w] < load[best]) best = w;
    mine[best].push_back(sl);
    load[best] += width;
  }
  for (int w = 0; w < kCheckWaves; ++w)
    if (load[w] > achr::kCheckResident) return;
  for (int w = 0; w < kCheckWaves; ++w) {
    plan->woff[w] = static_cast<int>(plan->ent.size());
    plan->roff[w] = static_cast<int>(plan->rowmap.size());
    plan->wlen[w] = load[w];
    int j = 0;
    for (int sl : mine[w]) {
      const int width = (eptr[sl + 1] - eptr[sl]) / 64;
      for (int e = 0; e < width; ++e, ++j) {
        for (int l = 0; l < 64; ++l) {
          const size_t at = static_cast<size_t>(eptr[sl]) + 64 * static_cast<size_t>(e) + l;
          const int col = ecol[at];
          if (col < 0) {
            plan->ent.push_back(achr::kCheckPad);
          } else {
            const int vi = index[__builtin_bit_cast(unsigned long long, eval[at])];
            plan->ent.push_back(static_cast<unsigned>(col) | (static_cast<unsigned>(vi) << 16));
          }
        }
        if (e == width - 1) plan->wends[w] |= 1ull << j;
      }
      for (int l = 0; l < 64; ++l) plan->rowmap.push_back(erow[static_cast<size_t>(sl) * 64 + l]);
    }
  }
  if (plan->ent.empty()) plan->ent.push_back(achr::kCheckPad);
  if (plan->rowmap.empty()) plan->rowmap.push_back(-1);
  plan->vtab = std::move(vtab);
  plan->ok = true;
}

}  // namespace achr_host
