// Host-side handle types and error helpers shared by the C ABI sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <new>
#include <algorithm>
#include <stdexcept>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/achr.h"
#include "achr_internal.h"
#include "achr_model.h"

namespace achr_host {

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), count(o.count) {
    o.p = nullptr;
    o.count = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      count = o.count;
      o.p = nullptr;
      o.count = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    count = 0;
  }
  hipError_t ensure(size_t n) {
    if (n <= count && p) return hipSuccess;
    release();
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T));
    if (e == hipSuccess) count = n;
    return e;
  }
  hipError_t upload(const T* h, size_t n, hipStream_t s) {
    hipError_t e = ensure(n);
    if (e != hipSuccess || n == 0) return e;
    e = hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
  }
};

}  // namespace achr_host

#define HIP_TRY(expr)                                                                                      \
  do {                                                                                                     \
    hipError_t e_ = (expr);                                                                                \
    if (e_ != hipSuccess) return achr_host::fail(ACHR_E_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

struct achr_sampler {
  const achr_model* model = nullptr;
  int n = 0, m = 0, W = 0, d = 0;
  achr_options opt{};
  achr::Schedule sched{};
  // Schedule of the last run: its lanes fix the norm's reduction tree, so a
  // continuation or checkpoint of that run keeps them (a batch may have run
  // the sampler's chains with other lanes than sched).
  achr::Schedule run_sched{};
  int mode = 1;
  hipStream_t stream = nullptr;
  std::vector<double> c0;
  achr_host::DevBuf<double> warm, lb, ub, c0d, vals, Ncm, Nrm;
  achr_host::DevBuf<int> rowptr, colidx;
  achr_host::DevBuf<int> ell_ptr, ell_col, ell_row;  // S as a sliced ELL table (StepArgs)
  achr_host::DevBuf<double> ell_val;
  int nslices = 0;
  // resident check plan (achr_host::CheckPlan, device copies)
  bool chk_ok = false;
  achr_host::DevBuf<unsigned> chk_ent;
  achr_host::DevBuf<int> chk_row;
  achr_host::DevBuf<double> chk_vtab;
  int chk_nv = 0;
  int chk_wlen[achr::kCheckWaves] = {}, chk_woff[achr::kCheckWaves] = {}, chk_roff[achr::kCheckWaves] = {};
  unsigned long long chk_wends[achr::kCheckWaves] = {};
  // reduced-precision projection operands (FP32 copies)
  achr_host::DevBuf<unsigned char> Bpk;  // pseudo-inverse S+ packed for the residual rounds (launch_pack_b)
  int r32_ld = 0;                        // row stride of R32: packed_k(m), zero beyond m
  achr_host::DevBuf<float> N32;       // N, n x d column major
  achr_host::DevBuf<float> R32, T32;  // per-chain residuals (P x r32_ld) and N^T x (P x d)
  int rounds_fp64 = 0, rounds_low = 0;
  // direction bank (learned sampler)
  achr_host::DevBuf<double> dirs;     // n x K column major, unit columns
  int K = 0;
  int use_dirs = 0;
  std::vector<double> s0;             // start of bank-mode chains (empty until computed)
  achr_host::DevBuf<double> s0d;
  // run state
  int P = 0, S = 0;
  uint64_t seed = 0;                  // of the last run (achr_sample_continue repeats them)
  uint32_t file = 0;
  int chain0 = 0;
  int run_use_dirs = 0, run_check_every = 0, run_trace_thin = 0, run_trace_chains = 0;
  achr_host::DevBuf<double> X, C, T, out, trace;
  achr_host::DevBuf<int4> st, stats;
  achr_host::DevBuf<int> flag, idx, count;
  int trace_len = 0;
  int trace_chains_run = 0;
  // timing
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_step, ev_proj, ev_check;  // indices into ev_pool
  bool have_run = false;
};

namespace achr_host {

// Resident form of the check kernel for a sliced ELL table of S (CheckArgs).
struct CheckPlan {
  bool ok = false;  // the table fits: entries per wavefront, value count, LDS
  std::vector<unsigned> ent;
  std::vector<int> rowmap;
  std::vector<double> vtab;
  int wlen[achr::kCheckWaves] = {}, woff[achr::kCheckWaves] = {}, roff[achr::kCheckWaves] = {};
  unsigned long long wends[achr::kCheckWaves] = {};
};
// eptr/erow/ecol/eval: the sliced ELL table (StepArgs); n reactions.
void build_check_plan(int n, int nslices, const std::vector<int>& eptr, const std::vector<int>& erow,
                      const std::vector<int>& ecol, const std::vector<double>& eval, CheckPlan* plan);

// Run set-up and projection rounds shared by single-model runs and batches.
enum { kProjNone = 0, kProjFP64 = 1, kProjLow = 2 };
// Rejects run sizes a schedule cannot address.
int run_limits(const achr_sampler* S, const achr::Schedule& sc, int P, int Ssteps);
// Kernel launches (segments) of a run: 1 in mode 1, one per check + 1 in mode 2.
int segments_for(const achr_sampler* S, int Ssteps);
// Sizes the run buffers for P chains, queues the reset of the chain state on
// st and fills the kernel arguments.
int prepare_run(achr_sampler* S, int P, int Ssteps, uint64_t seed, uint32_t file, int chain0, hipStream_t st,
                achr::StepArgs* args);
// Kernel arguments of a run of Ssteps steps over the sampler's current buffers.
void fill_args(const achr_sampler* S, int P, int Ssteps, uint64_t seed, uint32_t file, int chain0,
               achr::StepArgs* args);
// Queues the segments [0, segments) of a run with the schedule S->run_sched
// (step launch, then the projection round of checks already done + segment
// index), recording their events.
int enqueue_segments(achr_sampler* S, const achr::StepArgs& a, int segments, int checks_done, hipStream_t st);
// Whether batched rounds of this schedule leave the residual check to
// launch_check (StepArgs::check_out), and that kernel's arguments for a run.
bool use_check_kernel(const achr_sampler* S, const achr::Schedule& sc);
achr::CheckArgs check_args(const achr_sampler* S, const achr::StepArgs& a);
// Projection round after segment seg (kProjNone outside mode 2).
int projection_kind(const achr_sampler* S, int seg);
// Queues one projection round over the chains flagged by the last segment.
int enqueue_projection(achr_sampler* S, int kind, hipStream_t st);
// Records the settings of the run just queued (directions, check period,
// trace, schedule) that achr_sample_continue and the checkpoint must match.
void note_run(achr_sampler* S, const achr::Schedule& sc);

// Runs f(t, T) for t = 0 .. T-1, T = min(16, hardware threads), on host
// threads (t = 0 on the caller); a part whose thread cannot be started runs
// on the caller.
template <class F>
void on_threads(F&& f) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int T = static_cast<int>(std::max(1u, std::min(16u, hw)));
  std::vector<std::thread> pool;
  int started = 1;
  try {
    for (int t = 1; t < T; ++t) {
      pool.emplace_back(f, t, T);
      ++started;
    }
  } catch (const std::exception&) {
  }
  for (int t = started; t < T; ++t) f(t, T);
  f(0, T);
  for (auto& th : pool) th.join();
}

}  // namespace achr_host
