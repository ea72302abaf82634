// achr_sample_noalloc: one file of achr_sample for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). The chain state the step kernel and the
// FP64 projection rounds work on (points and centres, N^T x, phases,
// counters, round flags and the compacted index list) lives in one
// file-scope device array instead of buffers the sampler grows on demand;
// the caller passes the output buffers. No hipMalloc, hipFree or allocating
// library call is made, and no event is created or recorded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/achr.h"
#include "achr_host.h"
#include "achr_internal.h"

using achr_host::fail;

namespace {
constexpr size_t kScratchBytes = size_t(64) << 20;  // per GPU, held until the process ends
constexpr int kMaxDevices = 64;
}  // namespace

// The scratch array. In host code its name is a host-side stand-in, not a
// device address: only the address hipGetSymbolAddress returns for the
// current device (scratch_on_device) may reach a kernel or a copy.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The device address of the scratch array on the current device, looked up
// once per device (with that device current) and kept.
int scratch_on_device(unsigned char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ACHR_E_DEVICE, "device %d beyond the %d the library tracks", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(ACHR_E_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_ptr[dev]);
  return ACHR_OK;
}

__global__ __launch_bounds__(256) void zero_words_kernel(unsigned* __restrict__ p, size_t count) {
  for (size_t i = blockIdx.x * static_cast<size_t>(256) + threadIdx.x; i < count;
       i += static_cast<size_t>(gridDim.x) * 256)
    p[i] = 0u;
}

size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

}  // namespace

extern "C" {

int achr_sample_noalloc(achr_sampler* S, int P, int Ssteps, uint64_t seed, uint32_t file, int chain0, double* out,
                        int32_t* stats) {
  if (!S || !out) return fail(ACHR_E_ARG, "NULL argument");
  if (P <= 0 || Ssteps < 0 || chain0 < 0) return fail(ACHR_E_ARG, "bad P=%d S=%d chain0=%d", P, Ssteps, chain0);
  if (S->opt.proj_prec != ACHR_PROJ_FP64)
    return fail(ACHR_E_ARG, "achr_sample_noalloc runs FP64 projection rounds only (proj_prec %d)", S->opt.proj_prec);
  const int n = S->n, d = S->d;
  // carve the scratch: X, C (P x n), T (P x max(d,1)), st, stats (P int4),
  // flag, idx (P int), count (1 int), each part 256-byte aligned
  const size_t b_x = align256(sizeof(double) * static_cast<size_t>(P) * n);
  const size_t b_t = align256(sizeof(double) * static_cast<size_t>(P) * std::max(d, 1));
  const size_t b_st = align256(sizeof(int4) * static_cast<size_t>(P));
  const size_t b_i = align256(sizeof(int) * static_cast<size_t>(P));
  const size_t need = 2 * b_x + b_t + 2 * b_st + 2 * b_i + 256;
  if (need > kScratchBytes)
    return fail(ACHR_E_ARG, "P=%d chains of n=%d need %zu bytes of scratch; achr_sample_noalloc holds %zu "
                "(P * (16 n + 8 max(d, 1) + 40) + 2 KiB at most)", P, n, need, kScratchBytes);
  int rc = achr_host::run_limits(S, S->sched, P, Ssteps);
  if (rc != ACHR_OK) return rc;
  unsigned char* base = nullptr;
  rc = scratch_on_device(&base);
  if (rc != ACHR_OK) return rc;
  unsigned char* q = base;
  auto take = [&](size_t bytes) {
    unsigned char* r = q;
    q += bytes;
    return r;
  };
  double* X = reinterpret_cast<double*>(take(b_x));
  double* Cc = reinterpret_cast<double*>(take(b_x));
  double* T = reinterpret_cast<double*>(take(b_t));
  int4* st = reinterpret_cast<int4*>(take(b_st));
  int4* stt = stats ? reinterpret_cast<int4*>(stats) : reinterpret_cast<int4*>(take(b_st));
  int* flag = reinterpret_cast<int*>(take(b_i));
  int* idx = reinterpret_cast<int*>(take(b_i));
  int* count = reinterpret_cast<int*>(take(256));
  hipStream_t stream = S->stream;
  // phases (fresh = 0), counters and flags start at zero: a fill kernel
  // rather than hipMemsetAsync, so a caller may capture the call in a graph
  const size_t words = static_cast<size_t>(P) * 4;
  const unsigned blocks = static_cast<unsigned>(std::min<size_t>((words + 255) / 256, 4096));
  hipLaunchKernelGGL(zero_words_kernel, dim3(blocks), dim3(256), 0, stream, reinterpret_cast<unsigned*>(st), words);
  hipLaunchKernelGGL(zero_words_kernel, dim3(blocks), dim3(256), 0, stream, reinterpret_cast<unsigned*>(stt), words);
  hipLaunchKernelGGL(zero_words_kernel, dim3(blocks), dim3(256), 0, stream, reinterpret_cast<unsigned*>(flag),
                     static_cast<size_t>(P));
  HIP_TRY(hipGetLastError());
  achr::StepArgs a{};
  achr_host::fill_args(S, P, Ssteps, seed, file, chain0, &a);
  a.X = X;
  a.C = Cc;
  a.T = T;
  a.st = st;
  a.stats = stt;
  a.flag = flag;
  a.R32 = nullptr;
  a.out = out;
  a.trace = nullptr;
  a.trace_chains = 0;
  a.trace_len = 0;
  const int segments = achr_host::segments_for(S, Ssteps);
  for (int seg = 0; seg < segments; ++seg) {
    hipError_t e = achr::launch_step(a, S->sched, stream);
    if (e != hipSuccess) return fail(ACHR_E_DEVICE, "step kernel launch: %s", hipGetErrorString(e));
    if (S->mode == 2 && d > 0 && seg + 1 < segments) {
      e = achr::launch_compact(flag, P, idx, count, stream);
      if (e == hipSuccess) e = achr::launch_project(S->Ncm.p, n, d, X, T, idx, count, P, stream);
      if (e != hipSuccess) return fail(ACHR_E_DEVICE, "projection launch: %s", hipGetErrorString(e));
    }
  }
  return ACHR_OK;
}

}  // extern "C"
