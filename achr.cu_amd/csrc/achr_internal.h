// Internal declarations shared by the kernels and the C ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace achr {

// Arguments of the chain-step kernel. All pointers are device memory.
struct StepArgs {
  // model
  int n, m, W, d;
  const double* warm;    // n x W, point k at warm + k*n
  const double* lb;
  const double* ub;
  const double* c0;      // warmup centre
  const int* rowptr;     // CSR of S (m rows)
  const int* colidx;
  const double* vals;
  const double* Ncm;     // n x d column major: N[i + j*n]
  const double* Nrm;     // n x d row major:    N[i*d + j]
  // run
  int P, S;
  uint64_t seed;
  uint32_t file, model;
  int chain0;
  double dTol, uTol, maxMinTol, reprojTol;
  int check_every, max_misses;
  int mode;              // 1: reprojection inside the kernel; 2: stop at checks (batched GEMM)
  // state, chain major
  double* X;             // P x n: x of chain b at X + b*n
  double* C;             // P x n: running centre
  double* T;             // P x d: N^T x scratch (fused mode without LDS)
  int4* st;              // {attempt, s, phase, consecutive misses}
  int4* stats;           // {misses, reprojections, checks, flags}
  int* flag;             // batched mode: 1 if chain b asks for a reprojection this round
  float* R32;            // batched mode, optional: residual S x of each check, P x m
  double* out;           // n x P reaction major
  double* trace;
  int trace_chains, trace_thin, trace_len;
};

// Phases stored in st.z
enum Phase : int { kFresh = 0, kRunning = 1, kPendingCheck = 2, kDone = 3 };

// Launch shape chosen for a model size.
struct Schedule {
  int tpc;   // threads per chain
  int e;     // elements per thread (register slots)
  int cpb;   // chains per block
};

bool schedule_supported(int tpc, int e);
int elems_for(int n, int tpc);  // smallest instantiated E >= ceil(n / tpc), or -1
size_t step_lds_bytes(const StepArgs& a, const Schedule& s);
hipError_t launch_step(const StepArgs& a, const Schedule& s, hipStream_t stream);

// Batched reprojection: X[idx[q]] <- N (N^T X[idx[q]]) for q < *count.
hipError_t launch_compact(const int* flag, int P, int* idx, int* count, hipStream_t stream);
hipError_t launch_project(const double* Ncm, int n, int d, double* X, double* T, const int* idx,
                          const int* count, int P, hipStream_t stream);
// Elementwise FP64 -> FP32 copy of `count` values.
hipError_t launch_to_f32(const double* src, float* dst, size_t count, hipStream_t stream);

}  // namespace achr
