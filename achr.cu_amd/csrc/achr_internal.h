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
  // The same rows as a sliced ELL table for chains of 64 or more lanes: rows
  // sorted by length, 64 to a slice, entry e of the slice's row l at
  // ell_ptr[s] + 64 e + l (column -1 for padding), ell_row[64 s + l] the
  // row's index in S (-1 past m).
  int nslices;
  const int* ell_ptr;
  const int* ell_col;
  const double* ell_val;
  const int* ell_row;
  const double* Ncm;     // n x d column major: N[i + j*n]
  const double* Nrm;     // n x d row major:    N[i*d + j]
  // run
  int P, S;
  uint64_t seed;
  uint32_t file, model;  // Philox counter words 2 and 3 (model << 4 | tag)
  int chain0;
  double dTol, uTol, maxMinTol, reprojTol;
  int check_every, max_misses;
  int mode;              // 1: reprojection inside the kernel; 2: stop at checks (batched GEMM)
  int stop_flagged_only; // mode 2: a chain stops only at a check that flags it, or
  int check_budget;      //   after check_budget checks in one launch (launch balance)
  int check_out;         // mode 2: every chain stops at its check with x stored; the
                         //   residual is left to launch_check
  // state, chain major
  double* X;             // P x n: x of chain b at X + b*n
  double* C;             // P x n: running centre
  double* T;             // P x d: N^T x scratch (fused mode without LDS)
  int4* st;              // {attempt, s, phase, consecutive misses}
  int4* stats;           // {misses, reprojections, checks, flags}
  int* flag;             // batched mode: 1 if chain b asks for a reprojection this round
  float* R32;            // batched mode, optional: residual S x of each check, row b at R32 + b*r32_ld
  long r32_ld;
  double* out;           // n x P reaction major
  int dir_mode;          // 1: warm holds a bank of unit directions (no centre)
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
size_t step_lds_bytes(bool dir_mode, const Schedule& s);
hipError_t launch_step(const StepArgs& a, const Schedule& s, hipStream_t stream);
// Multi-model persistent form. prepare_step_batch opts the kernel of this
// schedule into its LDS size and returns how many of its workgroups stay
// resident on the device; launch_step_batch runs `grid` workgroups over
// `nitems` items {model index into tab, first chain} of s.cpb chains each.
hipError_t prepare_step_batch(const Schedule& s, bool dir_mode, int* resident);
hipError_t launch_step_batch(const StepArgs* tab, const int2* items, int nitems, int grid, const Schedule& s,
                             bool dir_mode, hipStream_t stream);

// Residual checks of the chains a segment stopped at their check (check_out):
// for every chain b in phase kPendingCheck, max_j |(S x_b)_j| from the sliced
// ELL table of S over x_b = X + b*n, the flag (residual > reprojTol and d > 0),
// the reprojection counter stats[b].y and, if R32 is set, the FP32 row sums.
constexpr int kMaxCheckReactions = 16384;
// The resident form: the slices of S dealt to the kernel's 16 wavefronts
// (longest first, each to the least loaded wavefront), each wavefront's
// entries held in registers for the whole launch as packed words
// col | value index << 16 (bit 31: padding), the values in an LDS table.
constexpr int kCheckWaves = 16;
constexpr int kCheckResident = 40;  // entries per lane a wavefront can hold
constexpr unsigned kCheckPad = 0x80000000u;
struct CheckArgs {
  int n, d, P;
  int nslices;
  const int* ell_ptr;
  const int* ell_col;
  const double* ell_val;
  const int* ell_row;
  const double* X;
  const int4* st;
  int4* stats;
  int* flag;
  float* R32;
  long r32_ld;
  double reprojTol;
  // resident form (ent != nullptr): wavefront w holds wlen[w] entries per
  // lane, entry j of lane l at ent[woff[w] + 64 j + l]; bit j of wends[w]
  // marks the last entry of a slice; the S row of lane l in the k-th slice
  // of wavefront w is rowmap[roff[w] + 64 k + l] (-1: none)
  const unsigned* ent;
  const int* rowmap;
  const double* vtab;
  int nv;
  int wlen[kCheckWaves], woff[kCheckWaves], roff[kCheckWaves];
  unsigned long long wends[kCheckWaves];
};
size_t check_lds_bytes(int n);
size_t resident_lds_bytes(int n, int nv);  // LDS of the resident form
hipError_t launch_check(const CheckArgs& a, hipStream_t stream);

// Batched reprojection: X[idx[q]] <- N (N^T X[idx[q]]) for q < *count.
hipError_t launch_compact(const int* flag, int P, int* idx, int* count, hipStream_t stream);
hipError_t launch_project(const double* Ncm, int n, int d, double* X, double* T, const int* idx,
                          const int* count, int P, hipStream_t stream);
// C(:, q) = A B(:, q) in FP64, k ascending from +0.0 (bitwise sequential):
// A(r, k) = A[r*sar + k*sak] (M x K), column q of B at B + q*ldb, column q of C at C + q*ldc.
hipError_t launch_dgemm_cols(int M, int K, const double* A, long sar, long sak, const double* B, long ldb,
                             double* C, long ldc, const int* count, int max_cols, hipStream_t stream);
// Normalises each of the K columns (length n) of A to unit Euclidean norm.
hipError_t launch_normalize_cols(double* A, int n, int K, hipStream_t stream);
// Column means of the rows x cols row-major FP64 matrix X (mean[j]), summed
// over kColMeanStripes row stripes; work holds kColMeanStripes * cols doubles.
constexpr int kColMeanStripes = 64;
hipError_t launch_col_means(const double* X, long rows, int cols, double* mean, double* work, hipStream_t stream);
// Y[r, j] -= v[j] for the rows x cols row-major FP32 matrix Y.
hipError_t launch_sub_row(float* Y, long rows, int cols, const float* v, hipStream_t stream);
// Standard normal FP32 values from the Philox stream.
hipError_t launch_gauss(float* z, size_t count, uint64_t seed, uint32_t stream_id, hipStream_t stream);
// y = alpha * y + beta * x (FP32), x and y of `count` values.
hipError_t launch_axpby(float* y, const float* x, float alpha, float beta, size_t count, hipStream_t stream);
// FP32 -> FP64 copy.
hipError_t launch_to_f64(const float* src, double* dst, size_t count, hipStream_t stream);
// Sum of squares of `count` FP32 values, accumulated in FP64 (synchronous).
hipError_t sum_squares(const float* x, size_t count, double* out);
// Y = X - v 1^T over the cols columns of the n x cols matrix X.
hipError_t launch_sub_cols(const double* X, const double* v, double* Y, int n, int cols, hipStream_t stream);
// X += v 1^T, then rows with lb == ub set to lb.
hipError_t launch_add_cols_fix(double* X, const double* v, const double* lb, const double* ub, int n, int cols,
                               hipStream_t stream);
// S+ = S^T G (n x m column major, FP32) from the CSC form of S (colptr n+1,
// rowidx / vals per column) and the symmetric m x m FP64 matrix G.
// B[i + r*n] = sum of the entries of S at (r, i): the dense S^T (n x m,
// column major) from the CSR rows, B zeroed by the caller.
hipError_t launch_dense_st(int m, int n, const int* rowptr, const int* colidx, const double* vals, double* B,
                           hipStream_t stream);
hipError_t launch_pinv_csc(int n, int m, const int* colptr, const int* rowidx, const double* vals, const double* G,
                           float* out, hipStream_t stream);
// Elementwise FP64 -> FP32 copy of `count` values.
hipError_t launch_to_f32(const double* src, float* dst, size_t count, hipStream_t stream);
// q[i] = a[i] / b[i] by the plain chord's reciprocal-and-correction (device arrays).
hipError_t launch_quotient(const double* a, const double* b, double* q, size_t count, hipStream_t stream);
// Read-only stream over `words` 16-byte words of `buf` (a grid-stride sum
// per thread, one double per workgroup into sums[blocks]); blocks = enough
// resident workgroups to fill the device.
hipError_t read_stream_blocks(int* blocks);
hipError_t launch_read_stream(const double* buf, size_t words, double* sums, int blocks, hipStream_t stream);
// r[i] = the bank-mode chord's reciprocal of u[i] (device arrays).
hipError_t launch_bank_reciprocal(const double* u, double* r, size_t count, hipStream_t stream);

}  // namespace achr
