/*
 * achr.h — C ABI of the MI355X ACHR sampler (libachr.so).
 *
 * The reference has no library API: its only boundary is the CLI
 *   ACHRCuda model.mps warmup.csv nFiles pointsPerFile stepsPerPoint
 * (marouenbg/ACHR.cu ACHRcu/ACHRCuda.cu:552-554) and, inside it, the
 * device entry stepPointProgress(...) (ACHRCuda.cu:361) that main launches
 * once per output file (:741). The functions below are that seam, made
 * callable: the CLI shipped with this library (ACHRCuda) is built on them,
 * and any host language binds them with plain pointers and sizes.
 *
 * Conventions
 *   - Every function returns ACHR_OK (0) or a negative ACHR_E_* code and
 *     never exits the process (the reference exits or asserts:
 *     ACHRcu/Utilities.cu:28-37, :76-85). achr_last_error() gives the text
 *     of the last failure on the calling thread.
 *   - Matrices are column major unless stated: the warmup set is n x W with
 *     each point contiguous (ACHRCuda.cu:409, :712-713); the null-space basis
 *     is n x d (ACHRCuda.cu:486-489). Sample output is reaction major,
 *     out[i*P + chain], the layout of the reference's points buffer and of
 *     its File<k> text files (ACHRCuda.cu:198-204, :744-756).
 *   - Host pointers are caller-owned and only read during the call. Device
 *     memory is owned by the handles. A handle is used by one host thread at
 *     a time; distinct handles may run concurrently on distinct streams.
 */
#ifndef ACHR_H_
#define ACHR_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ACHR_OK 0
#define ACHR_E_ARG (-1)      /* invalid argument or size */
#define ACHR_E_IO (-2)       /* file missing or unreadable */
#define ACHR_E_FORMAT (-3)   /* malformed MPS / CSV content */
#define ACHR_E_DEVICE (-4)   /* HIP runtime or kernel launch failure */
#define ACHR_E_SOLVER (-5)   /* rocSOLVER / null-space failure */
#define ACHR_E_STATE (-6)    /* call out of order (e.g. no null space yet) */
#define ACHR_E_NOMEM (-7)    /* host or device allocation failed */

typedef struct achr_model achr_model;
typedef struct achr_sampler achr_sampler;

/* Text of the last error on this thread ("" if none). */
const char *achr_last_error(void);
/* Library version string. */
const char *achr_version(void);

/* ------------------------------------------------------------- model (a1) */

/* Model from a sparse stoichiometric matrix S (m x n, coordinate format,
 * duplicates summed, explicit zeros dropped) and reaction bounds.
 * Replaces the dense-S + row-sorted COO assembly of ACHRCuda.cu:560-663;
 * S is kept as CSR (columns ascending within a row) on host and device. */
int achr_model_create(int n_rxns, int n_mets, int nnz, const int *rows, const int *cols,
                      const double *vals, const double *lb, const double *ub, achr_model **out);

/* Model from an MPS file as the reference's sampler sees it through
 * glp_read_mps(GLP_MPS_DECK) (ACHRCuda.cu:560-597): every row after the
 * objective (the first N row) is a row of S whatever its type, RHS and RANGES
 * are read and ignored (the sampled set is S x = 0), COLUMNS with one or two
 * entries per line, BOUNDS UP/LO/FX/FR/MI/PL/BV/LI/UI (default bounds
 * [0, +inf), an infinite bound is stored as +-DBL_MAX as GLPK reports it).
 * Free-format whitespace splitting, so fixed-format decks read the same. */
int achr_model_read_mps(const char *path, achr_model **out);

/* The same file with its rows as written, the view GLPK gives the reference's
 * warmup generator (createWarmupPtsGLPK.c:219): E rows S_i x = b_i, L rows
 * S_i x <= b_i, G rows S_i x >= b_i, RANGES giving the other side by the MPS
 * rules, free rows after the objective dropped. The model is returned as a
 * set {x' : S' x' = 0, lb' <= x' <= ub'} whose first n_file columns are the
 * file's and range over exactly its polytope: each row with two distinct
 * sides gets a slack column (S_i x - s_i = 0, bounds the row's sides, an
 * infinite side replaced by the row's activity range over the column bounds),
 * and equality rows with b_i != 0 take -b_i on one last column fixed at 1.
 * ACHR_E_FORMAT when a row is unbounded or unmeetable within the column
 * bounds. No counterpart in the reference's sampler, which ignores row types
 * and right-hand sides. */
int achr_model_read_mps_rows(const char *path, achr_model **out);

/* Number of leading columns that come from the file: n for every model but
 * those of achr_model_read_mps_rows. */
int achr_model_file_columns(const achr_model *model, int *n_file);

/* Fills the auxiliary columns of P points given by their file columns
 * (x_file: n_file doubles per point, points one after another; x: n doubles
 * per point): a slack takes its row's activity, the right-hand-side column
 * 1. achr_sampler_create_csv does this itself for a warmup file of n_file
 * lines. */
int achr_model_complete_points(const achr_model *model, const double *x_file, int P, double *x);

void achr_model_destroy(achr_model *model);

int achr_model_dims(const achr_model *model, int *n_rxns, int *n_mets, int *nnz, int *null_dim);

/* Copies bounds out (n doubles each). */
int achr_model_bounds(const achr_model *model, double *lb, double *ub);

/* Copies the CSR form of S out: rowptr (m+1), colidx (nnz), vals (nnz). */
int achr_model_csr(const achr_model *model, int *rowptr, int *colidx, double *vals);

/* ------------------------------------------------------- null space (a3) */

/* Null-space basis N of S computed on the GPU: eigenvectors of S^T S (FP64,
 * rocSOLVER syevd) whose eigenvalue is below n * lambda_max * 2.2204e-16 —
 * the singular-value rule of computeKernelCuda (ACHRCuda.cu:436-502, :474).
 * Writes d to *null_dim. */
int achr_model_compute_nullspace(achr_model *model, int *null_dim);

/* Supplies N (n x d, column major) computed elsewhere. d may be 0. */
int achr_model_set_nullspace(achr_model *model, const double *N, int d);

/* Copies N out (n x d column major). */
int achr_model_get_nullspace(const achr_model *model, double *N);

/* -------------------------------------------------------- sampler (a2-a14) */

typedef struct achr_options {
  double dTol;          /* 1e-14, ACHRCuda.cu:512 */
  double uTol;          /* 1e-9,  ACHRCuda.cu:508 */
  double maxMinTol;     /* 1e-9,  ACHRCuda.cu:505 */
  double reprojTol;     /* 1e-9,  ACHRCuda.cu:338 */
  int check_every;      /* 10,    ACHRCuda.cu:325 */
  int max_misses;       /* consecutive misses before a chain stops (reference: unbounded) */
  int lanes;            /* threads per chain: 0 = autotuned; else 16..1024, power of two */
  int reproject_mode;   /* 0 = auto, 1 = in-kernel per chain, 2 = batched GEMM rounds */
  int trace_chains;     /* chains whose thinned trajectory is recorded (0 = none) */
  int trace_thin;       /* record every trace_thin accepted steps */
  /* Arithmetic of the batched projection rounds (reproject_mode 2):
   *   ACHR_PROJ_FP64       x <- N (N^T x), FP64 FMA in the reference's loop
   *                        order (bit-identical to the oracle);
   *   ACHR_PROJ_RES_BF16X3 x <- x - S+ (S x): the FP64 residual S x of the
   *                        check, contracted with the pseudo-inverse S+ on the
   *                        matrix cores in split-BF16x3 with FP32 accumulation;
   *   ACHR_PROJ_RES_F32    the same contraction with FP32 MFMA;
   *   ACHR_PROJ_NULL_F32   x <- N (N^T x) with FP32 MFMA (x rounded to FP32);
   *   ACHR_PROJ_NULL_BF16X3 the same in split-BF16x3.
   * In exact arithmetic all five are the orthogonal projection onto null(S). */
  int proj_prec;
  int fp64_every;       /* with a reduced-precision proj_prec: every fp64_every-th
                           round uses ACHR_PROJ_FP64 (10 = every 100 steps) */
  int model_id;         /* Philox counter word 3 is model_id << 4 | tag: models
                           sampled together take distinct ids (0 .. 2^27 - 1) */
  int check_kernel;     /* where the residual check of batched rounds of plain
                           chains runs (reproject_mode 2): 0 = auto (its own
                           kernel for chains of 512 or more lanes), 1 = inside
                           the step kernel, 2 = its own kernel between
                           segments (every chain stops at its check; S held
                           in registers when it fits, else read from L2),
                           3 = its own kernel, S always read from L2. The
                           results are the same in every case. */
} achr_options;

#define ACHR_PROJ_FP64 0
#define ACHR_PROJ_RES_BF16X3 1
#define ACHR_PROJ_RES_F32 2
#define ACHR_PROJ_NULL_F32 3
#define ACHR_PROJ_NULL_BF16X3 4

/* Fills the reference defaults. */
void achr_options_default(achr_options *opts);

/* Sampler over a model with a null space and a warmup set (n x W column
 * major, each point contiguous). Computes the warmup centre on the host as
 * (1/W) * sum_k w_k (createCenterPt, ACHRCuda.cu:414-434) and uploads the
 * warmup set, bounds, S and N to device memory. opts may be NULL. The model
 * must outlive the sampler (learning a bank and bank mode read it). */
int achr_sampler_create(const achr_model *model, const double *warmup, int W, const achr_options *opts,
                        achr_sampler **out);

/* Same, reading the warmup CSV of the reference (one line per reaction, one
 * comma-separated column per point; computenWrmup/parseLine,
 * ACHRCuda.cu:387-412, :676-689). */
int achr_sampler_create_csv(const achr_model *model, const char *warmup_csv, const achr_options *opts,
                            achr_sampler **out);

void achr_sampler_destroy(achr_sampler *sampler);

/* Binds the sampler to a HIP stream (hipStream_t passed as void*; NULL = the
 * null stream). */
int achr_sampler_set_stream(achr_sampler *sampler, void *hip_stream);

/* Centre of the warmup set (n doubles). */
int achr_sampler_center(const achr_sampler *sampler, double *c0);

/* Runs one output file: P chains, each started at the warmup centre, each
 * making S accepted ACHR steps (stepPointProgress, ACHRCuda.cu:361-385 and
 * the per-file loop :736-743). Uniforms come from Philox4x32-10 keyed by
 * `seed` with counter {index, chain0 + chain, file, model_id << 4 | tag}. Results stay in
 * device memory until fetched; the call returns once the work is queued.
 */
int achr_sample(achr_sampler *sampler, int P, int S, uint64_t seed, uint32_t file, int chain0);

/* achr_sample for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * no counterpart in the reference. One file of P chains x S steps as
 * achr_sample runs it (same schedule, same Philox streams, same points and
 * counters, bit for bit), with the chain state in a static 64 MiB device
 * array that the library holds on each GPU it runs on until the process ends,
 * instead of buffers it allocates. The caller passes the outputs as device
 * memory: out (n x P doubles, reaction major) and stats (4 int32 per chain
 * as achr_sample_stats gives them, or NULL). The call makes no hipMalloc,
 * hipFree or other allocation, records no event and returns once the work
 * is queued on the sampler's stream. Limits: FP64 projection rounds only
 * (proj_prec ACHR_PROJ_FP64), no trace, and
 *   P * (16 n + 8 max(d, 1) + 40) + 2 KiB <= 64 MiB
 * (e.g. 39,000 chains at n = 95, 322 at n = 10,600); larger runs are refused
 * with ACHR_E_ARG. Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads, other samplers) race,
 * so a caller must order them. The run is not the sampler's last run:
 * achr_sample_points, _continue and _save_state keep referring to the last
 * achr_sample. */
int achr_sample_noalloc(achr_sampler *sampler, int P, int S, uint64_t seed, uint32_t file, int chain0, double *out,
                        int32_t *stats);

/* Continues every chain of the last run (or of a loaded state) for S_more
 * further accepted steps, as if the run had been started with S + S_more:
 * points, counters and traces equal those of the longer run bit for bit.
 * The run keeps its threads per chain, also when a batch chose them
 * (achr_batch_sample, achr_batch_set_lanes).
 * The directions (bank or ACHR), check period and trace settings must be
 * those of the run. No counterpart in the reference CLI (one fixed S per
 * file); it is the resume half of the MATLAB sampler's time-limited runs
 * (ACHRSamplerMOD.m:173-181, which save the last point and stop). */
int achr_sample_continue(achr_sampler *sampler, int S_more);

/* Checkpoint of the last run's chains (points, centres, counters, seed,
 * file, chain offset, step count and threads per chain) to a binary file,
 * and its reload into a sampler of the same model, warmup or bank, check
 * period, reprojection mode and model_id whose n the saved threads per chain
 * can hold (ACHR_E_STATE otherwise); after a reload
 * achr_sample_points returns the saved points and achr_sample_continue
 * resumes the chains exactly. Traces are not part of the state. */
int achr_sampler_save_state(const achr_sampler *sampler, const char *path);
int achr_sampler_load_state(achr_sampler *sampler, const char *path);

/* Waits for the last achr_sample and copies the points out: n x P reaction
 * major (out[i*P + chain]). */
int achr_sample_points(achr_sampler *sampler, double *out);

/* Device pointer (as void*) of the n x P reaction-major points of the last
 * run, valid until the next achr_sample or destroy. */
int achr_sample_points_device(achr_sampler *sampler, void **dev_ptr);

/* Per-chain counters of the last run, 4 int32 per chain: misses,
 * reprojections, residual checks, flags (bit 0: stopped after max_misses). */
int achr_sample_stats(achr_sampler *sampler, int32_t *stats);

/* Trajectory of the first trace_chains chains, every trace_thin accepted
 * steps: trace_chains x (S / trace_thin) x n doubles. */
int achr_sample_trace(achr_sampler *sampler, double *trace);

/* Milliseconds between HIP events recorded on the sampler's stream around
 * the last run's kernels, split as: total, step kernels, projection GEMMs.
 * Also the number of launches of each. */
int achr_sample_timing(achr_sampler *sampler, double *ms_total, double *ms_step, double *ms_proj,
                       int *n_step_launches, int *n_proj_launches);

/* Milliseconds and launches of the last run's residual-check kernels
 * (check_kernel 2; zero when the checks ran inside the step kernel). No
 * counterpart in the reference, whose check is a child launch per chain
 * (findMaxAbs, ACHRCuda.cu:232-240, :325-338). */
int achr_sample_timing_check(achr_sampler *sampler, double *ms_check, int *n_check_launches);

/* Schedule the autotuner picked for this sampler: threads per chain,
 * reprojection mode (1 in-kernel, 2 batched). */
int achr_sampler_schedule(const achr_sampler *sampler, int *lanes, int *reproject_mode);

/* Moves every warmup point onto the set the chains sample, {x : S x = 0,
 * x_i = lb_i for each fixed reaction (lb_i = ub_i)}: w <- s0 + N_f N_f^T
 * (w - s0), s0 a point of that set and N_f an orthonormal basis of its
 * directions (FP64), then recomputes the centre. ACHR directions w - c then
 * stay in that set, so the residual checks of a run flag only what the clip
 * to the bounds creates. The reference samples from its warmup file as read;
 * its "%f" text leaves every point off S x = 0 by its rounding, which is
 * what makes ACHRCuda reproject at every check (ACHRCuda.cu:325-343). */
int achr_sampler_project_warmup(achr_sampler *sampler);

/* Copies the sampler's warmup set out (n x W column major) and its W; with
 * warmup NULL only W. */
int achr_sampler_get_warmup(const achr_sampler *sampler, double *warmup, int *W);

/* ------------------------------------------------------------- batches */

/* Several samplers (one per model, any mix of reaction counts) run as one
 * job. Chains of the models that share a register schedule (threads per
 * chain, slots per thread) run in one launch per segment (one workgroup per
 * chain of 128-1,024 threads, persistent workgroups for smaller chains); a
 * segment of the whole batch (every such launch, then every model's
 * projection round) is captured once in a HIP graph and replayed, so a run
 * costs ceil(S / check_every) + 1 graph launches however many models it
 * holds. Every model keeps its own Philox stream (its options' model_id), so
 * its samples equal those of achr_sample on that sampler alone with the same
 * lanes, seed and file. Results are read per sampler (achr_sample_points,
 * _stats, _trace). No counterpart in the reference, which runs one model per
 * process with one child launch per step (ACHRCuda.cu:308-355). */
typedef struct achr_batch achr_batch;

/* The samplers stay owned by the caller and must outlive the batch; the
 * batch starts with each sampler's own schedule. */
int achr_batch_create(achr_sampler *const *samplers, int count, achr_batch **out);
void achr_batch_destroy(achr_batch *batch);

/* Work goes to this stream (hipStream_t as void*, not NULL: graphs are
 * captured on it). By default the batch owns a non-blocking stream. */
int achr_batch_set_stream(achr_batch *batch, void *hip_stream);

/* Autotuner: for every size class (n in [2^k, 2^(k+1))) times each candidate
 * threads-per-chain on probe_steps steps of P[k] chains of every model of the
 * class and keeps the fastest. ms, if not NULL, receives count x 7 probe times
 * (candidates 16, 32, ..., 1024 lanes; -1 where not a candidate). Chain
 * states of the probe are discarded. */
int achr_batch_autotune(achr_batch *batch, const int *P, int probe_steps, uint64_t seed, double *ms);

/* Threads per chain of every model (count ints); set_lanes pins them
 * (0 = the sampler's own schedule). */
int achr_batch_lanes(const achr_batch *batch, int *lanes);
int achr_batch_set_lanes(achr_batch *batch, const int *lanes);

/* One file of every model: P[k] chains of model k, S accepted steps each.
 * use_graph = 0 launches the same work eagerly. Returns when the run ends. */
int achr_batch_sample(achr_batch *batch, const int *P, int S, uint64_t seed, uint32_t file, int use_graph);

/* Milliseconds of the last run (events around all its work), its segments,
 * buckets (persistent launches per segment), distinct graphs and their
 * total node count. */
int achr_batch_timing(const achr_batch *batch, double *ms_total, int *segments, int *buckets, int *graphs,
                      int *graph_nodes);

/* ---------------------------------------- learned direction distribution */

/* Hit-and-run along a fixed bank of unit directions, each in the free null
 * space (S b = 0, b_i = 0 for fixed reactions), drawn uniformly: a symmetric,
 * state-independent direction law, so the uniform law on the polytope stays
 * stationary. It replaces ACHR's warmup-minus-centre direction (and its
 * adaptive, non-Markovian centre); bounds, chord, step, clip and the periodic
 * FP64 residual check are those of the reference loop. No counterpart in the
 * reference: it is the learned sampler of this library. */

typedef struct achr_dirnet achr_dirnet;

/* 3-layer MLP latent -> width -> width -> outdim (GELU), odd by
 * construction: v(z) = F(z) - F(-z). FP32 parameters, Philox-initialised. */
int achr_dirnet_create(int latent, int width, int outdim, uint64_t seed, achr_dirnet **out);
void achr_dirnet_destroy(achr_dirnet *net);
int achr_dirnet_param_count(const achr_dirnet *net, long *count);
int achr_dirnet_get_params(const achr_dirnet *net, float *params);
int achr_dirnet_set_params(achr_dirnet *net, const float *params);
/* Target second moment C (outdim x outdim, row major). */
int achr_dirnet_set_target(achr_dirnet *net, const float *C);
/* Loss |V^T V / B - C|_F^2 / |C|_F^2 on the given latents (B x latent) and
 * its gradient with respect to every parameter (param_count floats). */
int achr_dirnet_loss_grad(achr_dirnet *net, const float *Z, int B, float *loss, float *grads);
/* Adam on fresh Gaussian latents (Philox); losses[it] if not NULL. */
int achr_dirnet_train(achr_dirnet *net, int iters, int batch, float lr, uint64_t seed, float *losses);
/* V = v(Z) for B latents; prec 0 FP32, 2 BF16 matrix cores. */
int achr_dirnet_sample(achr_dirnet *net, const float *Z, int B, int prec, float *V);

typedef struct achr_learn_options {
  int rank;         /* principal subspace the network models (e.g. 512) */
  int width;        /* hidden width (512) */
  int iters;        /* Adam iterations */
  int batch;        /* latents per iteration */
  float lr;
  int bank_size;    /* directions in the bank */
  int isotropic;    /* 1: no network, isotropic directions in the free null space */
  uint64_t seed;
} achr_learn_options;

void achr_learn_options_default(achr_learn_options *opts);

/* Builds the direction bank from the samples of the last run's trace
 * (trace_chains x S/trace_thin points): their covariance in an orthonormal
 * basis of the free null space, reduced to its top `rank` principal
 * directions (randomised range finder), is the network's target; bank
 * directions are N_f (Q v + sigma_perp (I - Q Q^T) g), v from the trained
 * network, g Gaussian, assembled in FP64 and normalised. report (8 doubles):
 * free null dimension, samples used, loss before and after training, sigma_perp^2,
 * captured variance fraction, setup ms, training ms. Then switches the sampler to
 * bank directions (achr_sampler_use_directions(s, 1)). */
int achr_sampler_learn_directions(achr_sampler *sampler, const achr_learn_options *opts, double *report);

/* Changes the per-run settings of an existing sampler: trace_chains,
 * trace_thin (0, 0 = no trace) and check_every (residual check period). */
int achr_sampler_set_run_options(achr_sampler *sampler, int trace_chains, int trace_thin, int check_every);

/* Supplies the bank directly (n x K column major, host array). */
int achr_sampler_set_directions(achr_sampler *sampler, const double *bank, int K);
/* Copies the sampler's bank out (n x K column major) and its K; with bank
 * NULL only K. A learned bank saved this way can be reloaded with
 * achr_sampler_set_directions, skipping the training. */
int achr_sampler_get_directions(const achr_sampler *sampler, double *bank, int *K);
/* 1: bank directions; 0: ACHR directions (the reference's). Bank-mode
 * chains start from the point of {x : S x = 0, fixed reactions at their
 * bound} nearest the warmup centre (computed on the host when the bank is
 * first switched on), ACHR chains from the warmup centre itself. */
int achr_sampler_use_directions(achr_sampler *sampler, int on);
/* The point every chain of the next run starts from (n doubles): the warmup
 * centre with ACHR directions, the bank-mode start with bank directions. No
 * counterpart in the reference (its chains all start at the centre). */
int achr_sampler_start_point(const achr_sampler *sampler, double *x0);

/* -------------------------------------------------------------- file I/O */

/* Writes File<k>-style text: n lines of P comma-separated "%f" values
 * (ACHRCuda.cu:744-756). Rows are formatted by up to 16 host threads; the
 * bytes are those of a sequential fprintf. */
int achr_write_points_csv(const char *path, const double *points, int n, int P);

/* ------------------------------------------------------- warmup points */

typedef struct {
  int max_iters;    /* PDHG iterations at most (default 40,000) */
  int check_every;  /* iterations between restart / convergence checks (64) */
  double tol;       /* per LP, on the rescaled problem: |A x| and the duality gap relative to
                       1 + |objectives| (1e-9) */
  uint64_t seed;    /* random objectives (1) */
  int polish;       /* 1 (default): LPs still unconverged at max_iters are finished exactly on the host
                       by a primal active-set method in null-space coordinates (models with n x d up to
                       3e7); 0: left as the solver stopped */
} achr_warmup_options;

void achr_warmup_options_default(achr_warmup_options *opts);

/* The reference's warmup generator (VFWarmup/createWarmupPtsGLPK.c:67-182,
 * :283-297) on the device: nPts LPs (nPts even, >= 2n) min / max c^T x
 * subject to S x = 0, lb <= x <= ub, with c = e_i for the first n pairs (flux
 * variability) and a random c in [-0.5, 0.5)^n per LP after that, all solved
 * at once by a batched primal-dual hybrid gradient method (PDLP restarts);
 * each solution clipped to the bounds and moved toward the points' mean,
 * x <- 0.33 x + 0.67 mean. warmup: n x nPts column major (each point
 * contiguous, the sampler's layout). fva (n x 2, may be NULL): the minimum and
 * maximum of each reaction. For a model from achr_model_read_mps_rows the
 * objectives range over its n_file file columns only (nPts >= 2 n_file, the
 * reference's LPs) and fva holds NaN for the auxiliary columns. converged (nPts ints, may be NULL): 1 for each LP
 * whose iterate met tol, 0 for one stopped by max_iters (a first-order method:
 * its value is approximate, its point only approximately on S x = 0).
 * report (6 doubles, may be NULL): iterations, LPs converged (after the
 * polish), solve milliseconds, |A~|_2 of the rescaled matrix, LPs finished by
 * the polish, polish milliseconds. The reference solves one LP at
 * a time with GLPK's simplex over MPI ranks. */
int achr_warmup_generate(const achr_model *model, int nPts, const achr_warmup_options *opts, double *warmup,
                         double *fva, int *converged, double *report);

/* The reference generator's raw FVA file (createWarmupPtsGLPK.c:253-268):
 * "reaction,min,max" then one "name,%f,%f" line per reaction (names from the
 * MPS file; "?" for models built from arrays), the file's columns only.
 * fva: n x 2 as returned by achr_warmup_generate. */
int achr_write_fva_csv(const char *path, const achr_model *model, const double *fva);

/* Lossless counterpart of achr_write_points_csv (whose "%f" keeps six
 * decimals): the 8 bytes "ACHRPT01", int32 n, int32 P, then the n x P
 * doubles, reaction major, little endian. No counterpart in the reference. */
int achr_write_points_bin(const char *path, const double *points, int n, int P);
/* Reads such a file into a newly malloc'ed n x P array (release with achr_free). */
int achr_read_points_bin(const char *path, double **points, int *n, int *P);

/* Reads a warmup CSV into a newly malloc'ed n x W column-major array
 * (release with achr_free). The file is read whole and its lines parsed by
 * up to 16 host threads; an error names the first bad line. */
int achr_read_warmup_csv(const char *path, int n_expected, double **warmup, int *W);

void achr_free(void *p);

/* --------------------------------------------------- test/debug helpers */

/* Philox4x32-10 block of the library's stream (host implementation of the
 * same header the kernels use). */
void achr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* The library's matrix-core contraction on host arrays (copied to and from
 * the device): C = alpha * A B + beta * C with A(i, k) = A[i*sa_m + k*sa_k],
 * B(k, j) = B[k*sb_k + j*sb_n], C(i, j) = C[i*sc_m + j*sc_n] (one stride of
 * each operand must be 1). prec: 0 FP32 MFMA, 1 split-BF16x3, 2 BF16. */
int achr_gemm_f32(int M, int N, int K, const float *A, long sa_m, long sa_k, const float *B, long sb_k, long sb_n,
                  float *C, long sc_m, long sc_n, float alpha, float beta, int prec);

/* The residual form of the reduced-precision projection rounds on host
 * arrays (copied to and from the device): for q < M, row r = idx[q] (or q
 * when idx is NULL) of the FP64 rows x N matrix C becomes
 * C[r, :] - A[r, :] B, with A rows x K and B K x N row-major FP32, in the
 * kernel the sampler's rounds use (B packed once, 256 x 256 tiles, FP32
 * accumulation). prec: 0 FP32 MFMA, 1 split-BF16x3. kernel_ms, if not NULL,
 * receives the kernel's duration (HIP events; packing and copies excluded).
 * No reference counterpart: the reference reprojects in FP64
 * (ACHRcu/ACHRCuda.cu:338-343); tests hold this kernel to an FP64 product. */
int achr_gemm_sub_f64(int rows, int M, int N, int K, const float *A, const float *B, double *C, const int *idx,
                      int prec, float *kernel_ms);

/* The reciprocal the direction-bank chord computes on the device, r[i] =
 * 1/u[i], for `count` host values (copied to and from the device). It must
 * equal IEEE division for every |u| in [2^-30, 2^30]; tests hold it to that. */
int achr_bank_reciprocal(const double *u, double *r, long count);

/* The quotient the plain chord and the direction's normalisation compute on
 * the device in place of a division, q[i] = a[i] / b[i] for `count` host
 * values (copied to and from the device): the reciprocal above and one
 * corrected product. It must equal IEEE division, signed zeros included,
 * for |b| in [2^-30, 2^30] and quotients in the normal range; tests hold it
 * to that. */
int achr_quotient(const double *a, const double *b, double *q, long count);

/* Measurement aid, no reference counterpart: the HBM read rate a plain
 * streaming kernel reaches on this device (GB/s, best of `reps` passes over
 * `bytes` of device memory; every byte read once per pass, nothing written
 * but one double per workgroup). The yardstick for kernels that mostly
 * read, beside a device-to-device copy's rate. */
int achr_read_bandwidth(long bytes, int reps, double *gb_per_s);

/* The exact finish the warmup generator applies to LPs its first-order
 * iteration leaves unconverged: min c^T x s.t. x = N z, lb <= x <= ub, with N
 * (n x d, column major) an orthonormal basis of null(S), by a primal
 * active-set method in z on the host. x: in, a start; out, the solution.
 * xfeas: a point within the bounds on S x = 0 toward which an infeasible start
 * is pulled. Returns 0 at a KKT point, 1 at the iteration cap, 2 without a
 * feasible start, 3 for an unbounded direction, -1 for bad arguments, -7
 * out of memory. */
int achr_lp_polish(int n, int d, const double *N, const double *lb, const double *ub, const double *c,
                   const double *xfeas, double *x, int max_iters);

#ifdef __cplusplus
}
#endif

#endif /* ACHR_H_ */
