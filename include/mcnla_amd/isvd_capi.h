/* C ABI of the MI355X-native iSVD (integrated randomized SVD) library, libmcnla_amd.so.
 *
 * Everything here is plain C: opaque handles, plain pointers and sizes.  Matrix buffers passed to
 * the *_run / solve entry points are DEVICE pointers (hipMalloc'd memory, or anything the HIP
 * runtime accepts as device memory); `stream` is a hipStream_t passed as void* (NULL = the null
 * stream).  Every call returns an mcnla_status; mcnla_last_error() gives the message of the last
 * failure on the calling thread.
 *
 * Layouts follow the reference (emfomy/mcnla 0.7.2):
 *   - A: m x n, column-major (isvd_dense_cb, src/isvd_dense_cb.cpp:166-176) or row-major
 *     (isvd_dense_rb, src/isvd_dense_rb.cpp:166), leading dimension lda;
 *   - the sketch collection Y / Qs: ONE row-major m x (N*l) buffer, sketch i in columns
 *     [i*l, (i+1)*l) (include/mcnla/core/matrix/collection/dense/dense_matrix_collection_col_block.hpp:329-345);
 *   - Qbar: row-major m x l;  S: k values;  U: m x k (column-major for the SVD former,
 *     svd_former.hh:108, row-major for the Gramian formers);  V: row-major n x k.
 *
 * Each stage handle is the counterpart of one reference stage object: *_create = constructor +
 * initialize() (allocates the stage workspace, include/mcnla/isvd/core/stage_wrapper.hpp:37-46),
 * *_run = operator() (stage_wrapper.hpp:52-61), mcnla_stage_time = time()/times()
 * (stage_wrapper.hpp:115-164).
 */
#ifndef MCNLA_AMD_ISVD_CAPI_H_
#define MCNLA_AMD_ISVD_CAPI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------------
 * Status codes.  The reference is noexcept and only asserts (include/mcnla/def.hpp:71-107);
 * these codes surface the same conditions as errors instead. */
typedef enum {
  MCNLA_OK = 0,
  MCNLA_ERR_INVALID = 1,       /* bad shape/argument, e.g. k+p > min(m,n) */
  MCNLA_ERR_NOMEM = 2,         /* device allocation failed or the HBM budget is exceeded */
  MCNLA_ERR_HIP = 3,           /* a HIP runtime call failed */
  MCNLA_ERR_NOT_CONVERGED = 4, /* integrator hit maxiter; results are still returned */
  MCNLA_ERR_NUMERIC = 5,       /* NaN/Inf, or a Gram matrix that is not positive definite */
  MCNLA_ERR_STATE = 6,         /* call out of order (e.g. getter before run) */
  MCNLA_ERR_NO_DEVICE = 7
} mcnla_status;

typedef enum { MCNLA_F64 = 0, MCNLA_F32 = 1 } mcnla_dtype;
typedef enum { MCNLA_COL_MAJOR = 0, MCNLA_ROW_MAJOR = 1 } mcnla_order;

/* Arithmetic of the sketch GEMM (and of the collection it writes), and of the products that read
 * an FP32 A again (the former's Z = A'Qbar; Qbar = A X on the Gram-of-A route): FP32 MFMA under
 * the mixed precisions (FP32 sums within each split-K range, FP64 across them), FP64 arithmetic
 * under MCNLA_PREC_F64. */
typedef enum {
  MCNLA_PREC_F64 = 0,        /* everything FP64 (FP64 MFMA) */
  MCNLA_PREC_MIXED_BF16 = 1, /* sketch: BF16 MFMA, FP32 accumulate, FP32 Y; Gram/eig/KN/SVD FP64 */
  MCNLA_PREC_MIXED_BF16X3 = 2,/* as MIXED_BF16 with a 3-product hi/lo BF16 split of A and Omega */
  MCNLA_PREC_MIXED_F16 = 3,  /* sketch: FP16 MFMA with each row of A scaled by a power of two
                                (undone exactly in the output), FP32 accumulate */
  MCNLA_PREC_MIXED_F16X2 = 4,/* as MIXED_F16 with A split hi + lo (2 products, ~22 bits of A);
                                Omega enters rounded to FP16, so pair it with
                                MCNLA_OMEGA_PHILOX_F16, whose Omega is exactly FP16 */
  MCNLA_PREC_MIXED_F16X3 = 5 /* as MIXED_F16X2 plus the product A_hi Omega_lo (3 products; any
                                Omega) */
} mcnla_precision;

/* isvd::Parameters (include/mcnla/isvd/core/parameters.hh:31-119): m, n, k, p, N. */
typedef struct {
  int64_t nrow;
  int64_t ncol;
  int64_t rank;       /* k */
  int64_t over_rank;  /* p */
  int64_t num_sketch; /* N */
} mcnla_params;

/* ---------------------------------------------------------------------------------------------
 * Library / device plumbing. */
const char* mcnla_last_error(void);
const char* mcnla_version(void);
mcnla_status mcnla_device_count(int* count);
mcnla_status mcnla_set_device(int device);
mcnla_status mcnla_malloc(void** ptr, size_t bytes);
mcnla_status mcnla_free(void* ptr);
mcnla_status mcnla_memcpy_htod(void* dst, const void* src, size_t bytes);
mcnla_status mcnla_memcpy_dtoh(void* dst, const void* src, size_t bytes);
mcnla_status mcnla_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
mcnla_status mcnla_memset(void* ptr, int value, size_t bytes);
mcnla_status mcnla_stream_create(void** stream);
mcnla_status mcnla_stream_destroy(void* stream);
mcnla_status mcnla_stream_synchronize(void* stream);
mcnla_status mcnla_device_synchronize(void);
/* Events (hipEvent_t as void*, created without timing), pinned host memory and a stream-ordered
 * host-to-device copy: a client that produces a request's matrix on its own stream records an
 * event after it and hands the event to the server (mcnla_request.ready_event).  The copy is
 * asynchronous from pinned memory (mcnla_host_malloc); its source must stay untouched until the
 * stream has passed it. */
mcnla_status mcnla_event_create(void** event);
mcnla_status mcnla_event_destroy(void* event);
mcnla_status mcnla_event_record(void* event, void* stream);
mcnla_status mcnla_event_synchronize(void* event);
mcnla_status mcnla_host_malloc(void** ptr, size_t bytes);
mcnla_status mcnla_host_free(void* ptr);
mcnla_status mcnla_memcpy_htod_async(void* dst, const void* src, size_t bytes, void* stream);
mcnla_status mcnla_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ---------------------------------------------------------------------------------------------
 * Stage handles. */
typedef struct mcnla_stage_s* mcnla_stage;

/* Sketchers.  Counterparts of ColBlockGaussianProjectionSketcher
 * (include/mcnla/isvd/sketcher/col_block_gaussian_projection_sketcher.hpp:49-102),
 * RowBlockGaussianProjectionSketcher (row_block_gaussian_projection_sketcher.hpp:51-114),
 * GaussianProjectionSketcher with its power iterations (gaussian_projection_sketcher.hpp:66-103),
 * ColumnSamplingSketcher (column_sampling_sketcher.hpp:49-99) and
 * RowBlockColumnSamplingSketcher (row_block_column_sampling_sketcher.hpp:49-99).  The column
 * sampling kinds copy columns of A (MCNLA_PREC_F64 only) and are stage sketchers: the fused
 * solver rejects them. */
typedef enum {
  MCNLA_SKETCHER_COL_BLOCK_GAUSSIAN = 0,
  MCNLA_SKETCHER_ROW_BLOCK_GAUSSIAN = 1,
  MCNLA_SKETCHER_GAUSSIAN = 2,
  MCNLA_SKETCHER_COLUMN_SAMPLING = 3,
  MCNLA_SKETCHER_ROW_BLOCK_COLUMN_SAMPLING = 4
} mcnla_sketcher_kind;

/* Omega generation.  PHILOX: counter-based N(0,1) generated on the device (Philox4x32-10 +
 * Box-Muller; element e of Omega depends only on (seed, e)).  REFERENCE: the reference's own
 * stream (std::seed_seq -> std::mt19937 -> std::normal_distribution<>, split into `threads`
 * chunks; include/mcnla/core/random/streams/streams.hpp:118-150, routine/gaussian.hpp:39-56),
 * generated on the host and uploaded, as the reference's GPU sketcher does
 * (col_block_gaussian_projection_sketcher_gpu.hpp:110). */
typedef enum {
  MCNLA_OMEGA_PHILOX = 0,
  MCNLA_OMEGA_REFERENCE = 1,
  MCNLA_OMEGA_PHILOX_F16 = 2  /* the PHILOX stream with each value rounded to the nearest FP16
                                 number: an N(0,1) sketch that the FP16 sketch reads exactly */
} mcnla_omega_mode;

mcnla_status mcnla_sketcher_create(const mcnla_params* params, int kind, int precision,
                                   mcnla_stage* out);
/* setSeed (col_block_gaussian_projection_sketcher.hh:76).  `threads` is the OpenMP thread count
 * of the reference run whose stream should be reproduced (REFERENCE mode only). */
mcnla_status mcnla_sketcher_set_seed(mcnla_stage s, uint64_t seed, int omega_mode, int threads);
/* Y = A * Omega.  A: m x n (a_dtype, a_order, lda).  Y: row-major m x (N*l), FP64 for
 * MCNLA_PREC_F64, FP32 for the mixed precisions. */
mcnla_status mcnla_sketcher_run(mcnla_stage s, const void* A, int64_t lda, int a_dtype,
                                int a_order, void* Y, void* stream);
/* setExponent (gaussian_projection_sketcher.hh:87): q power iterations after Y = A Omega,
 * Omega := A'Y, Y := A Omega (gaussian_projection_sketcher.hpp:96-100), so Y = (A A')^q A Omega
 * and Omega ends as A' times the previous Y.  GAUSSIAN kind with MCNLA_PREC_F64 only; the other
 * kinds accept 0 (row_block_gaussian_projection_sketcher.hpp:158-168 allows no other value). */
mcnla_status mcnla_sketcher_set_exponent(mcnla_stage s, int64_t exponent);
/* Device pointer to the n x (N*l) row-major FP64 Omega of the last run (for checking). */
mcnla_status mcnla_sketcher_omega(mcnla_stage s, const double** omega);
/* Device pointer to the N*l column indices (int64) the last column-sampling run copied: column
 * j of Y is column idx[j] of A (column_sampling_sketcher.hpp:92-96). */
mcnla_status mcnla_sketcher_indices(mcnla_stage s, const int64_t** idx);
/* The column-sampling index stream on the host: `len` raw std::mt19937 draws of the seed chain
 * of `sketcher_kind` (col-block or row-block, as for Omega) in `threads` chunks, each stored in
 * the reference's 32-bit index_t and mapped to abs(draw) % ncol (streams.hpp:118-150,
 * uniform_bits.hpp:39-53, uniform_bits_detail_nomkl.hh:33-40, column_sampling_sketcher.hpp:95).
 * Every sketch draws its indices this way (both Omega modes). */
mcnla_status mcnla_column_indices_reference(uint64_t seed, int sketcher_kind, int threads,
                                            int64_t len, int64_t ncol, int64_t* out);
/* The REFERENCE-mode stream on the host (no device involved): `len` N(0,1) values for `seed`,
 * with the col-block (kind COL_BLOCK_GAUSSIAN / GAUSSIAN) or row-block seed chain and `threads`
 * chunks (streams.hpp:118-150, gaussian.hpp:39-56).  Used by the drivers and for checking. */
mcnla_status mcnla_omega_reference(uint64_t seed, int sketcher_kind, int threads, int64_t len,
                                   double* out);

/* Orthogonalizers: GramianOrthogonalizer (gramian_orthogonalizer.hpp:62-104) and
 * RowBlockGramianOrthogonalizer (row_block_gramian_orthogonalizer.hpp:63-105).
 * In place: Q_i = Y_i V_i S_i^{-1/2} with eig(Y_i' Y_i) = V_i S_i V_i'. */
typedef enum { MCNLA_ORTH_GRAMIAN = 0, MCNLA_ORTH_ROW_BLOCK_GRAMIAN = 1 } mcnla_orth_kind;
mcnla_status mcnla_orthogonalizer_create(const mcnla_params* params, int kind, int precision,
                                         mcnla_stage* out);
mcnla_status mcnla_orthogonalizer_run(mcnla_stage s, void* Qs, void* stream);

/* Integrators: KolmogorovNagumoIntegrator (kolmogorov_nagumo_integrator.hpp:77-185),
 * RowBlockKolmogorovNagumoIntegrator (row_block_kolmogorov_nagumo_integrator.hpp:80-208),
 * RowBlockGramianKolmogorovNagumoIntegrator (row_block_gramian_kolmogorov_nagumo_integrator.hpp:80-229),
 * RowBlockWenYinIntegrator (row_block_wen_yin_integrator.hpp:80-281),
 * RowBlockGramianWenYinIntegrator (row_block_gramian_wen_yin_integrator.hpp:87-331),
 * RowBlockReductionIntegrator (row_block_reduction_integrator.hpp:70-121) and
 * RowBlockExtrinsicMeanIntegrator (row_block_extrinsic_mean_integrator.hpp:90-163); the last two
 * do not iterate (maxiter and tol are ignored, iteration() is 0).
 * Defaults of the reference: maxiter 256; tol 1e-4 (KN) / 1e-3 (WY). */
typedef enum {
  MCNLA_INT_KN = 0,
  MCNLA_INT_ROW_BLOCK_KN = 1,
  MCNLA_INT_ROW_BLOCK_GRAMIAN_KN = 2,
  MCNLA_INT_ROW_BLOCK_WY = 3,
  MCNLA_INT_ROW_BLOCK_GRAMIAN_WY = 4,
  MCNLA_INT_ROW_BLOCK_REDUCTION = 5,
  MCNLA_INT_ROW_BLOCK_EXTRINSIC_MEAN = 6
} mcnla_integrator_kind;
mcnla_status mcnla_integrator_create(const mcnla_params* params, int kind, int precision,
                                     int64_t maxiter, double tol, mcnla_stage* out);
/* Qs: row-major m x (N*l) (FP64, or FP32 for mixed precisions); Qbar: row-major m x l FP64. */
mcnla_status mcnla_integrator_run(mcnla_stage s, const void* Qs, double* Qbar, void* stream);
/* iteration() (row_block_kolmogorov_nagumo_integrator.hh:121): iterations executed. */
mcnla_status mcnla_integrator_iteration(mcnla_stage s, int64_t* iteration);

/* Formers: SvdFormer (svd_former.hpp:71-116), GramianFormer (gramian_former.hpp),
 * ColBlockGramianFormer (col_block_gramian_former.hpp:74-131), RowBlockGramianFormer
 * (row_block_gramian_former.hpp:79-146) and RowBlockSymmetricFormer
 * (row_block_symmetric_former.hpp:98-131, square symmetric A: S holds the k algebraically
 * largest Ritz values in ascending order, U = Qbar W; V is not formed).  jobv != 0 also forms V. */
typedef enum {
  MCNLA_FORM_SVD = 0,
  MCNLA_FORM_GRAMIAN = 1,
  MCNLA_FORM_COL_BLOCK_GRAMIAN = 2,
  MCNLA_FORM_ROW_BLOCK_GRAMIAN = 3,
  MCNLA_FORM_ROW_BLOCK_SYMMETRIC = 4
} mcnla_former_kind;
mcnla_status mcnla_former_create(const mcnla_params* params, int kind, int jobv, int precision,
                                 mcnla_stage* out);
mcnla_status mcnla_former_run(mcnla_stage s, const void* A, int64_t lda, int a_dtype,
                              int a_order, const double* Qbar, void* stream);
/* vectorS() / matrixU() / matrixV(): device pointers owned by the stage.  U is column-major
 * m x k for MCNLA_FORM_SVD and row-major m x k otherwise; V is row-major n x k. */
mcnla_status mcnla_former_results(mcnla_stage s, const double** S, const double** U,
                                  const double** V);

/* Common to every stage. */
mcnla_status mcnla_stage_destroy(mcnla_stage s);
/* time() and the per-part times of the last run, in seconds (stage_wrapper.hpp:115-164).
 * Writes up to max_parts part times; *nparts receives the number of parts. */
mcnla_status mcnla_stage_time(mcnla_stage s, double* total, double* parts, int max_parts,
                              int* nparts);
const char* mcnla_stage_name(mcnla_stage s);
const char* mcnla_stage_part_names(mcnla_stage s);

/* ---------------------------------------------------------------------------------------------
 * Solver: sketch -> orthogonalize -> integrate -> form as one plan, with A resident on the
 * device and every intermediate (Y, Gram, integrator state, Qbar) kept in HBM.  This is the
 * isvd::Solver facade over the stage plugins; it never materializes Qs: the Gram matrix of Y
 * carries the orthogonalizer and the integrator, and Qbar is formed by one pass over Y. */
typedef struct mcnla_solver_s* mcnla_solver;

typedef struct {
  int sketcher;        /* mcnla_sketcher_kind */
  int orthogonalizer;  /* mcnla_orth_kind */
  int integrator;      /* mcnla_integrator_kind */
  int former;          /* mcnla_former_kind */
  int precision;       /* mcnla_precision */
  int jobv;            /* form V too */
  int64_t maxiter;
  double tol;
  uint64_t seed;
  int omega_mode;      /* mcnla_omega_mode */
  int omega_threads;   /* REFERENCE mode: OpenMP thread count of the stream to reproduce */
  int route;           /* mcnla_route: how G = Y'Y and Qbar are formed (results agree to rounding) */
  int graph;           /* 1: record the integrator's chunks of iterations as HIP graphs on the
                          first solve and replay them in later solves of this plan (same results) */
} mcnla_solver_config;

/* Two exact formulations of the fused solve.  SKETCH: Y = A Omega, G = Y'Y, Qbar = Y (T M), the
 * former reads A again (m Nl^2 FP64 MFMA work for G).  GRAM_OF_A: H = A'A, G = Omega' H Omega,
 * Qbar = A (Omega T M) and Z = A'Qbar = H (Omega T M), so Y is never formed (m n^2 FP64 MFMA work;
 * the better choice when n is small against Nl, e.g. 500000 x 2000 with Nl = 8192).  AUTO picks
 * the one with less work for the shape and precision. */
typedef enum { MCNLA_ROUTE_AUTO = 0, MCNLA_ROUTE_SKETCH = 1, MCNLA_ROUTE_GRAM_OF_A = 2 } mcnla_route;

/* Fills `cfg` with the defaults of the reference's isvd_dense_cb driver as used on this path:
 * col-block Gaussian sketcher, Gramian orthogonalizer, row-block KN (maxiter 256, tol 1e-4),
 * SVD former, FP64, Philox Omega, seed 0. */
void mcnla_solver_config_default(mcnla_solver_config* cfg);
mcnla_status mcnla_solver_create(const mcnla_params* params, const mcnla_solver_config* cfg,
                                 mcnla_solver* out);
mcnla_status mcnla_solver_destroy(mcnla_solver s);
/* Device bytes the solver holds (workspaces, not counting A). */
mcnla_status mcnla_solver_workspace_bytes(mcnla_solver s, size_t* bytes);
/* Device bytes a solver of this shape and configuration would hold, computed without allocating
 * and without a device (what the serving loop reserves per request).  No reference counterpart:
 * the reference sizes workspaces inside each stage's initialize() (stage_wrapper.hpp:52-61). */
mcnla_status mcnla_solver_required_bytes(const mcnla_params* params,
                                         const mcnla_solver_config* cfg, size_t* bytes);
mcnla_status mcnla_solver_set_seed(mcnla_solver s, uint64_t seed);
mcnla_status mcnla_solver_solve(mcnla_solver s, const void* A, int64_t lda, int a_dtype,
                                int a_order, void* stream);
/* The fused solve of a HOST matrix A (pageable or pinned host memory; lda, dtype, order as in
 * mcnla_solver_solve), streamed to the device in chunks of `chunk` columns (column-major A) or
 * rows (row-major A) on the solver's own copy stream (0: chunks of about 1 GiB, at most 64 of
 * them).  On the sketch route in FP64, each chunk's share of the sketch Y = A Omega runs as soon
 * as the chunk has landed, beside the copy of the next (column chunks: Y += A(:, K_k) Omega(K_k,
 * :), the reference GPU sketchers' chunked la::mm with beta = 1,
 * include/mcnla/isvd_gpu/sketcher/col_block_gaussian_projection_sketcher_gpu.hpp:108-134 and
 * row_block_gaussian_projection_sketcher_gpu.hpp:111-137; row chunks of a row-major A: their
 * rows of Y); the other routes and precisions take the copy first.
 * The solver keeps the device copy of A (m n elements) for the rest of the solve and for later
 * calls.  Results equal mcnla_solver_solve's on the same A to rounding (the chunked sums only
 * reorder the K sum); with a mixed precision or on the Gram-of-A route they are bit-identical.
 * A solver created by a server cannot grow this copy (MCNLA_ERR_STATE). */
mcnla_status mcnla_solver_solve_host(mcnla_solver s, const void* A, int64_t lda, int a_dtype,
                                     int a_order, int64_t chunk, void* stream);
/* Up to MCNLA_MAX_BATCH fused solves on one stream, the KN iterations of all of them in lockstep:
 * every iteration is one launch per kernel for the whole batch (not one per solve), and the host
 * reads the batch's convergence flags once per chunk of iterations.  Solver i solves A[i] (lda,
 * dtype, order as in mcnla_solver_solve); its results, iteration count and status[i] are
 * bit-identical to mcnla_solver_solve(solvers[i], ...) alone.  The solvers may differ in shape,
 * precision and configuration; Wen-Yin and combining integrators run one after another inside
 * the batch.  The call returns MCNLA_OK once every member has finished (each member's outcome is
 * in status[]; the error text is the first failing member's) or an error for invalid arguments,
 * before anything is enqueued.  Stage and kernel times of batched solves read zero; iteration
 * graphs (cfg.graph) are not used.  No reference counterpart: the reference runs one solve per
 * MPI job (src/isvd_dense_cb.cpp:46-302); this is for servers that hold many small requests. */
#define MCNLA_MAX_BATCH 16
mcnla_status mcnla_solver_solve_batch(int count, const mcnla_solver* solvers,
                                      const void* const* A, const int64_t* lda,
                                      const int* a_dtype, const int* a_order,
                                      mcnla_status* status, void* stream);
mcnla_status mcnla_solver_results(mcnla_solver s, const double** S, const double** U,
                                  const double** V, const double** Qbar);
mcnla_status mcnla_solver_iteration(mcnla_solver s, int64_t* iteration);
/* Per-stage device times of the last solve (seconds): [sketch, orthogonalize, integrate, form]
 * plus, in sketch_kernel_s, the duration of the sketch GEMM launch alone. */
mcnla_status mcnla_solver_stage_times(mcnla_solver s, double times[4], double* sketch_kernel_s);
/* Durations (seconds, HIP events on the solve's stream) of single kernel launches of the last
 * solve: times[0] = sketch GEMM, times[1] = Gram GEMM G = Y'Y (without its split-K reduction).
 * No reference counterpart (the reference times whole stages only, stage_wrapper.hpp:120-140);
 * this is for roofline accounting. */
mcnla_status mcnla_solver_kernel_times(mcnla_solver s, double times[2]);
/* The same for one kernel: MCNLA_KERNEL_SKETCH, MCNLA_KERNEL_GRAM, or MCNLA_KERNEL_FORMER_Z (the
 * former's Z = A'Qbar, its one pass over A; 0 on the Gram-of-A route, which forms Z from H).
 * No reference counterpart (roofline accounting, as above). */
#define MCNLA_KERNEL_SKETCH 0
#define MCNLA_KERNEL_GRAM 1
#define MCNLA_KERNEL_FORMER_Z 2
mcnla_status mcnla_solver_kernel_time(mcnla_solver s, int kernel, double* seconds);

/* ---------------------------------------------------------------------------------------------
 * Solve without allocating.  No counterpart in the reference (its stages allocate their
 * workspaces in initialize(), stage_wrapper.hpp:37-46); this is for callers that cannot allocate
 * device memory: code that must not call hipMalloc/hipFree (tight serving loops, harnesses that
 * forbid it).  Runs the fused solve of mcnla_solver_create + mcnla_solver_solve with every
 * workspace carved from a static device array of MCNLA_STATIC_WORKSPACE_BYTES per GPU (held from
 * the library's first use on that GPU until the process ends), and copies S (k), U (m x k), V
 * (n x k, with jobv) and Qbar (m x (k+p)) into the caller's device buffers (a NULL output is
 * skipped); layouts as mcnla_solver_results.  Synchronous on `stream`.  A problem whose
 * workspace (mcnla_solver_required_bytes) exceeds the array fails with MCNLA_ERR_NOMEM;
 * cfg->graph must be 0.  Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU race, sequential calls do not (each solve initializes what it reads).  Results are those of
 * mcnla_solver_solve with the same configuration, bit for bit. */
#define MCNLA_STATIC_WORKSPACE_BYTES ((size_t)256 << 20)
mcnla_status mcnla_solve_static(const mcnla_params* params, const mcnla_solver_config* cfg,
                                const void* A, int64_t lda, int a_dtype, int a_order, double* S,
                                double* U, double* V, double* Qbar, int64_t* iterations,
                                void* stream);

/* ---------------------------------------------------------------------------------------------
 * Autotuner.  No reference counterpart (the reference has no device plans to tune).  The FP64
 * GEMMs of a solve split K so that their grid holds about `target` workgroups, and each runs on
 * one of two kernels with different tile shapes (`kernels`); the defaults are the values
 * measured best at the 100000 x 10000 (Nl = 2048) and 500000 x 2000 (Nl = 8192) shapes.  A split
 * count or tile shape changes only the order of the sums, so a tuned plan gives the same results
 * to rounding, and the same plan always gives the same bits.
 *
 * Tuned plans are kept in a process-wide table keyed by (route, precision, m, n, l, N); a solver
 * created after an entry exists for its key uses it (servers too, per request shape). */
typedef struct {
  int64_t gram;     /* G = Y'Y on the sketch route; H = A'A on the Gram-of-A route */
  int64_t former;   /* Z = A'Qbar of the former (sketch route) */
  int64_t h_omega;  /* P = H Omega (Gram-of-A route) */
  int64_t omega_p;  /* G = Omega'P (Gram-of-A route) */
  int64_t n_by_l;   /* X = Omega (T M) and Z = H X (Gram-of-A route) */
  int64_t qbar;     /* Qbar = Y (T M) (sketch route) */
  int64_t kernels;  /* tile shape per GEMM, a bit each (set: the alternative kernel): 1 gram,
                       2 former, 4 h_omega, 8 omega_p, 16 n_by_l, 32 qbar.  Square products:
                       256 x 128 DMA tiles (clear) or register-staged 128 x 128 tiles (set);
                       tall-skinny products: A and B streamed into LDS by DMA, A staged through
                       registers, or A straight into MFMA registers (the default is the faster
                       per operand type and arithmetic, the bit takes another). */
} mcnla_tuning;

/* The defaults (no table entry). */
void mcnla_tuning_default(mcnla_tuning* t);
/* The plan this solver runs with. */
mcnla_status mcnla_solver_get_tuning(mcnla_solver s, mcnla_tuning* t);
/* Times candidate targets for each GEMM of this solver's route on A (as mcnla_solver_solve takes
 * it; a few solves per candidate, `reps` >= 1 of them timed), one GEMM at a time with the others
 * fixed, then that GEMM's other kernel (tile shape) at the chosen target, and keeps a candidate
 * only if its stage is faster than the incumbent's by more than 1 %,
 * applies the result to the solver and records it in the table.  The solver's results afterwards
 * are those of a solve with the tuned plan.  Synchronous.  MCNLA_ERR_STATE for a solver carved
 * from a reserved region (a server's), whose workspace cannot grow. */
mcnla_status mcnla_solver_autotune(mcnla_solver s, const void* A, int64_t lda, int a_dtype,
                                   int a_order, int reps, void* stream, mcnla_tuning* out);
/* Table access: record / look up the plan for a shape and configuration (route AUTO is resolved
 * as the solver would); *found = 0 and the defaults when there is no entry. */
mcnla_status mcnla_tune_set(const mcnla_params* params, const mcnla_solver_config* cfg,
                            const mcnla_tuning* t);
mcnla_status mcnla_tune_get(const mcnla_params* params, const mcnla_solver_config* cfg,
                            mcnla_tuning* t, int* found);
void mcnla_tune_clear(void);
/* Text persistence, one entry a line: "route precision m n l N gram former h_omega omega_p
 * n_by_l qbar kernels" (a line without `kernels` reads as 0).  Load merges into the table; a
 * malformed line is MCNLA_ERR_INVALID and loads nothing. */
mcnla_status mcnla_tune_save(const char* path);
mcnla_status mcnla_tune_load(const char* path);

/* ---------------------------------------------------------------------------------------------
 * Serving loop: continuous batching of low-rank-approximation requests on one GPU.  No reference
 * counterpart (the reference runs one solve per MPI job, src/isvd_dense_cb.cpp:46-302).
 *
 * The server owns one device arena of `hbm_budget` bytes, allocated at creation; nothing is
 * allocated afterwards.  A request is admitted when one of `slots` workers is free and the arena
 * holds a free region of the request's size (its solver's workspaces, plus a copy of A when A is
 * on the host, plus its results).  Requests that arrive while others run are admitted as slots
 * and memory free up; up to `max_bypass` later requests may be admitted ahead of a head request
 * that does not fit yet.  Each worker owns one HIP stream, so admitted solves overlap on the
 * device.  A finished request releases its workspaces at once and keeps only its results (S, U,
 * V, Qbar) in the arena until mcnla_server_release.  With batch_max > 1 a free worker admits the
 * queue head together with up to batch_max - 1 requests behind it (under the same bypass rule)
 * and runs their KN iterations in lockstep, one launch per kernel per iteration for all of them.
 * A solve is deterministic: its results do not depend on what else runs, alone or batched. */
typedef struct mcnla_server_s* mcnla_server;

typedef struct {
  mcnla_solver_config solver; /* per-request template (precision, integrator, former, tol, ...);
                                 the seed comes with each request */
  int slots;                  /* concurrent solves, each with a host thread and a stream; 1..16 */
  size_t hbm_budget;          /* bytes of the device arena */
  int max_bypass;             /* later requests admitted ahead of a blocked head, at most */
  int batch_max;              /* KN requests a worker admits at once and solves as one batch
                                 (mcnla_solver_solve_batch), 1..MCNLA_MAX_BATCH; 1 = one at a
                                 time.  Results do not depend on it. */
} mcnla_server_config;

typedef struct {
  mcnla_params params;
  const void* A;   /* device pointer, or host pointer when a_on_host (then copied into the arena) */
  int64_t lda;
  int a_dtype;     /* mcnla_dtype */
  int a_order;     /* mcnla_order */
  int a_on_host;
  uint64_t seed;   /* Omega seed */
  void* ready_event; /* hipEvent_t (or NULL): the solve waits for it on the device before it reads
                        A, so A may still be in production on another stream at submission */
} mcnla_request;

typedef struct {
  int status;           /* mcnla_status of the solve (MCNLA_ERR_NOT_CONVERGED still has results) */
  int64_t iterations;   /* integrator iterations */
  const double* S;      /* device pointers into the arena, valid until release: S (k) */
  const double* U;      /* m x k, column-major (row-major for a Gramian former) */
  const double* V;      /* n x k row-major, or NULL without jobv */
  const double* Qbar;   /* m x (k+p) row-major */
  double queue_s;       /* submit -> admission */
  double run_s;         /* admission -> results ready */
  double latency_s;     /* submit -> results ready */
  const char* error;    /* message of a failed solve, "" otherwise; valid until release */
} mcnla_served;

typedef struct {
  uint64_t submitted, completed, failed;
  int max_running;       /* most solves admitted at once */
  size_t peak_bytes;     /* most arena bytes reserved at once */
  size_t budget_bytes;
  double p50_s, p99_s, mean_s, max_s; /* latency (submit -> results ready) of completed requests */
  double span_s;         /* first submit -> last completion */
  uint64_t batches;      /* admissions: each admits one request or a batch of them */
  uint64_t kn_chunks;    /* KN-group chunks run (batch_max > 1) */
  double kn_group_mean;  /* members per KN-group chunk, on average */
  double kn_chunk_mean_s;/* wall time of a KN-group chunk (its iterations and read-back) */
} mcnla_server_stats;

void mcnla_server_config_default(mcnla_server_config* cfg);
mcnla_status mcnla_server_create(const mcnla_server_config* cfg, mcnla_server* out);
/* Waits for running solves; pending requests are dropped. */
mcnla_status mcnla_server_destroy(mcnla_server s);
/* Arena bytes the request would reserve. */
mcnla_status mcnla_server_request_bytes(mcnla_server s, const mcnla_request* r, size_t* bytes);
/* Queues a request (never blocks); MCNLA_ERR_NOMEM if it can never fit the budget. */
mcnla_status mcnla_server_submit(mcnla_server s, const mcnla_request* r, uint64_t* id);
/* Blocks until request `id` is finished or `timeout_s` passes (< 0: no limit); *done tells which. */
mcnla_status mcnla_server_wait(mcnla_server s, uint64_t id, double timeout_s, int* done);
/* Next finished request not yet reported by this call, in completion order; *id = 0 on timeout. */
mcnla_status mcnla_server_next_done(mcnla_server s, double timeout_s, uint64_t* id);
mcnla_status mcnla_server_result(mcnla_server s, uint64_t id, mcnla_served* out);
/* Frees the request's results in the arena; the id is then unknown. */
mcnla_status mcnla_server_release(mcnla_server s, uint64_t id);
mcnla_status mcnla_server_get_stats(mcnla_server s, mcnla_server_stats* out);
/* Clears the latency record and counters (not the requests). */
mcnla_status mcnla_server_reset_stats(mcnla_server s);

/* ---------------------------------------------------------------------------------------------
 * Synthetic test matrices (the reference demo's construction, demo/demo.cpp:250-293):
 * A = U diag(sigma) V' + noise * G, with U (m x r, row-major) and V (n x r, row-major) device
 * buffers, G ~ N(0,1) from Philox(seed).  A is written with the given dtype/order (lda). */
mcnla_status mcnla_synth_lowrank(int64_t m, int64_t n, int64_t r, const double* U, const double* V,
                                 const double* sigma, double noise, uint64_t seed, void* A,
                                 int64_t lda, int a_dtype, int a_order, void* stream);
/* A = (I - 2uu') [diag(sigma); 0] (I - 2vv') (m x n, n <= m) with unit device vectors u (m) and v
 * (n) and device sigma (n): a dense matrix whose singular values are exactly sigma, written in
 * O(mn) (the serving stream's per-request matrices).  Orthogonal factors do not change what the
 * method computes in distribution (Omega is rotation-invariant and the method commutes with an
 * orthogonal transform from the left), so these behave like Haar factors.  No reference
 * counterpart. */
mcnla_status mcnla_synth_reflected(int64_t m, int64_t n, const double* sigma, const double* u,
                                   const double* v, void* A, int64_t lda, int a_dtype,
                                   int a_order, void* stream);
/* out[e] = scale * g_e with g the Philox N(0,1) stream of `seed` (device buffer of `count`
 * doubles): random factors for the synthetic matrices.  No reference counterpart (the reference
 * demo draws its factors on the host, demo/demo.cpp:250-293). */
mcnla_status mcnla_fill_gaussian(double* out, int64_t count, uint64_t seed, double scale,
                                 void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MCNLA_AMD_ISVD_CAPI_H_ */
