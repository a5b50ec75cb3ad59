/*
 * spca.h -- C ABI of the MI355X alternating-maximization (AM) sparse-PCA solver.
 *
 * One library, libspca_mi355x.so, built with hipcc for gfx950.  Every entry point
 * returns an int status (SPCA_OK on success); no C++ exception crosses this boundary.
 * Buffers passed in are caller-owned host memory unless a name says "device";
 * device memory the solver uses is owned by the library.
 *
 * Matrix convention is the reference's: A has m rows (samples) and n columns
 * (variables), column major, leading dimension ld >= m.  An iterate x has length n.
 *
 * Reference interfaces this boundary replaces (optml/24am):
 *   - SPCASolver::GPUSolver::denseDataSolver<F>(cublasHandle_t&, m, n, d_B, h_x,
 *       OptimizationSettings*, OptimizationStatistics*, LD_M, LD_N)
 *       c++/src/gpugpower/gpu_sparse_PCA_solver.h:29-35      -> spca_dense_solver_device,
 *                                                                spca_solve
 *   - SPCASolver::MulticoreSolver::denseDataSolver<F>(B, ldB, x, m, n, settings, stats)
 *       c++/src/gpower/sparse_PCA_solver.h:36-40             -> spca_dense_solver_host
 *   - SolverStructures::OptimizationSettings / OptimizationStatistics
 *       c++/src/class/optimization_settings.h:71-168,
 *       c++/src/class/optimization_statistics.h:27-46        -> spca_settings / spca_stats
 *   - SPCA_Formulation  c++/src/class/optimization_settings.h:27-37 -> spca_formulation
 */
#ifndef SPCA_MI355X_H_
#define SPCA_MI355X_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define SPCA_OK 0
#define SPCA_EINVAL 1     /* bad argument (null pointer, bad enum, bad size) */
#define SPCA_ENOMEM 2     /* device allocation failed */
#define SPCA_EBUDGET 3    /* request exceeds the context's HBM budget */
#define SPCA_EDEVICE 4    /* HIP runtime error */
#define SPCA_ECANCELLED 5 /* job cancelled before completion */
#define SPCA_ENOTFOUND 6  /* unknown job or matrix handle */
#define SPCA_EAGAIN 7     /* job not finished yet (spca_poll) */

/* ---- enums -------------------------------------------------------------- */
/* reference SPCA_Formulation, same values (c++/src/class/optimization_settings.h:27-37) */
typedef enum spca_formulation {
    L0_constrained_L2_PCA = 0,
    L0_constrained_L1_PCA = 1,
    L1_constrained_L2_PCA = 2,
    L1_constrained_L1_PCA = 3,
    L0_penalized_L2_PCA = 4,
    L0_penalized_L1_PCA = 5,
    L1_penalized_L2_PCA = 6,
    L1_penalized_L1_PCA = 7
} spca_formulation;

/* arithmetic of the two products A*X and A^T*Z; everything else is FP32/FP64 */
typedef enum spca_precision {
    SPCA_PREC_FP32 = 0,   /* FP32 MFMA, FP32 accumulate */
    SPCA_PREC_BF16X3 = 1, /* A and iterates split hi+lo in bf16, three MFMA products, FP32 accumulate */
    SPCA_PREC_FP16 = 2    /* A and iterates in fp16, FP32 accumulate */
} spca_precision;

/* when a batch stops (reference batch mode: c++/src/gpower/sparse_PCA_solver.h:184-208) */
typedef enum spca_termination {
    SPCA_TERM_BATCH_MAX = 0, /* stop the batch when max_j err_j < tol (and it > 0) */
    SPCA_TERM_PER_SP = 1     /* freeze each starting point once its own err < tol */
} spca_termination;

/* where starting points come from */
typedef enum spca_start {
    SPCA_START_EXPLICIT = 0,     /* caller passes X0 (n x total_sp, column major) */
    SPCA_START_REFERENCE = 1,    /* reference rand_r generator, SP k seeded with k
                                    (c++/src/gpower/gpower_commons.h:21-47), in float */
    SPCA_START_REFERENCE_F64 = 2, /* the same generator evaluated in double, as the reference's
                                    double-precision runs draw it, then stored as float */
    SPCA_START_GAUSSIAN = 3      /* generated on the device: SP k is the unit-normalised vector of
                                    Box-Muller normals of Philox4x32-10 counters (i, k, 4) keyed
                                    by settings.seed (see spca_matrix_generate) */
} spca_start;

/* storage forms a matrix handle may carry (bit mask) */
#define SPCA_STORE_FP32 (1u << 0)
#define SPCA_STORE_BF16X3 (1u << 1)
#define SPCA_STORE_FP16 (1u << 2)
/* A in 25-bit fixed point per column (a power-of-two scale and four int8 pieces): the operand
   of A^T * sgn(z), the L1-variance formulations' second product, on the int8 MFMA with exact
   integer sums (exact_sign_products) */
#define SPCA_STORE_FIX8 (1u << 3)

/* ---- settings / statistics ---------------------------------------------- */
/* mirrors OptimizationSettings (c++/src/class/optimization_settings.h:71-123) */
typedef struct spca_settings {
    int formulation;          /* spca_formulation */
    unsigned constraint;      /* s: constraintParameter */
    double penalty;           /* gamma: penaltyParameter */
    double tolerance;         /* tolerance */
    int max_iter;             /* maximumIterations */
    int total_sp;             /* totalStartingPoints */
    int batch;                /* batchSize (0 -> total_sp) */
    int use_otf;              /* useOTF: a finished SP's slot takes the next SP on the device
                                 (c++/src/gpower/sparse_PCA_solver.h:83-169) */
    int termination;          /* spca_termination */
    int precision;            /* spca_precision */
    int start;                /* spca_start */
    int check_every;          /* host reads the convergence flag every k iterations (>=1) */
    int resolve_iters;        /* > 0: FP64 re-solve on the support of x_out, at most this many iterations */
    double resolve_tol;       /* relative objective change that stops the re-solve */
    uint64_t seed;            /* key of SPCA_START_GAUSSIAN starting points */
    int sparse_products;      /* 1 (default): A*X is gathered over the nonzeros of the iterates
                                 (FP64 accumulation) for each group of 256 SPs of a batch once
                                 every iterating SP of the group has at most n/16 (FP32) or
                                 n/128 (bf16x3, fp16) of them; 0: always the dense product
                                 of `precision`; 2: as 1, and in batch mode (use_otf = 0) on a
                                 dense matrix without deflation the other groups' A*X is the FP32
                                 product over U, the union of their iterates' supports, whenever
                                 |U| <= n/4 (FP32), n/12 (bf16x3) or n/24 (fp16): the host reads
                                 |U| after every iteration.  Fused solves treat 2 as 1. */
    int exact_sign_products;  /* 1 (default): for the L1-variance formulations (z = sgn(A x) in
                                 {-1, 0, +1}), FP32 and bf16x3 solves compute A^T z from A in
                                 25-bit fixed point (SPCA_STORE_FIX8, added on first use) on the
                                 int8 MFMA: exact integer sums, one rounding of each v_j to FP32,
                                 |error| <= 2^-26 * max_i |A_ij| * sum_i |z_i| (+ that rounding);
                                 0: the product of `precision` */
    int exact_ties;           /* 1 (default): L0 constraint with L2 variance in FP32 and bf16x3:
                                 when candidates for the s-th place lie within the A^T*Z
                                 product's rounding of each other (2 to 16 of them), their
                                 entries of v are recomputed in FP64 from the FP32 master and z
                                 and the place is decided on those; 0: on the product's values */
} spca_settings;

/* mirrors OptimizationStatistics (c++/src/class/optimization_statistics.h:27-46) */
typedef struct spca_stats {
    unsigned it;              /* total batch iterations done (sum over batches) */
    double fval;              /* best reference-native objective (lagged value) */
    double error;             /* final batch max error of the last batch */
    double loop_ms;           /* iteration loop time, hipEvents (totalTrueComputationTime) */
    double wall_ms;           /* whole call, host clock (totalElapsedTime) */
    unsigned n;               /* length of x */
    int best_sp;              /* global index of the SP that gave fval */
    int total_sp;             /* after the reference clamp/rounding */
    unsigned constraint;      /* s after the reference clamp (s <= n) */
    double ev_fp64;           /* FP64 explained variance of x_out: |A x|_2^2 or |A x|_1 */
    double gemm_ax_ms;        /* summed kernel time of the A*X products (hipEvents) */
    double gemm_atz_ms;       /* summed kernel time of the A^T*Z products */
    double step_ms;           /* summed time of epilogue + formulation step + reduction kernels
                                 (deflation projections excluded) */
    unsigned long long sp_iterations; /* sum over SPs of iterations while active */
    double ev_resolve_fp64;   /* explained variance after the FP64 support re-solve (0 if not run) */
    int resolve_iterations;   /* iterations the re-solve ran */
    unsigned sparse_iterations; /* iterations whose A*X was gathered over the iterates' nonzeros */
    unsigned union_iterations;  /* iterations whose A*X was the product over the union of supports
                                   (sparse_products = 2) */
    unsigned list_tries;        /* constrained SP steps that started from the previous threshold's
                                   candidate list (every step after an SP's first) */
    unsigned list_steps;        /* of those, the steps the list decided (the others re-ran the
                                   column passes of a first step) */
    double step_kernel_ms;      /* of step_ms, the formulation step kernel alone */
} spca_stats;

/* per starting point result (reference values / iters / cardinalities) */
typedef struct spca_sp_result {
    float value;              /* final reference-native objective */
    int iters;                /* iterations this SP performed */
    int nnz;                  /* nonzeros of its final iterate */
    int converged;            /* 1 if its own err < tol at the end */
    double ev_fp64;           /* FP64 explained variance of its normalised final iterate */
} spca_sp_result;

typedef struct spca_ctx_s* spca_ctx;
typedef uint64_t spca_matrix;
typedef uint64_t spca_job;

/* ---- context ------------------------------------------------------------ */
/* Select device `device`, create the context's stream and set the HBM budget in bytes
 * (0 = 90% of free device memory).  No reference counterpart: the reference creates its
 * cuBLAS handle in the front end (c++/src/frontends/gpu_console.cu:64-72). */
int spca_ctx_create(int device, uint64_t hbm_budget_bytes, spca_ctx* out);
int spca_ctx_destroy(spca_ctx ctx);
/* Device memory of the context: used = held (live + cached), cached = freed blocks kept for
 * reuse by size class (released when a request would exceed the budget), budget = the limit. */
int spca_ctx_memory(spca_ctx ctx, uint64_t* used, uint64_t* cached, uint64_t* budget);
/* last error message on this context (thread-local to the context), never NULL */
const char* spca_last_error(spca_ctx ctx);
/* library version string */
const char* spca_version(void);

/* ---- resident matrices ---------------------------------------------------- */
/* A sparse A (host CSC arrays, zero-based: col_ptr n+1, row_idx and vals col_ptr[n]), the
 * input of the reference's sparse solver (sparse_PCA_solver_CSC,
 * c++/src/gpower/sparse_PCA_solver_for_CSC.h:112-495).  col_means (n) and row_means (m), each
 * optional (NULL), centre A implicitly, A_c = A - 1 col_means^T - row_means 1^T, as its doMean
 * and doRowMean do (:215-250, :287-316): A is never densified.  The device keeps the CSC
 * arrays and a CSR copy; products are gathers with FP64 sums.  Solves of a sparse matrix run
 * in FP32 (spca_solve, spca_solve_components, the request server); the FP64 re-solve, fused
 * solves, dense storage forms and download return SPCA_EINVAL.  `key` is kept, not shared. */
int spca_matrix_upload_csc(spca_ctx ctx, const int* col_ptr, const int* row_idx, const float* vals, int m, int n,
                           const float* col_means, const float* row_means, uint64_t key, spca_matrix* out);
/* Copy A (host, FP32, m x n, column major, leading dimension ld) to the device and build
 * the operand layouts for the storage forms in `storage` (SPCA_STORE_* mask).  `key` lets
 * later uploads of the same data share the resident copy (0 = no sharing).
 * Replaces the front end's H2D copy, c++/src/frontends/gpu_console.cu:43-62. */
int spca_matrix_upload(spca_ctx ctx, const float* A, int m, int n, int ld, unsigned storage,
                       uint64_t key, spca_matrix* out);
/* Same, from a caller-owned device buffer (FP32, column major, leading dimension ld). */
int spca_matrix_from_device(spca_ctx ctx, const float* d_A, int m, int n, int ld, unsigned storage,
                            uint64_t key, spca_matrix* out);
/* Add storage forms to an existing matrix. */
int spca_matrix_add_storage(spca_ctx ctx, spca_matrix mat, unsigned storage);
int spca_matrix_release(spca_ctx ctx, spca_matrix mat);
/* Synthetic spiked matrix generated on the device (no host copy of A):
 * A = N + sum_k sqrt(theta[k]) g_k v_k^T, m x n.  N(0,1) noise N and g_k ~ N(0, I_m) are
 * Box-Muller normals of Philox4x32-10 outputs 0 and 1 (counter (r, c, 0) for N[r, c],
 * (r, k, 1) for g_k[r], key = seed); v_k has spike_nnz distinct columns (the t-th candidate
 * is output 0 of counter (t, k, 3) mod n, repeats skipped) with values from counters
 * (t, k, 2), normalised to unit 2-norm.  loadings (optional, n x nspikes column major)
 * receives the v_k.  key != 0 shares a resident matrix as spca_matrix_upload does.  For
 * benchmarks and tests at sizes a host copy would not serve; the reference has no
 * counterpart (its experiments build data on the host). */
int spca_matrix_generate(spca_ctx ctx, int m, int n, uint64_t seed, int nspikes, const double* theta,
                         int spike_nnz, unsigned storage, uint64_t key, float* loadings, spca_matrix* out);
/* copy the resident FP32 matrix back (column major, leading dimension ld >= m) */
int spca_matrix_download(spca_ctx ctx, spca_matrix mat, float* A, int ld);
/* device bytes held by a matrix (all its storage forms) */
int spca_matrix_bytes(spca_ctx ctx, spca_matrix mat, uint64_t* bytes);

/* ---- solve ---------------------------------------------------------------- */
/* Solve on a resident matrix.  X0: n x total_sp column-major host starting points when
 * settings->start == SPCA_START_EXPLICIT (else ignored, may be NULL).  x_out (n floats)
 * receives the best iterate normalised to unit 2-norm.  per_sp (total_sp entries) and
 * X_final (n x total_sp host, column major; final iterates as held, unnormalised for
 * penalised formulations) are optional.  Semantics: c++/src/gpower/gpower_commons.h:67-177
 * for the iteration, with the reference GPU path's thresholds
 * (c++/src/gpugpower/my_cublas_wrapper.h:240-255,335-395,571-634). */
int spca_solve(spca_ctx ctx, spca_matrix mat, const spca_settings* settings, const float* X0,
               float* x_out, spca_stats* stats, spca_sp_result* per_sp, float* X_final);

/* spca_solve without device allocation: every device buffer the solve uses (iterates, K-split
 * partial products, per-SP state, the FP64 explained-variance partials) lies in one fixed
 * array of 128 MiB per GPU that the library holds from load to process end (a file-scope
 * device array; its address is looked up at spca_ctx_create), so the call makes no
 * hipMalloc/hipFree.  The arguments mean what spca_solve's mean.  Differences:
 *   - a shape whose buffers exceed the array returns SPCA_EINVAL, saying how many bytes it needs
 *     (about 8 B (FP32) to 12 B (16-bit forms) per padded n x padded batch entry for the
 *     iterates, plus 4 B per padded m x padded batch entry for each K split of A*X and for z,
 *     plus the explained variance's partial sums);
 *   - the matrix must already hold every storage form the solve reads (SPCA_EINVAL otherwise;
 *     add them with spca_matrix_add_storage first): the precision's, plus SPCA_STORE_FIX8 and
 *     SPCA_STORE_FP32 for the L1-variance formulations' exact sign products in FP32/bf16x3;
 *   - batch mode on dense matrices only (use_otf, CSC matrices and resolve_iters > 0 return
 *     SPCA_EINVAL; sparse_products = 2 acts as 1); the product plan is the cached or heuristic
 *     one (no tuning run); the phase times in stats are 0.
 * Every call on one GPU uses that GPU's one array and leaves it holding its last contents (each
 * call clears what must start at zero).  Calls on one context are serialised; calls from
 * different contexts on the same GPU that overlap in time race on the array.
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc). */
int spca_solve_noalloc(spca_ctx ctx, spca_matrix mat, const spca_settings* settings, const float* X0,
                       float* x_out, spca_stats* stats, spca_sp_result* per_sp, float* X_final);

/* One job of a fused multi-job solve.  The fields mean what spca_solve's arguments mean;
 * status receives the job's own result code. */
typedef struct spca_fused_job {
    spca_matrix matrix;
    const spca_settings* settings; /* one batch (total_sp <= batch), use_otf = 0 */
    const float* X0;
    float* x_out;                  /* n floats */
    spca_stats* stats;             /* optional; loop_ms counts to the check at which the job had finished */
    spca_sp_result* per_sp;        /* optional, batch entries */
    int status;
} spca_fused_job;

#define SPCA_FUSED_GRAPH 1u /* replay the fused iteration from a captured hipGraph */

/* Solve uneven jobs (different matrices, shapes, formulations, s, batch sizes) together:
 * every AM iteration of all jobs is one fused launch per phase, and a chunk of
 * min(check_every) iterations is captured into a hipGraph (SPCA_FUSED_GRAPH).  All jobs
 * share one precision.  Each job terminates on its own rule; a job's FP32 results equal
 * those of a solo spca_solve of it with the same plan (128x128 product tiles).  No
 * counterpart in the reference, whose GPU solver runs one job per call
 * (c++/src/gpugpower/gpu_sparse_PCA_solver.h:29-211). */
int spca_solve_fused(spca_ctx ctx, spca_fused_job* jobs, int njobs, unsigned flags);

/* ---- product plans (autotuner) --------------------------------------------- */
/* A solve shape (precision, padded m, n, batch) runs its two products with a plan: the K
 * split of A*X and the tile kernel (128x128, or 256x256 for 16-bit shapes it fits).
 * mode 0 (default): a fixed heuristic.  mode 1: the first solve of a shape times every valid
 * plan on its own matrix and starting points and keeps the fastest.  cache_path (may be
 * NULL): measured plans are loaded from and saved to this JSON file.  A plan fixes the
 * summation order of A*X, so results are reproducible for a given plan; solo and fused
 * solves of a shape without a tile choice share one plan.  No counterpart in the reference
 * (cuBLAS chooses its own kernels, c++/src/gpugpower/my_cublas_wrapper.h:100-133). */
int spca_set_tuning(spca_ctx ctx, int mode, const char* cache_path);
/* The plan a solve of this shape would use now (fused != 0: the fused path's 128x128 tiles);
 * tuned_ms is the measured time of A*X + epilogue + A^T*Z, 0 for an untuned plan. */
int spca_get_plan(spca_ctx ctx, int precision, int m, int n, int batch, int fused, int* splits, int* big,
                  double* tuned_ms);

/* Drop-in for GPUSolver::denseDataSolver<float>(handle, m, n, d_B, h_x, settings, stats,
 * LD_M, LD_N) (c++/src/gpugpower/gpu_sparse_PCA_solver.h:29-35): d_B is a caller-owned
 * device matrix with leading dimension LD_M; h_x (n floats, host) receives the result.
 * The matrix is uploaded for the call and released on return, as the reference holds
 * nothing between calls.  LD_N is accepted for signature parity and ignored. */
int spca_dense_solver_device(spca_ctx ctx, unsigned m, unsigned n, const float* d_B, float* h_x,
                             const spca_settings* settings, spca_stats* stats, unsigned LD_M,
                             unsigned LD_N);

/* Drop-in for MulticoreSolver::denseDataSolver<float>(B, ldB, x, m, n, settings, stats)
 * (c++/src/gpower/sparse_PCA_solver.h:36-40), computed on the GPU: B is host memory.
 * Returns the status; *fval_out receives the best value the reference returns. */
int spca_dense_solver_host(spca_ctx ctx, const float* B, int ldB, float* x, unsigned m, unsigned n,
                           const spca_settings* settings, spca_stats* stats, double* fval_out);

/* Sparse principal components one after another, by deflation (reference
 * SparseDeflationCollection, c++/src/gpower/sparse_PCA_solver_for_CSC.h:56-106, used by its
 * sparse solver at :207, :311, :425, :438).  Component c is solved as spca_solve solves,
 * except that every iterate is projected off the components before it, x <- x - (p . x) p
 * for p = x_0 .. x_{c-1} in order, before each A*X and on v right after each A^T*Z; its
 * unit-norm result then joins the list.  The same settings and starting points serve every
 * component.  X_out: n x k column major; stats: k entries (may be NULL). */
int spca_solve_components(spca_ctx ctx, spca_matrix mat, const spca_settings* settings, const float* X0, int k,
                          float* X_out, spca_stats* stats);

/* ---- request server (continuous batching) ---------------------------------- */
/* No reference counterpart: the reference solves one request per call.  A server runs
 * `workers` worker threads, each with its own HIP stream; matrices stay resident and are
 * shared by handle.  Continuous batching: a short job (dense, one batch with total_sp <=
 * batch, no OTF, 128x128 product tiles, an estimated iteration of at most 200 us) runs in a
 * fused group, whose jobs iterate together as spca_solve_fused runs them (one launch per phase
 * for the group).  After every chunk of min(check_every, 64) iterations, finished and
 * cancelled jobs leave and return their memory, and queued short jobs of the group's precision
 * join while the group's estimated iteration stays within 1 ms and their workspaces fit the
 * HBM budget; they start at iteration 0 beside the others.  With several workers, worker 0 is
 * the group lane (it takes short jobs from anywhere in the queue); the others take jobs in
 * order and solve longer ones alone.  A single worker admits strictly in order.  Results are
 * independent of what runs beside a job: bitwise those of the same job submitted alone, and of
 * its spca_solve. */
#define SPCA_JOB_QUEUED 0
#define SPCA_JOB_RUNNING 1
#define SPCA_JOB_DONE 2
#define SPCA_JOB_CANCELLED 3

typedef struct spca_job_times {
    double queue_ms;          /* submit -> start */
    double run_ms;            /* start -> finish */
    double latency_ms;        /* submit -> finish */
} spca_job_times;

/* Counters of a running server. */
typedef struct spca_server_info {
    int groups;          /* fused groups started */
    int joined_running;  /* jobs admitted into a group while other members were iterating */
    int max_members;     /* most jobs iterating in one group at once */
    int solo_jobs;       /* jobs solved alone (sparse, several batches, OTF, or batching off) */
} spca_server_info;

int spca_server_start(spca_ctx ctx, int workers);
int spca_server_stop(spca_ctx ctx);  /* cancels queued jobs, waits for running ones */
/* Most jobs per fused group (default 64; 0: every job is solved alone, one per worker). */
int spca_server_set_batching(spca_ctx ctx, int max_group_jobs);
int spca_server_get_info(spca_ctx ctx, spca_server_info* info);
/* X0 is copied at submission; want_per_sp asks for per-SP results in spca_wait. */
int spca_submit(spca_ctx ctx, spca_matrix mat, const spca_settings* settings, const float* X0, int want_per_sp,
                spca_job* out);
int spca_poll(spca_ctx ctx, spca_job job, int* state);  /* SPCA_JOB_* */
int spca_cancel(spca_ctx ctx, spca_job job);
/* Blocks until the job ends, copies its results (x_out: n floats; per_sp: total_sp entries)
 * and retires the handle.  Returns the job's status (SPCA_ECANCELLED if cancelled). */
int spca_wait(spca_ctx ctx, spca_job job, float* x_out, spca_stats* stats, spca_sp_result* per_sp,
              spca_job_times* times);

/* FP64 re-solve restricted to the support of x (its nonzeros), on the GPU: the unit vector y
 * on that support maximising |A y|_2 (L2-variance formulations: power iteration on
 * A_S^T A_S) or |A y|_1 (L1-variance formulations: the AM sign iteration restricted to S),
 * started from x normalised.  Stops when the objective changes by <= tol (relative) or after
 * max_iter iterations.  y_out (n doubles, zero off the support) and *ev_out receive the
 * result.  No reference counterpart: it refines the solver's output on a fixed support. */
int spca_resolve_support(spca_ctx ctx, spca_matrix mat, int formulation, const float* x, int max_iter, double tol,
                         double* y_out, double* ev_out, int* iters_out);

/* Product forms a dense solve of these settings on an m x n matrix runs, decided on the host
 * alone (no device call).  *ax_form: the dense A*X form (the precision; sparse iterates may
 * still be gathered in FP64 at run time).  *atz_form: the A^T*Z form, SPCA_FORM_FIX8 for the
 * exact fixed-point A^T sgn(z) (L1 variance, FP32/bf16x3, exact_sign_products, and fewer
 * than 2^18 rows after padding to 256), else the precision.  Returns SPCA_EINVAL for a bad
 * formulation or precision, or n >= 2^23 (the limit of the step kernel's fixed-point sums),
 * which spca_solve rejects too.  No reference counterpart (cuBLAS picks its kernels). */
#define SPCA_FORM_FP32 0
#define SPCA_FORM_BF16X3 1
#define SPCA_FORM_FP16 2
#define SPCA_FORM_FIX8 4
int spca_product_forms(const spca_settings* settings, int m, int n, int* ax_form, int* atz_form);

/* ---- reference settings helpers ------------------------------------------- */
/* reference OptimizationSettings defaults (c++/src/class/optimization_settings.h:101-123) */
void spca_settings_default(spca_settings* st);
/* reference chceckInputAndModifyIt(n) (c++/src/class/optimization_settings.h:156-168):
 * clamp s <= n, total_sp rounded up to a multiple of batch; returns the batch count */
int spca_settings_check(spca_settings* st, unsigned n);
/* isConstrainedProblem / isL1ConstrainedProblem / isL1PenalizedProblem, :125-153 */
int spca_is_constrained(int formulation);
int spca_is_l1_constrained(int formulation);
int spca_is_l1_penalized(int formulation);
int spca_is_l1_variance(int formulation);

#ifdef __cplusplus
}
#endif
#endif /* SPCA_MI355X_H_ */
