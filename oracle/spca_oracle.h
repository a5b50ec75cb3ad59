/*
 * CPU oracle for the alternating-maximization (AM) sparse-PCA solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP solver
 * is compared against; it is never linked into, loaded by, or called from the
 * product path (24am_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * It restates the reference's CPU method (optml/24am, C++ multicore solver):
 *   - settings / clamp            c++/src/class/optimization_settings.h:27-37,125-168
 *   - starting points             c++/src/gpower/gpower_commons.h:21-64
 *   - constrained iteration       c++/src/gpower/gpower_commons.h:67-136
 *   - penalised iteration         c++/src/gpower/gpower_commons.h:139-177
 *   - penalised thresholds        c++/src/gpower/sparse_PCA_thresholding.h:25-102
 *   - top-s hard threshold        c++/src/utils/thresh_functions.h:91-145
 *   - soft threshold (lambda)     c++/src/utils/thresh_functions.h:160-254 (CPU rule)
 *                                 c++/src/gpugpower/my_cublas_wrapper.h:335-395 (GPU rule)
 *   - error / termination         c++/src/utils/termination_criteria.h:29-44
 *   - batch and OTF drivers       c++/src/gpower/sparse_PCA_solver.h:36-243
 *   - synthetic problem generator c++/src/problem_generators/gpower_problem_generator.h:51-76
 *
 * Matrix convention follows the reference: A has m rows and n columns, column
 * major with leading dimension ldA >= m.  An iterate x has length n, z = A x has
 * length m.  Every function exists for float (suffix _f32) and double (_f64).
 *
 * Pinning status of each function is documented in the test that pins it
 * (tests/test_oracle_*.py).  Functions without a pin say "parity unpinned" here:
 *   - oracle_hard_rule SET (std::set de-duplicating top-s): pinned only on tie-free
 *     data (where it equals the SORT rule) -- parity unpinned on data with ties.
 *   - soft threshold degenerate branch (s >= (|v|_1/|v|_2)^2, no root in any
 *     segment): parity unpinned, for both lambda rules.
 *   - soft_thresholding (CPU rule) also accepts a spurious first-segment root when
 *     the zero discriminant of the one-element quadratic rounds positive by enough
 *     (thresh_functions.h:205-216); the restatement reproduces it, and the tests pin
 *     when it happens, but no reference output shows it.
 */
#ifndef SPCA_ORACLE_H_
#define SPCA_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* reference enum SPCA_Formulation, c++/src/class/optimization_settings.h:27-37 */
enum oracle_formulation {
    OR_L0_CONS_L2 = 0, OR_L0_CONS_L1 = 1, OR_L1_CONS_L2 = 2, OR_L1_CONS_L1 = 3,
    OR_L0_PEN_L2 = 4, OR_L0_PEN_L1 = 5, OR_L1_PEN_L2 = 6, OR_L1_PEN_L1 = 7
};

/* hard-threshold tau rule */
enum oracle_hard_rule {
    OR_HARD_SORT = 0, /* tau = s-th largest |v| counting multiplicity (thresh_functions.h:91-98) */
    OR_HARD_SET = 1   /* std::set top-s, de-duplicates equal magnitudes (thresh_functions.h:100-118) */
};

/* L1-constraint lambda rule */
enum oracle_lambda_rule {
    OR_LAMBDA_CPU = 0, /* soft_thresholding scan (thresh_functions.h:160-254) */
    OR_LAMBDA_GPU = 1  /* compute_soft_thresholding_parameter (gpugpower/my_cublas_wrapper.h:335-375) */
};

enum oracle_termination {
    OR_TERM_BATCH_MAX = 0, /* batch stops when max_j err_j < tol (sparse_PCA_solver.h:184-208) */
    OR_TERM_PER_SP = 1     /* each SP freezes once its own err < tol (no slot refill) */
};

typedef struct oracle_settings {
    int formulation;        /* 0..7 */
    unsigned constraint;    /* s */
    double penalty;         /* gamma */
    double tolerance;
    int max_iter;
    int total_sp;
    int batch;              /* 0 -> total_sp */
    int use_otf;            /* reference on-the-fly slot refill (sparse_PCA_solver.h:83-169) */
    int termination;        /* oracle_termination (ignored when use_otf) */
    int hard_rule;          /* oracle_hard_rule */
    int threads;            /* OpenMP threads (results do not depend on it) */
    int gemm_backend;       /* 0: fixed-order loops, double accumulation; 1: host CBLAS (timing only) */
    int lambda_rule;        /* oracle_lambda_rule */
} oracle_settings;

typedef struct oracle_stats {
    unsigned it;            /* total batch iterations (reference OptimizationStatistics::it) */
    double fval;            /* best objective (reference-native, lagged value) */
    int best_sp;            /* global index of the SP that gave fval */
    double loop_s;          /* iteration-loop seconds (totalTrueComputationTime convention) */
    double wall_s;
    int threads_used;
    int total_sp;           /* after the reference clamp/rounding */
    unsigned constraint;    /* after the reference clamp */
} oracle_stats;

/* glibc rand_r restated is NOT used here: the oracle calls the C library's rand_r,
 * which is what the reference called.  The HIP library carries its own device
 * restatement, checked against this. */
unsigned oracle_rand_r(unsigned* seed);

/* reference generateProblem(n, m, h_B, LD_M, LD_N, uniformScale)
 * (gpower_problem_generator.h:51-76): column i is m values -1+2U, then multiplied
 * by its own 2-norm (uniformScale=0) or by U/norm (uniformScale=1).  Writes A[j+i*ldA]. */
void oracle_generate_problem_f64(int m, int n, double* A, int ldA, int uniform_scale);
void oracle_generate_problem_f32(int m, int n, float* A, int ldA, int uniform_scale);

/* reference getSignleStartingPoint (gpower_commons.h:21-47) for global SP index k:
 * writes V[0..n) (V must be zeroed by the caller for constrained formulations). */
void oracle_starting_point_f64(double* V, unsigned n, unsigned s, int constrained, unsigned k);
void oracle_starting_point_f32(float* V, unsigned n, unsigned s, int constrained, unsigned k);

/* Thresholding steps, in place, return the 2-norm of the result. */
double oracle_hard_threshold_f64(double* x, unsigned n, unsigned s, int rule, double* tau_out);
float  oracle_hard_threshold_f32(float* x, unsigned n, unsigned s, int rule, float* tau_out);
double oracle_soft_threshold_f64(double* x, unsigned n, unsigned s, double* lambda_out, int* found_out);
float  oracle_soft_threshold_f32(float* x, unsigned n, unsigned s, float* lambda_out, int* found_out);
double oracle_soft_threshold_gpu_f64(double* x, unsigned n, unsigned s, double* lambda_out, int* found_out);
float  oracle_soft_threshold_gpu_f32(float* x, unsigned n, unsigned s, float* lambda_out, int* found_out);
/* penalised thresholds return the penalised objective f (L1: sqrt of sum) */
double oracle_l0_pen_threshold_f64(double* x, unsigned n, double gamma);
float  oracle_l0_pen_threshold_f32(float* x, unsigned n, double gamma);
double oracle_l1_pen_threshold_f64(double* x, unsigned n, double gamma);
float  oracle_l1_pen_threshold_f32(float* x, unsigned n, double gamma);

/*
 * Whole solve.  X0: n x total_sp column-major starting points, or NULL for the
 * reference's rand_r starting points (SP k seeded with k).
 * Optional per-SP outputs (length total_sp after rounding; may be NULL):
 *   sp_val   final reference-native value f
 *   sp_iters iterations this SP performed
 *   sp_nnz   nonzeros of its final iterate
 *   sp_err   its last relative change |f - f_prev| / f_prev (reference termination_criteria.h:31-36)
 *   X_final  n x total_sp final iterates (unnormalised as the reference holds them)
 * x_out (length n) receives the best iterate normalised to unit 2-norm.
 * Returns the best value (as the reference's denseDataSolver does).
 */
double oracle_solve_f64(const double* A, int ldA, int m, int n, const oracle_settings* st,
                        const double* X0, double* x_out, oracle_stats* stats,
                        double* sp_val, int* sp_iters, int* sp_nnz, double* sp_err, double* X_final);
float oracle_solve_f32(const float* A, int ldA, int m, int n, const oracle_settings* st,
                       const float* X0, float* x_out, oracle_stats* stats,
                       float* sp_val, int* sp_iters, int* sp_nnz, float* sp_err, float* X_final);

/* Explained variance of a unit vector x in double: |A x|_2^2 (l2=1) or |A x|_1. */
double oracle_explained_variance_f64(const double* A, int ldA, int m, int n, const double* x, int l2);
double oracle_explained_variance_f32(const float* A, int ldA, int m, int n, const float* x, int l2);

/* host CBLAS availability for gemm_backend=1 (returns library name or NULL) */
const char* oracle_cblas_name(void);

#ifdef __cplusplus
}
#endif
#endif
