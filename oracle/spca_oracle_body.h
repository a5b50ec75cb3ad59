/*
 * Precision-generic body of the CPU oracle.  Included twice by spca_oracle.c
 * with R = float / double and SFX(x) appending _f32 / _f64.
 * TEST INFRASTRUCTURE ONLY (see spca_oracle.h).
 *
 * Arithmetic that the reference writes itself (threshold sums, lambda root,
 * error, scaling) is restated in R exactly as written there.  Arithmetic the
 * reference delegates to BLAS (gemm, nrm2, asum) is computed here with fixed
 * order and double accumulation; the reference's BLAS summation order is a
 * third-party detail no reference test pins.
 */

/* ---- small helpers ---------------------------------------------------- */

/* reference sgn(), c++/src/utils/various.h:44-46 : sgn(0) = 0 */
static inline int SFX(sgn)(R v) { return (v > (R)0) - (v < (R)0); }
/* reference myabs(), various.h:48-50 */
static inline R SFX(myabs)(R v) { return (v > 0) ? v : -v; }

/* reference computeTheError, c++/src/utils/termination_criteria.h:29-36 */
static inline R SFX(compute_error)(R f, R fprev) { return SFX(myabs)(f - fprev) / fprev; }

/* BLAS nrm2 / asum stand-ins: fixed order, double accumulation */
static R SFX(nrm2)(const R* x, int n) {
    double s = 0;
    for (int i = 0; i < n; i++) s += (double)x[i] * (double)x[i];
    return (R)sqrt(s);
}
static R SFX(asum)(const R* x, int n) {
    double s = 0;
    for (int i = 0; i < n; i++) s += fabs((double)x[i]);
    return (R)s;
}
/* cblas ?scal(n, alpha, x): x_i <- alpha * x_i */
static void SFX(scal)(R* x, int n, R alpha) {
    for (int i = 0; i < n; i++) x[i] = alpha * x[i];
}
static unsigned SFX(nnz)(const R* x, int n) {
    unsigned c = 0;
    for (int i = 0; i < n; i++) c += (x[i] != 0);
    return c;
}

/* ---- products ----------------------------------------------------------- */

#define ORACLE_ROW_BLOCK 512

/* Host-BLAS product over the active columns only (timing backend): the active columns of the
 * right operand are packed, multiplied in one gemm, and unpacked into their own columns, so
 * per-SP termination costs the active SPs' work alone.  trans: A^T (n x m) instead of A. */
static void SFX(gemm_packed)(int trans, const R* A, int ld, int rows, int inner, const R* X, int ldx, R* Z,
                             int ldz, int nb, const unsigned char* active) {
    int na = 0;
    for (int j = 0; j < nb; j++) na += active[j] != 0;
    if (na == 0) return;
    R* Xc = (R*)malloc(sizeof(R) * (size_t)inner * na);
    R* Zc = (R*)malloc(sizeof(R) * (size_t)rows * na);
    for (int j = 0, c = 0; j < nb; j++)
        if (active[j]) memcpy(Xc + (size_t)inner * c++, X + (long)j * ldx, sizeof(R) * inner);
    SFX(cblas_gemm)(ORACLE_COLMAJOR, trans ? ORACLE_TRANS : ORACLE_NOTRANS, ORACLE_NOTRANS, rows, na, inner, (R)1,
                    A, ld, Xc, inner, (R)0, Zc, rows);
    for (int j = 0, c = 0; j < nb; j++)
        if (active[j]) memcpy(Z + (long)j * ldz, Zc + (size_t)rows * c++, sizeof(R) * rows);
    free(Xc);
    free(Zc);
}

/* Z[:, j] = A * X[:, j] for columns j with active[j] (active may be NULL). */
static void SFX(gemm_ax)(const oracle_settings* st, const R* A, int ld, int m, int n,
                         const R* X, int ldx, R* Z, int ldz, int nb, const unsigned char* active) {
    if (st->gemm_backend == 1 && SFX(cblas_gemm) && active == NULL) {
        SFX(cblas_gemm)(ORACLE_COLMAJOR, ORACLE_NOTRANS, ORACLE_NOTRANS, m, nb, n, (R)1, A, ld, X, ldx,
                        (R)0, Z, ldz);
        return;
    }
    if (st->gemm_backend == 1 && SFX(cblas_gemm)) {
        SFX(gemm_packed)(0, A, ld, m, n, X, ldx, Z, ldz, nb, active);
        return;
    }
    int nblk = (m + ORACLE_ROW_BLOCK - 1) / ORACLE_ROW_BLOCK;
    long tasks = (long)nb * nblk;
#pragma omp parallel for schedule(dynamic, 1) num_threads(st->threads > 0 ? st->threads : 1)
    for (long t = 0; t < tasks; t++) {
        int j = (int)(t / nblk), b = (int)(t % nblk);
        if (active && !active[j]) continue;
        int i0 = b * ORACLE_ROW_BLOCK, i1 = i0 + ORACLE_ROW_BLOCK < m ? i0 + ORACLE_ROW_BLOCK : m;
        double acc[ORACLE_ROW_BLOCK];
        for (int i = 0; i < i1 - i0; i++) acc[i] = 0;
        const R* x = X + (long)j * ldx;
        for (int k = 0; k < n; k++) {
            double xk = (double)x[k];
            if (xk == 0) continue;
            const R* a = A + (long)k * ld + i0;
            for (int i = 0; i < i1 - i0; i++) acc[i] += (double)a[i] * xk;
        }
        R* z = Z + (long)j * ldz + i0;
        for (int i = 0; i < i1 - i0; i++) z[i] = (R)acc[i];
    }
}

/* V[:, j] = A^T * Z[:, j] */
static void SFX(gemm_atz)(const oracle_settings* st, const R* A, int ld, int m, int n,
                          const R* Z, int ldz, R* V, int ldv, int nb, const unsigned char* active) {
    if (st->gemm_backend == 1 && SFX(cblas_gemm) && active == NULL) {
        SFX(cblas_gemm)(ORACLE_COLMAJOR, ORACLE_TRANS, ORACLE_NOTRANS, n, nb, m, (R)1, A, ld, Z, ldz,
                        (R)0, V, ldv);
        return;
    }
    if (st->gemm_backend == 1 && SFX(cblas_gemm)) {
        SFX(gemm_packed)(1, A, ld, n, m, Z, ldz, V, ldv, nb, active);
        return;
    }
    int nblk = (n + ORACLE_ROW_BLOCK - 1) / ORACLE_ROW_BLOCK;
    long tasks = (long)nb * nblk;
#pragma omp parallel for schedule(dynamic, 1) num_threads(st->threads > 0 ? st->threads : 1)
    for (long t = 0; t < tasks; t++) {
        int j = (int)(t / nblk), b = (int)(t % nblk);
        if (active && !active[j]) continue;
        int k0 = b * ORACLE_ROW_BLOCK, k1 = k0 + ORACLE_ROW_BLOCK < n ? k0 + ORACLE_ROW_BLOCK : n;
        const R* z = Z + (long)j * ldz;
        for (int k = k0; k < k1; k++) {
            const R* a = A + (long)k * ld;
            double s = 0;
            for (int i = 0; i < m; i++) s += (double)a[i] * (double)z[i];
            V[(long)j * ldv + k] = (R)s;
        }
    }
}

/* ---- data and starting points ------------------------------------------ */

/* reference generateProblem(n, m, h_B, LD_M, LD_N, uniformScale),
 * c++/src/problem_generators/gpower_problem_generator.h:51-76 */
void SFX(oracle_generate_problem)(int m, int n, R* A, int ldA, int uniform_scale) {
    unsigned seed = 0;
    for (int i = 0; i < n; i++) {
        R total = 0;
        for (int j = 0; j < m; j++) {
            R tmp = (R)rand_r(&seed) / (R)RAND_MAX;
            tmp = -1 + 2 * tmp;
            A[j + (long)i * ldA] = tmp;
            total += tmp * tmp;
        }
        total = SFX(sqrt_r)(total);
        if (uniform_scale) {
            /* (F)rand_r / (RAND_MAX + 0.0) is a double expression */
            total = (R)(1 / total * (+((R)rand_r(&seed) / (RAND_MAX + 0.0))));
        }
        for (int j = 0; j < m; j++) A[j + (long)i * ldA] = A[j + (long)i * ldA] * total;
    }
}

/* reference getSignleStartingPoint, c++/src/gpower/gpower_commons.h:21-47,
 * with myseed = global SP index k. */
void SFX(oracle_starting_point)(R* V, unsigned n, unsigned s, int constrained, unsigned k) {
    unsigned seed = k;
    if (constrained) {
        R tmp_norm = 0;
        for (unsigned i = 0; i < s; i++) {
            unsigned idx = (unsigned)(int)((R)n * (R)rand_r(&seed) / (R)RAND_MAX);
            if (idx == n) idx--;
            R tmp = (R)rand_r(&seed) / (R)RAND_MAX;
            V[idx] = tmp;
            tmp_norm += tmp * tmp;
        }
        SFX(scal)(V, (int)n, (R)1 / SFX(sqrt_r)(tmp_norm));
    } else {
        for (unsigned i = 0; i < n; i++) {
            R tmp = (R)rand_r(&seed) / (R)RAND_MAX;
            tmp = -1 + 2 * tmp;
            V[i] = tmp;
        }
    }
}

/* ---- thresholds ---------------------------------------------------------- */

static int SFX(cmp_desc)(const void* a, const void* b) {
    R x = *(const R*)a, y = *(const R*)b;
    return (x < y) - (x > y);
}

/* tau by the SET rule: simulation of the reference std::set with comparator
 * lhs.value > rhs.value (c++/src/utils/thresh_functions.h:41-60,100-118);
 * 'set' holds distinct magnitudes in descending order. */
static R SFX(tau_set)(const R* x, unsigned n, unsigned k, R* set) {
    unsigned cnt = 0;
#define SET_INSERT(val)                                                        \
    do {                                                                       \
        R v_ = (val);                                                          \
        unsigned lo_ = 0, hi_ = cnt;                                           \
        while (lo_ < hi_) {                                                    \
            unsigned mid_ = (lo_ + hi_) / 2;                                   \
            if (set[mid_] > v_) lo_ = mid_ + 1; else hi_ = mid_;               \
        }                                                                      \
        if (!(lo_ < cnt && set[lo_] == v_)) {                                  \
            memmove(set + lo_ + 1, set + lo_, (cnt - lo_) * sizeof(R));        \
            set[lo_] = v_;                                                     \
            cnt++;                                                             \
        }                                                                      \
    } while (0)
    for (unsigned j = 0; j < k && j < n; j++) SET_INSERT(fabs(x[j]));
    for (unsigned j = k; j < n; j++) {
        R tmp = (R)fabs(x[j]);
        R thresh = set[cnt - 1];
        if (tmp > thresh) {
            cnt--; /* erase the smallest */
            SET_INSERT(tmp);
        }
    }
#undef SET_INSERT
    return set[cnt - 1];
}

/* reference k_hard_thresholding, c++/src/utils/thresh_functions.h:120-145 */
R SFX(oracle_hard_threshold)(R* x, unsigned n, unsigned s, int rule, R* tau_out) {
    R* buf = (R*)malloc(sizeof(R) * (n ? n : 1));
    R tau;
    if (s < 1) s = 1; /* the reference indexes out of range for s = 0 */
    if (s > n) s = n;
    if (rule == OR_HARD_SET) {
        tau = SFX(tau_set)(x, n, s, buf);
    } else {
        /* sort path: |sorted ascending[length - k]| = s-th largest |x| */
        for (unsigned i = 0; i < n; i++) buf[i] = (R)fabs(x[i]);
        qsort(buf, n, sizeof(R), SFX(cmp_desc));
        tau = buf[s - 1];
    }
    free(buf);
    R norm = 0;
    for (unsigned i = 0; i < n; i++) {
        R val = x[i];
        if ((R)fabs(val) < tau) x[i] = 0;
        else norm += val * val;
    }
    if (tau_out) *tau_out = tau;
    return SFX(sqrt_r)(norm);
}

/* reference soft_thresholding, c++/src/utils/thresh_functions.h:160-254 */
R SFX(oracle_soft_threshold)(R* x, unsigned length, unsigned constrain, R* lambda_out, int* found_out) {
    R* myvector = (R*)malloc(sizeof(R) * (length ? length : 1));
    for (unsigned i = 0; i < length; i++) myvector[i] = (R)fabs(x[i]);
    qsort(myvector, length, sizeof(R), SFX(cmp_desc)); /* myvector[i] = i-th largest |x| */
    R sq_constr = (R)sqrt(constrain + 0.0);
    R lambda_Low = 0;
    R lambda_High = myvector[0];
    R linfty = myvector[0];
    R sum_abs_x = 0, sum_abs_x2 = 0;
    R epsilon = (R)0.000001;
    R tmp;
    int total_elements = 0;
    R w = 0;
    R subgradRightOld = sq_constr;
    int found = 0;
    for (unsigned i = 0; i < length; i++) {
        tmp = myvector[i];
        sum_abs_x += tmp;
        sum_abs_x2 += tmp * tmp;
        total_elements++;
        lambda_High = tmp;
        if (i < length - 1) lambda_Low = myvector[i + 1];
        else lambda_Low = 0;
        R subgradLeft = (sum_abs_x - total_elements * tmp) /
                        SFX(sqrt_r)(sum_abs_x2 - 2 * tmp * sum_abs_x + total_elements * tmp * tmp);
        if (subgradLeft > sq_constr && subgradRightOld < sq_constr) {
            w = tmp;
            found = 1;
            break;
        }
        R subgradRight = (sum_abs_x - total_elements * lambda_Low) /
                         SFX(sqrt_r)(sum_abs_x2 - 2 * lambda_Low * sum_abs_x +
                                     total_elements * lambda_Low * lambda_Low);
        R a = total_elements * (total_elements - sq_constr * sq_constr);
        R b = 2 * sum_abs_x * sq_constr * sq_constr - 2 * total_elements * sum_abs_x;
        R c = sum_abs_x * sum_abs_x - sq_constr * sq_constr * sum_abs_x2;
        w = (-b - SFX(sqrt_r)(b * b - 4 * a * c)) / (2 * a);
        if (w > lambda_Low && w < lambda_High) { found = 1; break; }
        w = (-b + SFX(sqrt_r)(b * b - 4 * a * c)) / (2 * a);
        if (w > lambda_Low && w < lambda_High && w < linfty - epsilon) { found = 1; break; }
        subgradRightOld = subgradRight;
    }
    free(myvector);
    R norm = 0;
    for (unsigned i = 0; i < length; i++) {
        R t = x[i];
        R t2 = (R)fabs(t) - w;
        if (t2 > 0) {
            const R val = t2 * SFX(sgn)(t);
            x[i] = val;
            norm += val * val;
        } else {
            x[i] = 0;
        }
    }
    if (lambda_out) *lambda_out = w;
    if (found_out) *found_out = found;
    return SFX(sqrt_r)(norm);
}

/* reference GPU-path lambda, compute_soft_thresholding_parameter
 * (c++/src/gpugpower/my_cublas_wrapper.h:335-375), applied with gpu_soft_treshholding
 * (:377-395) as perform_hard_and_soft_thresholding_with_sorting does (:571-634):
 * scan |x| descending; at element i evaluate the ratio at lambda = |x|_(i) over elements
 * 0..i; at the first i where it reaches sqrt(s) drop element i and solve the quadratic on
 * elements 0..i-1 ("-" root).  When no element triggers, all n elements enter the quadratic. */
R SFX(oracle_soft_threshold_gpu)(R* x, unsigned length, unsigned constrain, R* lambda_out, int* found_out) {
    R* myvector = (R*)malloc(sizeof(R) * (length ? length : 1));
    for (unsigned i = 0; i < length; i++) myvector[i] = (R)fabs(x[i]);
    qsort(myvector, length, sizeof(R), SFX(cmp_desc));
    R w = (R)sqrt((double)constrain);
    R sum_abs_x = 0, sum_abs_x2 = 0;
    R tmp;
    int total_elements = 0;
    int found = 0;
    for (unsigned i = 0; i < length; i++) {
        tmp = myvector[i];
        sum_abs_x += tmp;
        sum_abs_x2 += tmp * tmp;
        total_elements++;
        R subgrad = (sum_abs_x - total_elements * tmp) /
                    SFX(sqrt_r)(sum_abs_x2 - 2 * tmp * sum_abs_x + total_elements * tmp * tmp);
        if (subgrad >= w) {
            sum_abs_x -= tmp;
            sum_abs_x2 -= tmp * tmp;
            total_elements--;
            found = 1;
            break;
        }
    }
    free(myvector);
    R a = total_elements * (total_elements - w * w);
    R b = 2 * sum_abs_x * w * w - 2 * total_elements * sum_abs_x;
    R c = sum_abs_x * sum_abs_x - w * w * sum_abs_x2;
    R lam = (-b - SFX(sqrt_r)(b * b - 4 * a * c)) / (2 * a);
    R norm = 0;
    for (unsigned i = 0; i < length; i++) {
        R t = x[i];
        R t2 = (R)fabs(t) - lam;
        if (t2 > 0) {
            const R val = t2 * SFX(sgn)(t);
            x[i] = val;
            norm += val * val;
        } else {
            x[i] = 0;
        }
    }
    if (lambda_out) *lambda_out = lam;
    if (found_out) *found_out = found;
    return SFX(sqrt_r)(norm);
}

/* reference L0_penalized_thresholding body, c++/src/gpower/sparse_PCA_thresholding.h:76-85;
 * penaltyParameter is a double, so the difference is formed in double. */
R SFX(oracle_l0_pen_threshold)(R* x, unsigned n, double gamma) {
    R f = 0;
    for (unsigned i = 0; i < n; i++) {
        R const tmp = x[i];
        R tmp2 = (R)((double)(tmp * tmp) - gamma);
        if (tmp2 > 0) f += tmp2;
        else x[i] = 0;
    }
    return f;
}

/* reference L1_penalized_thresholding body, sparse_PCA_thresholding.h:33-47 */
R SFX(oracle_l1_pen_threshold)(R* x, unsigned n, double gamma) {
    R f = 0;
    for (unsigned i = 0; i < n; i++) {
        R const tmp = x[i];
        R tmp2 = (R)((double)(R)fabs(tmp) - gamma);
        if (tmp2 > 0) {
            f += tmp2 * tmp2;
            x[i] = tmp2 * SFX(sgn)(tmp);
        } else {
            x[i] = 0;
        }
    }
    return SFX(sqrt_r)(f);
}

/* ---- one AM iteration over a batch of nb columns ---------------------- */

typedef struct {
    R val;      /* value of the last completed iteration */
    R tmp;      /* L1-variance objective held across the products */
    R err;      /* current_error */
} SFX(spstate);

static int SFX(is_constrained)(int f) { return f < 4; }
static int SFX(is_l1_constrained)(int f) { return !(f == OR_L0_CONS_L1 || f == OR_L0_CONS_L2); }
static int SFX(is_l1_penalized)(int f) { return !(f == OR_L0_PEN_L1 || f == OR_L0_PEN_L2); }
static int SFX(is_l1_variance)(int f) { return f == 1 || f == 3 || f == 5 || f == 7; }

/* reference perform_one_iteration_for_constrained_pca (gpower_commons.h:67-136) and
 * perform_one_iteration_for_penalized_pca (gpower_commons.h:139-177).
 * Returns the batch max error (NaN skipped, as the reference's '<' update does). */
static R SFX(iterate)(const oracle_settings* st, const R* A, int ld, int m, int n, R* V, R* Z, int nb,
                      SFX(spstate)* vals, const unsigned char* active, unsigned s) {
    const int f = st->formulation;
    SFX(gemm_ax)(st, A, ld, m, n, V, n, Z, m, nb, active);
    if (SFX(is_constrained)(f)) {
        if (SFX(is_l1_variance)(f)) {
            for (int j = 0; j < nb; j++) {
                if (active && !active[j]) continue;
                vals[j].tmp = SFX(asum)(Z + (long)m * j, m);
                for (int i = 0; i < m; i++) Z[(long)m * j + i] = (R)SFX(sgn)(Z[(long)m * j + i]);
            }
        }
    } else {
        for (int j = 0; j < nb; j++) {
            if (active && !active[j]) continue;
            R* z = Z + (long)m * j;
            if (SFX(is_l1_variance)(f)) {
                for (int i = 0; i < m; i++) z[i] = (R)SFX(sgn)(z[i]);
            } else {
                R tn = SFX(nrm2)(z, m);
                SFX(scal)(z, m, 1 / tn);
            }
        }
    }
    SFX(gemm_atz)(st, A, ld, m, n, Z, m, V, n, nb, active);
    R maxerr = 0;
    /* per-column step: independent across columns, parallel and deterministic */
    R* errs = (R*)malloc(sizeof(R) * (nb ? nb : 1));
#pragma omp parallel for schedule(dynamic, 1) num_threads(st->threads > 0 ? st->threads : 1)
    for (int j = 0; j < nb; j++) {
        if (active && !active[j]) continue;
        R* v = V + (long)n * j;
        R fval = 0;
        if (SFX(is_constrained)(f)) {
            if (f == OR_L0_CONS_L2 || f == OR_L1_CONS_L2) fval = SFX(nrm2)(Z + (long)m * j, m);
            R norm_of_x;
            if (SFX(is_l1_constrained)(f))
                norm_of_x = st->lambda_rule == OR_LAMBDA_GPU ? SFX(oracle_soft_threshold_gpu)(v, n, s, NULL, NULL)
                                                             : SFX(oracle_soft_threshold)(v, n, s, NULL, NULL);
            else norm_of_x = SFX(oracle_hard_threshold)(v, n, s, st->hard_rule, NULL);
            SFX(scal)(v, n, 1 / norm_of_x);
            if (SFX(is_l1_variance)(f)) fval = vals[j].tmp;
        } else {
            if (SFX(is_l1_penalized)(f)) fval = SFX(oracle_l1_pen_threshold)(v, n, st->penalty);
            else fval = SFX(oracle_l0_pen_threshold)(v, n, st->penalty);
        }
        errs[j] = SFX(compute_error)(fval, vals[j].val);
        vals[j].err = errs[j];
        vals[j].val = fval;
    }
    for (int j = 0; j < nb; j++) {
        if (active && !active[j]) continue;
        if (maxerr < errs[j]) maxerr = errs[j];
    }
    free(errs);
    return maxerr;
}

/* reference termination_criteria, termination_criteria.h:38-44 */
static int SFX(terminated)(R err, int it, double tol) { return it > 0 && err < tol; }

/* ---- drivers --------------------------------------------------------------- */

R SFX(oracle_solve)(const R* A, int ldA, int m, int n, const oracle_settings* st_in, const R* X0,
                    R* x_out, oracle_stats* stats, R* sp_val, int* sp_iters, int* sp_nnz, R* sp_err, R* X_final) {
    oracle_settings st = *st_in;
    double t_wall0 = oracle_now();
    if (st.gemm_backend == 1) oracle_blas_threads(st.threads);
    /* reference chceckInputAndModifyIt(n), optimization_settings.h:156-168 */
    unsigned s = st.constraint;
    if (s > (unsigned)n) s = (unsigned)n;
    int batch = st.batch == 0 ? st.total_sp : st.batch;
    int total_batches = st.total_sp / batch;
    if (total_batches * batch < st.total_sp) total_batches++;
    int total = total_batches * batch;
    const int constrained = SFX(is_constrained)(st.formulation);

    R* Z = (R*)calloc((size_t)m * batch, sizeof(R));
    R* V = (R*)calloc((size_t)n * batch, sizeof(R));
    SFX(spstate)* vals = (SFX(spstate)*)calloc(batch, sizeof(SFX(spstate)));
    unsigned char* active = (unsigned char*)calloc(batch, 1);
    int* cur_it = (int*)calloc(batch, sizeof(int));
    int* order = (int*)calloc(batch, sizeof(int));
    R the_best = -1;
    int best_sp = -1;
    unsigned total_iterations = 0;
    double loop_s = 0;
    memset(x_out, 0, sizeof(R) * n);

#define LOAD_SP(slot, k)                                                          \
    do {                                                                          \
        R* v_ = V + (long)n * (slot);                                             \
        if (X0) memcpy(v_, X0 + (long)n * (k), sizeof(R) * n);                    \
        else {                                                                    \
            if (constrained) memset(v_, 0, sizeof(R) * n);                        \
            SFX(oracle_starting_point)(v_, (unsigned)n, s, constrained, (unsigned)(k)); \
        }                                                                         \
    } while (0)

    if (st.use_otf) {
        /* reference OTF branch, sparse_PCA_solver.h:83-169 */
        for (int j = 0; j < batch; j++) { LOAD_SP(j, j); order[j] = j; }
        int generated = batch;
        int do_iterate = 1;
        double t0 = oracle_now();
        while (do_iterate) {
            total_iterations++;
            SFX(iterate)(&st, A, ldA, m, n, V, Z, batch, vals, NULL, s);
            do_iterate = 0;
            int new_points = 0;
            for (int i = 0; i < batch; i++) {
                cur_it[i]++;
                if (SFX(terminated)(vals[i].err, cur_it[i], st.tolerance) || cur_it[i] >= st.max_iter) {
                    if (the_best < vals[i].val) {
                        the_best = vals[i].val;
                        best_sp = order[i];
                        memcpy(x_out, V + (long)n * i, sizeof(R) * n);
                    }
                    if (sp_val) sp_val[order[i]] = vals[i].val;
                    if (sp_iters) sp_iters[order[i]] = cur_it[i];
                    if (sp_nnz) sp_nnz[order[i]] = (int)SFX(nnz)(V + (long)n * i, n);
                    if (sp_err) sp_err[order[i]] = vals[i].err;
                    if (X_final) memcpy(X_final + (long)n * order[i], V + (long)n * i, sizeof(R) * n);
                    if (generated < total) {
                        vals[i].val = 0; vals[i].tmp = 0; vals[i].err = 0;
                        order[i] = generated;
                        new_points++;
                        generated++;
                        cur_it[i] = 0;
                        do_iterate = 1;
                    }
                } else {
                    do_iterate = 1;
                }
            }
            if (new_points > 0)
                for (int j = 0; j < batch; j++)
                    if (cur_it[j] == 0) LOAD_SP(j, order[j]);
        }
        loop_s += oracle_now() - t0;
    } else {
        for (int bt = 0; bt < total_batches; bt++) {
            int shift = bt * batch;
            memset(V, 0, sizeof(R) * (size_t)n * batch);
            for (int j = 0; j < batch; j++) {
                LOAD_SP(j, shift + j);
                vals[j].val = 0; vals[j].tmp = 0; vals[j].err = 0;
                active[j] = 1;
                cur_it[j] = 0;
            }
            double t0 = oracle_now();
            for (int it = 0; it < st.max_iter; it++) {
                total_iterations++;
                if (st.termination == OR_TERM_PER_SP) {
                    SFX(iterate)(&st, A, ldA, m, n, V, Z, batch, vals, active, s);
                    int any = 0;
                    for (int j = 0; j < batch; j++) {
                        if (!active[j]) continue;
                        cur_it[j]++;
                        if (SFX(terminated)(vals[j].err, it, st.tolerance) || cur_it[j] >= st.max_iter)
                            active[j] = 0;
                        else any = 1;
                    }
                    if (!any) break;
                } else {
                    R error = SFX(iterate)(&st, A, ldA, m, n, V, Z, batch, vals, NULL, s);
                    for (int j = 0; j < batch; j++) cur_it[j]++;
                    if (SFX(terminated)(error, it, st.tolerance)) break;
                }
            }
            loop_s += oracle_now() - t0;
            /* best of batch, sparse_PCA_solver.h:212-231 */
            int sel = 0;
            R best_value = vals[0].val;
            for (int i = 1; i < batch; i++)
                if (vals[i].val > best_value) { best_value = vals[i].val; sel = i; }
            for (int i = 0; i < batch; i++) {
                if (sp_val) sp_val[shift + i] = vals[i].val;
                if (sp_iters) sp_iters[shift + i] = cur_it[i];
                if (sp_nnz) sp_nnz[shift + i] = (int)SFX(nnz)(V + (long)n * i, n);
                if (sp_err) sp_err[shift + i] = vals[i].err;
            }
            if (X_final) memcpy(X_final + (long)n * shift, V, sizeof(R) * (size_t)n * batch);
            if (the_best < best_value) {
                the_best = best_value;
                best_sp = shift + sel;
                memcpy(x_out, V + (long)n * sel, sizeof(R) * n);
            }
        }
    }
#undef LOAD_SP
    /* final x <- x / |x|, sparse_PCA_solver.h:236-237 */
    R nx = SFX(nrm2)(x_out, n);
    SFX(scal)(x_out, n, 1 / nx);
    if (stats) {
        stats->it = total_iterations;
        stats->fval = (double)the_best;
        stats->best_sp = best_sp;
        stats->loop_s = loop_s;
        stats->wall_s = oracle_now() - t_wall0;
        stats->threads_used = st.threads > 0 ? st.threads : 1;
        stats->total_sp = total;
        stats->constraint = s;
    }
    free(Z); free(V); free(vals); free(active); free(cur_it); free(order);
    return the_best;
}

double SFX(oracle_explained_variance)(const R* A, int ldA, int m, int n, const R* x, int l2) {
    double* z = (double*)calloc(m, sizeof(double));
    for (int k = 0; k < n; k++) {
        double xk = (double)x[k];
        if (xk == 0) continue;
        for (int i = 0; i < m; i++) z[i] += (double)A[i + (long)k * ldA] * xk;
    }
    double ev = 0;
    for (int i = 0; i < m; i++) ev += l2 ? z[i] * z[i] : fabs(z[i]);
    free(z);
    return ev;
}
