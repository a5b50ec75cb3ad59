/*
 * CPU oracle for the AM sparse-PCA solver -- TEST INFRASTRUCTURE ONLY.
 * See spca_oracle.h for what it restates and what pins it.
 */
#define _GNU_SOURCE
#include "spca_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double oracle_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

unsigned oracle_rand_r(unsigned* seed) { return (unsigned)rand_r(seed); }

/* ---- optional host CBLAS, used only for timing the CPU baseline -------- */
#define ORACLE_COLMAJOR 102
#define ORACLE_NOTRANS 111
#define ORACLE_TRANS 112
typedef void (*sgemm_fn)(int, int, int, int, int, int, float, const float*, int, const float*, int,
                         float, float*, int);
typedef void (*dgemm_fn)(int, int, int, int, int, int, double, const double*, int, const double*,
                         int, double, double*, int);
static sgemm_fn cblas_gemm_f32 = NULL;
static dgemm_fn cblas_gemm_f64 = NULL;
static const char* cblas_lib = NULL;
typedef void (*setthreads_fn)(int);
static setthreads_fn blas_set_threads = NULL;

__attribute__((constructor)) static void oracle_load_cblas(void) {
    const char* env = getenv("SPCA_ORACLE_CBLAS");
    const char* cands[] = {env, "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so.1",
                           "libopenblas.so.0", "libcblas.so.3", NULL};
    for (int i = 0; i < (int)(sizeof(cands) / sizeof(cands[0])); i++) {
        if (!cands[i]) continue;
        void* h = dlopen(cands[i], RTLD_NOW | RTLD_LOCAL);
        if (!h) continue;
        sgemm_fn s = (sgemm_fn)dlsym(h, "cblas_sgemm");
        dgemm_fn d = (dgemm_fn)dlsym(h, "cblas_dgemm");
        if (s && d) {
            cblas_gemm_f32 = s;
            cblas_gemm_f64 = d;
            cblas_lib = cands[i];
            blas_set_threads = (setthreads_fn)dlsym(h, "MKL_Set_Num_Threads");
            if (!blas_set_threads) blas_set_threads = (setthreads_fn)dlsym(h, "openblas_set_num_threads");
            return;
        }
        dlclose(h);
    }
}

const char* oracle_cblas_name(void) { return cblas_lib; }

static void oracle_blas_threads(int t) {
    if (blas_set_threads) blas_set_threads(t > 0 ? t : 1);
}

/* ---- instantiate the body for float and double --------------------------- */
#define R float
#define SFX(x) x##_f32
#define sqrt_r_f32 sqrtf
#include "spca_oracle_body.h"
#undef R
#undef SFX

#define R double
#define SFX(x) x##_f64
#define sqrt_r_f64 sqrt
#include "spca_oracle_body.h"
#undef R
#undef SFX
