// Host driver of the MI355X AM sparse-PCA solver and the C ABI (include/spca.h).
//
// The batched AM loop mirrors the reference GPU driver
// (c++/src/gpugpower/gpu_sparse_PCA_solver.h:29-211) and the CPU batch driver
// (c++/src/gpower/sparse_PCA_solver.h:170-243), re-planned for one MI355X:
//   - the matrix stays resident in HBM in the layouts the products want (row major and
//     column major, FP32 and/or bf16 hi+lo and/or fp16), keyed so jobs can share it;
//   - every per-SP reduction stays on the device; the host only reads a convergence flag
//     every `check_every` iterations (kernels enqueued after the batch has finished see
//     the device flag and return at once, so the iteration count is exact);
//   - per-phase kernel time is taken with hipEvents on the solver's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spca.h"
#include "spca_internal.h"

namespace spca {

struct Matrix {
    int m = 0, n = 0, mp = 0, np = 0;
    uint64_t key = 0;
    unsigned storage = 0;
    float* a32_cm = nullptr;  // mp x np column major: FP32 master, A^T*Z operand
    float* a32_rm = nullptr;  // mp x np row major (ld np): A*X operand
    uint16_t* hi_cm = nullptr;
    uint16_t* lo_cm = nullptr;
    uint16_t* hi_rm = nullptr;
    uint16_t* lo_rm = nullptr;
    uint16_t* h_cm = nullptr;
    uint16_t* h_rm = nullptr;
    int8_t* q_cm = nullptr;      // fixed-point pieces of A for the exact A^T * sgn(z) (4 x mp x np bytes)
    float* qscale = nullptr;     // their per-column power-of-two scales (np)
    float* rowmax = nullptr;     // max_k |A_ik| per row (mp): the sign refinement's error bound
    // sparse matrices (spca_matrix_upload_csc): CSC and CSR arrays, optional centring means;
    // no dense storage exists for them
    bool sparse = false;
    int nnz = 0;
    int* cptr = nullptr;
    int* ridx = nullptr;
    float* cval = nullptr;
    int* rptr = nullptr;
    int* cidx = nullptr;
    float* rval = nullptr;
    float* mu = nullptr;
    float* rmu = nullptr;
    uint64_t bytes = 0;
    int refs = 0;
};

struct Workspace {
    int np = 0, mp = 0, bp = 0, splits = 0;
    float* X = nullptr;
    uint16_t* Xhi = nullptr;
    uint16_t* Xlo = nullptr;
    float* Zpart = nullptr;
    float* Z = nullptr;
    uint16_t* Zhi = nullptr;
    uint16_t* Zlo = nullptr;
    float* V = nullptr;
    float* fstate = nullptr;  // val, fz, err, errmax, thr  (kStateFloats x bp floats)
    int* istate = nullptr;    // iters, nnz, active, nkept, lmax (bp ints each),
                              // done (state_flag_ints(bp))
    int* kept = nullptr;      // kListMax indices per SP: the nonzeros a list step wrote
    int* lcount = nullptr;                 // step_chunks(np) ints per SP: the chunk scans' entry counts
    unsigned long long* lstage = nullptr;  // step_chunks(np) * kStageCap entries per SP: their entries
    int* gl = nullptr;        // kGatherStride ints per SP: the gathered A*X's nonzero lists (null: scan x)
    double* ev = nullptr;
    double* corr = nullptr;  // per-SP centring scalars of sparse products (4 x bp)
    // union A*X (sparse_products = 2): index flags and list (np ints each), stat (4 ints),
    // Au = A[:, U] (mp x ldu, row major) and Xu = X[U, :] (bp x ldu); ldu = 0: not allocated
    int* uflag = nullptr;
    int* uidx = nullptr;
    int* ustat = nullptr;
    float* Au = nullptr;
    float* Xu = nullptr;
    int ldu = 0;
    uint64_t bytes = 0;
};

// Everything one solve needs besides the matrix: its stream, workspace, a pinned flag for
// the host's convergence checks and hipEvents for timing.  The context owns one engine for
// synchronous calls; the request server owns one per concurrent job.
struct Engine {
    hipStream_t stream = nullptr;
    Workspace ws;
    int* h_done = nullptr;
    std::vector<hipEvent_t> events;  // pool of (start, stop) pairs for phase timing
    hipEvent_t loop_ev[2] = {nullptr, nullptr};
    const std::atomic<int>* cancel = nullptr;  // set by the request server; checked at host syncs
};

// Launch plan of the two products of one solve shape: K split of A*X and the tile kernel.
struct Plan {
    int splits = 1;
    int big = 0;       // 1: 256x256-tile kernel for both products
    double ms = 0;     // measured A*X + epilogue + A^T*Z time when tuned (0: heuristic plan)
};
struct PlanKey {
    int precision, mp, np, bp, allow_big;
    bool operator<(const PlanKey& o) const {
        if (precision != o.precision) return precision < o.precision;
        if (mp != o.mp) return mp < o.mp;
        if (np != o.np) return np < o.np;
        if (bp != o.bp) return bp < o.bp;
        return allow_big < o.allow_big;
    }
};

}  // namespace spca

struct spca_server;

struct spca_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;  // the default engine's stream
    uint64_t budget = 0;
    uint64_t used = 0;
    std::mutex alloc_mu;  // guards `used`
    std::mutex mu;        // guards the matrix table and the default engine
    std::mutex err_mu;    // guards err
    std::string err;
    std::map<uint64_t, spca::Matrix*> mats;
    uint64_t next_handle = 1;
    spca::Engine eng;  // default engine (synchronous calls)
    spca_server* server = nullptr;
    // product plans: measured per shape when tuning is on, else the fixed heuristic
    std::mutex plan_mu;
    int tuning = 0;
    std::string tune_path;
    std::map<spca::PlanKey, spca::Plan> plans;
    // device blocks: live ones by pointer (size class), freed ones kept for reuse by class.
    // `used` counts both, so live + cached device memory stays within the budget.
    std::map<void*, uint64_t> live;
    std::multimap<uint64_t, void*> cache;
    uint64_t cached = 0;
};

using namespace spca;

namespace {

const char* kVersion = "spca_mi355x 0.1 (gfx950)";

thread_local std::string tl_err;  // this thread's last error (worker threads read their own)

int fail(spca_ctx ctx, int code, const std::string& msg) {
    tl_err = msg;
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->err_mu);
        ctx->err = msg;
    }
    return code;
}

#define HIP_TRY(ctx, expr)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(ctx, e_ == hipErrorOutOfMemory ? SPCA_ENOMEM : SPCA_EDEVICE,               \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                        \
    } while (0)

// Device memory goes through a size-class cache: a freed block is kept and handed to the
// next request of its class, so solves of changing shapes (the request server) do not pay
// hipMalloc and, above all, hipFree, which waits for the whole device and would serialise
// concurrent solves.  Classes are 4 per octave (at most 19% slack); cached blocks count
// against the HBM budget and are released (largest first) when a request would exceed it.
uint64_t size_class(uint64_t b) {
    if (b <= 4096) return 4096;
    int e = 63 - __builtin_clzll(b - 1);  // 2^e < b <= 2^(e+1)
    const uint64_t base = 1ull << e, q = base / 4;
    return base + (b - base + q - 1) / q * q;
}

// Caller holds ctx->alloc_mu.  Releases cached blocks until `need` more bytes fit the budget.
void evict_for(spca_ctx ctx, uint64_t need) {
    while (ctx->used + need > ctx->budget && !ctx->cache.empty()) {
        auto it = std::prev(ctx->cache.end());
        (void)hipFree(it->second);
        ctx->used -= it->first;
        ctx->cached -= it->first;
        ctx->cache.erase(it);
    }
}

void set_err(spca_ctx ctx, std::string* err, const std::string& m) {
    if (err) {
        *err = m;
        return;
    }
    tl_err = m;
    std::lock_guard<std::mutex> g2(ctx->err_mu);
    ctx->err = m;
}

int raw_alloc(spca_ctx ctx, void** p, uint64_t b, std::string* err) {
    const uint64_t cls = size_class(b);
    std::lock_guard<std::mutex> g(ctx->alloc_mu);
    auto it = ctx->cache.find(cls);
    if (it != ctx->cache.end()) {
        *p = it->second;
        ctx->cache.erase(it);
        ctx->cached -= cls;
        ctx->live[*p] = cls;
        return SPCA_OK;
    }
    evict_for(ctx, cls);
    if (ctx->used + cls > ctx->budget) {
        set_err(ctx, err, "allocation of " + std::to_string(b) + " bytes exceeds the HBM budget");
        return SPCA_EBUDGET;
    }
    hipError_t e = hipMalloc(p, cls);
    if (e == hipErrorOutOfMemory && ctx->cached) {
        (void)hipGetLastError();
        evict_for(ctx, ctx->budget);  // release every cached block and retry once
        e = hipMalloc(p, cls);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        set_err(ctx, err, std::string("hipMalloc: ") + hipGetErrorString(e));
        return SPCA_ENOMEM;
    }
    ctx->used += cls;
    ctx->live[*p] = cls;
    return SPCA_OK;
}

// The block must be idle: every stream that used it has been synchronised.
void raw_free(spca_ctx ctx, void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(ctx->alloc_mu);
    auto it = ctx->live.find(p);
    if (it == ctx->live.end()) return;
    ctx->cache.emplace(it->second, p);
    ctx->cached += it->second;
    ctx->live.erase(it);
}

void release_cache(spca_ctx ctx) {
    std::lock_guard<std::mutex> g(ctx->alloc_mu);
    for (auto& kv : ctx->cache) {
        (void)hipFree(kv.second);
        ctx->used -= kv.first;
    }
    ctx->cache.clear();
    ctx->cached = 0;
}

template <class T>
int dev_alloc(spca_ctx ctx, T** p, size_t count, uint64_t* acct, std::string* err = nullptr) {
    const uint64_t b = (uint64_t)count * sizeof(T);
    void* q = nullptr;
    const int rc = raw_alloc(ctx, &q, b, err);
    *p = (T*)q;
    if (rc == SPCA_OK) *acct += b;
    return rc;
}
template <class T>
void dev_free(spca_ctx ctx, T*& p, uint64_t /*bytes_of_p*/) {
    raw_free(ctx, (void*)p);
    p = nullptr;
}

bool valid_formulation(int f) { return f >= 0 && f <= 7; }

void free_matrix(spca_ctx ctx, Matrix* M) {
    (void)hipStreamSynchronize(ctx->stream);
    void* ptrs[] = {M->a32_cm, M->a32_rm, M->hi_cm, M->lo_cm, M->hi_rm, M->lo_rm, M->h_cm, M->h_rm, M->q_cm, M->qscale, M->rowmax,
                    M->cptr, M->ridx, M->cval, M->rptr, M->cidx, M->rval, M->mu, M->rmu};
    for (void* p : ptrs) raw_free(ctx, p);
    delete M;
}

// Build the operand layouts for `storage` from the FP32 column-major master.
int add_storage(spca_ctx ctx, Matrix* M, unsigned storage) {
    const size_t cnt = (size_t)M->mp * M->np;
    hipStream_t st = ctx->stream;
    unsigned want = storage & ~M->storage;
    if (!want) return SPCA_OK;
    float* rm_tmp = nullptr;
    uint64_t tmp_bytes = 0;
    int rc;
    auto need_rm = [&]() -> int {
        if (M->a32_rm) return SPCA_OK;
        if (rm_tmp) return SPCA_OK;
        if ((rc = dev_alloc(ctx, &rm_tmp, cnt, &tmp_bytes))) return rc;
        HIP_TRY(ctx, launch_transpose_f32(M->a32_cm, M->mp, M->np, M->mp, rm_tmp, M->np, st));
        return SPCA_OK;
    };
    if (want & SPCA_STORE_FP32) {
        if ((rc = dev_alloc(ctx, &M->a32_rm, cnt, &M->bytes))) return rc;
        HIP_TRY(ctx, launch_transpose_f32(M->a32_cm, M->mp, M->np, M->mp, M->a32_rm, M->np, st));
        M->storage |= SPCA_STORE_FP32;
    }
    if (want & SPCA_STORE_BF16X3) {
        if ((rc = dev_alloc(ctx, &M->hi_cm, cnt, &M->bytes)) || (rc = dev_alloc(ctx, &M->lo_cm, cnt, &M->bytes)))
            return rc;
        HIP_TRY(ctx, launch_to_blocked(M->a32_cm, M->mp, M->np, M->mp, 1, M->hi_cm, M->lo_cm, st));
        if ((rc = need_rm())) return rc;
        const float* rm = M->a32_rm ? M->a32_rm : rm_tmp;
        if ((rc = dev_alloc(ctx, &M->hi_rm, cnt, &M->bytes)) || (rc = dev_alloc(ctx, &M->lo_rm, cnt, &M->bytes)))
            return rc;
        HIP_TRY(ctx, launch_to_blocked(rm, M->np, M->mp, M->np, 1, M->hi_rm, M->lo_rm, st));
        M->storage |= SPCA_STORE_BF16X3;
    }
    if (want & SPCA_STORE_FIX8) {
        if ((rc = dev_alloc(ctx, &M->q_cm, 4 * cnt, &M->bytes)) || (rc = dev_alloc(ctx, &M->qscale, (size_t)M->np, &M->bytes)))
            return rc;
        HIP_TRY(ctx, launch_to_fix8(M->a32_cm, M->mp, M->np, M->q_cm, M->qscale, st));
        if ((rc = dev_alloc(ctx, &M->rowmax, (size_t)M->mp, &M->bytes))) return rc;
        HIP_TRY(ctx, launch_row_max(M->a32_cm, M->mp, M->np, M->rowmax, st));
        M->storage |= SPCA_STORE_FIX8;
    }
    if (want & SPCA_STORE_FP16) {
        if ((rc = need_rm())) return rc;
        const float* rm = M->a32_rm ? M->a32_rm : rm_tmp;
        if ((rc = dev_alloc(ctx, &M->h_cm, cnt, &M->bytes)) || (rc = dev_alloc(ctx, &M->h_rm, cnt, &M->bytes)))
            return rc;
        HIP_TRY(ctx, launch_to_blocked(M->a32_cm, M->mp, M->np, M->mp, 2, M->h_cm, nullptr, st));
        HIP_TRY(ctx, launch_to_blocked(rm, M->np, M->mp, M->np, 2, M->h_rm, nullptr, st));
        M->storage |= SPCA_STORE_FP16;
    }
    HIP_TRY(ctx, hipStreamSynchronize(st));
    if (rm_tmp) dev_free(ctx, rm_tmp, tmp_bytes);
    return SPCA_OK;
}

int make_matrix(spca_ctx ctx, const float* src, bool src_on_device, int m, int n, int ld, unsigned storage,
                uint64_t key, spca_matrix* out) {
    if (!ctx || !src || !out || m <= 0 || n <= 0 || ld < m) return fail(ctx, SPCA_EINVAL, "bad matrix arguments");
    if (key) {
        for (auto& kv : ctx->mats) {
            Matrix* M = kv.second;
            if (M->key == key && M->m == m && M->n == n) {
                int rc = add_storage(ctx, M, storage);
                if (rc) return rc;
                M->refs++;
                *out = kv.first;
                return SPCA_OK;
            }
        }
    }
    Matrix* M = new Matrix();
    M->m = m;
    M->n = n;
    M->mp = round_up(m, kPadMN);
    M->np = round_up(n, kPadMN);
    M->key = key;
    M->refs = 1;
    const size_t cnt = (size_t)M->mp * M->np;
    int rc = dev_alloc(ctx, &M->a32_cm, cnt, &M->bytes);
    if (rc) {
        delete M;
        return rc;
    }
    float* staging = nullptr;
    uint64_t sbytes = 0;
    const float* dsrc = src;
    if (!src_on_device) {
        rc = dev_alloc(ctx, &staging, (size_t)ld * n, &sbytes);
        if (rc) {
            free_matrix(ctx, M);
            return rc;
        }
        hipError_t e = hipMemcpyAsync(staging, src, (size_t)ld * n * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) {
            dev_free(ctx, staging, sbytes);
            free_matrix(ctx, M);
            return fail(ctx, SPCA_EDEVICE, std::string("H2D copy: ") + hipGetErrorString(e));
        }
        dsrc = staging;
    }
    hipError_t e = launch_pad_copy_f32(dsrc, m, n, ld, M->a32_cm, M->mp, M->mp, M->np, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (staging) dev_free(ctx, staging, sbytes);
    if (e != hipSuccess) {
        free_matrix(ctx, M);
        return fail(ctx, SPCA_EDEVICE, std::string("pad copy: ") + hipGetErrorString(e));
    }
    rc = add_storage(ctx, M, storage);
    if (rc) {
        free_matrix(ctx, M);
        return rc;
    }
    const uint64_t h = ctx->next_handle++;
    ctx->mats[h] = M;
    *out = h;
    return SPCA_OK;
}

// ---- synthetic spiked matrices (device noise, planted sparse loadings)
// Normal variate of counter (a, b, stream) in double: Box-Muller on Philox outputs 0 and 1.
double host_normal(uint32_t a, uint32_t b, uint32_t stream, uint64_t seed) {
    uint32_t c[4] = {a, b, stream, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u0 = ((double)c[0] + 0.5) * 2.3283064365386963e-10, u1 = ((double)c[1] + 0.5) * 2.3283064365386963e-10;
    return std::sqrt(-2.0 * std::log(u0)) * std::cos(6.283185307179586 * u1);
}

// A = N + sum_k sqrt(theta_k) g_k v_k^T: N(0,1) noise N from Philox counters (r, c, 0);
// loading k: spike_nnz distinct columns, the t-th candidate column being output 0 of
// counter (t, k, 3) mod n (repeats skipped), values v from counters (t, k, 2) normalised to
// unit 2-norm, g_k[r] from counters (r, k, 1).  loadings (n x nspikes, optional) receives v.
int make_generated(spca_ctx ctx, int m, int n, uint64_t seed, int nspikes, const double* theta, int spike_nnz,
                   unsigned storage, uint64_t key, float* loadings, spca_matrix* out) {
    if (!ctx || !out || m <= 0 || n <= 0 || nspikes < 0 || (nspikes > 0 && (!theta || spike_nnz < 1)))
        return fail(ctx, SPCA_EINVAL, "bad generator arguments");
    if (key && !loadings) {
        for (auto& kv : ctx->mats) {
            Matrix* M = kv.second;
            if (M->key == key && M->m == m && M->n == n) {
                int rc = add_storage(ctx, M, storage);
                if (rc) return rc;
                M->refs++;
                *out = kv.first;
                return SPCA_OK;
            }
        }
    }
    Matrix* M = new Matrix();
    M->m = m;
    M->n = n;
    M->mp = round_up(m, kPadMN);
    M->np = round_up(n, kPadMN);
    M->key = key;
    M->refs = 1;
    int rc = dev_alloc(ctx, &M->a32_cm, (size_t)M->mp * M->np, &M->bytes);
    if (rc) {
        delete M;
        return rc;
    }
    hipStream_t st = ctx->stream;
    const int nnz = std::min(spike_nnz, n);
    int* didx = nullptr;
    float* dbuf = nullptr;
    uint64_t tb = 0;
    auto done = [&](int code) {
        (void)hipStreamSynchronize(st);
        dev_free(ctx, didx, sizeof(int) * (uint64_t)std::max(nnz, 1));
        dev_free(ctx, dbuf, sizeof(float) * (uint64_t)(std::max(nnz, 1) + m));
        if (code != SPCA_OK) free_matrix(ctx, M);
        return code;
    };
    if (hipSuccess != launch_gen_noise(M->a32_cm, m, n, M->mp, M->np, seed, st))
        return done(fail(ctx, SPCA_EDEVICE, "noise generation failed"));
    if (loadings) std::memset(loadings, 0, sizeof(float) * (size_t)n * nspikes);
    if (nspikes > 0) {
        if ((rc = dev_alloc(ctx, &didx, (size_t)nnz, &tb)) || (rc = dev_alloc(ctx, &dbuf, (size_t)nnz + m, &tb)))
            return done(rc);
        for (int k = 0; k < nspikes; k++) {
            std::vector<int> idx;
            std::vector<char> used(n, 0);
            for (uint32_t t = 0; (int)idx.size() < nnz; t++) {
                uint32_t c[4] = {t, (uint32_t)k, 3u, 0u};
                philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
                const int col = (int)(c[0] % (uint32_t)n);
                if (!used[col]) {
                    used[col] = 1;
                    idx.push_back(col);
                }
            }
            std::vector<double> v(nnz);
            double nv = 0;
            for (int t = 0; t < nnz; t++) {
                v[t] = host_normal((uint32_t)t, (uint32_t)k, 2u, seed);
                nv += v[t] * v[t];
            }
            nv = std::sqrt(nv);
            std::vector<float> h((size_t)nnz + m);
            for (int t = 0; t < nnz; t++) {
                v[t] /= nv;
                h[t] = (float)(std::sqrt(theta[k]) * v[t]);
                if (loadings) loadings[(size_t)k * n + idx[t]] = (float)v[t];
            }
            for (int r = 0; r < m; r++) h[nnz + r] = (float)host_normal((uint32_t)r, (uint32_t)k, 1u, seed);
            // the copies are synchronous with respect to the host buffers, which change per loading
            if (hipMemcpyAsync(didx, idx.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(dbuf, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
                launch_plant(M->a32_cm, m, M->mp, didx, dbuf, nnz, dbuf + nnz, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return done(fail(ctx, SPCA_EDEVICE, "planting a loading failed"));
        }
    }
    if (hipStreamSynchronize(st) != hipSuccess) return done(fail(ctx, SPCA_EDEVICE, "generation failed"));
    if ((rc = add_storage(ctx, M, storage))) return done(rc);
    done(SPCA_OK);
    const uint64_t h = ctx->next_handle++;
    ctx->mats[h] = M;
    *out = h;
    return SPCA_OK;
}

void free_ws(spca_ctx ctx, Workspace& w, hipStream_t s) {
    (void)hipStreamSynchronize(s);
    void* ptrs[] = {w.X,      w.Xhi,   w.Xlo,    w.Zpart, w.Z,  w.Zhi, w.Zlo, w.V,    w.fstate,
                    w.istate, w.kept,  w.lcount, w.lstage, w.gl, w.ev,   w.corr,  w.uflag, w.uidx, w.ustat, w.Au, w.Xu};
    for (void* p : ptrs) raw_free(ctx, p);
    w = Workspace();
}
void free_workspace(spca_ctx ctx, Engine& E) { free_ws(ctx, E.ws, E.stream); }

// device bytes a solve of this shape holds in its workspace
uint64_t workspace_bytes(int np, int mp, int bp, int splits) {
    const uint64_t xs = (uint64_t)np * bp, zs = (uint64_t)mp * bp;
    return xs * (4 + 2 + 2 + 4) + zs * splits * 4 + zs * (4 + 2 + 2) + (uint64_t)kStateFloats * bp * 4 +
           ((uint64_t)kStateInts * bp + state_flag_ints(bp)) * 4 + (uint64_t)bp * kListMax * 4 +
           (uint64_t)bp * step_chunks(np) * (4 + 8 * kStageCap) +
           (uint64_t)bp * kGatherStride * 4 + (uint64_t)bp * 8 * 5;
}

int ensure_ws(spca_ctx ctx, Workspace& w, hipStream_t s, int np, int mp, int bp, int splits,
              std::string* err = nullptr) {
    if (w.X && w.np == np && w.mp == mp && w.bp == bp && w.splits >= splits) return SPCA_OK;
    free_ws(ctx, w, s);
    w.np = np;
    w.mp = mp;
    w.bp = bp;
    w.splits = splits;
    const size_t xs = (size_t)np * bp, zs = (size_t)mp * bp;
    int rc;
    if ((rc = dev_alloc(ctx, &w.X, xs, &w.bytes, err)) || (rc = dev_alloc(ctx, &w.Xhi, xs, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.Xlo, xs, &w.bytes, err)) || (rc = dev_alloc(ctx, &w.Zpart, zs * splits, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.Z, zs, &w.bytes, err)) || (rc = dev_alloc(ctx, &w.Zhi, zs, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.Zlo, zs, &w.bytes, err)) || (rc = dev_alloc(ctx, &w.V, xs, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.fstate, (size_t)kStateFloats * bp, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.istate, (size_t)kStateInts * bp + state_flag_ints(bp), &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.kept, (size_t)kListMax * bp, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.lcount, (size_t)step_chunks(np) * bp, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.lstage, (size_t)step_chunks(np) * kStageCap * bp, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.gl, (size_t)kGatherStride * bp, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.ev, (size_t)bp, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.corr, (size_t)4 * bp, &w.bytes, err))) {
        free_ws(ctx, w, s);
        return rc;
    }
    if (hipMemsetAsync(w.Z, 0, zs * sizeof(float), s) != hipSuccess ||
        hipMemsetAsync(w.Zhi, 0, zs * sizeof(uint16_t), s) != hipSuccess ||
        hipMemsetAsync(w.Zlo, 0, zs * sizeof(uint16_t), s) != hipSuccess) {
        free_ws(ctx, w, s);
        return SPCA_EDEVICE;
    }
    return SPCA_OK;
}
// The union A*X buffers of a workspace, for union sizes up to ldu (a multiple of 32 * splits).
int ensure_union(spca_ctx ctx, Workspace& w, hipStream_t s, int ldu) {
    if (w.ldu >= ldu) return SPCA_OK;
    if (w.ldu > 0 && hipStreamSynchronize(s) != hipSuccess) return SPCA_EDEVICE;  // old buffers idle
    int* ptrs_i[] = {w.uflag, w.uidx, w.ustat};
    for (int* p : ptrs_i) raw_free(ctx, p);
    raw_free(ctx, w.Au);
    raw_free(ctx, w.Xu);
    w.uflag = w.uidx = w.ustat = nullptr;
    w.Au = w.Xu = nullptr;
    w.ldu = 0;
    int rc;
    if ((rc = dev_alloc(ctx, &w.uflag, (size_t)w.np, &w.bytes)) || (rc = dev_alloc(ctx, &w.uidx, (size_t)w.np, &w.bytes)) ||
        (rc = dev_alloc(ctx, &w.ustat, (size_t)4, &w.bytes)) ||
        (rc = dev_alloc(ctx, &w.Au, (size_t)w.mp * ldu, &w.bytes)) ||
        (rc = dev_alloc(ctx, &w.Xu, (size_t)w.bp * ldu, &w.bytes)))
        return rc;
    w.ldu = ldu;
    return SPCA_OK;
}

int ensure_workspace(spca_ctx ctx, Engine& E, int np, int mp, int bp, int splits, std::string* err = nullptr) {
    return ensure_ws(ctx, E.ws, E.stream, np, mp, bp, splits, err);
}

// The allocation-free solve's memory: consecutive 256-byte aligned blocks of the fixed device
// array.  With base == nullptr, take() only counts, so the same code sizes a request.
struct FixedScratch {
    unsigned char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    template <class T>
    T* take(size_t count) {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + count * sizeof(T);
        return base && used <= cap ? (T*)(base + off) : nullptr;
    }
};

// The workspace of a solve of this shape in fixed scratch: the buffers ensure_ws allocates, minus
// the 16-bit operand forms the precision and A^T*Z form do not read.  Z and its operand forms
// are cleared (padded rows must read as zero in A^T*Z); load_batch clears the rest.  Returns false
// when the scratch is too small (sc.used then holds the bytes the shape needs).
bool carve_ws(Workspace& w, FixedScratch& sc, int np, int mp, int bp, int splits, int precision, int zprec,
              hipStream_t s, hipError_t* e) {
    w = Workspace();
    w.np = np;
    w.mp = mp;
    w.bp = bp;
    w.splits = splits;
    const size_t xs = (size_t)np * bp, zs = (size_t)mp * bp;
    w.X = sc.take<float>(xs);
    if (precision) w.Xhi = sc.take<uint16_t>(xs);
    if (precision == 1) w.Xlo = sc.take<uint16_t>(xs);
    w.Zpart = sc.take<float>(zs * splits);
    w.Z = sc.take<float>(zs);
    if (zprec) w.Zhi = sc.take<uint16_t>(zs);
    if (zprec == 1) w.Zlo = sc.take<uint16_t>(zs);
    w.V = sc.take<float>(xs);
    w.fstate = sc.take<float>((size_t)kStateFloats * bp);
    w.istate = sc.take<int>((size_t)kStateInts * bp + state_flag_ints(bp));
    w.kept = sc.take<int>((size_t)kListMax * bp);
    w.lcount = sc.take<int>((size_t)step_chunks(np) * bp);
    w.lstage = sc.take<unsigned long long>((size_t)step_chunks(np) * kStageCap * bp);
    w.ev = sc.take<double>((size_t)bp);
    w.corr = sc.take<double>((size_t)4 * bp);
    *e = hipSuccess;
    if (!sc.base || sc.used > sc.cap) return false;
    *e = launch_fill_zero(w.Z, zs * sizeof(float), s);
    if (*e == hipSuccess && w.Zhi) *e = launch_fill_zero(w.Zhi, zs * sizeof(uint16_t), s);
    if (*e == hipSuccess && w.Zlo) *e = launch_fill_zero(w.Zlo, zs * sizeof(uint16_t), s);
    return true;
}

int engine_init(Engine& E) {
    if (hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking) != hipSuccess) return SPCA_EDEVICE;
    if (hipHostMalloc((void**)&E.h_done, sizeof(int) * 8, hipHostMallocDefault) != hipSuccess) return SPCA_ENOMEM;
    if (hipEventCreate(&E.loop_ev[0]) != hipSuccess || hipEventCreate(&E.loop_ev[1]) != hipSuccess)
        return SPCA_EDEVICE;
    return SPCA_OK;
}

void engine_destroy(spca_ctx ctx, Engine& E) {
    if (E.stream) (void)hipStreamSynchronize(E.stream);
    free_workspace(ctx, E);
    for (hipEvent_t e : E.events) (void)hipEventDestroy(e);
    E.events.clear();
    for (hipEvent_t& e : E.loop_ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    if (E.h_done) (void)hipHostFree(E.h_done);
    E.h_done = nullptr;
    if (E.stream) (void)hipStreamDestroy(E.stream);
    E.stream = nullptr;
}

SparseMat sparse_of(const Matrix* M) {
    return SparseMat{M->cptr, M->ridx, M->cval, M->rptr, M->cidx, M->rval, M->mu, M->rmu, M->m, M->n};
}

SpState state_of(Workspace& w) {
    SpState s;
    s.val = w.fstate;
    s.fz = w.fstate + w.bp;
    s.err = w.fstate + 2 * w.bp;
    s.errmax = w.fstate + 3 * w.bp;
    s.thr = w.fstate + 4 * w.bp;
    s.iters = w.istate;
    s.nnz = w.istate + w.bp;
    s.active = w.istate + 2 * w.bp;
    s.nkept = w.istate + 3 * w.bp;
    s.lmax = (unsigned*)(w.istate + 4 * w.bp);
    s.lcount = w.lcount;
    s.done = w.istate + kStateInts * w.bp;
    s.kept = w.kept;
    s.lstage = w.lstage;
    return s;
}

// reference getSignleStartingPoint<F> (c++/src/gpower/gpower_commons.h:21-47) with
// myseed = the global SP index k, evaluated in F (the index floor(n * r / RAND_MAX)
// depends on F's rounding, so the float and double instantiations draw different SPs),
// then stored as float.
template <class F>
void reference_starting_point(float* out, unsigned n, unsigned s, bool constrained, unsigned k) {
    std::vector<F> V(n, (F)0);
    unsigned seed = k;
    if (constrained) {
        F tmp_norm = 0;
        for (unsigned i = 0; i < s; i++) {
            unsigned idx = (unsigned)(int)((F)n * (F)rand_r(&seed) / (F)RAND_MAX);
            if (idx == n) idx--;
            F tmp = (F)rand_r(&seed) / (F)RAND_MAX;
            V[idx] = tmp;
            tmp_norm += tmp * tmp;
        }
        const F sc = (F)1 / std::sqrt(tmp_norm);
        for (unsigned i = 0; i < n; i++) V[i] = sc * V[i];
    } else {
        for (unsigned i = 0; i < n; i++) {
            F tmp = (F)rand_r(&seed) / (F)RAND_MAX;
            V[i] = -1 + 2 * tmp;
        }
    }
    for (unsigned i = 0; i < n; i++) out[i] = (float)V[i];
}

// Largest nonzero count per iterate at which A*X is gathered over the nonzeros.  The gather
// reads 4 B per nonzero and row (14-17 TB/s at 4096 x 65536, B = 1024, where SPs share
// columns); the dense product costs 2n flops per row and SP at the form's MFMA rate.  At that
// shape the FP32 product takes 4.0 ms and the gather 2.6-3.0 ms at 2400-3100 nonzeros per SP,
// so FP32 gathers up to n/16.  The 16-bit products take 1.07 ms (bf16x3) and 0.47 ms (fp16)
// against 0.27-0.51 ms gathered at 256 nonzeros per SP: n/128.
int sparse_cap(const spca_settings& st, int n) {
    if (!st.sparse_products) return 0;
    const int div = st.precision == SPCA_PREC_FP32 ? 16 : 128;
    return std::max(1, n / div);
}

// Largest union of supports |U| at which A*X is the union product (sparse_products = 2): its
// FP32 product costs ~|U|/n of the dense FP32 one, so up to n/4 in FP32 solves, and n/12 and
// n/24 against the ~3x and ~6x faster dense bf16x3 and fp16 products.
int union_cap(const spca_settings& st, int n) {
    if (st.sparse_products != 2) return 0;
    const int div = st.precision == SPCA_PREC_FP32 ? 4 : st.precision == SPCA_PREC_BF16X3 ? 12 : 24;
    return std::max(1, n / div);
}

// K-splits of A*X: enough workgroups to fill 256 CUs for a couple of rounds (128x128
// tiles run ~3 per CU, 256x256 tiles one per CU), each split at least 512 deep.
int choose_splits(int precision, int mp, int np, int bp, bool allow_big = true) {
    const bool big = allow_big && gemm_big_tiles(precision, mp, bp);
    const int tiles = big ? (mp / 256) * (bp / 256) : (mp / kGemmBM) * (bp / kGemmBN);
    const int target = big ? 512 : 2048;
    int splits = 1;
    while (tiles * splits < target && np % (kGemmBK * splits * 2) == 0 && np / (splits * 2) >= 512) splits *= 2;
    return splits;
}

Operand x_operand(Workspace& w, int precision) {
    return Operand{w.X, precision ? w.Xhi : nullptr, precision == 1 ? w.Xlo : nullptr, w.bp};
}
// precision 4: the int8 sign vectors of the fixed-point product, in Zhi alone
Operand z_operand(Workspace& w, int precision) {
    return Operand{w.Z, precision ? w.Zhi : nullptr, precision == 1 ? w.Zlo : nullptr, w.bp};
}

// The second product, A^T*Z, of a solve: the precision's own operands, or, for the
// L1-variance formulations with exact_sign_products (FP32 and bf16x3), the fixed-point pieces
// of A against Z = sgn(A x) in int8 on the i8 MFMA: exact integer sums, one rounding per v_j.
// Its int32 accumulators hold |sum_i (64 q0 + q1) z_i| <= 8160 mp, so it needs mp < 2^18
// padded rows; taller matrices run the precision's own A^T*Z product.
constexpr int kFix8MaxRows = 1 << 18;
bool uses_sign_products(const spca_settings& st, int mp) {
    return st.exact_sign_products && spca_is_l1_variance(st.formulation) && (st.precision == 0 || st.precision == 1) &&
           mp < kFix8MaxRows;
}
// The step kernel sums |v|/max|v| in 2^-40 fixed point in u64 LDS counters: n < 2^23 columns.
constexpr int kMaxCols = 1 << 23;
struct AtzForm {
    int prec;            // launch_gemm_tn precision, or 4: the fixed-point product (launch_gemm_fix8)
    int zprec;           // form prep_z writes Z in
    const float* p32;
    const uint16_t* hi;
    const uint16_t* lo;
    const int8_t* q;     // fixed-point pieces (prec 4)
    const float* qscale;
    long ldp, ldq;       // FP32 row strides or 16-bit row counts
    int big;
};
AtzForm atz_form(const Matrix* M, const spca_settings& st, int bp, int plan_big) {
    AtzForm a;
    a.p32 = M->a32_cm;
    a.q = M->q_cm;
    a.qscale = M->qscale;
    if (uses_sign_products(st, M->mp)) {
        a.prec = a.zprec = 4;
        a.hi = a.lo = nullptr;
        a.big = 0;
    } else {
        a.prec = a.zprec = st.precision;
        a.hi = st.precision == 1 ? M->hi_cm : M->h_cm;
        a.lo = M->lo_cm;
        a.big = plan_big;
    }
    a.ldp = a.prec ? M->np : M->mp;
    a.ldq = a.prec ? bp : M->mp;
    return a;
}
// Exact signs of z after a dense A*X (L1 variance with exact_sign_products): entries within
// eps * rowmax_i * |x|_2 of zero are recomputed in FP64.  The bound is about a dozen standard
// deviations of the product's rounding error: FP32 MFMA accumulation over K/splits terms
// (~2^-18.5 |z|), bf16x3 products (~2^-17 per term); with rowmax_i >= the rms of the row's
// terms it flags ~1e-4 (FP32) to ~2e-4 (bf16x3) of the entries of Gaussian data.
SignRefine sign_refine(const Matrix* M, const spca_settings& st, Workspace& w) {
    SignRefine r{M->a32_rm, M->rowmax, w.X, M->np, 0.f};
    if (uses_sign_products(st, M->mp) && M->a32_rm && M->rowmax) r.eps = st.precision == 0 ? 0x1p-17f : 0x1p-14f;
    return r;
}

// Exact ties for the L0 constraint with L2 variance (FP32, bf16x3): the band half-width is
// 2 * eps * tau, eps a few standard deviations of the A^T*Z product's relative rounding at
// the top of |v| (FP32 accumulation over m terms ~1e-6; bf16x3 products ~2e-6 more).
void set_exact_ties(StepArgs& sa, const Matrix* M, const spca_settings& st, Workspace& w) {
    sa.acm = M->a32_cm;
    sa.zf = w.Z;
    sa.mp = M->mp;
    sa.band_eps = 0.f;
    if (!M->a32_cm) return;  // sparse matrices: the products already sum in FP64
    const bool l2_thresholded = st.formulation == L0_constrained_L2_PCA || st.formulation == L0_penalized_L2_PCA ||
                                st.formulation == L1_penalized_L2_PCA;
    if (st.exact_ties && l2_thresholded && (st.precision == 0 || st.precision == 1))
        sa.band_eps = st.precision == 0 ? 0x1p-17f : 0x1p-15f;
}

// V = A^T * Z of one iteration in the solve's A^T*Z form (sparse matrices: the CSC gather with
// the centring terms).
hipError_t launch_atz(const AtzForm& a, const Matrix* M, Workspace& w, int bp, int batch, SpState sp, hipStream_t s) {
    if (M->sparse) {
        hipError_t e = launch_sp_scalars(w.Z, M->mp, M->m, M->rmu, w.corr + 2 * bp, bp, batch, sp, s);
        if (e != hipSuccess) return e;
        return launch_csc_atz(sparse_of(M), w.Z, M->mp, w.V, M->np, M->np, w.corr + 2 * bp, bp, batch, sp, s);
    }
    if (a.prec == 4)
        return launch_gemm_fix8(a.q, a.qscale, M->np, (const int8_t*)w.Zhi, bp, w.V, M->np, M->np, bp, M->mp, sp.done,
                                s);
    if (a.prec == 0 && !a.big && batch <= 64)  // narrow tile: no padding of the batch to 128 columns
        return launch_gemm_f32_narrow(a.p32, a.ldp, w.Z, a.ldq, w.V, M->np, 0, M->np, batch, M->mp, 1, sp.done, nullptr,
                                      s);
    return launch_gemm_tn(a.prec, a.p32, a.hi, a.lo, a.ldp, w.Z, w.Zhi, w.Zlo, a.ldq, w.V, M->np, 0, M->np, bp, M->mp, 1,
                          a.big, sp.done, nullptr, s);
}

// The gathered A*X into partial 0 of Zpart for the groups flagged on the device: over nonzero
// lists built once per SP when the workspace holds them (the allocation-free solve's does not),
// else scanning x per row block.  Both sum in ascending index order: the same bits.
hipError_t launch_gather(const Matrix* M, Workspace& w, int batch, SpState sp, hipStream_t s) {
    if (w.gl)
        return launch_sparse_ax_listed(M->a32_cm, M->mp, M->mp, M->n, w.X, w.np, w.Zpart, M->mp, batch, sp, w.gl, s);
    return launch_sparse_ax(M->a32_cm, M->mp, M->mp, M->n, w.X, w.np, w.Zpart, M->mp, batch, sp, s);
}

// Zpart = A * X of one iteration: the dense product (K split over `splits`) and, with
// `gather`, the sparse gather beside it (one of them runs, on the device flag); sparse
// matrices: the CSR gather with the centring terms, into partial 0.
hipError_t launch_ax(const Matrix* M, const spca_settings& st, Workspace& w, int splits, int big, int bp, int batch,
                     long zsplit, SpState sp, bool gather, hipStream_t s) {
    if (M->sparse) {
        hipError_t e = launch_sp_scalars(w.X, M->np, M->n, M->mu, w.corr, bp, batch, sp, s);
        if (e != hipSuccess) return e;
        return launch_csr_ax(sparse_of(M), w.X, M->np, w.Zpart, M->mp, M->mp, w.corr, bp, batch, sp, s);
    }
    const float* P1_32 = M->a32_rm;
    const uint16_t* P1_hi = st.precision == 1 ? M->hi_rm : M->h_rm;
    const long ld1p = st.precision ? M->mp : M->np, ld1q = st.precision ? bp : M->np;
    hipError_t e;
    if (st.precision == 0 && !big && batch <= 64)  // narrow tile, same splits and bits
        e = launch_gemm_f32_narrow(P1_32, ld1p, w.X, ld1q, w.Zpart, M->mp, zsplit, M->mp, batch, M->np, splits, sp.done,
                                   gather ? sp.done + kSparseFlags : nullptr, s);
    else
        e = launch_gemm_tn(st.precision, P1_32, P1_hi, M->lo_rm, ld1p, w.X, w.Xhi, w.Xlo, ld1q, w.Zpart, M->mp, zsplit,
                           M->mp, bp, M->np, splits, big, sp.done, gather ? sp.done + kSparseFlags : nullptr, s);
    if (e == hipSuccess && gather) e = launch_gather(M, w, batch, sp, s);
    return e;
}

// Starting points with global indices shift .. shift+cnt-1 (n floats each) into out.
void make_starts(const spca_settings& st, int n, const float* X0, int shift, int cnt, float* out) {
    if (st.start == SPCA_START_GAUSSIAN) return;  // generated on the device by load_batch
    const bool constrained = st.formulation < 4;
    for (int jj = 0; jj < cnt; jj++) {
        float* col = out + (size_t)jj * n;
        if (st.start == SPCA_START_EXPLICIT)
            std::memcpy(col, X0 + (size_t)(shift + jj) * n, sizeof(float) * n);
        else if (st.start == SPCA_START_REFERENCE_F64)
            reference_starting_point<double>(col, (unsigned)n, st.constraint, constrained, (unsigned)(shift + jj));
        else
            reference_starting_point<float>(col, (unsigned)n, st.constraint, constrained, (unsigned)(shift + jj));
    }
}

// Start of a batch: X <- the starting points in hX (padded rows and columns zero), its
// operand forms, and the per-SP state reset (val = err = 0, iters = nnz = 0,
// active = (j < batch), done = 0).  Returns after the copies have completed.
int load_batch(spca_ctx ctx, hipStream_t s, Workspace& w, const spca_settings& st, int n, int batch,
               const float* hX, int shift = 0) {
    const Operand xop = x_operand(w, st.precision);
    const SpState sp = state_of(w);
    // cleared by a kernel: the fixed-scratch solve's buffers keep the previous call's contents
    HIP_TRY(ctx, launch_fill_zero(w.X, sizeof(float) * (size_t)w.np * w.bp, s));
    if (st.start == SPCA_START_GAUSSIAN)
        HIP_TRY(ctx, launch_gen_starts(w.X, w.np, n, batch, shift, st.seed, s));
    else
        HIP_TRY(ctx, hipMemcpy2DAsync(w.X, sizeof(float) * w.np, hX, sizeof(float) * n, sizeof(float) * n, batch,
                                      hipMemcpyHostToDevice, s));
    if (st.precision) {
        HIP_TRY(ctx, launch_fill_zero(w.Xhi, sizeof(uint16_t) * (size_t)w.np * w.bp, s));
        if (st.precision == 1) HIP_TRY(ctx, launch_fill_zero(w.Xlo, sizeof(uint16_t) * (size_t)w.np * w.bp, s));
        HIP_TRY(ctx, launch_x_operand(xop, st.precision, w.np, batch, s));
    }
    HIP_TRY(ctx, launch_fill_zero(w.fstate, sizeof(float) * kStateFloats * (size_t)w.bp, s));
    HIP_TRY(ctx, launch_fill_zero(w.istate, sizeof(int) * (kStateInts * (size_t)w.bp + state_flag_ints(w.bp)), s));
    std::vector<int> act(w.bp, 0);
    std::fill(act.begin(), act.begin() + batch, 1);
    HIP_TRY(ctx, hipMemcpyAsync(sp.active, act.data(), sizeof(int) * w.bp, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return SPCA_OK;
}

// FP64 explained variance of `batch` columns of X (device) into ev (device); the partial
// sums live in a scratch block that returns to the cache once the kernels are done.
int run_ev(spca_ctx ctx, hipStream_t s, Matrix* M, const float* X, int ldx, int batch, int formulation, double* ev,
           FixedScratch* fx = nullptr) {
    const int l1 = spca_is_l1_variance(formulation);
    if (fx) {
        // the fixed array past the workspace, in groups of SPs whose partial sums fit (dense only)
        hipError_t e = hipSuccess;
        for (int j0 = 0; j0 < batch && e == hipSuccess;) {
            int g = batch - j0;
            while (g > 1 && ev_scratch_doubles(M->m, M->n, g) * sizeof(double) + 256 > fx->cap - fx->used) g = (g + 1) / 2;
            FixedScratch rest = *fx;
            double* work = rest.take<double>(ev_scratch_doubles(M->m, M->n, g));
            if (!work) return fail(ctx, SPCA_EINVAL, "fixed scratch too small for the explained variance of one SP");
            e = launch_ev_fp64(M->a32_cm, M->mp, M->m, M->n, X + (size_t)j0 * ldx, ldx, g, l1, ev + j0, work, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);  // the next group reuses the same memory
            j0 += g;
        }
        if (e != hipSuccess) return fail(ctx, SPCA_EDEVICE, std::string("explained variance: ") + hipGetErrorString(e));
        return SPCA_OK;
    }
    double* work = nullptr;
    uint64_t sb = 0;
    int rc = dev_alloc(ctx, &work, M->sparse ? csr_ev_scratch_doubles(M->m, batch) : ev_scratch_doubles(M->m, M->n, batch),
                       &sb);
    if (rc) return rc;
    hipError_t e = M->sparse ? launch_csr_ev(sparse_of(M), X, ldx, batch, l1, ev, work, s)
                             : launch_ev_fp64(M->a32_cm, M->mp, M->m, M->n, X, ldx, batch, l1, ev, work, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dev_free(ctx, work, sb);
    if (e != hipSuccess) return fail(ctx, SPCA_EDEVICE, std::string("explained variance: ") + hipGetErrorString(e));
    return SPCA_OK;
}

// What a solve accumulates over its batches.
struct Results {
    double best = -1;
    int best_sp = -1;
    unsigned total_it = 0;
    unsigned sparse_it = 0;
    unsigned union_it = 0;
    unsigned list_tries = 0, list_steps = 0;
    unsigned long long sp_its = 0;
    double last_err = 0;
    double ev_best = 0;
    std::vector<float> x_best;
};

// End of a batch: per-SP values, iterations, nonzeros and FP64 explained variance, the
// final iterates if asked, and the best of the batch (first maximum, as
// sparse_PCA_solver.h:212-231) against the best so far.
int collect_batch(spca_ctx ctx, hipStream_t s, Matrix* M, Workspace& w, const spca_settings& st, int shift, int batch,
                  bool want_ev, spca_sp_result* per_sp, float* X_final, Results& R, FixedScratch* fx = nullptr) {
    const SpState sp = state_of(w);
    const int n = M->n;
    std::vector<float> hval(batch), herr(batch);
    std::vector<int> hit(batch), hnnz(batch);
    std::vector<double> hev(batch, 0.0);
    int done_h[5];
    float emax;
    HIP_TRY(ctx, hipMemcpyAsync(done_h, sp.done, sizeof(int) * 5, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(hval.data(), sp.val, sizeof(float) * batch, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(herr.data(), sp.err, sizeof(float) * batch, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(hit.data(), sp.iters, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(hnnz.data(), sp.nnz, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(&emax, sp.errmax, sizeof(float), hipMemcpyDeviceToHost, s));
    // per-SP explained variance when per-SP results are asked for; else only the batch's best
    // SP needs it (for stats), which is computed below once it is known
    if (want_ev && per_sp) {
        int rc = run_ev(ctx, s, M, w.X, w.np, batch, st.formulation, w.ev, fx);
        if (rc) return rc;
        HIP_TRY(ctx, hipMemcpyAsync(hev.data(), w.ev, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    R.total_it += (unsigned)done_h[1];
    R.sparse_it += (unsigned)done_h[4];
    R.list_tries += (unsigned)done_h[2];
    R.list_steps += (unsigned)done_h[3];
    R.last_err = emax;
    for (int jj = 0; jj < batch; jj++) R.sp_its += (unsigned long long)hit[jj];
    int sel = 0;
    float bv = hval[0];
    for (int jj = 1; jj < batch; jj++)
        if (hval[jj] > bv) {
            bv = hval[jj];
            sel = jj;
        }
    if (per_sp) {
        for (int jj = 0; jj < batch; jj++) {
            spca_sp_result& r = per_sp[shift + jj];
            r.value = hval[jj];
            r.iters = hit[jj];
            r.nnz = hnnz[jj];
            r.converged = (double)herr[jj] < st.tolerance;
            r.ev_fp64 = hev[jj];
        }
    }
    if (X_final)
        HIP_TRY(ctx, hipMemcpy2DAsync(X_final + (size_t)shift * n, sizeof(float) * n, w.X, sizeof(float) * w.np,
                                      sizeof(float) * n, batch, hipMemcpyDeviceToHost, s));
    if (R.best < (double)bv || R.best_sp < 0) {
        R.best = bv;
        R.best_sp = shift + sel;
        R.x_best.assign(n, 0.f);
        HIP_TRY(ctx, hipMemcpyAsync(R.x_best.data(), w.X + (size_t)sel * w.np, sizeof(float) * n,
                                    hipMemcpyDeviceToHost, s));
        if (want_ev && !per_sp) {
            int rc = run_ev(ctx, s, M, w.X + (size_t)sel * w.np, w.np, 1, st.formulation, w.ev, fx);
            if (rc) return rc;
            HIP_TRY(ctx, hipMemcpyAsync(&hev[sel], w.ev, sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_TRY(ctx, hipStreamSynchronize(s));
        }
        R.ev_best = hev[sel];
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return SPCA_OK;
}

// final x <- x/|x|_2 (sparse_PCA_solver.h:236-237) and the statistics
void finish_solve(const Results& R, int n, const spca_settings& st, float* x_out, spca_stats* stats, double loop_ms,
                  const double* phase_ms, std::chrono::steady_clock::time_point t_wall0) {
    double nx = 0;
    for (int i = 0; i < n; i++) nx += (double)R.x_best[i] * (double)R.x_best[i];
    const float inv = 1.0f / (float)std::sqrt(nx);
    for (int i = 0; i < n; i++) x_out[i] = R.x_best[i] * inv;
    if (stats) {
        stats->it = R.total_it;
        stats->fval = R.best;
        stats->error = R.last_err;
        stats->loop_ms = loop_ms;
        stats->wall_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall0).count();
        stats->n = (unsigned)n;
        stats->best_sp = R.best_sp;
        stats->total_sp = st.total_sp;
        stats->constraint = st.constraint;
        stats->gemm_ax_ms = phase_ms ? phase_ms[0] : 0.0;
        stats->gemm_atz_ms = phase_ms ? phase_ms[1] : 0.0;
        stats->step_ms = phase_ms ? phase_ms[2] + phase_ms[3] : 0.0;
        stats->step_kernel_ms = phase_ms ? phase_ms[3] : 0.0;
        stats->sp_iterations = R.sp_its;
        stats->ev_fp64 = R.ev_best;
        stats->sparse_iterations = R.sparse_it;
        stats->union_iterations = R.union_it;
        stats->list_tries = R.list_tries;
        stats->list_steps = R.list_steps;
    }
}

// ---- product plans and the autotuner
Plan heuristic_plan(int precision, int mp, int np, int bp, bool allow_big) {
    Plan p;
    p.big = allow_big && gemm_big_tiles(precision, mp, bp);
    p.splits = choose_splits(precision, mp, np, bp, p.big != 0);
    return p;
}

// Plan cache file: {"plans": [{"precision": P, "mp": .., "np": .., "bp": .., "allow_big": ..,
// "splits": .., "big": .., "ms": ..}, ...]}
void load_plans(spca_ctx ctx) {
    FILE* f = std::fopen(ctx->tune_path.c_str(), "r");
    if (!f) return;
    std::string txt;
    char buf[4096];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, r);
    std::fclose(f);
    size_t pos = 0;
    while ((pos = txt.find('{', pos + 1)) != std::string::npos) {
        const size_t end = txt.find('}', pos);
        if (end == std::string::npos) break;
        const std::string obj = txt.substr(pos, end - pos + 1);
        auto num = [&](const char* key, double* v) {
            const size_t k = obj.find(std::string("\"") + key + "\"");
            if (k == std::string::npos) return false;
            const size_t c = obj.find(':', k);
            return c != std::string::npos && std::sscanf(obj.c_str() + c + 1, "%lf", v) == 1;
        };
        double P, mp, np, bp, ab, sp, bg, ms = 0;
        if (num("precision", &P) && num("mp", &mp) && num("np", &np) && num("bp", &bp) && num("allow_big", &ab) &&
            num("splits", &sp) && num("big", &bg)) {
            num("ms", &ms);
            Plan p;
            p.splits = (int)sp;
            p.big = (int)bg;
            p.ms = ms;
            ctx->plans[PlanKey{(int)P, (int)mp, (int)np, (int)bp, (int)ab}] = p;
        }
        pos = end;
    }
}

void save_plans(spca_ctx ctx) {
    if (ctx->tune_path.empty()) return;
    const std::string tmp = ctx->tune_path + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return;
    std::fprintf(f, "{\"plans\": [");
    bool first = true;
    for (auto& kv : ctx->plans) {
        if (kv.second.ms <= 0) continue;  // only measured plans are worth keeping
        std::fprintf(f, "%s\n  {\"precision\": %d, \"mp\": %d, \"np\": %d, \"bp\": %d, \"allow_big\": %d, "
                        "\"splits\": %d, \"big\": %d, \"ms\": %.5f}",
                     first ? "" : ",", kv.first.precision, kv.first.mp, kv.first.np, kv.first.bp,
                     kv.first.allow_big, kv.second.splits, kv.second.big, kv.second.ms);
        first = false;
    }
    std::fprintf(f, "\n]}\n");
    std::fclose(f);
    std::rename(tmp.c_str(), ctx->tune_path.c_str());
}

// Times every valid plan of this shape on the job's own matrix and starting points (already
// in w.X): A*X, the z epilogue and A^T*Z, one warm-up and three timed runs each, and keeps the
// fastest.  Candidates: K splits 1..16 of A*X (each split at least 256 deep, partial sums
// within 512 MiB) and, where the shape allows, both tile kernels.
int tune_plan(spca_ctx ctx, Engine& E, Matrix* M, const spca_settings& st, Workspace& w, bool allow_big,
              Plan* out) {
    hipStream_t s = E.stream;
    const int mp = M->mp, np = M->np, bp = w.bp, prec = st.precision;
    const SpState sp = state_of(w);
    const Operand zop = z_operand(w, prec);
    const long zsplit = (long)mp * bp;
    const long ld1p = prec ? mp : np, ld1q = prec ? bp : np, ld2p = prec ? np : mp, ld2q = prec ? bp : mp;
    const float* P1_32 = M->a32_rm;
    const uint16_t* P1_hi = prec == 1 ? M->hi_rm : M->h_rm;
    const float* P2_32 = M->a32_cm;
    const uint16_t* P2_hi = prec == 1 ? M->hi_cm : M->h_cm;
    hipEvent_t e0, e1;
    HIP_TRY(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(ctx, SPCA_EDEVICE, "event create failed");
    }
    Plan best;
    best.ms = 1e30;
    int rc = SPCA_OK;
    const int bigs = allow_big && gemm_big_tiles(prec, mp, bp) ? 2 : 1;
    for (int bi = 0; bi < bigs && rc == SPCA_OK; bi++)
        for (int splits = 1; splits <= std::min(16, w.splits) && rc == SPCA_OK; splits *= 2) {
            if (np % (kGemmBK * splits) || np / splits < 256) break;
            float tot = 0;
            for (int r = 0; r < 4; r++) {
                hipError_t e = hipEventRecord(e0, s);
                if (e == hipSuccess)
                    e = launch_gemm_tn(prec, P1_32, P1_hi, M->lo_rm, ld1p, w.X, w.Xhi, w.Xlo, ld1q, w.Zpart,
                                       mp, zsplit, mp, bp, np, splits, bi, nullptr, nullptr, s);
                if (e == hipSuccess)
                    e = launch_prep_z(w.Zpart, splits, zsplit, mp, mp, zop, st.formulation, prec, st.batch, sp,
                                      SignRefine{}, s);
                if (e == hipSuccess)
                    e = launch_gemm_tn(prec, P2_32, P2_hi, M->lo_cm, ld2p, w.Z, w.Zhi, w.Zlo, ld2q, w.V, np,
                                       0, np, bp, mp, 1, bi, nullptr, nullptr, s);
                if (e == hipSuccess) e = hipEventRecord(e1, s);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float t = 0;
                if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
                if (e != hipSuccess) {
                    rc = fail(ctx, SPCA_EDEVICE, std::string("tuning: ") + hipGetErrorString(e));
                    break;
                }
                if (r > 0) tot += t;
            }
            const double ms = tot / 3.0;
            if (rc == SPCA_OK && ms < best.ms) {
                best.ms = ms;
                best.splits = splits;
                best.big = bi;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == SPCA_OK) *out = best;
    return rc;
}

// The plan of a solve: cached (tuned or loaded), else tuned now when tuning is on (with the
// first batch's starting points), else the heuristic.  Tuning leaves the workspace sized for
// the largest candidate split.
PlanKey plan_key(int precision, int mp, int np, int bp, bool allow_big) {
    // the tile choice only exists for 16-bit shapes the large tiles fit, so FP32 solves and
    // such shapes share one plan between solo and fused solves
    return PlanKey{precision, mp, np, bp, allow_big && gemm_big_tiles(precision, mp, bp) ? 1 : 0};
}

// The cached plan of a shape, else the heuristic (never tunes).
Plan peek_plan(spca_ctx ctx, int precision, int mp, int np, int bp, bool allow_big) {
    std::lock_guard<std::mutex> g(ctx->plan_mu);
    auto it = ctx->plans.find(plan_key(precision, mp, np, bp, allow_big));
    return it != ctx->plans.end() ? it->second : heuristic_plan(precision, mp, np, bp, allow_big);
}

int plan_for(spca_ctx ctx, Engine& E, Matrix* M, const spca_settings& st, int bp, bool allow_big, const float* X0,
             Plan* out, Workspace* wsp = nullptr) {
    if (M->sparse) {  // sparse products: no K split, no tiles to choose
        *out = Plan();
        return SPCA_OK;
    }
    const PlanKey key = plan_key(st.precision, M->mp, M->np, bp, allow_big);
    allow_big = key.allow_big != 0;
    {
        std::lock_guard<std::mutex> g(ctx->plan_mu);
        auto it = ctx->plans.find(key);
        if (it != ctx->plans.end()) {
            *out = it->second;
            return SPCA_OK;
        }
        if (!ctx->tuning) {
            *out = heuristic_plan(st.precision, M->mp, M->np, bp, allow_big);
            return SPCA_OK;
        }
    }
    Workspace& w = wsp ? *wsp : E.ws;
    int max_splits = 16;
    while (max_splits > 1 && (uint64_t)M->mp * bp * 4 * max_splits > (512ull << 20)) max_splits /= 2;
    int rc = ensure_ws(ctx, w, E.stream, M->np, M->mp, bp, max_splits);
    if (rc) return rc;
    std::vector<float> hX(st.start == SPCA_START_GAUSSIAN ? 0 : (size_t)M->n * st.batch);
    make_starts(st, M->n, X0, 0, st.batch, hX.data());
    if ((rc = load_batch(ctx, E.stream, w, st, M->n, st.batch, hX.data()))) return rc;
    Plan p;
    if ((rc = tune_plan(ctx, E, M, st, w, allow_big, &p))) return rc;
    std::lock_guard<std::mutex> g(ctx->plan_mu);
    ctx->plans[key] = p;
    save_plans(ctx);
    *out = p;
    return SPCA_OK;
}

// Kernel time per phase (0: A*X, 1: A^T*Z, 2: epilogue/step/convergence), from
// hipEvent pairs recorded on the solver stream and folded in after each host sync.
struct PhaseTimer {
    Engine* eng;
    bool on = true;   // off: no events are created or recorded (the allocation-free solve)
    size_t used = 0;  // events handed out since the last flush (always pairs)
    double ms[4] = {0, 0, 0, 0};
    std::vector<int> phase;
    hipError_t pair(hipEvent_t* a, hipEvent_t* b, int ph) {
        if (!on) {
            *a = *b = nullptr;
            return hipSuccess;
        }
        while (eng->events.size() < used + 2) {
            hipEvent_t e;
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            eng->events.push_back(e);
        }
        *a = eng->events[used];
        *b = eng->events[used + 1];
        used += 2;
        phase.push_back(ph);
        return hipSuccess;
    }
    void flush() {
        for (size_t i = 0; i < phase.size(); i++) {
            float t = 0;
            if (hipEventElapsedTime(&t, eng->events[2 * i], eng->events[2 * i + 1]) == hipSuccess) ms[phase[i]] += t;
        }
        phase.clear();
        used = 0;
    }
};

// Checks a request and normalises its settings (clamp of s, batch rounding); makes sure the
// matrix holds the storage the precision needs.  Returns the number of batches in *nb.
int prepare_settings(spca_ctx ctx, Matrix* M, const spca_settings* st_in, const float* X0, const float* x_out,
                     spca_settings* st_out, int* nb, bool may_add_storage = true) {
    spca_settings st = *st_in;
    if (!valid_formulation(st.formulation)) return fail(ctx, SPCA_EINVAL, "formulation must be 0..7");
    if (st.precision < 0 || st.precision > 2) return fail(ctx, SPCA_EINVAL, "precision must be 0..2");
    if (st.max_iter < 1 || st.total_sp < 1 || st.batch < 0) return fail(ctx, SPCA_EINVAL, "bad iteration/SP counts");
    if (st.start < SPCA_START_EXPLICIT || st.start > SPCA_START_GAUSSIAN)
        return fail(ctx, SPCA_EINVAL, "start must be a spca_start value");
    if (st.start == SPCA_START_EXPLICIT && !X0) return fail(ctx, SPCA_EINVAL, "X0 is required for explicit starts");
    if (!x_out) return fail(ctx, SPCA_EINVAL, "x_out is required");
    if (M->sparse && st.precision != SPCA_PREC_FP32)
        return fail(ctx, SPCA_EINVAL, "a sparse (CSC) matrix is solved in FP32 only");
    if (M->sparse && st.resolve_iters > 0)
        return fail(ctx, SPCA_EINVAL, "the FP64 support re-solve needs a dense matrix");
    if (M->n >= kMaxCols)
        return fail(ctx, SPCA_EINVAL, "n must be below 2^23 columns (the step kernel's fixed-point sums)");
    unsigned need = st.precision == 0 ? SPCA_STORE_FP32 : st.precision == 1 ? SPCA_STORE_BF16X3 : SPCA_STORE_FP16;
    // the fixed-point product, and the FP32 row-major copy for the exact signs of z
    if (!M->sparse && uses_sign_products(st, M->mp)) need |= SPCA_STORE_FIX8 | SPCA_STORE_FP32;
    if (!may_add_storage && !M->sparse && (M->storage & need) != need)
        return fail(ctx, SPCA_EINVAL, "the matrix lacks storage form(s) " + std::to_string(need & ~M->storage) +
                                          " this solve reads; add them (spca_matrix_add_storage) before the solve");
    int rc = M->sparse ? SPCA_OK : add_storage(ctx, M, need);
    if (rc) return rc;
    *nb = spca_settings_check(&st, (unsigned)M->n);
    if (st.batch > kMaxSparseGroups * 256)
        return fail(ctx, SPCA_EINVAL, "batch must be at most " + std::to_string(kMaxSparseGroups * 256) + " SPs");
    if (st.constraint < 1) st.constraint = 1;
    *st_out = st;
    return SPCA_OK;
}

// Earlier components a solve deflates by (device arrays; see launch_deflate).
struct Deflation {
    const int* idx = nullptr;
    const float* val = nullptr;
    const int* off = nullptr;
    int ncomp = 0;
};

// a phase-timing event, or nothing when phase timing is off
hipError_t record(hipEvent_t e, hipStream_t s) { return e ? hipEventRecord(e, s) : hipSuccess; }

// fx (the allocation-free solve): the workspace and the explained-variance partials come from
// the fixed device array; the plan is the cached or heuristic one (no tuning run), no storage
// form is added, phase times are not taken (no events are created), and batch mode only.
int solve_impl(spca_ctx ctx, Engine& E, Matrix* M, const spca_settings* st_in, const float* X0, float* x_out,
               spca_stats* stats, spca_sp_result* per_sp, float* X_final, const Deflation* defl = nullptr,
               FixedScratch* fx = nullptr) {
    const auto t_wall0 = std::chrono::steady_clock::now();
    spca_settings st;
    int total_batches = 0;
    int rc = prepare_settings(ctx, M, st_in, X0, x_out, &st, &total_batches, fx == nullptr);
    if (rc) return rc;
    if (fx && (st.use_otf || M->sparse || defl))
        return fail(ctx, SPCA_EINVAL, "the allocation-free solve runs dense matrices in batch mode (no OTF)");
    const int n = M->n, m = M->m;
    const int batch = st.batch;
    const int bp = round_up(batch, kPadB);
    Plan plan;
    if (fx) plan = peek_plan(ctx, st.precision, M->mp, M->np, bp, true);
    else if ((rc = plan_for(ctx, E, M, st, bp, true, X0, &plan))) return rc;
    const int splits = plan.splits;
    const AtzForm atz = atz_form(M, st, bp, plan.big);
    Workspace fixed_ws;
    if (fx) {
        hipError_t e = hipSuccess;
        FixedScratch probe{nullptr, fx->cap, 0};
        carve_ws(fixed_ws, probe, M->np, M->mp, bp, splits, st.precision, atz.zprec, E.stream, &e);
        const size_t ev_need = (ev_scratch_doubles(m, n, 1) + 32) * sizeof(double);
        if (probe.used + ev_need > fx->cap)
            return fail(ctx, SPCA_EINVAL, "an allocation-free solve of this shape needs " +
                                              std::to_string(probe.used + ev_need) + " bytes of scratch; the fixed array holds " +
                                              std::to_string(fx->cap));
        fx->used = 0;
        carve_ws(fixed_ws, *fx, M->np, M->mp, bp, splits, st.precision, atz.zprec, E.stream, &e);
        if (e != hipSuccess) return fail(ctx, SPCA_EDEVICE, std::string("scratch clear: ") + hipGetErrorString(e));
    } else if ((rc = ensure_workspace(ctx, E, M->np, M->mp, bp, splits))) {
        return rc;
    }
    Workspace& w = fx ? fixed_ws : E.ws;
    hipStream_t s = E.stream;
    SpState sp = state_of(w);
    const int check_every = st.check_every > 0 ? st.check_every : 1;

    const SignRefine rf = sign_refine(M, st, w);
    const Operand xop = x_operand(w, st.precision);
    const Operand zop = z_operand(w, atz.zprec);
    const long zsplit = (long)M->mp * bp;

    StepArgs sa;
    sa.formulation = st.formulation;
    sa.precision = st.precision;
    sa.n = n;
    sa.ldv = M->np;
    sa.s = st.constraint;
    sa.gamma = st.penalty;
    sa.batch = batch;
    sa.sparse_rewrite = defl ? 0 : 1;  // deflation changes x between steps
    set_exact_ties(sa, M, st, w);

    std::vector<float> hX(st.start == SPCA_START_GAUSSIAN ? 0 : (size_t)n * batch);
    Results R;
    double loop_ms = 0;
    PhaseTimer pt{&E, fx == nullptr};
    hipEvent_t* loop_ev = E.loop_ev;

    if (st.use_otf) {
        // the sparse gather runs for groups whose active slots (refills included) are all sparse
        const int otf_cap = M->sparse ? 0 : sparse_cap(st, n);
        // ---- on-the-fly refill (reference OTF driver, sparse_PCA_solver.h:83-169): every SP is
        // resident in a device pool; finished slots take the next SP on the device.
        const int total = st.total_sp;
        float* pool = nullptr;
        float* xbest = nullptr;
        int* ibuf = nullptr;
        float* fbuf = nullptr;
        uint64_t ob = 0;
        auto release = [&]() {
            (void)hipStreamSynchronize(s);
            dev_free(ctx, pool, (uint64_t)sizeof(float) * w.np * total);
            dev_free(ctx, xbest, (uint64_t)sizeof(float) * w.np);
            dev_free(ctx, ibuf, (uint64_t)sizeof(int) * (3 * (size_t)w.bp + 3 + 2 * (size_t)total));
            dev_free(ctx, fbuf, (uint64_t)sizeof(float) * (1 + (size_t)total));
        };
        if ((rc = dev_alloc(ctx, &pool, (size_t)w.np * total, &ob)) || (rc = dev_alloc(ctx, &xbest, (size_t)w.np, &ob)) ||
            (rc = dev_alloc(ctx, &ibuf, 3 * (size_t)w.bp + 3 + 2 * (size_t)total, &ob)) ||
            (rc = dev_alloc(ctx, &fbuf, 1 + (size_t)total, &ob))) {
            release();
            return rc;
        }
        OtfState o;
        o.order = ibuf;
        o.cur_it = ibuf + w.bp;
        o.refill = ibuf + 2 * w.bp;
        o.generated = ibuf + 3 * w.bp;
        o.sp_iters = ibuf + 3 * w.bp + 3;
        o.sp_nnz = o.sp_iters + total;
        o.best_val = fbuf;
        o.sp_val = fbuf + 1;
        o.pool = pool;
        o.xbest = xbest;
        // pool of all starting points, padded rows zero
        {
            std::vector<float> hp(st.start == SPCA_START_GAUSSIAN ? 0 : (size_t)n * total);
            make_starts(st, n, X0, 0, total, hp.data());
            HIP_TRY(ctx, hipMemsetAsync(pool, 0, sizeof(float) * (size_t)w.np * total, s));
            if (st.start == SPCA_START_GAUSSIAN)
                HIP_TRY(ctx, launch_gen_starts(pool, w.np, n, total, 0, st.seed, s));
            else
                HIP_TRY(ctx, hipMemcpy2DAsync(pool, sizeof(float) * w.np, hp.data(), sizeof(float) * n,
                                              sizeof(float) * n, total, hipMemcpyHostToDevice, s));
            HIP_TRY(ctx, hipMemsetAsync(w.X, 0, sizeof(float) * (size_t)w.np * w.bp, s));
            HIP_TRY(ctx, hipMemcpyAsync(w.X, pool, sizeof(float) * (size_t)w.np * batch, hipMemcpyDeviceToDevice, s));
            if (st.precision) {
                HIP_TRY(ctx, hipMemsetAsync(w.Xhi, 0, sizeof(uint16_t) * (size_t)w.np * w.bp, s));
                HIP_TRY(ctx, hipMemsetAsync(w.Xlo, 0, sizeof(uint16_t) * (size_t)w.np * w.bp, s));
                HIP_TRY(ctx, launch_x_operand(xop, st.precision, w.np, batch, s));
            }
            std::vector<int> hi(3 * (size_t)w.bp + 3 + 2 * (size_t)total, 0);
            for (int i = 0; i < batch; i++) hi[i] = i;  // order
            hi[3 * w.bp] = batch;                       // generated
            hi[3 * w.bp + 1] = -1;
            hi[3 * w.bp + 2] = -1;
            std::vector<float> hf(1 + (size_t)total, 0.f);
            hf[0] = -1.f;  // the_best_solution_value starts at -1
            HIP_TRY(ctx, hipMemcpyAsync(ibuf, hi.data(), sizeof(int) * hi.size(), hipMemcpyHostToDevice, s));
            HIP_TRY(ctx, hipMemcpyAsync(fbuf, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice, s));
            HIP_TRY(ctx, hipMemsetAsync(w.fstate, 0, sizeof(float) * kStateFloats * (size_t)w.bp, s));
            HIP_TRY(ctx, hipMemsetAsync(w.istate, 0, sizeof(int) * (kStateInts * (size_t)w.bp + state_flag_ints(w.bp)), s));
            std::vector<int> act(w.bp, 0);
            std::fill(act.begin(), act.begin() + batch, 1);
            HIP_TRY(ctx, hipMemcpyAsync(sp.active, act.data(), sizeof(int) * w.bp, hipMemcpyHostToDevice, s));
            HIP_TRY(ctx, hipStreamSynchronize(s));
        }
        // every SP needs at most max_iter iterations and every loop iteration advances all slots
        const long cap = (long)total * st.max_iter / batch + st.max_iter + 1;
        HIP_TRY(ctx, hipEventRecord(loop_ev[0], s));
        for (long it = 0; it < cap; it++) {
            hipEvent_t e0, e1;
            HIP_TRY(ctx, pt.pair(&e0, &e1, 0));
            HIP_TRY(ctx, record(e0, s));
            if (defl) HIP_TRY(ctx, launch_deflate(xop, st.precision, w.np, defl->idx, defl->val, defl->off, defl->ncomp,
                                                  batch, sp, s));
            HIP_TRY(ctx, launch_ax(M, st, w, splits, plan.big, bp, batch, zsplit, sp, otf_cap > 0, s));
            HIP_TRY(ctx, record(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 2));
            HIP_TRY(ctx, record(e0, s));
            HIP_TRY(ctx, launch_prep_z(w.Zpart, splits, zsplit, M->mp, M->mp, zop, st.formulation, atz.zprec, batch,
                                       sp, rf, s));
            HIP_TRY(ctx, record(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 1));
            HIP_TRY(ctx, record(e0, s));
            HIP_TRY(ctx, launch_atz(atz, M, w, bp, batch, sp, s));
            HIP_TRY(ctx, record(e1, s));
            if (defl) HIP_TRY(ctx, launch_deflate(Operand{w.V, nullptr, nullptr, w.bp}, 0, w.np, defl->idx, defl->val,
                                                  defl->off, defl->ncomp, batch, sp, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 3));
            HIP_TRY(ctx, record(e0, s));
            HIP_TRY(ctx, launch_step(w.V, xop, sa, sp, s));
            HIP_TRY(ctx, record(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 2));
            HIP_TRY(ctx, record(e0, s));
            HIP_TRY(ctx, launch_otf_converge(sp, o, batch, st.tolerance, st.max_iter, total, s));
            HIP_TRY(ctx, launch_otf_refill(sp, o, xop, st.precision, n, w.np, batch, otf_cap, s));
            HIP_TRY(ctx, record(e1, s));
            if ((it + 1) % check_every == 0 || it + 1 == cap) {
                HIP_TRY(ctx, hipMemcpyAsync(E.h_done, sp.done, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
                HIP_TRY(ctx, hipStreamSynchronize(s));
                pt.flush();
                if (E.h_done[0]) break;
                if (E.cancel && E.cancel->load()) {
                    release();
                    return fail(ctx, SPCA_ECANCELLED, "cancelled");
                }
            }
        }
        HIP_TRY(ctx, hipEventRecord(loop_ev[1], s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        pt.flush();
        {
            float t = 0;
            HIP_TRY(ctx, hipEventElapsedTime(&t, loop_ev[0], loop_ev[1]));
            loop_ms += t;
        }
        int done_h[5], gen_h[3];
        float bestv;
        std::vector<float> spv(total);
        std::vector<int> spi(total), spn(total);
        HIP_TRY(ctx, hipMemcpyAsync(done_h, sp.done, sizeof(int) * 5, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(gen_h, o.generated, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(&bestv, o.best_val, sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(spv.data(), o.sp_val, sizeof(float) * total, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(spi.data(), o.sp_iters, sizeof(int) * total, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(spn.data(), o.sp_nnz, sizeof(int) * total, hipMemcpyDeviceToHost, s));
        R.x_best.assign(n, 0.f);
        HIP_TRY(ctx, hipMemcpyAsync(R.x_best.data(), xbest, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        double evb = 0;
        if ((rc = run_ev(ctx, s, M, xbest, w.np, 1, st.formulation, w.ev))) {
            release();
            return rc;
        }
        HIP_TRY(ctx, hipMemcpyAsync(&evb, w.ev, sizeof(double), hipMemcpyDeviceToHost, s));
        float emax;
        HIP_TRY(ctx, hipMemcpyAsync(&emax, sp.errmax, sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        release();
        R.total_it = (unsigned)done_h[1];
        R.sparse_it = (unsigned)done_h[4];
        R.list_tries = (unsigned)done_h[2];
        R.list_steps = (unsigned)done_h[3];
        R.best = bestv;
        R.best_sp = gen_h[2];
        R.last_err = emax;
        R.ev_best = evb;
        const int best_sp = R.best_sp;
        for (int k = 0; k < total; k++) R.sp_its += (unsigned long long)spi[k];
        if (per_sp)
            for (int k = 0; k < total; k++) {
                per_sp[k].value = spv[k];
                per_sp[k].iters = spi[k];
                per_sp[k].nnz = spn[k];
                per_sp[k].converged = spi[k] < st.max_iter;
                per_sp[k].ev_fp64 = (k == best_sp) ? evb : std::nan("");
            }
        (void)X_final;  // slots are reused: final iterates per SP are not kept in OTF mode
    } else {
    // union A*X (sparse_products = 2, dense matrix, no deflation): after each iteration the
    // host reads |U| and the next A*X is the FP32 product over U when |U| <= ucap
    const int ucap = (M->sparse || defl || fx) ? 0 : union_cap(st, n);
    if (ucap > 0 && (rc = ensure_union(ctx, w, s, round_up(ucap, kGemmBK * splits)))) return rc;
    for (int bt = 0; bt < total_batches; bt++) {
        const int shift = bt * batch;
        const int cap = M->sparse ? 0 : sparse_cap(st, n);
        make_starts(st, n, X0, shift, batch, hX.data());
        if ((rc = load_batch(ctx, s, w, st, n, batch, hX.data(), shift))) return rc;
        int uK = 0;  // |U| for the next iteration's A*X (0: not known or empty; the dense product)
        HIP_TRY(ctx, hipEventRecord(loop_ev[0], s));
        for (int it = 0; it < st.max_iter; it++) {
            hipEvent_t e0, e1;
            HIP_TRY(ctx, pt.pair(&e0, &e1, 0));
            HIP_TRY(ctx, record(e0, s));
            if (defl) HIP_TRY(ctx, launch_deflate(xop, st.precision, w.np, defl->idx, defl->val, defl->off, defl->ncomp,
                                                  batch, sp, s));
            int ax_splits = splits;
            if (uK > 0 && uK <= ucap) {
                // K split so each split is at least 512 deep; the sparse gather keeps its groups
                int us = 1;
                while (us * 2 <= splits && round_up(std::max(uK, 1), kGemmBK * us * 2) / (us * 2) >= 512) us *= 2;
                const int kpad = round_up(std::max(uK, 1), kGemmBK * us);
                HIP_TRY(ctx, launch_union_gather(M->a32_cm, M->mp, M->mp, w.X, w.np, bp, w.uidx, uK, kpad, w.Au, w.Xu,
                                                 w.ldu, s));
                HIP_TRY(ctx, launch_gemm_tn(0, w.Au, nullptr, nullptr, w.ldu, w.Xu, nullptr, nullptr, w.ldu, w.Zpart,
                                            M->mp, zsplit, M->mp, bp, kpad, us, 0, sp.done, sp.done + kSparseFlags, s));
                HIP_TRY(ctx, launch_gather(M, w, batch, sp, s));
                ax_splits = us;
                R.union_it++;
            } else {
                // one of the two dense A*X forms runs; the other sees the device flag and returns
                HIP_TRY(ctx, launch_ax(M, st, w, splits, plan.big, bp, batch, zsplit, sp, true, s));
            }
            HIP_TRY(ctx, record(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 2));
            HIP_TRY(ctx, record(e0, s));
            HIP_TRY(ctx, launch_prep_z(w.Zpart, ax_splits, zsplit, M->mp, M->mp, zop, st.formulation, atz.zprec, batch,
                                       sp, rf, s));
            HIP_TRY(ctx, record(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 1));
            HIP_TRY(ctx, record(e0, s));
            HIP_TRY(ctx, launch_atz(atz, M, w, bp, batch, sp, s));
            HIP_TRY(ctx, record(e1, s));
            if (defl) HIP_TRY(ctx, launch_deflate(Operand{w.V, nullptr, nullptr, w.bp}, 0, w.np, defl->idx, defl->val,
                                                  defl->off, defl->ncomp, batch, sp, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 3));
            HIP_TRY(ctx, record(e0, s));
            HIP_TRY(ctx, launch_step(w.V, xop, sa, sp, s));
            HIP_TRY(ctx, record(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 2));
            HIP_TRY(ctx, record(e0, s));
            HIP_TRY(ctx, launch_converge(sp, batch, it, st.tolerance, st.max_iter,
                                         st.termination == SPCA_TERM_PER_SP, cap, s));
            if (ucap > 0 && it + 1 < st.max_iter)
                HIP_TRY(ctx, launch_union_list(w.X, w.np, n, batch, sp, w.uflag, w.uidx, w.ustat, s));
            HIP_TRY(ctx, record(e1, s));
            if (ucap > 0 && it + 1 < st.max_iter) {
                // stat = {|U|, done[0], done[1]}: the host decides the next A*X form
                HIP_TRY(ctx, hipMemcpyAsync(E.h_done + 4, w.ustat, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
                HIP_TRY(ctx, hipStreamSynchronize(s));
                pt.flush();
                if (E.h_done[5]) break;
                if (E.cancel && E.cancel->load()) return fail(ctx, SPCA_ECANCELLED, "cancelled");
                uK = E.h_done[4];
            } else if ((it + 1) % check_every == 0 || it + 1 == st.max_iter) {
                HIP_TRY(ctx, hipMemcpyAsync(E.h_done, sp.done, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
                HIP_TRY(ctx, hipStreamSynchronize(s));
                pt.flush();
                if (E.h_done[0]) break;
                if (E.cancel && E.cancel->load()) return fail(ctx, SPCA_ECANCELLED, "cancelled");
            }
        }
        HIP_TRY(ctx, hipEventRecord(loop_ev[1], s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        pt.flush();
        {
            float t = 0;
            HIP_TRY(ctx, hipEventElapsedTime(&t, loop_ev[0], loop_ev[1]));
            loop_ms += t;
        }
        if ((rc = collect_batch(ctx, s, M, w, st, shift, batch, per_sp != nullptr || stats != nullptr, per_sp, X_final,
                                R, fx)))
            return rc;
    }
    }
    finish_solve(R, n, st, x_out, stats, loop_ms, pt.ms, t_wall0);
    return SPCA_OK;
}

// FP64 re-solve on the support of x (host vector of n floats)
int resolve_impl(spca_ctx ctx, Engine& E, Matrix* M, int formulation, const float* x, int max_iter, double tol,
                 double* y_out, double* ev_out, int* iters_out) {
    if (!valid_formulation(formulation) || !x || max_iter < 1) return fail(ctx, SPCA_EINVAL, "bad re-solve arguments");
    if (M->sparse) return fail(ctx, SPCA_EINVAL, "the FP64 support re-solve needs a dense matrix");
    const int n = M->n, m = M->m;
    std::vector<int> idx;
    double nx = 0;
    for (int i = 0; i < n; i++)
        if (x[i] != 0.f) {
            idx.push_back(i);
            nx += (double)x[i] * (double)x[i];
        }
    const int k = (int)idx.size();
    if (k == 0 || !(nx > 0) || !std::isfinite(nx)) return fail(ctx, SPCA_EINVAL, "re-solve needs a finite nonzero x");
    std::vector<double> y0(k);
    for (int b = 0; b < k; b++) y0[b] = (double)x[idx[b]] / std::sqrt(nx);
    const int nb = (m + 255) / 256;
    double* dbuf = nullptr;
    int* ibuf = nullptr;
    uint64_t bytes = 0;
    const size_t nd = (size_t)m * k + k + m + k + nb + 1;
    int rc;
    if ((rc = dev_alloc(ctx, &dbuf, nd, &bytes)) || (rc = dev_alloc(ctx, &ibuf, (size_t)k + 1, &bytes))) {
        dev_free(ctx, dbuf, sizeof(double) * nd);
        return rc;
    }
    double* As = dbuf;
    double* y = As + (size_t)m * k;
    double* z = y + k;
    double* yp = z + m;
    double* part = yp + k;
    double* ev = part + nb;
    int* didx = ibuf;
    int* flag = ibuf + k;
    hipStream_t s = E.stream;
    auto done = [&](int code) {
        (void)hipStreamSynchronize(s);
        dev_free(ctx, dbuf, sizeof(double) * nd);
        dev_free(ctx, ibuf, sizeof(int) * ((size_t)k + 1));
        return code;
    };
    const int zero = 0;
    const double dzero = 0;
    if (hipMemcpyAsync(didx, idx.data(), sizeof(int) * k, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(y, y0.data(), sizeof(double) * k, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(flag, &zero, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(ev, &dzero, sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
        launch_resolve(M->a32_cm, M->mp, m, didx, k, spca_is_l1_variance(formulation), As, y, z, yp, part, ev, flag,
                       tol, max_iter, s) != hipSuccess)
        return done(fail(ctx, SPCA_EDEVICE, "re-solve launch failed"));
    std::vector<double> yh(k);
    double evh = 0;
    int fl = 0;
    if (hipMemcpyAsync(yh.data(), y, sizeof(double) * k, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&evh, ev, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(&fl, flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return done(fail(ctx, SPCA_EDEVICE, "re-solve copy failed"));
    if (y_out) {
        std::fill(y_out, y_out + n, 0.0);
        for (int b = 0; b < k; b++) y_out[idx[b]] = yh[b];
    }
    if (ev_out) *ev_out = evh;
    if (iters_out) *iters_out = fl ? -1 : max_iter;  // -1: converged before max_iter
    return done(SPCA_OK);
}

int solve_and_resolve(spca_ctx ctx, Engine& E, Matrix* M, const spca_settings* settings, const float* X0,
                      float* x_out, spca_stats* stats, spca_sp_result* per_sp, float* X_final,
                      const Deflation* defl = nullptr) {
    int rc = solve_impl(ctx, E, M, settings, X0, x_out, stats, per_sp, X_final, defl);
    if (rc == SPCA_OK && settings->resolve_iters > 0 && stats) {
        int its = 0;
        rc = resolve_impl(ctx, E, M, settings->formulation, x_out, settings->resolve_iters, settings->resolve_tol,
                          nullptr, &stats->ev_resolve_fp64, &its);
        stats->resolve_iterations = its;
    }
    return rc;
}


// ---------------------------------------------------------------- fused multi-job solve
// Jobs of different shapes and formulations iterate in lockstep: each AM iteration of all
// of them is one launch per phase (launch_fused_iteration), and a chunk of check_every
// iterations is captured once into a hipGraph and replayed.  Each job keeps its own
// workspace, state and termination; a finished job's work items return at once, and the
// host stops when every job has finished.  Per job the tile order, K split, partial-sum
// order and kernel bodies are those of a solo solve, so FP32 results equal a solo solve's
// bit for bit.
struct FusedRun {
    Matrix* M = nullptr;
    spca_settings st{};
    Workspace ws;
    int splits = 0;
};

// Loads one job into a fused group of `precision` (-1: the job's own): settings, plan,
// workspace, starting points.  may_add_storage false: the matrix must already hold the
// storage the precision reads (the request server builds it at submission).
int fused_member_init(spca_ctx ctx, Engine& E, Matrix* M, const spca_settings* st_in, const float* X0,
                      const float* x_out, int precision, bool may_add_storage, FusedRun& r) {
    r.M = M;
    if (!M || !st_in) return fail(ctx, SPCA_EINVAL, "bad matrix or settings");
    if (M->sparse) return fail(ctx, SPCA_EINVAL, "sparse matrices solve alone");
    int nb = 0, rc;
    if ((rc = prepare_settings(ctx, M, st_in, X0, x_out, &r.st, &nb, may_add_storage))) return rc;
    if (nb != 1 || r.st.use_otf)
        return fail(ctx, SPCA_EINVAL, "fused jobs run one batch (total_sp <= batch) without OTF");
    if (precision >= 0 && r.st.precision != precision) return fail(ctx, SPCA_EINVAL, "fused jobs must share one precision");
    const int bp = round_up(r.st.batch, kPadB);
    Plan plan;
    if ((rc = plan_for(ctx, E, M, r.st, bp, false, X0, &plan, &r.ws))) return rc;
    r.splits = plan.splits;
    if ((rc = ensure_ws(ctx, r.ws, E.stream, M->np, M->mp, bp, r.splits))) return rc;
    std::vector<float> hX(r.st.start == SPCA_START_GAUSSIAN ? 0 : (size_t)M->n * r.st.batch);
    make_starts(r.st, M->n, X0, 0, r.st.batch, hX.data());
    return load_batch(ctx, E.stream, r.ws, r.st, M->n, r.st.batch, hX.data());
}

// Work-item offsets of a fused group: each job's tiles and columns follow the previous job's.
struct FusedTotals {
    int t1 = 0, t2 = 0, c0 = 0, t3 = 0, t4 = 0, t5 = 0, any_fix8 = 0;
};

// The job-table entry of one member at the current offsets, which it then advances.
void fused_entry(FusedRun& r, int precision, FusedTotals& T, FusedJob& f) {
    Matrix* M = r.M;
    f = FusedJob();
    f.p1_32 = M->a32_rm;
    f.p1_hi = precision == 1 ? M->hi_rm : M->h_rm;
    f.p1_lo = M->lo_rm;
    const AtzForm atz = atz_form(M, r.st, r.ws.bp, 0);
    f.p2_32 = M->a32_cm;
    f.p2_hi = atz.hi;
    f.p2_lo = atz.lo;
    f.p2_q = atz.q;
    f.p2_qscale = atz.qscale;
    f.p2_prec = atz.prec;
    f.zprec = atz.zprec;
    T.any_fix8 |= atz.prec == 4;
    f.Zpart = r.ws.Zpart;
    f.V = r.ws.V;
    f.gl = r.ws.gl;
    f.xop = x_operand(r.ws, precision);
    f.zop = z_operand(r.ws, atz.zprec);
    f.rf = sign_refine(M, r.st, r.ws);
    f.sp = state_of(r.ws);
    f.sa.formulation = r.st.formulation;
    f.sa.precision = precision;
    f.sa.n = M->n;
    f.sa.ldv = M->np;
    f.sa.s = r.st.constraint;
    f.sa.gamma = r.st.penalty;
    f.sa.batch = r.st.batch;
    f.sa.sparse_rewrite = 1;
    set_exact_ties(f.sa, M, r.st, r.ws);
    f.mp = M->mp;
    f.np = M->np;
    f.bp = r.ws.bp;
    f.splits = r.splits;
    f.zsplit = (long)M->mp * f.bp;
    f.tol = r.st.tolerance;
    f.max_iter = r.st.max_iter;
    f.per_sp = r.st.termination == SPCA_TERM_PER_SP;
    f.t1 = T.t1;
    f.t2 = T.t2;
    f.c0 = T.c0;
    f.t3 = T.t3;
    f.sparse_cap = sparse_cap(r.st, M->n);
    // FP32 batches of at most 64 SPs: narrow product tiles, as in a solo solve
    const int nb = (precision == 0 && r.st.batch <= 64) ? (r.st.batch <= 32 ? 32 : 64) : 0;
    f.nb1 = nb;
    f.nb2 = f.p2_prec == 0 ? nb : 0;
    T.t1 += (f.mp / kGemmBM) * (f.nb1 ? 1 : f.bp / kGemmBN) * f.splits;
    T.t2 += (f.np / kGemmBM) * (f.nb2 ? 1 : f.bp / kGemmBN);
    T.c0 += r.st.batch;
    T.t3 += (f.mp + kSparseRows - 1) / kSparseRows * r.st.batch;
    // small batches: the K-split partials are summed over (row block, SP) items first, as
    // in a solo solve (the same order, so the same bits)
    f.t4 = T.t4;
    f.psum = f.splits > 1 && r.st.batch <= 256;
    if (f.psum) T.t4 += (f.mp + 255) / 256 * r.st.batch;
    // the steps' chunk workgroups: (SP, chunk) items
    f.t5 = T.t5;
    T.t5 += step_chunks(M->n) * r.st.batch;
}

// A fused member's results once its state says done: the best SP, the final normalisation,
// the statistics and (resolve_iters > 0) the FP64 support re-solve.
// loop_ms(total iterations) gives the loop time the statistics report.
template <class LoopMs>
int fused_member_finish(spca_ctx ctx, Engine& E, FusedRun& r, float* x_out, spca_stats* stats, spca_sp_result* per_sp,
                        LoopMs loop_ms, std::chrono::steady_clock::time_point t_wall0) {
    Results R;
    int rc;
    if ((rc = collect_batch(ctx, E.stream, r.M, r.ws, r.st, 0, r.st.batch, per_sp != nullptr || stats != nullptr, per_sp,
                            nullptr, R)))
        return rc;
    finish_solve(R, r.M->n, r.st, x_out, stats, loop_ms((long)R.total_it), nullptr, t_wall0);
    if (stats && r.st.resolve_iters > 0) {
        int its = 0;
        if ((rc = resolve_impl(ctx, E, r.M, r.st.formulation, x_out, r.st.resolve_iters, r.st.resolve_tol, nullptr,
                               &stats->ev_resolve_fp64, &its)))
            return rc;
        stats->resolve_iterations = its;
    }
    return SPCA_OK;
}

int fused_impl(spca_ctx ctx, Engine& E, spca_fused_job* jobs, int njobs, unsigned flags) {
    const auto t_wall0 = std::chrono::steady_clock::now();
    if (!jobs || njobs < 1) return fail(ctx, SPCA_EINVAL, "fused solve needs at least one job");
    hipStream_t s = E.stream;
    std::vector<FusedRun> run(njobs);
    for (int k = 0; k < njobs; k++) jobs[k].status = SPCA_EAGAIN;
    FusedJob* d_jobs = nullptr;
    int* d_done = nullptr;
    uint64_t tb = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<hipEvent_t> chunk_ev;
    auto cleanup = [&](int code) {
        (void)hipStreamSynchronize(s);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : chunk_ev) (void)hipEventDestroy(e);
        dev_free(ctx, d_jobs, sizeof(FusedJob) * (uint64_t)njobs);
        dev_free(ctx, d_done, sizeof(int) * 4);
        for (auto& r : run) free_ws(ctx, r.ws, s);
        for (int k = 0; k < njobs; k++)
            if (jobs[k].status == SPCA_EAGAIN) jobs[k].status = code;
        return code;
    };
    // ---- jobs: settings, matrices (held for the call), plans, workspaces, starting points
    int rc;
    int precision = -1, check_every = 1 << 30, max_iter = 0;
    for (int k = 0; k < njobs; k++) {
        spca_fused_job& J = jobs[k];
        FusedRun& r = run[k];
        auto it = ctx->mats.find(J.matrix);  // the caller holds ctx->mu
        if (J.stats) std::memset(J.stats, 0, sizeof(*J.stats));
        if ((rc = fused_member_init(ctx, E, it != ctx->mats.end() ? it->second : nullptr, J.settings, J.X0, J.x_out,
                                    precision, true, r))) {
            J.status = rc;
            return cleanup(fail(ctx, rc, "fused job " + std::to_string(k) + ": " + tl_err));
        }
        precision = r.st.precision;
        check_every = std::min(check_every, std::max(1, r.st.check_every));
        max_iter = std::max(max_iter, r.st.max_iter);
    }
    // ---- job table
    std::vector<FusedJob> tab(njobs);
    FusedTotals T;
    for (int k = 0; k < njobs; k++) fused_entry(run[k], precision, T, tab[k]);
    const int t1 = T.t1, t2 = T.t2, c0 = T.c0, t3 = T.t3, t4 = T.t4, t5 = T.t5, any_fix8 = T.any_fix8;
    if ((rc = dev_alloc(ctx, &d_jobs, (size_t)njobs, &tb)) || (rc = dev_alloc(ctx, &d_done, 4, &tb)))
        return cleanup(rc);
    if (hipMemcpyAsync(d_jobs, tab.data(), sizeof(FusedJob) * njobs, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemsetAsync(d_done, 0, sizeof(int) * 4, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return cleanup(fail(ctx, SPCA_EDEVICE, "fused job table upload failed"));
    const int chunk = std::min(check_every, max_iter);
    if (flags & SPCA_FUSED_GRAPH) {
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        for (int i = 0; i < chunk && e == hipSuccess; i++)
            e = launch_fused_iteration(precision, d_jobs, njobs, t1, t2, c0, t3, t4, t5, any_fix8, d_done, s);
        hipError_t e2 = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess)
            return cleanup(fail(ctx, SPCA_EDEVICE, std::string("fused iteration graph: ") + hipGetErrorString(e)));
    }
    // ---- iterate in chunks until every job has finished
    const int max_chunks = (max_iter + chunk - 1) / chunk + 1;
    auto new_event = [&](hipEvent_t* ev) {
        if (hipEventCreate(ev) != hipSuccess) return false;
        chunk_ev.push_back(*ev);
        return true;
    };
    hipEvent_t ev0;
    if (!new_event(&ev0) || hipEventRecord(ev0, s) != hipSuccess)
        return cleanup(fail(ctx, SPCA_EDEVICE, "event record failed"));
    int chunks = 0;
    for (; chunks < max_chunks; chunks++) {
        hipError_t e = hipSuccess;
        if (gexec)
            e = hipGraphLaunch(gexec, s);
        else
            for (int i = 0; i < chunk && e == hipSuccess; i++)
                e = launch_fused_iteration(precision, d_jobs, njobs, t1, t2, c0, t3, t4, t5, any_fix8, d_done, s);
        hipEvent_t ev;
        if (e == hipSuccess && !new_event(&ev)) e = hipErrorOutOfMemory;
        if (e == hipSuccess) e = hipEventRecord(ev, s);
        if (e == hipSuccess) e = hipMemcpyAsync(E.h_done, d_done, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return cleanup(fail(ctx, SPCA_EDEVICE, std::string("fused iteration: ") + hipGetErrorString(e)));
        if (E.h_done[0] >= njobs) {
            chunks++;
            break;
        }
    }
    if (E.h_done[0] < njobs) return cleanup(fail(ctx, SPCA_EDEVICE, "fused jobs did not finish"));
    // ---- per-job results
    for (int k = 0; k < njobs; k++) {
        spca_fused_job& J = jobs[k];
        FusedRun& r = run[k];
        // the chunk in which the job finished bounds its loop time
        auto loop_ms = [&](long its) {
            const int c = std::min(chunks, std::max(1, (int)((its + chunk - 1) / chunk)));
            float t = 0;
            (void)hipEventElapsedTime(&t, chunk_ev[0], chunk_ev[c]);
            return (double)t;
        };
        if ((rc = fused_member_finish(ctx, E, r, J.x_out, J.stats, J.per_sp, loop_ms, t_wall0))) return cleanup(rc);
        J.status = SPCA_OK;
    }
    return cleanup(SPCA_OK);
}

// ---------------------------------------------------------------- request server
// Continuous batching: jobs wait in a FIFO queue; each worker thread owns an engine (its
// own stream) and takes the next job as soon as it is free.  A short job (dense, one batch,
// no OTF, 128x128 tiles, an estimated iteration of at most kFuseJobUs) starts a fused group
// on that worker: the group's jobs iterate together, one launch per phase for all of them
// (as spca_solve_fused), and at every chunk boundary finished jobs leave (their memory
// returns at once) and queued short jobs of the group's precision join while the group's
// estimated iteration stays within kGroupUs and their workspaces fit the HBM budget.  Members
// iterate in lockstep, so the caps keep a short job's iterations short.  With several
// workers, worker 0 is the group lane: it takes only short jobs, from anywhere in the queue,
// and its groups admit them past longer jobs, which the other workers take in order; a
// single worker admits strictly in order.  Longer jobs (and sparse, several-batch and OTF
// jobs) are solved alone by the worker that takes them, beside the groups.  A job that does
// not fit waits for running jobs to return memory; it fails with SPCA_EBUDGET only when
// nothing else is running.  Kernels are deterministic, every job has its own buffers and a
// fused job's tiles, plan and summation order do not depend on its group, so a job's result
// does not depend on what runs beside it.
struct Job {
    uint64_t id = 0;
    uint64_t mat = 0;
    Matrix* M = nullptr;
    spca_settings st{};
    std::vector<float> X0;
    std::vector<float> x;
    spca_stats stats{};
    std::vector<spca_sp_result> per_sp;
    bool want_per_sp = false;
    int status = SPCA_EAGAIN;
    int state = SPCA_JOB_QUEUED;
    std::atomic<int> cancel{0};
    std::string err;
    std::chrono::steady_clock::time_point t_submit, t_start, t_end;
    int np = 0, mp = 0, bp = 0, splits = 0;
    bool fusable = false;  // a short job: may run in a fused group
    double est_us = 0;     // estimated time of one iteration (product flops at a nominal rate)
    int nsp = 0;           // starting points after the settings check
    // the result buffers, sized when the job starts (a long queue holds no result memory)
    void alloc_results() {
        x.assign(M->n, 0.f);
        if (want_per_sp) per_sp.resize(nsp);
    }
};

}  // namespace

// Fused groups: jobs whose iteration is estimated at most kFuseJobUs, groups whose members'
// iterations add up to at most kGroupUs (estimates: 4 m n B product flops per iteration at a
// nominal rate per precision).
constexpr double kFuseJobUs = 200.0;
constexpr double kGroupUs = 1000.0;
inline double est_iteration_us(int precision, int mp, int np, int bp) {
    const double tflops = precision == 0 ? 140.0 : precision == 1 ? 350.0 : 700.0;
    return 4.0 * mp * np * bp / (tflops * 1e6);
}

struct spca_server {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job*> queue;
    std::map<uint64_t, Job*> jobs;
    std::vector<std::thread> workers;
    std::vector<Engine> engines;
    bool stop = false;
    uint64_t next_id = 1;
    int running = 0;
    int group_cap = 64;  // most jobs in one fused group; 0: every job is solved alone
    spca_server_info info{};
};

namespace {

void matrix_unref(spca_ctx ctx, uint64_t h) {
    std::lock_guard<std::mutex> g(ctx->mu);
    auto itm = ctx->mats.find(h);
    if (itm == ctx->mats.end()) return;
    if (--itm->second->refs <= 0) {
        free_matrix(ctx, itm->second);
        ctx->mats.erase(itm);
    }
}

// The end of a job: its matrix reference is dropped and waiters are woken.
void retire_job(spca_ctx ctx, spca_server* S, Job* J, int rc) {
    J->t_end = std::chrono::steady_clock::now();
    matrix_unref(ctx, J->mat);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        J->status = rc;
        J->state = rc == SPCA_ECANCELLED ? SPCA_JOB_CANCELLED : SPCA_JOB_DONE;
        S->running--;
    }
    S->cv.notify_all();
}

// Jobs a group took but cannot admit go back to the front of the queue, in their order.
void requeue(spca_server* S, const std::vector<Job*>& back) {
    {
        std::lock_guard<std::mutex> lk(S->mu);
        for (auto q = back.rbegin(); q != back.rend(); ++q) {
            (*q)->state = SPCA_JOB_QUEUED;
            S->queue.push_front(*q);
            S->running--;
        }
    }
    S->cv.notify_all();
}

// One member of a running fused group.
struct Member {
    Job* J = nullptr;
    FusedRun r;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // before its first chunk; after the latest chunk
};

// A fused group started by `first` on this worker's stream.  Each chunk is min(check_every)
// iterations (at most 64) of every member, launched as one fused iteration per step; then the
// host reads every member's done flag, retires finished and cancelled members and admits
// queued jobs, which start at iteration 0 beside the others.
void run_group(spca_ctx ctx, spca_server* S, Engine& E, Job* first, bool lane) {
    hipStream_t s = E.stream;
    const int precision = first->st.precision;
    int cap;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        cap = std::max(1, S->group_cap);
        S->info.groups++;
    }
    free_workspace(ctx, E);  // members hold their own workspaces
    FusedJob* d_jobs = nullptr;
    int* d_done = nullptr;
    int* h_flags = nullptr;
    uint64_t tb = 0;
    std::vector<Member*> mem;
    std::vector<FusedJob> tab;
    FusedTotals T;
    bool changed = true;
    auto drop = [&](Member* m, int rc) {
        free_ws(ctx, m->r.ws, s);
        if (m->ev0) (void)hipEventDestroy(m->ev0);
        if (m->ev1) (void)hipEventDestroy(m->ev1);
        retire_job(ctx, S, m->J, rc);
        delete m;
    };
    int rc = dev_alloc(ctx, &d_jobs, (size_t)cap, &tb);
    if (!rc) rc = dev_alloc(ctx, &d_done, 4, &tb);
    if (!rc && hipHostMalloc((void**)&h_flags, sizeof(int) * cap, hipHostMallocDefault) != hipSuccess) {
        h_flags = nullptr;
        rc = fail(ctx, SPCA_EDEVICE, "pinned flags allocation failed");
    }
    // admits J (taken off the queue, counted as running); false: no memory for it now while
    // other members run, so the caller puts it back
    auto admit = [&](Job* J) -> bool {
        J->t_start = std::chrono::steady_clock::now();
        if (J->cancel.load()) {
            retire_job(ctx, S, J, SPCA_ECANCELLED);
            return true;
        }
        Member* m = new Member();
        m->J = J;
        J->alloc_results();
        const float* X0 = J->X0.empty() ? nullptr : J->X0.data();
        int e;
        for (;;) {
            e = fused_member_init(ctx, E, J->M, &J->st, X0, J->x.data(), precision, false, m->r);
            if (e != SPCA_EBUDGET) break;
            free_ws(ctx, m->r.ws, s);
            if (!mem.empty()) {
                delete m;
                return false;
            }
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->running <= 1 || J->cancel.load() || S->stop) break;
            S->cv.wait_for(lk, std::chrono::milliseconds(50));
        }
        if (e == SPCA_OK && (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess ||
                             hipEventRecord(m->ev0, s) != hipSuccess))
            e = fail(ctx, SPCA_EDEVICE, "event creation failed");
        if (e != SPCA_OK) {
            J->err = tl_err;
            drop(m, e);
            return true;
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (!mem.empty()) S->info.joined_running++;
            S->info.max_members = std::max(S->info.max_members, (int)mem.size() + 1);
        }
        mem.push_back(m);
        changed = true;
        return true;
    };
    if (rc) {
        first->err = tl_err;
        retire_job(ctx, S, first, rc);
    } else {
        admit(first);
    }
    while (!rc && !mem.empty()) {
        // ---- admission at the chunk boundary.  The group lane takes short jobs of its precision
        // from anywhere in the queue (other workers take the rest in order); a single worker stops
        // at the first job the group cannot take, so the group drains and that job runs next.
        std::vector<Job*> take;
        {
            double est = 0;
            for (Member* m : mem) est += m->J->est_us;
            std::lock_guard<std::mutex> lk(S->mu);
            for (auto q = S->queue.begin(); q != S->queue.end() && (int)(mem.size() + take.size()) < cap;) {
                Job* Q = *q;
                if (Q->fusable && Q->st.precision == precision && est + Q->est_us <= kGroupUs) {
                    q = S->queue.erase(q);
                    Q->state = SPCA_JOB_RUNNING;
                    S->running++;
                    take.push_back(Q);
                    est += Q->est_us;
                } else if (lane) {
                    ++q;
                } else {
                    break;
                }
            }
        }
        for (size_t t = 0; t < take.size(); t++)
            if (!admit(take[t])) {
                requeue(S, std::vector<Job*>(take.begin() + t, take.end()));
                break;
            }
        if (mem.empty()) break;
        if (changed) {
            tab.resize(mem.size());
            T = FusedTotals();
            for (size_t k = 0; k < mem.size(); k++) fused_entry(mem[k]->r, precision, T, tab[k]);
            if (hipMemcpyAsync(d_jobs, tab.data(), sizeof(FusedJob) * tab.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemsetAsync(d_done, 0, sizeof(int) * 4, s) != hipSuccess) {
                rc = fail(ctx, SPCA_EDEVICE, "fused job table upload failed");
                break;
            }
            changed = false;
        }
        // ---- one chunk
        int chunk = 64;
        for (Member* m : mem) chunk = std::min(chunk, std::max(1, m->r.st.check_every));
        hipError_t e = hipSuccess;
        for (int i = 0; i < chunk && e == hipSuccess; i++)
            e = launch_fused_iteration(precision, d_jobs, (int)mem.size(), T.t1, T.t2, T.c0, T.t3, T.t4, T.t5, T.any_fix8,
                                       d_done, s);
        for (size_t k = 0; k < mem.size() && e == hipSuccess; k++) {
            e = hipEventRecord(mem[k]->ev1, s);
            if (e == hipSuccess) e = hipMemcpyAsync(&h_flags[k], tab[k].sp.done, sizeof(int), hipMemcpyDeviceToHost, s);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            rc = fail(ctx, SPCA_EDEVICE, std::string("fused group iteration: ") + hipGetErrorString(e));
            break;
        }
        // ---- retire finished and cancelled members
        std::vector<Member*> keep;
        for (size_t k = 0; k < mem.size(); k++) {
            Member* m = mem[k];
            Job* J = m->J;
            if (J->cancel.load()) {
                drop(m, SPCA_ECANCELLED);
                changed = true;
            } else if (h_flags[k]) {
                float t = 0;
                (void)hipEventElapsedTime(&t, m->ev0, m->ev1);
                int r2 = fused_member_finish(ctx, E, m->r, J->x.data(), &J->stats,
                                             J->want_per_sp ? J->per_sp.data() : nullptr,
                                             [t](long) { return (double)t; }, J->t_start);
                if (r2 != SPCA_OK) J->err = tl_err;
                drop(m, r2);
                changed = true;
            } else {
                keep.push_back(m);
            }
        }
        mem.swap(keep);
    }
    if (rc)  // a device failure ends every member
        for (Member* m : mem) {
            m->J->err = tl_err;
            drop(m, rc);
        }
    (void)hipStreamSynchronize(s);
    dev_free(ctx, d_jobs, sizeof(FusedJob) * (uint64_t)cap);
    dev_free(ctx, d_done, sizeof(int) * 4);
    if (h_flags) (void)hipHostFree(h_flags);
}

void worker_main(spca_ctx ctx, spca_server* S, int w) {
    (void)hipSetDevice(ctx->device);
    Engine& E = S->engines[w];
    for (;;) {
        Job* J = nullptr;
        bool group = false, lane = false;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            // the job this worker takes now: the queue's front, or for the group lane its first
            // short job
            auto pick = [&]() -> std::deque<Job*>::iterator {
                lane = w == 0 && S->group_cap > 0 && S->workers.size() > 1;
                if (!lane) return S->queue.begin();
                return std::find_if(S->queue.begin(), S->queue.end(), [](const Job* Q) { return Q->fusable; });
            };
            S->cv.wait(lk, [&] { return S->stop || pick() != S->queue.end(); });
            auto it = pick();
            if (it == S->queue.end()) break;  // stop requested and nothing this worker takes
            J = *it;
            S->queue.erase(it);
            J->state = SPCA_JOB_RUNNING;
            S->running++;
            group = J->fusable && S->group_cap > 0;
            if (!group) S->info.solo_jobs++;
        }
        if (group) {
            run_group(ctx, S, E, J, lane);
            continue;
        }
        // admission by HBM footprint
        int rc;
        for (;;) {
            std::string err;
            rc = ensure_workspace(ctx, E, J->np, J->mp, J->bp, J->splits, &err);
            if (rc != SPCA_EBUDGET) break;
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->running <= 1 || J->cancel.load() || S->stop) {
                J->err = err;
                break;
            }
            S->cv.wait_for(lk, std::chrono::milliseconds(50));
        }
        J->t_start = std::chrono::steady_clock::now();
        if (rc == SPCA_OK && J->cancel.load()) rc = SPCA_ECANCELLED;
        if (rc == SPCA_OK) {
            E.cancel = &J->cancel;
            J->alloc_results();
            const float* X0 = J->X0.empty() ? nullptr : J->X0.data();
            rc = solve_and_resolve(ctx, E, J->M, &J->st, X0, J->x.data(), &J->stats,
                                   J->want_per_sp ? J->per_sp.data() : nullptr, nullptr);
            E.cancel = nullptr;
            if (rc != SPCA_OK) J->err = tl_err;
        }
        // keep the workspace only if the next queued job has the same shape
        bool keep = false;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (!S->queue.empty()) {
                const Job* Nx = S->queue.front();
                keep = Nx->np == E.ws.np && Nx->mp == E.ws.mp && Nx->bp == E.ws.bp && Nx->splits <= E.ws.splits;
            }
        }
        if (!keep) free_workspace(ctx, E);
        retire_job(ctx, S, J, rc);
    }
    free_workspace(ctx, E);
}

void server_stop(spca_ctx ctx) {
    spca_server* S = ctx->server;
    if (!S) return;
    std::vector<uint64_t> unref;  // matrices held by queued jobs, which no worker will release
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        for (Job* J : S->queue) {
            J->status = SPCA_ECANCELLED;
            J->state = SPCA_JOB_CANCELLED;
            J->cancel.store(1);
            if (J->M) unref.push_back(J->mat);
            J->M = nullptr;
        }
        S->queue.clear();
        for (auto& kv : S->jobs) kv.second->cancel.store(1);
    }
    for (uint64_t h : unref) matrix_unref(ctx, h);
    S->cv.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& E : S->engines) engine_destroy(ctx, E);
    for (auto& kv : S->jobs) delete kv.second;
    delete S;
    ctx->server = nullptr;
}

}  // namespace

extern "C" {


const char* spca_version(void) { return kVersion; }

void spca_settings_default(spca_settings* st) {
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->formulation = L0_constrained_L2_PCA;
    st->constraint = 10;
    st->penalty = 0;
    st->tolerance = 0.01;
    st->max_iter = 20;
    st->total_sp = 64;
    st->batch = 64;
    st->use_otf = 0;
    st->termination = SPCA_TERM_BATCH_MAX;
    st->precision = SPCA_PREC_FP32;
    st->start = SPCA_START_EXPLICIT;
    st->check_every = 1;
    st->resolve_iters = 0;
    st->resolve_tol = 1e-12;
    st->seed = 0;
    st->sparse_products = 1;
    st->exact_sign_products = 1;
    st->exact_ties = 1;
}

int spca_settings_check(spca_settings* st, unsigned n) {
    if (!st) return 0;
    if (st->constraint > n) st->constraint = n;
    if (st->batch == 0) st->batch = st->total_sp;
    int tb = st->total_sp / st->batch;
    if (tb * st->batch < st->total_sp) tb++;
    st->total_sp = tb * st->batch;
    return tb;
}

int spca_product_forms(const spca_settings* st, int m, int n, int* ax_form, int* atz_form) {
    if (!st || m <= 0 || n <= 0 || m > (1 << 30)) return SPCA_EINVAL;
    if (!valid_formulation(st->formulation) || st->precision < 0 || st->precision > 2) return SPCA_EINVAL;
    if (n >= kMaxCols) return SPCA_EINVAL;
    if (ax_form) *ax_form = st->precision;
    if (atz_form) *atz_form = uses_sign_products(*st, round_up(m, kPadMN)) ? SPCA_FORM_FIX8 : st->precision;
    return SPCA_OK;
}

int spca_is_constrained(int f) { return f >= 0 && f < 4; }
int spca_is_l1_constrained(int f) { return !(f == L0_constrained_L1_PCA || f == L0_constrained_L2_PCA); }
int spca_is_l1_penalized(int f) { return !(f == L0_penalized_L1_PCA || f == L0_penalized_L2_PCA); }
int spca_is_l1_variance(int f) { return f == 1 || f == 3 || f == 5 || f == 7; }

int spca_ctx_create(int device, uint64_t hbm_budget_bytes, spca_ctx* out) {
    if (!out) return SPCA_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return SPCA_EDEVICE;
    if (device < 0 || device >= count) return SPCA_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return SPCA_EDEVICE;
    spca_ctx c = new spca_ctx_s();
    c->device = device;
    int rc = engine_init(c->eng);
    if (rc) {
        engine_destroy(c, c->eng);
        delete c;
        return rc;
    }
    c->stream = c->eng.stream;
    // the fixed scratch's device address, looked up once per device (spca_solve_noalloc fails
    // with SPCA_EDEVICE if the lookup did not succeed; nothing else depends on it)
    (void)scratch_setup(device);
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(0.9 * (double)fr);
    *out = c;
    return SPCA_OK;
}

int spca_ctx_destroy(spca_ctx ctx) {
    if (!ctx) return SPCA_EINVAL;
    server_stop(ctx);
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        (void)hipSetDevice(ctx->device);
        for (auto& kv : ctx->mats) free_matrix(ctx, kv.second);
        ctx->mats.clear();
        engine_destroy(ctx, ctx->eng);
        release_cache(ctx);
        for (auto& kv : ctx->live) (void)hipFree(kv.first);  // none expected
        ctx->live.clear();
    }
    delete ctx;
    return SPCA_OK;
}

const char* spca_last_error(spca_ctx ctx) {
    static thread_local std::string copy;
    if (!ctx) return "null context";
    std::lock_guard<std::mutex> g(ctx->err_mu);
    copy = ctx->err;
    return copy.c_str();
}

int spca_matrix_upload(spca_ctx ctx, const float* A, int m, int n, int ld, unsigned storage, uint64_t key,
                       spca_matrix* out) {
    if (!ctx) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    return make_matrix(ctx, A, false, m, n, ld, storage, key, out);
}

int spca_matrix_from_device(spca_ctx ctx, const float* d_A, int m, int n, int ld, unsigned storage, uint64_t key,
                            spca_matrix* out) {
    if (!ctx) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    return make_matrix(ctx, d_A, true, m, n, ld, storage, key, out);
}

int spca_matrix_generate(spca_ctx ctx, int m, int n, uint64_t seed, int nspikes, const double* theta,
                         int spike_nnz, unsigned storage, uint64_t key, float* loadings, spca_matrix* out) {
    if (!ctx) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    return make_generated(ctx, m, n, seed, nspikes, theta, spike_nnz, storage, key, loadings, out);
}

int spca_matrix_download(spca_ctx ctx, spca_matrix mat, float* A, int ld) {
    if (!ctx || !A) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto it = ctx->mats.find(mat);
    if (it == ctx->mats.end()) return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
    Matrix* M = it->second;
    if (M->sparse) return fail(ctx, SPCA_EINVAL, "a sparse matrix has no dense copy to download");
    if (ld < M->m) return fail(ctx, SPCA_EINVAL, "ld < m");
    (void)hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipMemcpy2DAsync(A, sizeof(float) * ld, M->a32_cm, sizeof(float) * M->mp, sizeof(float) * M->m,
                                  M->n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SPCA_OK;
}

int spca_matrix_upload_csc(spca_ctx ctx, const int* col_ptr, const int* row_idx, const float* vals, int m, int n,
                           const float* col_means, const float* row_means, uint64_t key, spca_matrix* out) {
    if (!ctx) return SPCA_EINVAL;
    if (!col_ptr || !out || m <= 0 || n <= 0) return fail(ctx, SPCA_EINVAL, "bad sparse matrix arguments");
    const int nnz = col_ptr[n];
    if (col_ptr[0] != 0 || nnz < 0 || (nnz > 0 && (!row_idx || !vals)))
        return fail(ctx, SPCA_EINVAL, "col_ptr must start at 0 and end at nnz");
    for (int j = 0; j < n; j++)
        if (col_ptr[j + 1] < col_ptr[j]) return fail(ctx, SPCA_EINVAL, "col_ptr must be non-decreasing");
    for (int t = 0; t < nnz; t++)
        if (row_idx[t] < 0 || row_idx[t] >= m) return fail(ctx, SPCA_EINVAL, "row index out of range");
    // CSR copy for A*X: counting sort by row, columns in order within a row
    std::vector<int> rp((size_t)m + 1, 0), ci((size_t)std::max(nnz, 1));
    std::vector<float> rv((size_t)std::max(nnz, 1));
    for (int t = 0; t < nnz; t++) rp[(size_t)row_idx[t] + 1]++;
    for (int i = 0; i < m; i++) rp[(size_t)i + 1] += rp[i];
    {
        std::vector<int> fill(rp.begin(), rp.end() - 1);
        for (int j = 0; j < n; j++)
            for (int t = col_ptr[j]; t < col_ptr[j + 1]; t++) {
                const int d = fill[row_idx[t]]++;
                ci[d] = j;
                rv[d] = vals[t];
            }
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    Matrix* M = new Matrix();
    M->sparse = true;
    M->m = m;
    M->n = n;
    M->mp = round_up(m, kPadMN);
    M->np = round_up(n, kPadMN);
    M->nnz = nnz;
    M->key = key;
    M->refs = 1;
    M->storage = SPCA_STORE_FP32;
    const size_t cnt = (size_t)std::max(nnz, 1);
    int rc;
    if ((rc = dev_alloc(ctx, &M->cptr, (size_t)n + 1, &M->bytes)) || (rc = dev_alloc(ctx, &M->ridx, cnt, &M->bytes)) ||
        (rc = dev_alloc(ctx, &M->cval, cnt, &M->bytes)) || (rc = dev_alloc(ctx, &M->rptr, (size_t)m + 1, &M->bytes)) ||
        (rc = dev_alloc(ctx, &M->cidx, cnt, &M->bytes)) || (rc = dev_alloc(ctx, &M->rval, cnt, &M->bytes)) ||
        (col_means && (rc = dev_alloc(ctx, &M->mu, (size_t)n, &M->bytes))) ||
        (row_means && (rc = dev_alloc(ctx, &M->rmu, (size_t)m, &M->bytes)))) {
        free_matrix(ctx, M);
        return rc;
    }
    hipError_t e = hipMemcpy(M->cptr, col_ptr, sizeof(int) * ((size_t)n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz > 0) e = hipMemcpy(M->ridx, row_idx, sizeof(int) * nnz, hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz > 0) e = hipMemcpy(M->cval, vals, sizeof(float) * nnz, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(M->rptr, rp.data(), sizeof(int) * ((size_t)m + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz > 0) e = hipMemcpy(M->cidx, ci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz > 0) e = hipMemcpy(M->rval, rv.data(), sizeof(float) * nnz, hipMemcpyHostToDevice);
    if (e == hipSuccess && col_means) e = hipMemcpy(M->mu, col_means, sizeof(float) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess && row_means) e = hipMemcpy(M->rmu, row_means, sizeof(float) * m, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_matrix(ctx, M);
        return fail(ctx, SPCA_EDEVICE, std::string("sparse matrix upload: ") + hipGetErrorString(e));
    }
    const uint64_t h = ctx->next_handle++;
    ctx->mats[h] = M;
    *out = h;
    return SPCA_OK;
}

int spca_matrix_add_storage(spca_ctx ctx, spca_matrix mat, unsigned storage) {
    if (!ctx) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto itm = ctx->mats.find(mat);
    if (itm == ctx->mats.end()) return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
    (void)hipSetDevice(ctx->device);
    if (itm->second->sparse) return fail(ctx, SPCA_EINVAL, "a sparse matrix takes no dense storage forms");
    return add_storage(ctx, itm->second, storage);
}

int spca_matrix_release(spca_ctx ctx, spca_matrix mat) {
    if (!ctx) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto itm = ctx->mats.find(mat);
    if (itm == ctx->mats.end()) return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
    if (--itm->second->refs <= 0) {
        (void)hipSetDevice(ctx->device);
        free_matrix(ctx, itm->second);
        ctx->mats.erase(itm);
    }
    return SPCA_OK;
}

int spca_matrix_bytes(spca_ctx ctx, spca_matrix mat, uint64_t* bytes) {
    if (!ctx || !bytes) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto itm = ctx->mats.find(mat);
    if (itm == ctx->mats.end()) return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
    *bytes = itm->second->bytes;
    return SPCA_OK;
}

int spca_solve(spca_ctx ctx, spca_matrix mat, const spca_settings* settings, const float* X0, float* x_out,
               spca_stats* stats, spca_sp_result* per_sp, float* X_final) {
    if (!ctx || !settings) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto itm = ctx->mats.find(mat);
    if (itm == ctx->mats.end()) return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
    (void)hipSetDevice(ctx->device);
    if (stats) std::memset(stats, 0, sizeof(*stats));
    return solve_and_resolve(ctx, ctx->eng, itm->second, settings, X0, x_out, stats, per_sp, X_final);
}


int spca_solve_noalloc(spca_ctx ctx, spca_matrix mat, const spca_settings* settings, const float* X0, float* x_out,
                       spca_stats* stats, spca_sp_result* per_sp, float* X_final) {
    if (!ctx || !settings) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto itm = ctx->mats.find(mat);
    if (itm == ctx->mats.end()) return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
    (void)hipSetDevice(ctx->device);
    void* base = scratch_ptr(ctx->device);
    if (!base) return fail(ctx, SPCA_EDEVICE, "the fixed scratch of this device has no address (lookup failed)");
    if (settings->resolve_iters > 0)
        return fail(ctx, SPCA_EINVAL, "the allocation-free solve has no FP64 re-solve (resolve_iters must be 0)");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    FixedScratch fx{(unsigned char*)base, kScratchBytes, 0};
    return solve_impl(ctx, ctx->eng, itm->second, settings, X0, x_out, stats, per_sp, X_final, nullptr, &fx);
}

int spca_solve_components(spca_ctx ctx, spca_matrix mat, const spca_settings* settings, const float* X0, int k,
                          float* X_out, spca_stats* stats) {
    if (!ctx) return SPCA_EINVAL;
    if (!settings || !X_out || k < 1) return fail(ctx, SPCA_EINVAL, "settings, X_out and k >= 1 are required");
    std::lock_guard<std::mutex> g(ctx->mu);
    auto itm = ctx->mats.find(mat);
    if (itm == ctx->mats.end()) return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
    Matrix* M = itm->second;
    const int n = M->n;
    if (k > n) return fail(ctx, SPCA_EINVAL, "k must be at most n");
    (void)hipSetDevice(ctx->device);
    if (stats) std::memset(stats, 0, sizeof(*stats) * (size_t)k);
    // components 0 .. c-1 as a sparse list: nonzeros idx/val, component c at off[c] .. off[c+1]
    int* d_idx = nullptr;
    float* d_val = nullptr;
    int* d_off = nullptr;
    uint64_t db = 0;
    const size_t cap = (size_t)n * (size_t)(k - 1) + 1;
    int rc = SPCA_OK;
    if ((rc = dev_alloc(ctx, &d_idx, cap, &db)) || (rc = dev_alloc(ctx, &d_val, cap, &db)) ||
        (rc = dev_alloc(ctx, &d_off, (size_t)k + 1, &db))) {
        dev_free(ctx, d_idx, 0);
        dev_free(ctx, d_val, 0);
        dev_free(ctx, d_off, 0);
        return rc;
    }
    std::vector<int> h_idx, h_off(1, 0);
    std::vector<float> h_val;
    for (int c = 0; c < k && rc == SPCA_OK; c++) {
        Deflation d;
        if (c > 0) {
            const float* p = X_out + (size_t)(c - 1) * n;
            for (int i = 0; i < n; i++)
                if (p[i] != 0.f) {
                    h_idx.push_back(i);
                    h_val.push_back(p[i]);
                }
            h_off.push_back((int)h_idx.size());
            if (hipMemcpy(d_idx, h_idx.data(), sizeof(int) * h_idx.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_val, h_val.data(), sizeof(float) * h_val.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_off, h_off.data(), sizeof(int) * h_off.size(), hipMemcpyHostToDevice) != hipSuccess) {
                rc = fail(ctx, SPCA_EDEVICE, "deflation upload failed");
                break;
            }
            d.idx = d_idx;
            d.val = d_val;
            d.off = d_off;
            d.ncomp = c;
        }
        rc = solve_and_resolve(ctx, ctx->eng, M, settings, X0, X_out + (size_t)c * n, stats ? stats + c : nullptr,
                               nullptr, nullptr, c > 0 ? &d : nullptr);
    }
    (void)hipStreamSynchronize(ctx->eng.stream);
    dev_free(ctx, d_idx, 0);
    dev_free(ctx, d_val, 0);
    dev_free(ctx, d_off, 0);
    return rc;
}

int spca_ctx_memory(spca_ctx ctx, uint64_t* used, uint64_t* cached, uint64_t* budget) {
    if (!ctx) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->alloc_mu);
    if (used) *used = ctx->used;
    if (cached) *cached = ctx->cached;
    if (budget) *budget = ctx->budget;
    return SPCA_OK;
}

int spca_set_tuning(spca_ctx ctx, int mode, const char* cache_path) {
    if (!ctx || mode < 0 || mode > 1) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->plan_mu);
    ctx->tuning = mode;
    ctx->tune_path = cache_path ? cache_path : "";
    if (!ctx->tune_path.empty()) load_plans(ctx);
    return SPCA_OK;
}

int spca_get_plan(spca_ctx ctx, int precision, int m, int n, int batch, int fused, int* splits, int* big,
                  double* tuned_ms) {
    if (!ctx || m <= 0 || n <= 0 || batch <= 0 || precision < 0 || precision > 2) return SPCA_EINVAL;
    const Plan p = peek_plan(ctx, precision, round_up(m, kPadMN), round_up(n, kPadMN), round_up(batch, kPadB), !fused);
    if (splits) *splits = p.splits;
    if (big) *big = p.big;
    if (tuned_ms) *tuned_ms = p.ms;
    return SPCA_OK;
}

int spca_solve_fused(spca_ctx ctx, spca_fused_job* jobs, int njobs, unsigned flags) {
    if (!ctx) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    return fused_impl(ctx, ctx->eng, jobs, njobs, flags);
}

int spca_server_start(spca_ctx ctx, int workers) {
    if (!ctx || workers < 1 || workers > 16) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->server) return SPCA_OK;
    (void)hipSetDevice(ctx->device);
    spca_server* S = new spca_server();
    S->engines.resize(workers);
    for (auto& E : S->engines) {
        int rc = engine_init(E);
        if (rc) {
            for (auto& E2 : S->engines) engine_destroy(ctx, E2);
            delete S;
            return fail(ctx, rc, "engine creation failed");
        }
    }
    ctx->server = S;
    for (int w = 0; w < workers; w++) S->workers.emplace_back(worker_main, ctx, S, w);
    return SPCA_OK;
}

int spca_server_stop(spca_ctx ctx) {
    if (!ctx) return SPCA_EINVAL;
    server_stop(ctx);
    return SPCA_OK;
}

int spca_server_set_batching(spca_ctx ctx, int max_group_jobs) {
    if (!ctx || !ctx->server || max_group_jobs < 0 || max_group_jobs > 1024) return SPCA_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->server->mu);
    ctx->server->group_cap = max_group_jobs;
    return SPCA_OK;
}

int spca_server_get_info(spca_ctx ctx, spca_server_info* info) {
    if (!ctx || !ctx->server || !info) return SPCA_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->server->mu);
    *info = ctx->server->info;
    return SPCA_OK;
}

int spca_submit(spca_ctx ctx, spca_matrix mat, const spca_settings* settings, const float* X0, int want_per_sp,
                spca_job* out) {
    if (!ctx || !settings || !out) return SPCA_EINVAL;
    spca_server* S = ctx->server;
    if (!S) return fail(ctx, SPCA_EINVAL, "no request server: call spca_server_start first");
    Job* J = new Job();
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto itm = ctx->mats.find(mat);
        if (itm == ctx->mats.end()) {
            delete J;
            return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
        }
        Matrix* M = itm->second;
        J->st = *settings;
        if (!valid_formulation(J->st.formulation) || J->st.precision < 0 || J->st.precision > 2 || J->st.max_iter < 1 ||
            J->st.total_sp < 1 || J->st.batch < 0 || (J->st.start == SPCA_START_EXPLICIT && !X0)) {
            delete J;
            return fail(ctx, SPCA_EINVAL, "bad job settings");
        }
        // every storage form the solve reads is built here, under ctx->mu: workers solving
        // the same matrix side by side must find it complete (they never add storage)
        unsigned need = J->st.precision == 0 ? SPCA_STORE_FP32
                        : J->st.precision == 1 ? SPCA_STORE_BF16X3 : SPCA_STORE_FP16;
        if (!M->sparse && uses_sign_products(J->st, M->mp)) need |= SPCA_STORE_FIX8 | SPCA_STORE_FP32;
        (void)hipSetDevice(ctx->device);
        int rc = M->sparse ? SPCA_OK : add_storage(ctx, M, need);
        if (rc) {
            delete J;
            return rc;
        }
        spca_settings chk = J->st;
        const int nbatches = spca_settings_check(&chk, (unsigned)M->n);
        J->fusable = !M->sparse && nbatches == 1 && !chk.use_otf;
        J->M = M;
        J->mat = mat;
        J->np = M->np;
        J->mp = M->mp;
        J->bp = round_up(chk.batch, kPadB);
        const Plan plan = peek_plan(ctx, J->st.precision, M->mp, M->np, J->bp, true);
        J->splits = plan.splits;
        // a fused group runs 128x128 tiles: a shape whose solo plan takes the large tiles is
        // solved alone, so every job's result is its solo result
        J->est_us = est_iteration_us(J->st.precision, J->mp, J->np, J->bp);
        J->fusable = J->fusable && !plan.big && J->est_us <= kFuseJobUs;
        uint64_t need_bytes = workspace_bytes(J->np, J->mp, J->bp, J->splits);
        const int ucap = (M->sparse || chk.use_otf) ? 0 : union_cap(chk, M->n);
        if (ucap > 0) {  // the union A*X buffers (flags, list, Au, Xu)
            const uint64_t ldu = (uint64_t)round_up(ucap, kGemmBK * J->splits);
            need_bytes += ((uint64_t)2 * J->np + 4 + ((uint64_t)J->mp + J->bp) * ldu) * 4;
        }
        if (need_bytes + M->bytes > ctx->budget) {
            delete J;
            return fail(ctx, SPCA_EBUDGET, "job workspace exceeds the HBM budget");
        }
        if (J->st.start == SPCA_START_EXPLICIT) J->X0.assign(X0, X0 + (size_t)M->n * chk.total_sp);
        J->nsp = chk.total_sp;  // result buffers are allocated when the job starts
        J->want_per_sp = want_per_sp != 0;
        M->refs++;
    }
    J->t_submit = std::chrono::steady_clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        J->id = S->next_id++;
        S->jobs[J->id] = J;
        S->queue.push_back(J);
        *out = J->id;
    }
    S->cv.notify_all();
    return SPCA_OK;
}

int spca_poll(spca_ctx ctx, spca_job job, int* state) {
    if (!ctx || !ctx->server || !state) return SPCA_EINVAL;
    spca_server* S = ctx->server;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return SPCA_ENOTFOUND;
    *state = it->second->state;
    return SPCA_OK;
}

int spca_cancel(spca_ctx ctx, spca_job job) {
    if (!ctx || !ctx->server) return SPCA_EINVAL;
    spca_server* S = ctx->server;
    // A job taken off the queue here never reaches a worker, so its matrix reference is
    // dropped here; a running job's worker drops its own.  J is not touched after the lock is
    // released: a concurrent spca_wait may delete it as soon as it reads CANCELLED.
    uint64_t unref = 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->jobs.find(job);
        if (it == S->jobs.end()) return SPCA_ENOTFOUND;
        Job* J = it->second;
        J->cancel.store(1);
        for (auto q = S->queue.begin(); q != S->queue.end(); ++q)
            if (*q == J) {
                S->queue.erase(q);
                J->status = SPCA_ECANCELLED;
                J->state = SPCA_JOB_CANCELLED;
                if (J->M) unref = J->mat;
                J->M = nullptr;
                break;
            }
    }
    if (unref) matrix_unref(ctx, unref);
    S->cv.notify_all();
    return SPCA_OK;
}

int spca_wait(spca_ctx ctx, spca_job job, float* x_out, spca_stats* stats, spca_sp_result* per_sp,
              spca_job_times* times) {
    if (!ctx || !ctx->server) return SPCA_EINVAL;
    spca_server* S = ctx->server;
    Job* J = nullptr;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->jobs.find(job);
        if (it == S->jobs.end()) return SPCA_ENOTFOUND;
        J = it->second;
        S->cv.wait(lk, [&] { return J->state == SPCA_JOB_DONE || J->state == SPCA_JOB_CANCELLED; });
        S->jobs.erase(it);
    }
    const int rc = J->status;
    if (rc == SPCA_OK) {
        if (x_out && !J->x.empty()) std::memcpy(x_out, J->x.data(), sizeof(float) * J->x.size());
        if (stats) *stats = J->stats;
        if (per_sp && J->want_per_sp) std::memcpy(per_sp, J->per_sp.data(), sizeof(spca_sp_result) * J->per_sp.size());
    } else if (!J->err.empty()) {
        std::lock_guard<std::mutex> g(ctx->err_mu);
        ctx->err = J->err;
    }
    if (times) {
        using ms = std::chrono::duration<double, std::milli>;
        const bool started = J->t_start.time_since_epoch().count() != 0;
        times->queue_ms = started ? ms(J->t_start - J->t_submit).count() : 0.0;
        times->run_ms = started ? ms(J->t_end - J->t_start).count() : 0.0;
        times->latency_ms = started ? ms(J->t_end - J->t_submit).count() : 0.0;
    }
    delete J;
    return rc;
}

int spca_resolve_support(spca_ctx ctx, spca_matrix mat, int formulation, const float* x, int max_iter, double tol,
                         double* y_out, double* ev_out, int* iters_out) {
    if (!ctx) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto itm = ctx->mats.find(mat);
    if (itm == ctx->mats.end()) return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
    (void)hipSetDevice(ctx->device);
    return resolve_impl(ctx, ctx->eng, itm->second, formulation, x, max_iter, tol, y_out, ev_out, iters_out);
}

int spca_dense_solver_device(spca_ctx ctx, unsigned m, unsigned n, const float* d_B, float* h_x,
                             const spca_settings* settings, spca_stats* stats, unsigned LD_M, unsigned LD_N) {
    (void)LD_N;
    if (!ctx || !settings) return SPCA_EINVAL;
    spca_matrix h = 0;
    const unsigned need = settings->precision == 0 ? SPCA_STORE_FP32
                          : settings->precision == 1 ? SPCA_STORE_BF16X3 : SPCA_STORE_FP16;
    int rc = spca_matrix_from_device(ctx, d_B, (int)m, (int)n, (int)LD_M, need, 0, &h);
    if (rc) return rc;
    // the reference GPU solver draws its own starting points
    spca_settings st = *settings;
    if (st.start == SPCA_START_EXPLICIT) st.start = SPCA_START_REFERENCE;
    rc = spca_solve(ctx, h, &st, nullptr, h_x, stats, nullptr, nullptr);
    spca_matrix_release(ctx, h);
    return rc;
}

int spca_dense_solver_host(spca_ctx ctx, const float* B, int ldB, float* x, unsigned m, unsigned n,
                           const spca_settings* settings, spca_stats* stats, double* fval_out) {
    if (!ctx || !settings) return SPCA_EINVAL;
    spca_matrix h = 0;
    const unsigned need = settings->precision == 0 ? SPCA_STORE_FP32
                          : settings->precision == 1 ? SPCA_STORE_BF16X3 : SPCA_STORE_FP16;
    int rc = spca_matrix_upload(ctx, B, (int)m, (int)n, ldB, need, 0, &h);
    if (rc) return rc;
    spca_stats local;
    spca_stats* sp = stats ? stats : &local;
    spca_settings st = *settings;
    if (st.start == SPCA_START_EXPLICIT) st.start = SPCA_START_REFERENCE;
    rc = spca_solve(ctx, h, &st, nullptr, x, sp, nullptr, nullptr);
    if (fval_out) *fval_out = sp->fval;
    spca_matrix_release(ctx, h);
    return rc;
}

}  // extern "C"
