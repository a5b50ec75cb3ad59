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
    float* fstate = nullptr;  // val, fz, err, errmax  (4 x bp floats)
    int* istate = nullptr;    // iters, nnz, active, done (3 x bp + 4 ints)
    double* ev = nullptr;
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

template <class T>
int dev_alloc(spca_ctx ctx, T** p, size_t count, uint64_t* acct, std::string* err = nullptr) {
    const uint64_t b = (uint64_t)count * sizeof(T);
    {
        std::lock_guard<std::mutex> g(ctx->alloc_mu);
        if (ctx->used + b > ctx->budget) {
            const std::string m = "allocation of " + std::to_string(b) + " bytes exceeds the HBM budget";
            if (err) *err = m;
            else {
                tl_err = m;
                std::lock_guard<std::mutex> g2(ctx->err_mu);
                ctx->err = m;
            }
            return SPCA_EBUDGET;
        }
        ctx->used += b;
    }
    hipError_t e = hipMalloc((void**)p, b ? b : 16);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> g(ctx->alloc_mu);
        ctx->used -= b;
        *p = nullptr;
        const std::string m = std::string("hipMalloc: ") + hipGetErrorString(e);
        if (err) *err = m;
        else {
            tl_err = m;
            std::lock_guard<std::mutex> g2(ctx->err_mu);
            ctx->err = m;
        }
        return SPCA_ENOMEM;
    }
    *acct += b;
    return SPCA_OK;
}
template <class T>
void dev_free(spca_ctx ctx, T*& p, uint64_t bytes_of_p) {
    if (p) {
        (void)hipFree(p);
        std::lock_guard<std::mutex> g(ctx->alloc_mu);
        ctx->used -= bytes_of_p;
        p = nullptr;
    }
}

bool valid_formulation(int f) { return f >= 0 && f <= 7; }

void free_matrix(spca_ctx ctx, Matrix* M) {
    (void)hipStreamSynchronize(ctx->stream);
    void* ptrs[] = {M->a32_cm, M->a32_rm, M->hi_cm, M->lo_cm, M->hi_rm, M->lo_rm, M->h_cm, M->h_rm};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    std::lock_guard<std::mutex> g(ctx->alloc_mu);
    ctx->used -= M->bytes;
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
        if ((rc = need_rm())) return rc;
        const float* rm = M->a32_rm ? M->a32_rm : rm_tmp;
        if ((rc = dev_alloc(ctx, &M->hi_cm, cnt, &M->bytes)) || (rc = dev_alloc(ctx, &M->lo_cm, cnt, &M->bytes)) ||
            (rc = dev_alloc(ctx, &M->hi_rm, cnt, &M->bytes)) || (rc = dev_alloc(ctx, &M->lo_rm, cnt, &M->bytes)))
            return rc;
        HIP_TRY(ctx, launch_split_bf16(M->a32_cm, cnt, M->hi_cm, M->lo_cm, st));
        HIP_TRY(ctx, launch_split_bf16(rm, cnt, M->hi_rm, M->lo_rm, st));
        M->storage |= SPCA_STORE_BF16X3;
    }
    if (want & SPCA_STORE_FP16) {
        if ((rc = need_rm())) return rc;
        const float* rm = M->a32_rm ? M->a32_rm : rm_tmp;
        if ((rc = dev_alloc(ctx, &M->h_cm, cnt, &M->bytes)) || (rc = dev_alloc(ctx, &M->h_rm, cnt, &M->bytes)))
            return rc;
        HIP_TRY(ctx, launch_to_fp16(M->a32_cm, cnt, M->h_cm, st));
        HIP_TRY(ctx, launch_to_fp16(rm, cnt, M->h_rm, st));
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

void free_workspace(spca_ctx ctx, Engine& E) {
    Workspace& w = E.ws;
    (void)hipStreamSynchronize(E.stream);
    void* ptrs[] = {w.X, w.Xhi, w.Xlo, w.Zpart, w.Z, w.Zhi, w.Zlo, w.V, w.fstate, w.istate, w.ev};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    {
        std::lock_guard<std::mutex> g(ctx->alloc_mu);
        ctx->used -= w.bytes;
    }
    w = Workspace();
}

// device bytes a solve of this shape holds in its workspace
uint64_t workspace_bytes(int np, int mp, int bp, int splits) {
    const uint64_t xs = (uint64_t)np * bp, zs = (uint64_t)mp * bp;
    return xs * (4 + 2 + 2 + 4) + zs * splits * 4 + zs * (4 + 2 + 2) + (uint64_t)4 * bp * 4 +
           ((uint64_t)3 * bp + 4) * 4 + (uint64_t)bp * 8;
}

int ensure_workspace(spca_ctx ctx, Engine& E, int np, int mp, int bp, int splits, std::string* err = nullptr) {
    Workspace& w = E.ws;
    if (w.X && w.np == np && w.mp == mp && w.bp == bp && w.splits >= splits) return SPCA_OK;
    free_workspace(ctx, E);
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
        (rc = dev_alloc(ctx, &w.fstate, (size_t)4 * bp, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.istate, (size_t)3 * bp + 4, &w.bytes, err)) ||
        (rc = dev_alloc(ctx, &w.ev, (size_t)bp, &w.bytes, err))) {
        free_workspace(ctx, E);
        return rc;
    }
    if (hipMemsetAsync(w.Z, 0, zs * sizeof(float), E.stream) != hipSuccess ||
        hipMemsetAsync(w.Zhi, 0, zs * sizeof(uint16_t), E.stream) != hipSuccess ||
        hipMemsetAsync(w.Zlo, 0, zs * sizeof(uint16_t), E.stream) != hipSuccess) {
        free_workspace(ctx, E);
        return SPCA_EDEVICE;
    }
    return SPCA_OK;
}

int engine_init(Engine& E) {
    if (hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking) != hipSuccess) return SPCA_EDEVICE;
    if (hipHostMalloc((void**)&E.h_done, sizeof(int) * 4, hipHostMallocDefault) != hipSuccess) return SPCA_ENOMEM;
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

SpState state_of(Workspace& w) {
    SpState s;
    s.val = w.fstate;
    s.fz = w.fstate + w.bp;
    s.err = w.fstate + 2 * w.bp;
    s.errmax = w.fstate + 3 * w.bp;
    s.iters = w.istate;
    s.nnz = w.istate + w.bp;
    s.active = w.istate + 2 * w.bp;
    s.done = w.istate + 3 * w.bp;
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

// K-splits of A*X: enough workgroups to fill 256 CUs for a couple of rounds (128x128
// tiles run ~3 per CU, 256x256 tiles one per CU), each split at least 512 deep.
int choose_splits(int precision, int mp, int np, int bp) {
    const bool big = gemm_big_tiles(precision, mp, bp);
    const int tiles = big ? (mp / 256) * (bp / 256) : (mp / kGemmBM) * (bp / kGemmBN);
    const int target = big ? 512 : 2048;
    int splits = 1;
    while (tiles * splits < target && np % (kGemmBK * splits * 2) == 0 && np / (splits * 2) >= 512) splits *= 2;
    return splits;
}

// Kernel time per phase (0: A*X, 1: A^T*Z, 2: epilogue/step/convergence), from
// hipEvent pairs recorded on the solver stream and folded in after each host sync.
struct PhaseTimer {
    Engine* eng;
    size_t used = 0;  // events handed out since the last flush (always pairs)
    double ms[3] = {0, 0, 0};
    std::vector<int> phase;
    hipError_t pair(hipEvent_t* a, hipEvent_t* b, int ph) {
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

int solve_impl(spca_ctx ctx, Engine& E, Matrix* M, const spca_settings* st_in, const float* X0, float* x_out,
               spca_stats* stats, spca_sp_result* per_sp, float* X_final) {
    const auto t_wall0 = std::chrono::steady_clock::now();
    spca_settings st = *st_in;
    if (!valid_formulation(st.formulation)) return fail(ctx, SPCA_EINVAL, "formulation must be 0..7");
    if (st.precision < 0 || st.precision > 2) return fail(ctx, SPCA_EINVAL, "precision must be 0..2");
    if (st.max_iter < 1 || st.total_sp < 1 || st.batch < 0) return fail(ctx, SPCA_EINVAL, "bad iteration/SP counts");
    if (st.start < SPCA_START_EXPLICIT || st.start > SPCA_START_REFERENCE_F64)
        return fail(ctx, SPCA_EINVAL, "start must be a spca_start value");
    if (st.start == SPCA_START_EXPLICIT && !X0) return fail(ctx, SPCA_EINVAL, "X0 is required for explicit starts");
    if (!x_out) return fail(ctx, SPCA_EINVAL, "x_out is required");
    const unsigned need = st.precision == 0 ? SPCA_STORE_FP32 : st.precision == 1 ? SPCA_STORE_BF16X3 : SPCA_STORE_FP16;
    int rc = add_storage(ctx, M, need);
    if (rc) return rc;
    const int n = M->n, m = M->m;
    const int total_batches = spca_settings_check(&st, (unsigned)n);
    if (st.constraint < 1) st.constraint = 1;
    const int batch = st.batch;
    const int bp = round_up(batch, kPadB);
    const int splits = choose_splits(st.precision, M->mp, M->np, bp);
    if ((rc = ensure_workspace(ctx, E, M->np, M->mp, bp, splits))) return rc;
    Workspace& w = E.ws;
    hipStream_t s = E.stream;
    SpState sp = state_of(w);
    const bool constrained = st.formulation < 4;
    const int check_every = st.check_every > 0 ? st.check_every : 1;

    Operand xop{w.X, st.precision ? w.Xhi : nullptr, st.precision == 1 ? w.Xlo : nullptr};
    Operand zop{w.Z, st.precision ? w.Zhi : nullptr, st.precision == 1 ? w.Zlo : nullptr};
    const float* P1_32 = M->a32_rm;
    const uint16_t* P1_hi = st.precision == 1 ? M->hi_rm : M->h_rm;
    const uint16_t* P1_lo = M->lo_rm;
    const float* P2_32 = M->a32_cm;
    const uint16_t* P2_hi = st.precision == 1 ? M->hi_cm : M->h_cm;
    const uint16_t* P2_lo = M->lo_cm;
    const long zsplit = (long)M->mp * bp;

    StepArgs sa;
    sa.formulation = st.formulation;
    sa.precision = st.precision;
    sa.n = n;
    sa.ldv = M->np;
    sa.s = st.constraint;
    sa.gamma = st.penalty;
    sa.batch = batch;

    std::vector<float> hX((size_t)n * batch);
    std::vector<float> hval(batch), herr(batch);
    std::vector<int> hit(batch), hnnz(batch);
    std::vector<double> hev(batch);
    double best = -1;
    int best_sp = -1;
    unsigned total_it = 0;
    unsigned long long sp_its = 0;
    double loop_ms = 0, last_err = 0;
    PhaseTimer pt{&E};
    std::vector<float> x_best(n, 0.f);
    hipEvent_t* loop_ev = E.loop_ev;

    if (st.use_otf) {
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
            std::vector<float> hp((size_t)n * total);
            for (int k = 0; k < total; k++) {
                float* col = hp.data() + (size_t)k * n;
                if (st.start == SPCA_START_EXPLICIT)
                    std::memcpy(col, X0 + (size_t)k * n, sizeof(float) * n);
                else if (st.start == SPCA_START_REFERENCE_F64)
                    reference_starting_point<double>(col, (unsigned)n, st.constraint, constrained, (unsigned)k);
                else
                    reference_starting_point<float>(col, (unsigned)n, st.constraint, constrained, (unsigned)k);
            }
            HIP_TRY(ctx, hipMemsetAsync(pool, 0, sizeof(float) * (size_t)w.np * total, s));
            HIP_TRY(ctx, hipMemcpy2DAsync(pool, sizeof(float) * w.np, hp.data(), sizeof(float) * n, sizeof(float) * n,
                                          total, hipMemcpyHostToDevice, s));
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
            HIP_TRY(ctx, hipMemsetAsync(w.fstate, 0, sizeof(float) * 4 * (size_t)w.bp, s));
            HIP_TRY(ctx, hipMemsetAsync(w.istate, 0, sizeof(int) * (3 * (size_t)w.bp + 4), s));
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
            HIP_TRY(ctx, hipEventRecord(e0, s));
            HIP_TRY(ctx, launch_gemm_tn(st.precision, P1_32, P1_hi, P1_lo, M->np, w.X, w.Xhi, w.Xlo, M->np, w.Zpart,
                                        M->mp, zsplit, M->mp, bp, M->np, splits, sp.done, s));
            HIP_TRY(ctx, hipEventRecord(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 2));
            HIP_TRY(ctx, hipEventRecord(e0, s));
            HIP_TRY(ctx, launch_prep_z(w.Zpart, splits, zsplit, M->mp, M->mp, zop, st.formulation, st.precision, batch,
                                       sp, s));
            HIP_TRY(ctx, hipEventRecord(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 1));
            HIP_TRY(ctx, hipEventRecord(e0, s));
            HIP_TRY(ctx, launch_gemm_tn(st.precision, P2_32, P2_hi, P2_lo, M->mp, w.Z, w.Zhi, w.Zlo, M->mp, w.V,
                                        M->np, 0, M->np, bp, M->mp, 1, sp.done, s));
            HIP_TRY(ctx, hipEventRecord(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 2));
            HIP_TRY(ctx, hipEventRecord(e0, s));
            HIP_TRY(ctx, launch_step(w.V, xop, sa, sp, s));
            HIP_TRY(ctx, launch_otf_converge(sp, o, batch, st.tolerance, st.max_iter, total, s));
            HIP_TRY(ctx, launch_otf_refill(sp, o, xop, st.precision, n, w.np, batch, s));
            HIP_TRY(ctx, hipEventRecord(e1, s));
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
        int done_h[2], gen_h[3];
        float bestv;
        std::vector<float> spv(total);
        std::vector<int> spi(total), spn(total);
        HIP_TRY(ctx, hipMemcpyAsync(done_h, sp.done, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(gen_h, o.generated, sizeof(int) * 3, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(&bestv, o.best_val, sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(spv.data(), o.sp_val, sizeof(float) * total, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(spi.data(), o.sp_iters, sizeof(int) * total, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(spn.data(), o.sp_nnz, sizeof(int) * total, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(x_best.data(), xbest, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        double evb = 0;
        HIP_TRY(ctx, launch_ev_fp64(M->a32_cm, M->mp, m, n, xbest, w.np, 1, spca_is_l1_variance(st.formulation), w.ev, s));
        HIP_TRY(ctx, hipMemcpyAsync(&evb, w.ev, sizeof(double), hipMemcpyDeviceToHost, s));
        float emax;
        HIP_TRY(ctx, hipMemcpyAsync(&emax, sp.errmax, sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        release();
        total_it = (unsigned)done_h[1];
        best = bestv;
        best_sp = gen_h[2];
        last_err = emax;
        for (int k = 0; k < total; k++) sp_its += (unsigned long long)spi[k];
        if (per_sp)
            for (int k = 0; k < total; k++) {
                per_sp[k].value = spv[k];
                per_sp[k].iters = spi[k];
                per_sp[k].nnz = spn[k];
                per_sp[k].converged = spi[k] < st.max_iter;
                per_sp[k].ev_fp64 = (k == best_sp) ? evb : std::nan("");
            }
        if (stats) stats->ev_fp64 = evb;
        (void)X_final;  // slots are reused: final iterates per SP are not kept in OTF mode
    } else
    for (int bt = 0; bt < total_batches; bt++) {
        const int shift = bt * batch;
        // ---- starting points
        for (int jj = 0; jj < batch; jj++) {
            float* col = hX.data() + (size_t)jj * n;
            if (st.start == SPCA_START_EXPLICIT)
                std::memcpy(col, X0 + (size_t)(shift + jj) * n, sizeof(float) * n);
            else if (st.start == SPCA_START_REFERENCE_F64)
                reference_starting_point<double>(col, (unsigned)n, st.constraint, constrained, (unsigned)(shift + jj));
            else
                reference_starting_point<float>(col, (unsigned)n, st.constraint, constrained, (unsigned)(shift + jj));
        }
        HIP_TRY(ctx, hipMemsetAsync(w.X, 0, sizeof(float) * (size_t)w.np * w.bp, s));
        HIP_TRY(ctx, hipMemcpy2DAsync(w.X, sizeof(float) * w.np, hX.data(), sizeof(float) * n, sizeof(float) * n,
                                      batch, hipMemcpyHostToDevice, s));
        if (st.precision) {
            HIP_TRY(ctx, hipMemsetAsync(w.Xhi, 0, sizeof(uint16_t) * (size_t)w.np * w.bp, s));
            HIP_TRY(ctx, hipMemsetAsync(w.Xlo, 0, sizeof(uint16_t) * (size_t)w.np * w.bp, s));
            HIP_TRY(ctx, launch_x_operand(xop, st.precision, w.np, batch, s));
        }
        // ---- per-SP state: val = err = 0, iters = nnz = 0, active = (j < batch), done = 0
        HIP_TRY(ctx, hipMemsetAsync(w.fstate, 0, sizeof(float) * 4 * (size_t)w.bp, s));
        HIP_TRY(ctx, hipMemsetAsync(w.istate, 0, sizeof(int) * (3 * (size_t)w.bp + 4), s));
        {
            std::vector<int> act(w.bp, 0);
            std::fill(act.begin(), act.begin() + batch, 1);
            HIP_TRY(ctx, hipMemcpyAsync(sp.active, act.data(), sizeof(int) * w.bp, hipMemcpyHostToDevice, s));
            HIP_TRY(ctx, hipStreamSynchronize(s));
        }
        HIP_TRY(ctx, hipEventRecord(loop_ev[0], s));
        for (int it = 0; it < st.max_iter; it++) {
            hipEvent_t e0, e1;
            HIP_TRY(ctx, pt.pair(&e0, &e1, 0));
            HIP_TRY(ctx, hipEventRecord(e0, s));
            HIP_TRY(ctx, launch_gemm_tn(st.precision, P1_32, P1_hi, P1_lo, M->np, w.X, w.Xhi, w.Xlo, M->np, w.Zpart,
                                        M->mp, zsplit, M->mp, bp, M->np, splits, sp.done, s));
            HIP_TRY(ctx, hipEventRecord(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 2));
            HIP_TRY(ctx, hipEventRecord(e0, s));
            HIP_TRY(ctx, launch_prep_z(w.Zpart, splits, zsplit, M->mp, M->mp, zop, st.formulation, st.precision, batch,
                                       sp, s));
            HIP_TRY(ctx, hipEventRecord(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 1));
            HIP_TRY(ctx, hipEventRecord(e0, s));
            HIP_TRY(ctx, launch_gemm_tn(st.precision, P2_32, P2_hi, P2_lo, M->mp, w.Z, w.Zhi, w.Zlo, M->mp, w.V,
                                        M->np, 0, M->np, bp, M->mp, 1, sp.done, s));
            HIP_TRY(ctx, hipEventRecord(e1, s));
            HIP_TRY(ctx, pt.pair(&e0, &e1, 2));
            HIP_TRY(ctx, hipEventRecord(e0, s));
            HIP_TRY(ctx, launch_step(w.V, xop, sa, sp, s));
            HIP_TRY(ctx, launch_converge(sp, batch, it, st.tolerance, st.max_iter,
                                         st.termination == SPCA_TERM_PER_SP, s));
            HIP_TRY(ctx, hipEventRecord(e1, s));
            if ((it + 1) % check_every == 0 || it + 1 == st.max_iter) {
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
        // ---- batch results
        int done_h[2];
        float emax;
        HIP_TRY(ctx, hipMemcpyAsync(done_h, sp.done, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(hval.data(), sp.val, sizeof(float) * batch, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(herr.data(), sp.err, sizeof(float) * batch, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(hit.data(), sp.iters, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(hnnz.data(), sp.nnz, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(&emax, sp.errmax, sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        total_it += (unsigned)done_h[1];
        last_err = emax;
        for (int jj = 0; jj < batch; jj++) sp_its += (unsigned long long)hit[jj];
        // best of batch: first maximum, as sparse_PCA_solver.h:212-231
        int sel = 0;
        float bv = hval[0];
        for (int jj = 1; jj < batch; jj++)
            if (hval[jj] > bv) {
                bv = hval[jj];
                sel = jj;
            }
        const bool want_ev = per_sp != nullptr || stats != nullptr;
        if (want_ev) {
            HIP_TRY(ctx, launch_ev_fp64(M->a32_cm, M->mp, m, n, w.X, w.np, batch, spca_is_l1_variance(st.formulation),
                                        w.ev, s));
            HIP_TRY(ctx, hipMemcpyAsync(hev.data(), w.ev, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
            HIP_TRY(ctx, hipStreamSynchronize(s));
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
        if (best < (double)bv || best_sp < 0) {
            best = bv;
            best_sp = shift + sel;
            HIP_TRY(ctx, hipMemcpyAsync(x_best.data(), w.X + (size_t)sel * w.np, sizeof(float) * n,
                                        hipMemcpyDeviceToHost, s));
            if (stats) stats->ev_fp64 = hev[sel];
        }
        HIP_TRY(ctx, hipStreamSynchronize(s));
    }
    // final x <- x/|x|_2 (sparse_PCA_solver.h:236-237)
    double nx = 0;
    for (int i = 0; i < n; i++) nx += (double)x_best[i] * (double)x_best[i];
    const float inv = 1.0f / (float)std::sqrt(nx);
    for (int i = 0; i < n; i++) x_out[i] = x_best[i] * inv;
    if (stats) {
        stats->it = total_it;
        stats->fval = best;
        stats->error = last_err;
        stats->loop_ms = loop_ms;
        stats->wall_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_wall0).count();
        stats->n = (unsigned)n;
        stats->best_sp = best_sp;
        stats->total_sp = st.total_sp;
        stats->constraint = st.constraint;
        stats->gemm_ax_ms = pt.ms[0];
        stats->gemm_atz_ms = pt.ms[1];
        stats->step_ms = pt.ms[2];
        stats->sp_iterations = sp_its;
    }
    return SPCA_OK;
}

// FP64 re-solve on the support of x (host vector of n floats)
int resolve_impl(spca_ctx ctx, Engine& E, Matrix* M, int formulation, const float* x, int max_iter, double tol,
                 double* y_out, double* ev_out, int* iters_out) {
    if (!valid_formulation(formulation) || !x || max_iter < 1) return fail(ctx, SPCA_EINVAL, "bad re-solve arguments");
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
                      float* x_out, spca_stats* stats, spca_sp_result* per_sp, float* X_final) {
    int rc = solve_impl(ctx, E, M, settings, X0, x_out, stats, per_sp, X_final);
    if (rc == SPCA_OK && settings->resolve_iters > 0 && stats) {
        int its = 0;
        rc = resolve_impl(ctx, E, M, settings->formulation, x_out, settings->resolve_iters, settings->resolve_tol,
                          nullptr, &stats->ev_resolve_fp64, &its);
        stats->resolve_iterations = its;
    }
    return rc;
}


// ---------------------------------------------------------------- request server
// Continuous batching: jobs wait in a FIFO queue; each worker thread owns an engine (its
// own stream and workspace) and takes the next job as soon as it is free and the job's
// workspace fits the HBM budget.  A job that does not fit waits for a running job to
// return its memory; it fails with SPCA_EBUDGET only when nothing else is running.  Work
// from different workers overlaps on the device.  Kernels are deterministic and every job
// has its own buffers, so a job's result does not depend on what runs beside it.
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
};

}  // namespace

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

void worker_main(spca_ctx ctx, spca_server* S, int w) {
    (void)hipSetDevice(ctx->device);
    Engine& E = S->engines[w];
    for (;;) {
        Job* J = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv.wait(lk, [&] { return S->stop || !S->queue.empty(); });
            if (S->queue.empty()) break;  // stop requested and nothing queued
            J = S->queue.front();
            S->queue.pop_front();
            J->state = SPCA_JOB_RUNNING;
            S->running++;
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
            const float* X0 = J->X0.empty() ? nullptr : J->X0.data();
            rc = solve_and_resolve(ctx, E, J->M, &J->st, X0, J->x.data(), &J->stats,
                                   J->want_per_sp ? J->per_sp.data() : nullptr, nullptr);
            E.cancel = nullptr;
            if (rc != SPCA_OK) J->err = tl_err;
        }
        J->t_end = std::chrono::steady_clock::now();
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
        matrix_unref(ctx, J->mat);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            J->status = rc;
            J->state = rc == SPCA_ECANCELLED ? SPCA_JOB_CANCELLED : SPCA_JOB_DONE;
            S->running--;
        }
        S->cv.notify_all();
    }
    free_workspace(ctx, E);
}

void server_stop(spca_ctx ctx) {
    spca_server* S = ctx->server;
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        for (Job* J : S->queue) {
            J->status = SPCA_ECANCELLED;
            J->state = SPCA_JOB_CANCELLED;
            J->cancel.store(1);
        }
        S->queue.clear();
        for (auto& kv : S->jobs) kv.second->cancel.store(1);
    }
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

int spca_matrix_add_storage(spca_ctx ctx, spca_matrix mat, unsigned storage) {
    if (!ctx) return SPCA_EINVAL;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto itm = ctx->mats.find(mat);
    if (itm == ctx->mats.end()) return fail(ctx, SPCA_ENOTFOUND, "unknown matrix handle");
    (void)hipSetDevice(ctx->device);
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
        const unsigned need = J->st.precision == 0 ? SPCA_STORE_FP32
                              : J->st.precision == 1 ? SPCA_STORE_BF16X3 : SPCA_STORE_FP16;
        (void)hipSetDevice(ctx->device);
        int rc = add_storage(ctx, M, need);
        if (rc) {
            delete J;
            return rc;
        }
        spca_settings chk = J->st;
        spca_settings_check(&chk, (unsigned)M->n);
        J->M = M;
        J->mat = mat;
        J->np = M->np;
        J->mp = M->mp;
        J->bp = round_up(chk.batch, kPadB);
        J->splits = choose_splits(J->st.precision, M->mp, M->np, J->bp);
        const uint64_t need_bytes = workspace_bytes(J->np, J->mp, J->bp, J->splits);
        if (need_bytes + M->bytes > ctx->budget) {
            delete J;
            return fail(ctx, SPCA_EBUDGET, "job workspace exceeds the HBM budget");
        }
        if (J->st.start == SPCA_START_EXPLICIT) J->X0.assign(X0, X0 + (size_t)M->n * chk.total_sp);
        J->x.assign(M->n, 0.f);
        J->want_per_sp = want_per_sp != 0;
        if (J->want_per_sp) J->per_sp.resize(chk.total_sp);
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
    Job* J = nullptr;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->jobs.find(job);
        if (it == S->jobs.end()) return SPCA_ENOTFOUND;
        J = it->second;
        J->cancel.store(1);
        for (auto q = S->queue.begin(); q != S->queue.end(); ++q)
            if (*q == J) {
                S->queue.erase(q);
                J->status = SPCA_ECANCELLED;
                J->state = SPCA_JOB_CANCELLED;
                break;
            }
    }
    if (J->state == SPCA_JOB_CANCELLED && J->M) {
        // a queued job never reached a worker: drop its matrix reference here
        if (J->t_start.time_since_epoch().count() == 0) matrix_unref(ctx, J->mat);
        J->M = nullptr;
    }
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
        if (x_out) std::memcpy(x_out, J->x.data(), sizeof(float) * J->x.size());
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
