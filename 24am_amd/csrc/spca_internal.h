// Internal interface between the host driver (spca_host.cpp) and the HIP kernels
// (spca_kernels.hip).  Not part of the public C ABI (include/spca.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spca {

// Padding of the resident operand layouts: rows and columns of A are padded with zeros
// to multiples of kPadMN so every product tile is full; iterate batches are padded to
// multiples of kPadB starting points (padded columns stay zero and are never active).
constexpr int kPadMN = 256;
constexpr int kPadB = 128;

// GEMM tile: 128x128 output tile, 32-deep K stage, 256 threads (2x2 waves of 64x64).
constexpr int kGemmBM = 128;
constexpr int kGemmBN = 128;
constexpr int kGemmBK = 32;

// Select/formulation-step kernel: workgroups of kSelThreads; four share a CU (LDS about 37 KiB
// each), so the 1024 SPs of a large batch are all resident at once and one's sort and scan
// phases overlap the others' column reads.  A constrained step whose previous threshold left a
// list bound has its column scanned first, in chunks of kStepChunk elements, one workgroup per
// chunk.
constexpr int kSelThreads = 256;
constexpr int kSelDepth = 16;  // loads in flight per thread in each of the step kernel's column passes
constexpr int kStepScanRows = 8;  // 16-byte loads per thread of one chunk scan
constexpr int kStepChunk = kSelThreads * 4 * kStepScanRows;
constexpr int kStageCap = 1024;   // staged list entries per chunk (more: the step runs the column passes)
__host__ __device__ inline int step_chunks(int ldv) { return (ldv + kStepChunk - 1) / kStepChunk; }
// Step kind of a formulation: 0 L0 constraint (also the L1 constraint with s <= 1, which admits
// only 1-sparse x: the top-1 hard threshold), 1 L1 constraint, 2 L0 penalty, 3 L1 penalty.
__host__ __device__ inline int step_kind(int f, unsigned s) {
    return f >= 6 ? 3 : f >= 4 ? 2 : (f >= 2 && s > 1) ? 1 : 0;
}

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Per-SP state flag block (SpState::done): 8 counters/flags, then one sparse-A*X flag per
// group of 256 SPs.  Batches are limited to kMaxSparseGroups * 256 SPs.
constexpr int kSparseFlags = 8;
constexpr int kMaxSparseGroups = 256;
inline int state_flag_ints(int bp) { return kSparseFlags + (bp + 255) / 256; }

// Philox4x32-10 block (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC 2011): ten rounds on the 128-bit counter c with the 64-bit key (k0, k1).
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// Per-starting-point device state (structure of arrays, length = padded batch).
struct SpState {
    float* val;      // value of the last completed iteration (reference vals[i].val)
    float* fz;       // objective taken from z in this iteration (|z|_2 or |z|_1)
    float* err;      // current_error
    int* iters;      // iterations performed
    int* nnz;        // nonzeros of the iterate
    int* active;     // 1 while the SP iterates
    int* done;       // [0]: batch finished, [1]: iterations counted, [2]: constrained steps that
                     // tried the previous threshold's list, [3]: those it decided, [4]: iterations
                     // with a gathered A*X counted, [5..7]: spare, [kSparseFlags + g]: the next
                     // A*X of SPs 256g .. 256g+255 is the sparse gather (their dense tiles stand down)
    float* errmax;   // [0]: batch max error of the last iteration
    float* thr;      // constrained: the bound of the next step's candidate list, left by the last
                     // step (0: none)
    int* nkept;      // constrained: nonzeros the last step wrote through its candidate list, whose
                     // indices are kept[j * kListMax ..] (-1: the last step wrote the whole column)
    int* kept;       // kListMax indices per SP
    // the chunk scans of a constrained step: max |v| as float bits (zero between steps; the step
    // reads and clears it), and per chunk c of SP j (step_chunks(ldv) chunks per SP) its entry
    // count lcount[j * chunks + c] and its entries lstage[(j * chunks + c) * kStageCap ..]
    unsigned* lmax;
    int* lcount;
    unsigned long long* lstage;
};
// per-SP float state: val, fz, err, errmax, thr; int state: iters, nnz, active, nkept, lmax
constexpr int kStateFloats = 5;
constexpr int kStateInts = 5;
// Most entries of the step kernel's candidate list (a power of two, for its sort), and so of the
// nonzeros a step written through the list can leave.
constexpr int kListMax = 4096;

// Operand buffers of one iterate matrix (n or m rows, padded batch columns).  The FP32
// values are column major; the 16-bit forms the products read are K-blocked: element i of
// column j at ((i / 32) * bp + j) * 32 + i % 32.
struct Operand {
    float* f32;          // FP32 values (also the master copy of iterates)
    uint16_t* hi;        // bf16 hi part, or fp16 values, or (Z of the fixed-point product) int8 signs
    uint16_t* lo;        // bf16 lo part (bf16x3 only)
    int bp;              // columns (padded batch)
};

// On-the-fly (OTF) slot refill state (reference OTF driver, c++/src/gpower/sparse_PCA_solver.h:83-169).
struct OtfState {
    int* order;        // global SP index held by each slot
    int* cur_it;       // iterations of the slot's current SP
    int* refill;       // 1: the slot takes a new SP before the next iteration
    int* generated;    // [0]: SPs handed out so far; [1]: slot whose column becomes the best this
                       // iteration (-1 none); [2]: global index of the best SP so far
    float* best_val;   // [0]: best value so far
    float* sp_val;     // per global SP: value at its last termination
    int* sp_iters;     // per global SP: iterations at its last termination
    int* sp_nnz;       // per global SP: nonzeros at its last termination
    const float* pool; // all starting points, ld = padded n
    float* xbest;      // iterate of the best SP (padded n)
};

// Exact signs of z = A x for the L1-variance formulations after a dense A*X product: entries
// with |z_i| <= eps * rowmax_i * |x|_2 (inside the product's rounding of zero) are recomputed
// in FP64 from the FP32 master.  eps = 0: off.
struct SignRefine {
    const float* arm;     // A row major (mp x np, ld np)
    const float* rowmax;  // max_k |A_ik| per row (mp; zero rows and padding 0)
    const float* x;       // the iterates the product read (FP32, ld np)
    int np;
    float eps;
};
constexpr int kRefineMax = 256;  // flagged entries of one SP and iteration (more: none refined)

struct StepArgs {
    int formulation;
    int precision;       // spca_precision
    int n;               // true length of x (columns of A)
    int ldv;             // leading dimension of V and X (padded n)
    unsigned s;          // constraint after clamp
    double gamma;
    int batch;           // true number of SPs in this batch
    // Exact ties (L0 constraint, L2 variance): candidates for the s-th place within
    // 2 * band_eps * tau of tau are recomputed as FP64 dot products of A's column with z.
    const float* acm;    // A column major (ld mp)
    const float* zf;     // z of this iteration, FP32 (ld mp)
    int mp;
    float band_eps;      // 0: off
    int sparse_rewrite;  // 1: a step decided on its candidate list rewrites only the previous and new
                         // nonzeros of x (0 when something else changes x between steps: deflation)
};
constexpr int kBandMax = 16;  // most candidates one tie band re-evaluates (more: none)

// One job of a fused multi-job iteration: every field the five per-iteration kernels of a
// solo solve take, plus the job's first tile in each product's tile list and its first
// column in the list of starting points.  Jobs of one fused launch share the precision.
struct FusedJob {
    const float* p1_32;     // A row major (mp x np, ld np): A*X operand
    const uint16_t* p1_hi;
    const uint16_t* p1_lo;
    const float* p2_32;     // A column major (ld mp): A^T*Z operand
    const uint16_t* p2_hi;
    const uint16_t* p2_lo;
    const int8_t* p2_q;      // fixed-point pieces of A (p2_prec 4)
    const float* p2_qscale;  // their per-column scales
    float* Zpart;           // splits x (mp x bp) partial products of A*X
    float* V;               // np x bp
    Operand xop, zop;
    SpState sp;
    StepArgs sa;
    SignRefine rf;
    long zsplit;            // mp * bp
    double tol;
    int mp, np, bp, splits;
    int t1, t2, c0;         // first A*X tile, first A^T*Z tile, first column of this job
    int t3;                 // first (row block, SP) item of the sparse A*X product
    int max_iter, per_sp;
    int sparse_cap;         // largest nonzero count for the sparse A*X product (0: never)
    int p2_prec;            // A^T*Z operand form: the job's precision, or 4 (fixed point x sign vectors)
    int zprec;              // form prep_z writes Z in (as p2_prec)
    int nb1, nb2;           // FP32 A*X / A^T*Z in the narrow 128 x nb tile (32 or 64; 0: 128 x 128 tiles)
    int t4;                 // first (row block, SP) item of the partial-sum kernel
    int psum;               // 1: the partial-sum kernel adds the K splits into partial 0 before prep_z
    int t5;                 // first (SP, chunk) item of the steps' chunk workgroups
    int* gl;                // the gathered A*X's nonzero lists (kGatherStride ints per SP; null: scan x)
};

// ---- launchers (spca_kernels.hip); every one enqueues on `st` and returns a hip status
hipError_t launch_pad_copy_f32(const float* src, int m, int n, int lds, float* dst, int ldd, int mp,
                               int np, hipStream_t st);
hipError_t launch_transpose_f32(const float* src, int rows, int cols, int lds, float* dst, int ldd,
                                hipStream_t st);
// src: R rows x K (row r at src + r*lds) -> K-blocked 16-bit form (precision 1: bf16 hi + lo,
// 2: fp16 in hi): element (r, k) at ((k / 32) * R + r) * 32 + k % 32.  K % 32 == 0.
hipError_t launch_to_blocked(const float* src, long lds, int R, int K, int precision, uint16_t* hi, uint16_t* lo,
                             hipStream_t st);
// A (mp x np column major) -> the fixed-point form of the exact A^T * sgn(z) product: per
// column a power-of-two scale s_j (qscale) and four int8 pieces with
// A_ij = s_j (2^-7 q0 + 2^-13 q1 + 2^-19 q2 + 2^-25 q3) + e, |e| <= 2^-26 s_j; piece p at
// q + p * mp * np, element (j, i) K-blocked by 64 over np rows: ((i / 64) * np + j) * 64 + i % 64.
hipError_t launch_to_fix8(const float* A, int mp, int np, int8_t* q, float* qscale, hipStream_t st);

// C[r + c*ldc + split*split_stride] = sum_{k in split} P(r, k) * Q(c, k)
// FP32: P(r, k) = P32[r*ldp + k], Q(c, k) = Q32[c*ldq + k].  16-bit: K-blocked operands with
// ldp / ldq rows, P(r, k) = P[((k / 32) * ldp + r) * 32 + k % 32] (likewise Q).  precision
// 1: bf16x3 (Phi, Plo, Qhi, Qlo), 2: fp16 (Phi, Qhi).
// M % 128 == 0, N % 128 == 0, K % (32*splits) == 0.  big = 1 runs the pipelined 256x256-tile
// kernel, allowed for 16-bit products with M and N multiples of 256 (gemm_big_tiles); else
// 128x128 tiles.  done (nullable): the launch returns when *done; skip (nullable): tiles
// of 256-column group g return when skip[g].
int gemm_big_tiles(int precision, int M, int N);
hipError_t launch_gemm_tn(int precision, const float* P32, const uint16_t* Phi, const uint16_t* Plo, long ldp,
                          const float* Q32, const uint16_t* Qhi, const uint16_t* Qlo, long ldq, float* C, long ldc,
                          long split_stride, int M, int N, int K, int splits, int big, const int* done,
                          const int* skip, hipStream_t st);
// C[r + c*ldc] = s_r (2^-13 accA + 2^-25 accB): the exact A^T * sgn(z) from the fixed-point
// pieces P (launch_to_fix8 with np = ldp rows, mp = K) against int8 signs Q (K-blocked by 64
// over ldq rows, entries -1, 0, +1); int32 accumulation, so the result does not depend on
// tiling or order.  M % 128 == 0, N % 128 == 0, K % 64 == 0, K < 2^18.
hipError_t launch_gemm_fix8(const int8_t* P, const float* qscale, long ldp, const int8_t* Q, long ldq, float* C,
                            long ldc, int M, int N, int K, const int* done, hipStream_t st);

// FP32 product for batches of at most 64 SPs (columns 0 .. N-1 of C written; N <= 64): the
// 128 x 32 or 128 x 64 tile, bitwise equal to launch_gemm_tn's FP32 product with the same split
hipError_t launch_gemm_f32_narrow(const float* P, long ldp, const float* Q, long ldq, float* C, long ldc,
                                  long split_stride, int M, int N, int K, int splits, const int* done,
                                  const int* skip, hipStream_t st);

// Sum split partials of Z, per-column statistic and map (sign / normalise), write the
// operand of the second product.  Batches of at most 256 SPs first sum the partials into
// partial 0 over (row block, SP) workgroups, in the same order.
hipError_t launch_prep_z(float* Zpart, int splits, long split_stride, int m, int ldz,
                         Operand zop, int formulation, int precision, int batch, SpState sp,
                         SignRefine rf, hipStream_t st);
// rowmax[i] = max_k |A_ik| of A (mp x np column major).
hipError_t launch_row_max(const float* Acm, int mp, int np, float* out, hipStream_t st);

// Formulation step on V (A^T Z) for each active SP, writes X (+operand), updates state.
hipError_t launch_step(const float* V, Operand xop, StepArgs a, SpState sp, hipStream_t st);

// Batch convergence: batch max error / per-SP freezing, done flag; chooses the next A*X
// product (sparse gather when every iterating SP has at most sparse_cap nonzeros).
hipError_t launch_converge(SpState sp, int batch, int it, double tol, int max_iter, int per_sp, int sparse_cap,
                           hipStream_t st);

// OTF: per-slot termination, best-so-far, assignment of new SPs to finished slots in slot
// order; then the refill kernel saves the best column and loads the new SPs.
hipError_t launch_otf_converge(SpState sp, OtfState o, int batch, double tol, int max_iter, int total_sp,
                               hipStream_t st);
// refilled slots take their next SP; sparse_cap > 0: then each group of 256 slots gets its
// sparse-gather flag for the next A*X (every active slot at most sparse_cap nonzeros)
hipError_t launch_otf_refill(SpState sp, OtfState o, Operand xop, int precision, int n, int ldx, int batch,
                             int sparse_cap, hipStream_t st);

// Synthetic spiked matrices: A (mp x np column major) <- N(0,1) noise from Philox4x32-10
// counters (r, c, 0) keyed by seed (padding zero); then A[:, idx[t]] += coef[t] * g, one
// planted loading per call.
hipError_t launch_gen_noise(float* A, int m, int n, int mp, int np, uint64_t seed, hipStream_t st);
// Starting points shift .. shift+batch-1 generated in place (columns of X, ld ldx).
hipError_t launch_gen_starts(float* X, int ldx, int n, int batch, int shift, uint64_t seed, hipStream_t st);
hipError_t launch_plant(float* A, int m, int mp, const int* idx, const float* coef, int cnt, const float* g,
                        hipStream_t st);

// One AM iteration of every job in `jobs` (device array, sorted by t1/t2/c0/t3/t4): A*X tiles
// of all jobs, the sparse gather items, the partial-sum items of small-batch jobs (items4, may
// be 0), the z epilogue of all columns, A^T*Z tiles (any_fix8: plus a launch over the same
// tile list for the jobs on the fixed-point product), the formulation step of all columns and
// the per-job convergence (a finished job adds 1 to *all_done).  128x128 tiles (128 x 32 / 64
// for FP32 jobs of at most 64 SPs); per job the same tile order, K split, sum order and
// kernel bodies as a solo solve, so a job's FP32 results do not depend on the other jobs.
hipError_t launch_fused_iteration(int precision, const FusedJob* jobs, int njobs, int tiles1, int tiles2, int cols,
                                  int items3, int items4, int items5, int any_fix8, int* all_done, hipStream_t st);

// z_j = A x_j over the nonzeros of each active x_j (FP64 accumulation, fixed order) into Z
// (ld ldz), for the groups of 256 SPs whose sparse flag (sp.done[kSparseFlags + g]) is set; the
// other groups' workgroups return at once.  A: FP32 column major, ld lda.
constexpr int kSparseRows = 1024;  // rows per workgroup of the sparse product
// Gathered A*X over per-SP nonzero lists (solo solves): each gathered SP's list holds a count
// (int 0), then up to kGatherList ascending indices (from int kGatherHead) and their values (float
// bits, kGatherList further on).  An SP with more nonzeros is gathered by scanning x instead.
constexpr int kGatherList = 4096;
constexpr int kGatherHead = 64;
constexpr int kGatherStride = kGatherHead + 2 * kGatherList;
// Solo solves of at most kOrderMaxBatch SPs gather in the order of each SP's largest |x_i|
// (header int 1 of its list: that index; int 2 of SP i's header: the SP at position i).
constexpr int kOrderMaxBatch = 4096;
// union A*X: flags and the ascending list U of the indices nonzero in any iterate whose A*X is
// dense next iteration, stat = {|U|, done[0], done[1]}; then Au = A[:, U] (row major) and
// Xu = X[U, :] (row j = SP j), both zero-padded to kpad columns, for the FP32 product
hipError_t launch_union_list(const float* X, int ldx, int n, int batch, SpState sp, int* flag, int* U, int* stat,
                             hipStream_t st);
hipError_t launch_union_gather(const float* Acm, int lda, int mp, const float* X, int ldx, int rows, const int* U,
                               int K, int kpad, float* Au, float* Xu, int ldu, hipStream_t st);
hipError_t launch_sparse_ax(const float* Acm, int lda, int mp, int n, const float* X, int ldx, float* Z, int ldz,
                            int batch, SpState sp, hipStream_t st);
hipError_t launch_sparse_ax_listed(const float* Acm, int lda, int mp, int n, const float* X, int ldx, float* Z,
                                   int ldz, int batch, SpState sp, int* gl, hipStream_t st);

// Deflation by earlier components (reference SparseDeflationCollection,
// c++/src/gpower/sparse_PCA_solver_for_CSC.h:56-106): for c = 0 .. ncomp-1 in order, each
// active column of op (ld rows) loses its projection on component c, whose nonzeros are
// idx[off[c] .. off[c+1]) with values val; the operand forms of `precision` are rewritten at
// the changed entries.
hipError_t launch_deflate(Operand op, int precision, long ld, const int* idx, const float* val, const int* off,
                          int ncomp, int batch, SpState sp, hipStream_t st);

// Sparse A (reference sparse_PCA_solver_CSC, c++/src/gpower/sparse_PCA_solver_for_CSC.h:112-495):
// CSC arrays and their CSR copy (zero-based), with optional implicit centring
// A_c = A - 1 mu^T - r 1^T (mu: column means, the reference's doMean; r: row means, doRowMean).
struct SparseMat {
    const int* cptr;   // n + 1 column pointers
    const int* ridx;   // row of each nonzero, columns in order
    const float* cval;
    const int* rptr;   // m + 1 row pointers
    const int* cidx;   // column of each nonzero, rows in order
    const float* rval;
    const float* mu;   // n column means or null
    const float* rmu;  // m row means or null
    int m, n;
};
// Per-SP scalars of a batch of columns (ld ldx, len entries): out[b] = w . x_b (0 when w is
// null), out[bp + b] = 1 . x_b, in FP64 in a fixed order; inactive SPs are skipped.
hipError_t launch_sp_scalars(const float* X, long ldx, int len, const float* w, double* out, int bp, int batch,
                             SpState sp, hipStream_t st);
// Z[:, b] = A_c x_b (rows >= m zero) with xs = launch_sp_scalars(X, mu): FP64 sums per row.
hipError_t launch_csr_ax(SparseMat A, const float* X, int ldx, float* Z, int ldz, int mp, const double* xs, int bp,
                         int batch, SpState sp, hipStream_t st);
// V[:, b] = A_c^T z_b (rows >= n zero) with zs = launch_sp_scalars(Z, r): FP64 sums per column.
hipError_t launch_csc_atz(SparseMat A, const float* Z, int ldz, float* V, int ldv, int np, const double* zs, int bp,
                          int batch, SpState sp, hipStream_t st);
// FP64 explained variance |A_c xhat|_2^2 or |A_c xhat|_1 of `batch` columns of X (ld ldx).
// work: csr_ev_scratch_doubles(m, batch) doubles.
size_t csr_ev_scratch_doubles(int m, int batch);
hipError_t launch_csr_ev(SparseMat A, const float* X, int ldx, int batch, int l1, double* ev, double* work,
                         hipStream_t st);

// Write the operand forms of X from its FP32 values (for starting points).
hipError_t launch_x_operand(Operand xop, int precision, int ldx, int batch, hipStream_t st);

// FP64 explained variance of each normalised column of X: |A x|_2^2 or |A x|_1.  work:
// ev_scratch_doubles(m, n, batch) doubles of device memory.
size_t ev_scratch_doubles(int m, int n, int batch, int* chunks = nullptr, int* chunk_len = nullptr);
hipError_t launch_ev_fp64(const float* Acm, int lda, int m, int n, const float* X, int ldx, int batch,
                          int l1, double* ev, double* work, hipStream_t st);

// FP64 re-solve on a support (indices idx[0..k)): max_iter iterations of z = map(A_S y),
// y = A_S^T z / |A_S^T z|, stopping (flag) when the objective changes by <= tol relative.
// ev[0] receives the objective of the final y.  Buffers: As m*k, y k, z m, yp k,
// part ceil(m/256), ev 1, flag 1 (all device; y holds the unit start, flag zeroed).
hipError_t launch_resolve(const float* Acm, int lda, int m, const int* idx, int k, int l1, double* As, double* y,
                          double* z, double* yp, double* part, double* ev, int* flag, double tol, int max_iter,
                          hipStream_t st);

// Fixed device scratch of the allocation-free solve: kScratchBytes per GPU, in device memory
// from module load to process end.  scratch_setup(device), with `device` current, looks up the
// array's device address once per device (hipGetSymbolAddress) and returns that lookup's
// status; scratch_ptr(device) then gives the address (null if the lookup failed or never ran).
constexpr size_t kScratchBytes = 128ull << 20;
hipError_t scratch_setup(int device);
void* scratch_ptr(int device);
// p[0 .. bytes) = 0 by a kernel (no memset node where a caller captures); p 16-byte aligned
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t st);

}  // namespace spca
