// HIP kernels of the MI355X AM sparse-PCA solver (gfx950 / CDNA4).
//
// One batched AM iteration over B starting points is, on one stream:
//   gemm_tn  (Zpart = A * X, split over K)           reference cublasSgemm(N,N),
//                                                    c++/src/gpugpower/gpu_sparse_PCA_solver.h:111-112
//   prep_z   (sum splits, |z|_2 or |z|_1, sign / normalise map, operand write)
//                                                    :105-121,158-164 (nrm2/asum/scal/sgn)
//   gemm_tn  (V = A^T * Z)                           :125-126,168-169
//   step     (top-s select | lambda soft threshold | gamma threshold, norm, error)
//                                                    c++/src/gpugpower/my_cublas_wrapper.h:571-710
//   converge (batch max error, termination)         gpu_sparse_PCA_solver.h:176-186
// Iteration semantics follow the reference CPU method (c++/src/gpower/gpower_commons.h:67-177);
// the thresholds follow the reference GPU path (sorted tau, lambda scan).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "spca_internal.h"

namespace spca {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- conversions
__device__ __forceinline__ uint16_t f2bf(float f) {
    // round to nearest even; inputs here are finite
    uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2h(float f) {
    _Float16 h = (_Float16)f;
    return __builtin_bit_cast(uint16_t, h);
}

// Buffer resources: a bounded view of one column.  Loads past num_records return 0 and
// stores past it are dropped, which handles ragged column lengths without per-element
// address registers (offsets go in the instruction's scalar offset).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t col_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// ---------------------------------------------------------------- reductions
// Deterministic: fixed shuffle tree inside a wave, fixed order across waves.
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
    return v;
}
template <int NT>
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double r = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) r += red[i];
    return r;
}
template <int NT>
__device__ float block_max(float v, double* red) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = (double)v;
    __syncthreads();
    float r = 0.f;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) r = fmaxf(r, (float)red[i]);
    return r;
}
// Three sums reduced together (the same tree and order per value as block_sum, so the same
// bits; the three shuffle chains overlap).  red: 3 * NT / 64 doubles.
template <int NT>
__device__ void block_sum3(double& a, double& b, double& c, double* red) {
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int w = threadIdx.x >> 6;
    constexpr int W = NT / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[w] = a;
        red[W + w] = b;
        red[2 * W + w] = c;
    }
    __syncthreads();
    a = b = c = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        a += red[i];
        b += red[W + i];
        c += red[2 * W + i];
    }
}
template <int NT>
__device__ void block_sum2(double& a, double& b, double* red) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6;
    constexpr int W = NT / 64;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[w] = a;
        red[W + w] = b;
    }
    __syncthreads();
    a = b = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        a += red[i];
        b += red[W + i];
    }
}

// Integer scans over the wave's 64 lanes by DPP: row_shr 1, 2, 4, 8 within each row of 16
// lanes, then row_bcast:15 (into rows 1 and 3) and row_bcast:31 (into rows 2 and 3).  Integer
// sums and maxima are exact, so the result is that of any order of evaluation.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v, uint32_t ident) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)ident, (int)v, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ uint32_t wave_prefix_u32(uint32_t x) {
    x += dpp_u32<0x111, 0xF>(x, 0);
    x += dpp_u32<0x112, 0xF>(x, 0);
    x += dpp_u32<0x114, 0xF>(x, 0);
    x += dpp_u32<0x118, 0xF>(x, 0);
    x += dpp_u32<0x142, 0xA>(x, 0);
    x += dpp_u32<0x143, 0xC>(x, 0);
    return x;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    const uint32_t lo = dpp_u32<CTRL, ROWS>((uint32_t)v, 0), hi = dpp_u32<CTRL, ROWS>((uint32_t)(v >> 32), 0);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_prefix_u64(uint64_t x) {
    x += dpp_u64<0x111, 0xF>(x);
    x += dpp_u64<0x112, 0xF>(x);
    x += dpp_u64<0x114, 0xF>(x);
    x += dpp_u64<0x118, 0xF>(x);
    x += dpp_u64<0x142, 0xA>(x);
    x += dpp_u64<0x143, 0xC>(x);
    return x;
}
// inclusive suffix sum over lanes l..63 (lane 63 holds the top): the wave total minus the
// prefix below lane l
__device__ __forceinline__ uint32_t wave_suffix_u32(uint32_t v) {
    const uint32_t p = wave_prefix_u32(v);
    return (uint32_t)__builtin_amdgcn_readlane((int)p, 63) - p + v;
}
__device__ __forceinline__ uint64_t wave_suffix_u64(uint64_t v) {
    const uint64_t p = wave_prefix_u64(v);
    const uint64_t tot = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(p >> 32), 63) << 32) |
                         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)p, 63);
    return tot - p + v;
}
// maximum over the wave, for values >= -1
__device__ __forceinline__ int wave_max_i(int v) {
    auto mx = [](int x, uint32_t y) { return max(x, (int)y); };
    const uint32_t m1 = 0xFFFFFFFFu;  // -1: the identity for values >= -1
    v = mx(v, dpp_u32<0x111, 0xF>((uint32_t)v, m1));
    v = mx(v, dpp_u32<0x112, 0xF>((uint32_t)v, m1));
    v = mx(v, dpp_u32<0x114, 0xF>((uint32_t)v, m1));
    v = mx(v, dpp_u32<0x118, 0xF>((uint32_t)v, m1));
    v = mx(v, dpp_u32<0x142, 0xA>((uint32_t)v, m1));
    v = mx(v, dpp_u32<0x143, 0xC>((uint32_t)v, m1));
    return __builtin_amdgcn_readlane(v, 63);
}

// ---------------------------------------------------------------- data preparation
__global__ void pad_copy_kernel(const float* __restrict__ src, int m, int n, int lds, float* __restrict__ dst,
                                int ldd, int mp, int np) {
    const long total = (long)mp * np;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % mp), k = (int)(e / mp);
        dst[(long)k * ldd + i] = (i < m && k < n) ? src[(long)k * lds + i] : 0.f;
    }
}
hipError_t launch_pad_copy_f32(const float* src, int m, int n, int lds, float* dst, int ldd, int mp, int np,
                               hipStream_t st) {
    pad_copy_kernel<<<4096, 256, 0, st>>>(src, m, n, lds, dst, ldd, mp, np);
    return hipGetLastError();
}

// dst[c*ldd + r] = src[r + c*lds] transposed: src is rows x cols column major, dst is
// cols x rows column major (i.e. src row-major).  rows, cols multiples of 64.
__global__ void transpose_kernel(const float* __restrict__ src, int rows, int cols, int lds,
                                 float* __restrict__ dst, int ldd) {
    __shared__ float tile[64][65];
    const int r0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 256 threads: 4 rows per pass
    for (int c = ty; c < 64; c += 4) tile[c][tx] = src[(long)(c0 + c) * lds + r0 + tx];
    __syncthreads();
    for (int r = ty; r < 64; r += 4) dst[(long)(r0 + r) * ldd + c0 + tx] = tile[tx][r];
}
hipError_t launch_transpose_f32(const float* src, int rows, int cols, int lds, float* dst, int ldd,
                                hipStream_t st) {
    dim3 g(rows / 64, cols / 64);
    transpose_kernel<<<g, 256, 0, st>>>(src, rows, cols, lds, dst, ldd);
    return hipGetLastError();
}

// 16-bit product operands are stored K-blocked: element (r, k) of an operand with R rows
// (the non-reduction dimension, padded) sits at ((k / 32) * R + r) * 32 + k % 32, so the
// 32-deep K stage of any row range is one contiguous run of 64-byte rows.  The products'
// per-stage loads are then dense blocks instead of 64-byte pieces one operand row apart
// (128 KiB for a 65536-wide operand), which keeps L2 lines and TLB entries whole.
__device__ __forceinline__ long blk_index(int r, int k, int R) { return ((long)(k >> 5) * R + r) * 32 + (k & 31); }

// src: R rows x K (row r at src + r*lds, K contiguous) -> blocked bf16 hi+lo (precision 1)
// or fp16 (precision 2).  One thread per 32-element chunk; consecutive threads take
// consecutive rows of one K block, so the 64-byte output rows are written contiguously.
__global__ void to_blocked_kernel(const float* __restrict__ src, long lds, int R, int K, int precision,
                                  uint16_t* __restrict__ hi, uint16_t* __restrict__ lo) {
    const long chunks = (long)R * (K / 32);
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < chunks; c += (long)gridDim.x * blockDim.x) {
        const int r = (int)(c % R), kb = (int)(c / R);
        const float* sp = src + (long)r * lds + (long)kb * 32;
        const long o = c * 32;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const f4 v0 = *(const f4*)(sp + 8 * q), v1 = *(const f4*)(sp + 8 * q + 4);
            const float x[8] = {v0[0], v0[1], v0[2], v0[3], v1[0], v1[1], v1[2], v1[3]};
            u4 hv, lv;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                uint32_t h[2], l[2] = {0, 0};
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const float v = x[2 * e + u];
                    if (precision == 2) {
                        h[u] = f2h(v);
                        continue;
                    }
                    // v - hi is exact in FP32; lo is its bf16 rounding
                    h[u] = f2bf(v);
                    l[u] = f2bf(v - bf2f((uint16_t)h[u]));
                }
                hv[e] = h[0] | (h[1] << 16);
                lv[e] = l[0] | (l[1] << 16);
            }
            *(u4*)(hi + o + 8 * q) = hv;
            if (precision != 2) *(u4*)(lo + o + 8 * q) = lv;
        }
    }
}
hipError_t launch_to_blocked(const float* src, long lds, int R, int K, int precision, uint16_t* hi, uint16_t* lo,
                             hipStream_t st) {
    if (R <= 0 || K <= 0 || K % 32 || precision < 1 || precision > 2) return hipErrorInvalidValue;
    to_blocked_kernel<<<8192, 256, 0, st>>>(src, lds, R, K, precision, hi, lo);
    return hipGetLastError();
}

// ---------------------------------------------------------------- synthetic spiked matrices
// Counter-based normal variates: philox4x32_10 on counter (a, b, stream, 0) with the 64-bit
// seed as key; outputs 0 and 1, as u = (x + 0.5) * 2^-32 in (0, 1), go through Box-Muller:
// sqrt(-2 ln u0) * cos(2 pi u1).
__device__ inline float philox_normal(uint32_t a, uint32_t b, uint32_t stream, uint64_t seed) {
    uint32_t c[4] = {a, b, stream, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float u0 = ((float)c[0] + 0.5f) * 2.3283064365386963e-10f;
    const float u1 = ((float)c[1] + 0.5f) * 2.3283064365386963e-10f;
    return sqrtf(-2.0f * logf(u0)) * cosf(6.283185307179586f * u1);
}

// A (mp x np column major, zero padding) <- N(0,1) noise, element (r, c) from counter (r, c, 0).
__global__ void gen_noise_kernel(float* __restrict__ A, int m, int n, int mp, int np, uint64_t seed) {
    const long total = (long)mp * np;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int r = (int)(e % mp), c = (int)(e / mp);
        A[e] = (r < m && c < n) ? philox_normal((uint32_t)r, (uint32_t)c, 0u, seed) : 0.f;
    }
}
// A[:, idx[t]] += coef[t] * g for t < cnt (one planted loading; g has m entries)
__global__ void plant_kernel(float* __restrict__ A, int m, int mp, const int* __restrict__ idx,
                             const float* __restrict__ coef, int cnt, const float* __restrict__ g) {
    const long total = (long)m * cnt;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int r = (int)(e % m), t = (int)(e / m);
        float* a = A + (long)idx[t] * mp + r;
        *a = *a + coef[t] * g[r];
    }
}
// X[:, j] (j < batch, ld ldx) <- unit-normalised normals of counters (i, shift + j, 4), i < n;
// rows n..ldx stay as they are (zero).  One workgroup per column; the norm in FP64.
__global__ __launch_bounds__(256) void gen_starts_kernel(float* __restrict__ X, int ldx, int n, int shift,
                                                         uint64_t seed) {
    __shared__ double red[4];
    const int j = blockIdx.x;
    float* x = X + (long)j * ldx;
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const float v = philox_normal((uint32_t)i, (uint32_t)(shift + j), 4u, seed);
        x[i] = v;
        acc += (double)v * v;
    }
    const float inv = (float)(1.0 / sqrt(block_sum<256>(acc, red)));
    for (int i = threadIdx.x; i < n; i += 256) x[i] = x[i] * inv;
}
hipError_t launch_gen_starts(float* X, int ldx, int n, int batch, int shift, uint64_t seed, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    gen_starts_kernel<<<batch, 256, 0, st>>>(X, ldx, n, shift, seed);
    return hipGetLastError();
}
hipError_t launch_gen_noise(float* A, int m, int n, int mp, int np, uint64_t seed, hipStream_t st) {
    gen_noise_kernel<<<8192, 256, 0, st>>>(A, m, n, mp, np, seed);
    return hipGetLastError();
}
hipError_t launch_plant(float* A, int m, int mp, const int* idx, const float* coef, int cnt, const float* g,
                        hipStream_t st) {
    if (cnt <= 0) return hipSuccess;
    plant_kernel<<<1024, 256, 0, st>>>(A, m, mp, idx, coef, cnt, g);
    return hipGetLastError();
}

// ---------------------------------------------------------------- products
// C[r + c*ldc] = sum_k P[r*ldp + k] * Q[c*ldq + k]: both operands are contiguous along
// the reduction, because the matrix is resident in both orientations (row major for
// A*X, column major for A^T*Z).  One 128x128 tile per workgroup, 2x2 waves of 64x64,
// 16x16 MFMA blocks.  Tiles are ordered (split, M tile, N tile) and blockIdx is remapped
// so that each XCD (dispatch is round-robin over 8 XCDs) runs one contiguous range of
// that order: with 8 K-splits an XCD owns one K range, so its slices of A and X come
// from HBM once, and neighbouring N tiles share each P panel in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int bid, int total) {
    return (total & 7) ? bid : (bid & 7) * (total >> 3) + (bid >> 3);
}

// One 128x128 FP32 output tile: C[m0.., n0..] (+ split offset) over K range [k0, k0+ksplit).
// smem: kF32TileLds bytes of LDS (16-byte aligned) from the calling kernel.
constexpr int kF32TileLds = (kGemmBM + kGemmBN) * (kGemmBK + 4) * 4;
__device__ __forceinline__ void gemm_f32_tile(const float* __restrict__ P, long ldp, const float* __restrict__ Q,
                                              long ldq, float* __restrict__ C, long ldc, long split_stride, int sp,
                                              long m0, long n0, long k0, int ksplit, uint8_t* smem) {
    auto Ps = (float(*)[kGemmBK + 4])smem;
    auto Qs = (float(*)[kGemmBK + 4])(smem + kGemmBM * (kGemmBK + 4) * 4);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lrow = tid >> 3, lc4 = (tid & 7) * 4;
    const float* Pg = P + (m0 + lrow) * ldp + k0 + lc4;
    const float* Qg = Q + (n0 + lrow) * ldq + k0 + lc4;
    f4 pr[4], qr[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        pr[i] = *(const f4*)(Pg + (long)(32 * i) * ldp);
        qr[i] = *(const f4*)(Qg + (long)(32 * i) * ldq);
    }
    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};

    for (int kt = 0; kt < ksplit; kt += kGemmBK) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; i++) {
            *(f4*)&Ps[lrow + 32 * i][lc4] = pr[i];
            *(f4*)&Qs[lrow + 32 * i][lc4] = qr[i];
        }
        __syncthreads();
        if (kt + kGemmBK < ksplit) {
            Pg += kGemmBK;
            Qg += kGemmBK;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                pr[i] = *(const f4*)(Pg + (long)(32 * i) * ldp);
                qr[i] = *(const f4*)(Qg + (long)(32 * i) * ldq);
            }
        }
#pragma unroll
        for (int kk = 0; kk < kGemmBK; kk += 4) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = Ps[wm * 64 + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = Qs[wn * 64 + j * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    float* Cb = C + (long)sp * split_stride;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long row = m0 + wm * 64 + i * 16 + 4 * (lane >> 4);
            const long col = n0 + wn * 64 + j * 16 + (lane & 15);
            *(f4*)(Cb + row + col * ldc) = acc[i][j];
        }
}

__global__ __launch_bounds__(256) void gemm_tn_f32_kernel(const float* __restrict__ P, long ldp,
                                                          const float* __restrict__ Q, long ldq,
                                                          float* __restrict__ C, long ldc, long split_stride,
                                                          int tiles_m, int tiles_n, int splits, int ksplit,
                                                          const int* __restrict__ done,
    const int* __restrict__ skip) {
    if (done && *done) return;
    const int total = tiles_m * tiles_n * splits;
    const int t = xcd_remap(blockIdx.x, total);
    const int tn = t % tiles_n;
    const int rest = t / tiles_n;
    const int tm = rest % tiles_m;
    const int sp = rest / tiles_m;
    if (skip && skip[tn >> 1]) return;  // this 256-column group's A*X is the sparse gather
    __shared__ __attribute__((aligned(16))) uint8_t smem[kF32TileLds];
    gemm_f32_tile(P, ldp, Q, ldq, C, ldc, split_stride, sp, (long)tm * kGemmBM, (long)tn * kGemmBN,
                  (long)sp * ksplit, ksplit, smem);
}

// Narrow FP32 tile for batches of at most BN (32 or 64) SPs: 128 x BN outputs, wave w owns
// rows 32w .. 32w+31 and all BN columns.  A wide 128x128 tile would pad the batch to 128
// columns and multiply mostly zeros; at BN = 32 the product reads A once at 2*BN/4 = 16
// flop/byte, under the FP32 ridge, so it is bound by HBM.  Every output element sees the
// same 16x16x4 MFMA sequence over the same K range as in the wide tile, so both give the
// same bits.
template <int BN>
__device__ __forceinline__ void gemm_f32_narrow_tile(const float* __restrict__ P, long ldp,
                                                     const float* __restrict__ Q, long ldq, float* __restrict__ C,
                                                     long ldc, long split_stride, int sp, long m0, long k0, int ksplit,
                                                     uint8_t* smem) {
    constexpr int QI = BN / 32;  // 16-byte Q loads per thread per stage
    constexpr int NJ = BN / 16;  // 16-column MFMA blocks per wave
    auto Ps = (float(*)[kGemmBK + 4])smem;
    auto Qs = (float(*)[kGemmBK + 4])(smem + kGemmBM * (kGemmBK + 4) * 4);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lrow = tid >> 3, lc4 = (tid & 7) * 4;
    const float* Pg = P + (m0 + lrow) * ldp + k0 + lc4;
    const float* Qg = Q + (long)lrow * ldq + k0 + lc4;
    f4 pr[4], qr[QI];
#pragma unroll
    for (int i = 0; i < 4; i++) pr[i] = *(const f4*)(Pg + (long)(32 * i) * ldp);
#pragma unroll
    for (int i = 0; i < QI; i++) qr[i] = *(const f4*)(Qg + (long)(32 * i) * ldq);
    f4 acc[2][NJ];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};

    for (int kt = 0; kt < ksplit; kt += kGemmBK) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; i++) *(f4*)&Ps[lrow + 32 * i][lc4] = pr[i];
#pragma unroll
        for (int i = 0; i < QI; i++) *(f4*)&Qs[lrow + 32 * i][lc4] = qr[i];
        __syncthreads();
        if (kt + kGemmBK < ksplit) {
            Pg += kGemmBK;
            Qg += kGemmBK;
#pragma unroll
            for (int i = 0; i < 4; i++) pr[i] = *(const f4*)(Pg + (long)(32 * i) * ldp);
#pragma unroll
            for (int i = 0; i < QI; i++) qr[i] = *(const f4*)(Qg + (long)(32 * i) * ldq);
        }
#pragma unroll
        for (int kk = 0; kk < kGemmBK; kk += 4) {
            float a[2], b[NJ];
#pragma unroll
            for (int i = 0; i < 2; i++) a[i] = Ps[wave * 32 + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int j = 0; j < NJ; j++) b[j] = Qs[j * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < NJ; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    float* Cb = C + (long)sp * split_stride;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const long row = m0 + wave * 32 + i * 16 + 4 * (lane >> 4);
            const long col = j * 16 + (lane & 15);
            *(f4*)(Cb + row + col * ldc) = acc[i][j];
        }
}

template <int BN>
__global__ __launch_bounds__(256) void gemm_tn_f32_narrow_kernel(const float* __restrict__ P, long ldp,
                                                                 const float* __restrict__ Q, long ldq,
                                                                 float* __restrict__ C, long ldc, long split_stride,
                                                                 int tiles_m, int splits, int ksplit,
                                                                 const int* __restrict__ done,
                                                                 const int* __restrict__ skip) {
    if (done && *done) return;
    if (skip && skip[0]) return;  // the batch (one 256-SP group) is the sparse gather's
    const int t = xcd_remap(blockIdx.x, tiles_m * splits);
    const int tm = t % tiles_m, sp = t / tiles_m;
    __shared__ __attribute__((aligned(16))) uint8_t smem[(kGemmBM + BN) * (kGemmBK + 4) * 4];
    gemm_f32_narrow_tile<BN>(P, ldp, Q, ldq, C, ldc, split_stride, sp, (long)tm * kGemmBM, (long)sp * ksplit, ksplit,
                             smem);
}

// FP32 C[:, 0..N) (+ split offsets) = P^T Q for N <= 64 columns: the narrow tile
hipError_t launch_gemm_f32_narrow(const float* P, long ldp, const float* Q, long ldq, float* C, long ldc,
                                  long split_stride, int M, int N, int K, int splits, const int* done,
                                  const int* skip, hipStream_t st) {
    if (M % kGemmBM || N < 1 || N > 64 || splits < 1 || K % (kGemmBK * splits)) return hipErrorInvalidValue;
    const int tiles_m = M / kGemmBM, ksplit = K / splits;
    const int grid = tiles_m * splits;
    if (N <= 32)
        gemm_tn_f32_narrow_kernel<32><<<grid, 256, 0, st>>>(P, ldp, Q, ldq, C, ldc, split_stride, tiles_m, splits,
                                                            ksplit, done, skip);
    else
        gemm_tn_f32_narrow_kernel<64><<<grid, 256, 0, st>>>(P, ldp, Q, ldq, C, ldc, split_stride, tiles_m, splits,
                                                            ksplit, done, skip);
    return hipGetLastError();
}

// 16-bit operand products (operand parts K-blocked, in bf16 unless noted):
//   PREC 1  bf16x3: P = hi + lo, Q = hi + lo; Phi*Qlo + Plo*Qhi + Phi*Qhi
//   PREC 2  fp16:   Phi*Qhi in fp16
// Small terms are accumulated first.  One 16x16x32 MFMA covers a whole 32-deep K stage;
// lane l supplies row l%16 and k = 8*(l/16) .. +7 of each operand.
constexpr int tile16_parts_p(int prec) { return prec == 1 ? 2 : 1; }
constexpr int tile16_parts_q(int prec) { return prec == 1 ? 2 : 1; }
// LDS bytes of one 128x128 16-bit tile (80-byte rows)
constexpr int tile16_lds(int prec) {
    return (tile16_parts_p(prec) * kGemmBM + tile16_parts_q(prec) * kGemmBN) * (kGemmBK + 8) * 2;
}
template <int PREC>
__device__ __forceinline__ void gemm16_tile(const uint16_t* __restrict__ Phi, const uint16_t* __restrict__ Plo, long ldp,
                                            const uint16_t* __restrict__ Qhi, const uint16_t* __restrict__ Qlo,
                                            long ldq, float* __restrict__ C, long ldc, long split_stride, int sp,
                                            long m0, long n0, long k0, int ksplit, uint8_t* smem) {
    constexpr int NPP = tile16_parts_p(PREC), NPQ = tile16_parts_q(PREC);
    constexpr int LDK = kGemmBK + 8;  // 80-byte rows: 16-byte aligned

    auto Ps = (uint16_t(*)[kGemmBM][LDK])smem;
    auto Qs = (uint16_t(*)[kGemmBN][LDK])(smem + NPP * kGemmBM * LDK * 2);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // a 128 x 32 tile of 16-bit values is 512 chunks of 16 bytes: 2 per thread
    const int lrow = tid >> 2, lc8 = (tid & 3) * 8;
    const uint16_t* Pg[NPP];
    const uint16_t* Qg[NPQ];
    // K-blocked operands: ldp / ldq are their row counts; one K stage advances 32 rows' worth
    const long pstep = ldp * kGemmBK, qstep = ldq * kGemmBK;
    const uint16_t* pbase[2] = {Phi, Plo};
    const uint16_t* qbase[2] = {Qhi, Qlo};
#pragma unroll
    for (int p = 0; p < NPP; p++) Pg[p] = pbase[p] + ((k0 >> 5) * ldp + m0 + lrow) * 32 + lc8;
#pragma unroll
    for (int p = 0; p < NPQ; p++) Qg[p] = qbase[p] + ((k0 >> 5) * ldq + n0 + lrow) * 32 + lc8;
    u4 pr[NPP][2], qr[NPQ][2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
#pragma unroll
        for (int p = 0; p < NPP; p++) pr[p][i] = *(const u4*)(Pg[p] + 64 * 32 * i);
#pragma unroll
        for (int p = 0; p < NPQ; p++) qr[p][i] = *(const u4*)(Qg[p] + 64 * 32 * i);
    }
    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};

    for (int kt = 0; kt < ksplit; kt += kGemmBK) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; i++) {
#pragma unroll
            for (int p = 0; p < NPP; p++) *(u4*)&Ps[p][lrow + 64 * i][lc8] = pr[p][i];
#pragma unroll
            for (int p = 0; p < NPQ; p++) *(u4*)&Qs[p][lrow + 64 * i][lc8] = qr[p][i];
        }
        __syncthreads();
        if (kt + kGemmBK < ksplit) {
#pragma unroll
            for (int p = 0; p < NPP; p++) {
                Pg[p] += pstep;
#pragma unroll
                for (int i = 0; i < 2; i++) pr[p][i] = *(const u4*)(Pg[p] + 64 * 32 * i);
            }
#pragma unroll
            for (int p = 0; p < NPQ; p++) {
                Qg[p] += qstep;
#pragma unroll
                for (int i = 0; i < 2; i++) qr[p][i] = *(const u4*)(Qg[p] + 64 * 32 * i);
            }
        }
        const int kof = 8 * (lane >> 4);
        u4 a[NPP][4], b[NPQ][4];
#pragma unroll
        for (int p = 0; p < NPP; p++)
#pragma unroll
            for (int i = 0; i < 4; i++) a[p][i] = *(const u4*)&Ps[p][wm * 64 + i * 16 + (lane & 15)][kof];
#pragma unroll
        for (int p = 0; p < NPQ; p++)
#pragma unroll
            for (int j = 0; j < 4; j++) b[p][j] = *(const u4*)&Qs[p][wn * 64 + j * 16 + (lane & 15)][kof];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if constexpr (PREC == 1) {
                    // small terms first, then the leading product
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf8, a[0][i]), __builtin_bit_cast(bf8, b[1][j]), acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf8, a[1][i]), __builtin_bit_cast(bf8, b[0][j]), acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf8, a[0][i]), __builtin_bit_cast(bf8, b[0][j]), acc[i][j], 0, 0, 0);
                } else {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                        __builtin_bit_cast(h8, a[0][i]), __builtin_bit_cast(h8, b[0][j]), acc[i][j], 0, 0, 0);
                }
            }
    }
    float* Cb = C + (long)sp * split_stride;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long row = m0 + wm * 64 + i * 16 + 4 * (lane >> 4);
            const long col = n0 + wn * 64 + j * 16 + (lane & 15);
            *(f4*)(Cb + row + col * ldc) = acc[i][j];
        }
}

template <int PREC>
__global__ __launch_bounds__(256) void gemm_tn_16_kernel(const uint16_t* __restrict__ Phi,
                                                         const uint16_t* __restrict__ Plo, long ldp,
                                                         const uint16_t* __restrict__ Qhi,
                                                         const uint16_t* __restrict__ Qlo, long ldq,
                                                         float* __restrict__ C, long ldc, long split_stride,
                                                         int tiles_m, int tiles_n, int splits, int ksplit,
                                                         const int* __restrict__ done,
    const int* __restrict__ skip) {
    if (done && *done) return;
    const int total = tiles_m * tiles_n * splits;
    const int t = xcd_remap(blockIdx.x, total);
    const int tn = t % tiles_n;
    const int rest = t / tiles_n;
    const int tm = rest % tiles_m;
    const int sp = rest / tiles_m;
    if (skip && skip[tn >> 1]) return;  // this 256-column group's A*X is the sparse gather
    __shared__ __attribute__((aligned(16))) uint8_t smem[tile16_lds(PREC)];
    gemm16_tile<PREC>(Phi, Plo, ldp, Qhi, Qlo, ldq, C, ldc, split_stride, sp, (long)tm * kGemmBM,
                      (long)tn * kGemmBN, (long)sp * ksplit, ksplit, smem);
}

// Pipelined large-tile variant: 256x256 output tile per workgroup, 8 waves as 2 (M) x 4 (N),
// each wave 128 x 64 (8 x 4 MFMA blocks, 128 accumulators).  Stages are filled by direct
// global->LDS loads (global_load_lds_dwordx4: each lane's 16 bytes land at the wave's LDS
// base + lane*16) into a ring of NSTAGE buffers, so NSTAGE-1 stages of loads are in flight
// while one is computed and no registers stage the data; the K-blocked operands make each
// wave's 16-row load one contiguous KiB.  LDS rows are 64 bytes unpadded; the 16-byte chunks
// of row r are XOR-swizzled by (r>>1)&3, so the 16 rows x 4 chunks a fragment read touches
// fall in distinct banks.  One barrier per stage: each wave first waits (vmcnt) for its own
// loads of stage k, so after the barrier every wave's loads of stage k have landed and every
// wave has finished reading stage k-1, whose buffer the loads of stage k+NSTAGE-1 then reuse.
// The barrier is the bare s_barrier: __syncthreads() adds a workgroup release fence, which
// the compiler implements by waiting for ALL outstanding loads (vmcnt(0)), including the
// later stages still in flight, and that would cut the lookahead to one stage.  The explicit
// per-wave wait above provides the ordering the fence would.  Each wave reads all its A and B
// fragments of the stage first, then issues the stage's MFMAs, so LDS latency is paid once
// per stage and the MFMAs of the two waves sharing a SIMD overlap the other's reads.
__device__ __forceinline__ void wait_vmcnt(int n) {
    // s_waitcnt with vmcnt = n (6 bits) and expcnt/lgkmcnt left at their maxima
    switch (n) {
        case 0: __builtin_amdgcn_s_waitcnt(0x0F70); break;
        case 4: __builtin_amdgcn_s_waitcnt(0x0F74); break;
        case 8: __builtin_amdgcn_s_waitcnt(0x0F78); break;
        case 16: __builtin_amdgcn_s_waitcnt(0x4F70); break;
        default: __builtin_amdgcn_s_waitcnt(0x0F70); break;
    }
}

template <int PREC, int NSTAGE>
__global__ __launch_bounds__(512, 1) void gemm_tn_16_pipe_kernel(const uint16_t* __restrict__ Phi,
                                                                 const uint16_t* __restrict__ Plo, long ldp,
                                                                 const uint16_t* __restrict__ Qhi,
                                                                 const uint16_t* __restrict__ Qlo, long ldq,
                                                                 float* __restrict__ C, long ldc, long split_stride,
                                                                 int tiles_m, int tiles_n, int splits, int ksplit,
                                                                 const int* __restrict__ done,
    const int* __restrict__ skip) {
    if (done && *done) return;
    constexpr int NPP = tile16_parts_p(PREC), NPQ = tile16_parts_q(PREC), NPT = NPP + NPQ;
    constexpr int BM = 256, BN = 256, BK = kGemmBK;
    constexpr int ROWB = BK * 2;          // bytes per LDS row
    constexpr int OPB = BM * ROWB;        // bytes per operand part per stage (16 KiB)
    constexpr int STB = NPT * OPB;        // bytes per stage: P parts, then Q parts
    constexpr int LPS = NPT * 2;          // global_load_lds instructions per wave per stage
    const int total = tiles_m * tiles_n * splits;
    const int t = xcd_remap(blockIdx.x, total);
    const int tn = t % tiles_n;
    const int rest = t / tiles_n;
    const int tm = rest % tiles_m;
    const int sp = rest / tiles_m;
    if (skip && skip[tn]) return;  // this 256-column group's A*X is the sparse gather
    const long m0 = (long)tm * BM, n0 = (long)tn * BN, k0 = (long)sp * ksplit;

    __shared__ __attribute__((aligned(1024))) uint8_t lds[NSTAGE * STB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 2, wn = wave & 3;
    const int nk = ksplit / BK;

    // loader: wave w fills rows [32w, 32w+32) of every operand part, 16 rows per instruction
    const int lrow = lane >> 2;
    const int lslot = lane & 3;
    // part t < NPP is P part t, the rest Q parts
    const uint16_t* src[NPT];
    {
        const uint16_t* pbase[2] = {Phi, Plo};
        const uint16_t* qbase[2] = {Qhi, Qlo};
#pragma unroll
        for (int t = 0; t < NPT; t++) src[t] = t < NPP ? pbase[t] : qbase[t - NPP];
    }
    // K-blocked operands with ldp / ldq rows: one K stage of an operand is rows*32 elements
    const long stage_elems[2] = {ldp * BK, ldq * BK};
    long rowoff[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int row = 32 * wave + 16 * i + lrow;
        const int g = lslot ^ ((row >> 1) & 3);
        rowoff[0][i] = ((k0 >> 5) * ldp + m0 + row) * 32 + g * 8;
        rowoff[1][i] = ((k0 >> 5) * ldq + n0 + row) * 32 + g * 8;
    }
    auto issue = [&](int ks, int buf) {
#pragma unroll
        for (int t = 0; t < NPT; t++)
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int op = t < NPP ? 0 : 1;
                uint8_t* dst = lds + buf * STB + t * OPB + (32 * wave + 16 * i) * ROWB;
                __builtin_amdgcn_global_load_lds((const void*)(src[t] + rowoff[op][i] + (long)ks * stage_elems[op]),
                                                 (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
            }
    };
    f4 acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};

    const int fchunk = lane >> 4;  // logical 16-byte chunk a lane reads
#pragma unroll
    for (int s0 = 0; s0 < NSTAGE - 1; s0++)
        if (s0 < nk) issue(s0, s0);
    for (int k = 0; k < nk; k++) {
        if (k + NSTAGE - 2 < nk) wait_vmcnt((NSTAGE - 2) * LPS);
        else wait_vmcnt(0);
        __builtin_amdgcn_s_barrier();
        if (k + NSTAGE - 1 < nk) issue(k + NSTAGE - 1, (k + NSTAGE - 1) % NSTAGE);
        const uint8_t* sb = lds + (k % NSTAGE) * STB;
        u4 b[NPQ][4];
#pragma unroll
        for (int p = 0; p < NPQ; p++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int r = wn * 64 + j * 16 + (lane & 15);
                b[p][j] = *(const u4*)(sb + (NPP + p) * OPB + r * ROWB + ((fchunk ^ ((r >> 1) & 3)) << 4));
            }
        {
            // every A fragment of the stage is read first; the MFMAs then issue in two groups kept
            // apart by scheduling barriers, so the compiler neither sinks the reads between the
            // MFMAs nor interleaves the groups
            u4 afr[8][NPP];
            auto load_a = [&](int i) {
                const int r = wm * 128 + i * 16 + (lane & 15);
#pragma unroll
                for (int p = 0; p < NPP; p++)
                    afr[i][p] = *(const u4*)(sb + p * OPB + r * ROWB + ((fchunk ^ ((r >> 1) & 3)) << 4));
            };
            auto mma = [&](int i) {
                const u4* a = afr[i];
                if constexpr (PREC == 1) {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf8, a[0]), __builtin_bit_cast(bf8, b[1][j]), acc[i][j], 0, 0, 0);
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf8, a[1]), __builtin_bit_cast(bf8, b[0][j]), acc[i][j], 0, 0, 0);
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf8, a[0]), __builtin_bit_cast(bf8, b[0][j]), acc[i][j], 0, 0, 0);
                } else {
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                            __builtin_bit_cast(h8, a[0]), __builtin_bit_cast(h8, b[0][j]), acc[i][j], 0, 0, 0);
                }
            };
#pragma unroll
            for (int i = 0; i < 8; i++) load_a(i);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 0; i < 4; i++) mma(i);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int i = 4; i < 8; i++) mma(i);
        }
    }
    float* Cb = C + (long)sp * split_stride;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long row = m0 + wm * 128 + i * 16 + 4 * (lane >> 4);
            const long col = n0 + wn * 64 + j * 16 + (lane & 15);
            *(f4*)(Cb + row + col * ldc) = acc[i][j];
        }
}

int gemm_big_tiles(int precision, int M, int N) { return precision != 0 && M % 256 == 0 && N % 256 == 0; }

hipError_t launch_gemm_tn(int precision, const float* P32, const uint16_t* Phi, const uint16_t* Plo, long ldp,
                          const float* Q32, const uint16_t* Qhi, const uint16_t* Qlo, long ldq, float* C, long ldc,
                          long split_stride, int M, int N, int K, int splits, int big, const int* done,
                          const int* skip, hipStream_t st) {
    if (M % kGemmBM || N % kGemmBN || splits < 1 || K % (kGemmBK * splits)) return hipErrorInvalidValue;
    if (precision < 0 || precision > 2) return hipErrorInvalidValue;
    if (big && !gemm_big_tiles(precision, M, N)) return hipErrorInvalidValue;
    if (big) {
        const int tiles_m = M / 256, tiles_n = N / 256, ksplit = K / splits;
        const int grid = tiles_m * tiles_n * splits;
        if (precision == 1)
            gemm_tn_16_pipe_kernel<1, 2><<<grid, 512, 0, st>>>(Phi, Plo, ldp, Qhi, Qlo, ldq, C, ldc, split_stride,
                                                               tiles_m, tiles_n, splits, ksplit, done, skip);
        else
            gemm_tn_16_pipe_kernel<2, 4><<<grid, 512, 0, st>>>(Phi, Plo, ldp, Qhi, Qlo, ldq, C, ldc, split_stride,
                                                               tiles_m, tiles_n, splits, ksplit, done, skip);
        return hipGetLastError();
    }
    const int tiles_m = M / kGemmBM, tiles_n = N / kGemmBN, ksplit = K / splits;
    const int grid = tiles_m * tiles_n * splits;
    if (precision == 0) {
        gemm_tn_f32_kernel<<<grid, 256, 0, st>>>(P32, ldp, Q32, ldq, C, ldc, split_stride, tiles_m, tiles_n, splits,
                                                 ksplit, done, skip);
    } else if (precision == 1) {
        gemm_tn_16_kernel<1><<<grid, 256, 0, st>>>(Phi, Plo, ldp, Qhi, Qlo, ldq, C, ldc, split_stride, tiles_m,
                                                   tiles_n, splits, ksplit, done, skip);
    } else {
        gemm_tn_16_kernel<2><<<grid, 256, 0, st>>>(Phi, Plo, ldp, Qhi, Qlo, ldq, C, ldc, split_stride, tiles_m,
                                                   tiles_n, splits, ksplit, done, skip);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- exact A^T * sgn(z)
// In the L1-variance formulations z = sgn(A x) holds -1, 0 and +1, so v = A^T z is a signed
// sum of entries of A.  Column j of A is held as a 25-bit fixed-point number under a
// power-of-two scale s_j with max_i |A_ij| <= (127/128) s_j:
//   A_ij = s_j (2^-7 q0 + 2^-13 q1 + 2^-19 q2 + 2^-25 q3) + e,   |e| <= 2^-26 s_j,
// with int8 pieces q0 in [-127, 127] and q1, q2, q3 in [-32, 32], each the rounding of what
// the previous ones leave.  The i8 MFMA multiplies and accumulates in int32 exactly; two
// accumulators take
//   accA = sum_i (64 q0 + q1) z_i   and   accB = sum_i (64 q2 + q3) z_i,
// the factor 64 entering through a second sign operand 64 z (z's bytes shifted in registers),
// and v_j = s_j (2^-13 accA + 2^-25 accB) is formed in FP64 and rounded once to FP32.
// |accA| <= 8160 m < 2^31 for m < 2^18.  Integer sums do not depend on their order, so every
// tiling (solo, fused) gives the same bits; the only rounding left against the FP32 master
// is e, a 2^-26 s_j per entry, where an FP32 accumulation of the same sum rounds at
// 2^-24 |partial sum| per addition.
//
// int8 operands are K-blocked by 64: element (r, k) of an operand with R rows sits at byte
// ((k / 64) * R + r) * 64 + k % 64, so one 64-deep K stage of a row range is one contiguous
// run of 64-byte rows (the 16-bit operands' layout, byte for byte).
typedef int i4 __attribute__((ext_vector_type(4)));
constexpr int kFixBK = 64;
__device__ __forceinline__ long fix8_index(int r, int k, int R) { return ((long)(k >> 6) * R + r) * 64 + (k & 63); }

// One workgroup per column j of A (column major, mp rows, zero padded): s_j, then the four
// pieces of every entry into q (piece p at q + p * mp * np, K-blocked with np rows).
__global__ __launch_bounds__(256) void to_fix8_kernel(const float* __restrict__ A, int mp, int np,
                                                      int8_t* __restrict__ q, float* __restrict__ qscale) {
    __shared__ double red[4];
    const int j = blockIdx.x;
    const float* col = A + (long)j * mp;
    float mx = 0.f;
    for (int i = threadIdx.x; i < mp; i += 256) mx = fmaxf(mx, fabsf(col[i]));
    mx = block_max<256>(mx, red);
    double sc = 1.0;
    if (mx > 0.f) {
        int e;
        frexp((double)mx * (128.0 / 127.0), &e);  // 2^e > mx * 128/127
        sc = ldexp(1.0, e);
    }
    if (threadIdx.x == 0) qscale[j] = (float)sc;
    const double inv = 1.0 / sc;  // a power of two: every step below is exact in FP64
    const long piece = (long)mp * np;
    for (int i = threadIdx.x; i < mp; i += 256) {
        double u = (double)col[i] * inv * 128.0;
        const long o = fix8_index(j, i, np);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const double r = rint(u);
            q[p * piece + o] = (int8_t)(int)r;
            u = (u - r) * 64.0;
        }
    }
}
hipError_t launch_to_fix8(const float* A, int mp, int np, int8_t* q, float* qscale, hipStream_t st) {
    if (mp <= 0 || np <= 0 || mp % kFixBK) return hipErrorInvalidValue;
    to_fix8_kernel<<<np, 256, 0, st>>>(A, mp, np, q, qscale);
    return hipGetLastError();
}

// LDS of one 128x128 fixed-point tile: the four pieces and the sign operand, 128 rows of 64
// bytes each, the 16-byte chunks of row r XOR-swizzled by (r >> 1) & 3.
constexpr int kFixPart = kGemmBM * kFixBK;
constexpr int kFixLds = 5 * kFixPart;
__device__ __forceinline__ int fix8_lds(int r, int chunk) { return r * kFixBK + ((chunk ^ ((r >> 1) & 3)) << 4); }

// The MFMAs of one 64-deep K stage held in LDS (the four pieces, then the signs; fix8_lds
// layout): wave (wm, wn) of the 128x128 tile accumulates its 64x64 block.  One 16x16x64 MFMA
// covers a whole stage of a 16x16 block, lane l supplying row l%16 and the 16 bytes of chunk
// l/16 of each operand (the same chunk of both operands, so the products pair equal k).
__device__ __forceinline__ void fix8_stage(const uint8_t* sb, i4 (&accA)[4][4], i4 (&accB)[4][4], int wm, int wn,
                                           int lane) {
    const int fc = lane >> 4;
    i4 b[4], b64[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int r = wn * 64 + j * 16 + (lane & 15);
        const u4 v = *(const u4*)(sb + 4 * kFixPart + fix8_lds(r, fc));
        // 64 z per byte: bits 0-1 of each byte (z in {-1, 0, 1}) moved to bits 6-7
        const u4 v64 = (v << 6u) & 0xC0C0C0C0u;
        b[j] = __builtin_bit_cast(i4, v);
        b64[j] = __builtin_bit_cast(i4, v64);
    }
    // piece p + 1's fragments are read while piece p's MFMAs run (two fragment sets)
    i4 a[2][4];
    auto frags = [&](int p, i4 (&d)[4]) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = wm * 64 + i * 16 + (lane & 15);
            d[i] = __builtin_bit_cast(i4, *(const u4*)(sb + p * kFixPart + fix8_lds(r, fc)));
        }
    };
    frags(0, a[0]);
#pragma unroll
    for (int p = 0; p < 4; p++) {
        if (p < 3) frags(p + 1, a[(p + 1) & 1]);
        const i4(&cur)[4] = a[p & 1];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (p == 0) accA[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cur[i], b64[j], accA[i][j], 0, 0, 0);
                else if (p == 1) accA[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cur[i], b[j], accA[i][j], 0, 0, 0);
                else if (p == 2) accB[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cur[i], b64[j], accB[i][j], 0, 0, 0);
                else accB[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cur[i], b[j], accB[i][j], 0, 0, 0);
            }
    }
}

// V[r + c*ldc] = s_r (2^-13 accA + 2^-25 accB) for the wave's 64x64 block of the tile at (m0, n0).
__device__ __forceinline__ void fix8_store(const i4 (&accA)[4][4], const i4 (&accB)[4][4],
                                           const float* __restrict__ qscale, float* __restrict__ C, long ldc,
                                           long m0, long n0, int wm, int wn, int lane) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const long row = m0 + wm * 64 + i * 16 + 4 * (lane >> 4);
        const f4 sc = *(const f4*)(qscale + row);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long col = n0 + wn * 64 + j * 16 + (lane & 15);
            f4 o;
#pragma unroll
            for (int e = 0; e < 4; e++)
                o[e] = (float)((double)sc[e] * ((double)accA[i][j][e] * 0x1p-13 + (double)accB[i][j][e] * 0x1p-25));
            *(f4*)(C + row + col * ldc) = o;
        }
    }
}

// One 128x128 tile of V = s (2^-13 accA + 2^-25 accB).  P: the pieces (stride `piece` bytes,
// K-blocked with ldp rows), Q: int8 signs (K-blocked with ldq rows).  256 threads as 2x2 waves
// of 64x64.  Each stage is loaded between two barriers and then computed, so no load
// registers stay live beside the 128 accumulator registers; two workgroups per CU overlap
// one's loads with the other's MFMAs.  The fused iteration's tiles; solo solves run the
// pipelined kernel below (same sums, so the same bits).
__device__ __forceinline__ void gemm_fix8_tile(const int8_t* __restrict__ P, long piece,
                                               const float* __restrict__ qscale, long ldp,
                                               const int8_t* __restrict__ Q, long ldq, float* __restrict__ C,
                                               long ldc, long m0, long n0, int K, uint8_t* smem) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // loader: a 128-row stage of one part is 512 chunks of 16 bytes, 2 per thread
    const int lrow = tid >> 2, lslot = tid & 3;
    const long pst = ldp * kFixBK, qst = ldq * kFixBK;  // bytes per K stage
    const int8_t* pg = P + (m0 + lrow) * kFixBK + lslot * 16;
    const int8_t* qg = Q + (n0 + lrow) * kFixBK + lslot * 16;
    i4 accA[4][4], accB[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            accA[i][j] = (i4){0, 0, 0, 0};
            accB[i][j] = (i4){0, 0, 0, 0};
        }
    const int nk = K / kFixBK;
    for (int ks = 0; ks < nk; ks++) {
        __syncthreads();
        {
            u4 pr[4][2], qr[2];
#pragma unroll
            for (int p = 0; p < 4; p++)
#pragma unroll
                for (int i = 0; i < 2; i++) pr[p][i] = *(const u4*)(pg + p * piece + ks * pst + 64 * kFixBK * i);
#pragma unroll
            for (int i = 0; i < 2; i++) qr[i] = *(const u4*)(qg + ks * qst + 64 * kFixBK * i);
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int r = lrow + 64 * i;
#pragma unroll
                for (int p = 0; p < 4; p++) *(u4*)(smem + p * kFixPart + fix8_lds(r, lslot)) = pr[p][i];
                *(u4*)(smem + 4 * kFixPart + fix8_lds(r, lslot)) = qr[i];
            }
        }
        __syncthreads();
        fix8_stage(smem, accA, accB, wm, wn, lane);
    }
    fix8_store(accA, accB, qscale, C, ldc, m0, n0, wm, wn, lane);
}

// Solo solves: the same 128x128 tile with a two-stage LDS ring filled by direct global->LDS
// loads (global_load_lds_dwordx4: lane l's 16 bytes land at the wave's LDS base + 16 l), so
// the loads of stage k+1 are in flight while stage k is computed.  Wave w loads rows
// [32w, 32w+32) of every part, 16 rows per instruction; each lane's source chunk is chosen so
// that the LDS rows come out in the fix8_lds (XOR-swizzled) layout.  Per stage: each wave
// waits for its own loads, then one barrier; after it every wave's loads of stage k have
// landed and every wave has finished reading stage k-1, whose buffer stage k+1 then reuses.
__global__ __launch_bounds__(256, 2) void gemm_tn_fix8_kernel(const int8_t* __restrict__ P, long piece,
                                                           const float* __restrict__ qscale, long ldp,
                                                           const int8_t* __restrict__ Q, long ldq,
                                                           float* __restrict__ C, long ldc, int tiles_m, int tiles_n,
                                                           int K, const int* __restrict__ done) {
    if (done && *done) return;
    const int t = xcd_remap(blockIdx.x, tiles_m * tiles_n);
    const int tn = t % tiles_n, tm = t / tiles_n;
    const long m0 = (long)tm * kGemmBM, n0 = (long)tn * kGemmBN;
    __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * kFixLds];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const long pst = ldp * kFixBK, qst = ldq * kFixBK;  // bytes per K stage
    long src[2][2];                                   // [P, Q][16-row group]: this lane's bytes in stage 0
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const int row = 32 * wave + 16 * i + (lane >> 2);
        const int g = (lane & 3) ^ ((row >> 1) & 3);
        src[0][i] = (m0 + row) * kFixBK + g * 16;
        src[1][i] = (n0 + row) * kFixBK + g * 16;
    }
    auto issue = [&](int ks, int buf) {
#pragma unroll
        for (int part = 0; part < 5; part++)
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int8_t* g = part < 4 ? P + part * piece + ks * pst + src[0][i] : Q + ks * qst + src[1][i];
                uint8_t* dst = lds + buf * kFixLds + part * kFixPart + (32 * wave + 16 * i) * kFixBK;
                __builtin_amdgcn_global_load_lds((const void*)g, (__attribute__((address_space(3))) void*)dst, 16, 0,
                                                 0);
            }
    };
    i4 accA[4][4], accB[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            accA[i][j] = (i4){0, 0, 0, 0};
            accB[i][j] = (i4){0, 0, 0, 0};
        }
    const int nk = K / kFixBK;
    issue(0, 0);
    for (int ks = 0; ks < nk; ks++) {
        wait_vmcnt(0);
        __builtin_amdgcn_s_barrier();
        if (ks + 1 < nk) issue(ks + 1, (ks + 1) & 1);
        fix8_stage(lds + (ks & 1) * kFixLds, accA, accB, wm, wn, lane);
    }
    fix8_store(accA, accB, qscale, C, ldc, m0, n0, wm, wn, lane);
}
hipError_t launch_gemm_fix8(const int8_t* P, const float* qscale, long ldp, const int8_t* Q, long ldq, float* C,
                            long ldc, int M, int N, int K, const int* done, hipStream_t st) {
    if (!P || !qscale || !Q || M % kGemmBM || N % kGemmBN || K % kFixBK || K <= 0 || K >= (1 << 18) || ldp < M ||
        ldq < N)
        return hipErrorInvalidValue;
    const int tiles_m = M / kGemmBM, tiles_n = N / kGemmBN;
    gemm_tn_fix8_kernel<<<tiles_m * tiles_n, 256, 0, st>>>(P, ldp * (long)K, qscale, ldp, Q, ldq, C, ldc, tiles_m,
                                                           tiles_n, K, done);
    return hipGetLastError();
}

// ---------------------------------------------------------------- sparse A*X
// z_j = A x_j over the nonzeros of x_j, for iterates that are sparse (after a constrained or
// penalised step): workgroup (row block, SP) walks the SP's nonzeros in index order (ballot
// compaction of 2048-index slabs, eight loads of x in flight per thread) and accumulates its
// 1024 rows in FP64 from the FP32 master
// A (column major, so each nonzero's column segment is one coalesced read; 4 rows per thread
// as one 16-byte load, eight nonzeros' loads in flight).  Runs for the SPs of each group of
// 256 whose flag the previous iteration's convergence kernel set; the dense product's tiles
// of those groups stand down.
// The FP64 sum in a fixed order is deterministic and more accurate than any product form.
constexpr int kSpRows = kSparseRows;  // 4 rows per thread
// idxs, vals: LDS of the calling kernel, kSlabIndices entries each
constexpr int kSlabIndices = 2048;
__device__ __forceinline__ void sparse_ax_body(int rb, int j, const float* __restrict__ A, int lda, int n,
                                               const float* __restrict__ X, int ldx, float* __restrict__ Z, int ldz,
                                               int mp, SpState sp, int* idxs, float* vals) {
    if (sp.done[0] || !sp.done[kSparseFlags + (j >> 8)] || !sp.active[j]) return;
    constexpr int kE = 8;                 // x entries per thread per slab, loads in flight
    constexpr int kSlab = 256 * kE;       // indices per slab
    static_assert(kSlab == kSlabIndices, "slab size");
    __shared__ int wcount[kE][4];
    const __amdgpu_buffer_rsrc_t xr = col_rsrc(X + (long)j * ldx, (uint32_t)n * 4u);  // reads 0 past n
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // each thread owns 4 consecutive rows: one 16-byte load per nonzero column
    const int row0 = rb * kSpRows + 4 * threadIdx.x;
    const bool in = row0 < mp;  // mp and the row block are multiples of 4 rows
    double acc[4] = {0, 0, 0, 0};
    for (int k0 = 0; k0 < n; k0 += kSlab) {
        // entry e of this thread is index k0 + e*256 + tid: coalesced loads, and the slab's index
        // order is (e, wave, lane), so a nonzero's list position is the count of nonzeros in
        // earlier rows e, then earlier waves of row e, then earlier lanes
        float xv[kE];
#pragma unroll
        for (int e = 0; e < kE; e++)
            xv[e] = __uint_as_float(
                __builtin_amdgcn_raw_buffer_load_b32(xr, (uint32_t)(k0 + e * 256 + (int)threadIdx.x) * 4u, 0, 0));
        uint64_t bal[kE];
#pragma unroll
        for (int e = 0; e < kE; e++) {
            bal[e] = __ballot(xv[e] != 0.f);
            if (lane == 0) wcount[e][w] = __popcll(bal[e]);
        }
        __syncthreads();
        int cnt = 0;
#pragma unroll
        for (int e = 0; e < kE; e++) {
            int before = cnt;  // nonzeros ahead of (e, w, lane 0)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                before += i < w ? wcount[e][i] : 0;
                cnt += wcount[e][i];
            }
            if (xv[e] != 0.f) {
                const int at = before + __popcll(bal[e] & ((1ull << lane) - 1ull));
                idxs[at] = k0 + e * 256 + (int)threadIdx.x;
                vals[at] = xv[e];
            }
        }
        __syncthreads();
        if (in) {
            // eight columns' loads in flight before their FMAs; the sum stays in index order
            int t = 0;
            for (; t + 8 <= cnt; t += 8) {
                f4 c[8];
#pragma unroll
                for (int u = 0; u < 8; u++) c[u] = *(const f4*)(A + (long)idxs[t + u] * lda + row0);
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const double v = vals[t + u];
#pragma unroll
                    for (int r = 0; r < 4; r++) acc[r] += (double)c[u][r] * v;
                }
            }
            for (; t < cnt; t++) {
                const f4 c = *(const f4*)(A + (long)idxs[t] * lda + row0);
                const double v = vals[t];
#pragma unroll
                for (int r = 0; r < 4; r++) acc[r] += (double)c[r] * v;
            }
        }
        __syncthreads();
    }
    if (in) {
        f4 o;
#pragma unroll
        for (int r = 0; r < 4; r++) o[r] = (float)acc[r];
        *(f4*)(Z + (long)j * ldz + row0) = o;
    }
}
__global__ __launch_bounds__(256) void sparse_ax_kernel(const float* __restrict__ A, int lda, int n,
                                                        const float* __restrict__ X, int ldx, float* __restrict__ Z,
                                                        int ldz, int mp, SpState sp) {
    __shared__ int idxs[kSlabIndices];
    __shared__ float vals[kSlabIndices];
    sparse_ax_body(blockIdx.x, blockIdx.y, A, lda, n, X, ldx, Z, ldz, mp, sp, idxs, vals);
}
hipError_t launch_sparse_ax(const float* Acm, int lda, int mp, int n, const float* X, int ldx, float* Z, int ldz,
                            int batch, SpState sp, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    if ((long)n * 4 + 8192 > 0xffffffffL) return hipErrorInvalidValue;  // 32-bit offsets into x
    sparse_ax_kernel<<<dim3((mp + kSpRows - 1) / kSpRows, batch), 256, 0, st>>>(Acm, lda, n, X, ldx, Z, ldz, mp, sp);
    return hipGetLastError();
}

// The same gather over nonzero lists built once per SP.  nz_list_kernel: one workgroup of 1024
// threads per gathered SP scans x in 8192-index slabs (ballot compaction, index order (e, wave,
// lane) = ascending) and writes the count, then up to kGatherList indices and values, to
// gl + j * kGatherStride.  sparse_ax_listed_kernel: workgroup (row block, SP) stages the list in
// LDS and walks it with sixteen columns' loads in flight, accumulating in the same ascending
// index order as sparse_ax_body (so the same bits); an SP with more than kGatherList nonzeros
// scans x itself as sparse_ax_body does.  Each x is read once instead of once per row block,
// and no slab barriers sit between a workgroup's column loads.
__device__ __forceinline__ void nz_list_body(int j, const float* __restrict__ X, int ldx, int n, SpState sp,
                                             int* __restrict__ gl) {
    if (sp.done[0] || !sp.done[kSparseFlags + (j >> 8)] || !sp.active[j]) return;
    constexpr int kE = 8, kSlab = 1024 * kE;
    __shared__ int wcount[kE][16];
    const __amdgpu_buffer_rsrc_t xr = col_rsrc(X + (long)j * ldx, (uint32_t)n * 4u);  // reads 0 past n
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int* out = gl + (long)j * kGatherStride;
    int* oidx = out + kGatherHead;
    int* oval = out + kGatherHead + kGatherList;
    int total = 0;
    // the grouping key of the grouped gather: the index of the largest |x_i| (the lowest such
    // index on ties), as (|x| bits << 32) | ~index maximised
    uint64_t best = 0;
    for (int k0 = 0; k0 < n; k0 += kSlab) {
        float xv[kE];
#pragma unroll
        for (int e = 0; e < kE; e++)
            xv[e] = __uint_as_float(
                __builtin_amdgcn_raw_buffer_load_b32(xr, (uint32_t)(k0 + e * 1024 + (int)threadIdx.x) * 4u, 0, 0));
        uint64_t bal[kE];
#pragma unroll
        for (int e = 0; e < kE; e++) {
            bal[e] = __ballot(xv[e] != 0.f);
            if (lane == 0) wcount[e][w] = __popcll(bal[e]);
        }
        __syncthreads();
        int cnt = 0;
#pragma unroll
        for (int e = 0; e < kE; e++) {
            int before = cnt;  // nonzeros ahead of (e, w, lane 0)
#pragma unroll
            for (int i = 0; i < 16; i++) {
                before += i < w ? wcount[e][i] : 0;
                cnt += wcount[e][i];
            }
            if (xv[e] != 0.f) {
                const int idx = k0 + e * 1024 + (int)threadIdx.x;
                const int at = total + before + __popcll(bal[e] & ((1ull << lane) - 1ull));
                if (at < kGatherList) {
                    oidx[at] = idx;
                    oval[at] = __float_as_int(xv[e]);
                }
                const uint64_t key = ((uint64_t)__float_as_uint(fabsf(xv[e])) << 32) | (uint32_t)~idx;
                best = key > best ? key : best;
            }
        }
        total += cnt;
        __syncthreads();  // wcount is rewritten by the next slab
    }
    __shared__ uint64_t wbest[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)best, o, 64), hi = __shfl_xor((uint32_t)(best >> 32), o, 64);
        const uint64_t other = ((uint64_t)hi << 32) | lo;
        best = other > best ? other : best;
    }
    if (lane == 0) wbest[w] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < 16; i++) best = wbest[i] > best ? wbest[i] : best;
        out[0] = total;
        out[1] = best ? (int)~(uint32_t)best : 0x7fffffff;  // no nonzero: after the others
    }
}
__global__ __launch_bounds__(1024) void nz_list_kernel(const float* __restrict__ X, int ldx, int n, SpState sp,
                                                       int* __restrict__ gl) {
    nz_list_body(blockIdx.x, X, ldx, n, sp, gl);
}

// sidx, sval: LDS of the calling kernel, kGatherList entries each (shared with the slab gather)
__device__ __forceinline__ void sparse_ax_listed_body(int rb, int j, const float* __restrict__ A, int lda, int n,
                                                      const float* __restrict__ X, int ldx, float* __restrict__ Z,
                                                      int ldz, int mp, SpState sp, const int* __restrict__ gl,
                                                      int* sidx, float* sval) {
    static_assert(kGatherList >= kSlabIndices, "the slab gather runs in the list's LDS");
    if (sp.done[0] || !sp.done[kSparseFlags + (j >> 8)] || !sp.active[j]) return;
    const int* L = gl + (long)j * kGatherStride;
    const int cnt = L[0];
    if (cnt > kGatherList) {  // longer than the list: scan x as the slab gather does
        sparse_ax_body(rb, j, A, lda, n, X, ldx, Z, ldz, mp, sp, sidx, sval);
        return;
    }
    for (int t = threadIdx.x; t < cnt; t += 256) {
        sidx[t] = L[kGatherHead + t];
        sval[t] = __int_as_float(L[kGatherHead + kGatherList + t]);
    }
    __syncthreads();
    const int row0 = rb * kSpRows + 4 * threadIdx.x;
    if (row0 >= mp) return;  // mp and the row block are multiples of 4 rows
    double acc[4] = {0, 0, 0, 0};
    constexpr int kU = 16;  // columns' loads in flight before their FMAs; the sum stays in index order
    int t = 0;
    for (; t + kU <= cnt; t += kU) {
        f4 c[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) c[u] = *(const f4*)(A + (long)sidx[t + u] * lda + row0);
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const double v = sval[t + u];
#pragma unroll
            for (int r = 0; r < 4; r++) acc[r] += (double)c[u][r] * v;
        }
    }
    for (; t < cnt; t++) {
        const f4 c = *(const f4*)(A + (long)sidx[t] * lda + row0);
        const double v = sval[t];
#pragma unroll
        for (int r = 0; r < 4; r++) acc[r] += (double)c[r] * v;
    }
    f4 o;
#pragma unroll
    for (int r = 0; r < 4; r++) o[r] = (float)acc[r];
    *(f4*)(Z + (long)j * ldz + row0) = o;
}
__global__ __launch_bounds__(256) void sparse_ax_listed_kernel(const float* __restrict__ A, int lda, int n,
                                                               const float* __restrict__ X, int ldx,
                                                               float* __restrict__ Z, int ldz, int mp, SpState sp,
                                                               const int* __restrict__ gl) {
    __shared__ int sidx[kGatherList];
    __shared__ float sval[kGatherList];
    sparse_ax_listed_body(blockIdx.x, blockIdx.y, A, lda, n, X, ldx, Z, ldz, mp, sp, gl, sidx, sval);
}

// ---------------------------------------------------------------- z epilogue
__device__ __forceinline__ bool is_constrained(int f) { return f < 4; }
__device__ __forceinline__ bool is_l1_variance(int f) { return f == 1 || f == 3 || f == 5 || f == 7; }
__device__ __forceinline__ float sgnf(float v) { return (float)((v > 0.f) - (v < 0.f)); }

// Element i of column j: FP32 column major (ld), 16-bit forms K-blocked over op.bp columns,
// precision 4: the int8 sign operand of the fixed-point product (K-blocked by 64).
__device__ __forceinline__ void write_operand(Operand op, int precision, long ld, int i, int j, float v) {
    op.f32[(long)j * ld + i] = v;
    if (precision == 1) {
        const long b = blk_index(j, i, op.bp);
        const uint16_t h = f2bf(v);
        op.hi[b] = h;
        op.lo[b] = f2bf(v - bf2f(h));
    } else if (precision == 2) {
        op.hi[blk_index(j, i, op.bp)] = f2h(v);
    } else if (precision == 4) {
        ((int8_t*)op.hi)[fix8_index(j, i, op.bp)] = (int8_t)(int)v;  // sign vectors of the fixed-point product
    }
}

// sum_k a[k] * b[k] in FP64 over k < n (n % 4 == 0, 16-byte aligned rows) by a 256-thread
// block: 16-byte loads, eight in flight per thread, fixed summation order.
__device__ __forceinline__ double block_dot_f64(const float* __restrict__ a, const float* __restrict__ b, int n,
                                                double* red) {
    const int nq = n >> 2;
    const f4* a4 = (const f4*)a;
    const f4* b4 = (const f4*)b;
    double acc = 0;
    for (int q0 = threadIdx.x; q0 < nq; q0 += 256 * 8) {
        f4 va[8], vb[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int q = q0 + 256 * u;
            va[u] = q < nq ? a4[q] : (f4){0.f, 0.f, 0.f, 0.f};
            vb[u] = q < nq ? b4[q] : (f4){0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
#pragma unroll
            for (int e = 0; e < 4; e++) acc += (double)va[u][e] * (double)vb[u][e];
    }
    return block_sum<256>(acc, red);
}

// One workgroup per SP column.  z = sum of split partials (fixed order).  Constrained:
// L1 variance f = |z|_1 then z = sgn z; L2 variance f = |z|_2 (z unchanged)
// (c++/src/gpower/gpower_commons.h:75-101).  Penalised: z = sgn z or z/|z|_2 (:150-166).
// Exact signs (rf.eps > 0, L1 variance, z from the dense product): an entry with
// |z_i| <= eps * max_k |A_ik| * |x|_2, inside the product's rounding of zero, is recomputed as
// an FP64 dot product of row i of the FP32 master with x, so sgn z_i is the one exact
// arithmetic gives.  The flagged set does not depend on thread order; more than kRefineMax
// flags (degenerate data) leave the column as the product gave it.
__device__ __forceinline__ void prep_z_body(int j, const float* __restrict__ Zpart, int splits, long split_stride,
                                            int m, int ldz, Operand zop, int formulation, int precision,
                                            SpState sp, SignRefine rf) {
    if (*sp.done || !sp.active[j]) return;
    const bool gathered = sp.done[kSparseFlags + (j >> 8)] != 0;
    if (gathered) splits = 1;  // the sparse gather wrote all of z (FP64 sums) into partial 0
    __shared__ double red[4];
    __shared__ int nflag;
    __shared__ int flags[kRefineMax];
    const int tid = threadIdx.x;
    const float* zc = Zpart + (long)j * ldz;
    const bool l1v = is_l1_variance(formulation);
    const bool refine = l1v && rf.eps > 0.f && !gathered;
    const float* xj = rf.x + (long)j * rf.np;
    float thr = 0.f;
    if (refine) {
        thr = rf.eps * (float)sqrt(block_dot_f64(xj, xj, rf.np, red));
        if (tid == 0) nflag = 0;
        __syncthreads();
    }
    double acc = 0;
    for (int i = tid; i < m; i += 256) {
        float z = zc[i];
        for (int s = 1; s < splits; s++) z += zc[(long)s * split_stride + i];
        zop.f32[(long)j * ldz + i] = z;
        if (!l1v) acc += (double)z * (double)z;
        if (refine && fabsf(z) <= thr * rf.rowmax[i] && rf.rowmax[i] > 0.f) {
            const int slot = atomicAdd(&nflag, 1);
            if (slot < kRefineMax) flags[slot] = i;
        }
    }
    if (refine) {
        __syncthreads();
        const int nf = nflag;
        if (nf <= kRefineMax) {
            for (int f = 0; f < nf; f++) {
                const int i = flags[f];
                const double d = block_dot_f64(rf.arm + (long)i * rf.np, xj, rf.np, red);
                if (tid == 0) zop.f32[(long)j * ldz + i] = (float)d;
            }
            __syncthreads();
        }
    }
    const bool constrained = is_constrained(formulation);
    float inv = 0.f;
    if (!l1v) {
        const float stat = (float)sqrt(block_sum<256>(acc, red));
        if (tid == 0) sp.fz[j] = stat;
        inv = 1.0f / stat;
    }
    for (int i = tid; i < m; i += 256) {
        const long idx = (long)j * ldz + i;
        float z = zop.f32[idx];
        if (l1v) {
            acc += (double)fabsf(z);
            z = sgnf(z);
        } else if (!constrained) {
            z = z * inv;
        }
        write_operand(zop, precision, ldz, i, j, z);
    }
    if (l1v) {
        const double tot = block_sum<256>(acc, red);
        if (tid == 0) sp.fz[j] = (float)tot;
    }
}
__global__ __launch_bounds__(256) void prep_z_kernel(const float* __restrict__ Zpart, int splits,
                                                     long split_stride, int m, int ldz, Operand zop,
                                                     int formulation, int precision, SpState sp, SignRefine rf) {
    prep_z_body(blockIdx.x, Zpart, splits, split_stride, m, ldz, zop, formulation, precision, sp, rf);
}
// Partial 0 <- the sum of the K-split partials of z_j, in prep_z's order (partial 0, then 1,
// 2, ...: the same bits), over workgroups (row block, SP).  For small batches, where prep_z's
// one workgroup per SP would walk every partial of every row alone (a batch of 32 with 64
// splits: 0.35 ms, against ~0.03 ms here).  Gathered groups hold all of z in partial 0.
__device__ __forceinline__ void sum_partials_body(int rb, int j, float* __restrict__ Zpart, int splits,
                                                  long split_stride, int m, int ldz, SpState sp) {
    if (*sp.done || !sp.active[j] || sp.done[kSparseFlags + (j >> 8)]) return;
    float* zc = Zpart + (long)j * ldz;
    const int i = rb * 256 + threadIdx.x;
    if (i >= m) return;
    float z = zc[i];
#pragma unroll 8
    for (int s = 1; s < splits; s++) z += zc[(long)s * split_stride + i];
    zc[i] = z;
}
__global__ __launch_bounds__(256) void sum_partials_kernel(float* __restrict__ Zpart, int splits, long split_stride,
                                                           int m, int ldz, SpState sp) {
    sum_partials_body(blockIdx.x, blockIdx.y, Zpart, splits, split_stride, m, ldz, sp);
}

hipError_t launch_prep_z(float* Zpart, int splits, long split_stride, int m, int ldz, Operand zop, int formulation,
                         int precision, int batch, SpState sp, SignRefine rf, hipStream_t st) {
    if (rf.eps > 0.f && (!rf.arm || !rf.rowmax || !rf.x || rf.np <= 0)) return hipErrorInvalidValue;
    if (splits > 1 && batch <= 256) {
        sum_partials_kernel<<<dim3((m + 255) / 256, batch), 256, 0, st>>>(Zpart, splits, split_stride, m, ldz, sp);
        splits = 1;
    }
    prep_z_kernel<<<batch, 256, 0, st>>>(Zpart, splits, split_stride, m, ldz, zop, formulation, precision, sp, rf);
    return hipGetLastError();
}

// max_k |A_ik| of each row of A (mp x np column major): block (row block, column chunk),
// thread per row over the chunk's columns, then an integer atomic max of the bit patterns
// (monotone for non-negative floats), so the result does not depend on order.  out must
// hold zeros.
constexpr int kRowMaxChunk = 256;
__global__ __launch_bounds__(256) void row_max_kernel(const float* __restrict__ A, int mp, int np,
                                                      float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= mp) return;
    const long k0 = (long)blockIdx.y * kRowMaxChunk, k1 = min((long)np, k0 + kRowMaxChunk);
    float mx = 0.f;
    for (long k = k0; k < k1; k++) mx = fmaxf(mx, fabsf(A[k * mp + i]));
    atomicMax((unsigned int*)&out[i], __float_as_uint(mx));
}
hipError_t launch_row_max(const float* Acm, int mp, int np, float* out, hipStream_t st) {
    if (mp <= 0 || np <= 0) return hipErrorInvalidValue;
    hipError_t e = hipMemsetAsync(out, 0, sizeof(float) * (size_t)mp, st);
    if (e != hipSuccess) return e;
    row_max_kernel<<<dim3((mp + 255) / 256, (np + kRowMaxChunk - 1) / kRowMaxChunk), 256, 0, st>>>(Acm, mp, np, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------- union A*X
// z_j = A[:, U] x_j[U] for every SP of a batch, where U is the union of the supports of the
// iterates whose A*X would otherwise be the dense product: the K dimension of the dense FP32
// product shrinks from n to |U|.  Iterates that converge to a few supports share most of U
// (L1-constrained iterates hold thousands of nonzeros each, too many for the per-SP gather).
// Sums over U in FP32 on the MFMA, like the dense FP32 product, over the nonzero terms only.

// flag[k] = 1 iff x_j[k] != 0 for an SP j that iterates next and whose group is not gathered
__global__ __launch_bounds__(256) void union_flags_kernel(const float* __restrict__ X, int ldx, int n, int batch,
                                                          SpState sp, int* __restrict__ flag) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    // 16 SPs' loads in flight per step (a loop that waited on each load would serialise them)
    int any = 0;
    for (int j0 = 0; j0 < batch && !any; j0 += 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = j0 + u < batch ? X[(long)(j0 + u) * ldx + k] : 0.f;
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int j = j0 + u;
            // NaN counts as nonzero, as in the dense product
            if (j < batch && v[u] != 0.f && sp.active[j] && !sp.done[kSparseFlags + (j >> 8)]) any = 1;
        }
    }
    flag[k] = any;
}

// U = the flagged indices in ascending order; stat = {|U|, done[0], done[1]} for the host.
// One workgroup: thread t counts and then lists the flags of its own contiguous index range.
__global__ __launch_bounds__(1024) void union_list_kernel(const int* __restrict__ flag, int n, int* __restrict__ U,
                                                          const int* __restrict__ done, int* __restrict__ stat) {
    __shared__ int wsum[16];
    const int per = (n + 1023) / 1024;
    const int k0 = threadIdx.x * per, k1 = min(n, k0 + per);
    int c = 0;
    for (int k = k0; k < k1; k++) c += flag[k];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = c;
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int pos = incl - c, tot = 0;
    for (int i = 0; i < 16; i++) {
        pos += i < w ? wsum[i] : 0;
        tot += wsum[i];
    }
    for (int k = k0; k < k1; k++)
        if (flag[k]) U[pos++] = k;
    if (threadIdx.x == 0) {
        stat[0] = tot;
        stat[1] = done[0];
        stat[2] = done[1];
    }
}

// Xu row j (ld ldu) = x_j[U[0..K)], zero on [K, kpad); rows j < rows
__global__ __launch_bounds__(256) void union_gather_x_kernel(const float* __restrict__ X, int ldx,
                                                             const int* __restrict__ U, int K, int kpad,
                                                             float* __restrict__ Xu, int ldu) {
    const int j = blockIdx.y;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < kpad; k += gridDim.x * 256)
        Xu[(long)j * ldu + k] = k < K ? X[(long)j * ldx + U[k]] : 0.f;
}

// Au (row major, ld ldu) = A[:, U[0..K)], zero on columns [K, kpad): 32x32 tiles transposed
// through LDS, so both the column reads of A (column major, ld lda) and the row writes coalesce
__global__ __launch_bounds__(256) void union_gather_a_kernel(const float* __restrict__ Acm, int lda,
                                                             const int* __restrict__ U, int K,
                                                             float* __restrict__ Au, int ldu) {
    __shared__ float t[32][33];
    const int k0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int k = k0 + r;
        t[r][tx] = k < K ? Acm[(long)U[k] * lda + i0 + tx] : 0.f;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) Au[(long)(i0 + r) * ldu + k0 + tx] = t[tx][r];
}

hipError_t launch_union_list(const float* X, int ldx, int n, int batch, SpState sp, int* flag, int* U, int* stat,
                             hipStream_t st) {
    if (n <= 0 || batch <= 0) return hipErrorInvalidValue;
    union_flags_kernel<<<(n + 255) / 256, 256, 0, st>>>(X, ldx, n, batch, sp, flag);
    union_list_kernel<<<1, 1024, 0, st>>>(flag, n, U, sp.done, stat);
    return hipGetLastError();
}

hipError_t launch_union_gather(const float* Acm, int lda, int mp, const float* X, int ldx, int rows, const int* U,
                               int K, int kpad, float* Au, float* Xu, int ldu, hipStream_t st) {
    if (K < 0 || K > kpad || kpad % 32 || kpad > ldu || mp % 32 || rows <= 0) return hipErrorInvalidValue;
    union_gather_x_kernel<<<dim3(std::min((kpad + 255) / 256, 64), rows), 256, 0, st>>>(X, ldx, U, K, kpad, Xu, ldu);
    union_gather_a_kernel<<<dim3(kpad / 32, mp / 32), 256, 0, st>>>(Acm, lda, U, K, Au, ldu);
    return hipGetLastError();
}

// ---------------------------------------------------------------- formulation step
// Exact selection without sorting the column.  A first pass bins |v| linearly over [0, max]
// (kBins bins, LDS histogram); the bin holding the answer is then resolved exactly by sorting
// its elements, or by four 8-bit radix passes over the IEEE bits of |v| (monotone for
// non-negative floats) restricted to that bin when it holds too many.  Counts are integers and
// the sums the lambda search needs are accumulated in 2^-40-scaled fixed point, so every
// decision is independent of atomic order: results are bitwise reproducible.
// The column is re-read per pass through a bounded buffer view (coalesced; it stays in
// L2/MALL), so any n works with one code path.
constexpr int kBins = 1024;
constexpr double kFix = 1099511627776.0;  // 2^40: n * 2^40 < 2^63 for n < 2^23

constexpr int kCand = 2048;  // candidates of the answer bin resolved in LDS (else the radix fallback)
// The previous threshold's list: up to kList elements with |v| at or above a bound L the SP's
// previous step left: the lower edge of the histogram bin (of the list's, or of the column's)
// holding place 2s + 64 (L0) or 3/2 survivors + 64 (L1), or, after a radix fallback, a margin
// below the threshold it found.
constexpr int kList = kListMax;
constexpr int kLBins = 256;   // bins of the list's histogram
static_assert(kLBins == 256, "the list bins take the top 8 bits of the span");
constexpr int kLCand = 256;   // entries of the list's answer bin resolved in LDS (more: the column passes)
constexpr float kListMarginL0 = 0.9f;
constexpr float kListMarginL1 = 0.95f;
// The place whose value becomes the next step's list bound, for `keep` elements kept: L0 (keep =
// s) 2s + 64, capped at 3/4 of the list's room (not below s + s/8 + 64); L1 (keep = the survivors,
// whose count moves with lambda) keep + keep/8 + 64.  0 (no list: the column passes) past the room.
__device__ __forceinline__ uint32_t list_rank(uint32_t keep, bool l1c) {
    const uint32_t tight = keep + keep / 8 + 64;
    if (tight > (uint32_t)kList) return 0u;
    return l1c ? tight : max(min(2 * keep + 64, (uint32_t)(3 * kList / 4)), tight);
}

// LDS of one step workgroup.  Every step kind uses StepSmall (reductions, the exact-tie band);
// the constrained kinds also the histogram and list arrays of SelShared, under 40 KiB in all, so
// four 256-thread workgroups share a CU.
struct StepSmall {
    double red[3 * (kSelThreads / 64)];
    int ires[4];
    uint64_t ures[2];
    int ncand;
    int nkeep;
    int nlc;
    int wtot[kSelThreads / 64];  // per-wave totals (chunk places of the staged list)
    double dres[4];
    // exact ties: the re-evaluated band (sorted descending), its value range and keep flags
    int nband;
    float band_lo, band_hi;
    int band_idx[kBandMax];
    float band_val[kBandMax];
    int band_keep[kBandMax];
};
struct SelShared : StepSmall {
    union {
        struct {
            uint32_t cnt[kBins];
            uint64_t s1[kBins];
            uint64_t s2[kBins];
            uint64_t cand[kCand];  // (|v| bits << 32) | index, sorted descending
        };
        struct {
            // the previous threshold's list: (|v| bits << 32) | index | sign of v in bit 31
            uint64_t list[kList];
            uint32_t lcnt[kLBins];   // its histogram over [L, max |v|], fixed-point sums as above
            uint64_t ls1[kLBins];
            uint64_t ls2[kLBins];
            uint64_t lcand[kLCand];  // the list's answer bin, sorted descending
        };
    };
};
static_assert(sizeof(SelShared) <= 40 * 1024, "four step workgroups per CU");
// While the staged list is gathered, ls1 .. lcand hold the chunk counts and places (2 ints per
// chunk): at most kMaxListChunks chunks (n <= kMaxListChunks * kStepChunk); longer columns take
// no list.
constexpr int kMaxListChunks = (int)(2 * sizeof(uint64_t) * kLBins + sizeof(uint64_t) * kLCand) / (2 * (int)sizeof(int));

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    const uint32_t lo = __shfl_xor((uint32_t)v, m, 64), hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Registers r and r | T of each thread (r with bit T clear holds the lower element index) as one
// compare-exchange step of a bitonic sort: element i = (w * R + r) * 64 + lane.
template <int R, int T>
__device__ __forceinline__ void sort_cx_regs(uint64_t (&x)[R], int k, int w, int lane) {
    if constexpr (T < R) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            if (r & T) continue;
            const bool desc = ((((w * R + r) * 64 + lane) & k) == 0);
            const uint64_t a0 = x[r], a1 = x[r | T];
            const uint64_t hi = a0 > a1 ? a0 : a1, lo = a0 > a1 ? a1 : a0;
            x[r] = desc ? hi : lo;
            x[r | T] = desc ? lo : hi;
        }
    }
}

// a[0 .. nc) sorted descending as u64 keys by the whole block (NT threads, R <= 16), nc <= NT * R; a must
// hold NT * R entries (the tail is padded with 0 keys, which sort last).  A bitonic network over
// NT * R elements, element i = (w * R + r) * 64 + lane held in register r of lane `lane` of wave
// w: exchanges across fewer than 64 places are lane shuffles, across 64 .. 64R - 1 places
// register swaps, and only the widest (64R and more) go through LDS.  Keys are distinct in use
// (the index is in the low word), so the order is the unique descending one.
template <int NT, int R>
__device__ void sort_desc_reg(uint64_t* a, int nc) {
    static_assert(R == 1 || R == 2 || R == 4 || R == 8 || R == 16, "register exchanges across 1 .. 8 registers");
    constexpr int N = NT * R;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t x[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = (w * R + r) * 64 + lane;
        x[r] = i < nc ? a[i] : 0ull;
    }
    for (int k = 2; k <= N; k <<= 1)
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            if (jj >= 64 * R) {
                __syncthreads();  // every read of the previous exchange is done
#pragma unroll
                for (int r = 0; r < R; r++) a[(w * R + r) * 64 + lane] = x[r];
                __syncthreads();
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int i = (w * R + r) * 64 + lane;
                    const uint64_t p = a[i ^ jj];
                    const bool keep_hi = ((i & jj) == 0) == ((i & k) == 0);
                    x[r] = keep_hi ? (x[r] > p ? x[r] : p) : (x[r] < p ? x[r] : p);
                }
            } else if (jj >= 64) {
                const int t = jj >> 6;  // < R <= 16
                if (t == 1) sort_cx_regs<R, 1>(x, k, w, lane);
                else if (t == 2) sort_cx_regs<R, 2>(x, k, w, lane);
                else if (t == 4) sort_cx_regs<R, 4>(x, k, w, lane);
                else sort_cx_regs<R, 8>(x, k, w, lane);
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int i = (w * R + r) * 64 + lane;
                    const uint64_t p = shfl_xor_u64(x[r], jj);
                    const bool keep_hi = ((lane & jj) == 0) == ((i & k) == 0);
                    x[r] = keep_hi ? (x[r] > p ? x[r] : p) : (x[r] < p ? x[r] : p);
                }
            }
        }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) a[(w * R + r) * 64 + lane] = x[r];
    __syncthreads();
}

// a[0 .. nc) sorted descending (as u64 keys) in LDS by the whole block; a must hold the next
// multiple of NT that is NT times a power of two >= nc (padded with zeros, which sort last)
template <int NT>
__device__ void bitonic_desc(uint64_t* a, int nc) {
    if (nc <= NT) sort_desc_reg<NT, 1>(a, nc);
    else if (nc <= 2 * NT) sort_desc_reg<NT, 2>(a, nc);
    else if (nc <= 4 * NT) sort_desc_reg<NT, 4>(a, nc);
    else sort_desc_reg<NT, 8>(a, nc);
}

__device__ __forceinline__ int lin_bin(float a, float scale) {
    int b = (int)(a * scale);
    return b < kBins - 1 ? b : kBins - 1;
}

// lambda of the reference GPU scan (c++/src/gpugpower/my_cublas_wrapper.h:335-375):
// with r(l) = (S1 - c l)/sqrt(S2 - 2 l S1 + c l^2) over elements > l, the survivors are
// the elements a with r(a) < w (w = sqrt(s)); lambda is the "-" root of
// c(c-w^2) l^2 + (2 S1 w^2 - 2 c S1) l + (S1^2 - w^2 S2) = 0 over those survivors.
__device__ __forceinline__ bool pred_r(double c, double S1, double S2, double lam, double w) {
    const double r = (S1 - c * lam) / sqrt(S2 - 2.0 * lam * S1 + c * lam * lam);
    return r >= w;  // NaN (no element above lam) is false
}

// One wave (lanes own NB/64 consecutive bins) scans a histogram from the top.  L0 mode
// (l1c false): find the bin where the count from the top reaches `rank`.  Lambda mode:
// find the largest bin whose lower boundary satisfies pred_r.  Writes ires[0] = bin (or
// -1), ires[1] = count above it, ures = fixed-point sums above it.
template <int NB, class Boundary>
__device__ void scan_hist(StepSmall& sh, const uint32_t* cnt, const uint64_t* hs1, const uint64_t* hs2, bool l1c,
                          uint32_t rank, uint32_t cA, uint64_t s1A, uint64_t s2A, double M, double wd,
                          Boundary boundary) {
    const int lane = threadIdx.x & 63;
    constexpr int PER = NB / 64;
    uint32_t lc = 0;
    uint64_t l1 = 0, l2 = 0;
    for (int b = lane * PER; b < lane * PER + PER; b++) {
        lc += cnt[b];
        if (l1c) { l1 += hs1[b]; l2 += hs2[b]; }
    }
    const uint32_t ic = wave_suffix_u32(lc);
    uint64_t i1 = 0, i2 = 0;
    if (l1c) { i1 = wave_suffix_u64(l1); i2 = wave_suffix_u64(l2); }
    uint32_t c = ic - lc;
    uint64_t S1 = i1 - l1, S2 = i2 - l2;
    int found = -1;
    uint32_t fc = 0;
    uint64_t f1 = 0, f2 = 0;
    for (int b = lane * PER + PER - 1; b >= lane * PER; b--) {
        const uint32_t cb = cnt[b];
        if (!l1c) {
            if (found < 0 && c < rank && c + cb >= rank) { found = b; fc = c; }
        } else {
            const uint64_t b1 = hs1[b], b2 = hs2[b];
            double lam;
            const int forced = boundary(b, lam);  // 1: true, -1: false, 0: evaluate
            const bool p = forced > 0 || (forced == 0 && pred_r((double)(cA + c + cb), (double)(s1A + S1 + b1) * M / kFix,
                                                                 (double)(s2A + S2 + b2) * M * M / kFix, lam, wd));
            if (found < 0 && p) { found = b; fc = c; f1 = S1; f2 = S2; }
            S1 += b1;
            S2 += b2;
        }
        c += cb;
    }
    const int fmax = wave_max_i(found);
    if (found == fmax && found >= 0) {
        sh.ires[0] = found;
        sh.ires[1] = (int)fc;
        sh.ures[0] = f1;
        sh.ures[1] = f2;
    }
    if (lane == 0 && fmax < 0) sh.ires[0] = -1;
}

// One workgroup of NT threads per starting point.  Each pass streams the column through the
// bounded view, DEPTH loads in flight per thread; every pass visits a thread's elements in
// the same order, and reductions have a fixed order, so results are bitwise reproducible.
// The step of one SP (column j of V), specialised by kind (step_kind: 0 L0 constraint, 1 L1
// constraint, 2 L0 penalty, 3 L1 penalty) and iterate precision, so the per-element code of the
// column sweeps carries no formulation or precision branches.  sh: the calling kernel's LDS.
// Stores of an iterate's entry x_i (i < n) of SP j: FP32 and the operand forms.  16-bit forms are
// K-blocked: element i of column j at ((i/32)*bp + j)*32 + i%32; the views cover the whole operand
// (ldv * bp elements), and elements i >= n are dropped by the FP32 view's bound.
template <int PREC>
struct XPut {
    __amdgpu_buffer_rsrc_t xr, hr, lr;
    int j, bp;
    __device__ XPut(const Operand& xop, const StepArgs& a, int j_) : j(j_), bp(xop.bp) {
        xr = col_rsrc(xop.f32 + (long)j * a.ldv, (uint32_t)a.n * 4u);
        const uint32_t opb = PREC ? (uint32_t)((long)a.ldv * xop.bp * 2) : 0u;
        hr = col_rsrc(xop.hi ? (const void*)xop.hi : (const void*)xop.f32, opb);
        lr = col_rsrc(xop.lo ? (const void*)xop.lo : (const void*)xop.f32, PREC == 1 ? opb : 0u);
    }
    __device__ __forceinline__ void operator()(int i, float x) const {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(x), xr, (uint32_t)i * 4u, 0, 0);
        if constexpr (PREC != 0) {
            const uint32_t bo = (uint32_t)(blk_index(j, i, bp) * 2);
            if constexpr (PREC == 1) {
                const uint16_t h = f2bf(x);
                __builtin_amdgcn_raw_buffer_store_b16(h, hr, bo, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b16(f2bf(x - bf2f(h)), lr, bo, 0, 0);
            } else {
                __builtin_amdgcn_raw_buffer_store_b16(f2h(x), hr, bo, 0, 0);
            }
        }
    }
    // entries i .. i + 3 (i a multiple of 4, i + 3 < n): one 16-byte FP32 store and one 8-byte
    // store per 16-bit part (the four are consecutive in the K-blocked layout too)
    __device__ __forceinline__ void put4(int i, f4 x) const {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, x), xr, (uint32_t)i * 4u, 0, 0);
        if constexpr (PREC != 0) {
            const uint32_t bo = (uint32_t)(blk_index(j, i, bp) * 2);
            uint16_t h[4], l[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if constexpr (PREC == 1) {
                    h[k] = f2bf(x[k]);
                    l[k] = f2bf(x[k] - bf2f(h[k]));
                } else {
                    h[k] = f2h(x[k]);
                }
            }
            typedef unsigned int u2 __attribute__((ext_vector_type(2)));
            const u2 hv = {(uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16)};
            __builtin_amdgcn_raw_buffer_store_b64(hv, hr, bo, 0, 0);
            if constexpr (PREC == 1) {
                const u2 lv = {(uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16)};
                __builtin_amdgcn_raw_buffer_store_b64(lv, lr, bo, 0, 0);
            }
        }
    }
};

__device__ __forceinline__ void step_scan_chunk(int j, int c, float L, const float* __restrict__ V,
                                                const StepArgs& a, const SpState& sp);

template <int NT, int DEPTH, int KIND, int PREC, class SH>
__device__ __forceinline__ void step_body(SH& sh, int j, const float* __restrict__ V, Operand xop, StepArgs a,
                                          SpState sp) {
    if (*sp.done || !sp.active[j]) return;
    static_assert(NT <= kSelThreads, "SelShared holds kSelThreads / 64 waves");
    static_assert(KIND >= 0 && KIND < 4 && PREC >= 0 && PREC < 3, "step kind and precision");
    constexpr bool kConstrained = KIND < 2;
    const int tid = threadIdx.x;
    const int nq = (a.n + NT - 1) / NT;
    const __amdgpu_buffer_rsrc_t vr = col_rsrc(V + (long)j * a.ldv, (uint32_t)a.n * 4u);
#define LDV(q) __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(vr, tid * 4 + (q) * NT * 4, 0, 0))
    // fn(q, v) for this thread's elements q = 0 .. nq-1 (element index tid + q * NT; the last
    // row's entries past n read 0)
    // Rows are taken DEPTH at a time, alternating between two register buffers: the next DEPTH
    // rows are loaded (unconditionally: rows past the column read 0 without a memory access)
    // before the current ones are processed, so the column's loads stay in flight.
    auto sweep = [&](auto&& fn) {
        float ba[DEPTH], bb[DEPTH];
        auto load = [&](float (&vv)[DEPTH], int q0) {
#pragma unroll
            for (int u = 0; u < DEPTH; u++) vv[u] = LDV(q0 + u);
        };
        auto run = [&](const float (&vv)[DEPTH], int q0) {
#pragma unroll
            for (int u = 0; u < DEPTH; u++)
                if (q0 + u < nq) fn(q0 + u, vv[u]);
        };
        if constexpr (KIND != 1) {
            load(ba, 0);
            for (int q0 = 0; q0 < nq; q0 += 2 * DEPTH) {
                load(bb, q0 + DEPTH);
                run(ba, q0);
                if (q0 + DEPTH >= nq) break;
                load(ba, q0 + 2 * DEPTH);
                run(bb, q0 + DEPTH);
            }
        } else {
            // the L1 constraint's passes: one buffer of 2 DEPTH rows, the registers of the two
            // above (two buffers measured 10-20% slower here)
            float bc[2 * DEPTH];
            for (int q0 = 0; q0 < nq; q0 += 2 * DEPTH) {
#pragma unroll
                for (int u = 0; u < 2 * DEPTH; u++) bc[u] = LDV(q0 + u);
#pragma unroll
                for (int u = 0; u < 2 * DEPTH; u++)
                    if (q0 + u < nq) fn(q0 + u, bc[u]);
            }
        }
    };
    // the iterate's stores: x_i (i < n) in FP32 and the operand forms
    double nz = 0;
    const XPut<PREC> put(xop, a, j);
    float fval;
    // element transform applied when the iterate is written: keep |v| >= tau (L0),
    // soft map with lam (L1), penalised maps; then multiply by `scale_out`
    float tau = 0.f, lam = 0.f, scale_out = 1.f;
    if (tid == 0) sh.nband = 0;  // no re-evaluated tie band unless the L0 fast path makes one

    // L1 constraint with s <= 1 (|x|_1 <= |x|_2) admits only 1-sparse x: the top-1 hard
    // threshold, which the lambda scan cannot produce (its quadratic degenerates)
    constexpr bool l1c = KIND == 1;
    bool decided = false;  // constrained: the step was resolved on the previous threshold's list
    int list_cnt = 0;      // entries of that list
    if constexpr (kConstrained) {
        // ---- constrained: select, then x <- x/|x|_2
        const double wd = (double)(float)sqrt((double)a.s);
        double norm2 = 0;
        float next_bound = 0.f;  // the next step's list bound (0: no list)
        bool decided_column = false;  // the column passes' histogram set next_bound
        // The threshold from candidates cand[0..nc) (the answer bin of a histogram, or the whole
        // previous threshold's list) sorted descending by (|v|, index), with the elements above
        // them counted in cA_hist (L0) / summed in (cAd, S1A, S2A), both exact.  band_inside(lo,
        // hi): every value in [lo, hi] is a candidate.  list_mode (the candidates are the list of
        // every element >= lbound): an L1 trigger below the list (r(lbound) < w over it) or a
        // clamped lambda, whose norm needs the whole column, is left to the column passes;
        // returns false then, block-uniformly.  Otherwise the elements just below the candidates
        // trigger (the histogram scan chose the bin by its lower edge).
        auto resolve = [&](uint64_t* cand, int nc, uint32_t cA_hist, double cAd, double S1A, double S2A,
                           bool list_mode, bool closed_ok, double lbound, auto&& band_inside) -> bool {
            if (!l1c) {
                // ---- L0 constrained: tau = the (s - above)-th largest candidate (counting ties)
                const uint32_t r = a.s - cA_hist;
                tau = (r >= 1 && (int)r <= nc) ? __uint_as_float((uint32_t)(cand[r - 1] >> 32)) : 0.f;
                // exact ties: candidates within the product's rounding of tau could be in either
                // order; when every value of the band [lo, hi] is a candidate and the band holds
                // 2..kBandMax of them, they are recomputed in FP64 and the s-th place is decided on
                // the exact values
                const float lo = tau * (1.f - 2.f * a.band_eps), hi = tau * (1.f + 2.f * a.band_eps);
                const bool try_band = a.band_eps > 0.f && tau > 0.f && band_inside(lo, hi);
                int i0 = 0, i1 = 0;  // band = sorted candidates [i0, i1)
                if (try_band) {
                    double above = 0, inband = 0;
                    for (int i = tid; i < nc; i += NT) {
                        const float av = __uint_as_float((uint32_t)(cand[i] >> 32));
                        above += av > hi ? 1.0 : 0.0;
                        inband += (av <= hi && av >= lo) ? 1.0 : 0.0;
                    }
                    block_sum2<NT>(above, inband, sh.red);
                    i0 = (int)above;
                    i1 = i0 + (int)inband;
                }
                const int nb = i1 - i0;
                if (tid == 0) sh.nband = 0;
                if (try_band && nb >= 2 && nb <= kBandMax) {
                    const int lane = tid & 63;
                    for (int w = tid >> 6; w < nb; w += NT / 64) {
                        const int idx = (int)((uint32_t)cand[i0 + w] & 0x7fffffffu);  // bit 31: list sign
                        const f4* ac = (const f4*)(a.acm + (long)idx * a.mp);
                        const f4* zc = (const f4*)(a.zf + (long)j * a.mp);
                        double d = 0;
                        for (int q = lane; q < (a.mp >> 2); q += 64) {
                            const f4 x = ac[q], y = zc[q];
                            d += (double)x[0] * y[0] + (double)x[1] * y[1] + (double)x[2] * y[2] + (double)x[3] * y[3];
                        }
                        d = wave_sum(d);
                        if (lane == 0) {
                            sh.band_idx[w] = idx;
                            sh.band_val[w] = (float)d;
                        }
                    }
                    __syncthreads();
                    if (tid == 0) {
                        // descending by (|v|, index), the candidates' own order
                        for (int p = 1; p < nb; p++) {
                            const float v = sh.band_val[p];
                            const int ix = sh.band_idx[p];
                            const uint64_t key = ((uint64_t)__float_as_uint(fabsf(v)) << 32) | (uint32_t)ix;
                            int q = p - 1;
                            while (q >= 0 && (((uint64_t)__float_as_uint(fabsf(sh.band_val[q])) << 32) |
                                              (uint32_t)sh.band_idx[q]) < key) {
                                sh.band_val[q + 1] = sh.band_val[q];
                                sh.band_idx[q + 1] = sh.band_idx[q];
                                q--;
                            }
                            sh.band_val[q + 1] = v;
                            sh.band_idx[q + 1] = ix;
                        }
                        // the band supplies places i0+1 .. of the candidates: r - i0 of them
                        const float tb = fabsf(sh.band_val[(int)r - 1 - i0]);
                        for (int p = 0; p < nb; p++) sh.band_keep[p] = !(fabsf(sh.band_val[p]) < tb);
                        sh.band_lo = lo;
                        sh.band_hi = hi;
                        sh.nband = nb;
                    }
                }
                __syncthreads();
                const int nband = sh.nband;
                double cs = 0;
                for (int i = tid; i < nc; i += NT) {
                    const float av = __uint_as_float((uint32_t)(cand[i] >> 32));
                    if (nband ? i < i0 : !(av < tau)) cs += (double)av * (double)av;
                }
                if (tid < nband && sh.band_keep[tid]) cs += (double)sh.band_val[tid] * (double)sh.band_val[tid];
                norm2 = S2A + block_sum<NT>(cs, sh.red);
                return true;
            }
            // ---- L1 constrained: the first candidate (descending) where the reference scan
            // triggers, r(a) >= w over the elements above it; one wave walks 64 at a time
            if (tid < 64) {
                const int lane = tid;
                double cc = cAd, c1 = S1A, c2 = S2A;  // carried sums of everything above the chunk
                int found = -1;
                double fc = 0, f1 = 0, f2 = 0;
                for (int c0 = 0; c0 < nc && found < 0; c0 += 64) {
                    const int i = c0 + lane;
                    const double av = i < nc ? (double)__uint_as_float((uint32_t)(cand[i] >> 32)) : 0.0;
                    double p0 = i < nc ? 1.0 : 0.0, p1 = av, p2 = av * av;
                    // inclusive prefix over lanes 0..lane, fixed shuffle order
                    for (int o = 1; o < 64; o <<= 1) {
                        const double t0 = __shfl_up(p0, o, 64), t1 = __shfl_up(p1, o, 64), t2 = __shfl_up(p2, o, 64);
                        if (lane >= o) { p0 += t0; p1 += t1; p2 += t2; }
                    }
                    // sums over the elements before candidate i
                    const double e0 = cc + p0 - (i < nc ? 1.0 : 0.0), e1 = c1 + p1 - av, e2 = c2 + p2 - av * av;
                    const bool trig = i < nc && pred_r(e0 + 1.0, e1 + av, e2 + av * av, av, wd);
                    const uint64_t bal = __ballot(trig);
                    if (bal) {
                        const int first = __ffsll((long long)bal) - 1;
                        found = c0 + first;
                        fc = __shfl(e0, first, 64);
                        f1 = __shfl(e1, first, 64);
                        f2 = __shfl(e2, first, 64);
                    } else {
                        cc += __shfl(p0, 63, 64);
                        c1 += __shfl(p1, 63, 64);
                        c2 += __shfl(p2, 63, 64);
                    }
                }
                if (lane == 0) {
                    if (found < 0) { fc = cc; f1 = c1; f2 = c2; }
                    sh.dres[0] = fc;
                    sh.dres[1] = f1;
                    sh.dres[2] = f2;
                    sh.ires[3] = found >= 0;
                    sh.ires[2] = found;
                }
            }
            __syncthreads();
            const double c = sh.dres[0], S1 = sh.dres[1], S2 = sh.dres[2];
            // no trigger inside the list: the first element below lbound triggers iff r(lbound) >= w
            // over the list (r grows as the boundary falls); else the survivors reach below it
            if (list_mode && !sh.ires[3] && !pred_r(c, S1, S2, lbound, wd)) return false;
            const double w2 = wd * wd;
            const double qa = c * (c - w2);
            const double qb = 2.0 * S1 * w2 - 2.0 * c * S1;
            const double qc = S1 * S1 - w2 * S2;
            lam = (float)((-qb - sqrt(qb * qb - 4.0 * qa * qc)) / (2.0 * qa));
            // no positive root: the constraint is inactive (s >= (|v|_1/|v|_2)^2, e.g. s >= n)
            // and lambda = 0; the reference scan's quadratic degenerates there (NaN or a
            // negative root), a branch no fixture pins
            const bool clamped = !(lam > 0.f);
            if (clamped) {
                if (list_mode) return false;
                lam = 0.f;
            }
            if (closed_ok && !clamped) {
                // every element of the survivor set lies above lam: sum of (a - lam)^2 in closed form
                const double lf = (double)lam;
                norm2 = S2 - 2.0 * lf * S1 + c * lf * lf;
            } else {
                double acc = 0;
                sweep([&](int, float v) {
                    const float t2 = fabsf(v) - lam;
                    if (t2 > 0.f) acc += (double)t2 * (double)t2;
                });
                norm2 = block_sum<NT>(acc, sh.red);
            }
            return true;
        };
        // ---- the previous threshold's list: from the second iteration on, the SP's threshold
        // (tau, or lambda) moves little, so the elements with |v| >= L, the bound the previous
        // step left a little below its threshold, were collected by the chunk scans of the previous
        // launch (step_scan_body), with max |v|.  A histogram of the list finds the threshold's
        // bin, and only that bin is sorted.  Exact whenever it decides: the list holds every
        // element >= L, L0 needs >= s of them (tau >= L then), L1 a trigger boundary at or above
        // L; otherwise the column passes below run as on a first step, with max |v| known.
        float Mcol = 0.f;
        bool have_M = false;
        const float thr_prev = sp.thr[j];
        if (thr_prev > 0.f) {
            float L = thr_prev;
            Mcol = __uint_as_float(sp.lmax[j]);
            have_M = true;
            // the staged entries into LDS in chunk order (within a chunk, the scan's order), so the
            // list's order, and every FP64 sum over it below, is fixed
            const int nch = (a.n + kStepChunk - 1) / kStepChunk, chunks = step_chunks(a.ldv);
            const int* lc = sp.lcount + (long)j * chunks;
            int* rec = (int*)sh.ls1;  // [0, nch): counts; [nch, 2 nch): places (nch <= kMaxListChunks)
            const unsigned long long* stg = sp.lstage + (long)j * chunks * kStageCap;
            // the chunks' counts and their places in the list (chunk-order exclusive prefix) into
            // rec; the list's length, or kList + 1 where a chunk overflowed its stage slots
            auto chunk_counts = [&]() {
            int over = 0, total = 0;
            for (int c0 = 0; c0 < nch; c0 += NT) {
                const int c = c0 + tid;
                const int k = c < nch ? lc[c] : 0;
                over |= k > kStageCap;
                // chunk-order exclusive prefix of the counts: this block of NT chunks' prefix and total
                const int incl = (int)wave_prefix_u32((uint32_t)k), pre = incl - k;
                const int wt = __builtin_amdgcn_readlane(incl, 63);
                const int w = tid >> 6;
                if ((tid & 63) == 0) sh.wtot[w] = wt;
                __syncthreads();
                int before = 0, btot = 0;
#pragma unroll
                for (int q = 0; q < NT / 64; q++) {
                    before += q < w ? sh.wtot[q] : 0;
                    btot += sh.wtot[q];
                }
                if (c < nch) {
                    rec[c] = k;
                    rec[nch + c] = total + before + pre;
                }
                total += btot;
                __syncthreads();  // wtot is reused by the next block of chunks
            }
            over = __syncthreads_or(over);  // also publishes rec
            return over ? kList + 1 : total;
            };
            int cnt = chunk_counts();
            if (!l1c && cnt > kList) {
                // L0, too long (the column's bulk grew past the bound): the staged keys -- every entry
                // of the chunks that fit their slots, the first kStageCap of any that did not -- by the
                // level-0 bins over [L, max |v|] (as below), and a second round, here in the step's
                // own workgroup, rescans the column's chunks with the edge of the bin holding place
                // list_rank(s) as its bound.  Few SPs take it, so this costs less than another
                // launch of scans and steps for every SP.  max |v| stays (the rescan finds the same).
                const uint32_t Lb = __float_as_uint(L);
                const int shift0 = max(0, 24 - __clz(__float_as_uint(Mcol) - Lb));
                for (int b = tid; b < kLBins; b += NT) sh.lcnt[b] = 0;
                __syncthreads();
                constexpr int P = 16;
                for (int f0 = 0; f0 < nch * kStageCap; f0 += NT * P) {
                    unsigned long long r[P];
#pragma unroll
                    for (int u = 0; u < P; u++) {
                        const int f = f0 + u * NT + tid, c = f / kStageCap, e = f - c * kStageCap;
                        r[u] = (c < nch && e < rec[c]) ? stg[f] : 0ull;
                    }
#pragma unroll
                    for (int u = 0; u < P; u++) {
                        const uint32_t kb = (uint32_t)(r[u] >> 32);
                        if (r[u]) atomicAdd(&sh.lcnt[(kb - Lb) >> shift0], 1u);  // staged keys are >= L
                    }
                }
                __syncthreads();
                const uint32_t want = list_rank(a.s, false);
                if (tid < 64 && want)
                    scan_hist<kLBins>(sh, sh.lcnt, sh.ls1, sh.ls2, false, want, 0u, 0ull, 0ull, 0.0, 0.0,
                                      [&](int, double&) { return 0; });
                __syncthreads();
                const int rb = want ? sh.ires[0] : -1;
                if (rb > 0) {
                    L = __uint_as_float(Lb + ((uint32_t)rb << shift0));
                    for (int c = 0; c < nch; c++) {
                        step_scan_chunk(j, c, L, V, a, sp);
                        __syncthreads();  // the scan's LDS totals are reused by the next chunk
                    }
                    cnt = chunk_counts();
                }
            }
            list_cnt = cnt;
            if (cnt > 0 && cnt <= kList) {
                // entry e of chunk c at stage slot c * kStageCap + e; the flat slots are taken 16 per
                // thread at a time, their loads all issued before the stores
                const unsigned long long* st = stg;
                constexpr int P = 16;
                for (int f0 = 0; f0 < nch * kStageCap; f0 += NT * P) {
                    unsigned long long r[P];
#pragma unroll
                    for (int u = 0; u < P; u++) {
                        const int f = f0 + u * NT + tid, c = f / kStageCap, e = f - c * kStageCap;
                        r[u] = (c < nch && e < rec[c]) ? st[f] : 0ull;
                    }
#pragma unroll
                    for (int u = 0; u < P; u++) {
                        const int f = f0 + u * NT + tid, c = f / kStageCap, e = f - c * kStageCap;
                        if (c < nch && e < rec[c]) sh.list[rec[nch + c] + e] = r[u];
                    }
                }
            }
            __syncthreads();
            if (tid == 0) {  // the next scans take the max afresh
                sp.lmax[j] = 0u;
            }
            if (cnt > 0 && cnt <= kList && (l1c || cnt >= (int)a.s)) {
                // Histogram levels over the IEEE bits of |v| (monotone for non-negative floats),
                // kLBins bins each: level 0 spans [L, max |v|], each further level the answer bin of
                // the level before, 8 bits finer, until the answer bin holds at most kLCand entries.
                // Bins of nearly constant relative width keep the bins near L, where the threshold
                // lies and the entries crowd, small; every bin edge is an exact float.  Counts and
                // the fixed-point sums above the range carry from level to level.
                const double invM = 1.0 / (double)Mcol;
                const uint32_t Lb = __float_as_uint(L);
                const int shift0 = max(0, 24 - __clz(__float_as_uint(Mcol) - Lb));  // span >> shift0 < 2^8
                auto hist_level = [&](uint32_t base, int shift, uint32_t lo, uint32_t hi, bool sums) {
                    for (int b = tid; b < kLBins; b += NT) {
                        sh.lcnt[b] = 0;
                        sh.ls1[b] = 0;
                        sh.ls2[b] = 0;
                    }
                    __syncthreads();
                    for (int e = tid; e < cnt; e += NT) {
                        const uint32_t kb = (uint32_t)(sh.list[e] >> 32);
                        if (kb < lo || kb >= hi) continue;
                        const int bb = (int)((kb - base) >> shift);
                        atomicAdd(&sh.lcnt[bb], 1u);
                        if (sums) {
                            const double w = (double)__uint_as_float(kb) * invM;
                            atomicAdd((unsigned long long*)&sh.ls1[bb], (unsigned long long)(w * kFix + 0.5));
                            atomicAdd((unsigned long long*)&sh.ls2[bb], (unsigned long long)(w * w * kFix + 0.5));
                        }
                    }
                    __syncthreads();
                };
                uint32_t base = Lb, lo = Lb, hi = 0xFFFFFFFFu, rank = a.s, cA = 0;
                uint64_t s1A = 0, s2A = 0;
                int shift = shift0, lb = -1, level = 0;
                for (;; level++) {
                    hist_level(base, shift, lo, hi, l1c);
                    if (tid < 64)
                        scan_hist<kLBins>(sh, sh.lcnt, sh.ls1, sh.ls2, l1c, rank, cA, s1A, s2A, (double)Mcol, wd,
                                          [&](int bb, double& l) {
                                              l = (double)__uint_as_float(base + ((uint32_t)bb << shift));
                                              return 0;
                                          });
                    __syncthreads();
                    lb = sh.ires[0];
                    if (lb < 0) break;
                    const uint32_t inb = sh.lcnt[lb];
                    cA += (uint32_t)sh.ires[1];
                    s1A += sh.ures[0];
                    s2A += sh.ures[1];
                    rank -= (uint32_t)sh.ires[1];
                    lo = base + ((uint32_t)lb << shift);
                    hi = lo + (1u << shift);
                    if (inb <= (uint32_t)kLCand || shift == 0) break;
                    base = lo;
                    shift = max(0, shift - 8);
                    __syncthreads();  // every wave has read this level's bin before the next level
                }
                if (lb >= 0 && sh.lcnt[lb] <= (uint32_t)kLCand) {
                    // the answer bin [lo, hi) into lcand; exact FP64 sums of the entries above it
                    if (tid == 0) sh.nlc = 0;
                    __syncthreads();
                    double c_loc = 0, s1_loc = 0, s2_loc = 0;
                    for (int e0 = 0; e0 < cnt; e0 += NT) {  // uniform trip count for the ballot
                        const int e = e0 + tid;
                        const uint64_t key = e < cnt ? sh.list[e] : 0ull;
                        const uint32_t kb = (uint32_t)(key >> 32);
                        const float av = __uint_as_float(kb);
                        if (e < cnt && kb >= hi) {
                            c_loc += 1.0;
                            s1_loc += (double)av;
                            s2_loc += (double)av * (double)av;
                        }
                        const bool inb = e < cnt && kb >= lo && kb < hi;
                        const uint64_t bal = __ballot(inb);
                        if (bal) {
                            const int lane = tid & 63, first = __ffsll((long long)bal) - 1;
                            int bs = 0;
                            if (lane == first) bs = atomicAdd(&sh.nlc, (int)__popcll(bal));
                            bs = __shfl(bs, first, 64);
                            if (inb) sh.lcand[bs + (int)__popcll(bal & ((1ull << lane) - 1ull))] = key;
                        }
                    }
                    block_sum3<NT>(c_loc, s1_loc, s2_loc, sh.red);
                    const double cAd = c_loc, S1A = s1_loc, S2A = s2_loc;
                    const int nc = sh.nlc;  // <= kLCand: the bin's count
                    bitonic_desc<NT>(sh.lcand, nc);
                    decided = resolve(sh.lcand, nc, cA, cAd, S1A, S2A, true, true, (double)__uint_as_float(lo),
                                      [&](float flo, float fhi) {
                                          return flo >= L && __float_as_uint(flo) >= lo && __float_as_uint(fhi) < hi;
                                      });
                    if (decided) {
                        // the next bound: the lower edge of the level-0 bin holding place list_rank
                        const uint32_t nr = list_rank(!l1c ? a.s : (uint32_t)sh.dres[0], l1c);
                        __syncthreads();
                        if (nr == 0) {
                            next_bound = 0.f;
                        } else {
                        if (level > 0) hist_level(Lb, shift0, Lb, 0xFFFFFFFFu, false);
                        if (tid < 64)
                            scan_hist<kLBins>(sh, sh.lcnt, sh.ls1, sh.ls2, false, nr, 0u, 0ull, 0ull, 0.0, 0.0,
                                              [&](int, double&) { return 0; });
                        __syncthreads();
                        const int rb = sh.ires[0];
                        next_bound = rb >= 0 ? __uint_as_float(Lb + ((uint32_t)rb << shift0)) : L;
                        }
                    }
                }
            }
            __syncthreads();  // the column passes below reuse this LDS
            if (tid == 0) {  // integer counts of list steps tried and decided (order-free)
                atomicAdd(&sp.done[2], 1);
                if (decided) atomicAdd(&sp.done[3], 1);
            }
            // Not decided in the first round with a usable list whose bound was too high (L0: fewer
            // than s entries; L1: no decision): the scans' counts under L give the highest half-octave
            // edge L' with enough entries (L0: list_rank(s); L1: twice the list), and the second
            // round rescans the column with L' instead of running the column passes here.  The
            // count of entries >= L' is known exactly, so L' is taken only if they fit the list.
        }
        if (!decided) {
        float M = Mcol;
        if (!have_M) {
            float mloc = 0.f;
            sweep([&](int, float v) { mloc = fmaxf(mloc, fabsf(v)); });
            M = block_max<NT>(mloc, sh.red);
        }
        const float scale = M > 0.f ? (float)kBins / M : 0.f;
        const double invM = M > 0.f ? 1.0 / (double)M : 0.0;
        const double Md = (double)M;
        for (int b = tid; b < kBins; b += NT) {
            sh.cnt[b] = 0;
            sh.s1[b] = 0;
            sh.s2[b] = 0;
        }
        __syncthreads();
        sweep([&](int, float v) {
            const float av = fabsf(v);
            const int b = lin_bin(av, scale);
            atomicAdd(&sh.cnt[b], 1u);
            if (l1c && av > 0.f) {
                const double w = (double)av * invM;
                atomicAdd((unsigned long long*)&sh.s1[b], (unsigned long long)(w * kFix + 0.5));
                atomicAdd((unsigned long long*)&sh.s2[b], (unsigned long long)(w * w * kFix + 0.5));
            }
        });
        __syncthreads();
        if (tid < 64)
            scan_hist<kBins>(sh, sh.cnt, sh.s1, sh.s2, l1c, a.s, 0u, 0ull, 0ull, Md, wd, [&](int b, double& l) {
                l = (double)b * Md / (double)kBins;
                return 0;
            });
        __syncthreads();
        const int bstar = sh.ires[0];
        const uint32_t cA_hist = bstar >= 0 ? (uint32_t)sh.ires[1] : 0u;
        const uint64_t s1A_fix = bstar >= 0 ? sh.ures[0] : 0ull, s2A_fix = bstar >= 0 ? sh.ures[1] : 0ull;
        const uint32_t cnt_b = bstar >= 0 ? sh.cnt[bstar] : 0u;
        bool fast = bstar >= 0 && cnt_b <= (uint32_t)kCand;
        if (bstar < 0 && l1c) fast = true;  // no trigger: every element survives, sums over all
        if (fast) {
            // ---- compaction: the answer bin's elements into LDS; exact FP64 sums of the bins above
            if (tid == 0) sh.ncand = 0;
            __syncthreads();
            double c_loc = 0, s1_loc = 0, s2_loc = 0;
            sweep([&](int q, float v) {
                const int idx = tid + q * NT;
                if (idx >= a.n) return;
                const float av = fabsf(v);
                const int b = lin_bin(av, scale);
                if (b > bstar) {
                    c_loc += 1.0;
                    s1_loc += (double)av;
                    s2_loc += (double)av * (double)av;
                } else if (b == bstar) {
                    const int slot = atomicAdd(&sh.ncand, 1);
                    if (slot < kCand) sh.cand[slot] = ((uint64_t)__float_as_uint(av) << 32) | (uint32_t)idx;
                }
            });
            block_sum3<NT>(c_loc, s1_loc, s2_loc, sh.red);
            const double cAd = c_loc, S1A = s1_loc, S2A = s2_loc;
            const int nc = min(sh.ncand, kCand);
            bitonic_desc<NT>(sh.cand, nc);
            resolve(sh.cand, nc, cA_hist, cAd, S1A, S2A, false, bstar >= 0, 0.0, [&](float lo, float hi) {
                return lin_bin(lo, scale) == bstar && lin_bin(hi, scale) == bstar;
            });
            // the next step's list bound: the lower edge of the bin holding place 2s + 64 (L0) or
            // 9/8 survivors + 64 (L1); none for an inactive L1 constraint (every element survives), a
            // column shorter than that, or a list that would not fit in LDS
            const uint32_t rank = list_rank(!l1c ? a.s : (uint32_t)sh.dres[0], l1c);  // 0: no list
            __syncthreads();
            if (tid < 64)
                scan_hist<kBins>(sh, sh.cnt, sh.s1, sh.s2, false, rank, 0u, 0ull, 0ull, 0.0, 0.0,
                                 [&](int, double&) { return 0; });
            __syncthreads();
            const int rb = sh.ires[0];
            const bool fits = rank > 0 && rb > 0 && (uint32_t)sh.ires[1] + sh.cnt[rb] <= (uint32_t)kList;
            next_bound = (fits && !(l1c && !(lam > 0.f))) ? (float)((double)rb * Md / (double)kBins) : 0.f;
            decided_column = true;
        } else {
        // ---- fallback (answer bin larger than the LDS list): radix passes over the column
        uint32_t prefix = 0;  // key of the answer within bin bstar
        uint32_t cA = cA_hist, rank = a.s - cA_hist;
        uint64_t s1A = s1A_fix, s2A = s2A_fix;
        if (bstar >= 0) {
            const double lam_b = (double)bstar * Md / (double)kBins;
            for (int pass = 0; pass < 4; pass++) {
                const int shift = 24 - 8 * pass;
                __syncthreads();
                if (tid < 256) {
                    sh.cnt[tid] = 0;
                    sh.s1[tid] = 0;
                    sh.s2[tid] = 0;
                }
                __syncthreads();
                sweep([&](int, float v) {
                    const float av = fabsf(v);
                    const uint32_t key = __float_as_uint(av);
                    if (lin_bin(av, scale) == bstar && (pass == 0 || (key >> (shift + 8)) == prefix)) {
                        const int d = (key >> shift) & 255;
                        atomicAdd(&sh.cnt[d], 1u);
                        if (l1c && av > 0.f) {
                            const double u = (double)av * invM;
                            atomicAdd((unsigned long long*)&sh.s1[d], (unsigned long long)(u * kFix + 0.5));
                            atomicAdd((unsigned long long*)&sh.s2[d], (unsigned long long)(u * u * kFix + 0.5));
                        }
                    }
                });
                __syncthreads();
                if (tid < 64)
                    scan_hist<256>(sh, sh.cnt, sh.s1, sh.s2, l1c, rank, cA, s1A, s2A, Md, wd, [&](int d, double& l) {
                        const uint32_t K = (pass == 0 ? 0u : (prefix << (shift + 8))) | ((uint32_t)d << shift);
                        // keys from +inf up (inf, NaN, sign bit set) bound no finite |v|: false.
                        // A boundary at or below the bin's lower edge is true by the choice of bstar.
                        if (K >= 0x7F800000u) return -1;
                        l = (double)__uint_as_float(K);
                        return l <= lam_b ? 1 : 0;
                    });
                __syncthreads();
                const int d = sh.ires[0];
                if (d < 0) break;  // cannot happen: digit 0 of the current range always qualifies
                cA += (uint32_t)sh.ires[1];
                s1A += sh.ures[0];
                s2A += sh.ures[1];
                rank -= (uint32_t)sh.ires[1];
                prefix = (pass == 0 ? 0u : prefix << 8) | (uint32_t)d;
            }
        }
        __syncthreads();
        if (!l1c) {
            tau = bstar >= 0 ? __uint_as_float(prefix) : 0.f;
            sweep([&](int, float v) {
                if (!(fabsf(v) < tau)) norm2 += (double)v * (double)v;
            });
        } else {
            double c = 0, S1 = 0, S2 = 0;
            sweep([&](int q, float v) {
                const float av = fabsf(v);
                const int b = lin_bin(av, scale);
                const bool in = b > bstar || (b == bstar && __float_as_uint(av) > prefix);
                if (in && (tid + q * NT) < a.n) {
                    c += 1.0;
                    S1 += (double)av;
                    S2 += (double)av * (double)av;
                }
            });
            block_sum3<NT>(c, S1, S2, sh.red);
            const double w2 = wd * wd;
            const double qa = c * (c - w2);
            const double qb = 2.0 * S1 * w2 - 2.0 * c * S1;
            const double qc = S1 * S1 - w2 * S2;
            lam = (float)((-qb - sqrt(qb * qb - 4.0 * qa * qc)) / (2.0 * qa));
            if (!(lam > 0.f)) lam = 0.f;  // inactive constraint, as in the fast path
            sweep([&](int, float v) {
                const float t2 = fabsf(v) - lam;
                if (t2 > 0.f) norm2 += (double)t2 * (double)t2;
            });
        }
        norm2 = block_sum<NT>(norm2, sh.red);
        }
        }  // !decided
        if (!decided && !decided_column) next_bound = l1c ? lam * kListMarginL1 : tau * kListMarginL0;
        // the next step's list bound (none for columns longer than the chunk records allow)
        if (tid == 0) sp.thr[j] = a.n <= kMaxListChunks * kStepChunk ? next_bound : 0.f;
        const float nf = (float)sqrt(norm2);
        scale_out = 1.0f / nf;
        fval = sp.fz[j];
    } else {
        // ---- penalised (my_cublas_wrapper.h:170-238,636-710), gamma in the iterate's precision.
        // The threshold is gamma itself, so the chunk workgroups of the previous launch
        // (pen_chunk_body) already wrote x by the plain rule and left per chunk the FP64 sums of
        // the objective's terms (all, and those of the exact-tie band), the nonzero counts and
        // the band's indices.  Here the chunks are summed in order and the band is decided.
        const float g = (float)a.gamma;
        lam = g;
        constexpr bool l0p = KIND == 2;
        auto contrib = [&](float v) -> double {
            if (l0p) {
                const float t = v * v - g;
                return t > 0.f ? (double)t : 0.0;
            }
            const float t = fabsf(v) - g;
            return t > 0.f ? (double)t * (double)t : 0.0;
        };
        const int nch = (a.n + kStepChunk - 1) / kStepChunk, chunks = step_chunks(a.ldv);
        const unsigned long long* st = sp.lstage + (long)j * chunks * kStageCap;
        double tot = 0, tb = 0;
        int nzi = 0, nzb = 0, nb = 0;
        if (tid == 0) {  // chunk order: a fixed order for the FP64 sums
            for (int c = 0; c < nch; c++) {
                const unsigned long long* r = st + (long)c * kStageCap;
                tot += __longlong_as_double((long long)r[0]);
                tb += __longlong_as_double((long long)r[1]);
                nzi += (int)(r[2] >> 32);
                nzb += (int)(uint32_t)r[2];
                nb += (int)r[3];
            }
            sh.dres[0] = tot;
            sh.dres[1] = tb;
            sh.ires[0] = nzi;
            sh.ires[1] = nzb;
            sh.ires[2] = nb;
        }
        __syncthreads();
        tot = sh.dres[0];
        tb = sh.dres[1];
        nzi = sh.ires[0];
        nzb = sh.ires[1];
        nb = sh.ires[2];
        nz = (double)nzi;
        // exact ties (L2 variance): the entries whose |v| lies within 2 eps of the threshold
        // (sqrt(gamma) for the L0 penalty, gamma for the L1 penalty), 1 .. kBandMax of them in
        // all, are recomputed in FP64 and their contributions replaced; more: none
        if (nb >= 1 && nb <= kBandMax) {
            if (tid == 0) {  // the chunks' band indices, then index order for the replaced sum
                int k = 0;
                for (int c = 0; c < nch; c++) {
                    const unsigned long long* r = st + (long)c * kStageCap;
                    for (int e = 0; e < (int)r[3]; e++) sh.band_idx[k++] = (int)r[4 + e];
                }
                for (int p = 1; p < nb; p++) {
                    const int ix = sh.band_idx[p];
                    int q = p - 1;
                    while (q >= 0 && sh.band_idx[q] > ix) {
                        sh.band_idx[q + 1] = sh.band_idx[q];
                        q--;
                    }
                    sh.band_idx[q + 1] = ix;
                }
            }
            __syncthreads();
            const int lane = tid & 63;
            for (int w = tid >> 6; w < nb; w += NT / 64) {
                const f4* ac = (const f4*)(a.acm + (long)sh.band_idx[w] * a.mp);
                const f4* zc = (const f4*)(a.zf + (long)j * a.mp);
                double d = 0;
                for (int q = lane; q < (a.mp >> 2); q += 64) {
                    const f4 x = ac[q], y = zc[q];
                    d += (double)x[0] * y[0] + (double)x[1] * y[1] + (double)x[2] * y[2] + (double)x[3] * y[3];
                }
                d = wave_sum(d);
                if (lane == 0) sh.band_val[w] = (float)d;
            }
            __syncthreads();
            if (tid == 0) {
                double tr = 0;
                for (int p = 0; p < nb; p++) {
                    const float v = sh.band_val[p];
                    sh.band_keep[p] = contrib(v) > 0.0;
                    tr += contrib(v);
                }
                sh.dres[3] = tr;
            }
            __syncthreads();
            tot = tot - tb + sh.dres[3];
            // the band's entries by their exact values: kept ones take the recomputed value
            if (tid < nb) {
                const float ve = sh.band_val[tid];
                const float x = sh.band_keep[tid] ? (l0p ? ve : (fabsf(ve) - g) * sgnf(ve)) : 0.f;
                put(sh.band_idx[tid], x);
            }
            int kept_b = 0;
            for (int p = 0; p < nb; p++) kept_b += sh.band_keep[p] && sh.band_val[p] != 0.f;
            nz += (double)kept_b;
        } else {
            nz += (double)nzb;  // none re-evaluated: the band's entries keep the plain rule
        }
        fval = l0p ? (float)tot : (float)sqrt(tot);
    }
    // ---- write the iterate (+ operand forms), count nonzeros, update the SP record
    if constexpr (!kConstrained) {
        // the chunk workgroups wrote the iterate; nz is the chunks' count, block-uniform
    } else if (decided) {
        // ---- a list step: x's nonzeros are the kept list entries, so only they are written,
        // after the previous step's nonzeros are cleared: the indices a previous list step kept,
        // or (after a column step, a refill or with deflation) the whole column
        int* kept = sp.kept + (long)j * kListMax;
        const int pk = sp.nkept[j];
        if (pk >= 0 && a.sparse_rewrite) {
            for (int e = tid; e < pk; e += NT) put(kept[e], 0.f);
        } else {
            for (int q = 0; q < nq; q++)
                if (tid + q * NT < a.n) put(tid + q * NT, 0.f);
        }
        if (tid == 0) sh.nkeep = 0;
        __syncthreads();  // the clears land before this step's values, which reuse most indices
        const int nband = !l1c ? sh.nband : 0;
        const float blo = nband ? sh.band_lo : 0.f, bhi = nband ? sh.band_hi : -1.f;
        for (int e0 = 0; e0 < list_cnt; e0 += NT) {  // uniform trip count: the ballot below sees every lane
            const int e = e0 + tid;
            bool keep = false;
            float x = 0.f;
            int i = 0;
            if (e < list_cnt) {
                const uint64_t key = sh.list[e];
                const float av = __uint_as_float((uint32_t)(key >> 32));
                i = (int)((uint32_t)key & 0x7fffffffu);
                const float sg = ((uint32_t)key & 0x80000000u) ? -1.f : 1.f;  // the sign of v
                if (!l1c && av >= blo && av <= bhi) {
                    // inside the re-evaluated tie band: its decision and exact value
                    for (int p = 0; p < nband; p++)
                        if (sh.band_idx[p] == i && sh.band_keep[p]) {
                            keep = true;
                            x = sh.band_val[p];
                        }
                } else if (!l1c) {
                    keep = !(av < tau);
                    x = sg * av;
                } else {
                    const float t2 = av - lam;
                    keep = t2 > 0.f;
                    x = t2 * sg;
                }
                x = x * scale_out;
            }
            const uint64_t bal = __ballot(keep);
            if (bal) {
                const int lane = tid & 63, first = __ffsll((long long)bal) - 1;
                int base = 0;
                if (lane == first) base = atomicAdd(&sh.nkeep, (int)__popcll(bal));
                base = __shfl(base, first, 64);
                if (keep) {
                    kept[base + (int)__popcll(bal & ((1ull << lane) - 1ull))] = i;
                    put(i, x);
                }
            }
        }
        __syncthreads();
        nz = (double)sh.nkeep;  // every kept value is nonzero: |v| >= tau > 0, or |v| - lam > 0
        if (tid == 0) sp.nkept[j] = sh.nkeep;
    } else {
        // exact ties (L0 fast path): values inside the re-evaluated band take its decision and value
        const int nband = !l1c ? sh.nband : 0;
        const float blo = nband ? sh.band_lo : 0.f, bhi = nband ? sh.band_hi : -1.f;
        sweep([&](int q, float v) {
                float x;
                if (!l1c) {
                    const float av = fabsf(v);
                    if (av >= blo && av <= bhi) {
                        const int i = tid + q * NT;
                        x = 0.f;
                        for (int p = 0; p < nband; p++)
                            if (sh.band_idx[p] == i) x = sh.band_keep[p] ? sh.band_val[p] : 0.f;
                    } else {
                        x = av < tau ? 0.f : v;
                    }
                } else {
                    const float t2 = fabsf(v) - lam;
                    x = t2 > 0.f ? t2 * sgnf(v) : 0.f;
                }
                x = x * scale_out;
                const int i = tid + q * NT;
                nz += (x != 0.f && i < a.n) ? 1.0 : 0.0;
                if (i < a.n) put(i, x);
        });
        nz = block_sum<NT>(nz, sh.red);
        if (tid == 0) sp.nkept[j] = -1;  // the whole column was written
    }
#undef LDV
    if (tid == 0) {
        const float prev = sp.val[j];
        sp.err[j] = fabsf(fval - prev) / prev;  // computeTheError (termination_criteria.h:29-36)
        sp.val[j] = fval;
        sp.nnz[j] = (int)nz;
    }
}


// One workgroup per SP; the penalised kinds need only StepSmall's LDS, so more of them share a CU.
template <int NT, int DEPTH, int KIND, int PREC>
__global__ __launch_bounds__(NT, 4) void step_kernel(const float* __restrict__ V, Operand xop,
                                                                     StepArgs a, SpState sp) {
    __shared__ std::conditional_t<(KIND < 2), SelShared, StepSmall> sh;
    step_body<NT, DEPTH, KIND, PREC>(sh, blockIdx.x, V, xop, a, sp);
}

// Chunk scans of the constrained steps (kinds 0 and 1): workgroup (SP j, chunk c).  An SP whose
// previous step left a list bound L has its column read by all its chunk workgroups at once, each
// over kStepChunk consecutive elements with eight 16-byte loads in flight per thread: the chunk's
// max |v| (atomicMax on the float bits, which order as the values for |v| >= 0), its count of
// entries |v| >= L and the entries themselves, in a fixed order (wave, lane, element), in the
// chunk's own kStageCap slots.  The step runs in the next launch (step_kernel), one workgroup per
// SP, and reads the chunks in order, so no SP waits for a whole-column pass by one workgroup and
// no workgroup waits on another.
__device__ __forceinline__ void step_scan_chunk(int j, int c, float L, const float* __restrict__ V,
                                                const StepArgs& a, const SpState& sp) {
    __shared__ int wtot[2][kSelThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    constexpr int R = kStepScanRows;
    const int n4 = a.n & ~3, nch = (a.n + kStepChunk - 1) / kStepChunk;
    const float* col = V + (long)j * a.ldv;
    // 16-byte loads below n4 (each wholly inside or past the view), the last n - n4 elements by
    // 4-byte loads; element base + (u * NT + tid) * 4 + k of row u, lane part k
    const __amdgpu_buffer_rsrc_t r4 = col_rsrc(col, (uint32_t)n4 * 4u);
    const int base = c * kStepChunk;
    f4 v[R];
#pragma unroll
    for (int u = 0; u < R; u++) {
        const uint32_t o = (uint32_t)(base + (u * kSelThreads + tid) * 4) * 4u;
        v[u] = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r4, o, 0, 0));
    }
    const int tail_i = (c == nch - 1 && tid < a.n - n4) ? n4 + tid : -1;
    float tv = 0.f;
    if (tail_i >= 0)
        tv = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(col_rsrc(col, (uint32_t)a.n * 4u),
                                                                 (uint32_t)tail_i * 4u, 0, 0));
    float mloc = fabsf(tv);
    uint32_t in = 0;
#pragma unroll
    for (int u = 0; u < R; u++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float av = fabsf(v[u][k]);
            mloc = fmaxf(mloc, av);
            in |= (av >= L ? 1u : 0u) << (u * 4 + k);
        }
    const bool tin = tail_i >= 0 && fabsf(tv) >= L;
    const int cnt = __popc(in) + (tin ? 1 : 0);  // at most 33: six bits
    int excl = 0, wsum = 0;
#pragma unroll
    for (int bit = 0; bit < 6; bit++) {
        const uint64_t m = __ballot((cnt >> bit) & 1);
        excl += (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u)) << bit;
        wsum += __popcll(m) << bit;
    }
    const float wmax = wave_max(mloc);
    if (lane == 0) {
        wtot[0][w] = wsum;
        wtot[1][w] = (int)__float_as_uint(wmax);
    }
    __syncthreads();
    int before = 0, tot = 0;
    uint32_t mx = 0;
#pragma unroll
    for (int k = 0; k < kSelThreads / 64; k++) {
        const int t = wtot[0][k];
        before += k < w ? t : 0;
        tot += t;
        mx = max(mx, (uint32_t)wtot[1][k]);  // non-negative floats order as their bits
    }
    const int chunks = step_chunks(a.ldv);
    if (tid == 0) {
        if (mx) atomicMax(sp.lmax + j, mx);
        sp.lcount[(long)j * chunks + c] = tot;  // more than kStageCap: the step runs the column passes
    }
    if (!cnt) return;
    int slot = before + excl;
    unsigned long long* dst = sp.lstage + ((long)j * chunks + c) * kStageCap;
    auto stage = [&](float x, int i) {
        if (slot < kStageCap)
            dst[slot] = ((uint64_t)__float_as_uint(fabsf(x)) << 32) | (uint32_t)i | (x < 0.f ? 0x80000000u : 0u);
        slot++;
    };
#pragma unroll
    for (int u = 0; u < R; u++)
#pragma unroll
        for (int k = 0; k < 4; k++)
            if ((in >> (u * 4 + k)) & 1u) stage(v[u][k], base + (u * kSelThreads + tid) * 4 + k);
    if (tin) stage(tv, tail_i);
}

__device__ __forceinline__ void step_scan_body(int j, int c, const float* __restrict__ V, const StepArgs& a,
                                               const SpState& sp) {
    if (*sp.done || !sp.active[j]) return;
    const float L = sp.thr[j];
    if (!(L > 0.f)) return;  // no list: the step runs the column passes
    step_scan_chunk(j, c, L, V, a, sp);
}

// Chunk workgroups of the penalised steps (kinds 2 and 3): workgroup (SP j, chunk c), as the
// constrained scans.  The threshold is gamma itself, so each chunk writes its entries of x by the
// plain rule at once and leaves, in its first stage slots: the FP64 sums of the objective's terms
// over the chunk (all, and those inside the exact-tie band), the nonzero counts (outside and inside
// the band) and the band's indices (up to kBandMax).  The step (step_kernel) sums the chunks in
// order and decides the band.  A thread's elements and the block's reduction tree are fixed, so
// the sums are reproducible.
template <int KIND, int PREC>
__device__ __forceinline__ void pen_chunk_body(int j, int c, const float* __restrict__ V, const Operand& xop,
                                               const StepArgs& a, const SpState& sp) {
    if (*sp.done || !sp.active[j]) return;
    __shared__ double red[2 * (kSelThreads / 64)];
    __shared__ int ib[kSelThreads / 64][2];
    __shared__ int nband;
    __shared__ int bidx[kBandMax];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) nband = 0;
    constexpr int R = kStepScanRows;
    constexpr bool l0p = KIND == 2;
    const float g = (float)a.gamma;
    const float thr = l0p ? sqrtf(g) : g;
    const float lo = thr * (1.f - 2.f * a.band_eps), hi = thr * (1.f + 2.f * a.band_eps);
    const bool band = a.band_eps > 0.f && thr > 0.f;
    const int n4 = a.n & ~3, nch = (a.n + kStepChunk - 1) / kStepChunk;
    const float* col = V + (long)j * a.ldv;
    const __amdgpu_buffer_rsrc_t r4 = col_rsrc(col, (uint32_t)n4 * 4u);
    const int base = c * kStepChunk;
    f4 v[R];
#pragma unroll
    for (int u = 0; u < R; u++) {
        const uint32_t o = (uint32_t)(base + (u * kSelThreads + tid) * 4) * 4u;
        v[u] = __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r4, o, 0, 0));
    }
    const int tail_i = (c == nch - 1 && tid < a.n - n4) ? n4 + tid : -1;
    float tv = 0.f;
    if (tail_i >= 0)
        tv = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(col_rsrc(col, (uint32_t)a.n * 4u),
                                                                 (uint32_t)tail_i * 4u, 0, 0));
    __syncthreads();  // nband cleared
    const XPut<PREC> put(xop, a, j);
    double acc = 0, accb = 0;
    int nzo = 0, nzb = 0;
    auto one = [&](int i, float x) -> float {  // i < n: the entry's term and band; returns x_i
        const float t = l0p ? x * x - g : fabsf(x) - g;
        const double ct = t > 0.f ? (l0p ? (double)t : (double)t * (double)t) : 0.0;
        // the iterate's entry: L0 penalty kept iff v^2 - gamma > 0, so a NaN v (the 0/0 after an
        // SP was zeroed out) gives 0, as the reference CPU rule (sparse_PCA_thresholding.h:67-102);
        // L1 penalty the soft threshold at gamma
        const float xv = l0p ? ((x * x - g) > 0.f ? x : 0.f) : (t > 0.f ? t * sgnf(x) : 0.f);
        acc += ct;
        const float ax = fabsf(x);
        if (band && ax >= lo && ax <= hi) {
            accb += ct;
            nzb += xv != 0.f;
            const int slot = atomicAdd(&nband, 1);
            if (slot < kBandMax) bidx[slot] = i;
        } else {
            nzo += xv != 0.f;
        }
        return xv;
    };
#pragma unroll
    for (int u = 0; u < R; u++) {
        const int i0 = base + (u * kSelThreads + tid) * 4;
        if (i0 < n4) {  // the row's four entries are all below n4
            f4 x;
#pragma unroll
            for (int k = 0; k < 4; k++) x[k] = one(i0 + k, v[u][k]);
            put.put4(i0, x);
        }
    }
    if (tail_i >= 0) put(tail_i, one(tail_i, tv));
    acc = wave_sum(acc);
    accb = wave_sum(accb);
    const int wnz = __builtin_amdgcn_readlane((int)wave_prefix_u32((uint32_t)nzo), 63);
    const int wnb = __builtin_amdgcn_readlane((int)wave_prefix_u32((uint32_t)nzb), 63);
    if (lane == 0) {
        red[w] = acc;
        red[kSelThreads / 64 + w] = accb;
        ib[w][0] = wnz;
        ib[w][1] = wnb;
    }
    __syncthreads();
    if (tid == 0) {
        double ta = 0, tbb = 0;
        int za = 0, zb = 0;
        for (int q = 0; q < kSelThreads / 64; q++) {
            ta += red[q];
            tbb += red[kSelThreads / 64 + q];
            za += ib[q][0];
            zb += ib[q][1];
        }
        unsigned long long* r = sp.lstage + ((long)j * step_chunks(a.ldv) + c) * kStageCap;
        r[0] = (unsigned long long)__double_as_longlong(ta);
        r[1] = (unsigned long long)__double_as_longlong(tbb);
        r[2] = ((unsigned long long)(uint32_t)za << 32) | (uint32_t)zb;
        const int nb = nband;
        r[3] = (unsigned long long)nb;  // more than kBandMax in a chunk: the step re-evaluates none
        for (int e = 0; e < min(nb, kBandMax); e++) r[4 + e] = (unsigned long long)bidx[e];
    }
}

__global__ __launch_bounds__(kSelThreads) void step_scan_kernel(const float* __restrict__ V, StepArgs a, SpState sp,
                                                                int nch) {
    const int j = blockIdx.x / nch;
    step_scan_body(j, blockIdx.x - j * nch, V, a, sp);
}

template <int K, int P>
__global__ __launch_bounds__(kSelThreads) void pen_chunk_kernel(const float* __restrict__ V, Operand xop, StepArgs a,
                                                                SpState sp, int nch) {
    const int j = blockIdx.x / nch;
    pen_chunk_body<K, P>(j, blockIdx.x - j * nch, V, xop, a, sp);
}

template <int K, int P>
void launch_step_kind(const float* V, Operand xop, StepArgs a, SpState sp, int nch, hipStream_t st) {
    // constrained kinds: the scans, then the step (which rescans the column itself for an L0 list
    // found too long); penalised kinds: the chunk workgroups, then the step that sums them
    if constexpr (K < 2) step_scan_kernel<<<a.batch * nch, kSelThreads, 0, st>>>(V, a, sp, nch);
    else pen_chunk_kernel<K, P><<<a.batch * nch, kSelThreads, 0, st>>>(V, xop, a, sp, nch);
    step_kernel<kSelThreads, kSelDepth, K, P><<<a.batch, kSelThreads, 0, st>>>(V, xop, a, sp);
}
hipError_t launch_step(const float* V, Operand xop, StepArgs a, SpState sp, hipStream_t st) {
    if (a.n <= 0 || (long)a.n * 4 > 0x7fffffffL) return hipErrorInvalidValue;
    if (a.precision && (long)a.ldv * xop.bp * 2 > 0xffffffffL) return hipErrorInvalidValue;  // 32-bit store offsets
    if (a.precision < 0 || a.precision > 2) return hipErrorInvalidValue;
    const int nch = (a.n + kStepChunk - 1) / kStepChunk;
    if ((long)a.batch * nch > 0x7fffffffL) return hipErrorInvalidValue;
    switch (step_kind(a.formulation, a.s) * 3 + a.precision) {
#define SPCA_STEP_CASE(K, P)                         \
    case K * 3 + P:                                  \
        launch_step_kind<K, P>(V, xop, a, sp, nch, st); \
        break;
        SPCA_STEP_CASE(0, 0) SPCA_STEP_CASE(0, 1) SPCA_STEP_CASE(0, 2)
        SPCA_STEP_CASE(1, 0) SPCA_STEP_CASE(1, 1) SPCA_STEP_CASE(1, 2)
        SPCA_STEP_CASE(2, 0) SPCA_STEP_CASE(2, 1) SPCA_STEP_CASE(2, 2)
        SPCA_STEP_CASE(3, 0) SPCA_STEP_CASE(3, 1) SPCA_STEP_CASE(3, 2)
#undef SPCA_STEP_CASE
    }
    return hipGetLastError();
}

// Support-ordered gather (solo solves of at most kOrderMaxBatch SPs).  nz_list_kernel also
// keys each SP by the index of its largest |x_i|; order_kernel sorts the gathered SPs by that
// key, so neighbours in the order mostly share their supports, and the ordered gather runs the
// listed gather's workgroups as (row block, position in the order), remapped so that each XCD
// runs one contiguous range of that order.  An XCD's concurrent workgroups then gather mostly
// the same columns of the same row block, and the repeats are served by its L2 rather than
// from MALL or HBM.  Every SP's sum is the listed gather's, so the bits are the same.

// position i of the order in SP i's list header (int 2): an SP index, -1 past the gathered SPs
__global__ __launch_bounds__(1024) void order_kernel(int* __restrict__ gl, int batch, SpState sp) {
    __shared__ uint64_t a[kOrderMaxBatch];
    for (int j = threadIdx.x; j < kOrderMaxBatch; j += 1024) {
        uint64_t k = 0;  // padding: sorts last
        if (j < batch) {
            const bool g = !sp.done[0] && sp.done[kSparseFlags + (j >> 8)] && sp.active[j];
            const uint32_t key = g ? (uint32_t)gl[(long)j * kGatherStride + 1] : 0xfffffffeu;
            k = ~(((uint64_t)key << 32) | (uint32_t)j);  // descending on ~(key, j): ascending on (key, j)
        }
        a[j] = k;
    }
    __syncthreads();
    static_assert(kOrderMaxBatch == 4 * 1024, "the sort's register counts");
    if (batch <= 1024) sort_desc_reg<1024, 1>(a, batch);
    else if (batch <= 2048) sort_desc_reg<1024, 2>(a, batch);
    else sort_desc_reg<1024, 4>(a, batch);
    for (int i = threadIdx.x; i < batch; i += 1024) {
        const uint64_t k = ~a[i];
        gl[(long)i * kGatherStride + 2] = (uint32_t)(k >> 32) == 0xfffffffeu ? -1 : (int)(uint32_t)k;
    }
}

__global__ __launch_bounds__(256) void sparse_ax_ordered_kernel(const float* __restrict__ A, int lda, int n,
                                                                const float* __restrict__ X, int ldx,
                                                                float* __restrict__ Z, int ldz, int mp, SpState sp,
                                                                const int* __restrict__ gl, int batch) {
    const int t = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = t / batch, i = t - rb * batch;
    const int j = gl[(long)i * kGatherStride + 2];
    if (j < 0) return;
    __shared__ int sidx[kGatherList];
    __shared__ float sval[kGatherList];
    sparse_ax_listed_body(rb, j, A, lda, n, X, ldx, Z, ldz, mp, sp, gl, sidx, sval);
}

hipError_t launch_sparse_ax_listed(const float* Acm, int lda, int mp, int n, const float* X, int ldx, float* Z,
                                   int ldz, int batch, SpState sp, int* gl, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    if (!gl || (long)n * 4 + 32768 > 0xffffffffL) return hipErrorInvalidValue;  // 32-bit offsets into x
    nz_list_kernel<<<batch, 1024, 0, st>>>(X, ldx, n, sp, gl);
    const int rbs = (mp + kSpRows - 1) / kSpRows;
    if (batch <= kOrderMaxBatch) {
        order_kernel<<<1, 1024, 0, st>>>(gl, batch, sp);
        sparse_ax_ordered_kernel<<<rbs * batch, 256, 0, st>>>(Acm, lda, n, X, ldx, Z, ldz, mp, sp, gl, batch);
    } else {
        sparse_ax_listed_kernel<<<dim3(rbs, batch), 256, 0, st>>>(Acm, lda, n, X, ldx, Z, ldz, mp, sp, gl);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- convergence
// Batch rule (gpu_sparse_PCA_solver.h:176-186, sparse_PCA_solver.h:203-207): error =
// max_j err_j with NaN skipped by the '<' update; stop iff it > 0 and error < tol.
// Per-SP rule: each active SP stops on its own error or at max_iter.
// Also chooses the next iteration's A*X product per group of 256 SPs: the sparse gather
// (done[kSparseFlags + g] = 1) when every SP of the group that keeps iterating has at most
// sparse_cap nonzeros (0: never); the dense product's tiles of that group stand down.
__device__ __forceinline__ void converge_body(SpState sp, int batch, int it, double tol, int max_iter, int per_sp,
                                              int sparse_cap, int* all_done) {
    if (*sp.done) return;
    if (it < 0) it = sp.done[1];  // fused launches read the iteration index from the device
    __shared__ float wmax[4];
    __shared__ int wany[4];
    __shared__ int gdense[kMaxSparseGroups];
    const int groups = (batch + 255) >> 8;
    for (int g = threadIdx.x; g < groups; g += 256) gdense[g] = sparse_cap > 0 ? 0 : 1;
    __syncthreads();
    float e = 0.f;
    int any = 0;
    for (int j = threadIdx.x; j < batch; j += 256) {
        if (!sp.active[j]) continue;
        const float ej = sp.err[j];
        if (e < ej) e = ej;
        sp.iters[j] += 1;
        if (per_sp) {
            if ((it > 0 && (double)ej < tol) || sp.iters[j] >= max_iter) sp.active[j] = 0;
            else any = 1;
        }
        if (sp.active[j] && sp.nnz[j] > sparse_cap) gdense[j >> 8] = 1;  // same value from every writer
    }
    // max with the same NaN-skipping semantics: a NaN never wins a '<' comparison
    for (int o = 32; o > 0; o >>= 1) {
        const float t = __shfl_down(e, o, 64);
        if (e < t) e = t;
        any |= __shfl_down(any, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        wmax[threadIdx.x >> 6] = e;
        wany[threadIdx.x >> 6] = any;
    }
    __syncthreads();
    int nsparse = 0;
    for (int g = threadIdx.x; g < groups; g += 256) {
        sp.done[kSparseFlags + g] = !gdense[g];
        nsparse += !gdense[g];
    }
    nsparse = __syncthreads_or(nsparse);
    if (threadIdx.x == 0) {
        float m = 0.f;
        int a = 0;
        for (int w = 0; w < 4; w++) {
            if (m < wmax[w]) m = wmax[w];
            a |= wany[w];
        }
        sp.errmax[0] = m;
        sp.done[1] += 1;
        int stop;
        if (per_sp) stop = !a;
        else stop = it > 0 && (double)m < tol;
        if (stop || it + 1 >= max_iter) {
            sp.done[0] = 1;
            if (all_done) atomicAdd(all_done, 1);
        }
        if (nsparse && !sp.done[0]) sp.done[4] += 1;  // iterations whose A*X is gathered (some group)
    }
}
__global__ __launch_bounds__(256) void converge_kernel(SpState sp, int batch, int it, double tol, int max_iter,
                                                       int per_sp, int sparse_cap) {
    converge_body(sp, batch, it, tol, max_iter, per_sp, sparse_cap, nullptr);
}
hipError_t launch_converge(SpState sp, int batch, int it, double tol, int max_iter, int per_sp, int sparse_cap,
                           hipStream_t st) {
    converge_kernel<<<1, 256, 0, st>>>(sp, batch, it, tol, max_iter, per_sp, sparse_cap);
    return hipGetLastError();
}

// ---------------------------------------------------------------- fused multi-job iteration
// Each launch covers the work items of all jobs; a workgroup finds its job by binary search
// over the jobs' first item (the job table is a few KB and stays in cache) and runs the same
// body as the solo kernel on the job's own buffers.  Finished jobs return at once.
template <int FIELD>
__device__ __forceinline__ int job_first(const FusedJob* __restrict__ jobs, int k) {
    return FIELD == 0 ? jobs[k].t1 : FIELD == 1 ? jobs[k].t2 : FIELD == 2 ? jobs[k].c0 : FIELD == 3 ? jobs[k].t3
                                                                           : FIELD == 4 ? jobs[k].t4 : jobs[k].t5;
}
template <int FIELD>
__device__ __forceinline__ int find_job(const FusedJob* __restrict__ jobs, int njobs, int item) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (job_first<FIELD>(jobs, mid) <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// WHICH 0: Zpart = A*X (K split), WHICH 1: V = A^T*Z.
// Windowed XCD mapping for the fused tile list: the list is cut into windows of 8*kWin tiles
// and XCD x takes the x-th run of kWin consecutive tiles of each window (the last, partial
// window maps straight through).  All XCDs work inside one window at a time, so the tiles in
// flight span one or two jobs' matrices rather than eight regions of a list that covers
// every job (fewer distinct pages and L2 sets at once), and each XCD still gets runs of
// consecutive tiles that share operand panels.
constexpr int kWin = 96;
__device__ __forceinline__ int xcd_window_remap(int bid, int total) {
    const int win = 8 * kWin;
    const int base = bid / win * win;
    if (base + win > total) return bid;
    const int r = bid - base;
    return base + (r & 7) * kWin + (r >> 3);
}

template <int PREC, int WHICH>
__global__ __launch_bounds__(256) void fused_gemm_kernel(const FusedJob* __restrict__ jobs, int njobs, int total) {
    const int t = xcd_window_remap(blockIdx.x, total);
    const int k = find_job<WHICH>(jobs, njobs, t);
    const FusedJob& J = jobs[k];
    if (*J.sp.done) return;
    const int lt = t - (WHICH == 0 ? J.t1 : J.t2);
    const int nb = WHICH == 0 ? J.nb1 : J.nb2;  // narrow tiles: one column tile
    const int tiles_m = (WHICH == 0 ? J.mp : J.np) / kGemmBM, tiles_n = nb ? 1 : J.bp / kGemmBN;
    const int tn = lt % tiles_n;
    if (WHICH == 0 && J.sp.done[kSparseFlags + (tn >> 1)]) return;  // A*X by the sparse gather
    const int rest = lt / tiles_n;
    const int tm = rest % tiles_m;
    const int sp = rest / tiles_m;
    const int ksplit = WHICH == 0 ? J.np / J.splits : J.mp;
    const long m0 = (long)tm * kGemmBM, n0 = (long)tn * kGemmBN, k0 = (long)sp * ksplit;
    // one LDS block serves whichever tile form the job's product takes
    if (WHICH == 1 && J.p2_prec == 4) return;  // fused_fix8_kernel's tile
    constexpr int kLds = cmax(kF32TileLds, tile16_lds(PREC));
    __shared__ __attribute__((aligned(16))) uint8_t smem[kLds];
    if constexpr (PREC == 0) {
        if (nb) {  // the solo solve's narrow tile, the same bits as the 128 x 128 one
            const float* P = WHICH == 0 ? J.p1_32 : J.p2_32;
            const float* Q = WHICH == 0 ? J.xop.f32 : J.zop.f32;
            const long ld = WHICH == 0 ? J.np : J.mp;
            float* C = WHICH == 0 ? J.Zpart : J.V;
            const long ldc = WHICH == 0 ? J.mp : J.np, sstride = WHICH == 0 ? J.zsplit : 0;
            if (nb == 32)
                gemm_f32_narrow_tile<32>(P, ld, Q, ld, C, ldc, sstride, sp, m0, k0, ksplit, smem);
            else
                gemm_f32_narrow_tile<64>(P, ld, Q, ld, C, ldc, sstride, sp, m0, k0, ksplit, smem);
        } else if constexpr (WHICH == 0)
            gemm_f32_tile(J.p1_32, J.np, J.xop.f32, J.np, J.Zpart, J.mp, J.zsplit, sp, m0, n0, k0, ksplit, smem);
        else
            gemm_f32_tile(J.p2_32, J.mp, J.zop.f32, J.mp, J.V, J.np, 0, 0, m0, n0, k0, ksplit, smem);
    } else {
        if constexpr (WHICH == 0)
            gemm16_tile<PREC>(J.p1_hi, J.p1_lo, J.mp, J.xop.hi, J.xop.lo, J.bp, J.Zpart, J.mp, J.zsplit, sp,
                              m0, n0, k0, ksplit, smem);
        else
            gemm16_tile<PREC>(J.p2_hi, J.p2_lo, J.np, J.zop.hi, J.zop.lo, J.bp, J.V, J.np, 0, 0, m0, n0, k0,
                              ksplit, smem);
    }
}

// The A^T*Z tiles of the jobs whose second product is the fixed-point one (p2_prec 4), over
// the same tile list as fused_gemm_kernel<PREC, 1>, which returns for them; a launch of its
// own keeps the fixed-point tile's 180 registers out of the other tiles' occupancy.
__global__ __launch_bounds__(256, 2) void fused_fix8_kernel(const FusedJob* __restrict__ jobs, int njobs, int total) {
    const int t = xcd_window_remap(blockIdx.x, total);
    const FusedJob& J = jobs[find_job<1>(jobs, njobs, t)];
    if (*J.sp.done || J.p2_prec != 4) return;
    const int lt = t - J.t2, tiles_n = J.bp / kGemmBN;
    __shared__ __attribute__((aligned(16))) uint8_t smem[kFixLds];
    gemm_fix8_tile(J.p2_q, (long)J.np * J.mp, J.p2_qscale, J.np, (const int8_t*)J.zop.hi, J.bp, J.V, J.np,
                   (long)(lt / tiles_n) * kGemmBM, (long)(lt % tiles_n) * kGemmBN, J.mp, smem);
}

// sparse A*X items: (row block, SP) of each job, row block fastest
// The gathered SPs' nonzero lists (jobs whose workspace holds them), one workgroup per column
__global__ __launch_bounds__(1024) void fused_nz_list_kernel(const FusedJob* __restrict__ jobs, int njobs) {
    const int k = find_job<2>(jobs, njobs, blockIdx.x);
    const FusedJob& J = jobs[k];
    if (J.gl) nz_list_body(blockIdx.x - J.c0, J.xop.f32, J.np, J.sa.n, J.sp, J.gl);
}
__global__ __launch_bounds__(256) void fused_sparse_ax_kernel(const FusedJob* __restrict__ jobs, int njobs) {
    const int k = find_job<3>(jobs, njobs, blockIdx.x);
    const FusedJob& J = jobs[k];
    const int li = blockIdx.x - J.t3, rbs = (J.mp + kSpRows - 1) / kSpRows;
    __shared__ int sidx[kGatherList];
    __shared__ float sval[kGatherList];
    if (J.gl)
        sparse_ax_listed_body(li % rbs, li / rbs, J.p2_32, J.mp, J.sa.n, J.xop.f32, J.np, J.Zpart, J.mp, J.mp, J.sp,
                              J.gl, sidx, sval);
    else
        sparse_ax_body(li % rbs, li / rbs, J.p2_32, J.mp, J.sa.n, J.xop.f32, J.np, J.Zpart, J.mp, J.mp, J.sp, sidx,
                       sval);
}

__global__ __launch_bounds__(256) void fused_sum_partials_kernel(const FusedJob* __restrict__ jobs, int njobs) {
    const FusedJob& J = jobs[find_job<4>(jobs, njobs, blockIdx.x)];
    const int li = blockIdx.x - J.t4, rbs = (J.mp + 255) / 256;
    sum_partials_body(li % rbs, li / rbs, J.Zpart, J.splits, J.zsplit, J.mp, J.mp, J.sp);
}
__global__ __launch_bounds__(256) void fused_prep_z_kernel(const FusedJob* __restrict__ jobs, int njobs) {
    const int k = find_job<2>(jobs, njobs, blockIdx.x);
    const FusedJob& J = jobs[k];
    prep_z_body(blockIdx.x - J.c0, J.Zpart, J.psum ? 1 : J.splits, J.zsplit, J.mp, J.mp, J.zop, J.sa.formulation,
                J.zprec, J.sp, J.rf);
}

// The chunk workgroups of the jobs' steps (constrained scans, penalised chunks): item (job, SP,
// chunk), chunk fastest.
template <int PREC>
__global__ __launch_bounds__(kSelThreads) void fused_scan_kernel(const FusedJob* __restrict__ jobs, int njobs) {
    const FusedJob& J = jobs[find_job<5>(jobs, njobs, blockIdx.x)];
    const int nch = (J.sa.n + kStepChunk - 1) / kStepChunk, li = blockIdx.x - J.t5, j = li / nch;
    switch (step_kind(J.sa.formulation, J.sa.s)) {
        case 0:
        case 1: step_scan_body(j, li - j * nch, J.V, J.sa, J.sp); break;
        case 2: pen_chunk_body<2, PREC>(j, li - j * nch, J.V, J.xop, J.sa, J.sp); break;
        default: pen_chunk_body<3, PREC>(j, li - j * nch, J.V, J.xop, J.sa, J.sp); break;
    }
}

// Jobs of one fused launch share the precision (PREC); each job's step kind is chosen per workgroup.
template <int PREC>
__global__ __launch_bounds__(kSelThreads, 4) void fused_step_kernel(const FusedJob* __restrict__ jobs, int njobs) {
    __shared__ SelShared sh;
    const int k = find_job<2>(jobs, njobs, blockIdx.x);
    const FusedJob& J = jobs[k];
    const int col = blockIdx.x - J.c0;
    switch (step_kind(J.sa.formulation, J.sa.s)) {
        case 0: step_body<kSelThreads, kSelDepth, 0, PREC>(sh, col, J.V, J.xop, J.sa, J.sp); break;
        case 1: step_body<kSelThreads, kSelDepth, 1, PREC>(sh, col, J.V, J.xop, J.sa, J.sp); break;
        case 2: step_body<kSelThreads, kSelDepth, 2, PREC>(sh, col, J.V, J.xop, J.sa, J.sp); break;
        default: step_body<kSelThreads, kSelDepth, 3, PREC>(sh, col, J.V, J.xop, J.sa, J.sp); break;
    }
}

__global__ __launch_bounds__(256) void fused_converge_kernel(const FusedJob* __restrict__ jobs, int* all_done) {
    const FusedJob& J = jobs[blockIdx.x];
    converge_body(J.sp, J.sa.batch, -1, J.tol, J.max_iter, J.per_sp, J.sparse_cap, all_done);
}

template <int PREC>
hipError_t fused_launch(const FusedJob* jobs, int njobs, int tiles1, int tiles2, int cols, int items3, int items4,
                        int items5, int any_fix8, int* all_done, hipStream_t st) {
    fused_gemm_kernel<PREC, 0><<<tiles1, 256, 0, st>>>(jobs, njobs, tiles1);
    fused_nz_list_kernel<<<cols, 1024, 0, st>>>(jobs, njobs);
    fused_sparse_ax_kernel<<<items3, 256, 0, st>>>(jobs, njobs);
    if (items4 > 0) fused_sum_partials_kernel<<<items4, 256, 0, st>>>(jobs, njobs);
    fused_prep_z_kernel<<<cols, 256, 0, st>>>(jobs, njobs);
    fused_gemm_kernel<PREC, 1><<<tiles2, 256, 0, st>>>(jobs, njobs, tiles2);
    if (any_fix8) fused_fix8_kernel<<<tiles2, 256, 0, st>>>(jobs, njobs, tiles2);
    fused_scan_kernel<PREC><<<items5, kSelThreads, 0, st>>>(jobs, njobs);
    fused_step_kernel<PREC><<<cols, kSelThreads, 0, st>>>(jobs, njobs);
    fused_converge_kernel<<<njobs, 256, 0, st>>>(jobs, all_done);
    return hipGetLastError();
}

hipError_t launch_fused_iteration(int precision, const FusedJob* jobs, int njobs, int tiles1, int tiles2, int cols,
                                  int items3, int items4, int items5, int any_fix8, int* all_done, hipStream_t st) {
    if (njobs < 1 || tiles1 < 1 || tiles2 < 1 || cols < 1 || items3 < 1 || items4 < 0 || items5 < 1)
        return hipErrorInvalidValue;
    if (precision == 0)
        return fused_launch<0>(jobs, njobs, tiles1, tiles2, cols, items3, items4, items5, any_fix8, all_done, st);
    if (precision == 1)
        return fused_launch<1>(jobs, njobs, tiles1, tiles2, cols, items3, items4, items5, any_fix8, all_done, st);
    return fused_launch<2>(jobs, njobs, tiles1, tiles2, cols, items3, items4, items5, any_fix8, all_done, st);
}

// ---------------------------------------------------------------- OTF slot refill
// Reference OTF loop body (c++/src/gpower/sparse_PCA_solver.h:120-163): every slot counts
// an iteration; a slot terminates when err < tol (its count is >= 1, so the reference's
// "it > 0" holds) or at max_iter; terminated slots update the best with a strict '>' in
// slot order and take the next SP while any remain.  Slots with nothing left keep
// iterating, as in the reference.  One 1024-thread workgroup; slots in chunks of 1024
// with a carried prefix count, so the assignment order is the slot order.
__global__ __launch_bounds__(1024) void otf_converge_kernel(SpState sp, OtfState o, int batch, double tol,
                                                            int max_iter, int total_sp) {
    if (*sp.done) return;
    __shared__ int wsum[16];
    __shared__ float wbest[16];
    __shared__ int wbesti[16];
    __shared__ int wbestg[16];
    __shared__ int s_bglob;
    __shared__ int s_carry, s_any;
    __shared__ float s_best;
    __shared__ int s_bslot;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) {
        s_carry = 0;
        s_any = 0;
        s_best = o.best_val[0];
        s_bslot = -1;
        s_bglob = -1;
    }
    __syncthreads();
    const int gen0 = o.generated[0];
    for (int c0 = 0; c0 < batch; c0 += 1024) {
        const int i = c0 + tid;
        int term = 0;
        float v = 0.f;
        int g = -1;
        if (i < batch) {
            g = o.order[i];
            const int cur = o.cur_it[i] + 1;
            o.cur_it[i] = cur;
            const float e = sp.err[i];
            term = ((double)e < tol) || cur >= max_iter;
            v = sp.val[i];
            if (term) {
                o.sp_val[g] = v;
                o.sp_iters[g] = cur;
                o.sp_nnz[g] = sp.nnz[i];
            } else {
                s_any = 1;  // benign: every writer stores 1
            }
        }
        // best among the terminated slots of this chunk: largest value, first slot on ties
        float bv = (term && v == v) ? v : -INFINITY;
        int bi = (term && v == v) ? i : 0x7fffffff;
        int bg = g;
        for (int off = 32; off > 0; off >>= 1) {
            const float ov = __shfl_down(bv, off, 64);
            const int oi = __shfl_down(bi, off, 64);
            const int og = __shfl_down(bg, off, 64);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; bg = og; }
        }
        // exclusive prefix of terminated slots (slot order)
        const uint64_t bal = __ballot(term);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) { wsum[w] = __popcll(bal); wbest[w] = bv; wbesti[w] = bi; wbestg[w] = bg; }
        // the carried count of earlier chunks, read before the barrier: thread 0 adds this
        // chunk's total to s_carry after it, while other waves may still be ranking
        const int carry = s_carry;
        __syncthreads();
        int off = 0, tot = 0;
        for (int k = 0; k < 16; k++) {
            if (k < w) off += wsum[k];
            tot += wsum[k];
        }
        if (term) {
            const int rank = carry + off + before;
            if (gen0 + rank < total_sp) {
                o.order[i] = gen0 + rank;
                o.cur_it[i] = 0;
                o.refill[i] = 1;
                s_any = 1;
            }
        }
        if (tid == 0) {
            // fold this chunk's best into the running best (strict '>', earlier chunks first)
            float cb = -INFINITY;
            int ci = 0x7fffffff, cg = -1;
            for (int k = 0; k < 16; k++)
                if (wbest[k] > cb || (wbest[k] == cb && wbesti[k] < ci)) { cb = wbest[k]; ci = wbesti[k]; cg = wbestg[k]; }
            if (ci != 0x7fffffff && s_best < cb) { s_best = cb; s_bslot = ci; s_bglob = cg; }
            s_carry += tot;
        }
        __syncthreads();
    }
    if (tid == 0) {
        o.generated[0] = min(total_sp, gen0 + s_carry);
        o.generated[1] = s_bslot;
        if (s_bglob >= 0) o.generated[2] = s_bglob;
        o.best_val[0] = s_best;
        sp.done[1] += 1;
        if (!s_any) sp.done[0] = 1;
    }
}
hipError_t launch_otf_converge(SpState sp, OtfState o, int batch, double tol, int max_iter, int total_sp,
                               hipStream_t st) {
    otf_converge_kernel<<<1, 1024, 0, st>>>(sp, o, batch, tol, max_iter, total_sp);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void otf_refill_kernel(SpState sp, OtfState o, Operand xop, int precision, int n,
                                                          int ldx, int batch) {
    const int i = blockIdx.x;
    const long base = (long)i * ldx;
    if (o.generated[1] == i)
        for (int k = threadIdx.x; k < n; k += 256) o.xbest[k] = xop.f32[base + k];
    if (!o.refill[i]) return;
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const float* src = o.pool + (long)o.order[i] * ldx;
    int c = 0;
    for (int k = threadIdx.x; k < n; k += 256) {
        const float v = src[k];
        write_operand(xop, precision, ldx, k, i, v);
        c += v != 0.f;
    }
    atomicAdd(&cnt, c);  // integer sum: order-free
    __syncthreads();
    if (threadIdx.x == 0) {
        sp.val[i] = 0.f;
        sp.err[i] = 0.f;
        sp.thr[i] = 0.f;  // a new SP has no previous threshold
        sp.nnz[i] = cnt;  // the new iterate's nonzeros, for the sparse A*X decision
        o.refill[i] = 0;
    }
}

// OTF: the next A*X of each group of 256 slots is the sparse gather iff every active slot of
// the group (refilled ones included) has at most sparse_cap nonzeros; counted in done[4]
__global__ __launch_bounds__(256) void otf_sparse_flags_kernel(SpState sp, int batch, int sparse_cap) {
    const int groups = (batch + 255) >> 8;
    int nsparse = 0;
    for (int g = 0; g < groups; g++) {  // uniform trip count: every thread reaches each barrier
        const int j = g * 256 + threadIdx.x;
        const int ok = j >= batch || !sp.active[j] || sp.nnz[j] <= sparse_cap;
        const int flag = __syncthreads_and(ok) && sparse_cap > 0 && !sp.done[0];
        if (threadIdx.x == 0) sp.done[kSparseFlags + g] = flag;
        nsparse += flag;
    }
    if (threadIdx.x == 0 && nsparse) sp.done[4] += 1;  // iterations whose A*X is gathered (some group)
}
hipError_t launch_otf_refill(SpState sp, OtfState o, Operand xop, int precision, int n, int ldx, int batch,
                             int sparse_cap, hipStream_t st) {
    otf_refill_kernel<<<batch, 256, 0, st>>>(sp, o, xop, precision, n, ldx, batch);
    if (sparse_cap > 0) otf_sparse_flags_kernel<<<1, 256, 0, st>>>(sp, batch, sparse_cap);
    return hipGetLastError();
}

__global__ void x_operand_kernel(Operand xop, int precision, int ldx, int batch) {
    const long total = (long)ldx * batch;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
        write_operand(xop, precision, ldx, (int)(e % ldx), (int)(e / ldx), xop.f32[e]);
}
hipError_t launch_x_operand(Operand xop, int precision, int ldx, int batch, hipStream_t st) {
    if (precision == 0) return hipSuccess;
    x_operand_kernel<<<2048, 256, 0, st>>>(xop, precision, ldx, batch);
    return hipGetLastError();
}

// ---------------------------------------------------------------- sparse (CSC) matrices
// Reference sparse_PCA_solver_CSC (c++/src/gpower/sparse_PCA_solver_for_CSC.h:112-495) with
// its implicit centring: z = A x - (mu . x) 1 - (1 . x) r (:215-250) and v = A^T z - (1 . z) mu -
// (r . z) 1 (:287-316).  Each product is a gather per SP: thread per row (CSR) or per column
// (CSC), the SP's column of X or Z read where the nonzeros point, FP64 sums in the stored
// order, one rounding to FP32.
__global__ __launch_bounds__(256) void sp_scalars_kernel(const float* __restrict__ X, long ldx, int len,
                                                         const float* __restrict__ w, double* __restrict__ out, int bp,
                                                         SpState sp) {
    const int b = blockIdx.x;
    if (*sp.done || !sp.active[b]) return;
    __shared__ double red[4];
    const float* x = X + (long)b * ldx;
    double d = 0, s = 0;
    for (int k = threadIdx.x; k < len; k += 256) {
        const double v = (double)x[k];
        s += v;
        if (w) d += (double)w[k] * v;
    }
    d = block_sum<256>(d, red);
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) {
        out[b] = d;
        out[bp + b] = s;
    }
}
hipError_t launch_sp_scalars(const float* X, long ldx, int len, const float* w, double* out, int bp, int batch,
                             SpState sp, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    sp_scalars_kernel<<<batch, 256, 0, st>>>(X, ldx, len, w, out, bp, sp);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void csr_ax_kernel(SparseMat A, const float* __restrict__ X, int ldx,
                                                     float* __restrict__ Z, int ldz, int mp,
                                                     const double* __restrict__ xs, int bp, SpState sp) {
    const int b = blockIdx.y;
    if (*sp.done || !sp.active[b]) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= mp) return;
    double acc = 0;
    if (i < A.m) {
        const float* x = X + (long)b * ldx;
        for (int t = A.rptr[i]; t < A.rptr[i + 1]; t++) acc += (double)A.rval[t] * (double)x[A.cidx[t]];
        acc -= xs[b];
        if (A.rmu) acc -= (double)A.rmu[i] * xs[bp + b];
    }
    Z[(long)b * ldz + i] = (float)acc;
}
hipError_t launch_csr_ax(SparseMat A, const float* X, int ldx, float* Z, int ldz, int mp, const double* xs, int bp,
                         int batch, SpState sp, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    csr_ax_kernel<<<dim3((mp + 255) / 256, batch), 256, 0, st>>>(A, X, ldx, Z, ldz, mp, xs, bp, sp);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void csc_atz_kernel(SparseMat A, const float* __restrict__ Z, int ldz,
                                                      float* __restrict__ V, int ldv, int np,
                                                      const double* __restrict__ zs, int bp, SpState sp) {
    const int b = blockIdx.y;
    if (*sp.done || !sp.active[b]) return;
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= np) return;
    double acc = 0;
    if (j < A.n) {
        const float* z = Z + (long)b * ldz;
        for (int t = A.cptr[j]; t < A.cptr[j + 1]; t++) acc += (double)A.cval[t] * (double)z[A.ridx[t]];
        if (A.mu) acc -= (double)A.mu[j] * zs[bp + b];
        acc -= zs[b];
    }
    V[(long)b * ldv + j] = (float)acc;
}
hipError_t launch_csc_atz(SparseMat A, const float* Z, int ldz, float* V, int ldv, int np, const double* zs, int bp,
                          int batch, SpState sp, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    csc_atz_kernel<<<dim3((np + 255) / 256, batch), 256, 0, st>>>(A, Z, ldz, V, ldv, np, zs, bp, sp);
    return hipGetLastError();
}

// FP64 explained variance over the CSR rows: row block partials per SP, summed in block order.
__global__ __launch_bounds__(256) void csr_ev_partial_kernel(SparseMat A, const float* __restrict__ X, int ldx,
                                                             const double* __restrict__ xs, int bp,
                                                             const double* __restrict__ invn, int l1,
                                                             double* __restrict__ part, int nblk) {
    __shared__ double red[4];
    const int b = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
    double e = 0;
    if (i < A.m) {
        const float* x = X + (long)b * ldx;
        double acc = 0;
        for (int t = A.rptr[i]; t < A.rptr[i + 1]; t++) acc += (double)A.rval[t] * (double)x[A.cidx[t]];
        acc -= xs[b];
        if (A.rmu) acc -= (double)A.rmu[i] * xs[bp + b];
        acc *= invn[b];
        e = l1 ? fabs(acc) : acc * acc;
    }
    e = block_sum<256>(e, red);
    if (threadIdx.x == 0) part[(long)b * nblk + blockIdx.x] = e;
}
size_t csr_ev_scratch_doubles(int m, int batch) { return (size_t)batch * (3 + (m + 255) / 256); }

// ---------------------------------------------------------------- deflation
// Reference SparseDeflationCollection::deflateV (c++/src/gpower/sparse_PCA_solver_for_CSC.h:
// 69-92): for each earlier component p in order (its nonzeros idx[off[c] .. off[c+1]) with
// values val), c = p . col and col -= c p.  One workgroup per active column; the dot product
// is summed in FP64 in a fixed order and each entry is updated once, through write_operand so
// the operand forms of X follow (precision 0 writes FP32 only).
__global__ __launch_bounds__(256) void deflate_kernel(Operand op, int precision, long ld, const int* __restrict__ idx,
                                                      const float* __restrict__ val, const int* __restrict__ off,
                                                      int ncomp, SpState sp) {
    const int j = blockIdx.x;
    if (*sp.done || !sp.active[j]) return;
    __shared__ double red[4];
    const float* col = op.f32 + (long)j * ld;
    for (int c = 0; c < ncomp; c++) {
        const int t0 = off[c], t1 = off[c + 1];
        double d = 0;
        for (int t = t0 + threadIdx.x; t < t1; t += 256) d += (double)col[idx[t]] * (double)val[t];
        d = block_sum<256>(d, red);
        for (int t = t0 + threadIdx.x; t < t1; t += 256) {
            const int i = idx[t];
            write_operand(op, precision, ld, i, j, (float)((double)col[i] - d * (double)val[t]));
        }
        __syncthreads();
    }
}
hipError_t launch_deflate(Operand op, int precision, long ld, const int* idx, const float* val, const int* off,
                          int ncomp, int batch, SpState sp, hipStream_t st) {
    if (ncomp <= 0 || batch <= 0) return hipSuccess;
    if (!idx || !val || !off || precision < 0 || precision > 2) return hipErrorInvalidValue;
    deflate_kernel<<<batch, 256, 0, st>>>(op, precision, ld, idx, val, off, ncomp, sp);
    return hipGetLastError();
}

// ---------------------------------------------------------------- FP64 explained variance
// ev_j = |A xhat_j|_2^2 (L2 variance) or |A xhat_j|_1 (L1 variance), xhat = x/|x|_2, all in
// FP64 from the FP32 master copy of A.  Work is split over (row block, index chunk of x, SP):
// each workgroup walks its chunk's nonzeros in index order (ballot compaction per 256-index
// slab) and accumulates its rows of z = A xhat; the chunk partials are summed in chunk order,
// then rows and row blocks in a fixed order, so the result does not depend on scheduling.
constexpr int kEvThreads = 256;
constexpr int kEvRows = 4 * kEvThreads;  // rows per workgroup, 4 per thread

__global__ __launch_bounds__(1024) void ev_norm_kernel(const float* __restrict__ X, int ldx, int n,
                                                       double* __restrict__ invn) {
    __shared__ double red[1024 / 64];
    const float* x = X + (long)blockIdx.x * ldx;
    double nx = 0;
    for (int k = threadIdx.x; k < n; k += 1024) nx += (double)x[k] * (double)x[k];
    nx = block_sum<1024>(nx, red);
    if (threadIdx.x == 0) invn[blockIdx.x] = 1.0 / sqrt(nx);
}

__global__ __launch_bounds__(kEvThreads) void ev_partial_kernel(const float* __restrict__ A, int lda, int m, int n,
                                                                 const float* __restrict__ X, int ldx,
                                                                 const double* __restrict__ invn, int chunk_len,
                                                                 int chunks, double* __restrict__ part, int ldp) {
    __shared__ int idxs[kEvThreads];
    __shared__ double vals[kEvThreads];
    __shared__ int wcount[kEvThreads / 64];
    const int rb = blockIdx.x, c = blockIdx.y, j = blockIdx.z;
    const float* x = X + (long)j * ldx;
    const double inv = invn[j];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double acc[4] = {0, 0, 0, 0};
    const int k_end = min(n, (c + 1) * chunk_len);
    for (int k0 = c * chunk_len; k0 < k_end; k0 += kEvThreads) {
        const int k = k0 + threadIdx.x;
        const float xv = k < k_end ? x[k] : 0.f;
        const bool nzf = xv != 0.f;
        const uint64_t bal = __ballot(nzf);
        if (lane == 0) wcount[w] = __popcll(bal);
        __syncthreads();
        int off = 0, cnt = 0;
        for (int i = 0; i < kEvThreads / 64; i++) {
            if (i < w) off += wcount[i];
            cnt += wcount[i];
        }
        if (nzf) {
            const int at = off + __popcll(bal & ((1ull << lane) - 1ull));
            idxs[at] = k;
            vals[at] = (double)xv * inv;
        }
        __syncthreads();
        for (int t = 0; t < cnt; t++) {
            const float* col = A + (long)idxs[t] * lda;
            const double v = vals[t];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = rb * kEvRows + threadIdx.x + r * kEvThreads;
                if (i < m) acc[r] += (double)col[i] * v;
            }
        }
        __syncthreads();
    }
    double* pp = part + ((long)j * chunks + c) * ldp;
#pragma unroll
    for (int r = 0; r < 4; r++) pp[rb * kEvRows + threadIdx.x + r * kEvThreads] = acc[r];
}

// one row per thread: workgroup rb covers rows [256 rb, 256 rb + 256)
__global__ __launch_bounds__(kEvThreads) void ev_reduce_kernel(const double* __restrict__ part, int chunks, int ldp,
                                                                int m, int l1, double* __restrict__ rpart) {
    __shared__ double red[kEvThreads / 64];
    const int rb = blockIdx.x, j = blockIdx.y;
    const int i = rb * kEvThreads + threadIdx.x;
    double z = 0;
    for (int c = 0; c < chunks; c++) z += part[((long)j * chunks + c) * ldp + i];
    double s = i < m ? (l1 ? fabs(z) : z * z) : 0.0;
    s = block_sum<kEvThreads>(s, red);
    if (threadIdx.x == 0) rpart[(long)j * gridDim.x + rb] = s;
}

__global__ void ev_final_kernel(const double* __restrict__ rpart, int rbs, int batch, double* __restrict__ ev) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= batch) return;
    double s = 0;
    for (int b = 0; b < rbs; b++) s += rpart[(long)j * rbs + b];
    ev[j] = s;
}

size_t ev_scratch_doubles(int m, int n, int batch, int* chunks_out, int* chunk_len_out) {
    const int rbs = (m + kEvRows - 1) / kEvRows, ldp = rbs * kEvRows;
    // up to 64 chunks of at least 256 indices, within 64 MiB of partials
    const long cap = (64l << 20) / 8 / ((long)batch * ldp);
    int chunks = (int)std::max(1l, std::min(cap, 64l));
    chunks = std::min(chunks, (n + kEvThreads - 1) / kEvThreads);
    int chunk_len = (n + chunks - 1) / chunks;
    chunk_len = (chunk_len + kEvThreads - 1) / kEvThreads * kEvThreads;
    chunks = (n + chunk_len - 1) / chunk_len;
    if (chunks_out) *chunks_out = chunks;
    if (chunk_len_out) *chunk_len_out = chunk_len;
    return (size_t)batch * chunks * ldp + (size_t)batch * (ldp / kEvThreads) + (size_t)batch;
}

hipError_t launch_ev_fp64(const float* Acm, int lda, int m, int n, const float* X, int ldx, int batch, int l1,
                          double* ev, double* work, hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    int chunks = 0, chunk_len = 0;
    ev_scratch_doubles(m, n, batch, &chunks, &chunk_len);
    const int rbs = (m + kEvRows - 1) / kEvRows, ldp = rbs * kEvRows;
    const int rrb = ldp / kEvThreads;  // 256-row blocks of the reduction
    double* part = work;
    double* rpart = part + (size_t)batch * chunks * ldp;
    double* invn = rpart + (size_t)batch * rrb;
    ev_norm_kernel<<<batch, 1024, 0, st>>>(X, ldx, n, invn);
    ev_partial_kernel<<<dim3(rbs, chunks, batch), kEvThreads, 0, st>>>(Acm, lda, m, n, X, ldx, invn, chunk_len, chunks,
                                                                       part, ldp);
    ev_reduce_kernel<<<dim3(rrb, batch), kEvThreads, 0, st>>>(part, chunks, ldp, m, l1, rpart);
    ev_final_kernel<<<(batch + 255) / 256, 256, 0, st>>>(rpart, rrb, batch, ev);
    return hipGetLastError();
}

// mu . x_b and 1 . x_b for the explained variance (every column, FP64, fixed order)
__global__ __launch_bounds__(256) void ev_scalars_kernel(const float* __restrict__ X, int ldx, int n,
                                                         const float* __restrict__ w, double* __restrict__ out,
                                                         int batch) {
    __shared__ double red[4];
    const int b = blockIdx.x;
    const float* x = X + (long)b * ldx;
    double d = 0, s = 0;
    for (int k = threadIdx.x; k < n; k += 256) {
        const double v = (double)x[k];
        s += v;
        if (w) d += (double)w[k] * v;
    }
    d = block_sum<256>(d, red);
    s = block_sum<256>(s, red);
    if (threadIdx.x == 0) {
        out[b] = d;
        out[batch + b] = s;
    }
}
hipError_t launch_csr_ev(SparseMat A, const float* X, int ldx, int batch, int l1, double* ev, double* work,
                         hipStream_t st) {
    if (batch <= 0) return hipSuccess;
    const int nblk = (A.m + 255) / 256;
    double* xs = work;                  // 2 * batch
    double* invn = xs + 2 * batch;      // batch
    double* part = invn + batch;        // nblk * batch
    ev_scalars_kernel<<<batch, 256, 0, st>>>(X, ldx, A.n, A.mu, xs, batch);
    ev_norm_kernel<<<batch, 1024, 0, st>>>(X, ldx, A.n, invn);
    csr_ev_partial_kernel<<<dim3(nblk, batch), 256, 0, st>>>(A, X, ldx, xs, batch, invn, l1, part, nblk);
    ev_final_kernel<<<(batch + 255) / 256, 256, 0, st>>>(part, nblk, batch, ev);
    return hipGetLastError();
}

// ---------------------------------------------------------------- FP64 support re-solve
// y on the support S of the returned iterate: L2 variance maximises |A_S y|_2 (power
// iteration on A_S^T A_S), L1 variance maximises |A_S y|_1 (the AM sign iteration restricted
// to S).  A_S is gathered once into an m x k FP64 matrix; each iteration is three launches
// with fixed-order reductions, so the result is reproducible.
__global__ void gather_support_kernel(const float* __restrict__ A, int lda, int m, const int* __restrict__ idx, int k,
                                      double* __restrict__ As) {
    const long total = (long)m * k;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % m), b = (int)(e / m);
        As[e] = (double)A[(long)idx[b] * lda + i];
    }
}
// z = A_S y, mapped (sign for L1 variance), per-block partial of the objective
__global__ __launch_bounds__(256) void resolve_az_kernel(const double* __restrict__ As, int m, int k,
                                                         const double* __restrict__ y, int l1, double* __restrict__ z,
                                                         double* __restrict__ part) {
    __shared__ double red[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double zi = 0;
    if (i < m)
        for (int b = 0; b < k; b++) zi += As[(long)b * m + i] * y[b];
    const double contrib = i < m ? (l1 ? fabs(zi) : zi * zi) : 0.0;
    if (i < m) z[i] = l1 ? (double)((zi > 0) - (zi < 0)) : zi;
    const double t = block_sum<256>(contrib, red);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
// y' = A_S^T z, one block per support column
__global__ __launch_bounds__(256) void resolve_atz_kernel(const double* __restrict__ As, int m,
                                                          const double* __restrict__ z, double* __restrict__ yp) {
    __shared__ double red[4];
    const int b = blockIdx.x;
    double acc = 0;
    for (int i = threadIdx.x; i < m; i += 256) acc += As[(long)b * m + i] * z[i];
    const double t = block_sum<256>(acc, red);
    if (threadIdx.x == 0) yp[b] = t;
}
// objective of the current y (sum of the partials), normalise y' into y, convergence flag
__global__ __launch_bounds__(256) void resolve_norm_kernel(int k, int nparts, const double* __restrict__ part,
                                                           double* __restrict__ yp, double* __restrict__ y,
                                                           double* __restrict__ ev, int* __restrict__ flag, double tol,
                                                           int it) {
    __shared__ double red[4];
    if (flag[0]) return;
    double e = 0;
    for (int p = threadIdx.x; p < nparts; p += 256) e += part[p];
    e = block_sum<256>(e, red);
    double nn = 0;
    for (int b = threadIdx.x; b < k; b += 256) nn += yp[b] * yp[b];
    nn = block_sum<256>(nn, red);
    const double inv = 1.0 / sqrt(nn);
    if (threadIdx.x == 0) {
        const double prev = ev[0];
        ev[0] = e;  // objective of the y this iteration started from
        if (it > 0 && fabs(e - prev) <= tol * fabs(e)) flag[0] = 1;
    }
    __syncthreads();
    if (flag[0]) return;
    for (int b = threadIdx.x; b < k; b += 256) y[b] = yp[b] * inv;
}
__global__ __launch_bounds__(256) void resolve_sum_kernel(int nparts, const double* __restrict__ part,
                                                          double* __restrict__ ev) {
    __shared__ double red[4];
    double e = 0;
    for (int p = threadIdx.x; p < nparts; p += 256) e += part[p];
    e = block_sum<256>(e, red);
    if (threadIdx.x == 0) ev[0] = e;
}
hipError_t launch_resolve(const float* Acm, int lda, int m, const int* idx, int k, int l1, double* As, double* y,
                          double* z, double* yp, double* part, double* ev, int* flag, double tol, int max_iter,
                          hipStream_t st) {
    gather_support_kernel<<<1024, 256, 0, st>>>(Acm, lda, m, idx, k, As);
    const int nb = (m + 255) / 256;
    for (int it = 0; it < max_iter; it++) {
        resolve_az_kernel<<<nb, 256, 0, st>>>(As, m, k, y, l1, z, part);
        resolve_atz_kernel<<<k, 256, 0, st>>>(As, m, z, yp);
        resolve_norm_kernel<<<1, 256, 0, st>>>(k, nb, part, yp, y, ev, flag, tol, it);
    }
    // objective of the final y
    resolve_az_kernel<<<nb, 256, 0, st>>>(As, m, k, y, l1, z, part);
    resolve_sum_kernel<<<1, 256, 0, st>>>(nb, part, ev);
    return hipGetLastError();
}

// ---------------------------------------------------------------- fixed scratch
// The workspace of the allocation-free solve (spca_solve_noalloc): one array per GPU, in the
// code object's device memory from module load to process end.  Host code never passes the
// array's name to a kernel (in host code it names a host-side stand-in): it takes the device
// address from hipGetSymbolAddress, once per device with that device current.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_rc[kMaxDevices];
bool g_scratch_done[kMaxDevices];
}  // namespace

hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        g_scratch_rc[device] = e;
        g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
        g_scratch_done[device] = true;
    }
    return g_scratch_rc[device];
}

void* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return g_scratch_done[device] ? g_scratch_ptr[device] : nullptr;
}

// 16-byte stores over the first n16 chunks, then the bytes of the tail (< 16) by one thread
__global__ void fill_zero_kernel(uint4* __restrict__ p, size_t n16, int tail) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n16; e += (size_t)gridDim.x * blockDim.x)
        p[e] = make_uint4(0u, 0u, 0u, 0u);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int b = 0; b < tail; b++) ((unsigned char*)(p + n16))[b] = 0;
}
hipError_t launch_fill_zero(void* p, size_t bytes, hipStream_t st) {
    if (!bytes) return hipSuccess;
    if (!p || ((uintptr_t)p & 15)) return hipErrorInvalidValue;
    const size_t n16 = bytes / 16;
    const unsigned grid = (unsigned)std::max<size_t>(1, std::min<size_t>((n16 + 255) / 256, 4096));
    fill_zero_kernel<<<grid, 256, 0, st>>>((uint4*)p, n16, (int)(bytes & 15));
    return hipGetLastError();
}

}  // namespace spca
