/*
 * gemm.hip -- the two GEMMs of the AMICA EM iteration on gfx950 MFMA.
 *
 * Unmixing GEMM (fraimondo/cudamica src/amica.cu:1035-1102: per block and
 * model a cublasDgemm b = W x, a Dgemv W c and the `substract` kernel):
 *   b[t][h*NP + i] = sum_k x[t][k] W_h[i][k] - (W_h c_h)[i]
 * for all models at once, as one [T x NP] x [NP x M*NP] product, in FP64
 * (FP32 path) or split BF16 (BF16 path).
 *
 * Statistic GEMM (amica.cu:1533-1550, cublasDgemm phi = g b^T per block and
 * model).  The engine computes G = sum_t g_t x_t^T for all models at once
 * ([M*NP x T] x [T x NP]) and recovers phi_h = G_h W_h^T - gsum_h (W_h c_h)^T
 * in FP64 in the reduction kernel (b = W x - W c), so x -- shared by all
 * models -- is the second operand instead of the M times larger b.
 *
 * MFMA operand layouts used here (per wave of 64 lanes, l = lane):
 *   v_mfma_f32_16x16x4_f32:   lane l holds A[l%16][l/16], B[l/16][l%16];
 *                             D[(l/16)*4 + r][l%16] in register r.
 *   v_mfma_f64_16x16x4_f64:   same A and B; D[(l/16) + 4r][l%16] in register r
 *                             (measured on gfx950).
 *   v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8*(l/16) .. +7],
 *                             B[8*(l/16) .. +7][l%16]; D as above.
 * Both are checked end to end against the CPU oracle by the statistics
 * parity tests (tests/test_gpu_engine.py).
 */
#include <algorithm>

#include "common.h"
#include "engine_internal.h"

namespace amica {

namespace {

/* ---------------- unmixing GEMM, FP64 ---------------------------------- */
/* FP64 path: b and hence y are formed in FP64, so that statistics with a
 * singular weight at y = 0 (mu_denom sums u |y|^(rho-2)) see the same y as
 * an FP64 implementation.  Workgroup tile 64 samples x 64 columns, K staged
 * through LDS in chunks of 32; each of the 4 waves owns a 32 x 32 quarter
 * (2 x 2 v_mfma_f64_16x16x4 tiles). */
__global__ __launch_bounds__(256) void unmix_f64_kernel(const double* __restrict__ x,
                                                        const double* __restrict__ Wg,
                                                        const double* __restrict__ wc,
                                                        double* __restrict__ b, long T, int NP,
                                                        int ldb) {
    __shared__ __attribute__((aligned(16))) double xs[64][34];
    __shared__ __attribute__((aligned(16))) double ws[64][34];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wcn = w & 1;
    const long t0 = (long)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    f64x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = (f64x4){0.0, 0.0, 0.0, 0.0};

    for (int k0 = 0; k0 < NP; k0 += 32) {
#pragma unroll
        for (int rep = 0; rep < 4; ++rep) {
            const int q = tid + rep * 256;          /* 64 rows x 16 double2 */
            const int r = q >> 4, c2 = (q & 15) * 2;
            const long t = t0 + r;
            double2 xv = make_double2(0.0, 0.0);
            if (t < T) xv = *(const double2*)(x + (size_t)t * NP + k0 + c2);
            *(double2*)&xs[r][c2] = xv;
            *(double2*)&ws[r][c2] = *(const double2*)(Wg + (size_t)(c0 + r) * NP + k0 + c2);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 32; kk += 4) {
            const int kr = kk + (lane >> 4);
            double av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = xs[wr * 32 + a * 16 + (lane & 15)][kr];
#pragma unroll
            for (int c = 0; c < 2; ++c) bv[c] = ws[wcn * 32 + c * 16 + (lane & 15)][kr];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int c = 0; c < 2; ++c)
                    acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int col = c0 + wcn * 32 + c * 16 + (lane & 15);
        const double bias = wc[col];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                /* f64 MFMA: register r of lane l holds row (l/16) + 4r */
                const long t = t0 + wr * 32 + a * 16 + (lane >> 4) + 4 * r;
                if (t < T) b[(size_t)t * ldb + col] = acc[a][c][r] - bias;
            }
    }
}

/* ---------------- unmixing GEMM, split BF16 ---------------------------- */
/* x (float32) is split in registers into x_hi = bf16(x) and
 * x_lo = bf16(x - x_hi); W arrives as bf16 hi and lo tables.  Three MFMAs
 * per fragment (hi*hi, hi*lo, lo*hi) give b to ~2^-16 relative at BF16
 * MFMA rate.  Operands are K-contiguous: each lane loads its 8 elements of
 * a fragment straight from global memory (W is L2-resident).  Wave tile 16
 * samples x 64 columns; workgroup 64 x 64. */
__device__ __forceinline__ void split_bf16x8(const float4 a, const float4 b, i16x8& hi, i16x8& lo) {
    const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const uint16_t h = f32_to_bf16_bits(v[e]);
        hi[e] = (short)h;
        lo[e] = (short)f32_to_bf16_bits(v[e] - bf16_bits_to_f32(h));
    }
}

__global__ __launch_bounds__(256) void unmix_bf16x3_kernel(const float* __restrict__ x,
                                                           const uint16_t* __restrict__ Whi,
                                                           const uint16_t* __restrict__ Wlo,
                                                           const float* __restrict__ wc,
                                                           float* __restrict__ b, long T, int NP,
                                                           int ldb) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long t0 = (long)blockIdx.x * 64 + w * 16;
    const int c0 = blockIdx.y * 64;
    const long trow = t0 + (lane & 15);
    const int kofs = (lane >> 4) * 8;
    f32x4 acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < NP; k0 += 32) {
        float4 x0 = make_float4(0.f, 0.f, 0.f, 0.f), x1 = x0;
        if (trow < T) {
            const float* xp = x + (size_t)trow * NP + k0 + kofs;
            x0 = *(const float4*)xp;
            x1 = *(const float4*)(xp + 4);
        }
        i16x8 ahi, alo;
        split_bf16x8(x0, x1, ahi, alo);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t wofs = (size_t)(c0 + c * 16 + (lane & 15)) * NP + k0 + kofs;
            const i16x8 bhi = *(const i16x8*)(Whi + wofs);
            const i16x8 blo = *(const i16x8*)(Wlo + wofs);
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, alo),
                                                             __builtin_bit_cast(bf16x8, bhi), acc[c], 0, 0, 0);
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ahi),
                                                             __builtin_bit_cast(bf16x8, blo), acc[c], 0, 0, 0);
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ahi),
                                                             __builtin_bit_cast(bf16x8, bhi), acc[c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int col = c0 + c * 16 + (lane & 15);
        const float bias = wc[col];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long t = t0 + (lane >> 4) * 4 + r;
            if (t < T) b[(size_t)t * ldb + col] = acc[c][r] - bias;
        }
    }
}

/* 8 FP8 e4m3 values (two dwords) -> 8 bf16, exactly (e4m3 has 3 mantissa
 * bits; the FP32 value the hardware conversion gives is truncated to bf16
 * without loss). */
__device__ __forceinline__ i16x8 fp8x8_to_bf16x8(uint2 q) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 v0 = __builtin_amdgcn_cvt_pk_f32_fp8((int)q.x, false);
    const f2 v1 = __builtin_amdgcn_cvt_pk_f32_fp8((int)q.x, true);
    const f2 v2 = __builtin_amdgcn_cvt_pk_f32_fp8((int)q.y, false);
    const f2 v3 = __builtin_amdgcn_cvt_pk_f32_fp8((int)q.y, true);
    i16x8 r;
    r[0] = (short)(__float_as_uint(v0.x) >> 16); r[1] = (short)(__float_as_uint(v0.y) >> 16);
    r[2] = (short)(__float_as_uint(v1.x) >> 16); r[3] = (short)(__float_as_uint(v1.y) >> 16);
    r[4] = (short)(__float_as_uint(v2.x) >> 16); r[5] = (short)(__float_as_uint(v2.y) >> 16);
    r[6] = (short)(__float_as_uint(v3.x) >> 16); r[7] = (short)(__float_as_uint(v3.y) >> 16);
    return r;
}

/* FP8 unmixing for any NP (the register-resident kernel below covers
 * NP <= 256): b = q W_hi + q W_lo - W~ c, q the e4m3 data converted exactly
 * to bf16 in registers; W fragments straight from global memory (L2-
 * resident).  Wave tile 16 samples x 64 columns, workgroup 64 x 64. */
__global__ __launch_bounds__(256) void unmix_fp8_tiled_kernel(const uint8_t* __restrict__ xq,
                                                              const uint16_t* __restrict__ Whi,
                                                              const uint16_t* __restrict__ Wlo,
                                                              const float* __restrict__ wc,
                                                              float* __restrict__ b, long T, int NP,
                                                              int ldb) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long t0 = (long)blockIdx.x * 64 + w * 16;
    const int c0 = blockIdx.y * 64;
    const long trow = t0 + (lane & 15);
    const int kofs = (lane >> 4) * 8;
    f32x4 acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < NP; k0 += 32) {
        uint2 q = make_uint2(0u, 0u);
        if (trow < T) q = *(const uint2*)(xq + (size_t)trow * NP + k0 + kofs);
        const i16x8 a = fp8x8_to_bf16x8(q);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t wofs = (size_t)(c0 + c * 16 + (lane & 15)) * NP + k0 + kofs;
            const i16x8 bhi = *(const i16x8*)(Whi + wofs);
            const i16x8 blo = *(const i16x8*)(Wlo + wofs);
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                             __builtin_bit_cast(bf16x8, blo), acc[c], 0, 0, 0);
            acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                             __builtin_bit_cast(bf16x8, bhi), acc[c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int col = c0 + c * 16 + (lane & 15);
        const float bias = wc[col];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long t = t0 + (lane >> 4) * 4 + r;
            if (t < T) b[(size_t)t * ldb + col] = acc[c][r] - bias;
        }
    }
}

/* Register-resident unmixing, b = x W~^T - W~ c with W~ = W_hi + W_lo (bf16):
 *   Q8 = false: x pre-split into bf16 hi and lo (written once per recording
 *     by the whitening kernel), b = x_hi W_hi + x_hi W_lo + x_lo W_hi
 *     (~2^-16 relative);
 *   Q8 = true: x stored as FP8 e4m3 q with the per-channel scales folded
 *     into W (W~ = split(W diag(s))), b = q W_hi + q W_lo, q exact in bf16.
 * A workgroup of 8 waves covers one group of 128 CT columns (blockIdx.y);
 * wave w owns 16 CT of them and keeps their W_hi / W_lo fragments for all
 * NP = 32 KS inputs in registers for the whole launch (waves past ldb only
 * load).  The workgroup walks 16-sample tiles (tile = blockIdx.x, +=
 * gridDim.x): every tile is read from HBM once per column group, 16 bytes
 * per lane, kUnmixStages tiles ahead of its use, written to one of two LDS
 * slots (rows padded by 16 bytes, so that the 16 lanes reading one
 * fragment column hit distinct banks) and read from there by every wave.
 * One barrier per tile.  A wave's 16 x 16 CT results go through its own
 * LDS tile so that each store instruction writes whole 64-byte runs of
 * 16-byte lanes (16 dword stores of 4 rows x 64 bytes measured 1.21 ms at
 * 128 ch x 2M x 4 models, these 1.02 ms).  HBM traffic per launch: x (4 NP
 * or NP bytes per sample, per column group) + b (4 ldb bytes per sample). */
constexpr int kUnmixStages = 4;

/* The workgroup's tiles: tile = first + j stride for j < ntl; column group
 * cgroup (of 128 CT columns). */
template <int CT, int KS, bool Q8, int TPW = 1>
__device__ __forceinline__ void unmix_lds_body(const void* __restrict__ xa,
                                               const uint16_t* __restrict__ xlo,
                                               const uint16_t* __restrict__ Whi,
                                               const uint16_t* __restrict__ Wlo,
                                               const float* __restrict__ wc,
                                               float* __restrict__ b, const long T, const int ldb,
                                               const long first, const long stride, const long ntl,
                                               const int cgroup) {
    constexpr int NP = 32 * KS;
    /* TPW 16-sample tiles per step, each computed by WPT waves: recordings
     * whose columns fill fewer than 8 waves (one 64-channel model: ldb = 64)
     * put the other waves on further tiles instead of leaving them idle */
    constexpr int WPT = 8 / TPW;
    constexpr int ROWS = 16 * TPW;                 /* sample rows per step */
    constexpr int PARTS = Q8 ? 1 : 2;              /* fp8 q, or bf16 hi and lo */
    constexpr int ROWB = Q8 ? NP : 2 * NP;         /* bytes of one sample row of one part */
    constexpr int RS = ROWB + 16;                  /* LDS row stride */
    constexpr int SLOT = PARTS * ROWS * RS;        /* LDS bytes of one step */
    constexpr int CPR = ROWB / 16;                 /* 16-byte chunks per row */
    constexpr int CHUNKS = PARTS * ROWS * CPR;     /* chunks per step */
    constexpr int NCH = (CHUNKS + 511) / 512;      /* chunks per lane */
    /* stages in flight: 3 where the W fragments of a split-BF16 NP = 256
     * wave (2 x 8 pairs) leave no room for a fourth without spilling, 2
     * where a step holds several tiles */
    constexpr int D = TPW > 1 ? 2 : (!Q8 && KS == 8) ? kUnmixStages - 1 : kUnmixStages;
    __shared__ __align__(16) unsigned char lds[2 * SLOT];
    /* per wave: its 16 x 16 CT output tile, rows padded by 4 floats */
    constexpr int OST = 16 * CT + 4;
    __shared__ __align__(16) float ostage[8][16 * OST];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int tu = w / WPT;                        /* this wave's tile of the step */
    const int c0 = (cgroup * WPT + (w - tu * WPT)) * 16 * CT;
    const bool computes = c0 < ldb;                /* wave-uniform */
    const int kofs = (lane >> 4) * 8;
    i16x8 bhi[CT][KS], blo[CT][KS];
    float bias[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int col = c0 + c * 16 + (lane & 15);
        bias[c] = computes ? wc[col] : 0.f;
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const size_t wofs = (size_t)col * NP + k * 32 + kofs;
            bhi[c][k] = computes ? *(const i16x8*)(Whi + wofs) : (i16x8){0, 0, 0, 0, 0, 0, 0, 0};
            blo[c][k] = computes ? *(const i16x8*)(Wlo + wofs) : (i16x8){0, 0, 0, 0, 0, 0, 0, 0};
        }
    }
    auto gload = [&](long j, uint4 (&r)[NCH]) {
        const long tile = first + j * stride;
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
            const int c = tid + q * 512;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (c < CHUNKS && j < ntl) {
                const int part = c / (ROWS * CPR);
                const int rem = c - part * ROWS * CPR;
                const int srow = rem / CPR, cc = rem - srow * CPR;
                const long t = tile * ROWS + srow;
                if (t < T) {
                    const unsigned char* src =
                        Q8 ? (const unsigned char*)xa + (size_t)t * NP + cc * 16
                           : (const unsigned char*)(part ? xlo : (const uint16_t*)xa) + (size_t)t * NP * 2 + cc * 16;
                    v = *(const uint4*)src;
                }
            }
            r[q] = v;
        }
    };
    auto lwrite = [&](int slot, const uint4 (&r)[NCH]) {
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
            const int c = tid + q * 512;
            if (c < CHUNKS) {
                const int part = c / (ROWS * CPR);
                const int rem = c - part * ROWS * CPR;
                const int srow = rem / CPR, cc = rem - srow * CPR;
                *(uint4*)(lds + slot * SLOT + part * ROWS * RS + srow * RS + cc * 16) = r[q];
            }
        }
    };
    auto compute = [&](int slot, long j) {
        const long tile = first + j * stride;
        const unsigned char* base = lds + slot * SLOT + (tu * 16 + (lane & 15)) * RS;
        f32x4 acc[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[c] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            i16x8 ahi, alo;
            if constexpr (Q8) {
                ahi = fp8x8_to_bf16x8(*(const uint2*)(base + k * 32 + kofs));
            } else {
                ahi = *(const i16x8*)(base + (k * 32 + kofs) * 2);
                alo = *(const i16x8*)(base + ROWS * RS + (k * 32 + kofs) * 2);
            }
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                if constexpr (!Q8)
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, alo),
                                                                     __builtin_bit_cast(bf16x8, bhi[c][k]), acc[c], 0, 0, 0);
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ahi),
                                                                 __builtin_bit_cast(bf16x8, blo[c][k]), acc[c], 0, 0, 0);
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ahi),
                                                                 __builtin_bit_cast(bf16x8, bhi[c][k]), acc[c], 0, 0, 0);
            }
        }
        /* through LDS: each store instruction then writes 16 CT
         * consecutive floats of 64 / (4 CT) rows, 16 bytes per lane */
        float* os = ostage[w];
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                os[((lane >> 4) * 4 + r) * OST + c * 16 + (lane & 15)] = acc[c][r] - bias[c];
        constexpr int LPR = 4 * CT;                  /* lanes per row (16 bytes each) */
        constexpr int RPI = 64 / LPR;                /* rows per store instruction */
        const int rl = lane / LPR, cl = (lane - rl * LPR) * 4;
#pragma unroll
        for (int i = 0; i < 16 / RPI; ++i) {
            const int row = i * RPI + rl;
            const long t = tile * ROWS + tu * 16 + row;
            const f32x4 v = *(const f32x4*)(os + row * OST + cl);
            if (t < T) *(f32x4*)(b + (size_t)t * ldb + c0 + cl) = v;
        }
    };

    uint4 R[D][NCH];
#pragma unroll
    for (int d = 0; d < D; ++d) gload(d, R[d]);
    /* unrolled by D so that stage u always lives in R[u] */
    for (long j0 = 0; j0 < ntl; j0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const long j = j0 + u;
            if (j < ntl) {                         /* block-uniform */
                lwrite((int)(j & 1), R[u]);        /* tile j, loaded D tiles ago */
                gload(j + D, R[u]);
                __syncthreads();                   /* slot j&1 written; slot (j-1)&1 free */
                if (computes) compute((int)(j & 1), j);
            }
        }
    }
}

/* nsteps steps of TPW 16-sample tiles: step = blockIdx.x, += gridDim.x */
template <int CT, int KS, bool Q8, int TPW>
__global__ __launch_bounds__(512) void unmix_lds_kernel(const void* __restrict__ xa,
                                                         const uint16_t* __restrict__ xlo,
                                                         const uint16_t* __restrict__ Whi,
                                                         const uint16_t* __restrict__ Wlo,
                                                         const float* __restrict__ wc,
                                                         float* __restrict__ b, long T, int ldb,
                                                         long nsteps) {
    const long first = blockIdx.x, stride = gridDim.x;
    const long ntl = first < nsteps ? (nsteps - 1 - first) / stride + 1 : 0;   /* block-uniform */
    unmix_lds_body<CT, KS, Q8, TPW>(xa, xlo, Whi, Wlo, wc, b, T, ldb, first, stride, ntl, (int)blockIdx.y);
}

/* Many recordings in one launch (same NP, same column groups): workgroup b
 * runs work item b, the 16-sample tiles [item.t0, item.t1) of recording
 * item.rec. */
template <int CT, int KS, bool Q8>
__global__ __launch_bounds__(512) void unmix_lds_batch_kernel(const UnmixArgs* __restrict__ args,
                                                               const BatchItem* __restrict__ items) {
    const BatchItem it = items[blockIdx.x];
    const UnmixArgs u = args[it.rec];
    unmix_lds_body<CT, KS, Q8>(u.xa, u.xlo, u.Whi, u.Wlo, u.wc, u.b, u.T, u.ldb, it.t0, 1,
                               it.t1 - it.t0, (int)blockIdx.y);
}

/* FP64 unmixing with register-resident W (the FP32 path's GEMM1): the
 * layout of unmix_lds_kernel with v_mfma_f64_16x16x4f64.  Wave w of the
 * workgroup's NT / 64 owns 16 CT columns and keeps W[col][4 ks + l/16] for its
 * KS = NP / 4 k-steps in registers; 16-sample tiles of x (FP64) come from
 * HBM kUnmixStages tiles ahead, 16 bytes per lane, through a double-buffered
 * LDS slot (rows padded by 16 bytes: the 16 lanes reading one A fragment
 * column hit distinct banks); one barrier per tile.  FP64 MFMA output: register
 * r of lane l holds sample (l/16) + 4r of column l%16. */
template <int CT, int KS, int NT>
__global__ __launch_bounds__(NT) void unmix_f64_lds_kernel(const double* __restrict__ x,
                                                             const double* __restrict__ Wg,
                                                             const double* __restrict__ wc,
                                                             double* __restrict__ b, long T, int ldb,
                                                             long ntiles) {
    constexpr int NP = 4 * KS;
    constexpr int ROWB = 8 * NP;                   /* bytes of one sample row */
    constexpr int RS = ROWB + 16;
    constexpr int SLOT = 16 * RS;
    constexpr int CPR = ROWB / 16;
    constexpr int CHUNKS = 16 * CPR;
    constexpr int NCH = (CHUNKS + NT - 1) / NT;
    constexpr int D = kUnmixStages;
    __shared__ __align__(16) unsigned char lds[2 * SLOT];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int c0 = (blockIdx.y * (NT / 64) + w) * 16 * CT;
    const bool computes = c0 < ldb;                /* wave-uniform */
    double wf[CT][KS];
    double bias[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int col = c0 + c * 16 + (lane & 15);
        bias[c] = computes ? wc[col] : 0.0;
#pragma unroll
        for (int k = 0; k < KS; ++k)
            wf[c][k] = computes ? Wg[(size_t)col * NP + 4 * k + (lane >> 4)] : 0.0;
    }
    const long first = blockIdx.x, stride = gridDim.x;
    const long ntl = first < ntiles ? (ntiles - 1 - first) / stride + 1 : 0;   /* block-uniform */

    auto gload = [&](long j, uint4 (&r)[NCH]) {
        const long tile = first + j * stride;
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
            const int c = tid + q * NT;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (c < CHUNKS && j < ntl) {
                const int srow = c / CPR, cc = c - srow * CPR;
                const long t = tile * 16 + srow;
                if (t < T) v = *(const uint4*)((const unsigned char*)x + (size_t)t * ROWB + cc * 16);
            }
            r[q] = v;
        }
    };
    auto lwrite = [&](int slot, const uint4 (&r)[NCH]) {
#pragma unroll
        for (int q = 0; q < NCH; ++q) {
            const int c = tid + q * NT;
            if (c < CHUNKS) {
                const int srow = c / CPR, cc = c - srow * CPR;
                *(uint4*)(lds + slot * SLOT + srow * RS + cc * 16) = r[q];
            }
        }
    };
    auto compute = [&](int slot, long j) {
        const long tile = first + j * stride;
        const double* base = (const double*)(lds + slot * SLOT + (lane & 15) * RS) + (lane >> 4);
        f64x4 acc[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[c] = (f64x4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < KS; ++k) {
            const double av = base[4 * k];
#pragma unroll
            for (int c = 0; c < CT; ++c)
                acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, wf[c][k], acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const int col = c0 + c * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long t = tile * 16 + (lane >> 4) + 4 * r;
                if (t < T) b[(size_t)t * ldb + col] = acc[c][r] - bias[c];
            }
        }
    };

    uint4 R[D][NCH];
#pragma unroll
    for (int d = 0; d < D; ++d) gload(d, R[d]);
    for (long j0 = 0; j0 < ntl; j0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const long j = j0 + u;
            if (j < ntl) {                         /* block-uniform */
                lwrite((int)(j & 1), R[u]);
                gload(j + D, R[u]);
                __syncthreads();
                if (computes) compute((int)(j & 1), j);
            }
        }
    }
}

/* ---------------- plain FP32 unmixing (comparator path) ------------- */
/* b = x W^T - W c with v_mfma_f32_16x16x4f32 on FP32 operands: the path
 * whose GEMMs are plain FP32, against which the split-BF16 path is compared.
 * Register-resident W as in the FP64 kernel above, with the reduction index
 * permuted so that every load is 16 bytes: lane group q = l / 16 covers the
 * contiguous inputs [q KS, (q + 1) KS) (KS = NP / 4), one per k-step, so a
 * lane holds W[col][q KS + s] for its KS k-steps (CT column blocks of 16)
 * and reads x[sample][q KS + s .. + 3] of the staged tile with one LDS load
 * per 4 k-steps.  16-sample tiles of x come from HBM kUnmixStages tiles
 * ahead, 16 bytes per lane, through a double-buffered LDS slot (rows padded
 * by 16 bytes: the 16 rows of one lane group hit distinct banks); one
 * barrier per tile.  Output: register r of lane l holds sample 4 (l/16) + r
 * of column l % 16. */
template <int CT, int KS>
__global__ __launch_bounds__(512) void unmix_f32_lds_kernel(const float* __restrict__ x,
                                                            const float* __restrict__ Wg,
                                                            const float* __restrict__ wc,
                                                            float* __restrict__ b, long T, int ldb,
                                                            long ntiles) {
    constexpr int NP = 4 * KS;
    constexpr int ROWB = 4 * NP;                   /* bytes of one sample row */
    constexpr int RS = ROWB + 16;
    constexpr int SLOT = 16 * RS;
    constexpr int CPR = ROWB / 16;
    constexpr int CHUNKS = 16 * CPR;
    constexpr int NCH = (CHUNKS + 511) / 512;
    constexpr int D = kUnmixStages;
    __shared__ __align__(16) unsigned char lds[2 * SLOT];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int q = lane >> 4;
    const int c0 = (blockIdx.y * 8 + w) * 16 * CT;
    const bool computes = c0 < ldb;                /* wave-uniform */
    f32x4 wf[CT][KS / 4];
    float bias[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int col = c0 + c * 16 + (lane & 15);
        bias[c] = computes ? wc[col] : 0.f;
#pragma unroll
        for (int s4 = 0; s4 < KS / 4; ++s4)
            wf[c][s4] = computes ? *(const f32x4*)(Wg + (size_t)col * NP + q * KS + 4 * s4)
                                 : (f32x4){0.f, 0.f, 0.f, 0.f};
    }
    const long first = blockIdx.x, stride = gridDim.x;
    const long ntl = first < ntiles ? (ntiles - 1 - first) / stride + 1 : 0;   /* block-uniform */

    auto gload = [&](long j, uint4 (&r)[NCH]) {
        const long tile = first + j * stride;
#pragma unroll
        for (int u = 0; u < NCH; ++u) {
            const int c = tid + u * 512;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (c < CHUNKS && j < ntl) {
                const int srow = c / CPR, cc = c - srow * CPR;
                const long t = tile * 16 + srow;
                if (t < T) v = *(const uint4*)((const unsigned char*)x + (size_t)t * ROWB + cc * 16);
            }
            r[u] = v;
        }
    };
    auto lwrite = [&](int slot, const uint4 (&r)[NCH]) {
#pragma unroll
        for (int u = 0; u < NCH; ++u) {
            const int c = tid + u * 512;
            if (c < CHUNKS) {
                const int srow = c / CPR, cc = c - srow * CPR;
                *(uint4*)(lds + slot * SLOT + srow * RS + cc * 16) = r[u];
            }
        }
    };
    auto compute = [&](int slot, long j) {
        const long tile = first + j * stride;
        const unsigned char* base = lds + slot * SLOT + (lane & 15) * RS + q * KS * 4;
        f32x4 acc[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[c] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s4 = 0; s4 < KS / 4; ++s4) {
            const f32x4 av = *(const f32x4*)(base + 16 * s4);
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int c = 0; c < CT; ++c)
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[e], wf[c][s4][e], acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            const int col = c0 + c * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long t = tile * 16 + q * 4 + r;
                if (t < T) b[(size_t)t * ldb + col] = acc[c][r] - bias[c];
            }
        }
    };

    uint4 R[D][NCH];
#pragma unroll
    for (int d = 0; d < D; ++d) gload(d, R[d]);
    for (long j0 = 0; j0 < ntl; j0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const long j = j0 + u;
            if (j < ntl) {                         /* block-uniform */
                lwrite((int)(j & 1), R[u]);
                gload(j + D, R[u]);
                __syncthreads();
                if (computes) compute((int)(j & 1), j);
            }
        }
    }
}

/* Plain FP32 unmixing for any NP (the inputs past 256): one wave per 16
 * samples x 16 columns, both operands read from global memory (the W
 * columns stay in the caches), the same permuted reduction index. */
__global__ __launch_bounds__(256) void unmix_f32_kernel(const float* __restrict__ x,
                                                        const float* __restrict__ Wg,
                                                        const float* __restrict__ wc,
                                                        float* __restrict__ b, long T, int NP,
                                                        int ldb) {
    const int lane = threadIdx.x & 63, q = lane >> 4;
    const long t0 = (long)blockIdx.x * 16;
    const int col = (blockIdx.y * 4 + (threadIdx.x >> 6)) * 16 + (lane & 15);
    const int KS = NP / 4;
    const long ta = t0 + (lane & 15);
    const float* xr = x + (size_t)(ta < T ? ta : T - 1) * NP + q * KS;
    const float* wr = Wg + (size_t)col * NP + q * KS;
    f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int s4 = 0; s4 < KS; s4 += 4) {
        const f32x4 av = *(const f32x4*)(xr + s4);
        const f32x4 wv = *(const f32x4*)(wr + s4);
#pragma unroll
        for (int e = 0; e < 4; ++e) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[e], wv[e], acc, 0, 0, 0);
    }
    const float bias = wc[col];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const long t = t0 + q * 4 + r;
        if (t < T) b[(size_t)t * ldb + col] = acc[r] - bias;
    }
}

/* ---------------- statistic GEMM G = g^T x, FP32 --------------------- */
/* FP32 statistic GEMM in slabs: workgroup = ldb/64 waves over one chunk of
 * samples (at most a few thousand, so the FP32 accumulation inside the MFMA
 * stays short); wave w owns g rows [64w, 64w+64) x the 16 CT x columns of
 * slab blockIdx.y, 4 CT accumulators of v_mfma_f32_16x16x4f32; every g
 * element is read once per slab, x once per workgroup (the other waves hit
 * the caches), each lane's operands of one sample in 16-byte loads (164
 * VGPRs at CT = 8, 3 waves per SIMD: 3.10 -> 2.94 ms at 128 ch x 2M x 4
 * models against 4-byte loads).  FP64 partial per chunk.  Where g has
 * fewer than 512 columns (rw = ldb/64 row waves), a workgroup of 8 waves
 * runs S = 8 / rw sub-ranges of its chunk side by side (each as long as a
 * chunk of the one-wave-per-row layout) and adds their totals in FP64
 * through LDS in a fixed order before the one partial is written. */
template <int CT>
__global__ __launch_bounds__(512) void gx_f32_slab_kernel(const float* __restrict__ g,
                                                          const float* __restrict__ x,
                                                          double* __restrict__ part, long T,
                                                          int NP, int ldb, long chunk) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rw = min(ldb - (int)blockIdx.z * 512, 512) / 64;   /* row waves */
    /* sample sub-ranges: only below 512 columns, where the launch gives 8
     * waves for fewer row waves and the LDS tile with them */
    const int S = ldb < 512 ? max(1, (int)(blockDim.x >> 6) / rw) : 1;
    const int sr = w / rw, wr = w - sr * rw;
    const int r0 = blockIdx.z * 512 + wr * 64;     /* grid.z: groups of 512 g columns */
    if (S == 1 && r0 >= ldb) return;                /* no barriers below when S = 1 */
    const int k0 = blockIdx.y * 16 * CT;
    long ts = (long)blockIdx.x * chunk;
    long te = min(T, ts + chunk);
    {
        const long sub = ((chunk + S - 1) / S + 3) / 4 * 4;      /* whole 4-sample steps */
        ts = min(te, ts + sr * sub);
        te = min(te, ts + sub);
    }
    f32x4 acc[4][CT];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = (f32x4){0.f, 0.f, 0.f, 0.f};
    /* MFMA (rt, ct) row i and column j stand for g row r0 + 4 i + rt and x
     * column k0 + CT j + ct, so that a lane's 4 A values and CT B values of
     * one sample are consecutive in memory (16-byte loads) */
    f32x4 A;
    float B[CT];
    auto load = [&](long t4) {
        const long t = t4 + (lane >> 4);
        const bool ok = t < te;
        A = ok ? *(const f32x4*)(g + (size_t)t * ldb + r0 + 4 * (lane & 15)) : (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c4 = 0; c4 < CT; c4 += 4) {
            const f32x4 v = ok ? *(const f32x4*)(x + (size_t)t * NP + k0 + CT * (lane & 15) + c4)
                               : (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int e = 0; e < 4; ++e) B[c4 + e] = v[e];
        }
    };
    load(ts);
    for (long t4 = ts; t4 < te; t4 += 4) {
        const f32x4 Ac = A;
        float Bc[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) Bc[ct] = B[ct];
        load(t4 + 4);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ac[rt], Bc[ct], acc[rt][ct], 0, 0, 0);
    }
    /* register r of MFMA (rt, ct) in lane l: row r0 + 4 (4 (l / 16) + r) + rt,
     * column k0 + CT (l % 16) + ct: a lane writes CT consecutive columns */
    double* out = part + (size_t)blockIdx.x * ldb * NP;
    if (S == 1) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = r0 + 16 * (lane >> 4) + 4 * r + rt;
                double* o = out + (size_t)row * NP + k0 + CT * (lane & 15);
#pragma unroll
                for (int ct = 0; ct < CT; ct += 2)
                    *(double2*)(o + ct) = make_double2((double)acc[rt][ct][r], (double)acc[rt][ct + 1][r]);
            }
        return;
    }
    /* sub-ranges S-1 .. 0 add their totals in FP64 into the LDS tile of
     * their row wave (register layout: lane-consecutive doubles), in that
     * order; sub-range 0 then writes the partial from the tile */
    extern __shared__ double red64[];
    double* const tile = red64 + (size_t)wr * 16 * CT * 64;
    for (int q = S - 1; q >= 0; --q) {
        if (sr == q) {
#pragma unroll
            for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    /* converted here, in this wave's turn (not hoisted out
                     * of the q loop: 16 CT doubles beside the accumulators
                     * would not fit the registers) */
                    __asm__ volatile("" : "+v"(acc[rt][ct]));
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        double& d = tile[((rt * CT + ct) * 4 + r) * 64 + lane];
                        d = q == S - 1 ? (double)acc[rt][ct][r] : d + (double)acc[rt][ct][r];
                    }
                }
        }
        __syncthreads();
    }
    if (sr != 0) return;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + 16 * (lane >> 4) + 4 * r + rt;
            double* o = out + (size_t)row * NP + k0 + CT * (lane & 15);
#pragma unroll
            for (int ct = 0; ct < CT; ct += 2)
                *(double2*)(o + ct) = make_double2(tile[((rt * CT + ct) * 4 + r) * 64 + lane],
                                                   tile[((rt * CT + ct + 1) * 4 + r) * 64 + lane]);
        }
}

/* ---------------- statistic GEMM G = g^T x, BF16 --------------------- */
/* The reduction index (the sample) is the slow axis of g and x in sample-
 * major memory, and an MFMA fragment wants 8 consecutive samples of one
 * column per lane.  So both operands are stored packed 8 samples per column
 * ([t/8][col][8] bf16: the E-step writes g that way, x is packed once per
 * recording) and every A and B fragment is one 16-byte load straight from
 * HBM, no LDS.  Workgroup = ldb/64 waves over one chunk of samples; wave w
 * owns g rows [64w, 64w+64) x 16 CT x columns (grid.y splits NP into
 * 16 CT-column slabs), 4 CT 16x16 FP32 accumulators, and the next 32-sample
 * block's fragments are loaded while the current block's MFMAs run.  Every
 * g and x element is read from HBM once per slab; FP32 accumulation inside
 * the MFMA (32 products per rounding) over the chunk, FP64 partial per
 * chunk.  Where g has fewer than 512 columns (ldb/64 = rw row waves), a
 * workgroup of 8 waves runs S = 8 / rw sub-ranges of its chunk side by
 * side and adds their FP32 accumulators through LDS in a fixed order
 * (S - 1 first, 0 last) before the one FP64 partial is written: as many
 * waves in flight and per-wave accumulation lengths as with S times the
 * chunks, but S times fewer partials to write and reduce. */
/* 32-sample blocks [b0, b1) of one recording into its partial slot `slot`;
 * S = (waves of the workgroup) / rw sub-ranges, red: dynamic LDS of
 * rw 64 x 16 CT floats when S > 1 */
template <int CT>
__device__ __forceinline__ void gx_bf16_packed_body(const uint16_t* __restrict__ g8,
                                                    const uint16_t* __restrict__ x8,
                                                    double* __restrict__ part, const int NP,
                                                    const int ldb, long b0, long b1,
                                                    const int slot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rw = min(ldb - (int)blockIdx.z * 512, 512) / 64;   /* row waves */
    /* sample sub-ranges: only below 512 columns, where the launch gives 8
     * waves for fewer row waves and the LDS tile with them */
    const int S = ldb < 512 ? max(1, (int)(blockDim.x >> 6) / rw) : 1;
    const int sr = w / rw, wr = w - sr * rw;
    const int r0 = blockIdx.z * 512 + wr * 64;     /* grid.z: groups of 512 g columns */
    if (S == 1 && r0 >= ldb) return;                /* no barriers below when S = 1 */
    {
        const long sub = (b1 - b0 + S - 1) / S;     /* this wave's sub-range */
        const long s0 = min(b1, b0 + sr * sub);
        b1 = min(b1, s0 + sub);
        b0 = s0;
    }
    const int k0 = blockIdx.y * 16 * CT;
    f32x4 acc[4][CT];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = (f32x4){0.f, 0.f, 0.f, 0.f};
    i16x8 A[4], B[CT];
    auto load = [&](long blk) {
        const bool ok = blk < b1;
        const size_t grp = (size_t)(ok ? blk : b0) * 4 + (lane >> 4);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
            A[rt] = *(const i16x8*)(g8 + (grp * ldb + r0 + rt * 16 + (lane & 15)) * 8);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
            B[ct] = *(const i16x8*)(x8 + (grp * NP + k0 + ct * 16 + (lane & 15)) * 8);
    };
    if (b0 < b1) load(b0);
    for (long blk = b0; blk < b1; ++blk) {
        i16x8 Ac[4], Bc[CT];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) Ac[rt] = A[rt];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) Bc[ct] = B[ct];
        load(blk + 1);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, Ac[rt]),
                                                                      __builtin_bit_cast(bf16x8, Bc[ct]),
                                                                      acc[rt][ct], 0, 0, 0);
    }
    if (S > 1) {
        /* sub-ranges S-1 .. 1 into the LDS tile of their row wave, in order;
         * sub-range 0 adds the total to its own and writes the partial */
        extern __shared__ float red[];
        constexpr int RS = 16 * CT + 4;             /* padded row of the tile */
        float* const tile = red + (size_t)wr * 64 * RS;
        for (int q = S - 1; q >= 1; --q) {
            if (sr == q) {
#pragma unroll
                for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            float& d = tile[(rt * 16 + (lane >> 4) * 4 + r) * RS + ct * 16 + (lane & 15)];
                            d = q == S - 1 ? acc[rt][ct][r] : d + acc[rt][ct][r];
                        }
            }
            __syncthreads();
        }
        if (sr != 0) return;
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc[rt][ct][r] += tile[(rt * 16 + (lane >> 4) * 4 + r) * RS + ct * 16 + (lane & 15)];
    }
    double* out = part + (size_t)slot * ldb * NP;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = r0 + rt * 16 + (lane >> 4) * 4 + r;
                out[(size_t)row * NP + k0 + ct * 16 + (lane & 15)] = (double)acc[rt][ct][r];
            }
}

template <int CT>
__global__ __launch_bounds__(512) void gx_bf16_packed_kernel(const uint16_t* __restrict__ g8,
                                                             const uint16_t* __restrict__ x8,
                                                             double* __restrict__ part, long nblk,
                                                             int NP, int ldb, long cblk) {
    const long b0 = (long)blockIdx.x * cblk;
    gx_bf16_packed_body<CT>(g8, x8, part, NP, ldb, b0, min(nblk, b0 + cblk), (int)blockIdx.x);
}

/* Many recordings in one launch (same NP and ldb): workgroup (b, slab, row
 * group) runs work item b, the 32-sample blocks [item.t0, item.t1) of
 * recording item.rec into its partial slot item.slot. */
template <int CT>
__global__ __launch_bounds__(512) void gx_bf16_batch_kernel(const GxArgs* __restrict__ args,
                                                            const BatchItem* __restrict__ items,
                                                            int NP, int ldb) {
    const BatchItem it = items[blockIdx.x];
    const GxArgs g = args[it.rec];
    gx_bf16_packed_body<CT>(g.g8, g.x8, g.part, NP, ldb, it.t0, it.t1, it.slot);
}

__global__ void pack8_kernel(const uint16_t* __restrict__ xhi, long T, int NP,
                             uint16_t* __restrict__ x8) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   /* (group, column) */
    const size_t ngrp = (size_t)(T + 7) / 8;
    if (q >= ngrp * NP) return;
    const size_t grp = q / NP;
    const int i = (int)(q % NP);
    i16x8 v = (i16x8){0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const long t = (long)grp * 8 + e;
        if (t < T) v[e] = (short)xhi[(size_t)t * NP + i];
    }
    *(i16x8*)(x8 + q * 8) = v;
}

}  // namespace

int launch_unmix_f64(const double* x, const double* Wg, const double* wc, double* b, long T,
                     int NP, int ldb, int cus, hipStream_t st) {
    if (NP % 32 || ldb % 64) return -1;
    /* register-resident W: 16 CT columns x NP inputs per wave, 2 CT NP / 4 VGPRs
     * (at most 128), workgroups of NT threads whose waves all compute
     * (ldb >= NT / 4 CT); other shapes use the LDS-tiled kernel below */
    const int ctmax = NP == 64 ? 4 : NP == 128 ? 2 : NP == 192 ? 1 : 0;
    int CT = 0, NT = 0;
    for (int nt = 512; nt >= 256 && !CT; nt >>= 1)
        for (int ct = ctmax; ct >= 1; ct >>= 1)
            if (ldb >= nt / 4 * ct && ldb % (16 * ct) == 0) { CT = ct; NT = nt; break; }
    if (CT) {
        const int cols = NT / 4 * CT;                  /* columns per workgroup */
        const int groups = (ldb + cols - 1) / cols;
        const long ntiles = (T + 15) / 16;
#define AMICA_UNMIX_F64(C, K, N)                                                                \
        do {                                                                                    \
            static int bpc = 0;                                                                 \
            if (bpc == 0) {                                                                     \
                int nb = 0;                                                                     \
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                              \
                        &nb, unmix_f64_lds_kernel<C, K, N>, N, 0) != hipSuccess || nb < 1)     \
                    nb = 1;                                                                     \
                bpc = nb;                                                                       \
            }                                                                                   \
            const long gx = std::max(1L, std::min<long>(ntiles, (long)cus * bpc / groups));     \
            hipLaunchKernelGGL((unmix_f64_lds_kernel<C, K, N>), dim3((unsigned)gx, groups),     \
                               dim3(N), 0, st, x, Wg, wc, b, T, ldb, ntiles);                  \
        } while (0)
        if (NP == 64 && CT == 4) AMICA_UNMIX_F64(4, 16, 512);
        else if (NP == 64 && CT == 2) AMICA_UNMIX_F64(2, 16, 512);
        else if (NP == 64 && NT == 512) AMICA_UNMIX_F64(1, 16, 512);
        else if (NP == 64) AMICA_UNMIX_F64(1, 16, 256);
        else if (NP == 128 && CT == 2) AMICA_UNMIX_F64(2, 32, 512);
        else if (NP == 128 && NT == 512) AMICA_UNMIX_F64(1, 32, 512);
        else if (NP == 192 && NT == 512) AMICA_UNMIX_F64(1, 48, 512);
        else CT = 0;
#undef AMICA_UNMIX_F64
        if (CT) return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    dim3 grid((unsigned)((T + 63) / 64), ldb / 64);
    hipLaunchKernelGGL(unmix_f64_kernel, grid, dim3(256), 0, st, x, Wg, wc, b, T, NP, ldb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_unmix_f32(const float* x, const float* Wg, const float* wc, float* b, long T, int NP,
                     int ldb, int cus, hipStream_t st) {
    if (NP % 64 || ldb % 64) return -1;
    const long ntiles = (T + 15) / 16;
    /* register-resident W: CT NP / 4 floats per lane (at most 128) */
    const int CT = (NP == 64 || NP == 128) ? 4 : (NP == 192 || NP == 256) ? 2 : 0;
    if (CT && ldb % (16 * CT) == 0) {
        const int cols = 8 * 16 * CT;
        const int groups = (ldb + cols - 1) / cols;
#define AMICA_UNMIX_F32(C, K)                                                                   \
        do {                                                                                    \
            static int bpc = 0;                                                                 \
            if (bpc == 0) {                                                                     \
                int nb = 0;                                                                     \
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                              \
                        &nb, unmix_f32_lds_kernel<C, K>, 512, 0) != hipSuccess || nb < 1)      \
                    nb = 1;                                                                     \
                bpc = nb;                                                                       \
            }                                                                                   \
            const long gx = std::max(1L, std::min<long>(ntiles, (long)cus * bpc / groups));     \
            hipLaunchKernelGGL((unmix_f32_lds_kernel<C, K>), dim3((unsigned)gx, groups),        \
                               dim3(512), 0, st, x, Wg, wc, b, T, ldb, ntiles);                \
        } while (0)
        if (NP == 64) AMICA_UNMIX_F32(4, 16);
        else if (NP == 128) AMICA_UNMIX_F32(4, 32);
        else if (NP == 192) AMICA_UNMIX_F32(2, 48);
        else AMICA_UNMIX_F32(2, 64);
#undef AMICA_UNMIX_F32
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    hipLaunchKernelGGL(unmix_f32_kernel, dim3((unsigned)ntiles, ldb / 64), dim3(256), 0, st, x, Wg,
                       wc, b, T, NP, ldb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* Column tiles per wave and k-steps of the register-resident kernel, or
 * CT = 0 where none fits (W fragments: 8 CT KS VGPRs). */
static void unmix_shape(int NP, int& CT, int& KS) {
    KS = NP / 32;
    switch (NP) {
        case 64: CT = 4; break;
        case 128: CT = 4; break;
        case 192: CT = 2; break;
        case 256: CT = 2; break;
        default: CT = 0; break;
    }
}

template <bool Q8>
static int launch_unmix_reg(const void* xa, const uint16_t* xlo, const uint16_t* Whi,
                            const uint16_t* Wlo, const float* wc, float* b, long T, int NP, int ldb,
                            int cus, hipStream_t st) {
    int CT, KS;
    unmix_shape(NP, CT, KS);
    if (CT == 0 || ldb % (16 * CT)) return -1;
    /* tiles per step: the waves one column group leaves free take further
     * tiles (1, 2, 4 or 8 tiles of 16 samples per step) */
    const int need = (ldb + 16 * CT - 1) / (16 * CT);          /* waves for all columns */
    const int tpw = need >= 5 ? 1 : need >= 3 ? 2 : need == 2 ? 4 : 8;
    const int groups = (ldb + (8 / tpw) * 16 * CT - 1) / ((8 / tpw) * 16 * CT);
    const long nsteps = ((T + 15) / 16 + tpw - 1) / tpw;
    /* as many workgroups as the CUs hold at once (queried once per kernel) */
#define AMICA_UNMIX_REG(C, K, P)                                                                \
    do {                                                                                        \
        static int bpc = 0;                                                                     \
        if (bpc == 0) {                                                                         \
            int nb = 0;                                                                         \
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                                  \
                    &nb, unmix_lds_kernel<C, K, Q8, P>, 512, 0) != hipSuccess || nb < 1)       \
                nb = 1;                                                                         \
            bpc = nb;                                                                           \
        }                                                                                       \
        const long gx = std::max(1L, std::min<long>(nsteps, (long)cus * bpc / groups));         \
        hipLaunchKernelGGL((unmix_lds_kernel<C, K, Q8, P>), dim3((unsigned)gx, groups),         \
                           dim3(512), 0, st, xa, xlo, Whi, Wlo, wc, b, T, ldb, nsteps);         \
    } while (0)
    if (CT == 4 && KS == 2) {
        if (tpw == 8) AMICA_UNMIX_REG(4, 2, 8);
        else if (tpw == 4) AMICA_UNMIX_REG(4, 2, 4);
        else if (tpw == 2) AMICA_UNMIX_REG(4, 2, 2);
        else AMICA_UNMIX_REG(4, 2, 1);
    } else if (CT == 4 && KS == 4) {
        /* NP = 128: ldb >= 128 needs two waves, so at most 4 tiles per step
         * (8 would not fit the LDS) */
        if (tpw >= 4) AMICA_UNMIX_REG(4, 4, 4);
        else if (tpw == 2) AMICA_UNMIX_REG(4, 4, 2);
        else AMICA_UNMIX_REG(4, 4, 1);
    } else if (CT == 2 && KS == 6) AMICA_UNMIX_REG(2, 6, 1);
    else if (CT == 2 && KS == 8) AMICA_UNMIX_REG(2, 8, 1);
    else return -1;
#undef AMICA_UNMIX_REG
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_unmix_bf16x3(const float* x, const uint16_t* xbf, const uint16_t* Whi,
                        const uint16_t* Wlo, const float* wc, float* b, long T, int NP, int ldb,
                        int cus, hipStream_t st) {
    if (NP % 32 || ldb % 64) return -1;
    if (xbf && launch_unmix_reg<false>(xbf, xbf + (size_t)T * NP, Whi, Wlo, wc, b, T, NP, ldb, cus, st) == 0)
        return 0;
    dim3 grid((unsigned)((T + 63) / 64), ldb / 64);
    hipLaunchKernelGGL(unmix_bf16x3_kernel, grid, dim3(256), 0, st, x, Whi, Wlo, wc, b, T, NP, ldb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_unmix_batch(const UnmixArgs* args, const BatchItem* items, int nitems, int NP, int ldb,
                       hipStream_t st) {
    int CT, KS;
    unmix_shape(NP, CT, KS);
    if (CT == 0 || ldb % (16 * CT) || nitems < 1) return -1;
    const dim3 grid(nitems, (ldb + 128 * CT - 1) / (128 * CT));
    if (CT == 4 && KS == 2)
        hipLaunchKernelGGL((unmix_lds_batch_kernel<4, 2, false>), grid, dim3(512), 0, st, args, items);
    else if (CT == 4 && KS == 4)
        hipLaunchKernelGGL((unmix_lds_batch_kernel<4, 4, false>), grid, dim3(512), 0, st, args, items);
    else if (CT == 2 && KS == 6)
        hipLaunchKernelGGL((unmix_lds_batch_kernel<2, 6, false>), grid, dim3(512), 0, st, args, items);
    else if (CT == 2 && KS == 8)
        hipLaunchKernelGGL((unmix_lds_batch_kernel<2, 8, false>), grid, dim3(512), 0, st, args, items);
    else
        return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_unmix_fp8(const uint8_t* xq, const uint16_t* Whi, const uint16_t* Wlo, const float* wc,
                     float* b, long T, int NP, int ldb, int cus, hipStream_t st) {
    if (NP % 32 || ldb % 64) return -1;
    if (launch_unmix_reg<true>(xq, nullptr, Whi, Wlo, wc, b, T, NP, ldb, cus, st) == 0) return 0;
    dim3 grid((unsigned)((T + 63) / 64), ldb / 64);
    hipLaunchKernelGGL(unmix_fp8_tiled_kernel, grid, dim3(256), 0, st, xq, Whi, Wlo, wc, b, T, NP, ldb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static long gx_chunk(long T, int nchunks, int mult) {
    long c = (T + nchunks - 1) / nchunks;
    return ((c + mult - 1) / mult) * mult;
}

static int gx_f32_ct(int NP) { return NP % 128 == 0 ? 8 : 4; }

/* workgroups of 8 waves that split their chunk's samples (g narrower than
 * 512 columns, the FP64 LDS tile within 64 KB) */
static size_t gx_f32_split_lds(int NP, int ldb) {
    return (size_t)(ldb / 64) * 16 * gx_f32_ct(NP) * 64 * sizeof(double);
}

static bool gx_f32_split(int NP, int ldb) {
    return ldb < 512 && 8 % (ldb / 64) == 0 && gx_f32_split_lds(NP, ldb) <= 65536;
}

int gx_f32_chunks(long T, int NP, int ldb, int cus) {
    /* chunks of >= 512 samples (>= 256 per wave where S waves split a
     * chunk: twice the waves in flight for the same partial count), at
     * most 256 MB of partials, and at least one per CU of the budget where
     * T allows */
    const long per = gx_f32_split(NP, ldb) ? 256L * (8 / (ldb / 64)) : 512L;
    const long by_len = (T + per - 1) / per;
    const long by_mem = std::max(1L, (256L << 20) / ((long)ldb * NP * 8));
    const int slabs = NP / (16 * gx_f32_ct(NP));
    const long want = std::max(1L, (long)cus / slabs);
    return (int)std::max(1L, std::min(by_len, std::max(want, by_mem)));
}

int launch_gx_f32(const float* g, const float* x, double* part, long T, int NP, int ldb,
                  int nchunks, hipStream_t st) {
    if (NP % 64 || ldb % 64 || nchunks < 1) return -1;
    const int CT = gx_f32_ct(NP);
    dim3 grid(nchunks, NP / (16 * CT), (ldb + 511) / 512), block(std::min(ldb, 512));
    const long chunk = gx_chunk(T, nchunks, 4);
    const bool split = gx_f32_split(NP, ldb);
    if (split) block = dim3(512);
    const size_t shm = split ? gx_f32_split_lds(NP, ldb) : 0;
    if (CT == 8)
        hipLaunchKernelGGL(gx_f32_slab_kernel<8>, grid, block, shm, st, g, x, part, T, NP, ldb, chunk);
    else
        hipLaunchKernelGGL(gx_f32_slab_kernel<4>, grid, block, shm, st, g, x, part, T, NP, ldb, chunk);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static int gx_bf16_ct(int NP) { return NP % 128 == 0 ? 8 : 4; }

/* workgroups of 8 waves that split their chunk's samples (g narrower than
 * 512 columns, the LDS tile within 64 KB), or of min(ldb, 512) / 64 waves */
static size_t gx_bf16_split_lds(int NP, int ldb) {
    return (size_t)(ldb / 64) * 64 * (16 * gx_bf16_ct(NP) + 4) * sizeof(float);
}

static bool gx_bf16_split(int NP, int ldb) {
    return ldb < 512 && 8 % (ldb / 64) == 0 && gx_bf16_split_lds(NP, ldb) <= 65536;
}

int gx_bf16_chunks(long T, int NP, int ldb, int cus) {
    /* about 8 waves per CU of the budget where T allows (a workgroup has
     * 8 waves, or min(ldb, 512) / 64 where they do not split the samples),
     * and at most 256 MB of partials */
    const long nblk = (T + 31) / 32;
    const int slabs = NP / (16 * gx_bf16_ct(NP)) * ((ldb + 511) / 512);
    const int wpg = gx_bf16_split(NP, ldb) ? 8 : std::min(ldb, 512) / 64;
    const long by_mem = std::max(1L, (256L << 20) / ((long)ldb * NP * 8));
    const long want = std::max(1L, (long)cus * 8 / ((long)slabs * wpg));
    return (int)std::max(1L, std::min(std::min(nblk, by_mem), want));
}

int launch_gx_bf16(const uint16_t* g8, const uint16_t* x8, double* part, long T, int NP, int ldb,
                   int nchunks, hipStream_t st) {
    if (NP % 64 || ldb % 64 || nchunks < 1) return -1;
    const int CT = gx_bf16_ct(NP);
    const long nblk = (T + 31) / 32;
    const long cblk = (nblk + nchunks - 1) / nchunks;
    /* up to 8 waves of 64 g columns per workgroup, grid.z over the rest;
     * narrower g: 8 waves over sub-ranges of the chunk */
    const bool split = gx_bf16_split(NP, ldb);
    dim3 grid(nchunks, NP / (16 * CT), (ldb + 511) / 512), block(split ? 512 : std::min(ldb, 512));
    const size_t shm = split ? gx_bf16_split_lds(NP, ldb) : 0;
    if (CT == 8)
        hipLaunchKernelGGL(gx_bf16_packed_kernel<8>, grid, block, shm, st, g8, x8, part, nblk, NP, ldb, cblk);
    else
        hipLaunchKernelGGL(gx_bf16_packed_kernel<4>, grid, block, shm, st, g8, x8, part, nblk, NP, ldb, cblk);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_gx_bf16_batch(const GxArgs* args, const BatchItem* items, int nitems, int NP, int ldb,
                         hipStream_t st) {
    if (NP % 64 || ldb % 64 || nitems < 1) return -1;
    const int CT = gx_bf16_ct(NP);
    /* the sub-range split of the single-recording launch, so that an item
     * covering a recording's own chunk sums exactly as that launch does */
    const bool split = gx_bf16_split(NP, ldb);
    const dim3 grid(nitems, NP / (16 * CT), (ldb + 511) / 512), block(split ? 512 : std::min(ldb, 512));
    const size_t shm = split ? gx_bf16_split_lds(NP, ldb) : 0;
    if (CT == 8)
        hipLaunchKernelGGL(gx_bf16_batch_kernel<8>, grid, block, shm, st, args, items, NP, ldb);
    else
        hipLaunchKernelGGL(gx_bf16_batch_kernel<4>, grid, block, shm, st, args, items, NP, ldb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_pack8(const uint16_t* xhi, long T, int NP, uint16_t* x8, hipStream_t st) {
    const size_t cnt = (size_t)((T + 7) / 8) * NP;
    hipLaunchKernelGGL(pack8_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, xhi, T, NP, x8);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace amica
