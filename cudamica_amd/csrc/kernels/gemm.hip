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

/* ---------------- statistic GEMM G = g^T x, FP32 --------------------- */
/* The reduction index is the sample, the row index of both operands, so
 * the 16x16x4 fragments are coalesced rows read straight from global
 * memory.  Workgroup tile 64 (g columns) x 64 (x columns) over a chunk of
 * samples; wave w owns x columns [16w, 16w+16) and all four 16-row tiles.
 * FP32 accumulation over 64 samples, then FP64. */
__global__ __launch_bounds__(256) void gx_f32_kernel(const float* __restrict__ g,
                                                     const float* __restrict__ x,
                                                     double* __restrict__ part, long T, int NP,
                                                     int ldb, long chunk) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = blockIdx.x * 64;
    const int k0 = blockIdx.y * 64 + w * 16;
    const long ts = (long)blockIdx.z * chunk;
    const long te = min(T, ts + chunk);
    f32x4 acc[4];
    f64x4 tot[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        acc[a] = (f32x4){0.f, 0.f, 0.f, 0.f};
        tot[a] = (f64x4){0.0, 0.0, 0.0, 0.0};
    }
    int steps = 0;
    for (long t = ts; t < te; t += 4) {
        const long tr = t + (lane >> 4);
        const bool ok = tr < te;
        const float bv = ok ? x[(size_t)tr * NP + k0 + (lane & 15)] : 0.f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const float av = ok ? g[(size_t)tr * ldb + r0 + a * 16 + (lane & 15)] : 0.f;
            acc[a] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[a], 0, 0, 0);
        }
        if (++steps == 16) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
#pragma unroll
                for (int r = 0; r < 4; ++r) tot[a][r] += (double)acc[a][r];
                acc[a] = (f32x4){0.f, 0.f, 0.f, 0.f};
            }
            steps = 0;
        }
    }
    double* out = part + (size_t)blockIdx.z * ldb * NP;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + a * 16 + (lane >> 4) * 4 + r;
            out[(size_t)row * NP + k0 + (lane & 15)] = tot[a][r] + (double)acc[a][r];
        }
}

/* ---------------- statistic GEMM G = g^T x, BF16 --------------------- */
/* The reduction index (sample) is the row index of g and x, so fragments of
 * 8 consecutive samples are transposed through LDS: 32 samples x 64 columns
 * of g and of x per stage, stored [column][sample] so every lane reads its
 * 8 bf16 with one 16-byte LDS load. */
__global__ __launch_bounds__(256) void gx_bf16_kernel(const uint16_t* __restrict__ g,
                                                      const uint16_t* __restrict__ x,
                                                      double* __restrict__ part, long T, int NP,
                                                      int ldb, long chunk) {
    __shared__ __attribute__((aligned(16))) uint16_t gs[64][40];
    __shared__ __attribute__((aligned(16))) uint16_t xs[64][40];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int r0 = blockIdx.x * 64;
    const int k0 = blockIdx.y * 64;
    const long ts = (long)blockIdx.z * chunk;
    const long te = min(T, ts + chunk);
    f32x4 acc[4];
    f64x4 tot[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        acc[a] = (f32x4){0.f, 0.f, 0.f, 0.f};
        tot[a] = (f64x4){0.0, 0.0, 0.0, 0.0};
    }
    int steps = 0;
    /* load mapping: thread -> (sample row lr = tid/8, 8 columns at lc = (tid%8)*8) */
    const int lr = tid >> 3, lc = (tid & 7) * 8;
    for (long t = ts; t < te; t += 32) {
        const long tr = t + lr;
        i16x8 gv = (i16x8){0, 0, 0, 0, 0, 0, 0, 0}, xv = gv;
        if (tr < te) {
            gv = *(const i16x8*)(g + (size_t)tr * ldb + r0 + lc);
            xv = *(const i16x8*)(x + (size_t)tr * NP + k0 + lc);
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            gs[lc + e][lr] = (uint16_t)gv[e];
            xs[lc + e][lr] = (uint16_t)xv[e];
        }
        __syncthreads();
        /* wave w: x columns [16w, 16w+16), all four 16-row tiles of g */
        const i16x8 bv = *(const i16x8*)&xs[w * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const i16x8 av = *(const i16x8*)&gs[a * 16 + (lane & 15)][(lane >> 4) * 8];
            acc[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av),
                                                             __builtin_bit_cast(bf16x8, bv),
                                                             acc[a], 0, 0, 0);
        }
        __syncthreads();
        if (++steps == 4) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {
#pragma unroll
                for (int r = 0; r < 4; ++r) tot[a][r] += (double)acc[a][r];
                acc[a] = (f32x4){0.f, 0.f, 0.f, 0.f};
            }
            steps = 0;
        }
    }
    double* out = part + (size_t)blockIdx.z * ldb * NP;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = r0 + a * 16 + (lane >> 4) * 4 + r;
            out[(size_t)row * NP + k0 + w * 16 + (lane & 15)] = tot[a][r] + (double)acc[a][r];
        }
}

}  // namespace

int launch_unmix_f64(const double* x, const double* Wg, const double* wc, double* b, long T,
                     int NP, int ldb, hipStream_t st) {
    if (NP % 32 || ldb % 64) return -1;
    dim3 grid((unsigned)((T + 63) / 64), ldb / 64);
    hipLaunchKernelGGL(unmix_f64_kernel, grid, dim3(256), 0, st, x, Wg, wc, b, T, NP, ldb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_unmix_bf16x3(const float* x, const uint16_t* Whi, const uint16_t* Wlo, const float* wc,
                        float* b, long T, int NP, int ldb, hipStream_t st) {
    if (NP % 32 || ldb % 64) return -1;
    dim3 grid((unsigned)((T + 63) / 64), ldb / 64);
    hipLaunchKernelGGL(unmix_bf16x3_kernel, grid, dim3(256), 0, st, x, Whi, Wlo, wc, b, T, NP, ldb);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static long gx_chunk(long T, int nchunks, int mult) {
    long c = (T + nchunks - 1) / nchunks;
    return ((c + mult - 1) / mult) * mult;
}

int launch_gx_f32(const float* g, const float* x, double* part, long T, int NP, int ldb,
                  int nchunks, hipStream_t st) {
    if (NP % 64 || ldb % 64) return -1;
    dim3 grid(ldb / 64, NP / 64, nchunks);
    hipLaunchKernelGGL(gx_f32_kernel, grid, dim3(256), 0, st, g, x, part, T, NP, ldb,
                       gx_chunk(T, nchunks, 4));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_gx_bf16(const uint16_t* g, const uint16_t* x, double* part, long T, int NP, int ldb,
                   int nchunks, hipStream_t st) {
    if (NP % 64 || ldb % 64) return -1;
    dim3 grid(ldb / 64, NP / 64, nchunks);
    hipLaunchKernelGGL(gx_bf16_kernel, grid, dim3(256), 0, st, g, x, part, T, NP, ldb,
                       gx_chunk(T, nchunks, 32));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace amica
