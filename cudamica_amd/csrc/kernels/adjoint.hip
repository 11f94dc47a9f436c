/*
 * adjoint.hip -- the per-sample AMICA log-likelihood and its adjoint
 * (include/amica_adjoint.h).
 *
 * The forward terms are the reference's E-step quantities
 * (fraimondo/cudamica src/kernels.cu:251-450: y, the generalized-Gaussian
 * log density Q, the log-sum-exp over components and over models).  The
 * adjoint weights each sample's model posterior v_ht by the upstream
 * gradient dll_t and re-evaluates the responsibilities u = dll v z (nothing
 * is stored between the passes); the gradients are then the reference's
 * sufficient statistics of those weighted responsibilities:
 *   dalpha = sum u / alpha                  dmu  = sum sqrt(beta) u f
 *   dbeta  = sum u (1 - f y) / (2 beta)     drho = sum u (psi(1+1/rho)/rho^2 - |y|^rho ln|y|)
 *   dgm    = sum dll v / gm                 g    = sum_j sqrt(beta) u f   (per sample)
 *   dW_h   = (sum_t dll v_h) W_h^-T - sum_t g_t (x_t - c_h)^T
 *   dc_h   = W_h^T sum_t g_t
 * with f = rho sign(y)|y|^(rho-1).
 *
 * Kernels: prep (per model: W^-1 by LU, log|det W|, W c, the constant part
 * of Q), the unmixing GEMM b = W x - W c on MFMA (the engine's FP64 or plain
 * FP32 kernels, into a [T][M NP] buffer), ll (one wave per (model, 64-channel
 * block) over a range of samples, all M ceil(n/64) waves in one workgroup
 * where they fit in 8, else three passes through per-sample buffers; b read
 * from the GEMM's buffer -- the allocation-free forward instead forms it per
 * lane from x tiles in LDS), gx (sum_t g_t x_t^T: FP32 MFMA on the FP32
 * variant, 64 x 64 FP64 tiles on the FP64 one, over sample chunks), and
 * finish (fixed-order sums over workgroups and the assembly of the
 * gradients).
 * Every reduction is in a fixed order, so results repeat bit for bit.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../../include/amica_adjoint.h"
#include "../../../include/amica_engine.h"
#include "common.h"
#include "engine_internal.h"
#include "linalg_dev.h"

namespace amica {

namespace {

struct AdjArgs {
    const void* x;
    long T;
    int n, NP, M, m;
    const double *W, *c, *alpha, *mu, *beta, *rho, *gm;
    double ldetS;
    const double* dll;   /* null: forward only */
    double* ll;          /* [T] or null */
    /* workspace */
    double* Winv;        /* [M][n][n] column-major */
    double* ldet;        /* [M] log|det W_h| */
    double* cq;          /* [M][m][n] log alpha + .5 log beta - log2 - lgamma(1+1/rho) */
    double* wc;          /* [M][NP] W_h c_h */
    double* lu;          /* [M][n][n] LU scratch */
    double* X;           /* [M][n][n] inverse scratch */
    void* g;             /* [T][M NP] (adjoint) */
    double* part;        /* [nblocks][pstride] */
    long chunk;          /* samples per workgroup */
    size_t pstride;
    /* split passes (M ceil(n/64) > 8 waves): workgroups of pg (model,
     * channel-block) pairs, blockIdx.y = pair group; qbuf [T][M CW] the
     * pairs' channel sums of q, vbuf [T][M] dll_t v_ht */
    double* qbuf;
    double* vbuf;
    int pg;
    /* b = W x - W c from the unmixing GEMM, [T][M NP] Real, or null: each
     * lane then forms its b from an LDS tile of x (the allocation-free
     * forward) */
    const void* bbuf;
};

enum { kLlFused = 0, kLlQ = 1, kLlGrad = 2 };

/* per pair (h, i): dalpha_j, dmu_j, dbeta_j, drho_j (j < m), gsum; then Omega_h */
__host__ __device__ inline int pair_stride(int m) { return 4 * m + 1; }

/* Per model: W^-1 and log|det W| -- here by LU in this workgroup (the
 * allocation-free forward, a.lu set), else already computed by the
 * Gauss-Jordan launches, which leave -log|det W| in ldet -- then W c and
 * the constant part of Q. */
__global__ __launch_bounds__(256) void adj_prep_kernel(AdjArgs a) {
    __shared__ int perm[1024];
    const int h = blockIdx.x, n = a.n, m = a.m;
    const long nn = (long)n * n;
    const double* W = a.W + (size_t)h * nn;
    if (a.lu) {
        const double lad = block_lu_inverse(W, n, a.lu + (size_t)h * nn, a.X + (size_t)h * nn,
                                            a.Winv + (size_t)h * nn, perm);
        if (threadIdx.x == 0) a.ldet[h] = lad;
    } else if (threadIdx.x == 0) {
        a.ldet[h] = -a.ldet[h];
    }
    const double* c = a.c + (size_t)h * n;
    for (int i = threadIdx.x; i < a.NP; i += blockDim.x) {
        double s = 0.0;
        if (i < n)
            for (int k = 0; k < n; ++k) s += W[i + (long)k * n] * c[k];
        a.wc[(size_t)h * a.NP + i] = s;
    }
    const long nm = (long)n * m;
    for (long q = threadIdx.x; q < nm; q += blockDim.x) {
        const size_t e = (size_t)h * nm + q;
        a.cq[e] = log(a.alpha[e]) + 0.5 * log(a.beta[e]) - AMICA_LOG2_REF - lgamma(1.0 + 1.0 / a.rho[e]);
    }
}

/* FP32: the hardware base-2 logarithm and exponential (v_log_f32,
 * v_exp_f32; arguments here are normal or underflow to 0) and reciprocal. */
template <typename Real> __device__ __forceinline__ Real r_log(Real v);
template <> __device__ __forceinline__ float r_log<float>(float v) {
    return __builtin_amdgcn_logf(v) * 0.6931471805599453f;
}
template <> __device__ __forceinline__ double r_log<double>(double v) { return log(v); }
template <typename Real> __device__ __forceinline__ Real r_exp(Real v);
template <> __device__ __forceinline__ float r_exp<float>(float v) {
    return __builtin_amdgcn_exp2f(v * 1.4426950408889634f);
}
template <> __device__ __forceinline__ double r_exp<double>(double v) { return exp(v); }
template <typename Real> __device__ __forceinline__ Real r_rcp(Real v);
template <> __device__ __forceinline__ float r_rcp<float>(float v) { return __builtin_amdgcn_rcpf(v); }
template <> __device__ __forceinline__ double r_rcp<double>(double v) { return 1.0 / v; }

template <typename Real> struct RealTiny;
template <> struct RealTiny<float> { static constexpr float v = 1.17549435e-38f; };
template <> struct RealTiny<double> { static constexpr double v = 2.2250738585072014e-308; };

/* Samples per tile held in LDS and registers. */
template <typename Real> struct Tile { static constexpr int S = sizeof(Real) == 8 ? 4 : 8; };

/* Component j of channel i of model h at sample value b: y, Q, and (for the
 * adjoint) f y, f and |y|^rho ln|y|. */
template <typename Real>
__device__ __forceinline__ void density(Real b, Real sb, Real mu, Real rho, Real cq, Real& y,
                                        Real& Q, Real& yr, Real& lny) {
    y = sb * (b - mu);
    const Real ay = fmax(fabs(y), RealTiny<Real>::v);
    lny = r_log<Real>(ay);
    yr = r_exp<Real>(rho * lny);     /* |y|^rho */
    Q = cq - yr;
}

/* MODE kLlFused: a workgroup holds all M ceil(n/64) (model, channel-block)
 * pairs of a sample range and exchanges the channel and model sums in LDS.
 * Where they do not fit one workgroup the same body runs in three passes:
 * kLlQ writes each pair's channel sum of q per sample (qbuf),
 * ll_post_kernel forms L_h, ll_t and dll_t v_ht (vbuf), and kLlGrad
 * re-evaluates the densities and accumulates the gradients. */
template <typename Real, int MJ, bool ADJ, int MODE>
__global__ __launch_bounds__(512) void ll_kernel(AdjArgs a) {
    constexpr int S = Tile<Real>::S;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n = a.n, NP = a.NP, M = a.M;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
    const int CW = NP / 64;
    int h, pair = 0;
    bool wreal = true;   /* wave-uniform: the wave owns a real pair */
    if constexpr (MODE == kLlFused) {
        h = wv / CW;
        pair = wv;
    } else {
        pair = blockIdx.y * a.pg + wv;
        wreal = pair < M * CW;
        if (!wreal) pair = M * CW - 1;
        h = pair / CW;
    }
    const int i = (pair - h * CW) * 64 + lane;
    const bool act = wreal && i < n;
    /* LDS: xs[S][NP] (Real), qp[S][nw], Lh[S][M], vh[S][M] (double) */
    Real* xs = (Real*)smem_raw;
    double* qp = (double*)(smem_raw + (size_t)S * NP * sizeof(Real));
    double* Lh = qp + S * nw;
    double* vh = Lh + S * M;

    Real sb[MJ], mu[MJ], rho[MJ], cq[MJ];
    double psi1[ADJ ? MJ : 1], al[ADJ ? MJ : 1], be[ADJ ? MJ : 1], sbd[ADJ ? MJ : 1], rhod[ADJ ? MJ : 1];
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
        const size_t e = ((size_t)h * MJ + j) * n + (act ? i : 0);
        sb[j] = (Real)sqrt(a.beta[e]);
        mu[j] = (Real)a.mu[e];
        rho[j] = (Real)a.rho[e];
        cq[j] = (Real)a.cq[e];
        if constexpr (ADJ) {
            psi1[j] = psi_dev(1.0 + 1.0 / a.rho[e]);
            al[j] = a.alpha[e];
            be[j] = a.beta[e];
            sbd[j] = sqrt(a.beta[e]);
            rhod[j] = a.rho[e];
        }
    }
    const Real wci = (Real)a.wc[(size_t)h * NP + i];
    const double* Wh = a.W + (size_t)h * n * n;
    /* per component: sum u, sum u f, sum u f y, sum u |y|^rho ln|y|; the
     * gradients are these times per-(h, j, i) constants, applied at the end */
    double su[ADJ ? MJ : 1], suf[ADJ ? MJ : 1], sufy[ADJ ? MJ : 1], surl[ADJ ? MJ : 1];
    double gsum = 0.0, om = 0.0;
#pragma unroll
    for (int j = 0; j < (ADJ ? MJ : 1); ++j) su[j] = suf[j] = sufy[j] = surl[j] = 0.0;

    const long t0 = (long)blockIdx.x * a.chunk;
    const long t1 = min(a.T, t0 + a.chunk);
    const Real* x = (const Real*)a.x;
    const Real* const bb = (const Real*)a.bbuf;
    const size_t bcol = (size_t)h * NP + i, ldb = (size_t)M * NP;
    for (long tb = t0; tb < t1; tb += S) {
        Real bv[S];
        if (bb) {                                    /* block-uniform */
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const long t = tb + s;
                bv[s] = (act && t < t1) ? bb[(size_t)t * ldb + bcol] : (Real)0;
            }
        } else {
            for (int q = tid; q < S * NP; q += blockDim.x) {
                const int s = q / NP, k = q - (q / NP) * NP;
                const long t = tb + s;
                xs[q] = (t < t1 && k < n) ? x[(size_t)t * n + k] : (Real)0;
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < S; ++s) bv[s] = (Real)0;
            if (act)
                for (int k = 0; k < n; ++k) {
                    const Real wik = (Real)Wh[i + (long)k * n];
#pragma unroll
                    for (int s = 0; s < S; ++s) bv[s] = fma(wik, xs[s * NP + k], bv[s]);
                }
#pragma unroll
            for (int s = 0; s < S; ++s) bv[s] -= wci;
        }
        if constexpr (MODE != kLlGrad) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            double q = 0.0;
            if (act) {
                Real Qj[MJ], mx = (Real)0;
#pragma unroll
                for (int j = 0; j < MJ; ++j) {
                    Real y, yr, lny;
                    density<Real>(bv[s], sb[j], mu[j], rho[j], cq[j], y, Qj[j], yr, lny);
                    mx = (j == 0) ? Qj[j] : fmax(mx, Qj[j]);
                }
                Real se = (Real)0;
#pragma unroll
                for (int j = 0; j < MJ; ++j) se += r_exp<Real>(Qj[j] - mx);
                q = (double)mx + (double)r_log<Real>(se);
            }
            q = wave_sum_f64(q);
            if constexpr (MODE == kLlQ) {
                const long t = tb + s;
                if (lane == 0 && wreal && t < t1) a.qbuf[(size_t)t * M * CW + pair] = q;
            } else if constexpr (MODE == kLlFused) {
                if (lane == 0) qp[s * nw + wv] = q;
            }
        }
        }
        if constexpr (MODE == kLlQ) {
            __syncthreads();            /* xs is rewritten by the next tile */
            continue;
        }
        if constexpr (MODE == kLlFused) {
        __syncthreads();
        if (tid < S * M) {
            const int s = tid / M, hh = tid - (tid / M) * M;
            double L = log(a.gm[hh]) + a.ldet[hh] + a.ldetS;
            for (int c = 0; c < CW; ++c) L += qp[s * nw + hh * CW + c];
            Lh[s * M + hh] = L;
        }
        __syncthreads();
        if (tid < S) {
            const int s = tid;
            const long t = tb + s;
            double mx = Lh[s * M];
            for (int hh = 1; hh < M; ++hh) mx = fmax(mx, Lh[s * M + hh]);
            double se = 0.0;
            for (int hh = 0; hh < M; ++hh) se += exp(Lh[s * M + hh] - mx);
            const double P = mx + log(se);
            const bool valid = t < t1;
            if (valid && a.ll) a.ll[t] = P;
            const double w = (ADJ && valid) ? a.dll[t] : 0.0;
            for (int hh = 0; hh < M; ++hh) vh[s * M + hh] = w * exp(Lh[s * M + hh] - P);
        }
        __syncthreads();
        }
        if constexpr (ADJ) {
            Real* g = (Real*)a.g;
            const size_t ldg = (size_t)M * NP;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const long t = tb + s;
                if (t >= t1) break;
                double omh;
                if constexpr (MODE == kLlGrad) omh = a.vbuf[(size_t)t * M + h];
                else omh = vh[s * M + h];
                double gv = 0.0;
                if (act) {
                    Real Qj[MJ], Yj[MJ], YRj[MJ], LNj[MJ], mx = (Real)0;
#pragma unroll
                    for (int j = 0; j < MJ; ++j) {
                        density<Real>(bv[s], sb[j], mu[j], rho[j], cq[j], Yj[j], Qj[j], YRj[j], LNj[j]);
                        mx = (j == 0) ? Qj[j] : fmax(mx, Qj[j]);
                    }
                    Real ez[MJ], se = (Real)0;
#pragma unroll
                    for (int j = 0; j < MJ; ++j) {
                        ez[j] = r_exp<Real>(Qj[j] - mx);
                        se += ez[j];
                    }
                    const Real rse = r_rcp<Real>(se);
#pragma unroll
                    for (int j = 0; j < MJ; ++j) {
                        const double u = omh * (double)(ez[j] * rse);
                        /* f y = rho |y|^rho and f = f y / y (0 at y = 0) */
                        const Real fy = rho[j] * YRj[j];
                        const Real f = (Yj[j] == (Real)0) ? (Real)0 : fy * r_rcp<Real>(Yj[j]);
                        const double uf = u * (double)f;
                        su[j] += u;
                        suf[j] += uf;
                        sufy[j] += u * (double)fy;
                        surl[j] += u * ((double)YRj[j] * (double)LNj[j]);
                        gv += sbd[j] * uf;
                    }
                }
                if (wreal) g[(size_t)t * ldg + (size_t)h * NP + i] = (Real)gv;
                gsum += gv;
                if (i == 0) om += omh;
            }
        }
        __syncthreads();
    }
    if constexpr (ADJ && MODE != kLlQ) {
        if (!wreal) return;
        double* base = a.part + (size_t)blockIdx.x * a.pstride;
        double* pp = base + ((size_t)h * NP + i) * pair_stride(MJ);
        /* dalpha = sum u / alpha, dmu = sqrt(beta) sum u f,
         * dbeta = sum u (1 - f y) / (2 beta),
         * drho = sum u (psi(1 + 1/rho) / rho^2 - |y|^rho ln|y|) */
#pragma unroll
        for (int j = 0; j < MJ; ++j) {
            pp[j] = su[j] / al[j];
            pp[MJ + j] = sbd[j] * suf[j];
            pp[2 * MJ + j] = (su[j] - sufy[j]) / (2.0 * be[j]);
            pp[3 * MJ + j] = su[j] * psi1[j] / (rhod[j] * rhod[j]) - surl[j];
        }
        pp[4 * MJ] = gsum;
        if (i == 0) base[(size_t)M * NP * pair_stride(MJ) + h] = om;
    }
}

/* The split passes' middle step, one thread per sample: L_h = log gm_h +
 * log|det W_h| + ldetS + the pairs' sums of q, ll_t = logsumexp_h L_h and
 * (adjoint) dll_t v_ht. */
__global__ __launch_bounds__(256) void ll_post_kernel(AdjArgs a) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.T) return;
    const int M = a.M, CW = a.NP / 64;
    const double* q = a.qbuf + (size_t)t * M * CW;
    double mx = -HUGE_VAL;
    for (int hh = 0; hh < M; ++hh) {
        double L = log(a.gm[hh]) + a.ldet[hh] + a.ldetS;
        for (int c = 0; c < CW; ++c) L += q[hh * CW + c];
        mx = fmax(mx, L);
    }
    double se = 0.0;
    for (int hh = 0; hh < M; ++hh) {
        double L = log(a.gm[hh]) + a.ldet[hh] + a.ldetS;
        for (int c = 0; c < CW; ++c) L += q[hh * CW + c];
        se += exp(L - mx);
    }
    const double P = mx + log(se);
    if (a.ll) a.ll[t] = P;
    if (a.dll) {
        const double w = a.dll[t];
        for (int hh = 0; hh < M; ++hh) {
            double L = log(a.gm[hh]) + a.ldet[hh] + a.ldetS;
            for (int c = 0; c < CW; ++c) L += q[hh * CW + c];
            a.vbuf[(size_t)t * M + hh] = w * exp(L - P);
        }
    }
}

/* part[chunk][r][k] = sum_{t in chunk} g[t][r] x[t][k]   (r < M NP, k < n) */
template <typename Real>
__global__ __launch_bounds__(256) void adj_gx_kernel(const Real* __restrict__ g,
                                                     const Real* __restrict__ x, long T, int n,
                                                     int ldg, long chunk, double* __restrict__ part) {
    __shared__ double ga[32][64];
    __shared__ double xb[32][64];
    const int tid = threadIdx.x;
    const int rb = blockIdx.y * 64, kb = blockIdx.z * 64;
    const long ts = (long)blockIdx.x * chunk, te = min(T, ts + chunk);
    const int ti = (tid & 15) * 4, tk = (tid >> 4) * 4;
    double acc[4][4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = 0.0;
    for (long t0 = ts; t0 < te; t0 += 32) {
        for (int q = tid; q < 32 * 64; q += 256) {
            const int s = q >> 6, col = q & 63;
            const long t = t0 + s;
            double va = 0.0, vb = 0.0;
            if (t < te) {
                if (rb + col < ldg) va = (double)g[(size_t)t * ldg + rb + col];
                if (kb + col < n) vb = (double)x[(size_t)t * n + kb + col];
            }
            ga[s][col] = va;
            xb[s][col] = vb;
        }
        __syncthreads();
        for (int s = 0; s < 32; ++s) {
            double av[4], bv[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) av[p] = ga[s][ti + p];
#pragma unroll
            for (int q = 0; q < 4; ++q) bv[q] = xb[s][tk + q];
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] = fma(av[p], bv[q], acc[p][q]);
        }
        __syncthreads();
    }
    double* out = part + (size_t)blockIdx.x * ldg * n;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = rb + ti + p, k = kb + tk + q;
            if (r < ldg && k < n) out[(size_t)r * n + k] = acc[p][q];
        }
}

/* The unmixing GEMM's operands: Wg[h NP + i][k] = W_h[i][k] (zero past n)
 * and W c in the GEMM's precision. */
template <typename Real>
__global__ __launch_bounds__(256) void adj_wg_kernel(AdjArgs a, Real* __restrict__ Wg,
                                                     Real* __restrict__ wcr) {
    const int NP = a.NP, n = a.n;
    const long total = (long)a.M * NP * NP;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        const long r = q / NP;
        const int k = (int)(q - r * NP), h = (int)(r / NP), i = (int)(r - (long)h * NP);
        Wg[q] = (i < n && k < n) ? (Real)a.W[(size_t)h * n * n + (size_t)k * n + i] : (Real)0;
        if (k == 0) wcr[r] = (Real)a.wc[r];
    }
}

/* x [T][n] into [T][NP] with zero columns past n (the GEMMs' row layout) */
template <typename Real>
__global__ __launch_bounds__(256) void adj_pad_kernel(const Real* __restrict__ x, Real* __restrict__ xp,
                                                      long T, int n, int NP) {
    const long total = T * NP;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        const long t = q / NP;
        const int k = (int)(q - t * NP);
        xp[q] = k < n ? x[(size_t)t * n + k] : (Real)0;
    }
}

/* Fixed-order sums over slots: out[e] = sum_{b < nslots} in[b stride + e]
 * for e < count, one thread per entry (consecutive entries coalesce). */
__global__ __launch_bounds__(256) void adj_slot_sum_kernel(const double* __restrict__ in, size_t stride,
                                                           int nslots, size_t count,
                                                           double* __restrict__ out) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (size_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int b = 0; b < nslots; ++b) acc += in[(size_t)b * stride + e];
        out[e] = acc;
    }
}

struct AdjOut {
    double *dW, *dc, *dalpha, *dmu, *dbeta, *drho, *dgm;
};

/* One workgroup per model: fixed-order sums over the ll workgroups and the
 * gx chunks, then the gradients. */
__global__ __launch_bounds__(256) void adj_finish_kernel(AdjArgs a, int nblocks,
                                                         const double* __restrict__ gxpart,
                                                         int nchunks, int gxs, AdjOut o) {
    extern __shared__ double gs[];   /* gsum_h[n] */
    __shared__ double om_sh;
    const int h = blockIdx.x, n = a.n, m = a.m, NP = a.NP, M = a.M;
    const int ps = pair_stride(m);
    const long nm = (long)n * m;
    for (long q = threadIdx.x; q < nm + n; q += blockDim.x) {
        /* q < nm: (j, i) of the four per-component gradients; then gsum_i */
        const int j = q < nm ? (int)(q / n) : 0;
        const int i = q < nm ? (int)(q % n) : (int)(q - nm);
        const size_t off = ((size_t)h * NP + i) * ps;
        if (q < nm) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            for (int b = 0; b < nblocks; ++b) {
                const double* pp = a.part + (size_t)b * a.pstride + off;
                s0 += pp[j];
                s1 += pp[m + j];
                s2 += pp[2 * m + j];
                s3 += pp[3 * m + j];
            }
            const size_t e = (size_t)h * nm + q;
            o.dalpha[e] = s0;
            o.dmu[e] = s1;
            o.dbeta[e] = s2;
            o.drho[e] = s3;
        } else {
            double s = 0.0;
            for (int b = 0; b < nblocks; ++b) s += a.part[(size_t)b * a.pstride + off + 4 * m];
            gs[i] = s;
        }
    }
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nblocks; ++b) s += a.part[(size_t)b * a.pstride + (size_t)M * NP * ps + h];
        om_sh = s;
        o.dgm[h] = s / a.gm[h];
    }
    __syncthreads();
    const double om = om_sh;
    const long nn = (long)n * n;
    const double* W = a.W + (size_t)h * nn;
    const double* Wi = a.Winv + (size_t)h * nn;
    const double* c = a.c + (size_t)h * n;
    const size_t ldg = (size_t)M * NP;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += W[i + (long)k * n] * gs[i];
        o.dc[(size_t)h * n + k] = s;
    }
    for (long q = threadIdx.x; q < nn; q += blockDim.x) {
        const int i = (int)(q % n), k = (int)(q / n);
        double G = 0.0;
        for (int ch = 0; ch < nchunks; ++ch) G += gxpart[((size_t)ch * ldg + (size_t)h * NP + i) * gxs + k];
        o.dW[(size_t)h * nn + q] = om * Wi[k + (long)i * n] - (G - gs[i] * c[k]);
    }
}

template <typename Real, bool ADJ, int MODE>
int launch_ll_mode(int MJ, const AdjArgs& a, dim3 grid, dim3 block, size_t shm, hipStream_t st) {
#define AMICA_LL_CASE(J)                                                                           \
    case J:                                                                                        \
        if (hipFuncSetAttribute((const void*)ll_kernel<Real, J, ADJ, MODE>,                        \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) != hipSuccess) \
            return -1;                                                                             \
        hipLaunchKernelGGL((ll_kernel<Real, J, ADJ, MODE>), grid, block, shm, st, a);             \
        break;
    switch (MJ) {
        AMICA_LL_CASE(1)
        AMICA_LL_CASE(2)
        AMICA_LL_CASE(3)
        AMICA_LL_CASE(4)
        AMICA_LL_CASE(5)
        AMICA_LL_CASE(6)
        AMICA_LL_CASE(7)
        AMICA_LL_CASE(8)
        default: return -1;
    }
#undef AMICA_LL_CASE
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename Real, bool ADJ>
int launch_ll(int MJ, const AdjArgs& a, int nblocks, size_t shm, hipStream_t st) {
    const int pairs = a.M * (a.NP / 64);
    if (!a.qbuf)
        return launch_ll_mode<Real, ADJ, kLlFused>(MJ, a, dim3(nblocks), dim3(64 * pairs), shm, st);
    const dim3 grid(nblocks, (pairs + a.pg - 1) / a.pg), block(64 * a.pg);
    const size_t shx = (size_t)Tile<Real>::S * a.NP * sizeof(Real);
    if (launch_ll_mode<Real, false, kLlQ>(MJ, a, grid, block, shx, st)) return -1;
    hipLaunchKernelGGL(ll_post_kernel, dim3((unsigned)((a.T + 255) / 256)), dim3(256), 0, st, a);
    if (hipGetLastError() != hipSuccess) return -1;
    return ADJ ? launch_ll_mode<Real, true, kLlGrad>(MJ, a, grid, block, shx, st) : 0;
}

/* Workspace of one call, freed on every exit path. */
struct Workspace {
    static constexpr int kMax = 20;
    void* ptrs[kMax] = {nullptr};
    int n = 0;
    ~Workspace() {
        for (int k = 0; k < n; ++k) (void)hipFree(ptrs[k]);
    }
    bool alloc(void** p, size_t bytes) {
        if (n == kMax) return false;
        if (hipMalloc(p, std::max<size_t>(bytes, 8)) != hipSuccess) return false;
        ptrs[n++] = *p;
        return true;
    }
};

/* The fixed device workspace of amica_ll_forward_noalloc (callers that
 * cannot allocate: code under hipGraph capture, tight serving loops).  It
 * holds, in doubles: W^-1, the LU and inverse scratch (M n n each), the
 * per-(h, j, i) constants (M m n), W c (M NP) and log|det W| (M).  Limits:
 * M n n <= kNoAllocNN (e.g. 256 channels x 1 model, 128 x 4), plus the
 * forward's own M ceil(n/64) 64 <= 512, m <= 8.  1.6 MB per device, held
 * until the process ends; every call on a device shares it. */
constexpr size_t kNoAllocNN = 65536;
constexpr size_t kNoAllocNM = 4096;        /* 8 components x 512 padded channels */
constexpr size_t kNoAllocNP = 512;
constexpr size_t kNoAllocM = 8;
constexpr size_t kNoAllocDoubles = 3 * kNoAllocNN + kNoAllocNM + kNoAllocNP + kNoAllocM;

__device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {nullptr};

/* The device address of `scratch` on the current device, looked up once per
 * device (each GPU holds its own copy of the array at its own address; the
 * array's name in host code is only a handle for the lookup). */
int noalloc_scratch(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return report_error(AMICA_E_DEVICE, "no HIP device");
    if (dev < 0 || dev >= kMaxDevices) return report_error(AMICA_E_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        double* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return report_error(AMICA_E_DEVICE, (std::string("hipGetSymbolAddress(scratch): ") +
                                                 hipGetErrorString(e)).c_str());
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

int ll_call(int real_bits, const void* x, long T, int n, int M, int m, const double* W,
            const double* c, const double* alpha, const double* mu, const double* beta,
            const double* rho, const double* gm, double ldetS, const double* dll, double* ll,
            const AdjOut* out, void* stream, bool noalloc = false) {
    const bool adj = out != nullptr;
    if (real_bits != 32 && real_bits != 64) return report_error(AMICA_E_INVALID, "real_bits must be 32 or 64");
    if (!x || !W || !c || !alpha || !mu || !beta || !rho || !gm || (adj && !dll))
        return report_error(AMICA_E_INVALID, "null argument");
    if (adj && (!out->dW || !out->dc || !out->dalpha || !out->dmu || !out->dbeta || !out->drho || !out->dgm))
        return report_error(AMICA_E_INVALID, "null gradient output");
    if (!adj && !ll) return report_error(AMICA_E_INVALID, "null ll output");
    if (T < 1 || n < 1 || M < 1 || m < 1 || m > 8)
        return report_error(AMICA_E_INVALID, "need T >= 1, n >= 1, M >= 1, 1 <= m <= 8");
    if (n > 1024 || M > 16) return report_error(AMICA_E_INVALID, "need n <= 1024 and M <= 16");
    const int NP = (n + 63) / 64 * 64;
    /* all (model, channel-block) pairs in one workgroup, else the split passes */
    const bool split = (long)M * NP > 512;
    if (split && noalloc)
        return report_error(AMICA_E_INVALID, "M * ceil(n/64) * 64 must be <= 512 without allocation");
    hipStream_t st = (hipStream_t)stream;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return report_error(AMICA_E_DEVICE, "no HIP device");

    const size_t rb = real_bits == 64 ? 8 : 4;
    const int S = real_bits == 64 ? 4 : 8;
    const int nw = M * NP / 64;
    const long nblocks = std::max(1L, std::min((long)cus * 2, (T + S - 1) / S));
    long chunk = (T + nblocks - 1) / nblocks;
    chunk = (chunk + S - 1) / S * S;
    const int nb = (int)((T + chunk - 1) / chunk);

    AdjArgs a = {};
    a.x = x; a.T = T; a.n = n; a.NP = NP; a.M = M; a.m = m;
    a.W = W; a.c = c; a.alpha = alpha; a.mu = mu; a.beta = beta; a.rho = rho; a.gm = gm;
    a.ldetS = ldetS; a.dll = adj ? dll : nullptr; a.ll = ll; a.chunk = chunk;
    a.pstride = ((size_t)M * NP * pair_stride(m) + M + 31) & ~(size_t)31;
    Workspace ws;
    const size_t nn = (size_t)n * n;
    if (noalloc) {
        /* forward only, workspace carved from the static device array */
        if ((size_t)M * nn > kNoAllocNN)
            return report_error(AMICA_E_INVALID, "M * n * n exceeds the allocation-free workspace (65536)");
        double* sp = nullptr;
        const int rc = noalloc_scratch(&sp);
        if (rc) return rc;
        a.Winv = sp;
        a.lu = sp + kNoAllocNN;
        a.X = sp + 2 * kNoAllocNN;
        a.cq = sp + 3 * kNoAllocNN;
        a.wc = a.cq + kNoAllocNM;
        a.ldet = a.wc + kNoAllocNP;
    } else if (!ws.alloc((void**)&a.Winv, M * nn * 8) || !ws.alloc((void**)&a.ldet, M * 8) ||
               !ws.alloc((void**)&a.cq, (size_t)M * m * n * 8) || !ws.alloc((void**)&a.wc, (size_t)M * NP * 8)) {
        return report_error(AMICA_E_NOMEM, "workspace allocation failed");
    }
    if (!noalloc) {
        /* W^-1 and log|det W| by the M-step's blocked Gauss-Jordan */
        double* gjw = nullptr;
        if (!ws.alloc((void**)&gjw, (8 + M * nn + inverse_work_doubles(n, M)) * 8))
            return report_error(AMICA_E_NOMEM, "workspace allocation failed");
        if (launch_gj_inverse(W, a.Winv, a.ldet, n, M, gjw, st))
            return report_error(AMICA_E_DEVICE, "inverse launch failed");
    }
    if (split) {
        a.pg = 8;
        if (!ws.alloc((void**)&a.qbuf, (size_t)T * (M * NP / 64) * 8) ||
            (adj && !ws.alloc((void**)&a.vbuf, (size_t)T * M * 8)))
            return report_error(AMICA_E_NOMEM, "workspace allocation failed");
    }
    const int ldb = M * NP;
    /* the statistic GEMM sum_t g_t x_t^T: FP32 MFMA over padded rows of x
     * (gx_f32_slab_kernel, row stride NP) on the FP32 variant, FP64 tiles
     * over the unpadded rows (row stride n) on the FP64 one */
    const bool gx_mfma = adj && real_bits == 32;
    const int gchunks = gx_mfma ? gx_f32_chunks(T, NP, ldb, cus)
                                : (int)std::max(1L, std::min((long)cus, (T + 255) / 256));
    const int gxs = gx_mfma ? NP : n;
    double* gxpart = nullptr;
    if (adj) {
        if (!ws.alloc(&a.g, (size_t)T * ldb * rb) ||
            !ws.alloc((void**)&a.part, (size_t)nb * a.pstride * 8) ||
            !ws.alloc((void**)&gxpart, (size_t)gchunks * ldb * gxs * 8))
            return report_error(AMICA_E_NOMEM, "workspace allocation failed");
    }
    hipLaunchKernelGGL(adj_prep_kernel, dim3(M), dim3(256), 0, st, a);
    if (hipGetLastError() != hipSuccess) return report_error(AMICA_E_DEVICE, "prep launch failed");
    /* b = W x - W c by the unmixing GEMM on MFMA (FP64 or FP32 operands)
     * into a [T][M NP] buffer; the allocation-free forward forms b per lane */
    const void* xp = x;
    if (!noalloc) {
        void *Wg = nullptr, *wcr = nullptr, *bb = nullptr;
        if (!ws.alloc(&Wg, (size_t)ldb * NP * rb) || !ws.alloc(&wcr, (size_t)ldb * rb) ||
            !ws.alloc(&bb, (size_t)T * ldb * rb))
            return report_error(AMICA_E_NOMEM, "workspace allocation failed");
        const unsigned eg = (unsigned)std::min<long>(4096, ((long)ldb * NP + 255) / 256);
        if (n != NP) {
            void* xpad = nullptr;
            if (!ws.alloc(&xpad, (size_t)T * NP * rb))
                return report_error(AMICA_E_NOMEM, "workspace allocation failed");
            const unsigned pg = (unsigned)std::min<long>(65536, (T * NP + 255) / 256);
            if (real_bits == 64)
                hipLaunchKernelGGL(adj_pad_kernel<double>, dim3(pg), dim3(256), 0, st, (const double*)x,
                                   (double*)xpad, T, n, NP);
            else
                hipLaunchKernelGGL(adj_pad_kernel<float>, dim3(pg), dim3(256), 0, st, (const float*)x,
                                   (float*)xpad, T, n, NP);
            xp = xpad;
        }
        int urc;
        if (real_bits == 64) {
            hipLaunchKernelGGL(adj_wg_kernel<double>, dim3(eg), dim3(256), 0, st, a, (double*)Wg, (double*)wcr);
            urc = launch_unmix_f64((const double*)xp, (const double*)Wg, (const double*)wcr, (double*)bb, T,
                                   NP, ldb, cus, st);
        } else {
            hipLaunchKernelGGL(adj_wg_kernel<float>, dim3(eg), dim3(256), 0, st, a, (float*)Wg, (float*)wcr);
            urc = launch_unmix_f32((const float*)xp, (const float*)Wg, (const float*)wcr, (float*)bb, T, NP,
                                   ldb, cus, st);
        }
        if (urc || hipGetLastError() != hipSuccess)
            return report_error(AMICA_E_DEVICE, "unmixing GEMM launch failed");
        a.bbuf = bb;
    }
    const size_t shm = (size_t)S * NP * rb + (size_t)S * nw * 8 + 2 * (size_t)S * M * 8;
    int rc;
    if (real_bits == 64)
        rc = adj ? launch_ll<double, true>(m, a, nb, shm, st) : launch_ll<double, false>(m, a, nb, shm, st);
    else
        rc = adj ? launch_ll<float, true>(m, a, nb, shm, st) : launch_ll<float, false>(m, a, nb, shm, st);
    if (rc) return report_error(AMICA_E_DEVICE, "likelihood kernel launch failed");
    if (adj) {
        int gc = gchunks;
        if (gx_mfma) {
            if (launch_gx_f32((const float*)a.g, (const float*)xp, gxpart, T, NP, ldb, gchunks, st))
                return report_error(AMICA_E_DEVICE, "statistic GEMM launch failed");
        } else {
            long gch = (T + gchunks - 1) / gchunks;
            gch = (gch + 31) / 32 * 32;
            gc = (int)((T + gch - 1) / gch);
            const dim3 grid(gc, (ldb + 63) / 64, (n + 63) / 64);
            hipLaunchKernelGGL(adj_gx_kernel<double>, grid, dim3(256), 0, st, (const double*)a.g,
                               (const double*)x, T, n, ldb, gch, gxpart);
        }
        /* the partial slots summed across the GPU, then one workgroup per
         * model assembles the gradients from the totals */
        double *ptot = nullptr, *gtot = nullptr;
        const size_t pcount = (size_t)ldb * pair_stride(m) + M, gcount = (size_t)ldb * gxs;
        if (!ws.alloc((void**)&ptot, pcount * 8) || !ws.alloc((void**)&gtot, gcount * 8))
            return report_error(AMICA_E_NOMEM, "workspace allocation failed");
        hipLaunchKernelGGL(adj_slot_sum_kernel, dim3((unsigned)std::min<size_t>(4096, (pcount + 255) / 256)),
                           dim3(256), 0, st, a.part, a.pstride, nb, pcount, ptot);
        hipLaunchKernelGGL(adj_slot_sum_kernel, dim3((unsigned)std::min<size_t>(4096, (gcount + 255) / 256)),
                           dim3(256), 0, st, gxpart, gcount, gc, gcount, gtot);
        AdjArgs at = a;
        at.part = ptot;
        at.pstride = pcount;
        hipLaunchKernelGGL(adj_finish_kernel, dim3(M), dim3(256), (size_t)n * 8, st, at, 1, gtot, 1, gxs,
                           *out);
        if (hipGetLastError() != hipSuccess) return report_error(AMICA_E_DEVICE, "adjoint launch failed");
    }
    if (noalloc) return 0;     /* asynchronous: the caller synchronises `stream` */
    const hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return report_error(AMICA_E_DEVICE, hipGetErrorString(e));
    return 0;
}

}  // namespace

}  // namespace amica

extern "C" int amica_ll_forward(int real_bits, const void* x, long T, int n, int M, int m,
                                const double* W, const double* c, const double* alpha,
                                const double* mu, const double* beta, const double* rho,
                                const double* gm, double ldetS, double* ll, void* stream) {
    return amica::ll_call(real_bits, x, T, n, M, m, W, c, alpha, mu, beta, rho, gm, ldetS, nullptr,
                          ll, nullptr, stream);
}

extern "C" int amica_ll_forward_noalloc(int real_bits, const void* x, long T, int n, int M, int m,
                                        const double* W, const double* c, const double* alpha,
                                        const double* mu, const double* beta, const double* rho,
                                        const double* gm, double ldetS, double* ll, void* stream) {
    return amica::ll_call(real_bits, x, T, n, M, m, W, c, alpha, mu, beta, rho, gm, ldetS, nullptr,
                          ll, nullptr, stream, true);
}

extern "C" int amica_ll_adjoint(int real_bits, const void* x, long T, int n, int M, int m,
                                const double* W, const double* c, const double* alpha,
                                const double* mu, const double* beta, const double* rho,
                                const double* gm, double ldetS, const double* dll, double* ll,
                                double* dW, double* dc, double* dalpha, double* dmu, double* dbeta,
                                double* drho, double* dgm, void* stream) {
    const amica::AdjOut out = {dW, dc, dalpha, dmu, dbeta, drho, dgm};
    return amica::ll_call(real_bits, x, T, n, M, m, W, c, alpha, mu, beta, rho, gm, ldetS, dll, ll,
                          &out, stream);
}
