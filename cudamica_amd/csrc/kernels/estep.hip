/*
 * estep.hip -- fused AMICA E-step for gfx950.
 *
 * One kernel computes, for every sample t, model h, channel i and mixture
 * component j, everything the reference spreads over ten kernels and two
 * cuBLAS calls per block (fraimondo/cudamica src/amica.cu:1109-1606):
 *   y    = sqrt(beta) (b - mu)                         betabyxminusmu   kernels.cu:251-272
 *   Q    = log a + .5 log beta - log2 - lgamma(1+1/rho) - |y|^rho
 *                                                      logalogblogpfun  kernels.cu:282-307
 *   q    = logsumexp_j Q, z = exp(Q - q), L_h += sum_i q     updateltz  kernels.cu:317-366
 *   P    = logsumexp_h L, v = exp(L - P), LL += P             updatell  kernels.cu:376-450
 *   u    = v z, f = rho sign(y)|y|^(rho-1), g = sum_j sqrt(beta) u f
 *                                   computeu / computefp / computeg  kernels.cu:473-574
 *   usum, kappa, lambda, mu, beta, rho numerators         getusum ... getdrhonumer
 *   vsum, sigma2 = sum v b^2, cnew = sum v x               amica.cu:1280-1323, 1584
 * and writes g for the statistic GEMM.
 *
 * Layout: b and g are [T][ldb] with column h*NP + i; x is [T][NP]; NP is n
 * rounded up to 64.  Lane = channel, wave = (model h, 64-channel block cb);
 * a workgroup of M*CW waves owns a contiguous range of samples, so the
 * cross-channel sum (wave butterfly + LDS) and the cross-model logsumexp
 * stay on chip.  Per-element transcendentals are cut to three:
 *   la  = log2|y|
 *   ym2 = 2^((rho-2) la) = |y|^(rho-2),  sy = ym2 y = sign(y)|y|^(rho-1),
 *   yr  = sy y = |y|^rho
 *   e_j = 2^(Q2_j - max Q2)  with Q2 = Q log2(e)
 * and every statistic becomes a product of these with u, with the constant
 * per-(h,j,i) factors (rho, sqrt(beta), ln 2) applied once in the reduction:
 *   usum   = sum u                       mu_numer  = rho  sum u sy
 *   kappa  = beta rho^2 sum u sy^2        mu_denom  = sqrt(beta) rho sum u ym2   (rho <= 2)
 *   beta_d = rho sum u yr                            sqrt(beta) rho^2 sum u sy^2 (rho > 2)
 *   lambda = rho^2 sum u yr^2 - 2 rho sum u yr + usum
 *   rho_n  = rho ln2 sum u yr la
 * Statistics are summed in FP32 over 32 samples, then in FP64 per lane (in
 * LDS), and written as one FP64 partial per workgroup (no atomics: the
 * reduction kernel sums partials in a fixed order).
 *
 * B64 = true (FP32 path): b is FP64 and y = sqrt(beta)(b - mu) is formed in
 * FP64 before the FP32 density math, so |y| near 0 -- where mu_denom's
 * weight |y|^(rho-2) is singular for rho < 2 -- keeps full relative
 * precision.
 */
#include "common.h"
#include "engine_internal.h"

namespace amica {

namespace {

constexpr int kS = 4;          /* samples per barrier tile */
constexpr int kFlush = 8;      /* tiles summed in FP32 before the FP64 flush */

__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

template <int MJ, bool B64>
__global__ __launch_bounds__(512) void estep_kernel(EstepArgs a) {
    constexpr int NSJ = kStatsPerJ;
    constexpr int NS = NSJ * MJ + kStatsPerI;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int h = w / a.CW;
    const int cb = w - h * a.CW;
    const int i = cb * 64 + lane;
    const bool act = i < a.n;
    const int nw = a.M * a.CW;

    __shared__ float Lp[2][kMaxEstepWaves][kS];
    __shared__ float dc2[AMICA_MAX_M];
    /* per-lane FP64 totals, [NS][blockDim.x] */
    extern __shared__ double tot[];

    /* L'_h = (log gm + ldet + ldetS) - c0 with c0 = max_h of that constant,
     * so that only differences enter FP32 (in log2 units). */
    if (threadIdx.x < a.M) {
        double c0 = a.lconst[0];
        for (int hh = 1; hh < a.M; ++hh) c0 = fmax(c0, a.lconst[hh]);
        dc2[threadIdx.x] = (float)((a.lconst[threadIdx.x] - c0) * 1.4426950408889634);
    }

    float sb[MJ], nsbmu[MJ], rm2[MJ], cq2[MJ], rsb[MJ];
    double sb64[B64 ? MJ : 1], nsbmu64[B64 ? MJ : 1];
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
        const float* cj = a.cst + ((size_t)h * kNumCst * MJ + j) * a.NP + i;
        const size_t ks = (size_t)MJ * a.NP;
        if (act) {
            sb[j] = cj[kCstSb * ks];
            nsbmu[j] = cj[kCstNsbmu * ks];
            rm2[j] = cj[kCstRm2 * ks];
            cq2[j] = cj[kCstCq2 * ks];
            rsb[j] = cj[kCstRsb * ks];
        } else {
            sb[j] = 0.f; nsbmu[j] = 1.f; rm2[j] = 0.f; cq2[j] = 0.f; rsb[j] = 0.f;
        }
        if constexpr (B64) {
            const double* dj = a.cst64 + ((size_t)h * kNumCst64 * MJ + j) * a.NP + i;
            sb64[j] = act ? dj[kCst64Sb * ks] : 0.0;
            nsbmu64[j] = act ? dj[kCst64Nsbmu * ks] : 1.0;
        }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) tot[s * blockDim.x + threadIdx.x] = 0.0;
    __syncthreads();

    float acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = 0.f;
    double vsum_acc = 0.0, ll_acc = 0.0;

    const long t0 = (long)blockIdx.x * a.chunk;
    const long t1 = min(a.T, t0 + a.chunk);
    const size_t bcol = (size_t)h * a.NP + i;
    int buf = 0, tiles = 0;

    for (long tb = t0; tb < t1; tb += kS) {
        float E[kS][MJ], SY[kS][MJ], YM2[kS][MJ], YR[kS][MJ], LA[kS][MJ];
        float RS[kS], BV[kS];
        /* ---- phase 1: densities, component posteriors, channel sums ---- */
#pragma unroll
        for (int s = 0; s < kS; ++s) {
            const long t = tb + s;
            const bool valid = t < t1;
            float bv;
            double bv64 = 0.0;
            if constexpr (B64) {
                bv64 = (act && valid) ? a.b64[(size_t)t * a.ldb + bcol] : 0.0;
                bv = (float)bv64;
            } else {
                bv = (act && valid) ? a.b[(size_t)t * a.ldb + bcol] : 0.f;
            }
            BV[s] = bv;
            float q2[MJ];
            float mx = 0.f;
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                float y;
                if constexpr (B64) y = (float)fma(sb64[j], bv64, nsbmu64[j]);
                else y = fmaf(sb[j], bv, nsbmu[j]);
                const float la = fast_log2(fmaxf(fabsf(y), 1.17549435e-38f));
                const float ym2 = fast_exp2(rm2[j] * la);
                const float sy = ym2 * y;
                const float yr = sy * y;
                q2[j] = fmaf(yr, -AMICA_L2E, cq2[j]);
                mx = (j == 0) ? q2[j] : fmaxf(mx, q2[j]);
                SY[s][j] = sy; YM2[s][j] = ym2; YR[s][j] = yr; LA[s][j] = la;
            }
            float se = 0.f;
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                E[s][j] = fast_exp2(q2[j] - mx);
                se += E[s][j];
            }
            RS[s] = fast_rcp(se);
            const float qtot = mx + fast_log2(se);
            float qs = act ? qtot : 0.f;
            qs = wave_sum_f32(qs);
            if (lane == 0) Lp[buf][w][s] = qs;
        }
        __syncthreads();
        /* ---- model posteriors: lane s < kS handles sample tb+s ---- */
        float vl = 0.f;
        if (lane < kS) {
            const long t = tb + lane;
            const bool valid = t < t1;
            float Lmine = 0.f, Lmax = -3.0e38f;
            for (int hh = 0; hh < a.M; ++hh) {
                float l = dc2[hh];
                for (int c = 0; c < a.CW; ++c) l += Lp[buf][hh * a.CW + c][lane];
                Lmax = fmaxf(Lmax, l);
                if (hh == h) Lmine = l;
            }
            float sum = 0.f;
            for (int hh = 0; hh < a.M; ++hh) {
                float l = dc2[hh];
                for (int c = 0; c < a.CW; ++c) l += Lp[buf][hh * a.CW + c][lane];
                sum += fast_exp2(l - Lmax);
            }
            const float P2 = Lmax + fast_log2(sum);
            vl = valid ? fast_exp2(Lmine - P2) : 0.f;
            if (valid) {
                if (cb == 0) vsum_acc += (double)vl;
                if (h == 0 && cb == 0) ll_acc += (double)P2 * AMICA_LN2;
                if (a.ltall && cb == 0) a.ltall[(size_t)h * a.T + t] = Lmine * (float)AMICA_LN2;
            }
        }
        /* ---- phase 2: responsibilities and statistics ---- */
#pragma unroll
        for (int s = 0; s < kS; ++s) {
            const long t = tb + s;
            const float v = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(vl), s));
            const float vrs = v * RS[s];
            float g = 0.f;
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                const float u = E[s][j] * vrs;
                const float uf = u * SY[s][j];
                g = fmaf(uf, rsb[j], g);
                float* aj = acc + j * NSJ;
                aj[0] += u;
                aj[1] += uf;
                aj[2] = fmaf(uf, SY[s][j], aj[2]);
                aj[3] = fmaf(u, YM2[s][j], aj[3]);
                const float uyr = u * YR[s][j];
                aj[4] += uyr;
                aj[5] = fmaf(uyr, YR[s][j], aj[5]);
                aj[6] = fmaf(uyr, LA[s][j], aj[6]);
            }
            if (t < t1) {
                const float vb = v * BV[s];
                acc[NSJ * MJ + 0] = fmaf(vb, BV[s], acc[NSJ * MJ + 0]);
                const float xv = act ? a.x[(size_t)t * a.NP + i] : 0.f;
                acc[NSJ * MJ + 1] = fmaf(v, xv, acc[NSJ * MJ + 1]);
                if (!act) g = 0.f;
                acc[NSJ * MJ + 2] += g;
                if (a.g_bf16) {
                    ((uint16_t*)a.g)[(size_t)t * a.ldb + bcol] = f32_to_bf16_bits(g);
                } else {
                    ((float*)a.g)[(size_t)t * a.ldb + bcol] = g;
                }
            }
        }
        buf ^= 1;
        if (++tiles == kFlush) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                tot[s * blockDim.x + threadIdx.x] += (double)acc[s];
                acc[s] = 0.f;
            }
            tiles = 0;
        }
    }
    double* base = a.part + (size_t)blockIdx.x * a.pstride;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double v = tot[s * blockDim.x + threadIdx.x] + (double)acc[s];
        base[((size_t)h * NS + s) * a.NP + i] = act ? v : 0.0;
    }
    /* per-model vsum and the log-likelihood sum from lanes 0..kS-1 */
    double vs = wave_sum_f64(lane < kS ? vsum_acc : 0.0);
    double ls = wave_sum_f64(lane < kS ? ll_acc : 0.0);
    const size_t tail = (size_t)a.M * NS * a.NP;
    if (lane == 0 && cb == 0) base[tail + h] = vs;
    if (lane == 0 && cb == 0 && h == 0) base[tail + a.M] = ls;
    (void)nw;
}

}  // namespace

int estep_configure(int MJ, size_t lds_bytes) {
    hipError_t r = hipSuccess;
#define AMICA_ESTEP_ATTR(J)                                                                       \
    case J:                                                                                       \
        r = hipFuncSetAttribute((const void*)estep_kernel<J, true>,                              \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);      \
        if (r == hipSuccess)                                                                      \
            r = hipFuncSetAttribute((const void*)estep_kernel<J, false>,                         \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);  \
        break;
    switch (MJ) {
        AMICA_ESTEP_ATTR(1)
        AMICA_ESTEP_ATTR(2)
        AMICA_ESTEP_ATTR(3)
        AMICA_ESTEP_ATTR(4)
        AMICA_ESTEP_ATTR(5)
        AMICA_ESTEP_ATTR(6)
        AMICA_ESTEP_ATTR(7)
        AMICA_ESTEP_ATTR(8)
        default: return -1;
    }
#undef AMICA_ESTEP_ATTR
    return r == hipSuccess ? 0 : -1;
}

int launch_estep(const EstepArgs& a, int MJ, int nblocks, hipStream_t st) {
    dim3 grid(nblocks), block(64 * a.M * a.CW);
    const size_t shm = (size_t)(kStatsPerJ * MJ + kStatsPerI) * block.x * sizeof(double);
    const bool b64 = a.b64 != nullptr;
#define AMICA_ESTEP_CASE(J)                                                                   \
    case J:                                                                                   \
        if (b64) hipLaunchKernelGGL((estep_kernel<J, true>), grid, block, shm, st, a);        \
        else hipLaunchKernelGGL((estep_kernel<J, false>), grid, block, shm, st, a);           \
        break;
    switch (MJ) {
        AMICA_ESTEP_CASE(1)
        AMICA_ESTEP_CASE(2)
        AMICA_ESTEP_CASE(3)
        AMICA_ESTEP_CASE(4)
        AMICA_ESTEP_CASE(5)
        AMICA_ESTEP_CASE(6)
        AMICA_ESTEP_CASE(7)
        AMICA_ESTEP_CASE(8)
        default: return -1;
    }
#undef AMICA_ESTEP_CASE
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace amica
