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
 * and writes g for the statistic GEMM (float32 [T][ldb], or bf16 packed 8
 * samples per column, [T/8][ldb][8], the layout the BF16 statistic GEMM
 * reads its MFMA fragments in).
 *
 * Layout: b and g are [T][ldb] with column h*NP + i; x is [T][NP]; NP is n
 * rounded up to 64.  Lane = channel, wave = (model h, 64-channel block cb);
 * a workgroup of M*CW waves owns a contiguous range of samples, so the
 * cross-channel sum (DPP row sums + LDS) and the cross-model logsumexp
 * stay on chip.  Per-element transcendentals are cut to three:
 *   la  = log2|y|
 *   ym2 = 2^((rho-2) la) = |y|^(rho-2),  sy = ym2 y = sign(y)|y|^(rho-1),
 *   yr  = sy y = |y|^rho
 *   e_j = 2^Q2_j  with Q2 = Q log2(e) (shifted by max_j Q2 only when
 *         every component underflows)
 * and every statistic becomes a product of these with u, with the constant
 * per-(h,j,i) factors (rho, sqrt(beta), ln 2) applied once in the reduction:
 *   usum   = sum u                       mu_numer  = rho  sum u sy
 *   kappa  = beta rho^2 sum u sy^2        mu_denom  = sqrt(beta) rho sum u ym2   (rho <= 2)
 *   beta_d = rho sum u yr                            sqrt(beta) rho^2 sum u sy^2 (rho > 2)
 *   lambda = rho^2 sum u yr^2 - 2 rho sum u yr + usum
 *   rho_n  = rho ln2 sum u yr la
 * Statistics are summed in FP32 registers over 32 samples, those partials in
 * FP32 per lane in LDS over 1024 samples, and those in FP64 in the
 * workgroup's own slot of the partial buffer, which the reduction kernel
 * sums over workgroups in a fixed order (no atomics, run-to-run identical).
 *
 * Latency: the b and x loads of tile k+1 are issued before tile k's math
 * (raw buffer loads, scalar row offsets), the per-sample channel sums use
 * DPP row shifts instead of LDS permutes, two barriers per tile of 2
 * samples (1 for m >= 4) publish the per-row channel sums and then the
 * model posteriors, which one wave evaluates for the workgroup, and the
 * per-lane LDS footprint (FP32 level-2 totals) leaves room for two
 * 512-thread workgroups per CU.
 *
 * B64 = true (FP32 path): b is FP64 and y = sqrt(beta)(b - mu) is formed in
 * FP64 before the FP32 density math, so |y| near 0 -- where mu_denom's
 * weight |y|^(rho-2) is singular for rho < 2 -- keeps full relative
 * precision.
 */
#include <algorithm>

#include "common.h"
#include "engine_internal.h"

namespace amica {

namespace {

constexpr int kBufferRsrcWord3 = 0x00020000;   /* raw 32-bit buffer, gfx9-family format */
constexpr int kFlush2 = 32;    /* register flushes summed in FP32 LDS */

__device__ __forceinline__ float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

template <int CTRL, int ROWS, int BANKS, bool BOUND>
__device__ __forceinline__ float dpp(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 CTRL, ROWS, BANKS, BOUND));
}

/* max of v and the value a row rotation by R (1, 2, 4, 8) brings in; every
 * lane of every row has a source, so the DPP move keeps no old value */
template <int R>
__device__ __forceinline__ float row_ror_max(float v) {
    const float o = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v),
                                                                       0x120 + R, 0xf, 0xf, false));
    return fmaxf(v, o);
}

/* Sum over each row of 16 lanes; the row total lands in the row's lane 15
 * (the other lanes hold partial sums).  Row shifts by 1, 2, 4 and 8 with
 * zeros shifted in, each one v_add_f32 with a DPP source. */
__device__ __forceinline__ float row_sum_lane15(float v) {
    v += dpp<0x111, 0xf, 0xf, true>(v);
    v += dpp<0x112, 0xf, 0xf, true>(v);
    v += dpp<0x114, 0xf, 0xf, true>(v);
    v += dpp<0x118, 0xf, 0xf, true>(v);
    return v;
}

/* Floor of |y| before log2|y|.  FP32 path (y formed in FP64 from FP64 b):
 * the smallest normal float.  Split-BF16 and FP8 paths (y = sb b - sb mu in
 * FP32 from FP32 b): 2^-24, the FP32 spacing at |sb mu| ~ 1.  There b can
 * round exactly onto mu, and the y that cancels to 0 would give
 * |y|^(rho-2) = 2^63 for rho = 1.5 and swamp the mu denominator
 * sum(u |y|^(rho-2)), which is singular at y = 0 for rho < 2; at the FP32
 * spacing the weight stays of the size FP64 gives its smallest |y|.  (A
 * floor at the ~2^-17 error of the BF16 GEMM would cap the weights of all
 * |y| below it and bias mu_denom low.) */
template <bool B64>
__device__ __forceinline__ constexpr float y_floor() { return B64 ? 1.17549435e-38f : 0x1p-24f; }

/* two floats -> two round-to-nearest-even bf16 in one dword (a low): one
 * v_cvt_pk_bf16_f32 */
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f2){a, b}, bf2));
}

/* The FP64 level of a lane's NS statistic sums: l3[s] = l3[s] + add(s)
 * (add(s) the first time), or 0 where keep is false.  The loads go out B
 * at a time before their adds, so a flush costs ceil(NS / B) memory round
 * trips instead of one per sum, for 2 B transient registers. */
template <int NS, int B, typename F>
__device__ __forceinline__ void l3_accumulate(double* l3, bool started, bool keep, F add) {
#pragma unroll
    for (int s0 = 0; s0 < NS; s0 += B) {
        double cur[B];
        if (started) {                  /* wave-uniform */
#pragma unroll
            for (int k = 0; k < B; ++k) cur[k] = s0 + k < NS ? l3[s0 + k] : 0.0;
        } else {
#pragma unroll
            for (int k = 0; k < B; ++k) cur[k] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (s0 + k < NS) l3[s0 + k] = keep ? cur[k] + add(s0 + k) : 0.0;
    }
}

/* MODE selects the launch the kernel serves:
 *   kEstepFused  all M models x CW channel blocks in one workgroup (M CW <= 8):
 *                the cross-channel and cross-model sums stay on chip;
 *   kEstepLsum   split launch, first pass: a workgroup holds a group of models
 *                x channel blocks (blockIdx.y), and writes each model's sum of
 *                q over its channel blocks per sample (lq);
 *   kEstepStats  split launch, last pass: the densities again, the model
 *                posteriors read from vq (written between the passes by
 *                estep_post_kernel), then the statistics and g. */
template <int MJ, bool B64, int MODE>
__global__ __launch_bounds__(512) void estep_kernel(EstepArgs a) {
    /* samples per barrier tile: 2, or 1 where m >= 4 so that the per-sample
     * values stay within 128 VGPRs (two workgroups per CU) */
    constexpr int kS = MJ >= 4 ? 1 : 2;
    constexpr int kFlush = 32 / kS;     /* tiles summed in FP32 registers (32 samples) */
    /* FP64 flushes with batched loads, except in the split launch's
     * statistics pass, which they would push past 128 VGPRs on the FP32
     * path */
    constexpr bool kL3Batched = MODE != kEstepStats;
    constexpr int NSJ = kStatsPerJ;
    constexpr int NS = NSJ * MJ + kStatsPerI;
    constexpr bool kSplit = MODE != kEstepFused;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   /* wave-uniform */
    int h, cb, hg = 0, cg = 0;
    bool wreal = true;                  /* wave-uniform: the wave owns real columns */
    if constexpr (kSplit) {
        hg = blockIdx.y / a.ncg;
        cg = blockIdx.y - hg * a.ncg;
        const int hl = w / a.cwg;
        h = hg * a.mg + hl;
        cb = cg * a.cwg + (w - hl * a.cwg);
        wreal = h < a.M && cb < a.CW;
        /* a wave past the last model or channel block reads a real column and
         * runs as padding channels (q = 0, no statistics, nothing written) */
        h = min(h, a.M - 1);
        cb = min(cb, a.CW - 1);
    } else {
        h = w / a.CW;
        cb = w - h * a.CW;
    }
    const int i = cb * 64 + lane;
    const bool act = wreal && i < a.n;

    /* per-row channel sums of one tile, double buffered:
     * [buf][wave][sample][row of 16 lanes] (a sample's 4 rows in 16 bytes),
     * and the tile's model posteriors [buf][model h * kS + sample] */
    __shared__ __align__(16) float Lp[2][kMaxEstepWaves][kS][4];
    __shared__ __align__(8) float Vp[2][AMICA_MAX_M * 2];
    __shared__ float dc2[AMICA_MAX_M];
    /* per-lane FP32 level-2 totals, [blockDim.x][NSP] (odd stride: the
     * lanes of a wave hit distinct banks) */
    constexpr int NSP = NS | 1;
    extern __shared__ float tot2_all[];
    float* const tot2 = tot2_all + threadIdx.x * NSP;

    /* L'_h = (log gm + ldet + ldetS) - c0 with c0 = max_h of that constant,
     * so that only differences enter FP32 (in log2 units). */
    if ((int)threadIdx.x < a.M) {
        double c0 = a.lconst[0];
        for (int hh = 1; hh < a.M; ++hh) c0 = fmax(c0, a.lconst[hh]);
        dc2[threadIdx.x] = (float)((a.lconst[threadIdx.x] - c0) * 1.4426950408889634);
    }

    float sb[MJ], nsbmu[MJ], rm2[MJ], cq2[MJ], rsb[MJ];
    double sb64[B64 ? MJ : 1], nsbmu64[B64 ? MJ : 1];
    /* FP32 path, m = 3: the FP64 constants live in LDS ([j][2][thread],
     * conflict-free) and are read once per tile (the barrier orders them),
     * which keeps the kernel within 128 VGPRs: 4 waves per SIMD */
    constexpr bool kC64Lds = B64 && MJ == 3;
    __shared__ double c64s[kC64Lds ? 2 * MJ * 512 : 1];
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
            /* a padding channel: b = 0 gives y = 1, |y|^rho = 1, so
             * q2 = (0, -1e30, ...) and the component sum is exactly 1: its
             * log (the channel's share of L) is exactly 0, and g = 0 */
            sb[j] = 0.f; nsbmu[j] = 1.f; rm2[j] = 0.f; cq2[j] = j == 0 ? AMICA_L2E : -1e30f;
            rsb[j] = 0.f;
        }
        if constexpr (B64) {
            const double* dj = a.cst64 + ((size_t)h * kNumCst64 * MJ + j) * a.NP + i;
            sb64[j] = act ? dj[kCst64Sb * ks] : 0.0;
            nsbmu64[j] = act ? dj[kCst64Nsbmu * ks] : 1.0;
            if constexpr (kC64Lds) {
                c64s[(2 * j) * 512 + threadIdx.x] = sb64[j];
                c64s[(2 * j + 1) * 512 + threadIdx.x] = nsbmu64[j];
            }
        }
    }
    if constexpr (MODE != kEstepLsum) {
#pragma unroll
        for (int s = 0; s < NS; ++s) tot2[s] = 0.f;
    }

    float acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = 0.f;
    double vsum_acc = 0.0, ll_acc = 0.0;
    double* const base = a.part + (size_t)blockIdx.x * a.pstride;
    /* this lane's FP64 totals: [M][NP][NS], contiguous per lane */
    double* const l3 = base + ((size_t)h * a.NP + i) * NS;
    bool l3_started = false;

    const long t0 = (long)blockIdx.x * a.chunk;
    const long t1 = min(a.T, t0 + a.chunk);
    const size_t bcol = (size_t)h * a.NP + i;
    int buf = 0, tiles = 0, flushes = 0;
    /* bf16 g, 8 samples packed per column: [t/8][ldb][8], one 16-byte store
     * per 8-sample group (t0 is a multiple of 8) */
    uint4* const g8 = (uint4*)a.g;

    /* the loads of the first tile.  Loads carry no channel check: lanes
     * i in [n, NP) read the padding columns of b and x, which hold zeros
     * (the padding rows of W and entries of W c are zero, and x's padding
     * columns are written as zero), so every load is unconditional and the
     * wait for a prefetch can sit where its value is first used. */
    float bn[kS], xn[kS];
    double bn64[B64 ? kS : 1];
    /* b and x through buffer resources based at row t0: the lane's byte
     * offset is fixed, the row's a scalar (launch_estep bounds it below
     * 2^31), so a load costs no address arithmetic on the vector ALU */
    constexpr int kEb = B64 ? 8 : 4;
    const char* const bbase = (const char*)(B64 ? (const void*)a.b64 : (const void*)a.b);
    auto rsrc_b = [&](long t) {
        return __builtin_amdgcn_make_buffer_rsrc((void*)(bbase + (size_t)t * a.ldb * kEb), (short)0,
                                                 (int)0xffffffff, kBufferRsrcWord3);
    };
    auto rsrc_x = [&](long t) {
        return __builtin_amdgcn_make_buffer_rsrc((void*)(a.x + (size_t)t * a.NP), (short)0,
                                                 (int)0xffffffff, kBufferRsrcWord3);
    };
    /* rows are addressed relative to tbase: the workgroup's first row (fused
     * launch, whose row offsets launch_estep bounds), or in the split
     * launches -- whose rows may be far wider -- the current 8-sample group */
    long tbase = t0;
    __amdgpu_buffer_rsrc_t rb = rsrc_b(t0), rx = rsrc_x(t0);
    const int vob = (int)bcol * kEb, vox = i * 4;
    const int browb = a.ldb * kEb, xrowb = a.NP * 4;
    auto load_tile = [&](long t, int s) {
        const int r = (int)(t - tbase);
        if constexpr (B64)
            bn64[s] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rb, vob, r * browb, 0));
        else
            bn[s] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rb, vob, r * browb, 0));
        xn[s] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, vox, r * xrowb, 0));
    };
#pragma unroll
    for (int s = 0; s < kS; ++s) load_tile(t0 + s, s);

    constexpr int kTPG = 8 / kS;               /* tiles per packed 8-sample group */
    for (long tg = t0; tg < t1; tg += 8) {
        if constexpr (kSplit) {
            /* loads already in flight keep the resources they were issued with */
            tbase = tg;
            rb = rsrc_b(tg);
            rx = rsrc_x(tg);
        }
        /* the group's packed g words (2 samples each); tiles past t1 in the
         * last group run with posterior 0 and so give g = 0 */
        uint32_t gw[4] = {0u, 0u, 0u, 0u};
        float gpend = 0.f;
#pragma unroll
        for (int sub = 0; sub < kTPG; ++sub) {
            const long tb = tg + sub * kS;
            float E[kS][MJ], SY[kS][MJ], YM2[kS][MJ], YR[kS][MJ], LA[kS][MJ];
            float RS[kS], BV[kS], XV[kS];
            double BV64[B64 ? kS : 1];
#pragma unroll
            for (int s = 0; s < kS; ++s) {
                if constexpr (B64) { BV64[s] = bn64[s]; BV[s] = (float)bn64[s]; }
                else BV[s] = bn[s];
                XV[s] = xn[s];
            }
            /* issue the next tile's b and x loads now; they complete under this
             * tile's math.  Rows past t1 are read without a check (the next
             * range's samples, or the kEstepRowPad zero rows after the last
             * one) and get posterior 0. */
#pragma unroll
            for (int s = 0; s < kS; ++s) load_tile(tb + kS + s, s);
            /* ---- phase 1: densities, component posteriors, channel sums ---- */
            float QS[kS];
#pragma unroll
            for (int s = 0; s < kS; ++s) {
                float q2[MJ];
#pragma unroll
                for (int j = 0; j < MJ; ++j) {
                    float y;
                    if constexpr (kC64Lds)
                        y = (float)fma(c64s[(2 * j) * 512 + threadIdx.x], BV64[s],
                                       c64s[(2 * j + 1) * 512 + threadIdx.x]);
                    else if constexpr (B64) y = (float)fma(sb64[j], BV64[s], nsbmu64[j]);
                    else y = fmaf(sb[j], BV[s], nsbmu[j]);
                    const float la = fast_log2(fmaxf(fabsf(y), y_floor<B64>()));
                    const float ym2 = fast_exp2(rm2[j] * la);
                    const float sy = ym2 * y;
                    const float yr = sy * y;
                    q2[j] = fmaf(yr, -AMICA_L2E, cq2[j]);
                    SY[s][j] = sy; YM2[s][j] = ym2; YR[s][j] = yr; LA[s][j] = la;
                }
                /* q2 <= log2(alpha sqrt(beta) / (2 Gamma(1 + 1/rho))), a few
                 * units, so 2^q2 needs no max shift unless every component
                 * underflows (a sample far from all of them): then, for the
                 * whole wave, shift by the per-lane max */
#pragma unroll
                for (int j = 0; j < MJ; ++j) E[s][j] = fast_exp2(q2[j]);
                float se = E[s][0];
#pragma unroll
                for (int j = 1; j < MJ; ++j) se += E[s][j];
                const bool under = __any(se < 0x1p-100f);
                QS[s] = fast_log2(se);            /* exactly 0 on padding channels */
                if (__builtin_expect(under, 0)) {
                    float off = q2[0];
#pragma unroll
                    for (int j = 1; j < MJ; ++j) off = fmaxf(off, q2[j]);
#pragma unroll
                    for (int j = 0; j < MJ; ++j) E[s][j] = fast_exp2(q2[j] - off);
                    se = E[s][0];
#pragma unroll
                    for (int j = 1; j < MJ; ++j) se += E[s][j];
                    QS[s] = off + fast_log2(se);
                }
                RS[s] = fast_rcp(se);
            }
            const int nval = (int)min(t1 - tb, (long)kS);    /* samples of the tile before t1 */
            float GV[kS], VV[kS];
            if constexpr (MODE == kEstepStats) {
                /* the posteriors of the split launch's middle pass */
#pragma unroll
                for (int s = 0; s < kS; ++s) VV[s] = s < nval ? a.vq[(size_t)(tb + s) * a.M + h] : 0.f;
            } else {
#pragma unroll
            for (int s = 0; s < kS; ++s) QS[s] = row_sum_lane15(QS[s]);
            if ((lane & 15) == 15) {
#pragma unroll
                for (int s = 0; s < kS; ++s) Lp[buf][w][s][lane >> 4] = QS[s];
            }
            __syncthreads();
            if constexpr (MODE == kEstepLsum) {
                /* wave 0: lane 16 s + hl adds the row sums of local model hl
                 * for sample s and writes them for the posterior pass */
                if (w == 0) {
                    const int hl = lane & 15, s = lane >> 4;
                    const int hh = hg * a.mg + hl;
                    if (s < kS && s < nval && hl < a.mg && hh < a.M) {
                        float l = 0.f;
                        for (int c = 0; c < a.cwg; ++c)
#pragma unroll
                            for (int row = 0; row < 4; ++row) l += Lp[buf][hl * a.cwg + c][s][row];
                        a.lq[((size_t)(tb + s) * a.M + hh) * a.ncg + cg] = l;
                    }
                }
                buf ^= 1;
                continue;
            } else {
            /* ---- model posteriors, by wave 0 alone: lane hh*kS + s holds model
             * hh, sample s; the other waves read them from LDS after a second
             * barrier ---- */
            if (w == 0) {
                /* the other waves wait for this one at the next barrier: let it
                 * issue ahead of the other workgroup's waves on its SIMD */
                __builtin_amdgcn_s_setprio(3);
                float vl;
                const int hh = lane / kS;
                const int s = lane - hh * kS;
                const long t = tb + s;
                const bool valid = s < nval;
                float l = -3.0e38f;
                if (hh < a.M) {
                    l = dc2[hh];
                    for (int c = 0; c < a.CW; ++c)
#pragma unroll
                        for (int row = 0; row < 4; ++row) l += Lp[buf][hh * a.CW + c][s][row];
                }
                /* logsumexp over the models: the lanes of one sample are kS
                 * apart inside the first row of 16 lanes (M <= 8), so row
                 * rotations by (1,) 2, 4 and 8 reach all of them */
                float lmax = l;
                if constexpr (kS == 1) lmax = row_ror_max<1>(lmax);
                lmax = row_ror_max<2>(lmax);
                lmax = row_ror_max<4>(lmax);
                lmax = row_ror_max<8>(lmax);
                /* lanes hh >= M hold l = -3e38: e = 0 (the first row's lmax is
                 * a finite model value; lanes past it are never read) */
                const float e = fast_exp2(l - lmax);
                float esum = e;
                if constexpr (kS == 1) esum += dpp<0x121, 0xf, 0xf, false>(esum);
                esum += dpp<0x122, 0xf, 0xf, false>(esum);
                esum += dpp<0x124, 0xf, 0xf, false>(esum);
                esum += dpp<0x128, 0xf, 0xf, false>(esum);
                vl = valid ? e * fast_rcp(esum) : 0.f;
                if (lane < a.M * kS) Vp[buf][lane] = vl;
                if (valid && hh < a.M) {
                    vsum_acc += (double)vl;
                    if (hh == 0) ll_acc += (double)(lmax + fast_log2(esum)) * AMICA_LN2;
                    if (a.ltall) a.ltall[(size_t)hh * a.T + t] = l * (float)AMICA_LN2;
                }
                __builtin_amdgcn_s_setprio(0);
            }
            __syncthreads();
            if constexpr (kS == 2) {                      /* both samples in one 8-byte read */
                const float2 vp = *(const float2*)&Vp[buf][h * 2];
                VV[0] = vp.x;
                VV[1] = vp.y;
            } else {
                VV[0] = Vp[buf][h];
            }
            }
            }
            /* ---- phase 2: responsibilities and statistics ---- */
#pragma unroll
            for (int s = 0; s < kS; ++s) {
                const long t = tb + s;
                const float v = VV[s];
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
                /* g is 0 on padding channels (rsb = 0) and past t1 (v = 0) */
                const float vb = v * BV[s];
                acc[NSJ * MJ + 0] = fmaf(vb, BV[s], acc[NSJ * MJ + 0]);
                acc[NSJ * MJ + 1] = fmaf(v, XV[s], acc[NSJ * MJ + 1]);
                acc[NSJ * MJ + 2] += g;
                GV[s] = g;
                if (!a.g_bf16 && s < nval && wreal) ((float*)a.g)[(size_t)t * a.ldb + bcol] = g;
            }
            if constexpr (kS == 2) gw[sub] = pack_bf16x2(GV[0], GV[1]);
            else if (sub & 1) gw[sub >> 1] = pack_bf16x2(gpend, GV[0]);
            else gpend = GV[0];
            buf ^= 1;
            if (++tiles == kFlush) {
                tiles = 0;
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    tot2[s] += acc[s];
                    acc[s] = 0.f;
                }
                if (++flushes == kFlush2) {
                    flushes = 0;
                    if constexpr (kL3Batched) {
                        if (wreal)
                            l3_accumulate<NS, 8>(l3, l3_started, true, [&](int s) { return (double)tot2[s]; });
#pragma unroll
                        for (int s = 0; s < NS; ++s) tot2[s] = 0.f;
                    } else {
#pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            const double v = (double)tot2[s];
                            if (wreal) l3[s] = l3_started ? l3[s] + v : v;
                            tot2[s] = 0.f;
                        }
                    }
                    l3_started = true;
                }
            }
        }
        if constexpr (MODE != kEstepLsum)
            if (a.g_bf16 && wreal) g8[(size_t)(tg >> 3) * a.ldb + bcol] = make_uint4(gw[0], gw[1], gw[2], gw[3]);
    }
    if constexpr (MODE == kEstepLsum) return;
    if constexpr (kL3Batched) {
        if (wreal)
            l3_accumulate<NS, 8>(l3, l3_started, act, [&](int s) { return (double)tot2[s] + (double)acc[s]; });
    } else {
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double v = (double)tot2[s] + (double)acc[s];
            const double r = l3_started ? l3[s] + v : v;
            if (wreal) l3[s] = act ? r : 0.0;
        }
    }
    if constexpr (MODE == kEstepStats) return;   /* vsum and the LL sum: estep_post_kernel */
    /* per-model vsum (lanes hh*kS + s of wave 0) and the LL sum (lanes s) */
    if (w == 0) {
        const size_t tail = (size_t)a.M * NS * a.NP;
        for (int hh = 0; hh < a.M; ++hh) {
            const double vs = wave_sum_f64((lane / kS == hh) ? vsum_acc : 0.0);
            if (lane == 0) base[tail + hh] = vs;
        }
        const double ls = wave_sum_f64(ll_acc);
        if (lane == 0) base[tail + a.M] = ls;
    }
}

/* M = 1: the model posterior v is 1, so the waves of a workgroup (the
 * 64-channel blocks of one range of samples) never exchange anything: no
 * barrier, no LDS, and phase 2 of a sample follows its phase 1 in
 * registers.  One sample per iteration, loads 3-4 samples ahead; the
 * statistics are summed in FP32 registers over 32 samples, in FP32 per lane
 * in LDS over 1024 samples (odd stride: distinct banks) and in FP64 in the
 * workgroup's slot of the partial buffer.  The log-likelihood sum is sum_t sum_i q_ti (the
 * model constant is added by the reduction); vsum is the sample count.
 * Not used when per-sample L is recorded (ltall). */
template <int MJ, bool B64>
__device__ __forceinline__ void estep_m1_body(const EstepArgs& a, const long t0, const long t1,
                                              const int slot) {
    constexpr int NSJ = kStatsPerJ;
    constexpr int NS = NSJ * MJ + kStatsPerI;
    constexpr int G = 8;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = w * 64 + lane;
    const bool act = i < a.n;
    __shared__ double llw[kMaxEstepWaves];

    constexpr int NSP = NS | 1;
    extern __shared__ float tot2_all[];
    float* const tot2 = tot2_all + threadIdx.x * NSP;
    /* FP32 path, m = 3: the FP64 constants live in LDS ([2 j + k][thread]
     * after the totals, each thread its own slots), read once per sample:
     * 12 VGPRs fewer, so that 4 waves per SIMD fit */
    constexpr bool kC64Lds = B64 && MJ == 3;
    double* const c64 = (double*)((char*)tot2_all + ((size_t)blockDim.x * NSP * sizeof(float) + 7) / 8 * 8);

    float sb[MJ], nsbmu[MJ], rm2[MJ], cq2[MJ], rsb[MJ];
    double sb64[B64 ? MJ : 1], nsbmu64[B64 ? MJ : 1];
#pragma unroll
    for (int j = 0; j < MJ; ++j) {
        const float* cj = a.cst + (size_t)j * a.NP + i;
        const size_t ks = (size_t)MJ * a.NP;
        /* a padding channel: y = 1, so q2 = (0, -1e30, ...) and the
         * component sum is exactly 1 (log 0 to the LL, g = 0) */
        sb[j] = act ? cj[kCstSb * ks] : 0.f;
        nsbmu[j] = act ? cj[kCstNsbmu * ks] : 1.f;
        rm2[j] = act ? cj[kCstRm2 * ks] : 0.f;
        cq2[j] = act ? cj[kCstCq2 * ks] : (j == 0 ? AMICA_L2E : -1e30f);
        rsb[j] = act ? cj[kCstRsb * ks] : 0.f;
        if constexpr (B64) {
            const double* dj = a.cst64 + (size_t)j * a.NP + i;
            sb64[j] = act ? dj[kCst64Sb * ks] : 0.0;
            nsbmu64[j] = act ? dj[kCst64Nsbmu * ks] : 1.0;
            if constexpr (kC64Lds) {
                c64[(size_t)(2 * j) * blockDim.x + threadIdx.x] = sb64[j];
                c64[(size_t)(2 * j + 1) * blockDim.x + threadIdx.x] = nsbmu64[j];
            }
        }
    }
    float acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        acc[s] = 0.f;
        tot2[s] = 0.f;
    }
    float qacc = 0.f;
    double qtot = 0.0;
    double* const base = a.part + (size_t)slot * a.pstride;
    double* const l3 = base + (size_t)i * NS;
    bool l3_started = false;

    const size_t bcol = (size_t)i;
    uint4* const g8 = (uint4*)a.g;

    /* a ring of PF prefetched samples; the loop runs whole packed
     * 8-sample groups with the sample loop unrolled, so that ring slots and
     * packed g words are fixed registers (PF divides 8) */
    constexpr int PF = 4;
    static_assert(G % PF == 0, "ring slots must repeat every group");
    float bq[PF], xq[PF];
    double dq[PF];
    /* Unconditional loads through buffer resources based at row t0 (scalar
     * row offsets, bounded by launch_estep): samples past T read row T, the
     * first of the zero padding rows (a range's samples past t1 below T are
     * only ever prefetched, never used), and lanes i in [n, NP) the zero
     * padding columns of b and x (the padding rows of W and entries of W c
     * are zero, and x's padding columns are written as zero). */
    constexpr int kEb = B64 ? 8 : 4;
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        (void*)((const char*)(B64 ? (const void*)a.b64 : (const void*)a.b) + (size_t)t0 * a.ldb * kEb),
        (short)0, (int)0xffffffff, kBufferRsrcWord3);
    const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.x + (size_t)t0 * a.NP), (short)0, (int)0xffffffff, kBufferRsrcWord3);
    const int vob = (int)bcol * kEb, vox = i * 4;
    auto ld = [&](long t, float& bqv, double& dqv, float& xqv) {
        const int r = (int)(min(t, a.T) - t0);
        if constexpr (B64)
            dqv = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rb, vob, r * a.ldb * kEb, 0));
        else
            bqv = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rb, vob, r * a.ldb * kEb, 0));
        xqv = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rx, vox, r * a.NP * 4, 0));
    };
#pragma unroll
    for (int p = 0; p < PF; ++p) {
        bq[p] = 0.f;
        dq[p] = 0.0;
        ld(t0 + p, bq[p], dq[p], xq[p]);
    }
    const long tend = t0 + (t1 - t0 + G - 1) / G * G;   /* whole 8-sample groups */
    int grp = 0;
#pragma unroll 1
    for (long tg = t0; tg < tend; tg += G, ++grp) {
        uint32_t gw[G / 2];                           /* packed bf16 g, 2 samples a word */
        float gprev = 0.f;
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const long t = tg + u;
            const bool valid = t < t1;                /* wave-uniform */
            const int sl = u % PF;
            const float bcur = B64 ? 0.f : bq[sl];
            const double dcur = dq[sl];
            const float xcur = xq[sl];
            ld(t + PF, bq[sl], dq[sl], xq[sl]);
            const float bv = B64 ? (float)dcur : bcur;
            /* the constants' LDS reads stay in the loop (not hoisted back into
             * registers): a compiler-only fence, no instruction */
            if constexpr (kC64Lds) __asm__ volatile("" ::: "memory");
            float q2[MJ], E[MJ], SY[MJ], YM2[MJ], YR[MJ], LA[MJ];
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                float y;
                if constexpr (kC64Lds)
                    y = (float)fma(c64[(size_t)(2 * j) * blockDim.x + threadIdx.x], dcur,
                                   c64[(size_t)(2 * j + 1) * blockDim.x + threadIdx.x]);
                else if constexpr (B64) y = (float)fma(sb64[j], dcur, nsbmu64[j]);
                else y = fmaf(sb[j], bv, nsbmu[j]);
                LA[j] = fast_log2(fmaxf(fabsf(y), y_floor<B64>()));
                YM2[j] = fast_exp2(rm2[j] * LA[j]);
                SY[j] = YM2[j] * y;
                YR[j] = SY[j] * y;
                q2[j] = fmaf(YR[j], -AMICA_L2E, cq2[j]);
            }
            float se = 0.f, off = 0.f;
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                E[j] = fast_exp2(q2[j]);
                se += E[j];
            }
            if (__builtin_expect(__any(se < 0x1p-100f), 0)) {
                off = q2[0];
#pragma unroll
                for (int j = 1; j < MJ; ++j) off = fmaxf(off, q2[j]);
                se = 0.f;
#pragma unroll
                for (int j = 0; j < MJ; ++j) {
                    E[j] = fast_exp2(q2[j] - off);
                    se += E[j];
                }
            }
            /* samples past t1 (the last range's last group) count nothing */
            const float rs = valid ? fast_rcp(se) : 0.f;
            qacc += valid ? off + fast_log2(se) : 0.f;
            float g = 0.f;
#pragma unroll
            for (int j = 0; j < MJ; ++j) {
                const float u_ = E[j] * rs;
                const float uf = u_ * SY[j];
                g = fmaf(uf, rsb[j], g);
                float* aj = acc + j * NSJ;
                aj[0] += u_;
                aj[1] += uf;
                aj[2] = fmaf(uf, SY[j], aj[2]);
                aj[3] = fmaf(u_, YM2[j], aj[3]);
                const float uyr = u_ * YR[j];
                aj[4] += uyr;
                aj[5] = fmaf(uyr, YR[j], aj[5]);
                aj[6] = fmaf(uyr, LA[j], aj[6]);
            }
            acc[NSJ * MJ + 0] = fmaf(bv, bv, acc[NSJ * MJ + 0]);
            acc[NSJ * MJ + 1] += xcur;
            acc[NSJ * MJ + 2] += g;
            if (a.g_bf16) {
                if (u & 1) gw[u >> 1] = pack_bf16x2(gprev, g);
                gprev = g;
            } else if (valid) {
                ((float*)a.g)[(size_t)t * a.ldb + bcol] = g;
            }
            /* one sample at a time: the sums are complete here (empty asm
             * with them as operands, no instruction) and the scheduler may
             * not interleave the unrolled samples -- deferring the sums would
             * keep every sample's values live, beyond 128 VGPRs */
#pragma unroll
            for (int s = 0; s < NS; ++s) __asm__ volatile("" : "+v"(acc[s]));
            __asm__ volatile("" : "+v"(qacc));
            __builtin_amdgcn_sched_barrier(0);
        }
        /* t0 is a multiple of 8: the group's packed word */
        if (a.g_bf16) g8[(size_t)((tg) >> 3) * a.ldb + bcol] = make_uint4(gw[0], gw[1], gw[2], gw[3]);
        if ((grp & 3) == 3) {                         /* 32 samples */
#pragma unroll
            for (int s = 0; s < NS; ++s) { tot2[s] += acc[s]; acc[s] = 0.f; }
            qtot += (double)qacc;
            qacc = 0.f;
            if ((grp & 127) == 127) {                 /* 1024 samples */
                l3_accumulate<NS, 8>(l3, l3_started, true, [&](int s) { return (double)tot2[s]; });
#pragma unroll
                for (int s = 0; s < NS; ++s) tot2[s] = 0.f;
                l3_started = true;
            }
        }
    }
    l3_accumulate<NS, 8>(l3, l3_started, act, [&](int s) { return (double)tot2[s] + (double)acc[s]; });
    qtot += (double)qacc;
    const double ls = wave_sum_f64(qtot);
    if (lane == 0) llw[w] = ls;
    __syncthreads();
    if (threadIdx.x == 0) {
        double l = 0.0;
        for (int c = 0; c < (int)(blockDim.x >> 6); ++c) l += llw[c];
        const size_t tail = (size_t)NS * a.NP;
        base[tail] = (double)(t1 - t0);
        base[tail + 1] = l * AMICA_LN2;
    }
}

template <int MJ, bool B64>
__global__ __launch_bounds__(512) void estep_m1_kernel(EstepArgs a) {
    const long t0 = (long)blockIdx.x * a.chunk;
    estep_m1_body<MJ, B64>(a, t0, min(a.T, t0 + a.chunk), (int)blockIdx.x);
}

/* Many recordings in one launch: workgroup b runs work item b, a range of
 * samples of one recording (args[item.rec]) into that recording's partial
 * slot item.slot.  Every recording of a launch has the same NP (blockDim). */
template <int MJ, bool B64>
__global__ __launch_bounds__(512) void estep_m1_batch_kernel(const EstepArgs* __restrict__ args,
                                                             const BatchItem* __restrict__ items) {
    const BatchItem it = items[blockIdx.x];
    const EstepArgs a = args[it.rec];
    estep_m1_body<MJ, B64>(a, it.t0, it.t1, it.slot);
}

/* The split launch's middle pass: per sample, the model log-likelihoods
 * L_h = (log gm_h + ldet_h + ldetS) + sum of q over all channels (the
 * channel-group sums of the first pass), their log-sum-exp (the sample's
 * log-likelihood, as updatell, kernels.cu:376-450) and the posteriors
 * v_h = exp(L_h - logsumexp) for the statistics pass.  One workgroup per
 * sample chunk (the chunks of the other passes): vsum per model and the LL
 * sum go to the tail of the chunk's partial slot, in FP64. */
__global__ __launch_bounds__(256) void estep_post_kernel(EstepArgs a) {
    __shared__ float dc2[AMICA_MAX_M];
    __shared__ double red[4][AMICA_MAX_M + 1];
    const int M = a.M, ncg = a.ncg;
    if ((int)threadIdx.x < M) {
        double c0 = a.lconst[0];
        for (int hh = 1; hh < M; ++hh) c0 = fmax(c0, a.lconst[hh]);
        dc2[threadIdx.x] = (float)((a.lconst[threadIdx.x] - c0) * 1.4426950408889634);
    }
    __syncthreads();
    const long t0 = (long)blockIdx.x * a.chunk;
    const long t1 = min(a.T, t0 + a.chunk);
    double ll = 0.0, vs[AMICA_MAX_M];
#pragma unroll
    for (int hh = 0; hh < AMICA_MAX_M; ++hh) vs[hh] = 0.0;
    for (long t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
        float l[AMICA_MAX_M];
        float lmax = -3.0e38f;
#pragma unroll
        for (int hh = 0; hh < AMICA_MAX_M; ++hh) {
            if (hh < M) {
                float v = dc2[hh];
                const float* q = a.lq + ((size_t)t * M + hh) * ncg;
                for (int c = 0; c < ncg; ++c) v += q[c];
                l[hh] = v;
                lmax = fmaxf(lmax, v);
            }
        }
        float esum = 0.f;
#pragma unroll
        for (int hh = 0; hh < AMICA_MAX_M; ++hh)
            if (hh < M) esum += fast_exp2(l[hh] - lmax);
        const float r = fast_rcp(esum);
#pragma unroll
        for (int hh = 0; hh < AMICA_MAX_M; ++hh) {
            if (hh < M) {
                const float v = fast_exp2(l[hh] - lmax) * r;
                a.vq[(size_t)t * M + hh] = v;
                vs[hh] += (double)v;
                if (a.ltall) a.ltall[(size_t)hh * a.T + t] = l[hh] * (float)AMICA_LN2;
            }
        }
        ll += (double)(lmax + fast_log2(esum)) * AMICA_LN2;
    }
    /* fixed-order block sums: waves, then the 4 wave totals */
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double lls = wave_sum_f64(ll);
    if (lane == 0) red[w][M] = lls;
#pragma unroll
    for (int hh = 0; hh < AMICA_MAX_M; ++hh) {
        if (hh < M) {
            const double v = wave_sum_f64(vs[hh]);
            if (lane == 0) red[w][hh] = v;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x <= M) {
        const double tot = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
        const int NS = kStatsPerJ * a.mj + kStatsPerI;
        a.part[(size_t)blockIdx.x * a.pstride + (size_t)M * NS * a.NP + threadIdx.x] = tot;
    }
}

}  // namespace

void estep_split_shape(int M, int CW, int* mg, int* cwg, int* ncg, int* ngroups) {
    /* channel blocks per group: all of a model's if they fit one workgroup,
     * else the fewest groups of at most kMaxEstepWaves blocks, balanced */
    const int g = (CW + kMaxEstepWaves - 1) / kMaxEstepWaves;
    const int c = (CW + g - 1) / g;
    const int models = std::max(1, std::min(M, kMaxEstepWaves / c));
    *cwg = c;
    *ncg = g;
    *mg = models;
    *ngroups = ((M + models - 1) / models) * g;
}

int estep_configure(int MJ, size_t lds_bytes) {
    /* one attribute per kernel for the whole process: keep the largest that
     * any engine asked for (engines of other shapes launch with less) */
    static size_t granted[AMICA_MAX_MJ + 1] = {0};
    if (MJ < 1 || MJ > AMICA_MAX_MJ) return -1;
    if (lds_bytes <= granted[MJ]) return 0;
    granted[MJ] = lds_bytes;
    hipError_t r = hipSuccess;
#define AMICA_ESTEP_ATTR(J)                                                                       \
    case J:                                                                                       \
        r = hipFuncSetAttribute((const void*)estep_kernel<J, true, kEstepFused>,                 \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);      \
        if (r == hipSuccess)                                                                      \
            r = hipFuncSetAttribute((const void*)estep_kernel<J, false, kEstepFused>,            \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);  \
        if (r == hipSuccess)                                                                      \
            r = hipFuncSetAttribute((const void*)estep_kernel<J, true, kEstepStats>,             \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);  \
        if (r == hipSuccess)                                                                      \
            r = hipFuncSetAttribute((const void*)estep_kernel<J, false, kEstepStats>,            \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);  \
        if (r == hipSuccess)                                                                      \
            r = hipFuncSetAttribute((const void*)estep_m1_batch_kernel<J, false>,                \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);  \
        if (r == hipSuccess)                                                                      \
            r = hipFuncSetAttribute((const void*)estep_m1_kernel<J, true>,                       \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);  \
        if (r == hipSuccess)                                                                      \
            r = hipFuncSetAttribute((const void*)estep_m1_kernel<J, false>,                      \
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

size_t estep_lds_bytes(int MJ, int threads, bool m1b64) {
    const size_t tot = (size_t)((kStatsPerJ * MJ + kStatsPerI) | 1) * threads * sizeof(float);
    /* the single-model FP32-path kernel at m = 3 also keeps its FP64
     * constants there (2 m doubles per thread after the totals) */
    return (m1b64 && MJ == 3) ? (tot + 7) / 8 * 8 + (size_t)2 * MJ * threads * sizeof(double) : tot;
}

int estep_blocks_per_cu(int MJ, bool b64, int threads, bool m1, bool split) {
    int nb = 0;
    const size_t shm = estep_lds_bytes(MJ, threads, m1 && b64);
    hipError_t r = hipErrorInvalidValue;
#define AMICA_ESTEP_OCC(J)                                                                        \
    case J:                                                                                       \
        if (m1)                                                                                   \
            r = b64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, estep_m1_kernel<J, true>, \
                                                                   threads, shm)                  \
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, estep_m1_kernel<J, false>,\
                                                                   threads, shm);                 \
        else if (split)                                                                           \
            r = b64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(                               \
                          &nb, estep_kernel<J, true, kEstepStats>, threads, shm)                  \
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(                               \
                          &nb, estep_kernel<J, false, kEstepStats>, threads, shm);                \
        else                                                                                      \
            r = b64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(                               \
                          &nb, estep_kernel<J, true, kEstepFused>, threads, shm)                  \
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(                               \
                          &nb, estep_kernel<J, false, kEstepFused>, threads, shm);                \
        break;
    switch (MJ) {
        AMICA_ESTEP_OCC(1)
        AMICA_ESTEP_OCC(2)
        AMICA_ESTEP_OCC(3)
        AMICA_ESTEP_OCC(4)
        AMICA_ESTEP_OCC(5)
        AMICA_ESTEP_OCC(6)
        AMICA_ESTEP_OCC(7)
        AMICA_ESTEP_OCC(8)
        default: return -1;
    }
#undef AMICA_ESTEP_OCC
    return r == hipSuccess ? nb : -1;
}

static int launch_estep_split(const EstepArgs& a, int MJ, int nblocks, hipStream_t st) {
    int mg = 0, cwg = 0, ncg = 0, ngroups = 0;
    estep_split_shape(a.M, a.CW, &mg, &cwg, &ncg, &ngroups);
    if (a.mg != mg || a.cwg != cwg || a.ncg != ncg || !a.lq || !a.vq || a.mj != MJ || (a.chunk % 8) != 0)
        return -1;
    const dim3 grid(nblocks, ngroups), block(64 * mg * cwg);
    const size_t shm = estep_lds_bytes(MJ, (int)block.x);
    const bool b64 = a.b64 != nullptr;
#define AMICA_ESTEP_SPLIT(J)                                                                          \
    case J:                                                                                           \
        if (b64) hipLaunchKernelGGL((estep_kernel<J, true, kEstepLsum>), grid, block, 0, st, a);      \
        else hipLaunchKernelGGL((estep_kernel<J, false, kEstepLsum>), grid, block, 0, st, a);         \
        hipLaunchKernelGGL(estep_post_kernel, dim3(nblocks), dim3(256), 0, st, a);                   \
        if (b64) hipLaunchKernelGGL((estep_kernel<J, true, kEstepStats>), grid, block, shm, st, a);   \
        else hipLaunchKernelGGL((estep_kernel<J, false, kEstepStats>), grid, block, shm, st, a);      \
        break;
    switch (MJ) {
        AMICA_ESTEP_SPLIT(1)
        AMICA_ESTEP_SPLIT(2)
        AMICA_ESTEP_SPLIT(3)
        AMICA_ESTEP_SPLIT(4)
        AMICA_ESTEP_SPLIT(5)
        AMICA_ESTEP_SPLIT(6)
        AMICA_ESTEP_SPLIT(7)
        AMICA_ESTEP_SPLIT(8)
        default: return -1;
    }
#undef AMICA_ESTEP_SPLIT
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_estep_m1_batch(const EstepArgs* args, const BatchItem* items, int nitems, int MJ, int NP,
                          hipStream_t st) {
    if (nitems < 1 || NP % 64 || NP > 64 * kMaxEstepWaves) return -1;
    const dim3 grid(nitems), block(NP);
    const size_t shm = estep_lds_bytes(MJ, NP);
#define AMICA_ESTEP_BATCH(J)                                                                        \
    case J:                                                                                         \
        hipLaunchKernelGGL((estep_m1_batch_kernel<J, false>), grid, block, shm, st, args, items); \
        break;
    switch (MJ) {
        AMICA_ESTEP_BATCH(1)
        AMICA_ESTEP_BATCH(2)
        AMICA_ESTEP_BATCH(3)
        AMICA_ESTEP_BATCH(4)
        AMICA_ESTEP_BATCH(5)
        AMICA_ESTEP_BATCH(6)
        AMICA_ESTEP_BATCH(7)
        AMICA_ESTEP_BATCH(8)
        default: return -1;
    }
#undef AMICA_ESTEP_BATCH
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_estep(const EstepArgs& a, int MJ, int nblocks, hipStream_t st) {
    if (a.split) return launch_estep_split(a, MJ, nblocks, st);
    if (a.M * a.CW > kMaxEstepWaves || a.M > 8 || (a.chunk % 8) != 0) return -1;
    /* byte offsets of the rows a workgroup reads (its chunk and the next
     * tile past it) must fit the 31-bit scalar offsets of its buffer loads */
    if ((a.chunk + kEstepRowPad) * (long)a.ldb * 8 >= (1L << 31)) return -1;
    dim3 grid(nblocks), block(64 * a.M * a.CW);
    const bool b64 = a.b64 != nullptr;
    const bool m1 = a.M == 1 && a.ltall == nullptr;
    const size_t shm = estep_lds_bytes(MJ, (int)block.x, m1 && b64);
#define AMICA_ESTEP_CASE(J)                                                                   \
    case J:                                                                                   \
        if (m1) {                                                                             \
            if (b64) hipLaunchKernelGGL((estep_m1_kernel<J, true>), grid, block, shm, st, a); \
            else hipLaunchKernelGGL((estep_m1_kernel<J, false>), grid, block, shm, st, a);    \
        } else {                                                                              \
            if (b64) hipLaunchKernelGGL((estep_kernel<J, true, kEstepFused>), grid, block, shm, st, a); \
            else hipLaunchKernelGGL((estep_kernel<J, false, kEstepFused>), grid, block, shm, st, a);    \
        }                                                                                     \
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
