/*
 * mstep.hip -- statistic reduction, M-step, Newton step, reparameterisation
 * and the per-iteration preparation (W = A^-1, log|det W|, E-step constants)
 * of the AMICA EM iteration, all in FP64 on the device so that one iteration
 * never returns to the host.
 *
 * Follows fraimondo/cudamica:
 *   cross-block reduction      src/amica.cu:1618-1725 (acumulate)
 *   LL, lrate on decrease      src/amica.cu:1963-1993
 *   gm, alpha, sigma2, cnew, phi scaling   src/amica.cu:1995-2053
 *   kappa, lambda              src/kernels.cu:848-867 (updatekappalambda)
 *   mu, beta                   src/amica.cu:2095-2137 (divide / acumulate / multiply)
 *   rho                        src/kernels.cu:906-954 (psi, updaterho)
 *   Newton / natural gradient  src/amica.cu:2156-2279, kernels.cu:961-1014
 *   lrate growth, A update     src/amica.cu:2281-2364
 *   reparameterisation         src/amica.cu:2367-2447, kernels.cu:1023-1065
 *   W = pinv(A), ldet          src/amica.cu:27-127, 835-924
 * The reference computes pinv with a host SVD; for the non-singular A of an
 * EM run pinv(A) = A^-1, computed here by Gauss-Jordan elimination with
 * partial pivoting (in the registers of one workgroup per model for
 * n <= 192, blocked by 16-column panels in an L2-resident copy above), and
 * log|det W| = -sum log|pivot| (the reference multiplies the diagonal of the
 * LU of W, which can overflow).
 */
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "common.h"
#include "engine_internal.h"
#include "linalg_dev.h"

namespace amica {

namespace {

constexpr int kBlk = 256;
constexpr int kInvMax = 1024;     /* largest n of the inverse and the M-step */

/* canonical statistics offsets within one model block */
struct COff {
    long usum, kn, ln, mn, md, bn, bd, rn, s2, cnew, phi;
    __device__ COff(int n, int m) {
        long nm = (long)n * m;
        usum = 1; kn = usum + nm; ln = kn + nm; mn = ln + nm; md = mn + nm;
        bn = md + nm; bd = bn + nm; rn = bd + nm; s2 = rn + nm; cnew = s2 + n; phi = cnew + n;
    }
};

__device__ double block_sum(double v, double* sh) {
    v = wave_sum_f64(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) r += sh[i];
    __syncthreads();
    return r;
}

/* ------------------------------------------------------------------------ */
/* reductions                                                                 */
/* ------------------------------------------------------------------------ */

/* out[y][e] = sum of part[p][e] over the partials p of group y = blockIdx.y
 * ([y PG, (y + 1) PG)).  A workgroup covers 64 elements (lanes); its
 * kSumWaves waves take the group's partials p = w, w + kSumWaves, ... and
 * their totals are added in wave order, so the result does not depend on
 * timing. */
constexpr int kSumWaves = 16;

__global__ __launch_bounds__(64 * kSumWaves) void sum_parts_kernel(const double* __restrict__ part,
                                                                    int nparts, size_t pstride,
                                                                    size_t len, int PG,
                                                                    double* __restrict__ out) {
    __shared__ double acc[kSumWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t e = (size_t)blockIdx.x * 64 + lane;
    const int p0 = (int)blockIdx.y * PG, p1 = min(nparts, p0 + PG);
    double s = 0.0;
    if (e < len) {
#pragma unroll 4
        for (int p = p0 + w; p < p1; p += kSumWaves) s += part[(size_t)p * pstride + e];
    }
    acc[w][lane] = s;
    __syncthreads();
    if (w == 0 && e < len) {
        double t = 0.0;
        for (int q = 0; q < kSumWaves; ++q) t += acc[q][lane];
        out[(size_t)blockIdx.y * len + e] = t;
    }
}

/* Canonical statistics of model blockIdx.x (layout of the oracle, see the
 * engine header), by the workgroups of blockIdx.y = 0; blockIdx.y is the
 * 16 x 16 tile of phi. */
__global__ __launch_bounds__(kBlk) void finalize_kernel(const double* __restrict__ raw,
                                                        const double* __restrict__ gx,
                                                        MstepArgs a) {
    const int h = blockIdx.x;
    const int n = a.n, m = a.m, NP = a.NP;
    const int NS = kStatsPerJ * m + kStatsPerI;
    const COff o(n, m);
    double* st = a.stats + (size_t)h * a.stride;
    /* raw statistics are [M][NP][NS]: stat s of (h, i) at (h NP + i) NS + s */
    const double* rh = raw + (size_t)h * NP * NS;
    const long nm = (long)n * m;
    if (blockIdx.y == 0) {
        for (long q = threadIdx.x; q < nm; q += blockDim.x) {
            const int j = (int)((unsigned)q / (unsigned)n), i = (int)((unsigned)q % (unsigned)n);
            const double be = a.beta[(size_t)h * nm + q], r = a.rho[(size_t)h * nm + q];
            const double sb = sqrt(be);
            const double* rj = rh + (size_t)i * NS + j * kStatsPerJ;
            const double U = rj[kRawU], Us = rj[kRawUs], Us2 = rj[kRawUs2];
            const double Um2 = rj[kRawUm2], Ur = rj[kRawUr], Ur2 = rj[kRawUr2];
            const double Url = rj[kRawUrl];
            st[o.usum + q] = U;
            st[o.kn + q] = be * r * r * Us2;
            st[o.ln + q] = r * r * Ur2 - 2.0 * r * Ur + U;
            st[o.mn + q] = r * Us;
            st[o.md + q] = (r <= 2.0) ? sb * r * Um2 : sb * r * r * Us2;
            st[o.bn + q] = (r <= 2.0) ? U : 0.0;
            st[o.bd + q] = r * Ur;
            st[o.rn + q] = r * AMICA_LN2 * Url;
        }
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const double* ri = rh + (size_t)i * NS + kStatsPerJ * m;
            st[o.s2 + i] = ri[kRawS2];
            st[o.cnew + i] = ri[kRawCnew];
        }
        if (threadIdx.x == 0) {
            const size_t tail = (size_t)a.M * NS * NP;
            st[0] = raw[tail + h];
            if (h == 0) {
                double c0 = a.lconst[0];
                for (int hh = 1; hh < a.M; ++hh) c0 = fmax(c0, a.lconst[hh]);
                a.stats[(size_t)a.M * a.stride] = raw[tail + a.M] + (double)a.T * c0;
                a.work[0] = c0;   /* offset of the ltall record */
            }
        }
    }
    /* phi_h[i][k] = sum_l G[h*NP+i][l] W~_h[k][l] - gsum_h[i] wc_h[k]: one
     * 16 x 16 (i, k) tile per workgroup (blockIdx.y), the l sum through
     * 16-wide LDS tiles of G and W~ (W~ joined to FP64 as it is staged: the
     * unmixing GEMM's W_hi + W_lo, fp32(W) or W), one FMA per l in
     * increasing l */
    const double* gsum_base = rh + kStatsPerJ * m + kRawGsum;
    const int ntk = (n + 15) / 16;
    const int tib = (int)blockIdx.y / ntk, tkb = (int)blockIdx.y - tib * ntk;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    __shared__ double Gs[16][17];                  /* [i][l] */
    __shared__ double Ws[16][17];                  /* [k][l] */
    double acc = 0.0;
    for (int l0 = 0; l0 < n; l0 += 16) {
        const int l = l0 + tx, gi = tib * 16 + ty, wk = tkb * 16 + ty;
        Gs[ty][tx] = (gi < n && l < n) ? gx[((size_t)h * NP + gi) * NP + l] : 0.0;
        double wv = 0.0;
        if (wk < n && l < n) {
            const size_t wi = ((size_t)h * NP + wk) * NP + l;
            if (a.prec_bf16) wv = (double)bf16_bits_to_f32(a.Whi[wi]) + (double)bf16_bits_to_f32(a.Wlo[wi]);
            else if (a.Wg32) wv = (double)a.Wg32[wi];
            else wv = a.Wg64[wi];
        }
        Ws[ty][tx] = wv;
        __syncthreads();
        const int lend = min(16, n - l0);
        for (int q = 0; q < lend; ++q) acc += Gs[ty][q] * Ws[tx][q];
        __syncthreads();
    }
    const int i = tib * 16 + ty, k = tkb * 16 + tx;
    if (i < n && k < n) {
        const bool wcf = a.prec_bf16 || a.Wg32;       /* W~ c held in FP32 */
        const double wck = wcf ? (double)a.wc[(size_t)h * NP + k] : a.wc64[(size_t)h * NP + k];
        st[o.phi + i + (long)k * n] = acc - gsum_base[(size_t)i * NS] * wck;
    }
}

/* ------------------------------------------------------------------------ */
/* M-step                                                                     */
/* ------------------------------------------------------------------------ */

/* Per model: gm, alpha, sigma2, cnew, phi/vsum, kappa, lambda, mu, beta,
 * rho, the Newton test and the matrix B.  (Flat indices below n n <= 2^20
 * are split with 32-bit division, here and in the kernels below.) */
__global__ __launch_bounds__(kBlk) void mstep_a_kernel(MstepArgs a) {
    __shared__ double sh[16];
    __shared__ double ks[kInvMax];          /* kappa_i sigma2_i */
    const int h = blockIdx.x, n = a.n, m = a.m;
    const long nm = (long)n * m;
    const COff o(n, m);
    const double* st = a.stats + (size_t)h * a.stride;
    const int k_it = a.ds->iter + 1;
    const double vsum = st[0];
    double* alpha = a.alpha + (size_t)h * nm;
    double* mu = a.mu + (size_t)h * nm;
    double* beta = a.beta + (size_t)h * nm;
    double* rho = a.rho + (size_t)h * nm;
    double* sig = a.sigma2 + (size_t)h * n;
    double* cnw = a.cnew + (size_t)h * n;
    double* kap = a.kappa + (size_t)h * n;
    double* lam = a.lambda + (size_t)h * n;

    if (threadIdx.x == 0 && a.update_gm) a.gm[h] = vsum / (double)a.T;
    if (a.update_alpha)
        for (long q = threadIdx.x; q < nm; q += blockDim.x) alpha[q] = st[o.usum + q] / vsum;
    __syncthreads();
    /* updatekappalambda with the new alpha and the old mu */
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double sg = st[o.s2 + i] / vsum;
        sig[i] = sg;
        cnw[i] = st[o.cnew + i] / vsum;
        double kk = 0.0, ll = 0.0;
        for (int j = 0; j < m; ++j) {
            const long q = (long)j * n + i;
            const double us = st[o.usum + q];
            const double dk = st[o.kn + q] / us;
            const double al = alpha[q];
            const double mv = mu[q];
            kk += dk * al;
            ll += (mv * mv * dk + st[o.ln + q] / us) * al;
        }
        kap[i] = kk;
        lam[i] = ll;
        ks[i] = kk * sg;
    }
    __syncthreads();
    for (long q = threadIdx.x; q < nm; q += blockDim.x) {
        if (a.update_mu) mu[q] += st[o.mn + q] / st[o.md + q];
        if (a.update_beta) beta[q] *= st[o.bn + q] / st[o.bd + q];
        if (a.update_rho && k_it >= a.rho_start_iter) {
            double r = rho[q];
            const double ps = psi_dev(1.0 + 1.0 / r);
            r = r + a.rholrate * (1.0 - (r / ps) * st[o.rn + q] / st[o.usum + q]);
            if (r > a.rhomax) r = a.rhomax;
            else if (r < a.rhomin) r = a.rhomin;
            rho[q] = r;
        }
    }
    /* getbflag: D_ik = kappa_i sigma2_i kappa_k sigma2_k <= 1, diagonal
     * included, counted over all (i, k) from the block's copy of ks; part of
     * the A update, so neither it nor a Newton step happens without one */
    int bflag = 0, newt = 0;
    if (a.update_A) {
        double cnt = 0.0;
        for (long q = threadIdx.x; q < (long)n * n; q += blockDim.x) {
            const int i = (int)((unsigned)q % (unsigned)n), k = (int)((unsigned)q / (unsigned)n);
            if (ks[i] * ks[k] <= 1.0) cnt += 1.0;
        }
        bflag = (int)block_sum(cnt, sh);
        newt = (bflag == 0 && a.do_newton == 1 && k_it >= a.newt_start_iter) ? 1 : 0;
    }
    if (threadIdx.x == 0) {
        a.bflag[h] = bflag;
        a.newton[h] = newt;
        if (k_it - 1 < a.log_cap) {
            a.log_bflag[(size_t)(k_it - 1) * a.M + h] = bflag;
            a.log_newton[(size_t)(k_it - 1) * a.M + h] = newt;
            a.log_vsum[(size_t)(k_it - 1) * a.M + h] = vsum;
        }
    }
}

/* phi_h = Phi_h / vsum_h and the direction B_h, one thread per (i, k) over
 * grid (., M): Newton's B (updateBDiagonal / getB) where model h takes a
 * Newton step, I - phi (geteyemphi) otherwise. */
__global__ __launch_bounds__(256) void mstep_bmat_kernel(MstepArgs a) {
    const int h = blockIdx.y, n = a.n;
    const long nn = (long)n * n;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nn) return;
    const int i = (int)((unsigned)q % (unsigned)n), k = (int)((unsigned)q / (unsigned)n);
    const COff o(n, a.m);
    const double* st = a.stats + (size_t)h * a.stride;
    const double vsum = st[0];
    const double ph = st[o.phi + q] / vsum;
    a.phis[(size_t)h * nn + q] = ph;
    double bv;
    if (a.newton[h]) {
        const double* sig = a.sigma2 + (size_t)h * n;
        const double* kap = a.kappa + (size_t)h * n;
        if (i == k) {
            bv = (1.0 - ph) / a.lambda[(size_t)h * n + i];          /* updateBDiagonal */
        } else {
            const double D = (kap[i] * sig[i]) * (kap[k] * sig[k]);
            const double nom = -sig[i] * kap[k];
            const double pt = st[o.phi + k + (long)i * n] / vsum;
            bv = (nom * -ph + -pt) / (D - 1.0);                    /* getB */
        }
    } else {
        bv = (i == k ? 1.0 : 0.0) - ph;                             /* geteyemphi */
    }
    a.B[(size_t)h * nn + q] = bv;
}

/* The scalar part: LL, the decrease rule and the lrate schedule. */
__global__ void mstep_b_kernel(MstepArgs a) {
    if (threadIdx.x != 0) return;
    DevState* ds = a.ds;
    const int k_it = ds->iter + 1;
    const double LL = a.stats[(size_t)a.M * a.stride] / ((double)a.n * (double)a.T);
    int lldec = 0;
    if (k_it - 1 < a.log_cap) {
        a.log_ll[k_it - 1] = LL;
        a.log_lrate[k_it - 1] = ds->lrate;
    }
    if (k_it > 1 && LL < ds->ll_prev) {
        lldec = 1;
        ds->lrate = ds->lrate * a.lratefact;
        ds->numdec += 1;
        if (ds->numdec > a.maxdec) {
            ds->lratemax_local *= a.lratefact;
            ds->numdec = 0;
        }
    }
    ds->ll_prev = LL;
    if (k_it - 1 < a.log_cap) a.log_lldec[k_it - 1] = lldec;
    if (a.update_A) {
        int no_newt = 0;
        for (int h = 0; h < a.M; ++h) if (a.bflag[h] > 0) no_newt = 1;
        ds->no_newt = no_newt;
        const double lr = ds->lrate;
        if (k_it >= a.newt_start_iter && a.do_newton && no_newt == 0)
            ds->lrate = fmin(a.lratemax, lr + fmin(0.1, lr));
        else
            ds->lrate = fmin(a.lnatrate, lr + fmin(0.1, lr));
    }
    ds->iter = k_it;
}

/* AB_h = A_h B_h for every model (column-major n x n, FP64): 16 x 16
 * output tiles per workgroup over grid (n/16, n/16, M), the l sum through
 * 16-wide LDS tiles of A and B, one FMA per l in increasing l. */
constexpr int kAbTile = 16;

__global__ __launch_bounds__(kAbTile * kAbTile) void ab_kernel(MstepArgs a) {
    __shared__ double As[kAbTile][kAbTile + 1];   /* [l][i] */
    __shared__ double Bs[kAbTile][kAbTile + 1];   /* [k][l] */
    const int h = blockIdx.z, n = a.n;
    const long nn = (long)n * n;
    const int ti = threadIdx.x % kAbTile, tk = threadIdx.x / kAbTile;
    const int i = blockIdx.x * kAbTile + ti, k = blockIdx.y * kAbTile + tk;
    const double* A = a.A + (size_t)h * nn;
    const double* B = a.B + (size_t)h * nn;
    double s = 0.0;
    for (int l0 = 0; l0 < n; l0 += kAbTile) {
        /* A tile: lanes along i (consecutive in memory); B tile: along l */
        const int ai = blockIdx.x * kAbTile + ti, al = l0 + tk;
        As[tk][ti] = (ai < n && al < n) ? A[ai + (long)al * n] : 0.0;
        const int bl = l0 + ti, bk = blockIdx.y * kAbTile + tk;
        Bs[tk][ti] = (bl < n && bk < n) ? B[bl + (long)bk * n] : 0.0;
        __syncthreads();
        const int lend = min(kAbTile, n - l0);
        for (int l = 0; l < lend; ++l) s += As[l][ti] * Bs[tk][l];
        __syncthreads();
    }
    if (i < n && k < n) a.work[8 + (size_t)h * nn + i + (long)k * n] = s;
}

/* A += +-lrate A B and the reparameterisation, one wave per column k of
 * model blockIdx.y (source k): normalizeAsave divides column k by its norm
 * nu, normalizemubeta scales mu_jk by nu and beta_jk by 1/nu^2, and the mu
 * shift subtracts (W_old (cnew - c))_k.  c itself is replaced by cnew in
 * copy_c_kernel afterwards, once every column has read the old c. */
constexpr int kColWaves = 4;
constexpr int kColRegs = kInvMax / 64;

__global__ __launch_bounds__(64 * kColWaves) void mstep_col_kernel(MstepArgs a) {
    const int h = blockIdx.y, n = a.n, m = a.m;
    const int lane = threadIdx.x & 63;
    const int k = blockIdx.x * kColWaves + (threadIdx.x >> 6);
    if (!a.update_A || k >= n) return;
    const long nm = (long)n * m;
    const long nn = (long)n * n;
    double* A = a.A + (size_t)h * nn + (long)k * n;
    const double* AB = a.work + 8 + (size_t)h * nn + (long)k * n;
    const double lr = a.ds->lrate;
    const double f = a.newton[h] ? lr : -lr;
    double col[kColRegs];
    double ss = 0.0;
#pragma unroll
    for (int r = 0; r < kColRegs; ++r) {
        const int i = lane + 64 * r;
        if (i < n) {
            col[r] = A[i] + f * AB[i];
            ss += col[r] * col[r];
        }
    }
    if (!a.do_reparm) {
#pragma unroll
        for (int r = 0; r < kColRegs; ++r) {
            const int i = lane + 64 * r;
            if (i < n) A[i] = col[r];
        }
        return;
    }
    const double nu = sqrt(wave_sum_f64(ss));
#pragma unroll
    for (int r = 0; r < kColRegs; ++r) {
        const int i = lane + 64 * r;
        if (i < n) A[i] = col[r] / nu;
    }
    /* (W_old (cnew - c))_k: row k of W, lanes over l */
    const double* W = a.W + (size_t)h * nn;
    const double* c = a.c + (size_t)h * n;
    const double* cnw = a.cnew + (size_t)h * n;
    double d = 0.0;
    for (int l = lane; l < n; l += 64) d += W[k + (long)l * n] * (cnw[l] - c[l]);
    d = wave_sum_f64(d);
    double* mu = a.mu + (size_t)h * nm;
    double* beta = a.beta + (size_t)h * nm;
    for (int j = lane; j < m; j += 64) {
#pragma clang fp contract(off)
        const double ms = mu[(long)j * n + k] * nu;      /* rounded, then shifted */
        mu[(long)j * n + k] = ms - d;
        beta[(long)j * n + k] = beta[(long)j * n + k] / (nu * nu);
    }
}

/* c <- cnew for every model (grid (., M)) */
__global__ __launch_bounds__(256) void copy_c_kernel(MstepArgs a) {
    const int h = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n) a.c[(size_t)h * a.n + i] = a.cnew[(size_t)h * a.n + i];
}

/* ------------------------------------------------------------------------ */
/* prepare: W = A^-1, ldet, GEMM copies of W, W c, E-step constants          */
/* ------------------------------------------------------------------------ */

/* bf16 hi and lo parts of one W entry, as the split-BF16 unmixing GEMM
 * multiplies by them */
__device__ __forceinline__ void split_w(double w, uint16_t& hi, uint16_t& lo) {
    hi = f32_to_bf16_bits((float)w);
    lo = f32_to_bf16_bits((float)(w - (double)bf16_bits_to_f32(hi)));
}

__device__ __forceinline__ double joined_w(uint16_t hi, uint16_t lo) {
    return (double)bf16_bits_to_f32(hi) + (double)bf16_bits_to_f32(lo);
}

/* After W = A^-1, for all models at once (grid (., M)): the GEMM copies of
 * W (rows h NP + i, columns k) and the per-(h, j, i) E-step constants.
 * Flat index per model: [0, n n) W entries, then the n m constants. */
__global__ __launch_bounds__(256) void prepare_tail_kernel(MstepArgs a) {
    const int h = blockIdx.y, n = a.n, m = a.m, NP = a.NP;
    const long nn = (long)n * n, nm = (long)n * m;
    const double* W = a.W + (size_t)h * nn;
    const long total = nn + nm;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        if (q < nn) {
            const int i = (int)((unsigned)q % (unsigned)n), k = (int)((unsigned)q / (unsigned)n);
            const double wv2 = W[q];
            const size_t dst = ((size_t)h * NP + i) * NP + k;
            if (a.prec_bf16) {
                /* FP8 storage: the GEMM multiplies q = x / s, so it gets W diag(s) */
                uint16_t hi, lo;
                split_w(a.xscale ? wv2 * (double)a.xscale[k] : wv2, hi, lo);
                a.Whi[dst] = hi;
                a.Wlo[dst] = lo;
            } else if (a.Wg32) {
                a.Wg32[dst] = (float)wv2;
            } else {
                a.Wg64[dst] = wv2;
            }
        } else {
            const long e = q - nn;
            const int j = (int)((unsigned)e / (unsigned)n), i = (int)((unsigned)e % (unsigned)n);
            const double al = a.alpha[(size_t)h * nm + e], mv = a.mu[(size_t)h * nm + e];
            const double be = a.beta[(size_t)h * nm + e], r = a.rho[(size_t)h * nm + e];
            const double sb = sqrt(be);
            const double cq = log(al) + 0.5 * log(be) - AMICA_LOG2_REF - lgamma(1.0 + 1.0 / r);
            float* cj = a.cst + ((size_t)h * kNumCst * m + j) * NP + i;
            const size_t ks = (size_t)m * NP;
            cj[kCstSb * ks] = (float)sb;
            cj[kCstNsbmu * ks] = (float)(-sb * mv);
            cj[kCstRm2 * ks] = (float)(r - 2.0);
            cj[kCstCq2 * ks] = (float)(cq * 1.4426950408889634);
            cj[kCstRsb * ks] = (float)(r * sb);
            if (a.cst64) {
                double* dj = a.cst64 + ((size_t)h * kNumCst64 * m + j) * NP + i;
                dj[kCst64Sb * ks] = sb;
                dj[kCst64Nsbmu * ks] = -sb * mv;
            }
        }
    }
}

/* W~ c_h, one wave per row i over grid (n/4, M), lanes over k: the W~ the
 * unmixing GEMM applies (W~ = split(W diag(s)) diag(1/s) on the BF16 and FP8
 * paths, fp32(W) on the plain FP32 path, W itself on the FP32 path). */
__global__ __launch_bounds__(256) void wc_kernel(MstepArgs a) {
    const int h = blockIdx.y, n = a.n, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double* W = a.W + (size_t)h * n * n;
    const double* c = a.c + (size_t)h * n;
    double s = 0.0;
    for (int k = lane; k < n; k += 64) {
        double wr = W[i + (long)k * n];
        if (a.prec_bf16) {
            const double sk = a.xscale ? (double)a.xscale[k] : 1.0;
            uint16_t hi, lo;
            split_w(wr * sk, hi, lo);
            wr = joined_w(hi, lo) / sk;
        } else if (a.Wg32) {
            wr = (double)(float)wr;
        }
        s += wr * c[k];
    }
    s = wave_sum_f64(s);
    if (lane == 0) {
        if (a.prec_bf16 || a.Wg32) a.wc[(size_t)h * a.NP + i] = (float)s;
        else a.wc64[(size_t)h * a.NP + i] = s;
    }
}

__device__ __forceinline__ void set_ldet(const MstepArgs& a, int h, double lad) {
    const double ldet = -lad;          /* log|det W| = -log|det A| */
    a.ldet[h] = ldet;
    if (a.lconst) a.lconst[h] = log(a.gm[h]) + ldet + a.ds->ldetS;   /* null: the inverse alone */
}

/* W = A^-1 by blocked Gauss-Jordan with implicit partial pivoting.
 *
 * Rows are never moved: step k takes the pivot p_k with the largest |a_ik|
 * among the rows not pivoted yet (ties to the smallest index) and eliminates
 * in place,
 *   a[p][j] <- a[p][j] / d (j != k),   a[p][k] <- 1 / d,
 *   a[i][j] <- a[i][j] - a[i][k] a[p][j] / d,   a[i][k] <- -a[i][k] / d,
 * with d = a[p][k], so that afterwards a[p_r][j] = (A^-1)[r][p_j]; W is
 * gathered from that and log|det A| = sum_k log|d_k|.
 *
 * Blocked over panels of NB columns K = [k0, k0 + NB) with pivot rows P:
 * NB steps restricted to the panel turn its columns into A_PK^-1 (rows P)
 * and -A_RK A_PK^-1 (the other rows R), and the remaining columns J then
 * take all NB steps at once,
 *   a'[i][j] = (i in P ? 0 : a[i][j]) + sum_q a'[i][k0 + q] a[p_q][j],
 * which is the Schur complement form of the same elimination.  A panel is
 * factored by one wave with its n x NB values in registers (lane l holds
 * rows l + 64 r), so a step needs no barrier and no LDS; the trailing
 * update runs one wave per column over the whole GPU.  Every panel after the first is factored by block 0 of the launch
 * that applies the previous panel, right after it has updated the new
 * panel's columns.
 *
 * Scratch per model h: the n x n matrix (column-major) at
 * work + 8 + M n n + h n n, then the pivot bookkeeping after all matrices:
 * step[n] (the step that pivoted row i, -1 while unused) and piv[n] (the
 * pivot row of step k) as ints, then |d_k| per step. */

struct GjView {
    double* aw;
    int* step;
    int* piv;
    double* dk;     /* |pivot| of step k */
};

__device__ __forceinline__ GjView gj_view(const MstepArgs& a, int h) {
    const size_t nn = (size_t)a.n * a.n;
    double* const tail = a.work + 8 + 2 * (size_t)a.M * nn;
    GjView g;
    g.aw = a.work + 8 + (size_t)a.M * nn + (size_t)h * nn;
    g.step = (int*)tail + (size_t)h * 2 * a.n;
    g.piv = g.step + a.n;
    g.dk = tail + (size_t)a.M * a.n + (size_t)h * a.n;
    return g;
}

/* 64-bit DPP move: both halves through the same lane pattern */
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

/* 32-bit maximum over the wave, in every lane: row maxima by DPP (quad
 * permutes, then the half-row and row mirrors), then the four rows' */
template <typename I>
__device__ __forceinline__ I wave_max_bcast(I v) {
    auto dmax = [](I x, auto ctrl) {
        const I o = (I)__builtin_amdgcn_mov_dpp((int)x, decltype(ctrl)::value, 0xf, 0xf, false);
        return o > x ? o : x;
    };
    v = dmax(v, std::integral_constant<int, 0xB1>());
    v = dmax(v, std::integral_constant<int, 0x4E>());
    v = dmax(v, std::integral_constant<int, 0x141>());
    v = dmax(v, std::integral_constant<int, 0x140>());
    const I r0 = (I)__builtin_amdgcn_readlane((int)v, 0), r1 = (I)__builtin_amdgcn_readlane((int)v, 16);
    const I r2 = (I)__builtin_amdgcn_readlane((int)v, 32), r3 = (I)__builtin_amdgcn_readlane((int)v, 48);
    const I m01 = r0 > r1 ? r0 : r1, m23 = r2 > r3 ? r2 : r3;
    return m01 > m23 ? m01 : m23;
}

/* Pivot row of the column held as col[r] (rows lane + 64 r): the largest
 * |value| among the unused rows (st < 0; NaN counts as infinite, so that a
 * NaN column still names an unused row), ties to the smallest row index;
 * wave-uniform.  The bit patterns of non-negative doubles order like their
 * values, so the maximum is found on the high words (used rows -1) and,
 * only where several rows share the largest high word, on their low words:
 * 32-bit DPP maxima instead of 64-bit value-and-index compares. */
template <int R>
__device__ __forceinline__ int gj_pivot(const double (&col)[R], const int (&st)[R]) {
    int hi[R];
    unsigned lo[R];
    int mh = -1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double x = fabs(col[r]);
        if (x != x) x = __builtin_inf();
        const uint64_t u = __builtin_bit_cast(uint64_t, x);
        hi[r] = st[r] < 0 ? (int)(u >> 32) : -1;
        lo[r] = (unsigned)u;
        mh = max(mh, hi[r]);
    }
    mh = wave_max_bcast<int>(mh);
    unsigned long long cand[R];
    int nc = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        cand[r] = __ballot(hi[r] == mh);
        nc += __builtin_popcountll(cand[r]);
    }
    if (nc > 1) {                           /* equal high words: compare the low words */
        unsigned ml = 0;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (hi[r] == mh) ml = max(ml, lo[r]);
        ml = wave_max_bcast<unsigned>(ml);
#pragma unroll
        for (int r = 0; r < R; ++r) cand[r] = __ballot(hi[r] == mh && lo[r] == ml);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (cand[r]) return 64 * r + __builtin_ctzll(cand[r]);
    return 0;       /* not reached: an unused row holds the maximum */
}

/* argmax of |v[.][c]| over this wave's unused rows (the pivot of column c) */
template <int R, int NB>
__device__ __forceinline__ int gj_col_argmax(const double (&v)[R][NB], const int (&st)[R], int c) {
    double col[R];
#pragma unroll
    for (int r = 0; r < R; ++r) col[r] = v[r][c];
    return gj_pivot<R>(col, st);
}

/* NB Gauss-Jordan steps on the panel v (columns k0 .., rows lane + 64 r),
 * one wave, no LDS: the pivot row reaches every lane by readlane, and each
 * step updates the next pivot column first and takes its argmax before the
 * other columns, so that the reduction overlaps their FMAs. */
template <int R, int NB>
__device__ __forceinline__ void gj_factor(const GjView& g, int n, int k0, double (&v)[R][NB]) {
    const int lane = threadIdx.x;
    const int nc = min(NB, n - k0);
    int st[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        st[r] = i < n ? g.step[i] : n;          /* padding rows count as used */
    }
    int p = gj_col_argmax<R, NB>(v, st, 0);
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c >= nc) continue;                      /* wave-uniform; keeps the loop unrollable */
        const int k = k0 + c;
        const int pl = p & 63, pr = p >> 6;
        double prow[NB];
#pragma unroll
        for (int c2 = 0; c2 < NB; ++c2) prow[c2] = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r == pr) {                          /* wave-uniform */
#pragma unroll
                for (int c2 = 0; c2 < NB; ++c2) prow[c2] = readlane_f64(v[r][c2], pl);
            }
        const double d = prow[c];
        /* 1/d: hardware estimate and two Newton steps (within an ulp) */
        double inv = __builtin_amdgcn_rcp(d);
        inv = fma(fma(-d, inv, 1.0), inv, inv);
        inv = fma(fma(-d, inv, 1.0), inv, inv);
        double prw[NB];
#pragma unroll
        for (int c2 = 0; c2 < NB; ++c2) prw[c2] = (c2 == c) ? inv : prow[c2] * inv;
        double pc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            pc[r] = v[r][c];
            if (lane + 64 * r == p) st[r] = k;
        }
        /* every row as a non-pivot row (the pivot row is rewritten below):
         * the next pivot column first, then its argmax (the pivot row is
         * marked used), then the other columns */
        int pn = 0;
        if (c + 1 < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r][c + 1] -= pc[r] * prw[c + 1];
            if (c + 1 < nc) pn = gj_col_argmax<R, NB>(v, st, c + 1);
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c2 = 0; c2 < NB; ++c2) {
                if (c2 == c + 1) continue;
                v[r][c2] = ((c2 == c) ? 0.0 : v[r][c2]) - pc[r] * prw[c2];
            }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r == pr && lane == pl) {            /* one lane of one slot */
#pragma unroll
                for (int c2 = 0; c2 < NB; ++c2) v[r][c2] = prw[c2];
            }
        if (lane == 0) {
            g.piv[k] = p;
            g.dk[k] = fabs(d);
        }
        p = pn;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < n) g.step[i] = st[r];
    }
}

template <int R, int NB>
__device__ __forceinline__ void gj_store(const GjView& g, int n, int k0, const double (&v)[R][NB]) {
#pragma unroll
    for (int c2 = 0; c2 < NB; ++c2) {
        if (k0 + c2 >= n) continue;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = threadIdx.x + 64 * r;
            if (i < n) g.aw[i + (size_t)(k0 + c2) * n] = v[r][c2];
        }
    }
}

/* A -> scratch, every row unused.  grid (x, M) */
__global__ __launch_bounds__(256) void gj_init_kernel(MstepArgs a) {
    const int h = blockIdx.y;
    const GjView g = gj_view(a, h);
    const long nn = (long)a.n * a.n;
    const double* A = a.A + (size_t)h * nn;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (long)gridDim.x * blockDim.x)
        g.aw[q] = A[q];
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < a.n; i += blockDim.x) g.step[i] = -1;
    }
}

/* the first panel: grid (M), one wave */
template <int R, int NB>
__global__ __launch_bounds__(64) void gj_panel_kernel(MstepArgs a) {
    const GjView g = gj_view(a, blockIdx.x);
    const int n = a.n;
    double v[R][NB];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = threadIdx.x + 64 * r;
#pragma unroll
        for (int c2 = 0; c2 < NB; ++c2)
            v[r][c2] = (i < n && c2 < n) ? g.aw[i + (size_t)c2 * n] : 0.0;
    }
    gj_factor<R, NB>(g, n, 0, v);
    gj_store<R, NB>(g, n, 0, v);
}

/* Applies panel [kp, kp + NB) to every other column, one wave per column
 * (grid (n + 1, M)).  With kn >= 0, block 0 instead updates the columns of
 * the next panel [kn, kn + NB) in registers and factors them.  Every load a
 * wave needs (its rows of the panel, the pivot-row entries) is issued before
 * the first FMA: the q loop runs over all NB panel columns, those past the
 * matrix (a last, partial panel) with multiplier 0. */
template <int R, int NB>
__global__ __launch_bounds__(64) void gj_update_kernel(MstepArgs a, int kp, int kn) {
    const int h = blockIdx.y, n = a.n, lane = threadIdx.x;
    const GjView g = gj_view(a, h);
    const int ncp = min(NB, n - kp);
    __shared__ int pv[NB];
    __shared__ double on[NB][NB];                   /* block 0: pivot rows x next panel */
    if (lane < NB) pv[lane] = g.piv[kp + min(lane, ncp - 1)];
    __syncthreads();
    bool inP[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        bool f = false;
#pragma unroll
        for (int q = 0; q < NB; ++q) f |= (q < ncp && pv[q] == i);
        inP[r] = f;
    }
    const double* pan = g.aw + (size_t)kp * n;      /* the applied panel's columns */
    if (kn >= 0 && blockIdx.x == 0) {
        /* the panel's rows and the pivot rows of the next panel through LDS */
        __shared__ double ps[NB][R * 64];
        const int ncn = min(NB, n - kn);
        double v[R][NB];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane + 64 * r;
#pragma unroll
            for (int c2 = 0; c2 < NB; ++c2)
                v[r][c2] = (i < n && c2 < ncn && !inP[r]) ? g.aw[i + (size_t)(kn + min(c2, ncn - 1)) * n] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NB; ++q)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = lane + 64 * r;
                ps[q][i] = (i < n && q < ncp) ? pan[i + (size_t)q * n] : 0.0;
            }
        for (int e = lane; e < NB * NB; e += 64) {
            const int q = e / NB, c2 = e - q * NB;
            on[q][c2] = (q < ncp && c2 < ncn) ? g.aw[pv[q] + (size_t)(kn + c2) * n] : 0.0;
        }
        __syncthreads();
#pragma unroll 2
        for (int q = 0; q < NB; ++q) {
            double pq[R];
#pragma unroll
            for (int r = 0; r < R; ++r) pq[r] = ps[q][lane + 64 * r];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int c2 = 0; c2 < NB; ++c2) v[r][c2] += pq[r] * on[q][c2];
        }
        gj_factor<R, NB>(g, n, kn, v);
        gj_store<R, NB>(g, n, kn, v);
        return;
    }
    const int j = (int)blockIdx.x - (kn >= 0 ? 1 : 0);
    if (j < 0 || j >= n || (j >= kp && j < kp + NB) || (kn >= 0 && j >= kn && j < kn + NB)) return;
    /* this lane's rows of the panel; columns past n read column kp */
    double pq[NB][R];
#pragma unroll
    for (int q = 0; q < NB; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane + 64 * r;
            pq[q][r] = i < n ? pan[i + (size_t)(q < ncp ? q : 0) * n] : 0.0;
        }
    double* col = g.aw + (size_t)j * n;
    double o[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) o[q] = q < ncp ? col[pv[q]] : 0.0;
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        acc[r] = (i < n && !inP[r]) ? col[i] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NB; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] += pq[q][r] * o[q];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        if (i < n) col[i] = acc[r];
    }
}

/* W[r][m] = a[piv[r]][pinv[m]] and the log-determinant terms; grid (M),
 * 1024 threads */
__global__ __launch_bounds__(1024) void gj_gather_kernel(MstepArgs a) {
    const int h = blockIdx.x, n = a.n, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nw = blockDim.x >> 6;
    const GjView g = gj_view(a, h);
    __shared__ int piv[kInvMax], pinv[kInvMax];
    __shared__ double lsum[16];
    double l = 0.0;
    for (int r = tid; r < n; r += blockDim.x) {
        piv[r] = g.piv[r];
        l += log(g.dk[r]);
    }
    l = wave_sum_f64(l);
    if (lane == 0) lsum[wv] = l;
    __syncthreads();
    for (int r = tid; r < n; r += blockDim.x) pinv[piv[r]] = r;
    __syncthreads();
    double* W = a.W + (size_t)h * n * n;
    for (int mcol = wv; mcol < n; mcol += nw) {
        const double* src = g.aw + (size_t)pinv[mcol] * n;
        for (int r = lane; r < n; r += 64) W[r + (size_t)mcol * n] = src[piv[r]];
    }
    if (tid == 0) {
        double lad = 0.0;                           /* log|det A| = sum_k log|d_k| */
        for (int w = 0; w < nw; ++w) lad += lsum[w];
        set_ldet(a, h, lad);
    }
}

/* n <= 192: the whole inverse of one model by one workgroup, the matrix in
 * registers.  Wave w holds columns [w C, w C + C), lane l rows l + 64 r.
 * A step's only exchange is the pivot column -- its values, pivot row and
 * 1/d -- which the wave that owns it publishes in LDS (double buffered)
 * before the step's one barrier; each wave reads the pivot row's entries
 * of its own columns by readlane.  The owner of the next pivot column
 * updates that column first, at raised issue priority, and finds its
 * pivot, so that the next step can start as soon as the other waves have
 * finished this one.  Same elimination and pivot rule as the blocked
 * kernels above, except that a pivot row is not divided by its d: it
 * keeps a_pj (and 1 in column k) with multiplier 0 in its own step, and
 * in later steps takes the same update as every row, which holds the row
 * up to the factor 1/d, applied once when W is written -- so a step costs
 * one FMA per element and no per-column work.  W is written through the
 * pivot permutation directly and log|det A| summed from the pivots kept in
 * LDS: one launch in all. */
template <int R, int C>
__global__ __launch_bounds__(1024) void gj_small_kernel(MstepArgs a) {
    const int h = blockIdx.x, n = a.n, lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    __shared__ double pcol[2][R * 64];      /* pivot column before its step */
    __shared__ double pinv[2];              /* 1 / d */
    __shared__ int prow[2];                 /* pivot row */
    __shared__ int pivk[R * 64];            /* pivot row of step k */
    __shared__ double dk[R * 64];           /* |d| of step k */
    __shared__ double invk[R * 64];         /* 1 / d of step k */
    const double* A = a.A + (size_t)h * n * n;
    double v[R][C];
    int st[R];                              /* step that pivoted row lane + 64 r, -1 before */
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        st[r] = i < n ? -1 : n;             /* padding rows count as used */
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int j = w * C + c;
            v[r][c] = (i < n && j < n) ? A[i + (size_t)j * n] : 0.0;
        }
    }
    /* pivot of column k, this wave's local column cl, published in buffer b */
    auto publish = [&](int cl, int k, int b) {
        double col[R];
#pragma unroll
        for (int r = 0; r < R; ++r) col[r] = 0.0;
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (c == cl) {                  /* wave-uniform */
#pragma unroll
                for (int r = 0; r < R; ++r) col[r] = v[r][c];
            }
        const int p = gj_pivot<R>(col, st);
        double d = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (r == (p >> 6)) d = readlane_f64(col[r], p & 63);
        /* 1/d: hardware estimate and two Newton steps (within an ulp) */
        double inv = __builtin_amdgcn_rcp(d);
        inv = fma(fma(-d, inv, 1.0), inv, inv);
        inv = fma(fma(-d, inv, 1.0), inv, inv);
#pragma unroll
        for (int r = 0; r < R; ++r) pcol[b][lane + 64 * r] = col[r];
        if (lane == 0) {
            prow[b] = p;
            pinv[b] = inv;
            pivk[k] = p;
            dk[k] = fabs(d);
            invk[k] = inv;
        }
    };
    if (w == 0) publish(0, 0, 0);
    __syncthreads();
    /* C steps per pass of the outer loop, so that the local columns of a
     * step and of the next one are compile-time indices: wave ow holds
     * columns k0 .. k0 + C - 1 */
    for (int k0 = 0; k0 < n; k0 += C) {
        const int ow = k0 / C;
#pragma unroll
        for (int cc = 0; cc < C; ++cc) {
            const int k = k0 + cc;
            if (k >= n) break;
            const int b = k & 1;
            const int p = __builtin_amdgcn_readfirstlane(prow[b]);
            const double inv = pinv[b];
            const int pl = p & 63, pr = p >> 6;
            /* the multipliers a_ik / d, 0 in the pivot row, which keeps its
             * values: row p stands for a_pj / d until W is written */
            double l[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const bool piv_lane = lane + 64 * r == p;
                if (piv_lane) st[r] = k;
                l[r] = piv_lane ? 0.0 : pcol[b][lane + 64 * r] * inv;
            }
            const bool own = w == ow;           /* column k is this wave's column cc */
            /* column k + 1: local column cl1 of wave ow or ow + 1 (never
             * column k of the same wave).  Its owner updates it first, needing
             * only its pivot-row entry, and publishes its pivot. */
            const int cl1 = cc + 1 < C ? cc + 1 : 0;
            const bool next = k + 1 < n && w == (cc + 1 < C ? ow : ow + 1);
            if (next) {
                __builtin_amdgcn_s_setprio(3);
                double pv = 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (r == pr) pv = readlane_f64(v[r][cl1], pl);
#pragma unroll
                for (int r = 0; r < R; ++r) v[r][cl1] = fma(-l[r], pv, v[r][cl1]);
                publish(cl1, k + 1, b ^ 1);
                __builtin_amdgcn_s_setprio(0);
            }
            /* the pivot row's entries of this wave's other columns (uniform:
             * scalar registers) */
            double prv[C];
#pragma unroll
            for (int c = 0; c < C; ++c) prv[c] = 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (r == pr) {
#pragma unroll
                    for (int c = 0; c < C; ++c) prv[c] = readlane_f64(v[r][c], pl);
                }
            /* a_ij - l_i a_pj, and -l_i in column k, whose pivot-row entry
             * becomes 1 (standing for 1 / d) */
#pragma unroll
            for (int c = 0; c < C; ++c) {
                if (c == cl1 && next) continue;
                if (c == cc && own) {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[r][c] = -l[r];
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (r == pr && lane == pl) v[r][c] = 1.0;
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[r][c] = fma(-l[r], prv[c], v[r][c]);
                }
            }
            __syncthreads();
        }
    }
    /* W[r][m] = a[p_r][pinv[m]]: element a[i][j] (row i scaled by the 1 / d
     * of its step) goes to W[step of row i][p_j] */
    double* W = a.W + (size_t)h * n * n;
    double sc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) sc[r] = lane + 64 * r < n ? invk[st[r]] : 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int j = w * C + c;
        if (j >= n) continue;               /* wave-uniform */
        const size_t m = (size_t)pivk[j] * n;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (lane + 64 * r < n) W[m + st[r]] = v[r][c] * sc[r];
    }
    if (w == 0) {
        double l = 0.0;                     /* log|det A| = sum_k log|d_k| */
        for (int k = lane; k < n; k += 64) l += log(dk[k]);
        l = wave_sum_f64(l);
        if (lane == 0) set_ldet(a, h, l);
    }
}

template <int R, int C>
static void launch_gj_small(const MstepArgs& ma, hipStream_t st) {
    const int nw = (ma.n + C - 1) / C;      /* <= 16 waves */
    hipLaunchKernelGGL((gj_small_kernel<R, C>), dim3(ma.M), dim3(64 * nw), 0, st, ma);
}

template <int R, int NB>
static void launch_gj(const MstepArgs& ma, hipStream_t st) {
    const int n = ma.n;
    const long nn = (long)n * n;
    const unsigned gi = (unsigned)std::min<long>((nn + 255) / 256, 64);
    hipLaunchKernelGGL(gj_init_kernel, dim3(gi, ma.M), dim3(256), 0, st, ma);
    hipLaunchKernelGGL((gj_panel_kernel<R, NB>), dim3(ma.M), dim3(64), 0, st, ma);
    const int npan = (n + NB - 1) / NB;
    for (int b = 1; b < npan; ++b)
        hipLaunchKernelGGL((gj_update_kernel<R, NB>), dim3(n + 1, ma.M), dim3(64), 0, st, ma,
                           (b - 1) * NB, b * NB);
    if (npan > 1)
        hipLaunchKernelGGL((gj_update_kernel<R, NB>), dim3(n, ma.M), dim3(64), 0, st, ma,
                           (npan - 1) * NB, -1);
    hipLaunchKernelGGL(gj_gather_kernel, dim3(ma.M), dim3(1024), 0, st, ma);
}

/* W_h = A_h^-1 and log|det W_h| of every model: one workgroup per model,
 * the matrix in registers, up to 192 channels; the blocked kernels beyond
 * (registers per lane of a panel: R rows x NB columns <= 64 doubles) */
static void launch_inverse(const MstepArgs& ma, hipStream_t st) {
    const int n = ma.n;
    if (n <= 64) launch_gj_small<1, 4>(ma, st);
    else if (n <= 128) launch_gj_small<2, 8>(ma, st);
    else if (n <= 192) launch_gj_small<3, 12>(ma, st);
    else if (n <= 256) launch_gj<4, 16>(ma, st);
    else if (n <= 512) launch_gj<8, 8>(ma, st);
    else launch_gj<16, 4>(ma, st);
}

}  // namespace

/* groups of the first pass: about 4 workgroups per CU in all, at least 32
 * partials per group, at most kSumGroups */
static int sum_groups(int nparts, size_t len) {
    const long xb = (long)((len + 63) / 64);
    long g = (1024 + xb - 1) / xb;
    g = std::min<long>(g, std::max(1, nparts / 32));
    return (int)std::max(1L, std::min<long>(g, kSumGroups));
}

/* out = sum over nparts partials: one pass, or a pass over groups into tmp
 * and a second pass over the groups */
static void sum_parts(const double* part, int nparts, size_t pstride, size_t len, double* tmp,
                      double* out, hipStream_t st) {
    const dim3 blk(64 * kSumWaves);
    const unsigned xb = (unsigned)((len + 63) / 64);
    const int G = sum_groups(nparts, len);
    if (G == 1) {
        hipLaunchKernelGGL(sum_parts_kernel, dim3(xb, 1), blk, 0, st, part, nparts, pstride, len, nparts, out);
        return;
    }
    const int PG = (nparts + G - 1) / G;
    hipLaunchKernelGGL(sum_parts_kernel, dim3(xb, G), blk, 0, st, part, nparts, pstride, len, PG, tmp);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(xb, 1), blk, 0, st, (const double*)tmp, G, len, len, G, out);
}

int launch_reduce(const double* part, int nparts, size_t pstride, const double* gxpart,
                  int nchunks, double* raw, double* gxtot, double* tmp, const MstepArgs& ma,
                  hipStream_t st) {
    const int NS = kStatsPerJ * ma.m + kStatsPerI;
    const size_t plen = (size_t)ma.M * NS * ma.NP + ma.M + 1;
    sum_parts(part, nparts, pstride, plen, tmp, raw, st);
    const size_t glen = (size_t)ma.ldb * ma.NP;
    sum_parts(gxpart, nchunks, glen, glen, tmp, gxtot, st);
    /* one 16 x 16 tile of phi per workgroup (the statistics by the first) */
    const unsigned ntk = (unsigned)((ma.n + 15) / 16);
    hipLaunchKernelGGL(finalize_kernel, dim3(ma.M, ntk * ntk), dim3(kBlk), 0, st, raw, gxtot, ma);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mstep(const MstepArgs& ma, hipStream_t st) {
    const int n = ma.n;
    if (n < 1 || n > kInvMax) return -1;
    const long nn = (long)n * n;
    hipLaunchKernelGGL(mstep_a_kernel, dim3(ma.M), dim3(kBlk), 0, st, ma);
    hipLaunchKernelGGL(mstep_b_kernel, dim3(1), dim3(64), 0, st, ma);
    if (ma.update_A) {
        hipLaunchKernelGGL(mstep_bmat_kernel, dim3((unsigned)((nn + 255) / 256), ma.M), dim3(256), 0,
                           st, ma);
        const unsigned nt = (unsigned)((n + kAbTile - 1) / kAbTile);
        hipLaunchKernelGGL(ab_kernel, dim3(nt, nt, ma.M), dim3(kAbTile * kAbTile), 0, st, ma);
        hipLaunchKernelGGL(mstep_col_kernel, dim3((unsigned)((n + kColWaves - 1) / kColWaves), ma.M),
                           dim3(64 * kColWaves), 0, st, ma);
    }
    hipLaunchKernelGGL(copy_c_kernel, dim3((unsigned)((n + 255) / 256), ma.M), dim3(256), 0, st, ma);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mstep_configure(int n) { return (n >= 1 && n <= kInvMax) ? 0 : -1; }

/* scratch doubles of the inverse after the first 8 + M n n of work */
size_t inverse_work_doubles(int n, int M) {
    return (size_t)M * n * n + 2 * (size_t)M * n;
}

int launch_gj_inverse(const double* A, double* Ainv, double* ldet_inv, int n, int M, double* work,
                      hipStream_t st) {
    if (n < 1 || n > kInvMax || M < 1) return -1;
    MstepArgs ma;
    std::memset(&ma, 0, sizeof(ma));
    ma.n = n;
    ma.M = M;
    ma.A = const_cast<double*>(A);
    ma.W = Ainv;
    ma.ldet = ldet_inv;
    ma.work = work;
    launch_inverse(ma, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_prepare(const MstepArgs& ma, hipStream_t st) {
    const int n = ma.n;
    if (n < 1 || n > kInvMax) return -1;
    launch_inverse(ma, st);
    const long total = (long)ma.n * ma.n + (long)ma.n * ma.m;
    const unsigned gx = (unsigned)std::min<long>((total + 255) / 256, 256);
    hipLaunchKernelGGL(prepare_tail_kernel, dim3(gx, ma.M), dim3(256), 0, st, ma);
    hipLaunchKernelGGL(wc_kernel, dim3((unsigned)((n + 3) / 4), ma.M), dim3(256), 0, st, ma);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace amica
