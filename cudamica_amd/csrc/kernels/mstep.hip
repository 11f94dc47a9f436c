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
 * partial pivoting (in LDS for n <= 128, in an L2-resident copy above), and
 * log|det W| = -sum log|pivot| (the reference multiplies the diagonal of the
 * LU of W, which can overflow).
 */
#include <algorithm>

#include "common.h"
#include "engine_internal.h"
#include "linalg_dev.h"

namespace amica {

namespace {

constexpr int kBlk = 256;

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

/* out[e] = sum_p part[p][e].  A workgroup covers 64 elements (lanes); its
 * kSumWaves waves take the partials p = w, w + kSumWaves, ... and their
 * totals are added in wave order, so the result does not depend on timing. */
constexpr int kSumWaves = 16;

__global__ __launch_bounds__(64 * kSumWaves) void sum_parts_kernel(const double* __restrict__ part,
                                                                    int nparts, size_t pstride,
                                                                    size_t len, double* __restrict__ out) {
    __shared__ double acc[kSumWaves][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t e = (size_t)blockIdx.x * 64 + lane;
    double s = 0.0;
    if (e < len)
        for (int p = w; p < nparts; p += kSumWaves) s += part[(size_t)p * pstride + e];
    acc[w][lane] = s;
    __syncthreads();
    if (w == 0 && e < len) {
        double t = 0.0;
        for (int q = 0; q < kSumWaves; ++q) t += acc[q][lane];
        out[e] = t;
    }
}

/* Canonical statistics of model blockIdx.x (layout of the oracle, see the
 * engine header); blockIdx.y splits the phi rows. */
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
            const int j = (int)(q / n), i = (int)(q % n);
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
    /* phi_h[i][k] = sum_l G[h*NP+i][l] W~_h[k][l] - gsum_h[i] wc_h[k] */
    const double* gsum_base = rh + kStatsPerJ * m + kRawGsum;
    const int rows_per = (n + gridDim.y - 1) / gridDim.y;
    const int ib = blockIdx.y * rows_per, ie = min(n, ib + rows_per);
    const long cnt = (long)(ie - ib) * n;
    for (long q = threadIdx.x; q < cnt; q += blockDim.x) {
        const int i = ib + (int)(q % (ie - ib)), k = (int)(q / (ie - ib));
        const double* grow = gx + ((size_t)h * NP + i) * NP;
        const size_t wrow = ((size_t)h * NP + k) * NP;
        double s = 0.0, wck;
        if (a.prec_bf16) {
            /* the unmixing GEMM used W~ = W_hi + W_lo */
            for (int l = 0; l < n; ++l)
                s += grow[l] * ((double)bf16_bits_to_f32(a.Whi[wrow + l]) +
                                (double)bf16_bits_to_f32(a.Wlo[wrow + l]));
            wck = (double)a.wc[(size_t)h * NP + k];
        } else {
            for (int l = 0; l < n; ++l) s += grow[l] * a.Wg64[wrow + l];
            wck = a.wc64[(size_t)h * NP + k];
        }
        s -= gsum_base[(size_t)i * NS] * wck;
        st[o.phi + i + (long)k * n] = s;
    }
}

/* ------------------------------------------------------------------------ */
/* M-step                                                                     */
/* ------------------------------------------------------------------------ */

/* Per model: gm, alpha, sigma2, cnew, phi/vsum, kappa, lambda, mu, beta,
 * rho, the Newton test and the matrix B. */
__global__ __launch_bounds__(kBlk) void mstep_a_kernel(MstepArgs a) {
    __shared__ double sh[16];
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
    double* phs = a.phis + (size_t)h * n * n;
    double* kap = a.kappa + (size_t)h * n;
    double* lam = a.lambda + (size_t)h * n;
    double* B = a.B + (size_t)h * n * n;

    if (threadIdx.x == 0 && a.update_gm) a.gm[h] = vsum / (double)a.T;
    if (a.update_alpha)
        for (long q = threadIdx.x; q < nm; q += blockDim.x) alpha[q] = st[o.usum + q] / vsum;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        sig[i] = st[o.s2 + i] / vsum;
        cnw[i] = st[o.cnew + i] / vsum;
    }
    for (long q = threadIdx.x; q < (long)n * n; q += blockDim.x) phs[q] = st[o.phi + q] / vsum;
    __syncthreads();
    /* updatekappalambda with the new alpha and the old mu */
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
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
    /* getbflag: D_ik = kappa_i sigma2_i kappa_k sigma2_k <= 1, diagonal included */
    double cnt = 0.0;
    for (long q = threadIdx.x; q < (long)n * n; q += blockDim.x) {
        const int i = (int)(q % n), k = (int)(q / n);
        const double D = (kap[i] * sig[i]) * (kap[k] * sig[k]);
        if (D <= 1.0) cnt += 1.0;
    }
    const double bsum = block_sum(cnt, sh);
    const int bflag = (int)bsum;
    const int newt = (bflag == 0 && a.do_newton == 1 && k_it >= a.newt_start_iter) ? 1 : 0;
    for (long q = threadIdx.x; q < (long)n * n; q += blockDim.x) {
        const int i = (int)(q % n), k = (int)(q / n);
        double bv;
        if (newt) {
            if (i == k) {
                bv = (1.0 - phs[q]) / lam[i];                       /* updateBDiagonal */
            } else {
                const double D = (kap[i] * sig[i]) * (kap[k] * sig[k]);
                const double nom = -sig[i] * kap[k];
                bv = (nom * -phs[q] + -phs[k + (long)i * n]) / (D - 1.0);   /* getB */
            }
        } else {
            bv = (i == k ? 1.0 : 0.0) - phs[q];                     /* geteyemphi */
        }
        B[q] = bv;
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

/* A update, reparameterisation and c <- cnew for model blockIdx.x. */
/* AB_h = A_h B_h for every model, one thread per output element (the
 * threads of a wave share the column of B and read consecutive rows of A). */
__global__ __launch_bounds__(256) void ab_kernel(MstepArgs a) {
    const int h = blockIdx.y, n = a.n;
    const long nn = (long)n * n;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (!a.update_A || q >= nn) return;
    const int i = (int)(q % n), k = (int)(q / n);
    const double* A = a.A + (size_t)h * nn;
    const double* B = a.B + (size_t)h * nn + (long)k * n;
    double s = 0.0;
    for (int l = 0; l < n; ++l) s += A[i + (long)l * n] * B[l];
    a.work[8 + (size_t)h * nn + q] = s;
}

/* A += +-lrate A B, reparameterisation, mu shift with the old W, c = cnew. */
__global__ __launch_bounds__(1024) void mstep_c_kernel(MstepArgs a) {
    const int h = blockIdx.x, n = a.n, m = a.m;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const long nm = (long)n * m;
    const long nn = (long)n * n;
    double* A = a.A + (size_t)h * nn;
    double* c = a.c + (size_t)h * n;
    double* mu = a.mu + (size_t)h * nm;
    double* beta = a.beta + (size_t)h * nm;
    const double* cnw = a.cnew + (size_t)h * n;
    const double* W = a.W + (size_t)h * nn;
    const double* AB = a.work + 8 + (size_t)h * nn;
    if (a.update_A) {
        const double lr = a.ds->lrate;
        const double f = a.newton[h] ? lr : -lr;
        for (long q = threadIdx.x; q < nn; q += blockDim.x) A[q] += f * AB[q];
        __syncthreads();
        if (a.do_reparm) {
            /* normalizeAsave + normalizemubeta: one wave per column */
            for (int k = wv; k < n; k += nw) {
                double s = 0.0;
                for (int i = lane; i < n; i += 64) s += A[i + (long)k * n] * A[i + (long)k * n];
                const double nu = sqrt(wave_sum_f64(s));
                for (int i = lane; i < n; i += 64) A[i + (long)k * n] /= nu;
                for (int j = lane; j < m; j += 64) {
                    mu[(long)j * n + k] *= nu;
                    beta[(long)j * n + k] = beta[(long)j * n + k] / (nu * nu);
                }
            }
            __syncthreads();
            /* mu -= W_old (cnew - c) */
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                double d = 0.0;
                for (int l = 0; l < n; ++l) d += W[i + (long)l * n] * (cnw[l] - c[l]);
                for (int j = 0; j < m; ++j) mu[(long)j * n + i] -= d;
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) c[i] = cnw[i];
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

/* Everything after W = A^-1, for all models at once (grid (., M)): the GEMM
 * copies of W (rows h NP + i, columns k), W~ c with the W~ the unmixing GEMM
 * multiplies by, and the per-(h, j, i) E-step constants.  Flat index per
 * model: [0, n n) W entries, [n n, n n + n) rows of W~ c, then the n m
 * constants. */
__global__ __launch_bounds__(256) void prepare_tail_kernel(MstepArgs a) {
    const int h = blockIdx.y, n = a.n, m = a.m, NP = a.NP;
    const long nn = (long)n * n, nm = (long)n * m;
    const double* W = a.W + (size_t)h * nn;
    const long total = nn + n + nm;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        if (q < nn) {
            const int i = (int)(q % n), k = (int)(q / n);
            const double wv2 = W[q];
            const size_t dst = ((size_t)h * NP + i) * NP + k;
            if (a.prec_bf16) {
                /* FP8 storage: the GEMM multiplies q = x / s, so it gets W diag(s) */
                uint16_t hi, lo;
                split_w(a.xscale ? wv2 * (double)a.xscale[k] : wv2, hi, lo);
                a.Whi[dst] = hi;
                a.Wlo[dst] = lo;
            } else {
                a.Wg64[dst] = wv2;
            }
        } else if (q < nn + n) {
            const int i = (int)(q - nn);
            const double* c = a.c + (size_t)h * n;
            double s = 0.0;
            for (int k = 0; k < n; ++k) {
                double wr = W[i + (long)k * n];
                if (a.prec_bf16) {
                    /* the W~ the unmixing GEMM applies (W~ = split(W diag(s)) diag(1/s)) */
                    const double sk = a.xscale ? (double)a.xscale[k] : 1.0;
                    uint16_t hi, lo;
                    split_w(wr * sk, hi, lo);
                    wr = joined_w(hi, lo) / sk;
                }
                s += wr * c[k];
            }
            if (a.prec_bf16) a.wc[(size_t)h * NP + i] = (float)s;
            else a.wc64[(size_t)h * NP + i] = s;
        } else {
            const long e = q - nn - n;
            const int j = (int)(e / n), i = (int)(e % n);
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
            if (!a.prec_bf16) {
                double* dj = a.cst64 + ((size_t)h * kNumCst64 * m + j) * NP + i;
                dj[kCst64Sb * ks] = sb;
                dj[kCst64Nsbmu * ks] = -sb * mv;
            }
        }
    }
}

__device__ __forceinline__ void set_ldet(const MstepArgs& a, int h, double lad) {
    const double ldet = -lad;          /* log|det W| = -log|det A| */
    a.ldet[h] = ldet;
    a.lconst[h] = log(a.gm[h]) + ldet + a.ds->ldetS;
}

/* W = A^-1 by Gauss-Jordan with implicit partial pivoting, one workgroup per
 * model, the matrix in LDS (n <= kLdsInvMax) or, for larger n, in a global
 * scratch copy (L2-resident: n n doubles per model).  Rows are never moved: step k takes
 * the pivot p_k with the largest |a_ik| among rows not pivoted yet (ties to
 * the smallest index) and eliminates in place, so that afterwards
 * a[p_r][j] = (A^-1)[r][p_j]; W is gathered from that.  The wave that
 * updates the next pivot column takes its argmax while it writes it, which
 * leaves two barriers per step. */
constexpr int kLdsInvMax = 128;
constexpr int kInvMax = 1024;

template <bool GLOBAL>
__global__ __launch_bounds__(1024) void inverse_gj_kernel(MstepArgs a) {
    constexpr int NMAX = GLOBAL ? kInvMax : kLdsInvMax;
    extern __shared__ double sm[];               /* LDS variant: a[n n] (column-major) */
    __shared__ double prow[NMAX], pcol[NMAX];
    __shared__ int piv[NMAX], pinv[NMAX], used[NMAX];
    __shared__ int cand_i[16];
    __shared__ double lad_sh;
    const int h = blockIdx.x, n = a.n, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int bd = blockDim.x, nw = bd >> 6;
    const long nn = (long)n * n;
    double* aw = GLOBAL ? a.work + 8 + (size_t)a.M * nn + (size_t)h * nn : sm;
    const double* A = a.A + (size_t)h * nn;
    for (long q = tid; q < nn; q += bd) aw[q] = A[q];
    for (int i = tid; i < n; i += bd) used[i] = 0;
    if (tid == 0) lad_sh = 0.0;
    __syncthreads();
    /* lanes run over rows, waves over columns (no index division); the
     * candidates of column c come from wave c % nw */
    if (wv == 0) {
        double cv = -1.0;
        int ci = n;
        for (int i = lane; i < n; i += 64) {
            const double av = fabs(aw[i]);
            if (av > cv) { cv = av; ci = i; }
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const double ov = __shfl_xor(cv, off, 64);
            const int oi = __shfl_xor(ci, off, 64);
            if (ov > cv || (ov == cv && oi < ci)) { cv = ov; ci = oi; }
        }
        if (lane == 0) cand_i[0] = ci;
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        const int p = cand_i[k % nw];
        const double d = aw[p + (long)k * n];
        const double inv = 1.0 / d;
        if (tid < n) {
            prow[tid] = (tid == k) ? inv : aw[p + (long)tid * n] * inv;
        } else if (tid < 2 * n) {
            const int i = tid - n;
            pcol[i] = (i == p) ? 0.0 : aw[i + (long)k * n];
        }
        if (tid == 0) {
            lad_sh += log(fabs(d));
            piv[k] = p;
            used[p] = 1;
        }
        __syncthreads();
        for (int j = wv; j < n; j += nw) {
            const double pj = prow[j];
            double* col = aw + (long)j * n;
            double cv = -1.0;
            int ci = n;
            for (int i = lane; i < n; i += 64) {
                const double v = (i == p) ? pj : ((j == k) ? 0.0 : col[i]) - pcol[i] * pj;
                col[i] = v;
                if (j == k + 1 && !used[i]) {
                    const double av = fabs(v);
                    if (av > cv) { cv = av; ci = i; }
                }
            }
            if (j == k + 1) {   /* wave-uniform: this wave holds the next pivot column */
                for (int off = 32; off >= 1; off >>= 1) {
                    const double ov = __shfl_xor(cv, off, 64);
                    const int oi = __shfl_xor(ci, off, 64);
                    if (ov > cv || (ov == cv && oi < ci)) { cv = ov; ci = oi; }
                }
                if (lane == 0) cand_i[j % nw] = ci;
            }
        }
        __syncthreads();
    }
    for (int r = tid; r < n; r += bd) pinv[piv[r]] = r;
    __syncthreads();
    double* W = a.W + (size_t)h * nn;
    for (int mcol = wv; mcol < n; mcol += nw) {
        const double* src = aw + (long)pinv[mcol] * n;
        for (int r = lane; r < n; r += 64) W[r + (long)mcol * n] = src[piv[r]];
    }
    if (tid == 0) set_ldet(a, h, lad_sh);
}

}  // namespace

int launch_reduce(const double* part, int nparts, size_t pstride, const double* gxpart,
                  int nchunks, double* raw, double* gxtot, const MstepArgs& ma, hipStream_t st) {
    const int NS = kStatsPerJ * ma.m + kStatsPerI;
    const size_t plen = (size_t)ma.M * NS * ma.NP + ma.M + 1;
    hipLaunchKernelGGL(sum_parts_kernel, dim3((unsigned)((plen + 63) / 64)), dim3(64 * kSumWaves), 0,
                       st, part, nparts, pstride, plen, raw);
    const size_t glen = (size_t)ma.ldb * ma.NP;
    hipLaunchKernelGGL(sum_parts_kernel, dim3((unsigned)((glen + 63) / 64)), dim3(64 * kSumWaves), 0,
                       st, gxpart, nchunks, glen, glen, gxtot);
    /* one phi row per workgroup (threads over its columns) */
    hipLaunchKernelGGL(finalize_kernel, dim3(ma.M, ma.n), dim3(kBlk), 0, st, raw, gxtot, ma);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mstep(const MstepArgs& ma, hipStream_t st) {
    hipLaunchKernelGGL(mstep_a_kernel, dim3(ma.M), dim3(kBlk), 0, st, ma);
    hipLaunchKernelGGL(mstep_b_kernel, dim3(1), dim3(64), 0, st, ma);
    const long nn = (long)ma.n * ma.n;
    hipLaunchKernelGGL(ab_kernel, dim3((unsigned)((nn + 255) / 256), ma.M), dim3(256), 0, st, ma);
    hipLaunchKernelGGL(mstep_c_kernel, dim3(ma.M), dim3(1024), 0, st, ma);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static size_t inverse_lds_bytes(int n) { return (size_t)n * n * sizeof(double); }

int mstep_configure(int n) {
    if (n > kLdsInvMax) return 0;
    return hipFuncSetAttribute((const void*)inverse_gj_kernel<false>,
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)inverse_lds_bytes(n)) == hipSuccess ? 0 : -1;
}

int launch_prepare(const MstepArgs& ma, hipStream_t st) {
    if (ma.n > 1024) return -1;
    if (ma.n <= kLdsInvMax)
        hipLaunchKernelGGL(inverse_gj_kernel<false>, dim3(ma.M), dim3(1024), inverse_lds_bytes(ma.n), st, ma);
    else
        hipLaunchKernelGGL(inverse_gj_kernel<true>, dim3(ma.M), dim3(1024), 0, st, ma);
    const long total = (long)ma.n * ma.n + ma.n + (long)ma.n * ma.m;
    const unsigned gx = (unsigned)std::min<long>((total + 255) / 256, 256);
    hipLaunchKernelGGL(prepare_tail_kernel, dim3(gx, ma.M), dim3(256), 0, st, ma);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace amica
