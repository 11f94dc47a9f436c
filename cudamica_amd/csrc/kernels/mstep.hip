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
 * EM run pinv(A) = A^-1, computed here by LU with partial pivoting, and
 * log|det W| = -sum log|u_kk| (the reference multiplies the diagonal of the
 * LU of W, which can overflow).
 */
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

/* Digamma by the Stirling asymptotic series (Bernoulli coefficients
 * B_2k / 2k) at x + 6 and the recurrence psi(x) = psi(x+1) - 1/x. */
__device__ double psi_dev(double x0) {
    const double x = x0 + 6.0;
    const double inv = 1.0 / (x * x);
    const double coef[9] = {1.0 / 12.0, -1.0 / 120.0, 1.0 / 252.0, -1.0 / 240.0, 1.0 / 132.0,
                            -691.0 / 32760.0, 1.0 / 12.0, -3617.0 / 8160.0, 43867.0 / 14364.0};
    double r = log(x) - 0.5 / x;
    double p = inv;
#pragma unroll
    for (int k = 0; k < 9; ++k) { r -= coef[k] * p; p *= inv; }
#pragma unroll
    for (int k = 5; k >= 0; --k) r -= 1.0 / (x0 + k);
    return r;
}

/* ------------------------------------------------------------------------ */
/* reductions                                                                 */
/* ------------------------------------------------------------------------ */

__global__ void sum_parts_kernel(const double* __restrict__ part, int nparts, size_t pstride,
                                 size_t len, double* __restrict__ out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= len) return;
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += part[(size_t)p * pstride + e];
    out[e] = s;
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
    const double* rh = raw + (size_t)h * NS * NP;
    const long nm = (long)n * m;
    if (blockIdx.y == 0) {
        for (long q = threadIdx.x; q < nm; q += blockDim.x) {
            const int j = (int)(q / n), i = (int)(q % n);
            const double be = a.beta[(size_t)h * nm + q], r = a.rho[(size_t)h * nm + q];
            const double sb = sqrt(be);
            const double* rj = rh + (size_t)j * kStatsPerJ * NP + i;
            const double U = rj[kRawU * NP], Us = rj[kRawUs * NP], Us2 = rj[kRawUs2 * NP];
            const double Um2 = rj[kRawUm2 * NP], Ur = rj[kRawUr * NP], Ur2 = rj[kRawUr2 * NP];
            const double Url = rj[kRawUrl * NP];
            st[o.usum + q] = U;
            st[o.kn + q] = be * r * r * Us2;
            st[o.ln + q] = r * r * Ur2 - 2.0 * r * Ur + U;
            st[o.mn + q] = r * Us;
            st[o.md + q] = (r <= 2.0) ? sb * r * Um2 : sb * r * r * Us2;
            st[o.bn + q] = (r <= 2.0) ? U : 0.0;
            st[o.bd + q] = r * Ur;
            st[o.rn + q] = r * AMICA_LN2 * Url;
        }
        const double* ri = rh + (size_t)kStatsPerJ * m * NP;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            st[o.s2 + i] = ri[kRawS2 * NP + i];
            st[o.cnew + i] = ri[kRawCnew * NP + i];
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
    const double* gsum = rh + (size_t)(kStatsPerJ * m + kRawGsum) * NP;
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
        s -= gsum[i] * wck;
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
__global__ __launch_bounds__(kBlk) void mstep_c_kernel(MstepArgs a) {
    const int h = blockIdx.x, n = a.n, m = a.m;
    const long nm = (long)n * m;
    const long nn = (long)n * n;
    double* A = a.A + (size_t)h * nn;
    double* c = a.c + (size_t)h * n;
    double* mu = a.mu + (size_t)h * nm;
    double* beta = a.beta + (size_t)h * nm;
    const double* cnw = a.cnew + (size_t)h * n;
    const double* W = a.W + (size_t)h * nn;
    double* AB = a.work + 8 + (size_t)h * nn;
    if (a.update_A) {
        const double lr = a.ds->lrate;
        const double f = a.newton[h] ? lr : -lr;
        const double* B = a.B + (size_t)h * nn;
        for (long q = threadIdx.x; q < nn; q += blockDim.x) {
            const int i = (int)(q % n), k = (int)(q / n);
            double s = 0.0;
            for (int l = 0; l < n; ++l) s += A[i + (long)l * n] * B[l + (long)k * n];
            AB[q] = s;
        }
        __syncthreads();
        for (long q = threadIdx.x; q < nn; q += blockDim.x) A[q] += f * AB[q];
        __syncthreads();
        if (a.do_reparm) {
            /* normalizeAsave + normalizemubeta */
            for (int k = threadIdx.x; k < n; k += blockDim.x) {
                double s = 0.0;
                for (int i = 0; i < n; ++i) s += A[i + (long)k * n] * A[i + (long)k * n];
                const double nu = sqrt(s);
                for (int i = 0; i < n; ++i) A[i + (long)k * n] /= nu;
                for (int j = 0; j < m; ++j) {
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

__global__ __launch_bounds__(kBlk) void prepare_kernel(MstepArgs a) {
    __shared__ int perm[1024];
    const int h = blockIdx.x, n = a.n, m = a.m, NP = a.NP;
    const long nn = (long)n * n, nm = (long)n * m;
    const int tid = threadIdx.x;
    double* W = a.W + (size_t)h * nn;
    double* X = a.work + 8 + (size_t)a.M * nn + (size_t)h * nn;
    const double lad = block_lu_inverse(a.A + (size_t)h * nn, n, a.lu + (size_t)h * nn, X, W, perm);
    /* GEMM copy of W (rows h*NP + i, columns k) and W~ c with the W~ the
     * unmixing GEMM multiplies by: FP64 W, or bf16 hi + lo parts of W */
    for (long q = tid; q < nn; q += blockDim.x) {
        const int i = (int)(q / n), k = (int)(q % n);
        const double wv2 = W[i + (long)k * n];
        const size_t dst = ((size_t)h * NP + i) * NP + k;
        if (a.prec_bf16) {
            const uint16_t hi = f32_to_bf16_bits((float)wv2);
            a.Whi[dst] = hi;
            a.Wlo[dst] = f32_to_bf16_bits((float)(wv2 - (double)bf16_bits_to_f32(hi)));
        } else {
            a.Wg64[dst] = wv2;
        }
    }
    __syncthreads();
    const double* c = a.c + (size_t)h * n;
    for (int i = tid; i < n; i += blockDim.x) {
        double s = 0.0;
        const size_t row = ((size_t)h * NP + i) * NP;
        for (int k = 0; k < n; ++k) {
            const double wr = a.prec_bf16 ? (double)bf16_bits_to_f32(a.Whi[row + k]) +
                                                (double)bf16_bits_to_f32(a.Wlo[row + k])
                                          : a.Wg64[row + k];
            s += wr * c[k];
        }
        if (a.prec_bf16) a.wc[(size_t)h * NP + i] = (float)s;
        else a.wc64[(size_t)h * NP + i] = s;
    }
    /* E-step constants */
    for (long q = tid; q < nm; q += blockDim.x) {
        const int j = (int)(q / n), i = (int)(q % n);
        const double al = a.alpha[(size_t)h * nm + q], mv = a.mu[(size_t)h * nm + q];
        const double be = a.beta[(size_t)h * nm + q], r = a.rho[(size_t)h * nm + q];
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
    if (tid == 0) {
        const double ldet = -lad;
        a.ldet[h] = ldet;
        a.lconst[h] = log(a.gm[h]) + ldet + a.ds->ldetS;
    }
}

}  // namespace

int launch_reduce(const double* part, int nparts, size_t pstride, const double* gxpart,
                  int nchunks, double* raw, double* gxtot, const MstepArgs& ma, hipStream_t st) {
    const int NS = kStatsPerJ * ma.m + kStatsPerI;
    const size_t plen = (size_t)ma.M * NS * ma.NP + ma.M + 1;
    hipLaunchKernelGGL(sum_parts_kernel, dim3((unsigned)((plen + 255) / 256)), dim3(256), 0, st,
                       part, nparts, pstride, plen, raw);
    const size_t glen = (size_t)ma.ldb * ma.NP;
    hipLaunchKernelGGL(sum_parts_kernel, dim3((unsigned)((glen + 255) / 256)), dim3(256), 0, st,
                       gxpart, nchunks, glen, glen, gxtot);
    const int ysplit = (ma.n + 31) / 32;
    hipLaunchKernelGGL(finalize_kernel, dim3(ma.M, ysplit), dim3(kBlk), 0, st, raw, gxtot, ma);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_mstep(const MstepArgs& ma, hipStream_t st) {
    hipLaunchKernelGGL(mstep_a_kernel, dim3(ma.M), dim3(kBlk), 0, st, ma);
    hipLaunchKernelGGL(mstep_b_kernel, dim3(1), dim3(64), 0, st, ma);
    hipLaunchKernelGGL(mstep_c_kernel, dim3(ma.M), dim3(kBlk), 0, st, ma);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_prepare(const MstepArgs& ma, hipStream_t st) {
    if (ma.n > 1024) return -1;
    hipLaunchKernelGGL(prepare_kernel, dim3(ma.M), dim3(kBlk), 0, st, ma);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace amica
