/*
 * linalg_dev.h -- workgroup-level FP64 dense linear algebra used by the
 * engine's small per-model kernels (n <= 1024, column-major), and the
 * digamma function of the rho update and the likelihood adjoint.
 */
#ifndef AMICA_LINALG_DEV_H_
#define AMICA_LINALG_DEV_H_

#include "common.h"

namespace amica {

/* Digamma by the Stirling asymptotic series (Bernoulli coefficients
 * B_2k / 2k) at x + 6 and the recurrence psi(x) = psi(x+1) - 1/x. */
__device__ inline double psi_dev(double x0) {
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

/* Inverse of the n x n matrix A by LU with partial pivoting, computed by
 * one workgroup (blockDim.x a multiple of 64, <= 1024).  lu and X are n*n
 * scratch arrays in global memory, perm an n-int array in LDS.  Writes
 * Ainv (column-major) and returns sum_k log|u_kk| = log|det A| to every
 * thread.  Ties in the pivot search go to the smallest row index. */
__device__ inline double block_lu_inverse(const double* __restrict__ A, int n,
                                          double* __restrict__ lu, double* __restrict__ X,
                                          double* __restrict__ Ainv, int* perm) {
    __shared__ double shv[16];
    __shared__ int shi[16];
    __shared__ double lad_sh;
    const long nn = (long)n * n;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
    for (long q = tid; q < nn; q += blockDim.x) lu[q] = A[q];
    for (int i = tid; i < n; i += blockDim.x) perm[i] = i;
    if (tid == 0) lad_sh = 0.0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bi = n;
        for (int i = k + tid; i < n; i += blockDim.x) {
            const double v = fabs(lu[i + (long)k * n]);
            if (v > best) { best = v; bi = i; }
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (lane == 0) { shv[wv] = best; shi[wv] = bi; }
        __syncthreads();
        best = shv[0]; bi = shi[0];
        for (int q = 1; q < nw; ++q)
            if (shv[q] > best || (shv[q] == best && shi[q] < bi)) { best = shv[q]; bi = shi[q]; }
        const int p = bi;
        if (p != k) {
            for (int j = tid; j < n; j += blockDim.x) {
                const double t1 = lu[k + (long)j * n];
                lu[k + (long)j * n] = lu[p + (long)j * n];
                lu[p + (long)j * n] = t1;
            }
            if (tid == 0) { const int tp = perm[k]; perm[k] = perm[p]; perm[p] = tp; }
        }
        __syncthreads();
        const double d = lu[k + (long)k * n];
        if (tid == 0) lad_sh += log(fabs(d));
        for (int i = k + 1 + tid; i < n; i += blockDim.x) lu[i + (long)k * n] /= d;
        __syncthreads();
        const int r = n - k - 1;
        for (long q = tid; q < (long)r * r; q += blockDim.x) {
            const int i = k + 1 + (int)(q % r), j = k + 1 + (int)(q / r);
            lu[i + (long)j * n] -= lu[i + (long)k * n] * lu[k + (long)j * n];
        }
        __syncthreads();
    }
    /* L U X = P, one right-hand side per thread; X row-major so that the
     * threads of a wave touch consecutive words. */
    for (int j = tid; j < n; j += blockDim.x) {
        for (int i = 0; i < n; ++i) {
            double s = (perm[i] == j) ? 1.0 : 0.0;
            for (int l = 0; l < i; ++l) s -= lu[i + (long)l * n] * X[(long)l * n + j];
            X[(long)i * n + j] = s;
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = X[(long)i * n + j];
            for (int l = i + 1; l < n; ++l) s -= lu[i + (long)l * n] * X[(long)l * n + j];
            X[(long)i * n + j] = s / lu[i + (long)i * n];
        }
    }
    __syncthreads();
    for (long q = tid; q < nn; q += blockDim.x) {
        const int i = (int)(q % n), k = (int)(q / n);
        Ainv[q] = X[(long)i * n + k];
    }
    __syncthreads();
    return lad_sh;
}

}  // namespace amica

#endif
