/*
 * sphere.hip -- preprocessing on the GPU (fraimondo/cudamica
 * src/preprocess.cu:31-420, which uses cublasDsyrk + magma_dsyevd_gpu +
 * magma_dgesv_gpu) and the final output transform (src/amica.cu:2564-2668).
 *
 *   mean     = sum_t x_t / T                      FP64 partial sums per chunk
 *   C        = sum_t (x_t - mean)(x_t - mean)^T / T   FP64 tiles
 *   C        = V diag(d) V^T                      parallel cyclic Jacobi (FP64)
 *   S        = V diag(2/sqrt(d)) V^T = 2 C^-1/2   (do_sphere 1)
 *              diag(1/std) (do_sphere 2), I otherwise
 *   ldetS    = -log|det S|
 *   x_white  = S (x - mean), stored float32 [T][NP] (+ bf16 copy)
 */
#include "common.h"
#include "engine_internal.h"
#include "linalg_dev.h"

namespace amica {

namespace {

/* per-chunk FP64 column sums of x (T x n float32) */
__global__ __launch_bounds__(256) void colsum_kernel(const float* __restrict__ x, long T, int n,
                                                     long chunk, double* __restrict__ part) {
    const long ts = (long)blockIdx.x * chunk, te = min(T, ts + chunk);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double s = 0.0;
        for (long t = ts; t < te; ++t) s += (double)x[(size_t)t * n + i];
        part[(size_t)blockIdx.x * n + i] = s;
    }
}

__global__ void finish_mean_kernel(const double* __restrict__ part, int nchunks, int n, long T,
                                   double* __restrict__ mean) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += part[(size_t)c * n + i];
    mean[i] = s / (double)T;
}

/* C tile [ib, ib+64) x [kb, kb+64) summed over one chunk of samples.  Each
 * of the 256 threads owns a 4 x 4 block of the tile; 32 centred samples are
 * staged in LDS per step. */
__global__ __launch_bounds__(256) void cov_kernel(const float* __restrict__ x,
                                                  const double* __restrict__ mean, long T, int n,
                                                  long chunk, double* __restrict__ part) {
    __shared__ double xa[32][64];
    __shared__ double xb[32][64];
    const int tid = threadIdx.x;
    const int ib = blockIdx.y * 64, kb = blockIdx.z * 64;
    const long ts = (long)blockIdx.x * chunk, te = min(T, ts + chunk);
    const int ti = (tid & 15) * 4, tk = (tid >> 4) * 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (long t0 = ts; t0 < te; t0 += 32) {
        for (int q = tid; q < 32 * 64; q += 256) {
            const int s = q >> 6, c = q & 63;
            const long t = t0 + s;
            double va = 0.0, vb = 0.0;
            if (t < te) {
                if (ib + c < n) va = (double)x[(size_t)t * n + ib + c] - mean[ib + c];
                if (kb + c < n) vb = (double)x[(size_t)t * n + kb + c] - mean[kb + c];
            }
            xa[s][c] = va;
            xb[s][c] = vb;
        }
        __syncthreads();
        for (int s = 0; s < 32; ++s) {
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = xa[s][ti + a];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = xb[s][tk + b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
        }
        __syncthreads();
    }
    double* out = part + (size_t)blockIdx.x * n * n;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int i = ib + ti + a, k = kb + tk + b;
            if (i < n && k < n) out[i + (size_t)k * n] = acc[a][b];
        }
}

__global__ void finish_cov_kernel(const double* __restrict__ part, int nchunks, int n, long T,
                                  double* __restrict__ C) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += part[(size_t)c * n * n + e];
    C[e] = s / (double)T;
}

/* Parallel cyclic Jacobi (Brent-Luk round-robin ordering): each round
 * rotates N/2 disjoint (p, q) pairs at once.  A (n x n, overwritten) and V
 * are column-major in global memory; one workgroup. */
__global__ __launch_bounds__(256) void eigh_kernel(double* __restrict__ A, double* __restrict__ V,
                                                   double* __restrict__ d, int n) {
    __shared__ double cs[512], sn[512];
    __shared__ int pp[512], qq[512];
    __shared__ double red[8];
    __shared__ int done;
    const int tid = threadIdx.x;
    const int N = n + (n & 1);
    const int half = N / 2;
    for (long q = tid; q < (long)n * n; q += blockDim.x) V[q] = ((q % n) == (q / n)) ? 1.0 : 0.0;
    __syncthreads();
    for (int sweep = 0; sweep < 40; ++sweep) {
        /* convergence: off-diagonal Frobenius norm relative to the diagonal */
        double off = 0.0, dia = 0.0;
        for (long q = tid; q < (long)n * n; q += blockDim.x) {
            const double v = A[q];
            if ((q % n) == (q / n)) dia += v * v; else off += v * v;
        }
        off = wave_sum_f64(off);
        dia = wave_sum_f64(dia);
        if ((tid & 63) == 0) { red[tid >> 6] = off; red[4 + (tid >> 6)] = dia; }
        __syncthreads();
        if (tid == 0) {
            double o = 0.0, dd = 0.0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { o += red[w]; dd += red[4 + w]; }
            done = (o <= 1e-30 * dd || o == 0.0) ? 1 : 0;
        }
        __syncthreads();
        if (done) break;
        for (int r = 0; r < N - 1; ++r) {
            for (int k = tid; k < half; k += blockDim.x) {
                /* circle method: position 0 fixed, the others rotate */
                const int a0 = (k == 0) ? 0 : ((k - 1 + r) % (N - 1)) + 1;
                const int b0 = ((N - 1 - k - 1 + r) % (N - 1)) + 1;
                int p = min(a0, b0), q = max(a0, b0);
                double c = 1.0, s = 0.0;
                if (q < n) {
                    const double apq = A[p + (long)q * n];
                    if (apq != 0.0) {
                        const double app = A[p + (long)p * n], aqq = A[q + (long)q * n];
                        const double th = (aqq - app) / (2.0 * apq);
                        const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        s = t * c;
                    }
                } else {
                    q = p;   /* pair with the padding index: identity */
                }
                cs[k] = c; sn[k] = s; pp[k] = p; qq[k] = q;
            }
            __syncthreads();
            /* A <- A J (columns p, q) and V <- V J */
            for (long e = tid; e < (long)half * n; e += blockDim.x) {
                const int k = (int)(e / n), i = (int)(e % n);
                const int p = pp[k], q = qq[k];
                if (p == q) continue;
                const double c = cs[k], s = sn[k];
                const double ap = A[i + (long)p * n], aq = A[i + (long)q * n];
                A[i + (long)p * n] = c * ap - s * aq;
                A[i + (long)q * n] = s * ap + c * aq;
                const double vp = V[i + (long)p * n], vq = V[i + (long)q * n];
                V[i + (long)p * n] = c * vp - s * vq;
                V[i + (long)q * n] = s * vp + c * vq;
            }
            __syncthreads();
            /* A <- J^T A (rows p, q) */
            for (long e = tid; e < (long)half * n; e += blockDim.x) {
                const int k = (int)(e / n), j = (int)(e % n);
                const int p = pp[k], q = qq[k];
                if (p == q) continue;
                const double c = cs[k], s = sn[k];
                const double ap = A[p + (long)j * n], aq = A[q + (long)j * n];
                A[p + (long)j * n] = c * ap - s * aq;
                A[q + (long)j * n] = s * ap + c * aq;
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < n; i += blockDim.x) d[i] = A[i + (long)i * n];
}

/* S from the eigen-decomposition (do_sphere 1), ldetS = -log|det S|. */
__global__ __launch_bounds__(256) void sphere_from_eig_kernel(const double* __restrict__ V,
                                                              const double* __restrict__ d, int n,
                                                              double* __restrict__ S,
                                                              double* __restrict__ ldetS) {
    for (long q = threadIdx.x; q < (long)n * n; q += blockDim.x) {
        const int i = (int)(q % n), k = (int)(q / n);
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += V[i + (long)j * n] * (2.0 / sqrt(d[j])) * V[k + (long)j * n];
        S[q] = s;
    }
    if (threadIdx.x == 0) {
        double l = 0.0;
        for (int j = 0; j < n; ++j) l += log(2.0 / sqrt(d[j]));
        *ldetS = -l;
    }
}

/* per-chunk FP64 sums of (x - mean) and (x - mean)^2 for do_sphere 2 */
__global__ __launch_bounds__(256) void var_kernel(const float* __restrict__ x,
                                                  const double* __restrict__ mean, long T, int n,
                                                  long chunk, double* __restrict__ part) {
    const long ts = (long)blockIdx.x * chunk, te = min(T, ts + chunk);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double s1 = 0.0, s2 = 0.0;
        for (long t = ts; t < te; ++t) {
            const double v = (double)x[(size_t)t * n + i] - mean[i];
            s1 += v;
            s2 += v * v;
        }
        part[(size_t)blockIdx.x * 2 * n + i] = s1;
        part[(size_t)blockIdx.x * 2 * n + n + i] = s2;
    }
}

/* do_sphere 2: S = diag(1/sqrt(var)) with the unbiased variance of the
 * centred data (getvariance, kernels.cu:148-161); otherwise S = I. */
__global__ __launch_bounds__(256) void sphere_diag_kernel(const double* __restrict__ part,
                                                          int nchunks, int n, long T, int do_var,
                                                          double* __restrict__ S,
                                                          double* __restrict__ ldetS) {
    __shared__ double lsum[256];
    double acc = 0.0;
    for (long q = threadIdx.x; q < (long)n * n; q += blockDim.x) S[q] = 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double sii = 1.0;
        if (do_var) {
            double s1 = 0.0, s2 = 0.0;
            for (int c = 0; c < nchunks; ++c) {
                s1 += part[(size_t)c * 2 * n + i];
                s2 += part[(size_t)c * 2 * n + n + i];
            }
            const double m2 = s2 - s1 * s1 / (double)T;
            sii = 1.0 / sqrt(m2 / (double)(T - 1));
        }
        S[i + (long)i * n] = sii;
        acc += log(fabs(sii));
    }
    lsum[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double l = 0.0;
        for (int k = 0; k < (int)blockDim.x; ++k) l += lsum[k];
        *ldetS = -l;
    }
}

/* x_white[t][i] = sum_k S[i][k] (x[t][k] - mean[k]) in FP64, stored float32
 * with the channel dimension padded to NP (zeros), plus optional split bf16
 * copies: xbf[0 .. T NP) = bf16(x), xbf[T NP .. 2 T NP) = bf16(x - bf16(x)). */
__device__ __forceinline__ void store_split_bf16(uint16_t* xbf, size_t lo_ofs, size_t e, float f) {
    const uint16_t hi = f32_to_bf16_bits(f);
    xbf[e] = hi;
    xbf[lo_ofs + e] = f32_to_bf16_bits(f - bf16_bits_to_f32(hi));
}

__global__ __launch_bounds__(256) void whiten_kernel(const float* __restrict__ x,
                                                     const double* __restrict__ mean,
                                                     const double* __restrict__ S, long T, int n,
                                                     int NP, int TS, float* __restrict__ xw,
                                                     double* __restrict__ xw64,
                                                     uint16_t* __restrict__ xbf) {
    extern __shared__ double xc[];
    const long t0 = (long)blockIdx.x * TS;
    for (int q = threadIdx.x; q < TS * n; q += blockDim.x) {
        const int s = q / n, k = q % n;
        const long t = t0 + s;
        xc[q] = (t < T) ? (double)x[(size_t)t * n + k] - mean[k] : 0.0;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < TS * NP; q += blockDim.x) {
        const int s = q / NP, i = q % NP;
        const long t = t0 + s;
        if (t >= T) continue;
        double v = 0.0;
        if (i < n)
            for (int k = 0; k < n; ++k) v += S[i + (long)k * n] * xc[s * n + k];
        const float f = (float)v;
        xw[(size_t)t * NP + i] = f;
        if (xw64) xw64[(size_t)t * NP + i] = v;
        if (xbf) store_split_bf16(xbf, (size_t)T * NP, (size_t)t * NP + i, f);
    }
}

__global__ void white_from_f64_kernel(const double* __restrict__ xin, long T, int n, int NP,
                                      float* __restrict__ xw, double* __restrict__ xw64,
                                      uint16_t* __restrict__ xbf) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)T * NP) return;
    const long t = (long)(e / NP);
    const int i = (int)(e % NP);
    const double d = (i < n) ? xin[(size_t)t * n + i] : 0.0;
    const float f = (float)d;
    xw[e] = f;
    if (xw64) xw64[e] = d;
    if (xbf) store_split_bf16(xbf, (size_t)T * NP, e, f);
}

/* FP8 storage of the whitened data (per-channel scales s_i = max_t |x_ti| /
 * 448, the largest finite e4m3 value): colmaxabs over sample chunks, then
 * the scales (order-free max: deterministic), then per element q =
 * e4m3(x / s) (hardware conversion, round to nearest even), the data the
 * algorithm runs on x^ = s dequant(q) written back into xw, and bf16(q)
 * (exact) packed 8 samples per channel for the statistic GEMM. */
__global__ __launch_bounds__(256) void colmaxabs_kernel(const float* __restrict__ xw, long T, int NP,
                                                         long chunk, float* __restrict__ part) {
    const int i = blockIdx.y * 256 + threadIdx.x;
    if (i >= NP) return;
    const long t0 = (long)blockIdx.x * chunk, t1 = min(T, t0 + chunk);
    float mx = 0.f;
    for (long t = t0; t < t1; ++t) mx = fmaxf(mx, fabsf(xw[(size_t)t * NP + i]));
    part[(size_t)blockIdx.x * NP + i] = mx;
}

__global__ void fp8_scale_kernel(const float* __restrict__ part, int nchunks, int NP,
                                 float* __restrict__ scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NP) return;
    float mx = 0.f;
    for (int c = 0; c < nchunks; ++c) mx = fmaxf(mx, part[(size_t)c * NP + i]);
    scale[i] = mx > 0.f ? mx / 448.f : 1.f;
}

__global__ void quantize_fp8_kernel(float* __restrict__ xw, long T, int NP,
                                    const float* __restrict__ scale, uint8_t* __restrict__ xq,
                                    uint16_t* __restrict__ x8) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)T * NP) return;
    const long t = (long)(e / NP);
    const int i = (int)(e % NP);
    const float s = scale[i];
    const int packed = __builtin_amdgcn_cvt_pk_fp8_f32(xw[e] / s, 0.f, 0, false);
    const f2 d = __builtin_amdgcn_cvt_pk_f32_fp8(packed, false);
    xq[e] = (uint8_t)(packed & 0xff);
    xw[e] = d.x * s;
    x8[((size_t)(t >> 3) * NP + i) * 8 + (t & 7)] = (uint16_t)(__float_as_uint(d.x) >> 16);
}

__global__ void bf16_from_f32_kernel(const float* __restrict__ x, size_t count,
                                     uint16_t* __restrict__ out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) out[e] = f32_to_bf16_bits(x[e]);
}

/* c_out = S^-1 c_h + mean; A_out = colnorm(S^-1 A_h). */
__global__ __launch_bounds__(256) void output_transform_kernel(const double* __restrict__ S,
                                                               const double* __restrict__ mean,
                                                               const double* __restrict__ A,
                                                               const double* __restrict__ c,
                                                               int n, int M, double* lu,
                                                               double* X, double* Sinv,
                                                               double* A_out, double* c_out) {
    __shared__ int perm[1024];
    block_lu_inverse(S, n, lu, X, Sinv, perm);
    const long nn = (long)n * n;
    for (int h = 0; h < M; ++h) {
        const double* Ah = A + (size_t)h * nn;
        double* Ao = A_out + (size_t)h * nn;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            double s = mean[i];
            for (int k = 0; k < n; ++k) s += Sinv[i + (long)k * n] * c[(size_t)h * n + k];
            c_out[(size_t)h * n + i] = s;
        }
        for (long q = threadIdx.x; q < nn; q += blockDim.x) {
            const int i = (int)(q % n), k = (int)(q / n);
            double s = 0.0;
            for (int l = 0; l < n; ++l) s += Sinv[i + (long)l * n] * Ah[l + (long)k * n];
            Ao[q] = s;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += Ao[i + (long)k * n] * Ao[i + (long)k * n];
            const double nu = sqrt(s);
            for (int i = 0; i < n; ++i) Ao[i + (long)k * n] /= nu;
        }
        __syncthreads();
    }
}

long chunk_of(long T, int nchunks) { return (T + nchunks - 1) / nchunks; }

}  // namespace

int launch_colmean(const float* x, long T, int n, double* part, int nchunks, double* mean,
                   hipStream_t st) {
    hipLaunchKernelGGL(colsum_kernel, dim3(nchunks), dim3(256), 0, st, x, T, n,
                       chunk_of(T, nchunks), part);
    hipLaunchKernelGGL(finish_mean_kernel, dim3((n + 255) / 256), dim3(256), 0, st, part, nchunks,
                       n, T, mean);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_cov(const float* x, const double* mean, long T, int n, double* part, int nchunks,
               double* C, hipStream_t st) {
    const int nt = (n + 63) / 64;
    hipLaunchKernelGGL(cov_kernel, dim3(nchunks, nt, nt), dim3(256), 0, st, x, mean, T, n,
                       chunk_of(T, nchunks), part);
    const size_t nn = (size_t)n * n;
    hipLaunchKernelGGL(finish_cov_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st,
                       part, nchunks, n, T, C);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_eigh(double* C, double* V, double* d, int n, double* work, hipStream_t st) {
    (void)work;
    if (n > 1024) return -1;
    hipLaunchKernelGGL(eigh_kernel, dim3(1), dim3(256), 0, st, C, V, d, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_sphere_matrix(const double* V, const double* d, const double* C, int n,
                         const float* x, const double* mean, long T, int do_sphere, double* S,
                         double* ldetS_out, double* part, int nchunks, hipStream_t st) {
    (void)C;
    if (do_sphere == 1) {
        hipLaunchKernelGGL(sphere_from_eig_kernel, dim3(1), dim3(256), 0, st, V, d, n, S,
                           ldetS_out);
    } else if (do_sphere == 2) {
        hipLaunchKernelGGL(var_kernel, dim3(nchunks), dim3(256), 0, st, x, mean, T, n,
                           chunk_of(T, nchunks), part);
        hipLaunchKernelGGL(sphere_diag_kernel, dim3(1), dim3(256), 0, st, part, nchunks, n, T, 1,
                           S, ldetS_out);
    } else {
        hipLaunchKernelGGL(sphere_diag_kernel, dim3(1), dim3(256), 0, st, part, 0, n, T, 0, S,
                           ldetS_out);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_whiten(const float* x, const double* mean, const double* S, long T, int n, int NP,
                  float* xw, double* xw64, uint16_t* xbf, hipStream_t st) {
    int TS = 4096 / n;
    if (TS < 1) TS = 1;
    const size_t shm = (size_t)TS * n * sizeof(double);
    hipLaunchKernelGGL(whiten_kernel, dim3((unsigned)((T + TS - 1) / TS)), dim3(256), shm, st, x,
                       mean, S, T, n, NP, TS, xw, xw64, xbf);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_white_from_f64(const double* xin, long T, int n, int NP, float* xw, double* xw64,
                          uint16_t* xbf, hipStream_t st) {
    const size_t cnt = (size_t)T * NP;
    hipLaunchKernelGGL(white_from_f64_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0,
                       st, xin, T, n, NP, xw, xw64, xbf);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_quantize_fp8(float* xw, long T, int NP, float* scale, uint8_t* xq, uint16_t* x8,
                        float* part, int nchunks, hipStream_t st) {
    const long chunk = (T + nchunks - 1) / nchunks;
    hipLaunchKernelGGL(colmaxabs_kernel, dim3(nchunks, (NP + 255) / 256), dim3(256), 0, st, xw, T, NP,
                       chunk, part);
    hipLaunchKernelGGL(fp8_scale_kernel, dim3((NP + 255) / 256), dim3(256), 0, st, part, nchunks, NP,
                       scale);
    const size_t cnt = (size_t)T * NP;
    hipLaunchKernelGGL(quantize_fp8_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st, xw,
                       T, NP, scale, xq, x8);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_bf16_from_f32(const float* x, size_t count, uint16_t* out, hipStream_t st) {
    hipLaunchKernelGGL(bf16_from_f32_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       st, x, count, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_output_transform(const double* S, const double* mean, const double* A,
                            const double* c, int n, int M, double* lu, double* Sinv,
                            double* A_out, double* c_out, hipStream_t st) {
    if (n > 1024) return -1;
    double* X = lu + (size_t)n * n;
    hipLaunchKernelGGL(output_transform_kernel, dim3(1), dim3(256), 0, st, S, mean, A, c, n, M,
                       lu, X, Sinv, A_out, c_out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace amica
