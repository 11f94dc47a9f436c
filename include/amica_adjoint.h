/*
 * amica_adjoint.h -- the per-sample AMICA log-likelihood and its adjoint on
 * the GPU (C ABI, libcudamica_amd).
 *
 * No counterpart in the reference: fraimondo/cudamica evaluates the same
 * per-sample terms only inside its EM statistics (src/kernels.cu:251-450,
 * src/amica.cu:1009-1606).  Here they are exposed as a differentiable
 * function of the parameters, for gradient checks and for callers that
 * optimise the AMICA likelihood by other means.
 *
 * For whitened data x_t (rows of x, t < T) and models h < M:
 *   b_h   = W_h (x_t - c_h)
 *   y_hji = sqrt(beta_hji) (b_hi - mu_hji)
 *   Q_hji = log alpha_hji + 0.5 log beta_hji - log 2 - lgamma(1 + 1/rho_hji) - |y_hji|^rho_hji
 *   L_h   = log gm_h + log|det W_h| + ldetS + sum_i logsumexp_j Q_hji
 *   ll_t  = logsumexp_h L_h
 * (log 2 is the reference's 0.693147180559945, src/config.h:25).  The sum of
 * ll_t over t divided by n T is the EM log-likelihood of these parameters.
 *
 * amica_ll_adjoint returns the gradients of sum_t dll_t ll_t with respect to
 * W, c, alpha, mu, beta, rho and gm; the responsibilities are recomputed,
 * never stored.  With dll = 1 they equal the reference's statistics:
 *   dW_h W_h^T = vsum_h I - Phi_h            dalpha = usum / alpha
 *   dc_h = W_h^T sum_t g_t                   dmu = sqrt(beta) mu_numer
 *   dbeta = (usum - sum u f y) / (2 beta)    drho = usum psi(1+1/rho)/rho^2 - rho_numer/rho
 *
 * Arguments are device pointers (HIP) and sizes, no library types:
 *   real_bits  32 or 64: x is float or double, and the density math runs in
 *              that precision (sums in FP64 either way)
 *   x          [T][n] whitened data, sample-major (channels contiguous)
 *   W          [M][n][n] column-major (W_h[i][k] at h n n + k n + i)
 *   c          [M][n]
 *   alpha, mu, beta, rho   [M][m][n] (component j of channel i at h m n + j n + i)
 *   gm         [M]
 *   ldetS      log|det S| of the sphering (a constant of every L_h)
 *   ll         [T] FP64 output, or NULL (adjoint only)
 *   dll        [T] FP64 upstream gradient
 *   dW ... dgm FP64 outputs, the shapes of their parameters
 *   stream     hipStream_t the work is enqueued on (NULL: the default stream)
 * Limits: 1 <= m <= 8, n <= 1024, M <= 16 (the reference's range).  Where
 * M * ceil(n/64) * 64 > 512 the likelihood runs in three passes through
 * per-sample buffers of (M ceil(n/64) + M) doubles.  The call allocates its
 * workspace, synchronises `stream` before it returns and returns 0, or a
 * negative AMICA_E_* code of amica_engine.h with amica_last_error() set.
 */
#ifndef AMICA_ADJOINT_H_
#define AMICA_ADJOINT_H_

#ifdef __cplusplus
extern "C" {
#endif

int amica_ll_forward(int real_bits, const void* x, long T, int n, int M, int m,
                     const double* W, const double* c, const double* alpha, const double* mu,
                     const double* beta, const double* rho, const double* gm, double ldetS,
                     double* ll, void* stream);

/* amica_ll_forward without allocating: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as amica_ll_forward, up to the rounding of b (it
 * forms b = W x - W c per lane, amica_ll_forward on MFMA into a buffer);
 * its workspace is a static device array of 1.6 MB per GPU (held until the
 * process ends), so it also requires M * n * n <= 65536 and
 * M * ceil(n/64) * 64 <= 512 and rejects larger models with
 * AMICA_E_INVALID.  The call only enqueues work on `stream` and returns:
 * the caller synchronises it.  Calls on one GPU share the workspace, so
 * calls that overlap on one GPU (other streams, other host threads) race. */
int amica_ll_forward_noalloc(int real_bits, const void* x, long T, int n, int M, int m,
                             const double* W, const double* c, const double* alpha,
                             const double* mu, const double* beta, const double* rho,
                             const double* gm, double ldetS, double* ll, void* stream);

int amica_ll_adjoint(int real_bits, const void* x, long T, int n, int M, int m,
                     const double* W, const double* c, const double* alpha, const double* mu,
                     const double* beta, const double* rho, const double* gm, double ldetS,
                     const double* dll, double* ll, double* dW, double* dc, double* dalpha,
                     double* dmu, double* dbeta, double* drho, double* dgm, void* stream);

#ifdef __cplusplus
}
#endif

#endif
