/*
 * amica_engine.h -- C ABI of the MI355X AMICA EM engine.
 *
 * This is the thin layer between the host program (runamica(), the
 * cudaamica CLI, Python tests) and the hand-written HIP kernels.  It has no
 * counterpart in the reference: there the host driver src/amica.cu:624-2843
 * calls cuBLAS and the kernels of src/kernels.cu directly.  The engine owns
 * all device memory of one recording on one GPU and runs the whole EM
 * iteration on the device (no host round trip per iteration).
 *
 * Layouts (host side, all FP64 unless stated):
 *   x         float32, T x n, sample-major (x[t*n + i]) -- the .fdt layout
 *   A, W      n x n per model, column-major, models stacked (A[h*n*n + i + k*n])
 *   c         n per model
 *   alpha, mu, beta, rho   n x m per model, element (i, j) at [h*n*m + j*n + i]
 *   gm        M
 * Every function returns 0 on success and a negative AMICA_E* code on error;
 * amica_last_error() gives a message.
 */
#ifndef AMICA_ENGINE_H_
#define AMICA_ENGINE_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMICA_OK               0
#define AMICA_E_INVALID      (-3)   /* same value as ERRORINVALIDPARAM */
#define AMICA_E_NOMEM        (-1)   /* same value as ERRORNODEVICEMEM */
#define AMICA_E_DEVICE      (-10)   /* HIP runtime error */
#define AMICA_E_STATE       (-11)   /* call out of order (e.g. no data loaded) */

#define AMICA_PREC_FP32       32    /* FP32 density math on FP64 b (FP64 unmixing GEMM), FP32
                                       statistic GEMM: the parity path */
#define AMICA_PREC_FP32_PLAIN 33    /* both GEMMs on FP32 operands (FP32 MFMA), FP32 b and g,
                                       FP32 density math: the plain FP32 path the mixed
                                       precision ones are compared with */
#define AMICA_PREC_BF16       16    /* BF16 GEMM operands, FP32 accumulation and density math */
#define AMICA_PREC_FP8         8    /* as BF16, with the whitened data stored as FP8 e4m3 with
                                       per-channel scales (the algorithm then runs on the
                                       dequantized data) */

typedef struct amica_engine amica_engine;

typedef struct {
    int    n;                /* channels */
    long   T;                /* samples */
    int    M;                /* models */
    int    m;                /* mixture components per source (1..8) */
    int    precision;        /* AMICA_PREC_FP32, AMICA_PREC_FP32_PLAIN, AMICA_PREC_BF16 or
                                AMICA_PREC_FP8 */
    int    max_iter;         /* capacity of the per-iteration logs */
    int    device;           /* HIP device ordinal */
    /* EM options, same meaning and defaults as the reference (config.h:46-83) */
    int    do_newton;        /* 1 on, 0 off, 2 "none" (counts as on for the lrate rule) */
    int    update_rho;
    int    rho_start_iter;
    int    newt_start_iter;
    int    update_gm, update_alpha, update_A, update_mu, update_beta, do_reparm;
    double lrate0, lratemax, lnatrate, lratefact;
    double rholrate, rhomin, rhomax;
    int    maxdec;
} amica_options;

/* Fills opts with the reference defaults for an n x T recording. */
void amica_options_default(amica_options* opts, int n, long T, int M, int m);

int  amica_engine_create(const amica_options* opts, amica_engine** out);
int  amica_engine_destroy(amica_engine* eng);

/* Raw float32 data (T x n sample-major), host or device pointer. */
int  amica_engine_load_raw(amica_engine* eng, const float* x, int x_on_device);
/* FP64 channel means of the raw data (center(), preprocess.cu:31-85). */
int  amica_engine_center(amica_engine* eng);
/* Sphering matrix S, ldetS and x_white = S (x - mean) (whiten(),
 * preprocess.cu:89-420); do_sphere 1: 2 C^-1/2, 2: diag(1/std), else I. */
int  amica_engine_whiten(amica_engine* eng, int do_sphere);
/* center + whiten. */
int  amica_engine_preprocess(amica_engine* eng, int do_sphere);
/* Sphering results: S (n x n col-major), mean (n), ldetS. */
int  amica_engine_get_sphere(amica_engine* eng, double* S, double* mean, double* ldetS);
/* Loads already-whitened FP64 data (T x n) and ldetS, skipping preprocessing. */
int  amica_engine_load_white(amica_engine* eng, const double* x_white, double ldetS);
/* The whitened data the engine holds, as float32 T x n. */
int  amica_engine_get_white(amica_engine* eng, float* x_white);

/* Model state.  set_state resets the iteration counter and the lrate
 * schedule (lrate = lrate0). */
int  amica_engine_set_state(amica_engine* eng, const double* A, const double* c,
                            const double* alpha, const double* mu, const double* beta,
                            const double* rho, const double* gm);
/* W is pinv(A) as used by the last E-step (the reference's weights output). */
int  amica_engine_get_state(amica_engine* eng, double* A, double* c, double* alpha,
                            double* mu, double* beta, double* rho, double* gm, double* W);

/* Enqueues k full EM iterations (E-step, M-step, Newton, reparameterisation)
 * on the engine's stream.  Asynchronous; use amica_engine_sync. */
int  amica_engine_iterate(amica_engine* eng, int k);
/* Same, replaying one captured hipGraph per iteration. */
int  amica_engine_iterate_graph(amica_engine* eng, int k);
int  amica_engine_sync(amica_engine* eng);
/* Number of iterations run since set_state. */
int  amica_engine_iterations(amica_engine* eng);

/* Per-iteration logs of the first n_iter iterations (any pointer may be
 * NULL): LL, lrate printed at that iteration, bflag and Newton flag per
 * model (n_iter x M), likelihood-decrease flag, vsum per model (n_iter x M). */
int  amica_engine_get_log(amica_engine* eng, int n_iter, double* ll, double* lrate,
                          int* bflag, int* newton, int* lldec, double* vsum);

/* The log entry of iteration `it` (1-based) alone: LL, lrate, lldec and per
 * model bflag, Newton flag and vsum (M each); any pointer may be NULL. */
int  amica_engine_get_log_at(amica_engine* eng, int it, double* ll, double* lrate, int* bflag,
                             int* newton, int* lldec, double* vsum);

/* Runs only the E-step with the current state and returns the canonical
 * statistics (per model: vsum, usum, kappa_numer, lambda_numer, mu_numer,
 * mu_denom, beta_numer, beta_denom, rho_numer (m*n each), sigma2_numer,
 * cnew (n each), phi (n*n col-major); then sum_t P_t).  stats holds
 * M*(1 + 8mn + 2n + n*n) + 1 doubles.  The state is not modified. */
int  amica_engine_estep_stats(amica_engine* eng, double* stats);
/* Enables recording of L_ht by the following E-steps (off by default). */
int  amica_engine_record_ltall(amica_engine* eng, int on);
/* L_ht of the last recorded E-step (M x T, model-major), the reference's
 * ltall output. */
int  amica_engine_get_ltall(amica_engine* eng, double* ltall);
/* Output transform (amica.cu:2564-2668): c_out = S^+ c + mean,
 * A_out = colnorm(S^+ A), computed on the GPU from the current state. */
int  amica_engine_output(amica_engine* eng, double* A_out, double* c_out);

/* Average duration (ms) of each kernel class over the last timed
 * iterations, measured with HIP events on the engine stream (ms[5]):
 * [0] W = A^-1 and the E-step constants, [1] unmixing GEMM, [2] fused
 * E-step, [3] statistic GEMM, [4] reductions + M-step.  Enabled by
 * amica_engine_set_timing(eng, 1). */
int  amica_engine_set_timing(amica_engine* eng, int on);
int  amica_engine_get_timing(amica_engine* eng, double* ms, int* count);

/* Diagnostic: copies one internal device buffer to host memory (names: x,
 * x64, xbf, b, b64, g, gx, raw, Wg64, wc64, wc); bytes must equal its size. */
int  amica_engine_debug_buffer(amica_engine* eng, const char* name, void* out, size_t bytes);

const char* amica_last_error(void);
/* Number of compute units of the engine's device. */
int  amica_engine_cu_count(amica_engine* eng);
/* Sizes the engine's grids for `cus` compute units (<= 0: the whole device;
 * the E-step then runs (resident workgroups per CU) x cus workgroups, which
 * the hardware spreads over all CUs).  No counterpart in the reference; the
 * single-recording form of amica_batch_set_budgets, for occupancy studies. */
int  amica_engine_set_cu_budget(amica_engine* eng, int cus);

/* Achievable HBM bandwidth on `device`, measured with this library's own
 * streaming kernels over two buffers of `bytes` each (best of `reps`
 * timed passes after one warm-up): copy_gbs counts bytes read + written,
 * read_gbs bytes read by a read-only pass.  No counterpart in the
 * reference (it prints only per-iteration times, amica.cu:2538-2540);
 * the yardstick beside the nominal HBM peak for the kernel figures. */
int  amica_hbm_probe(int device, size_t bytes, int reps, double* copy_gbs, double* read_gbs);

#ifdef __cplusplus
}
#endif

#endif
