/*
 * oracle.h -- CPU oracle for the AMICA EM iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain C++ restatement of the
 * reference's EM update equations (fraimondo/cudamica, src/amica.cu and
 * src/kernels.cu) used as the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (cudamica_amd/) never links or calls it.
 *
 * All matrices are column-major like the reference (element (i,k) of an n x n
 * matrix at [i + k*n]).  Parameter arrays alpha/mu/beta/rho are n x m per
 * model, element (i=channel, j=component) at [i + j*n], models stacked.
 * Sample data are sample-major: x[t*n + i].
 */
#ifndef AMICA_ORACLE_H_
#define AMICA_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;                 /* channels (sources) */
    long T;                /* samples */
    int M;                 /* ICA models */
    int m;                 /* mixture components per source */
    int maxiter;
    int do_newton;         /* 0, 1, or 2 ("none"); see orc_run */
    int update_rho;
    int rho_start_iter;
    int newt_start_iter;
    int update_gm, update_alpha, update_A, update_mu, update_beta, do_reparm;
    double lrate0, lratemax, lnatrate, lratefact;
    double rholrate, rhomin, rhomax;
    int maxdec;
} orc_config;

/* Fill cfg with the reference's fixed defaults (config.h:46-83). */
void orc_default_config(orc_config* cfg);

/* Number of doubles of the canonical statistics block of one model:
 * 1 + 8*m*n + 2*n + n*n.  Layout per model:
 *   [0]                vsum
 *   usum, kappa_numer, lambda_numer, mu_numer, mu_denom, beta_numer,
 *   beta_denom, rho_numer                                  (8 arrays of m*n)
 *   sigma2_numer[n], cnew[n]
 *   phi[n*n]           column-major, phi[i + k*n] = sum_t g_i b_k
 * The full statistics buffer holds M such blocks followed by one double,
 * sum_t P_t (the un-normalised log-likelihood). */
long orc_stats_stride(int n, int m);

/* Preprocessing (preprocess.cu:31-420): FP64 mean removal, sphering
 * (do_sphere 1: S = 2 C^-1/2; 2: diag(1/std); else identity),
 * ldetS = -log|det S|, x_white = S (x - mean).  x_raw is float32 T x n
 * sample-major; x_white is double T x n sample-major. */
int orc_preprocess(int n, long T, const float* x_raw, int do_sphere,
                   double* x_white, double* S, double* mean, double* ldetS);

/* One E-step (amica.cu:996-1611) with the given parameters.  W and ldet are
 * computed from A.  stats receives M*stride + 1 doubles.  update_rho
 * controls whether rho_numer is accumulated. */
int orc_estep(int n, long T, int M, int m, int update_rho,
              const double* x_white, double ldetS,
              const double* A, const double* c, const double* alpha,
              const double* mu, const double* beta, const double* rho,
              const double* gm, double* stats, int nthreads);

/* Same E-step computed in float32 arithmetic (statistics summed in FP64
 * across chunks). */
int orc_estep_f32(int n, long T, int M, int m, int update_rho,
                  const double* x_white, double ldetS,
                  const double* A, const double* c, const double* alpha,
                  const double* mu, const double* beta, const double* rho,
                  const double* gm, double* stats, int nthreads);

/* Full EM loop (amica.cu:820-2541).  Parameters are in/out.  Per-iteration
 * logs (each maxiter long, bflag/newton maxiter*M): ll, the lrate printed at
 * that iteration, bflag per model, whether the model took the Newton
 * branch, and the "Likelihood decreasing" flag.  W_last receives
 * pinv(A) computed at the start of the last iteration (the reference's
 * weights output).  ltall (optional, T*M, model-major) receives L_ht of the
 * last iteration.  precision: 64 = FP64 E-step, 32 = float32 E-step. */
int orc_run(const orc_config* cfg, const double* x_white, double ldetS,
            double* A, double* c, double* alpha, double* mu, double* beta,
            double* rho, double* gm, double* W_last,
            double* ll, double* lrate_log, int* bflag_log, int* newton_log,
            int* lldec_log, double* ltall, int precision, int nthreads);

/* Small FP64 linear algebra used by the oracle, exported for its own tests. */
int orc_pinv(int n, const double* A, double* W);             /* Jacobi SVD */
double orc_logabsdet(int n, const double* A);                /* LU, partial pivoting */
int orc_eigh(int n, const double* C, double* evals, double* evecs); /* cyclic Jacobi, ascending */
double orc_psi(double x);                                    /* digamma, asymptotic series */

/* Output transform of the reference (amica.cu:2564-2668): c_out = S^+ c + mean,
 * A_out = colnorm(S^+ A).  In place on A (n*n*M) and c (n*M). */
int orc_output_transform(int n, int M, const double* S, const double* mean,
                         double* A, double* c);

#ifdef __cplusplus
}
#endif

#endif
