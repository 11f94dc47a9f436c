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

/* Intermediates of one M-step, each optional (NULL: not written):
 * kappa, lambda, sigma2, cnew [M*n]; phi, B [M*n*n] column-major.
 * phi is Phi_h / vsum_h; B is the Newton matrix or I - phi of each model
 * (written only when update_A is on). */
typedef struct {
    double *kappa, *lambda, *sigma2, *cnew, *phi, *B;
} orc_mstep_trace;

/* One M-step (amica.cu:1963-2512) from canonical statistics (as orc_estep
 * returns them) at 1-based iteration `iter`.  W is pinv(A) of this
 * iteration's E-step (it shifts mu in the reparameterisation).  ll_hist
 * holds the log-likelihoods of iterations iter-1 and iter (ll_hist[0] is
 * ignored when iter == 1).  Parameters and *lrate are in/out; bflag and
 * newton (M each) and *lldec receive the decisions.  tr may be NULL. */
int orc_mstep(const orc_config* cfg, int iter, const double* stats, const double* W,
              const double* ll_hist, double* A, double* c, double* alpha, double* mu,
              double* beta, double* rho, double* gm, double* lrate, int* bflag, int* newton,
              int* lldec, orc_mstep_trace* tr);

/* One full EM iteration (E-step, LL, M-step) at 1-based iteration `iter`
 * from the given state, lrate and the previous iteration's log-likelihood
 * ll_prev (unused when iter == 1): for teacher-forced comparisons that
 * start every step from another implementation's state.  Outputs as
 * orc_run's for one iteration; W receives pinv(A) of this iteration.
 * precision: 64 or 32 as in orc_run. */
int orc_step(const orc_config* cfg, int iter, const double* x_white, double ldetS,
             double* A, double* c, double* alpha, double* mu, double* beta, double* rho,
             double* gm, double* lrate, double ll_prev, double* W, double* ll, int* bflag,
             int* newton, int* lldec, int precision, int nthreads);

/* Test-only: plant one known mistake in the oracle's equations (0 removes
 * it), so that the tests can show that each of their checks fails on a
 * plausible error.  Never set outside those tests. */
enum {
    ORC_PLANT_NONE = 0,
    ORC_PLANT_KAPPA_NUMER_NO_BETA = 1,   /* kappa_numer without the beta factor */
    ORC_PLANT_LAMBDA_NUMER_PLUS = 2,     /* (f y + 1)^2 in lambda_numer */
    ORC_PLANT_MU_DENOM_NO_SQRT_BETA = 3, /* mu_denom without sqrt(beta) */
    ORC_PLANT_RHO_NUMER_NO_LOG = 4,      /* rho_numer = sum u |y|^rho */
    ORC_PLANT_SIGMA2_FROM_X = 5,         /* sigma2_numer = sum v x^2 */
    ORC_PLANT_CNEW_NO_V = 6,             /* cnew = sum x */
    ORC_PLANT_G_NO_SQRT_BETA = 7,        /* g = sum_j u f */
    ORC_PLANT_KAPPA_OLD_ALPHA = 8,       /* kappa, lambda weighted by the old alpha */
    ORC_PLANT_LAMBDA_NO_MU = 9,          /* lambda without mu^2 dk */
    ORC_PLANT_BETA_SQRT_RATIO = 10,      /* beta *= sqrt(beta_numer / beta_denom) */
    ORC_PLANT_MU_STEP_SIGN = 11,         /* mu -= mu_numer / mu_denom */
    ORC_PLANT_RHO_STEP_SIGN = 12,        /* rho step 1 + (rho / psi) R / usum */
    ORC_PLANT_RHO_PSI_ARG = 13,          /* psi(1 / rho) instead of psi(1 + 1 / rho) */
    ORC_PLANT_NEWTON_TRANSPOSE = 14,     /* phi_ik and phi_ki swapped in the Newton B */
    ORC_PLANT_AB_ORDER = 15,             /* A + lrate B A instead of A + lrate A B */
    ORC_PLANT_NATGRAD_SIGN = 16,         /* natural gradient A + lrate A (I - phi) */
    ORC_PLANT_REPARAM_BETA_NU = 17,      /* beta / nu instead of beta / nu^2 */
    ORC_PLANT_MU_SHIFT_WT = 18,          /* mu shift with W^T */
    ORC_PLANT_LRATE_GROWTH = 19,         /* lrate + 0.1 instead of lrate + min(0.1, lrate) */
    ORC_PLANT_LLDEC_NO_HALVE = 20,       /* no lratefact on a log-likelihood decrease */
    ORC_PLANT_BFLAG_NO_DIAG = 21,        /* the Newton test skips i == k */
    ORC_PLANT_ALPHA_OVER_T = 22,         /* alpha = usum / T */
    ORC_PLANT_PHI_NO_NORM = 23           /* phi not divided by vsum */
};
void orc_plant(int code);

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
