/*
 * amica_synth.h -- deterministic synthetic recordings and initial states.
 *
 * No counterpart in the reference (its initial state comes from cuRAND,
 * amica.cu:272-394, which cannot be reproduced off NVIDIA hardware; the
 * reference's NORAND mode loads it from files instead, amica.cu:314-322).
 * Everything is drawn from xoshiro256** streams seeded through splitmix64
 * (Blackman & Vigna), so a seed gives the same bytes on every machine and
 * for any thread count.
 */
#ifndef AMICA_SYNTH_H_
#define AMICA_SYNTH_H_

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int    n;                  /* channels = sources */
    long   T;                  /* samples */
    int    nmix;               /* mixing matrices (1: stationary) */
    long   block;              /* samples per block when nmix > 1 */
    int    alternate;          /* 1: block b uses matrix b % nmix; 0: uniform draw */
    double rho_lo, rho_hi;     /* generalized-Gaussian shapes of the sources */
    double sub_frac;           /* fraction of sources with shape in [sub_lo, sub_hi] */
    double sub_lo, sub_hi;
    double cond_max;           /* redraw A until cond(A) <= cond_max (0: no limit) */
    double noise_std;          /* additive N(0, noise_std^2) channel noise */
    unsigned long long seed;
} amica_synth_spec;

/* x: float32 T x n sample-major (the .fdt layout).  A_true (nmix*n*n,
 * column-major) and rho_true (n) may be NULL.  nthreads <= 0: all cores. */
int amica_synth_data(const amica_synth_spec* spec, float* x, double* A_true, double* rho_true,
                     int nthreads);

/* Initial model state by the reference's recipe (amica.cu:272-394):
 * A_h = colnorm(I + 0.05 (U - 1/2)), c_h ~ N(0, 1), gm = 1/M, alpha = 1/m,
 * fix_init: mu_j = j - (m-1)/2, beta = 1; otherwise mu ~ N(0, 0.1^2)
 * (0 when m = 1), beta ~ N(1, 0.1^2); rho = rho0.  Layouts as in
 * amica_engine.h. */
int amica_synth_init(int n, int M, int m, unsigned long long seed, int fix_init, double rho0,
                     double* A, double* c, double* alpha, double* mu, double* beta, double* rho,
                     double* gm);

/* Condition number of an n x n column-major matrix (Jacobi eigenvalues of A^T A). */
double amica_synth_cond(int n, const double* A);

#ifdef __cplusplus
}
#endif

#endif
