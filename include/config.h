/*
 * config.h -- run configuration of the AMICA host program.
 * Same struct fields, defaults and entry points as the reference
 * (fraimondo/cudamica include/config.h:46-168); the fields after
 * `maxdec` are additions of this engine and are filled from keys the
 * reference parser does not know (it skips unknown keys).
 */
#ifndef CONFIG_H_
#define CONFIG_H_

#include <stdio.h>
#include "tipos.h"

#define MAX_DEVS        8
#define MAX_CHANNELS    1024
#define MAX_MODELS      16

#define LOG2            0.693147180559945

#define DEFAULT_COMP_THRESH     0.99
#define DEFAULT_IDENT_INTVL     100
#define DEFAULT_UPDATE_GM       1
#define DEFAULT_UPDATE_ALPHA    1
#define DEFAULT_UPDATE_A        1
#define DEFAULT_UPDATE_MU       1
#define DEFAULT_UPDATE_BETA     1
#define DEFAULT_UPDATE_RHO      1
#define DEFAULT_DO_REPARM       1
#define DEFAULT_FIX_INIT        1
#define DEFAULT_DO_SPHERE       1
#define DEFAULT_MAXITER         200
#define DEFAULT_NSDM            3
#define DEFAULT_NMMODELS        1
#define DEFAULT_DO_NEWTON       1
#define DEFAULT_LRATE0          0.1
#define DEFAULT_LRATEMAX        1.0
#define DEFAULT_LNATRATE        0.1
#define DEFAULT_NEWT_START_ITER 50
#define DEFAULT_RHO_START_ITER  1
#define DEFAULT_LRATEFACT       0.5
#define DEFAULT_BLOCK_SIZE      5000
#define DEFAULT_RHO0            1.5
#define DEFAULT_RHOLRATE        0.1
#define DEFAULT_RHOMIN          1
#define DEFAULT_RHOMAX          2
#define DEFAULT_SHOWLL          1
#define DEFAULT_NUMDEC          0
#define DEFAULT_MAXDEC          3
/* additions */
#define DEFAULT_PRECISION       32      /* 32: FP32, 16: BF16 GEMM operands */
#define DEFAULT_SEED            1234    /* the reference seeds cuRAND with 1234 + device */

typedef struct config_t {
    natural devs[MAX_DEVS];
    natural devcount;

    char*   datafile;
    natural nchannels;
    natural nsamples;
    natural nmmodels;
    natural nsdm;
    natural maxiter;
    natural do_sphere;
    natural do_newton;
    char*   ainitfile;      /* float64 n x n x M, column-major (initial A) */
    char*   cinitfile;      /* float64 n x M (initial c) */
    char*   kinitfile;      /* unused, as in the reference */

    char*   spherefile;
    char*   afile;
    char*   weightsfile;
    char*   khindsfile;
    char*   cfile;
    char*   llfile;
    char*   ltallfile;
    char*   gmfile;
    char*   alphafile;
    char*   mufile;
    char*   betafile;
    char*   rhofile;

    real    comp_thresh;
    natural ident_intvl;

    natural update_gm;
    natural update_alpha;
    natural update_A;
    natural update_mu;
    natural update_beta;
    natural update_rho;

    natural do_reparm;
    natural fix_init;

    real    lrate0;
    real    lratemax;
    real    lnatrate;
    natural newt_start_iter;
    natural rho_start_iter;
    real    lratefact;
    natural block_size;

    natural blocks_per_gpu_per_model;

    real    rho0;
    real    rholrate;
    real    rhomin;
    real    rhomax;

    natural showLL;

    natural numdec;
    natural maxdec;

    /* additions of this engine */
    natural precision;      /* key `precision`: fp32 (32), fp32plain (33), bf16 (16) or fp8 (8) */
    natural seed;           /* key `seed`: PRNG seed of the initial state */
    char*   jsonlogfile;    /* key `jsonlogfile`: one JSON line per iteration, or NULL */
} config;

#ifdef __cplusplus
extern "C" {
#endif
    error   getConfig(char*);
    void    printConfig(void);
    error   freeConfig(void);
    void    help(void);

    extern config currentConfig;
    extern char*  programName;
#ifdef __cplusplus
}
#endif

#endif /* CONFIG_H_ */
