/*
 * config.cpp -- run configuration from the parameter file, with the keys,
 * defaults, mandatory keys and messages of fraimondo/cudamica
 * src/config.c:12-267.
 *
 * Kept as the reference has it: a missing `update_rho` key sets do_newton
 * to DEFAULT_UPDATE_RHO and leaves update_rho at 0 (config.c:146-153), so a
 * file without `update_rho` runs with rho fixed.  Keys this engine adds
 * (unknown to the reference parser, which ignores them): `precision`
 * (fp32 | bf16 | fp8), `seed` (initial-state PRNG), `jsonlogfile` (a JSON
 * line per iteration), `ainitfile` / `cinitfile`
 * (float64 initial A and c; the reference declares these fields but forces
 * them to NULL, config.c:131-134).
 */
#include "../../../include/config.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../../include/params.h"

extern "C" {

config currentConfig;
char* programName;

error getConfig(char* filename) {
    error retorno = SUCCESS;
    error newerr = SUCCESS;
    parsedconfigs* pc = NULL;
    std::memset(&currentConfig, 0, sizeof(currentConfig));

    retorno = parseConfig(filename, &pc);
    if (retorno == SUCCESS) {
        natural devs[64];
        natural ndev = 0;
        if ((newerr = getNaturalList(pc, (char*)"devices", devs, &ndev)) != SUCCESS) {
            if (newerr == ERRORNOPARAM) {
                currentConfig.devcount = 1;
                currentConfig.devs[0] = 0;
            } else {
                fprintf(stderr, "ERROR: Invalid devices specification (devices)\n");
                retorno = newerr;
            }
        } else {
            if (ndev > MAX_DEVS) ndev = MAX_DEVS;
            currentConfig.devcount = ndev;
            for (natural i = 0; i < ndev; ++i) currentConfig.devs[i] = devs[i];
        }

        if ((newerr = getString(pc, (char*)"datafile", &currentConfig.datafile)) != SUCCESS) {
            fprintf(stderr, "ERROR: Invalid data file (datafile)\n");
            retorno = newerr;
        }
        if ((newerr = getNatural(pc, (char*)"nchannels", &currentConfig.nchannels)) != SUCCESS) {
            fprintf(stderr, "ERROR: Invalid number of channels (nchannels)\n");
            retorno = newerr;
        }
        if ((newerr = getNatural(pc, (char*)"nsamples", &currentConfig.nsamples)) != SUCCESS) {
            fprintf(stderr, "ERROR: Invalid number of samples (nsamples)\n");
            retorno = newerr;
        }

#define OPT_NATURAL(key, field, def, msg)                                            \
        if ((newerr = getNatural(pc, (char*)key, &currentConfig.field)) != SUCCESS) { \
            if (newerr == ERRORNOPARAM) {                                             \
                currentConfig.field = def;                                            \
            } else {                                                                  \
                fprintf(stderr, "ERROR: %s\n", msg);                                  \
                retorno = newerr;                                                     \
            }                                                                         \
        }
#define OPT_BOOL(key, field, def, msg)                                               \
        if ((newerr = getBool(pc, (char*)key, &currentConfig.field)) != SUCCESS) {    \
            if (newerr == ERRORNOPARAM) {                                             \
                currentConfig.field = def;                                            \
            } else {                                                                  \
                fprintf(stderr, "ERROR: %s\n", msg);                                  \
                retorno = newerr;                                                     \
            }                                                                         \
        }

        OPT_NATURAL("do_sphere", do_sphere, DEFAULT_DO_SPHERE, "Invalid sphering value (do_sphere)")
        OPT_NATURAL("maxiter", maxiter, DEFAULT_MAXITER, "Invalid number of maximum iterations (maxiter)")
        OPT_NATURAL("nsdm", nsdm, DEFAULT_NSDM, "Invalid number of source density modles (nsdm)")
        OPT_NATURAL("nmmodels", nmmodels, DEFAULT_NMMODELS, "Invalid number of ICA mixture models (nmmodels)")

        if ((newerr = getString(pc, (char*)"spherefile", &currentConfig.spherefile)) != SUCCESS) {
            fprintf(stderr, "ERROR: Invalid sphere file (spherefile)\n");
            retorno = newerr;
        }

        OPT_NATURAL("block_size", block_size, DEFAULT_BLOCK_SIZE, "Invalid block size (block_size)")
        OPT_BOOL("fix_init", fix_init, DEFAULT_FIX_INIT, "Invalid fix_init value (fix_init)")

#define PARSE_FILE(file)                                                             \
        if ((newerr = getString(pc, (char*)#file, &currentConfig.file)) != SUCCESS) {  \
            if (newerr == ERRORNOPARAM) {                                             \
                currentConfig.file = NULL;                                            \
            } else {                                                                  \
                fprintf(stderr, "ERROR: Invalid file (" #file ")\n");                 \
                retorno = newerr;                                                     \
            }                                                                         \
        }
        PARSE_FILE(afile)
        PARSE_FILE(weightsfile)
        PARSE_FILE(khindsfile)
        PARSE_FILE(cfile)
        PARSE_FILE(llfile)
        PARSE_FILE(ltallfile)
        PARSE_FILE(gmfile)
        PARSE_FILE(alphafile)
        PARSE_FILE(mufile)
        PARSE_FILE(betafile)
        PARSE_FILE(rhofile)
        PARSE_FILE(ainitfile)
        PARSE_FILE(cinitfile)
        PARSE_FILE(jsonlogfile)
        currentConfig.kinitfile = NULL;

        OPT_BOOL("do_newton", do_newton, DEFAULT_DO_NEWTON, "Invalid update policy (doNewton)")

        if ((newerr = getNatural(pc, (char*)"update_rho", &currentConfig.update_rho)) != SUCCESS) {
            if (newerr == ERRORNOPARAM) {
                currentConfig.do_newton = DEFAULT_UPDATE_RHO;   /* as config.c:149 */
            } else {
                fprintf(stderr, "ERROR: Invalid rho update policy (update_rho)\n");
                retorno = newerr;
            }
        }

        OPT_NATURAL("rho_start_iter", rho_start_iter, DEFAULT_RHO_START_ITER,
                    "Invalid number of iterations to update rho (rho_start_iter)")
        OPT_NATURAL("newt_start_iter", newt_start_iter, DEFAULT_NEWT_START_ITER,
                    "Invalid number of iterations to do newton (newt_start_iter)")
        OPT_NATURAL("seed", seed, DEFAULT_SEED, "Invalid seed (seed)")

        char* prec = NULL;
        if ((newerr = getString(pc, (char*)"precision", &prec)) == SUCCESS) {
            if (!strcmp(prec, "fp32") || !strcmp(prec, "32")) currentConfig.precision = 32;
            else if (!strcmp(prec, "bf16") || !strcmp(prec, "16")) currentConfig.precision = 16;
            else if (!strcmp(prec, "fp8") || !strcmp(prec, "8")) currentConfig.precision = 8;
            else {
                fprintf(stderr, "ERROR: Invalid precision (precision): %s\n", prec);
                retorno = ERRORINVALIDPARAM;
            }
            free(prec);
        } else if (newerr == ERRORNOPARAM) {
            currentConfig.precision = DEFAULT_PRECISION;
        } else {
            fprintf(stderr, "ERROR: Invalid precision (precision)\n");
            retorno = newerr;
        }
#undef OPT_NATURAL
#undef OPT_BOOL
#undef PARSE_FILE
        freeParsedConfig(pc);
    }

    /* fixed defaults (config.c:181-210) */
    currentConfig.comp_thresh = DEFAULT_COMP_THRESH;
    currentConfig.ident_intvl = DEFAULT_IDENT_INTVL;
    currentConfig.update_gm = DEFAULT_UPDATE_GM;
    currentConfig.update_alpha = DEFAULT_UPDATE_ALPHA;
    currentConfig.update_A = DEFAULT_UPDATE_A;
    currentConfig.update_mu = DEFAULT_UPDATE_MU;
    currentConfig.update_beta = DEFAULT_UPDATE_BETA;
    currentConfig.do_reparm = DEFAULT_DO_REPARM;
    currentConfig.lrate0 = DEFAULT_LRATE0;
    currentConfig.lratemax = DEFAULT_LRATEMAX;
    currentConfig.lnatrate = DEFAULT_LNATRATE;
    currentConfig.lratefact = DEFAULT_LRATEFACT;
    currentConfig.rho0 = DEFAULT_RHO0;
    currentConfig.rholrate = DEFAULT_RHOLRATE;
    currentConfig.rhomin = DEFAULT_RHOMIN;
    currentConfig.rhomax = DEFAULT_RHOMAX;
    currentConfig.showLL = DEFAULT_SHOWLL;
    currentConfig.numdec = DEFAULT_NUMDEC;
    currentConfig.maxdec = DEFAULT_MAXDEC;
    currentConfig.blocks_per_gpu_per_model = (natural)-1;
    return retorno;
}

void printConfig(void) {
    fprintf(stdout, "====================================\n");
    fprintf(stdout, "Configuration parsed\n");
    fprintf(stdout, "====================================\n");
    fprintf(stdout, "devices ");
    for (natural i = 0; i < currentConfig.devcount; i++) fprintf(stdout, "%lu ", currentConfig.devs[i]);
    fprintf(stdout, "(%lu)\n", currentConfig.devcount);
    fprintf(stdout, "datafile %s\n", currentConfig.datafile ? currentConfig.datafile : "(null)");
    fprintf(stdout, "nchannels %lu\n", currentConfig.nchannels);
    fprintf(stdout, "nsamples %lu\n", currentConfig.nsamples);
    fprintf(stdout, "do_sphere %lu\n", currentConfig.do_sphere);
    fprintf(stdout, "spherefile %s\n", currentConfig.spherefile ? currentConfig.spherefile : "(null)");
    fprintf(stdout, "weightsfile %s\n", currentConfig.weightsfile ? currentConfig.weightsfile : "(null)");
    fprintf(stdout, "maxiter %lu\n", currentConfig.maxiter);
    fprintf(stdout, "nsdm %lu\n", currentConfig.nsdm);
    fprintf(stdout, "nmmodels %lu\n", currentConfig.nmmodels);
    fprintf(stdout, "do_newton %lu\n", currentConfig.do_newton);
    fprintf(stdout, "fix_init %lu\n", currentConfig.fix_init);
    fprintf(stdout, "block_size %lu\n", currentConfig.block_size);
    fprintf(stdout, "precision %s\n", currentConfig.precision == 16 ? "bf16"
                                      : currentConfig.precision == 8 ? "fp8" : "fp32");
    fprintf(stdout, "====================================\n\n");
}

#define SAFE_FREE(p) do { if (p) { free(p); (p) = NULL; } } while (0)

error freeConfig(void) {
    SAFE_FREE(currentConfig.datafile);
    SAFE_FREE(currentConfig.spherefile);
    SAFE_FREE(currentConfig.afile);
    SAFE_FREE(currentConfig.weightsfile);
    SAFE_FREE(currentConfig.khindsfile);
    SAFE_FREE(currentConfig.cfile);
    SAFE_FREE(currentConfig.llfile);
    SAFE_FREE(currentConfig.ltallfile);
    SAFE_FREE(currentConfig.gmfile);
    SAFE_FREE(currentConfig.alphafile);
    SAFE_FREE(currentConfig.mufile);
    SAFE_FREE(currentConfig.betafile);
    SAFE_FREE(currentConfig.rhofile);
    SAFE_FREE(currentConfig.ainitfile);
    SAFE_FREE(currentConfig.jsonlogfile);
    SAFE_FREE(currentConfig.cinitfile);
    SAFE_FREE(currentConfig.kinitfile);
    return SUCCESS;
}

void help(void) {
    printf("usage: %s -f <parameter file>\n", programName ? programName : "cudaamica");
    printf("  required keys: datafile nchannels nsamples spherefile\n");
    printf("  optional keys: devices do_sphere maxiter nsdm nmmodels block_size fix_init\n");
    printf("                 do_newton update_rho rho_start_iter newt_start_iter\n");
    printf("                 afile weightsfile cfile llfile ltallfile gmfile alphafile\n");
    printf("                 mufile betafile rhofile ainitfile cinitfile precision seed jsonlogfile\n");
}

static const char* kErrors[] = {"Success!",     "No device memory", "No parameter",
                                "Invalid parameter", "Invalid config", "No file",
                                "Input/output error", "No host memory"};

const char* cGetErrorString(natural value) {
    long v = -(long)value;
    if (v < 0 || v > 7) return "Unknown error";
    return kErrors[v];
}

}  // extern "C"
