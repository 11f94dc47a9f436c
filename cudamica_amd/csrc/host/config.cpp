/*
 * config.cpp -- run configuration from the parameter file, with the keys,
 * defaults, mandatory keys and messages of fraimondo/cudamica
 * src/config.c:12-267.  One table describes every key (how it is read,
 * what its absence means); getConfig walks it.
 *
 * Kept as the reference has it: a missing `update_rho` key sets do_newton
 * to DEFAULT_UPDATE_RHO and leaves update_rho at 0 (config.c:146-153), so a
 * file without `update_rho` runs with rho fixed.  Keys this engine adds
 * (unknown to the reference parser, which ignores them): `precision`
 * (fp32 | fp32plain | bf16 | fp8), `seed` (initial-state PRNG), `jsonlogfile` (a JSON
 * line per iteration), `ainitfile` / `cinitfile`
 * (float64 initial A and c; the reference declares these fields but forces
 * them to NULL, config.c:131-134).
 */
#include "../../../include/config.h"

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../../include/params.h"

namespace {

/* How a key is read and what its absence means. */
enum KeyKind {
    kRequiredString,   /* absent: error */
    kRequiredNatural,  /* absent: error */
    kNatural,          /* absent: the default */
    kBool,             /* absent: the default */
    kFile,             /* absent: NULL */
    kRhoPolicy         /* update_rho, kept as the reference has it: its absence
                        * sets do_newton to the default and leaves update_rho at
                        * 0 (config.c:146-153) */
};

struct KeySpec {
    const char* key;
    KeyKind kind;
    size_t offset;      /* of the config field */
    natural fallback;   /* kNatural / kBool */
    const char* message;
};

#define AMICA_FIELD(f) offsetof(config, f)

/* Keys in the order their errors are reported (config.c:27-176); the
 * messages are the reference's, so scripts that match them keep working. */
const KeySpec kKeys[] = {
    {"datafile", kRequiredString, AMICA_FIELD(datafile), 0, "Invalid data file (datafile)"},
    {"nchannels", kRequiredNatural, AMICA_FIELD(nchannels), 0, "Invalid number of channels (nchannels)"},
    {"nsamples", kRequiredNatural, AMICA_FIELD(nsamples), 0, "Invalid number of samples (nsamples)"},
    {"do_sphere", kNatural, AMICA_FIELD(do_sphere), DEFAULT_DO_SPHERE, "Invalid sphering value (do_sphere)"},
    {"maxiter", kNatural, AMICA_FIELD(maxiter), DEFAULT_MAXITER,
     "Invalid number of maximum iterations (maxiter)"},
    {"nsdm", kNatural, AMICA_FIELD(nsdm), DEFAULT_NSDM, "Invalid number of source density modles (nsdm)"},
    {"nmmodels", kNatural, AMICA_FIELD(nmmodels), DEFAULT_NMMODELS,
     "Invalid number of ICA mixture models (nmmodels)"},
    {"spherefile", kRequiredString, AMICA_FIELD(spherefile), 0, "Invalid sphere file (spherefile)"},
    {"block_size", kNatural, AMICA_FIELD(block_size), DEFAULT_BLOCK_SIZE, "Invalid block size (block_size)"},
    {"fix_init", kBool, AMICA_FIELD(fix_init), DEFAULT_FIX_INIT, "Invalid fix_init value (fix_init)"},
    {"afile", kFile, AMICA_FIELD(afile), 0, "Invalid file (afile)"},
    {"weightsfile", kFile, AMICA_FIELD(weightsfile), 0, "Invalid file (weightsfile)"},
    {"khindsfile", kFile, AMICA_FIELD(khindsfile), 0, "Invalid file (khindsfile)"},
    {"cfile", kFile, AMICA_FIELD(cfile), 0, "Invalid file (cfile)"},
    {"llfile", kFile, AMICA_FIELD(llfile), 0, "Invalid file (llfile)"},
    {"ltallfile", kFile, AMICA_FIELD(ltallfile), 0, "Invalid file (ltallfile)"},
    {"gmfile", kFile, AMICA_FIELD(gmfile), 0, "Invalid file (gmfile)"},
    {"alphafile", kFile, AMICA_FIELD(alphafile), 0, "Invalid file (alphafile)"},
    {"mufile", kFile, AMICA_FIELD(mufile), 0, "Invalid file (mufile)"},
    {"betafile", kFile, AMICA_FIELD(betafile), 0, "Invalid file (betafile)"},
    {"rhofile", kFile, AMICA_FIELD(rhofile), 0, "Invalid file (rhofile)"},
    {"ainitfile", kFile, AMICA_FIELD(ainitfile), 0, "Invalid file (ainitfile)"},
    {"cinitfile", kFile, AMICA_FIELD(cinitfile), 0, "Invalid file (cinitfile)"},
    {"jsonlogfile", kFile, AMICA_FIELD(jsonlogfile), 0, "Invalid file (jsonlogfile)"},
    {"do_newton", kBool, AMICA_FIELD(do_newton), DEFAULT_DO_NEWTON, "Invalid update policy (doNewton)"},
    {"update_rho", kRhoPolicy, AMICA_FIELD(update_rho), DEFAULT_UPDATE_RHO,
     "Invalid rho update policy (update_rho)"},
    {"rho_start_iter", kNatural, AMICA_FIELD(rho_start_iter), DEFAULT_RHO_START_ITER,
     "Invalid number of iterations to update rho (rho_start_iter)"},
    {"newt_start_iter", kNatural, AMICA_FIELD(newt_start_iter), DEFAULT_NEWT_START_ITER,
     "Invalid number of iterations to do newton (newt_start_iter)"},
    {"seed", kNatural, AMICA_FIELD(seed), DEFAULT_SEED, "Invalid seed (seed)"},
};

#undef AMICA_FIELD

/* Reads one key into currentConfig; returns SUCCESS or the parser's error
 * (after printing the key's message). */
error read_key(parsedconfigs* pc, const KeySpec& k) {
    char* field = (char*)&currentConfig + k.offset;
    error rc;
    switch (k.kind) {
        case kRequiredString:
        case kFile: rc = getString(pc, (char*)k.key, (char**)field); break;
        case kBool: rc = getBool(pc, (char*)k.key, (natural*)field); break;
        default: rc = getNatural(pc, (char*)k.key, (natural*)field); break;
    }
    if (rc == SUCCESS) return SUCCESS;
    const bool optional = k.kind == kNatural || k.kind == kBool || k.kind == kFile;
    if (rc == ERRORNOPARAM && optional) {
        if (k.kind == kFile) *(char**)field = NULL;
        else *(natural*)field = k.fallback;
        return SUCCESS;
    }
    if (rc == ERRORNOPARAM && k.kind == kRhoPolicy) {
        currentConfig.do_newton = k.fallback;
        return SUCCESS;
    }
    fprintf(stderr, "ERROR: %s\n", k.message);
    return rc;
}

/* `devices`: a list of ordinals (at most 64), default device 0. */
error read_devices(parsedconfigs* pc) {
    natural list[64];
    natural count = 64;                 /* the capacity of list */
    const error rc = getNaturalList(pc, (char*)"devices", list, &count);
    if (rc == ERRORNOPARAM) {
        currentConfig.devcount = 1;
        currentConfig.devs[0] = 0;
        return SUCCESS;
    }
    if (rc != SUCCESS) {
        fprintf(stderr, "ERROR: Invalid devices specification (devices)\n");
        return rc;
    }
    if (count > MAX_DEVS) count = MAX_DEVS;
    currentConfig.devcount = count;
    for (natural i = 0; i < count; ++i) currentConfig.devs[i] = list[i];
    return SUCCESS;
}

/* `precision`: fp32 | fp32plain | bf16 | fp8 (or 32 | 33 | 16 | 8). */
error read_precision(parsedconfigs* pc) {
    char* text = NULL;
    const error rc = getString(pc, (char*)"precision", &text);
    if (rc == ERRORNOPARAM) {
        currentConfig.precision = DEFAULT_PRECISION;
        return SUCCESS;
    }
    if (rc != SUCCESS) {
        fprintf(stderr, "ERROR: Invalid precision (precision)\n");
        return rc;
    }
    static const struct { const char* a; const char* b; natural bits; } kNames[] = {
        {"fp32", "32", 32}, {"fp32plain", "33", 33}, {"bf16", "16", 16}, {"fp8", "8", 8}};
    error out = ERRORINVALIDPARAM;
    for (const auto& e : kNames)
        if (!strcmp(text, e.a) || !strcmp(text, e.b)) {
            currentConfig.precision = e.bits;
            out = SUCCESS;
        }
    if (out != SUCCESS) fprintf(stderr, "ERROR: Invalid precision (precision): %s\n", text);
    free(text);
    return out;
}

/* The settings the parameter file cannot change (config.c:181-210). */
void set_fixed_defaults() {
    currentConfig.kinitfile = NULL;
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
}

}  // namespace

extern "C" {

config currentConfig;
char* programName;

/* Every key is read and every bad one reported; the result is the last
 * error (SUCCESS when all keys were good). */
error getConfig(char* filename) {
    std::memset(&currentConfig, 0, sizeof(currentConfig));
    parsedconfigs* pc = NULL;
    error result = parseConfig(filename, &pc);
    if (result == SUCCESS) {
        auto note = [&result](error rc) {
            if (rc != SUCCESS) result = rc;
        };
        note(read_devices(pc));
        for (const KeySpec& k : kKeys) note(read_key(pc, k));
        note(read_precision(pc));
        freeParsedConfig(pc);
    }
    set_fixed_defaults();
    return result;
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
                                      : currentConfig.precision == 8 ? "fp8"
                                      : currentConfig.precision == 33 ? "fp32plain" : "fp32");
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
