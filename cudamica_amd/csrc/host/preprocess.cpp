/*
 * preprocess.cpp -- center() and whiten() of the AMICA host program
 * (fraimondo/cudamica src/preprocess.cu:31-420), run by the engine of
 * gpus[0] on the GPU: FP64 means, FP64 covariance, FP64 Jacobi
 * eigen-decomposition (in place of magma_dsyevd_gpu + magma_dgesv_gpu),
 * S = 2 C^-1/2 (do_sphere 1), diag(1/std) (2) or I, ldetS = -log|det S|,
 * x <- S (x - mean).  whiten() writes S to spherefile like the reference.
 */
#include "../../../include/preprocess.h"

#include <cstdio>
#include <vector>

#include "../../../include/device.h"
#include "../../../include/helpers.h"

extern "C" {

error center(void) {
    if (!gpus[0].engine) return ERRORNODEVICEMEM;
    if (amica_engine_center(gpus[0].engine) != 0) {
        fprintf(stderr, "ERROR: centering failed: %s\n", amica_last_error());
        return ERRORNODEVICEMEM;
    }
    return SUCCESS;
}

error whiten(void) {
    if (!gpus[0].engine) return ERRORNODEVICEMEM;
    if (amica_engine_whiten(gpus[0].engine, (int)currentConfig.do_sphere) != 0) {
        fprintf(stderr, "ERROR: whitening failed: %s\n", amica_last_error());
        return ERRORNODEVICEMEM;
    }
    const natural n = currentConfig.nchannels;
    std::vector<double> S(n * n), mean(n);
    double ldetS = 0.0;
    if (amica_engine_get_sphere(gpus[0].engine, S.data(), mean.data(), &ldetS) != 0)
        return ERRORNODEVICEMEM;
    gpus[0].ldetS = ldetS;
    if ((currentConfig.do_sphere == 1 || currentConfig.do_sphere == 2) && currentConfig.spherefile) {
        amica_pitched p = {S.data(), n * sizeof(double), n, n};
        error e = writeMatrix(p, currentConfig.spherefile);
        if (e != SUCCESS) return e;
    }
    return SUCCESS;
}

}  // extern "C"
