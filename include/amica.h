/*
 * amica.h -- the AMICA EM driver entry point (fraimondo/cudamica
 * include/amica.h:119-123).  runamica() reads currentConfig, loads the
 * data file, preprocesses it, runs maxiter EM iterations on the GPU and
 * writes the configured output files in the reference's formats.
 *
 * `blocks`/`nblocks` describe the sample partition, as in the reference;
 * the engine streams all samples through one fused kernel, so the table is
 * informational (start, end and size of each block of block_size samples).
 */
#ifndef AMICA_H_
#define AMICA_H_

#include "error.h"
#include "helpers.h"
#include "device.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    natural device;
    natural models;
    natural start;
    natural end;
    natural block_size;
    real    ll;
} model_block_t;

extern model_block_t* blocks;
extern natural        nblocks;

error runamica(void);

#ifdef __cplusplus
}
#endif

#endif /* AMICA_H_ */
