/*
 * preprocess.h -- centering and sphering of the recording on the GPU
 * (fraimondo/cudamica include/preprocess.h:17-18, src/preprocess.cu).
 * Both act on gpus[0].engine after its raw data are loaded.  whiten()
 * writes the sphere matrix to currentConfig.spherefile, as the reference
 * does (preprocess.cu:287, 331).
 */
#ifndef PREPROCESS_H_
#define PREPROCESS_H_

#include "error.h"
#include "config.h"

#ifdef __cplusplus
extern "C" {
#endif

error center(void);
error whiten(void);

#ifdef __cplusplus
}
#endif

#endif /* PREPROCESS_H_ */
