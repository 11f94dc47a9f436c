/*
 * amica_batch.h -- many recordings per EM iteration on one GPU (C ABI,
 * libcudamica_amd).
 *
 * No counterpart in the reference: fraimondo/cudamica runs one recording
 * per process and sizes its launches with a fixed heuristic table
 * (src/helpers.cu:55-511, autotune(); include/device.h:22-228).  A batch
 * holds one engine per recording (any mix of n, T, M, m and precision); one
 * EM iteration of every recording is a single hipGraph with one branch per
 * recording, so the recordings' kernels run side by side.  The grids of
 * each recording are sized for a budget of CUs, picked by a measured tuner.
 *
 * Use: amica_batch_create; for every recording amica_batch_engine, then the
 * engine calls of amica_engine.h (load_raw + preprocess or load_white,
 * set_state); optionally amica_batch_tune; amica_batch_iterate;
 * amica_batch_sync; per recording amica_engine_get_log / get_state /
 * output.  Returns 0 or a negative AMICA_E_* code (amica_last_error()).
 */
#ifndef AMICA_BATCH_H_
#define AMICA_BATCH_H_

#include "amica_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct amica_batch amica_batch;

/* Policies the tuner measures.  p < 8: p = 4 * launch mode + CU budget.
 * Launch modes: 0, the whole batch as one hipGraph (one branch per
 * recording); 1, every recording replays its own one-iteration graph on its
 * own stream (the streams join at the end of amica_batch_iterate).
 * CU budgets: every recording sized for the whole GPU; budgets proportional
 * to each recording's E-step work (M m n T); twice that; half the GPU each.
 * p = 8 .. 11: one hipGraph in which the unmixing GEMM, the E-step and the
 * statistic GEMM of all single-model BF16 recordings of one channel width
 * run as one launch each over a work list of (recording, sample range) items: each recording's own grid
 * partition (p = 8, results identical to the recording's own launches), or
 * items of 512, 1024, 2048 samples (the tile shape the tuner picks); the
 * per-recording inverse, reductions and M-step are graph branches; other
 * recordings run as in mode 0. */
#define AMICA_BATCH_POLICIES 12

/* nrec recordings; opts[r] as for amica_engine_create (same device). */
int amica_batch_create(const amica_options* opts, int nrec, amica_batch** out);
int amica_batch_destroy(amica_batch* b);
int amica_batch_size(amica_batch* b);
/* The engine of recording r (owned by the batch). */
int amica_batch_engine(amica_batch* b, int r, amica_engine** eng);
/* Grids of recording r sized for cu_budget[r] CUs (NULL: the whole GPU). */
int amica_batch_set_budgets(amica_batch* b, const int* cu_budget);
/* Applies policy p (0 .. AMICA_BATCH_POLICIES-1). */
int amica_batch_set_policy(amica_batch* b, int p);
/* Times `iters` batched iterations under every policy (after one untimed
 * iteration each), applies the fastest and returns it in *policy; ms[p]
 * receives the mean milliseconds per batched iteration (may be NULL).
 * Call after every recording has data and a state (set_state) and before
 * iterating: the tuner restores the state each recording had on entry. */
int amica_batch_tune(amica_batch* b, int iters, int* policy, double* ms);
/* k EM iterations of every recording. */
int amica_batch_iterate(amica_batch* b, int k);
int amica_batch_sync(amica_batch* b);

#ifdef __cplusplus
}
#endif

#endif
