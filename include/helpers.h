/*
 * helpers.h -- sample partition, launch tuning and matrix file I/O of the
 * AMICA host program (fraimondo/cudamica include/helpers.h:19-28,
 * src/helpers.cu).  The reference passes cudaPitchedPtr; the same four
 * fields (ptr, pitch in bytes, xsize elements per row, ysize rows) travel
 * in amica_pitched here.  A "Dev" matrix is a device pointer.
 * File formats: loadMatrix reads float32 (the .fdt data files, channels
 * fastest) or float64 when isdouble; the write* functions write raw
 * float64 rows of xsize elements.  open() failures are reported as
 * ERRORNOFILE (the reference tests fd == 0, helpers.cu:515) and short files
 * as ERRORIO.
 */
#ifndef HELPERS_H_
#define HELPERS_H_

#include <stddef.h>
#include "error.h"
#include "device.h"

typedef struct {
    void*  ptr;
    size_t pitch;    /* bytes between rows */
    size_t xsize;    /* elements per row */
    size_t ysize;    /* rows */
} amica_pitched;

#ifdef __cplusplus
extern "C" {
#endif

/* Block partition of the samples (helpers.cu:19-49).  The fused E-step has
 * no sample blocks of its own; block_size and blocks_per_gpu_per_model are
 * still computed for the reference's `blocks` table. */
error recalc_config(void);
/* The reference fills its heuristic launch-shape table here
 * (helpers.cu:55-511).  This engine sizes every grid at creation from the
 * occupancy of its kernels on the selected device, so there is nothing to
 * fill: kept for the reference's call sequence, returns SUCCESS. */
error autotune(natural device);
#ifdef __cplusplus
error loadMatrix(real* data, natural rows, natural elems, char* filename, bool isdouble = false);
error loadDevMatrix(amica_pitched ptr, char* filename, bool isdouble = false);
#else
error loadMatrix(real* data, natural rows, natural elems, char* filename, int isdouble);
error loadDevMatrix(amica_pitched ptr, char* filename, int isdouble);
#endif
error writeValue(real value, char* filename);
error writeMatrix(amica_pitched ptr, char* filename);
error writeDevMatrix(amica_pitched ptr, char* filename);
error writeDevMatrix3d(amica_pitched ptr, natural zsize, char* filename);
error loadDevMatrix3d(amica_pitched ptr, natural index, char* filename);

/* float32 T x n data straight into a float buffer (no FP64 copy). */
error loadMatrixF32(float* data, natural rows, natural elems, const char* filename);

#ifdef __cplusplus
}
#endif

#endif /* HELPERS_H_ */
