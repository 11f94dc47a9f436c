/*
 * device.h -- GPU selection and per-device state of the AMICA host
 * program (fraimondo/cudamica include/device.h:236-315, src/device.cu).
 * One GPU owns all samples: the engine (include/amica_engine.h) holds the
 * data, the model and every work buffer of the device, replacing the
 * reference's per-device streams, cuBLAS/cuRAND handles and work arrays.
 */
#ifndef DEVICE_H_
#define DEVICE_H_

#include "config.h"
#include "error.h"
#include "amica_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int           realdevice;     /* HIP device ordinal */
    char          name[256];
    char          arch[64];       /* gcnArchName, e.g. gfx950 */
    int           cu_count;
    size_t        total_mem;
    amica_engine* engine;         /* created by initializeDevs */
    real          ldetS;
} device_t;

error getDevices(void);
error selectDevice(natural* deviceNums, natural count);
void  printCapabilities(int device);
error initializeDevs(void);
error finalizeDevs(void);

extern natural  gpuCount;
extern device_t gpus[MAX_DEVS];

#ifdef __cplusplus
}
#endif

#endif /* DEVICE_H_ */
