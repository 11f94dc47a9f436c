/*
 * device.hip -- GPU discovery and selection, engine set-up and the
 * device-matrix file helpers of the AMICA host program (fraimondo/cudamica
 * src/device.cu:17-426, src/helpers.cu:589-747).
 *
 * One GPU owns all samples: selectDevice keeps the first listed device
 * (the reference replicates the data on every listed device and splits the
 * sample blocks round-robin, amica.cu:406-421, 744-789).
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "../../../include/amica.h"
#include "../../../include/config.h"
#include "../../../include/device.h"
#include "../../../include/helpers.h"

extern "C" {

natural gpuCount = 0;
device_t gpus[MAX_DEVS];

void printCapabilities(int device) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return;
    fprintf(stdout, "Device %d: %s (%s)\n", device, p.name, p.gcnArchName);
    fprintf(stdout, "  compute units       %d\n", p.multiProcessorCount);
    fprintf(stdout, "  global memory       %zu MB\n", (size_t)(p.totalGlobalMem >> 20));
    fprintf(stdout, "  LDS per workgroup   %zu KB\n", (size_t)(p.sharedMemPerBlock >> 10));
    fprintf(stdout, "  wavefront size      %d\n", p.warpSize);
    fprintf(stdout, "  clock               %d MHz\n", p.clockRate / 1000);
}

error getDevices(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        fprintf(stderr, "ERROR: no HIP device found\n");
        return ERRORNODEVICEMEM;
    }
    fprintf(stdout, "====================================\n");
    fprintf(stdout, "%d HIP device(s) found\n", count);
    for (int d = 0; d < count; ++d) printCapabilities(d);
    fprintf(stdout, "====================================\n\n");
    return SUCCESS;
}

error selectDevice(natural* deviceNums, natural count) {
    int avail = 0;
    if (hipGetDeviceCount(&avail) != hipSuccess) return ERRORNODEVICEMEM;
    if (count == 0) return ERRORINVALIDPARAM;
    for (natural i = 0; i < count; ++i)
        if ((int)deviceNums[i] >= avail) {
            fprintf(stderr, "ERROR: device %lu does not exist\n", deviceNums[i]);
            return ERRORINVALIDPARAM;
        }
    if (count > 1)
        fprintf(stdout, "One GPU owns all samples: using device %lu of the %lu listed\n",
                deviceNums[0], count);
    std::memset(gpus, 0, sizeof(gpus));
    gpuCount = 1;
    device_t& g = gpus[0];
    g.realdevice = (int)deviceNums[0];
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, g.realdevice) != hipSuccess) return ERRORNODEVICEMEM;
    std::snprintf(g.name, sizeof(g.name), "%s", p.name);
    std::snprintf(g.arch, sizeof(g.arch), "%s", p.gcnArchName);
    g.cu_count = p.multiProcessorCount;
    g.total_mem = p.totalGlobalMem;
    return SUCCESS;
}

error autotune(natural device) {
    (void)device;
    return SUCCESS;
}

error initializeDevs(void) {
    if (gpuCount == 0) return ERRORNODEVICEMEM;
    amica_options o;
    amica_options_default(&o, (int)currentConfig.nchannels, (long)currentConfig.nsamples,
                          (int)currentConfig.nmmodels, (int)currentConfig.nsdm);
    o.precision = currentConfig.precision == 16 ? AMICA_PREC_BF16
                  : currentConfig.precision == 8 ? AMICA_PREC_FP8
                  : currentConfig.precision == 33 ? AMICA_PREC_FP32_PLAIN : AMICA_PREC_FP32;
    o.max_iter = currentConfig.maxiter > 0 ? (int)currentConfig.maxiter : 1;
    o.device = gpus[0].realdevice;
    o.do_newton = (int)currentConfig.do_newton;
    o.update_rho = (int)currentConfig.update_rho;
    o.rho_start_iter = (int)currentConfig.rho_start_iter;
    o.newt_start_iter = (int)currentConfig.newt_start_iter;
    o.update_gm = (int)currentConfig.update_gm;
    o.update_alpha = (int)currentConfig.update_alpha;
    o.update_A = (int)currentConfig.update_A;
    o.update_mu = (int)currentConfig.update_mu;
    o.update_beta = (int)currentConfig.update_beta;
    o.do_reparm = (int)currentConfig.do_reparm;
    o.lrate0 = currentConfig.lrate0;
    o.lratemax = currentConfig.lratemax;
    o.lnatrate = currentConfig.lnatrate;
    o.lratefact = currentConfig.lratefact;
    o.rholrate = currentConfig.rholrate;
    o.rhomin = currentConfig.rhomin;
    o.rhomax = currentConfig.rhomax;
    o.maxdec = (int)currentConfig.maxdec;
    C_CHECK_RETURN(autotune(0));
    int rc = amica_engine_create(&o, &gpus[0].engine);
    if (rc != 0) {
        fprintf(stderr, "ERROR: engine creation failed: %s\n", amica_last_error());
        return rc == AMICA_E_NOMEM ? ERRORNODEVICEMEM : ERRORINVALIDPARAM;
    }
    return SUCCESS;
}

error finalizeDevs(void) {
    for (natural d = 0; d < gpuCount; ++d) {
        if (gpus[d].engine) amica_engine_destroy(gpus[d].engine);
        gpus[d].engine = nullptr;
    }
    return SUCCESS;
}

error writeDevMatrix(amica_pitched ptr, char* filename) {
    return writeDevMatrix3d(ptr, 1, filename);
}

error writeDevMatrix3d(amica_pitched ptr, natural zsize, char* filename) {
    std::vector<double> tmp(ptr.xsize * ptr.ysize * zsize);
    if (hipMemcpy2D(tmp.data(), ptr.xsize * sizeof(double), ptr.ptr, ptr.pitch,
                    ptr.xsize * sizeof(double), ptr.ysize * zsize, hipMemcpyDeviceToHost) != hipSuccess)
        return ERRORIO;
    amica_pitched h = {tmp.data(), ptr.xsize * sizeof(double), ptr.xsize, ptr.ysize * zsize};
    return writeMatrix(h, filename);
}

error loadDevMatrix(amica_pitched ptr, char* filename, bool isdouble) {
    std::vector<double> tmp(ptr.xsize * ptr.ysize);
    error e = loadMatrix(tmp.data(), ptr.ysize, ptr.xsize, filename, isdouble);
    if (e != SUCCESS) return e;
    if (hipMemcpy2D(ptr.ptr, ptr.pitch, tmp.data(), ptr.xsize * sizeof(double),
                    ptr.xsize * sizeof(double), ptr.ysize, hipMemcpyHostToDevice) != hipSuccess)
        return ERRORIO;
    return SUCCESS;
}

error loadDevMatrix3d(amica_pitched ptr, natural index, char* filename) {
    /* slab `index` of a float32 file of ysize x xsize slabs */
    const size_t cnt = ptr.xsize * ptr.ysize;
    std::vector<float> all(cnt * (index + 1));
    error e = loadMatrixF32(all.data(), ptr.ysize * (index + 1), ptr.xsize, filename);
    if (e != SUCCESS) return e;
    std::vector<double> tmp(cnt);
    for (size_t k = 0; k < cnt; ++k) tmp[k] = all[index * cnt + k];
    if (hipMemcpy2D(ptr.ptr, ptr.pitch, tmp.data(), ptr.xsize * sizeof(double),
                    ptr.xsize * sizeof(double), ptr.ysize, hipMemcpyHostToDevice) != hipSuccess)
        return ERRORIO;
    return SUCCESS;
}

}  // extern "C"
