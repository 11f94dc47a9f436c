/*
 * probe.hip -- achievable HBM bandwidth on the selected device, measured
 * with this library's own streaming kernels (a copy and a read-only sum),
 * the yardstick beside the nominal 8 TB/s for the E-step and GEMM figures.
 * No counterpart in the reference (it prints only per-iteration wall
 * time, src/amica.cu:2538-2540).
 */
#include <hip/hip_runtime.h>

#include <algorithm>

#include "amica_engine.h"

namespace {

constexpr int kProbeThreads = 256;
constexpr int kProbeUnroll = 4;      /* float4 per thread per pass */

/* dst = src, kProbeUnroll float4 per thread: all loads of a thread are
 * issued before its stores; the grid covers the buffer once */
__global__ __launch_bounds__(kProbeThreads) void probe_copy_kernel(const float4* __restrict__ src,
                                                                   float4* __restrict__ dst,
                                                                   size_t n4) {
    const size_t base = (size_t)blockIdx.x * kProbeThreads * kProbeUnroll + threadIdx.x;
    float4 v[kProbeUnroll];
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) {
        const size_t k = base + (size_t)u * kProbeThreads;
        if (k < n4) v[u] = src[k];
    }
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) {
        const size_t k = base + (size_t)u * kProbeThreads;
        if (k < n4) dst[k] = v[u];
    }
}

/* per-workgroup sums of the buffer (the read stream of the copy alone) */
__global__ __launch_bounds__(kProbeThreads) void probe_read_kernel(const float4* __restrict__ src,
                                                                   size_t n4, float* __restrict__ out) {
    const size_t base = (size_t)blockIdx.x * kProbeThreads * kProbeUnroll + threadIdx.x;
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < kProbeUnroll; ++u) {
        const size_t k = base + (size_t)u * kProbeThreads;
        if (k < n4) {
            const float4 v = src[k];
            s += (v.x + v.y) + (v.z + v.w);
        }
    }
    __shared__ float red[kProbeThreads / 64];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

/* fill with a pattern (no hipMemset, so nothing here depends on the
 * runtime's own kernels) */
__global__ __launch_bounds__(kProbeThreads) void probe_fill_kernel(float4* __restrict__ dst, size_t n4) {
    const size_t k = (size_t)blockIdx.x * kProbeThreads + threadIdx.x;
    if (k < n4) {
        const float f = (float)(k & 1023) * 0x1p-10f;
        dst[k] = make_float4(f, f + 1.f, f + 2.f, f + 3.f);
    }
}

}  // namespace

extern "C" int amica_hbm_probe(int device, size_t bytes, int reps, double* copy_gbs,
                               double* read_gbs) {
    if (bytes < (1u << 20) || reps < 1) return AMICA_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return AMICA_E_DEVICE;
    const size_t n4 = bytes / sizeof(float4);
    const unsigned nblk = (unsigned)((n4 + kProbeThreads * kProbeUnroll - 1) / (kProbeThreads * kProbeUnroll));
    float4 *a = nullptr, *b = nullptr;
    float* sums = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = AMICA_E_DEVICE;
    double best_copy = 0.0, best_read = 0.0;
    do {
        if (hipMalloc((void**)&a, n4 * sizeof(float4)) != hipSuccess) break;
        if (hipMalloc((void**)&b, n4 * sizeof(float4)) != hipSuccess) break;
        if (hipMalloc((void**)&sums, (size_t)nblk * sizeof(float)) != hipSuccess) break;
        if (hipStreamCreate(&st) != hipSuccess) break;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) break;
        const unsigned fblk = (unsigned)((n4 + kProbeThreads - 1) / kProbeThreads);
        probe_fill_kernel<<<fblk, kProbeThreads, 0, st>>>(a, n4);
        probe_fill_kernel<<<fblk, kProbeThreads, 0, st>>>(b, n4);
        bool ok = hipGetLastError() == hipSuccess;
        for (int r = 0; ok && r < reps + 1; ++r) {      /* the first copy and read warm up */
            float ms = 0.f;
            ok = hipEventRecord(e0, st) == hipSuccess;
            probe_copy_kernel<<<nblk, kProbeThreads, 0, st>>>(a, b, n4);
            ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess
                 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (ok && r > 0 && ms > 0.f) best_copy = std::max(best_copy, 2.0 * (double)bytes / (ms * 1e-3) / 1e9);
            ok = ok && hipEventRecord(e0, st) == hipSuccess;
            probe_read_kernel<<<nblk, kProbeThreads, 0, st>>>(b, n4, sums);
            ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess
                 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (ok && r > 0 && ms > 0.f) best_read = std::max(best_read, (double)bytes / (ms * 1e-3) / 1e9);
        }
        if (ok && hipStreamSynchronize(st) == hipSuccess) rc = AMICA_OK;
    } while (0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(sums);
    if (copy_gbs) *copy_gbs = best_copy;
    if (read_gbs) *read_gbs = best_read;
    return rc;
}
