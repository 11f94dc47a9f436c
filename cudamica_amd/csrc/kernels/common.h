/*
 * common.h -- device helpers shared by the engine's HIP kernels (gfx950).
 *
 * Wave64 reductions, MFMA operand vector types and the layout constants of
 * the engine's device buffers.
 */
#ifndef AMICA_KERNELS_COMMON_H_
#define AMICA_KERNELS_COMMON_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef float  f32x4  __attribute__((ext_vector_type(4)));
typedef double f64x4  __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short  i16x8  __attribute__((ext_vector_type(8)));

#define AMICA_WAVE 64
#define AMICA_MAX_M 16      /* MAX_MODELS of the reference (config.h:17) */
#define AMICA_MAX_MJ 8      /* mixture components supported by the fused E-step */

/* log2(e) and ln(2) */
#define AMICA_L2E 1.4426950408889634f
#define AMICA_LN2 0.6931471805599453

/* The reference's truncated log(2) constant (config.h:25). */
#define AMICA_LOG2_REF 0.693147180559945

__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

/* Sum over groups of `width` consecutive lanes (width a power of two <= 64). */
__device__ __forceinline__ float group_sum_f32(float v, int width) {
    for (int off = width >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

/* Round-to-nearest-even float -> bf16 bit pattern. */
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
    uint32_t u = __float_as_uint(f);
    uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    if ((u & 0x7F800000u) == 0x7F800000u) r = u;   /* inf / nan pass through */
    return (uint16_t)(r >> 16);
}

__device__ __forceinline__ float bf16_bits_to_f32(uint16_t b) {
    return __uint_as_float(((uint32_t)b) << 16);
}

#endif
