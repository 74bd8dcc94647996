// FP32 <-> BF16 conversion shared by the MLP and the dynamical core.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace thor {
namespace mlp {

// Round to nearest even; NaN stays a quiet NaN.
__host__ __device__ inline uint16_t f2bf(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// The same rounding (to nearest even) by gfx950's conversion instruction
// (v_cvt_pk_bf16_f32): one instruction for two values. The device kernels of
// the network use these throughout, so training and inference round alike.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ inline uint32_t f2bf2(float a, float b) {
    const bf16x2 v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(uint32_t, v);
}
__device__ inline uint16_t f2bf_dev(float a) { return (uint16_t)(f2bf2(a, 0.0f) & 0xffffu); }

} // namespace mlp
} // namespace thor
