// BF16 helpers of the MFMA kernels (bit-level, no library rounding calls).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hivf {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float floatx16;
typedef __attribute__((ext_vector_type(4))) int i32x4;     // 16 x int8 MFMA operand
typedef __attribute__((ext_vector_type(16))) int i32x16;   // int32 MFMA accumulator

constexpr uint16_t kBf16One = 0x3F80;

__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC0;  // NaN
    u += 0x7FFFu + ((u >> 16) & 1u);                      // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// v = hi + mid + lo exactly, each piece a BF16 value (top 8 significant bits,
// next 8, last 8 of the FP32 significand): truncations and the two
// subtractions are exact.
__device__ inline void split3(float v, uint16_t* out) {
    const uint16_t hi = (uint16_t)(__float_as_uint(v) >> 16);
    const float r = v - bf2f(hi);
    const uint16_t mid = (uint16_t)(__float_as_uint(r) >> 16);
    const float r2 = r - bf2f(mid);
    out[0] = hi;
    out[1] = mid;
    out[2] = f2bf(r2);
}


}  // namespace hivf
