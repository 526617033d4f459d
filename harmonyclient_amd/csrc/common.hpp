// Shared definitions of the hivf HIP library (device and host).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace hivf {

constexpr int kWave = 64;        // CDNA wavefront width
constexpr int kTileRows = 64;    // list rows per tile: one row per lane
constexpr int kMaxK = 1024;      // largest k of the per-query scans
constexpr uint32_t kNoId = 0xFFFFFFFFu;

// A result entry packed into one 64-bit word: order-preserving bits of the
// ranking key (the L2 distance, or the negated inner product) in the high
// half, the row id in the low half.  Unsigned comparison of two words is the
// (distance, id) total order of a result row: smaller distance first (L2),
// larger inner product first (IP), ascending id among equal distances
// (src/heap.hpp:35-90 ordering, made total as in faiss's cmp2,
// third_party/faiss/faiss/utils/Heap.h:127-129).
__host__ __device__ inline uint32_t f2ord(float f) {
    if (f == 0.0f) f = 0.0f;  // -0 and +0 rank equal
#ifdef __HIP_DEVICE_COMPILE__
    uint32_t u = __float_as_uint(f);
#else
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
#endif
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__host__ __device__ inline float ord2f(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float(u);
#else
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
#endif
}

__host__ __device__ inline uint64_t make_key(float rank, uint32_t id) {
    return (static_cast<uint64_t>(f2ord(rank)) << 32) | id;
}

__host__ __device__ inline float key_rank(uint64_t key) {
    return ord2f(static_cast<uint32_t>(key >> 32));
}

// Any key below this one has a rank < FLT_MAX: the admission rule of the
// reference's heap initialised with FLT_MAX (src/heap.hpp:13-25,
// src/IVFScan.hpp:131), which never admits inf or NaN.
__host__ __device__ constexpr uint64_t key_limit() {
    return static_cast<uint64_t>(0xFF7FFFFFu) << 32;  // f2ord(FLT_MAX) << 32
}

enum Metric : int { kIP = 0, kL2 = 1 };

// Four more dims of the canonical distance chain: acc = fma(q_i - x_i,
// q_i - x_i, acc) for L2 (src/utils.h:480-487), acc = fma(q_i, x_i, acc) for
// IP (src/utils.h:512-519), dims in ascending order.
template <int METRIC>
__device__ inline float accum(float acc, float4 q, float4 x) {
    if constexpr (METRIC == kL2) {
        float t;
        t = q.x - x.x; acc = __builtin_fmaf(t, t, acc);
        t = q.y - x.y; acc = __builtin_fmaf(t, t, acc);
        t = q.z - x.z; acc = __builtin_fmaf(t, t, acc);
        t = q.w - x.w; acc = __builtin_fmaf(t, t, acc);
    } else {
        acc = __builtin_fmaf(q.x, x.x, acc);
        acc = __builtin_fmaf(q.y, x.y, acc);
        acc = __builtin_fmaf(q.z, x.z, acc);
        acc = __builtin_fmaf(q.w, x.w, acc);
    }
    return acc;
}

// Ranking key of a distance: smaller ranks first for both metrics.
template <int METRIC>
__device__ inline float rank_of(float dist) {
    return METRIC == kL2 ? dist : -dist;
}

// Inverse of rank_of for a rank read back from a packed key.  Keys rank -0
// and +0 equal and give back +0, so an inner product of zero returns as
// +0 (0 - r, never -r), which is also what the FMA chain yields for it.
__device__ inline float dist_of_rank(int metric, float r) {
    return metric == kL2 ? r : 0.0f - r;
}

template <int METRIC>
__device__ inline float sentinel() {
    return METRIC == kL2 ? FLT_MAX : -FLT_MAX;
}

}  // namespace hivf
