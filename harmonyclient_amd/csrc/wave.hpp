// Wave-level helpers (one 64-lane wavefront, no workgroup barrier): min /
// max / prefix sums over the lanes, 256-bin linear histograms in LDS, and a
// bitonic sort of 64-bit keys in LDS.
#pragma once

#include "common.hpp"

namespace hivf {

__device__ inline float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
    return v;
}

__device__ inline float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
    return v;
}

// Linear bin of v in [lo, lo + 256 w): monotone non-decreasing in v.
__device__ inline int lin_bin(float v, float lo, float inv_w) {
    const float t = (v - lo) * inv_w;
    return t < 0.0f ? 0 : (t >= 255.0f ? 255 : (int)t);
}

__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off);
        if (lane >= off) v += t;
    }
    return v;
}

// Lanes below this one whose bit is set in m.
__device__ inline uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Bin of the need-th (1-based) entry of a 256-bin histogram held in LDS,
// found by the whole wave; *need becomes the rank inside that bin.
__device__ inline int wave_find_bin(const uint32_t* hist, uint32_t* need, int lane) {
    const uint32_t h0 = hist[lane * 4], h1 = hist[lane * 4 + 1], h2 = hist[lane * 4 + 2], h3 = hist[lane * 4 + 3];
    const uint32_t sum = h0 + h1 + h2 + h3;
    const uint32_t incl = wave_incl_scan(sum, lane);
    const uint64_t hit = __ballot(incl >= *need);
    const int L = hit ? __builtin_ctzll(hit) : 63;
    uint32_t before = __shfl(incl - sum, L);
    uint32_t hl[4] = {__shfl(h0, L), __shfl(h1, L), __shfl(h2, L), __shfl(h3, L)};
    int b = L * 4;
    for (int i = 0; i < 3; ++i) {
        if (before + hl[i] >= *need) break;
        before += hl[i];
        ++b;
    }
    *need -= before;
    return b;
}

// One stage of a bitonic sort by one wave: the pairs (lo, lo + stride) of
// the stage, each lane reading all of its (at most P) pairs before it
// compares and writes any, so the reads are in flight together and the
// stage waits for the LDS once.
template <int P>
__device__ inline void wave_sort_stage(uint64_t* kk, uint32_t np, uint32_t size, uint32_t stride, int lane) {
    uint32_t lo[P];
    uint64_t x[P], y[P];
#pragma unroll
    for (int t = 0; t < P; ++t) {
        const uint32_t i = (uint32_t)lane + 64u * t;
        lo[t] = 2 * i - (i & (stride - 1));
        if (i < np) {
            x[t] = kk[lo[t]];
            y[t] = kk[lo[t] + stride];
        }
    }
#pragma unroll
    for (int t = 0; t < P; ++t) {
        const uint32_t i = (uint32_t)lane + 64u * t;
        const bool asc = (lo[t] & size) == 0;
        if (i < np && (x[t] > y[t]) == asc) {
            kk[lo[t]] = y[t];
            kk[lo[t] + stride] = x[t];
        }
    }
}

// Bitonic sort, ascending, of n keys in LDS (n a power of two, >= 2) by one
// wave.  LDS operations of one wave complete in order, so the wave barrier
// (a compiler fence) is all the ordering the stages need.  A stage's pairs
// are disjoint; up to n = 512 each lane holds all of its pairs of a stage
// in registers (as many as the size needs: 1, 2 or 4).
__device__ inline void wave_sort(uint64_t* kk, uint32_t n, int lane) {
    const uint32_t np = n >> 1;
    for (uint32_t size = 2; size <= n; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            if (np <= 64u) {
                wave_sort_stage<1>(kk, np, size, stride, lane);
            } else if (np <= 128u) {
                wave_sort_stage<2>(kk, np, size, stride, lane);
            } else if (np <= 256u) {
                wave_sort_stage<4>(kk, np, size, stride, lane);
            } else {
                for (uint32_t i = lane; i < np; i += 64) {
                    const uint32_t lo = 2 * i - (i & (stride - 1));
                    const uint32_t hi = lo + stride;
                    const bool asc = (lo & size) == 0;
                    const uint64_t x = kk[lo], y = kk[hi];
                    if ((x > y) == asc) {
                        kk[lo] = y;
                        kk[hi] = x;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

}  // namespace hivf
