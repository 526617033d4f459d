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

// Linear bin of v among 1024 bins of width 1 / inv_w from lo.
__device__ inline int lin_bin1024(float v, float lo, float inv_w) {
    const float t = (v - lo) * inv_w;
    return t < 0.0f ? 0 : (t >= 1023.0f ? 1023 : (int)t);
}

// Bin of the need-th (1-based, need <= the total) entry of a 1024-bin
// histogram in LDS (16-byte aligned), found by the whole wave: each lane
// sums its 16 bins, a scan over the lanes finds the lane, that lane's
// running sum the bin.
__device__ inline int wave_find_bin1024(const uint32_t* hist, uint32_t need, int lane) {
    const uint4* h4 = reinterpret_cast<const uint4*>(hist);
    uint32_t h[16];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const uint4 x = h4[lane * 4 + v];
        h[v * 4 + 0] = x.x;
        h[v * 4 + 1] = x.y;
        h[v * 4 + 2] = x.z;
        h[v * 4 + 3] = x.w;
    }
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) sum += h[j];
    const uint32_t incl = wave_incl_scan(sum, lane);
    const uint64_t hit = __ballot(incl >= need);
    const int L = hit ? __builtin_ctzll(hit) : 63;
    uint32_t run = incl - sum;
    int bj = 15;
    bool found = false;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (!found && run + h[j] >= need) {
            bj = j;
            found = true;
        }
        run += h[j];
    }
    return L * 16 + __shfl(bj, L);
}

// Clear a 1024-bin LDS histogram (16-byte aligned) by one wave.
__device__ inline void wave_clear1024(uint32_t* hist, int lane) {
    uint4* h4 = reinterpret_cast<uint4*>(hist);
    for (int i = lane; i < 256; i += 64) h4[i] = make_uint4(0u, 0u, 0u, 0u);
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

// The value of lane (lane ^ J) for a lane's 64-bit v.  J < 16: DPP moves
// inside rows of 16 lanes (quad permutations for 1 and 2; row rotations for
// 4 and 8: row_ror:N gives lane i the value of lane (i - N) mod 16, so the
// partner i ^ 4 is ror:4 when bit 2 is set and ror:12 when it is clear, and
// i ^ 8 is ror:8); J = 16, 32: LDS permutes.
template <int J>
__device__ inline uint64_t lane_xor(uint64_t v, int lane) {
    const int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
    int rl, rh;
    if constexpr (J == 1 || J == 2 || J == 8) {
        constexpr int ctrl = J == 1 ? 0xB1 : J == 2 ? 0x4E : 0x128;
        rl = __builtin_amdgcn_update_dpp(0, lo, ctrl, 0xF, 0xF, false);
        rh = __builtin_amdgcn_update_dpp(0, hi, ctrl, 0xF, 0xF, false);
    } else if constexpr (J == 4) {
        const int l4 = __builtin_amdgcn_update_dpp(0, lo, 0x124, 0xF, 0xF, false);
        const int l12 = __builtin_amdgcn_update_dpp(0, lo, 0x12C, 0xF, 0xF, false);
        const int h4 = __builtin_amdgcn_update_dpp(0, hi, 0x124, 0xF, 0xF, false);
        const int h12 = __builtin_amdgcn_update_dpp(0, hi, 0x12C, 0xF, 0xF, false);
        rl = (lane & 4) ? l4 : l12;
        rh = (lane & 4) ? h4 : h12;
    } else {
        rl = __shfl_xor(lo, J);
        rh = __shfl_xor(hi, J);
    }
    return ((uint64_t)(uint32_t)rh << 32) | (uint32_t)rl;
}

// One compare-exchange stage (block size K, distance J) of a bitonic sort
// held in registers: key v at position i of the sequence, its partner at
// i ^ J in lane lane ^ J; the lower position keeps the smaller key in an
// ascending block ((i & K) == 0), the larger in a descending one.
template <int K, int J>
__device__ inline uint64_t reg_stage(uint64_t v, uint32_t i, int lane) {
    const uint64_t p = lane_xor<J>(v, lane);
    const bool keep_min = ((i & J) == 0) == ((i & K) == 0);
    return keep_min ? (v < p ? v : p) : (v < p ? p : v);
}

// Bitonic sort of 64 (one per lane) or 128 keys (lane holds positions lane
// and lane + 64) in registers, ascending.
template <int K, int J>
__device__ inline void reg_sort_step(uint64_t& a, uint64_t& b, int lane, bool two) {
    if constexpr (J == 64) {
        const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;  // K = 128: ascending
        a = lo;
        b = hi;
    } else {
        a = reg_stage<K, J>(a, (uint32_t)lane, lane);
        if (two) b = reg_stage<K, J>(b, (uint32_t)lane + 64u, lane);
    }
    if constexpr (J > 1) reg_sort_step<K, J / 2>(a, b, lane, two);
}

template <int K>
__device__ inline void reg_sort_from(uint64_t& a, uint64_t& b, int lane, bool two) {
    reg_sort_step<K, K / 2>(a, b, lane, two);
    if constexpr (K < 128) {
        if (K < 64 || two) reg_sort_from<K * 2>(a, b, lane, two);
    }
}

// Bitonic sort, ascending, of n keys in LDS (n a power of two, >= 2) by one
// wave.  n = 64 and 128 (the selections' usual sizes) sort in registers
// (one LDS read and write per key); larger n stage by stage in LDS.  LDS operations of one wave complete in order, so the wave barrier
// (a compiler fence) is all the ordering the stages need.  A stage's pairs
// are disjoint; up to n = 512 each lane holds all of its pairs of a stage
// in registers (as many as the size needs: 1, 2 or 4).
__device__ inline void wave_sort(uint64_t* kk, uint32_t n, int lane) {
    if (n == 64 || n == 128) {
        const bool two = n == 128;
        uint64_t a = kk[lane], b = two ? kk[lane + 64] : 0ull;
        reg_sort_from<2>(a, b, lane, two);
        kk[lane] = a;
        if (two) kk[lane + 64] = b;
        __builtin_amdgcn_wave_barrier();
        return;
    }
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
