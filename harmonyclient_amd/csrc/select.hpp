// Workgroup top-k selection in LDS.
//
// Replaces the reference's per-query binary heap (src/heap.hpp:35-90) with a
// threshold filter and periodic compaction, which suits a 256-lane
// workgroup: every lane compares its candidate's packed (distance, id) key
// against the current k-th key held in a register and appends the survivors
// to an LDS buffer with one LDS atomic; when the buffer could overflow in the
// next round the workgroup bitonic-sorts it, keeps the best k and tightens
// the threshold to the new k-th key.  The result is the exact top-k in the
// (distance, id) total order.
#pragma once

#include "common.hpp"

namespace hivf {

// Bitonic sort of n (a power of two, >= 2) keys in LDS, ascending, by all NT
// threads of the workgroup.  Ends with a barrier.
template <int NT>
__device__ inline void lds_sort(uint64_t* s, int n) {
    for (int size = 2; size <= n; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < (n >> 1); i += NT) {
                int lo = 2 * i - (i & (stride - 1));
                int hi = lo + stride;
                bool asc = (lo & size) == 0;
                uint64_t a = s[lo], b = s[hi];
                if ((a > b) == asc) {
                    s[lo] = b;
                    s[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

__device__ inline int pow2_at_least(int n) {
    int p = 64;
    while (p < n) p <<= 1;
    return p;
}

// State lives in LDS; every member function is called by all NT threads.
template <int NT, int B>
struct LdsTopK {
    static_assert((B & (B - 1)) == 0, "B must be a power of two");
    static_assert(B >= 2 * NT, "buffer must hold two rounds");
    uint64_t buf[B];
    int cnt;
    uint64_t thr;  // admission: key < thr

    __device__ void init() {
        if (threadIdx.x == 0) {
            cnt = 0;
            thr = key_limit();
        }
        __syncthreads();
    }

    // Any thread, any number of times per round (at most NT per round in
    // total across the workgroup).
    __device__ void offer(uint64_t key, uint64_t thr_reg) {
        if (key < thr_reg) {
            int s = atomicAdd(&cnt, 1);
            buf[s] = key;
        }
    }

    // Sort the first pow2(cnt) entries and keep the best k.
    __device__ void compact(int k) {
        int c = cnt;
        int n = pow2_at_least(c);
        for (int i = c + threadIdx.x; i < n; i += NT) buf[i] = ~0ull;
        __syncthreads();
        lds_sort<NT>(buf, n);
        if (threadIdx.x == 0) {
            if (c >= k) {
                // keys are unique (distinct ids): admission becomes key < k-th
                thr = buf[k - 1];
                cnt = k;
            }
        }
        __syncthreads();
    }

    // Called after the round's barrier: compacts when the next round could
    // overflow the buffer.  `incoming` = most keys one round can add.  The
    // count is read by every thread before the barrier that lets any thread
    // offer again, so all threads take the same branch.
    __device__ void maybe_compact(int k, int incoming) {
        int c = cnt;
        __syncthreads();
        if (c > B - incoming) compact(k);
    }

    __device__ void finish(int k) { compact(k); }
};

}  // namespace hivf
