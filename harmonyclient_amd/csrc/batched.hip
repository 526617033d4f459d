// List-major batched scan: many queries against each inverted list at once,
// with the query x list-tile contraction on BF16 MFMA.
//
// The reference scans every (query, probed list) pair on its own
// (src/Index.cpp:1432-1549); here the probe lists of a whole batch are
// inverted into per-list query sets, so each list tile is read once per work
// item of up to 128 queries instead of once per query.  For a list tile X (32 rows) and 32
// queries Q the wave forms  ||q||^2 + ||x||^2 - 2 q.x  from one
// v_mfma_f32_32x32x16_bf16 chain over the dims and keeps every row whose
// distance can be <= the query's bound tau (the warm-up bound,
// src/Index.cpp:999-1051, or the learned one).  A per-query selection then
// returns the exact top-k of those survivors:
//   - when the query and the lists hold integers in [0, 255] and d <= 129,
//     every product, sum and norm is an integer below 2^24, so the MFMA
//     distance is the exact distance, bit-identical to the FP32 FMA chain;
//   - otherwise a row survives when  dist - E <= tau  with E a proven bound
//     of the BF16 error (margin_coef below), and the survivors that can
//     reach the top-k (lower bound <= the k-th smallest upper bound) are
//     rescored with the exact FP32 FMA chain from the original vectors.
// Every row of the true top-k has exact distance <= D_k <= tau, so it
// survives; the selection over survivors is therefore the exact top-k.  A
// query whose survivors overflow the buffer, or whose learned bound is
// below its k-th distance, is rescanned by the exact per-query kernel.
#include <algorithm>
#include <climits>
#include <cstring>

#include "bf16.hpp"
#include "index_impl.hpp"
#include "prep.hpp"
#include "select.hpp"
#include "wave.hpp"

namespace hivf {

namespace {


// A work item is (list, up to 32 * QG queries, up to item_tiles 32-row
// tiles); QG is the kernel variant's number of 32-query MFMA column groups.
constexpr int kNT = 256;



// BF16 shadow of the FP32 tile arena in MFMA operand order:
// [32-row tile T][16-dim chunk kc][lane][8 bf16], lane l holding row l % 32,
// dims kc*16 + (l / 32) * 8 .. +7; `stride` chunks per tile.  With `aug`, the
// extra chunk kc = nkc of a row holds [n_hi, n_mid, n_lo, 1, 1, 1, 0, 0] in
// its first half (n = the row's squared norm split by split3) and zeros in
// the second, so one more MFMA adds ||x||^2 + ||q||^2 to -2 q.x.
__global__ void build_bf16_kernel(const float4* __restrict__ codes, const float* __restrict__ norms, int nquad,
                                  int64_t tiles32, int nkc, int stride, uint16_t* __restrict__ out) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= tiles32 * stride * 64) return;
    int lane = (int)(gid % 64);
    int kc = (int)((gid / 64) % stride);
    int64_t T = gid / (64 * (int64_t)stride);
    int64_t t64 = T / 2;
    int r64 = (int)(T % 2) * 32 + lane % 32;
    uint16_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (kc < nkc) {
        int c0 = kc * 16 + (lane / 32) * 8;
        for (int h = 0; h < 2; ++h) {
            int qd = c0 / 4 + h;
            float4 f = qd < nquad ? codes[(t64 * nquad + qd) * 64 + r64] : make_float4(0.f, 0.f, 0.f, 0.f);
            v[h * 4 + 0] = f2bf(f.x);
            v[h * 4 + 1] = f2bf(f.y);
            v[h * 4 + 2] = f2bf(f.z);
            v[h * 4 + 3] = f2bf(f.w);
        }
    } else if (lane < 32) {
        split3(norms[T * 32 + lane], v);
        v[3] = v[4] = v[5] = kBf16One;
    }
    uint4 pk;
    pk.x = v[0] | ((uint32_t)v[1] << 16);
    pk.y = v[2] | ((uint32_t)v[3] << 16);
    pk.z = v[4] | ((uint32_t)v[5] << 16);
    pk.w = v[6] | ((uint32_t)v[7] << 16);
    reinterpret_cast<uint4*>(out)[gid] = pk;
}

// Int8 shadow of an integer arena in MFMA operand order: [32-row tile]
// [32-dim chunk kc][lane][16 x int8], lane l holding row l % 32, dims
// kc * 32 + (l / 32) * 16 .. +15, value x - 128 (0 past d).  Shadow row p
// holds arena row perm[p] (the same list; every list's rows in increasing
// distance to the centroid).
__global__ void build_i8_kernel(const float4* __restrict__ codes, int nquad, int d, int64_t tiles32, int nkc8,
                                const int32_t* __restrict__ perm, i32x4* __restrict__ out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= tiles32 * nkc8 * 64) return;
    const int lane = (int)(gid % 64);
    const int kc = (int)((gid / 64) % nkc8);
    const int64_t T = gid / (64 * (int64_t)nkc8);
    const int64_t src = perm[T * 32 + lane % 32];
    const int64_t t64 = src / 64;
    const int r64 = (int)(src % 64);
    const int c0 = kc * 32 + (lane / 32) * 16;
    uint32_t wds[4];
    for (int h = 0; h < 4; ++h) {
        const int qd = c0 / 4 + h;
        const float4 f = qd < nquad ? codes[(t64 * nquad + qd) * 64 + r64] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float e[4] = {f.x, f.y, f.z, f.w};
        uint32_t wd = 0;
        for (int i = 0; i < 4; ++i) {
            const int v = c0 + h * 4 + i < d ? (int)e[i] - 128 : 0;
            wd |= ((uint32_t)v & 0xFFu) << (8 * i);
        }
        wds[h] = wd;
    }
    out[gid] = i32x4{(int)wds[0], (int)wds[1], (int)wds[2], (int)wds[3]};
}

// -sum (x - 128)^2 over the first d dims of every arena row (one thread per row).
__global__ void norms_i8_kernel(const float4* __restrict__ codes, int nquad, int d, int64_t rows,
                                const int32_t* __restrict__ perm, int32_t* __restrict__ xn) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int64_t src = perm[r];
    const float4* p = codes + (src / 64) * nquad * 64 + (src % 64);
    int s = 0;
    for (int j = 0; j < nquad; ++j) {
        const float4 f = p[(size_t)j * 64];
        const float e[4] = {f.x, f.y, f.z, f.w};
        for (int i = 0; i < 4; ++i)
            if (j * 4 + i < d) {
                const int v = (int)e[i] - 128;
                s += v * v;
            }
    }
    xn[r] = -s;  // negated: the scan adds it to 2 q'.x'
}


// out[p] = ids[perm[p]]: the ids in int8 shadow order.
__global__ void gather_ids_kernel(const uint32_t* __restrict__ ids, const int32_t* __restrict__ perm, int64_t rows,
                                  uint32_t* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < rows) out[p] = ids[perm[p]];
}

// Radius bounds of every 32-row tile of the arena: the smallest and largest
// distance ||x - c|| of its rows to their list's centroid c, widened by a
// relative 1e-4 (FP32 chain and square root errors are below 2e-5 for d <=
// 512).  One wave per two tiles; tiles without rows get [inf, -inf].
__global__ void tile_radius_kernel(const float4* __restrict__ codes, const float4* __restrict__ cent, int nquad,
                                   const int32_t* __restrict__ tile_list, const int32_t* __restrict__ list_tile_off,
                                   const int32_t* __restrict__ list_size, int64_t tiles32,
                                   float2* __restrict__ out, float* __restrict__ row_r) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = gid / 32;
    const int r = (int)(gid % 32);
    float lo = INFINITY, hi = -INFINITY;
    if (T < tiles32) {
        const int l = tile_list[T];
        const int row = (int)(T - 2 * (int64_t)list_tile_off[l]) * 32 + r;
        if (l >= 0 && row < list_size[l]) {
            const int64_t t64 = T / 2;
            const int r64 = (int)(T % 2) * 32 + r;
            const float4* xp = codes + t64 * nquad * 64 + r64;
            const float4* cp = cent + (int64_t)(l / 64) * nquad * 64 + (l % 64);
            float acc = 0.0f;
            for (int j = 0; j < nquad; ++j) acc = accum<kL2>(acc, xp[(size_t)j * 64], cp[(size_t)j * 64]);
            const float rr = sqrtf(acc);
            if (row_r) row_r[gid] = rr;
            lo = rr * (1.0f - 1e-4f);
            hi = rr * (1.0f + 1e-4f) + 1e-30f;
        }
    }
    for (int off = 16; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off));
        hi = fmaxf(hi, __shfl_xor(hi, off));
    }
    if (r == 0 && T < tiles32) out[T] = make_float2(lo, hi);
}

// Radius interval of every list: the union of its tiles' intervals.
__global__ void list_radius_kernel(const float2* __restrict__ tb, const int32_t* __restrict__ list_tile_off,
                                   const int32_t* __restrict__ list_size, int nlist, float2* __restrict__ out) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlist) return;
    float lo = INFINITY, hi = -INFINITY;
    const int64_t t0 = 2 * (int64_t)list_tile_off[l], t1 = t0 + (list_size[l] + 31) / 32;
    for (int64_t t = t0; t < t1; ++t) {
        lo = fminf(lo, tb[t].x);
        hi = fmaxf(hi, tb[t].y);
    }
    out[l] = make_float2(lo, hi);
}

// The preparation on its own: one 64-thread workgroup per query.
__global__ void prep_kernel(PrepArgs a) {
    float qv[kQRegs];
    query_regs(a.q, blockIdx.x, a.ld, a.d, (int)threadIdx.x, qv);
    prep_query(a, (int)blockIdx.x, (int)threadIdx.x, qv);
}

// Queries per list (only the `active` ones when non-null); optionally the
// candidate count sum_q C_q into *cand.
// List-level radius bound on a (query, list) pair: every row x of the list
// (centroid c) has ||x - c|| in [Rmin, Rmax], so | ||q - c|| - ||x - c|| | >
// sqrt(tau) for all of them when ||q - c|| lies farther than sqrt(tau) from
// that interval; then no row can pass the query's bound and the pair is
// left out of the plan.  Same widened arithmetic as the scan's tile bound.
struct PairFilter {
    const float* cd;       // [nq x P] squared coarse distances (null: keep every pair)
    const float* tau;      // [nq] the pass's bounds
    const float2* lbound;  // [nlist] radius interval of each list
};

__device__ inline bool pair_may_pass(const PairFilter& f, int64_t i, int q, int l) {
    if (!f.cd) return true;
    const float t = f.tau[q];
    if (!(t >= 0.0f)) return false;  // no distance is below a negative bound
    const float dc = f.cd[i];
    const float dlo = sqrtf(fmaxf(dc, 0.0f) * (1.0f - 1e-4f));
    const float dhi = sqrtf(dc * (1.0f + 1e-4f));
    const float sq = sqrtf(t * (1.0f + 1e-4f));
    const float2 rb = f.lbound[l];
    return !(dlo - rb.y > sq) && !(rb.x - dhi > sq);
}

__global__ void count_kernel(const int32_t* __restrict__ probes, int nq, int P, const int32_t* __restrict__ npq,
                             const int32_t* __restrict__ active, const int32_t* __restrict__ list_size,
                             PairFilter pf, uint32_t* __restrict__ lcnt, unsigned long long* cand) {
    __shared__ unsigned long long s_c;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int64_t)nq * P) {
        int q = (int)(i / P), j = (int)(i % P);
        int Pq = npq ? min(npq[q], P) : P;
        int l = probes[i];
        if (j < Pq && l >= 0 && (!active || active[q])) {
            // the candidate count is every probed row (the algorithmic work)
            if (cand) atomicAdd(&s_c, (unsigned long long)list_size[l]);
            if (pair_may_pass(pf, i, q, l)) atomicAdd(&lcnt[l], 1u);
        }
    }
    __syncthreads();
    if (cand && threadIdx.x == 0 && s_c) atomicAdd(cand, s_c);
}

// Work-queue groups: one per XCD of the MI355X (8 XCDs, each with its own
// L2); workgroups are dispatched to XCDs round-robin, so workgroup b runs on
// XCD b % kXcds and its waves take the items of that group first.
constexpr int kXcds = 8;
// meta layout (32-bit words): [0] items; [8 .. 16] group item offsets (read
// only during the scan); queue head of group x at kMetaHead + x * kHeadStride,
// each in its own 256-byte line, so the claims of one group never contend
// with another group's or with the offsets' reads
constexpr int kMetaOff = 8;
constexpr int kMetaHead = 64;
constexpr int kHeadStride = 64;
constexpr int kMetaWords = kMetaHead + kXcds * kHeadStride;
constexpr uint64_t kNoKey = ~0ull;  // survivor buffer padding: above key_limit(), skipped by the selections
constexpr int kStage = 32;     // survivor keys staged per query slot of the batched scan
// A slot is flushed before a tile whose survivors would not fit, so it must
// hold a whole tile's: up to 32 rows (int8 and BF16 tiles) per query.
static_assert(kStage >= 32, "a survivor slot holds one tile's rows");

// Work items are grouped by XCD.  In list order (a list's near-query items
// first, then its others), the items are cut into kXcds ranges of about
// equal estimated cost (every item counts 1, a near-query item kNearCost),
// at any item: a list much costlier than 1 / kXcds of the batch is shared by
// consecutive groups rather than left to one.  Group x is the contiguous
// range [S_x, S_x+1) of that order; its queue takes the range's near-query
// items first.  (With near-query items split, equal item counts balance
// best: weights 2 and 3 measured 3 and 7 us slower.)
constexpr uint32_t kNearCost = 1;

// Items of a list that hold its near queries (those for which it is one of
// their kNearProbes nearest lists, placed first among the list's queries):
// the query chunks that hold any, each split into tile chunks of rt /
// kNearSplit tiles.  Most of a query's survivors lie in its nearest lists:
// these items cost the most (the heaviest, where every row passes for most
// of the item's queries, staged and flushed tile after tile, would
// otherwise take the whole scan's time on one wave).
constexpr uint32_t kNearSplit = 2;  // (4 and 8 measured 6 and 18 us slower)
__device__ __host__ inline uint32_t near_tiles(uint32_t rt) { return rt > kNearSplit ? rt / kNearSplit : 1u; }
__device__ inline uint32_t near_chunks_of(uint32_t c, uint32_t near, uint32_t qc) {
    return (c == 0 || near == 0) ? 0u : min((c + qc - 1) / qc, (near + qc - 1) / qc);
}
__device__ inline uint32_t near_items_of(uint32_t c, uint32_t near, int rows, uint32_t qc, uint32_t rt) {
    if (rows == 0) return 0;
    const uint32_t t32 = (uint32_t)(rows + 31) / 32, rn = near_tiles(rt);
    return near_chunks_of(c, near, qc) * ((t32 + rn - 1) / rn);
}
// All items of a list: its near-query items, then its other query chunks
// with rt tiles per item.
__device__ inline uint32_t list_items_of(uint32_t c, uint32_t near, int rows, uint32_t qc, uint32_t rt) {
    if (c == 0 || rows == 0) return 0;
    const uint32_t t32 = (uint32_t)(rows + 31) / 32;
    const uint32_t nc = near_chunks_of(c, near, qc);
    return near_items_of(c, near, rows, qc, rt) + ((c + qc - 1) / qc - nc) * ((t32 + rt - 1) / rt);
}

// One workgroup of 1024 threads: exclusive scans of the per-list query
// counts, item counts and near-query item counts (lnear: near queries per
// list, or null), and the cut of the items into the groups.  meta[0] =
// items, meta[kMetaOff + x] = S_x, the first item of group x (x <= kXcds),
// meta[kMetaHead + x] = queue head of group x.  item_off[0 .. nlist] = items
// before each list, item_off[nlist + 1 .. 2 nlist + 1] = near-query items
// before each list.
__global__ __launch_bounds__(1024) void plan_kernel(const uint32_t* __restrict__ lcnt,
                                                    const uint32_t* __restrict__ lnear,
                                                    const int32_t* __restrict__ list_size, int nlist, uint32_t qc,
                                                    uint32_t rt,
                                                    uint32_t* __restrict__ qoff, uint32_t* __restrict__ lcur,
                                                    uint32_t* __restrict__ lend, uint32_t* __restrict__ item_off,
                                                    uint32_t* __restrict__ meta, const int32_t* __restrict__ qint,
                                                    int nq, uint32_t* __restrict__ all_int) {
    __shared__ uint32_t s1[1024], s2[1024], s3[1024];
    const int tid = threadIdx.x;
    // every query of the batch integer-valued in [0, 255] (8 loads in
    // flight per thread)
    int ok = 1;
    for (int i0 = 0; i0 < nq; i0 += 1024 * 8) {
        int32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = qint[min(i0 + j * 1024 + tid, nq - 1)];
#pragma unroll
        for (int j = 0; j < 8; ++j) ok &= v[j] != 0;  // (a clamped index repeats query nq - 1)
    }
    ok = __syncthreads_and(ok);
    if (tid == 0) *all_int = ok ? 1u : 0u;
    const int per = (nlist + 1023) / 1024;
    const int lo = min(nlist, tid * per), hi = min(nlist, lo + per);
    uint32_t* near_off = item_off + nlist + 1;
    auto near_of = [&](int l) { return lnear ? near_items_of(lcnt[l], lnear[l], list_size[l], qc, rt) : 0u; };
    auto all_of = [&](int l) { return list_items_of(lcnt[l], lnear ? lnear[l] : 0u, list_size[l], qc, rt); };
    uint32_t a = 0, b = 0, e = 0;
    for (int l = lo; l < hi; ++l) {
        a += lcnt[l];
        b += all_of(l);
        e += near_of(l);
    }
    s1[tid] = a;
    s2[tid] = b;
    s3[tid] = e;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
        uint32_t x = tid >= off ? s1[tid - off] : 0u, y = tid >= off ? s2[tid - off] : 0u,
                 z = tid >= off ? s3[tid - off] : 0u;
        __syncthreads();
        s1[tid] += x;
        s2[tid] += y;
        s3[tid] += z;
        __syncthreads();
    }
    a = s1[tid] - a;  // exclusive
    b = s2[tid] - b;
    e = s3[tid] - e;
    for (int l = lo; l < hi; ++l) {
        item_off[l] = b;
        near_off[l] = e;
        qoff[l] = a;
        lcur[l] = a;
        lend[l] = a + lcnt[l];
        a += lcnt[l];
        b += all_of(l);
        e += near_of(l);
    }
    if (tid == 1023) {
        item_off[nlist] = s2[1023];
        near_off[nlist] = s3[1023];
        meta[0] = s2[1023];
        meta[1] = 0;
        meta[kMetaOff + kXcds] = s2[1023];
    }
    if (tid < kXcds) meta[kMetaHead + tid * kHeadStride] = 0;
    __syncthreads();  // item_off and near_off complete
    if (tid < kXcds) {
        // S_x: the first item with at least x / kXcds of the total cost
        // before it (cost before list l: items + (kNearCost - 1) near items)
        using u64 = unsigned long long;
        auto cost_thr = [&](int t) {  // cost of the lists of threads 0 .. t
            return t < 0 ? 0ull : (u64)s2[t] + (kNearCost - 1) * (u64)s3[t];
        };
        auto cost_list = [&](int l) { return (u64)item_off[l] + (kNearCost - 1) * (u64)near_off[l]; };
        const u64 T = (cost_thr(1023) * (u64)tid + kXcds - 1) / kXcds;
        uint32_t S = 0;
        if (tid > 0 && T > 0) {
            int lo_t = -1, k = 1023;  // the first thread whose lists reach T: cost(lo_t) < T <= cost(k)
            while (k - lo_t > 1) {
                const int mid = (lo_t + k) / 2;
                if (cost_thr(mid) < T) lo_t = mid;
                else k = mid;
            }
            const int l0 = min(nlist, k * per), l1 = min(nlist, l0 + per);
            // the list holding the threshold: cost_list(l) < T <= cost_list(l + 1)
            int l = l0;
            while (l + 1 < l1 && cost_list(l + 1) < T) ++l;
            const u64 c0 = cost_list(l);
            const uint32_t nv = near_off[l + 1] - near_off[l], nit = item_off[l + 1] - item_off[l];
            const u64 x = T - c0;  // cost to take from list l (0 < x <= its cost)
            const u64 take = x <= (u64)nv * kNearCost ? (x + kNearCost - 1) / kNearCost : nv + (x - (u64)nv * kNearCost);
            S = item_off[l] + (uint32_t)min(take, (u64)nit);
        }
        meta[kMetaOff + tid] = S;
    }
}

// A list's queries are placed in two groups: those for which it is one of
// their kNearProbes nearest lists first (from the front cursor), the others
// after them (from the back cursor), each with its coarse distance to the
// list's centroid.  Far queries then share work items, whose tiles the
// scan's radius bound can skip as a whole.
constexpr int kNearProbes = 2;

__global__ void fill_kernel(const int32_t* __restrict__ probes, const float* __restrict__ cd, int nq, int P,
                            const int32_t* __restrict__ npq, const int32_t* __restrict__ active,
                            PairFilter pf, uint32_t* __restrict__ lcur, uint32_t* __restrict__ lend,
                            int32_t* __restrict__ lq, float* __restrict__ lqd) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nq * P) return;
    int q = (int)(i / P), j = (int)(i % P);
    int Pq = npq ? min(npq[q], P) : P;
    int l = probes[i];
    if (j < Pq && l >= 0 && (!active || active[q]) && pair_may_pass(pf, i, q, l)) {
        const uint32_t pos = j < kNearProbes ? atomicAdd(&lcur[l], 1u) : atomicSub(&lend[l], 1u) - 1u;
        lq[pos] = q;
        lqd[pos] = cd[i];
    }
}

// Lists up to this many use the LDS-histogram planning kernels below.
constexpr int kLdsLists = 8192;
constexpr int kPlanEntries = 4096;  // (query, probe) entries per planning workgroup

// count_kernel with a per-workgroup LDS histogram: one global atomic per
// (workgroup, list) instead of one per entry.
// The (query, probe) entries of a pass: every query's (those flagged in
// `active` when non-null), or, with a query list, the listed queries' only.
// Entry e is probe slot j of query q, at index gi = q * P + j of the probe
// table; q < 0 for an entry outside the pass.
struct PairSet {
    const int32_t* npq;
    const int32_t* active;
    const int32_t* qlist;
    const uint32_t* qcount;
    int nq, P;
    __device__ int64_t size() const { return qlist ? (int64_t)*qcount * P : (int64_t)nq * P; }
    __device__ int query(int64_t e, int& j, int64_t& gi) const {
        const int w = (int)(e / P);
        j = (int)(e % P);
        const int q = qlist ? qlist[w] : w;
        gi = (int64_t)q * P + j;
        const int Pq = npq ? min(npq[q], P) : P;
        return j < Pq && (qlist || !active || active[q]) ? q : -1;
    }
};

// Entry e of a pass, evaluated with unconditional loads (every index is
// valid whether or not the entry is in the pass), so that the unrolled loops
// below overlap the latencies of a thread's consecutive entries: the query
// q, probe slot j, its probe-table index i and list l; `in` = part of the
// pass; `pass` = in and the list-level radius bound keeps it; dc = the
// query's coarse distance to the list.
struct Entry {
    int q, j, l;
    int64_t i;
    bool in, pass;
    float dc;
    float fr;  // where the query's reach begins in the list's radius interval, in [0, 1]
};

// Where a query's reach in radius begins, ||q - c|| - sqrt(tau) (both
// widened as in the pair filter), as a fraction of the list's radius
// interval [rb.x, rb.y], clamped to [0, 1]: the rows of the list's tiles
// below it are out of the query's reach.
__device__ inline float reach_frac(float dlo, float sq, float2 rb) {
    const float fr = (dlo - sq - rb.x) / fmaxf(rb.y - rb.x, 1e-30f);
    return fminf(fmaxf(fr, 0.0f), 1.0f);  // (fmaxf takes 0 for NaN)
}

__device__ inline Entry entry_at(const PairSet& ps, const PairFilter& f, const int32_t* __restrict__ probes,
                                 int64_t e) {
    Entry x;
    const int w = (int)(e / ps.P);
    x.j = (int)(e % ps.P);
    x.q = ps.qlist ? ps.qlist[w] : w;
    x.i = (int64_t)x.q * ps.P + x.j;
    const int Pq = ps.npq ? min(ps.npq[x.q], ps.P) : ps.P;
    const bool act = ps.qlist || !ps.active || ps.active[x.q];
    x.l = probes[x.i];
    x.in = x.j < Pq && act && x.l >= 0;
    if (!f.cd) {
        x.pass = x.in;
        x.dc = 0.0f;
        x.fr = 0.0f;
        return x;
    }
    // pair_may_pass, branch-free
    const float t = f.tau[x.q];
    const float dc = f.cd[x.i];
    const float2 rb = f.lbound[max(x.l, 0)];
    const float dlo = sqrtf(fmaxf(dc, 0.0f) * (1.0f - 1e-4f));
    const float dhi = sqrtf(dc * (1.0f + 1e-4f));
    const float sq = sqrtf(t * (1.0f + 1e-4f));
    x.dc = dc;
    x.pass = x.in && t >= 0.0f && !(dlo - rb.y > sq) && !(rb.x - dhi > sq);
    x.fr = reach_frac(dlo, sq, rb);
    return x;
}

// The entries e0 + tid + 256 u (u < kEntryBatch) of a pass over every query
// (no query list, no flags), evaluated together: all their loads are issued
// before any is used.  f(x) is called for each entry below e1.
constexpr int kEntryBatch = 8;

template <bool NPQ, bool CD, class F>
__device__ inline void entry_batch(const PairSet& ps, const PairFilter& pf, const int32_t* __restrict__ probes,
                                   int64_t e, int64_t e1, F&& f) {
    Entry x[kEntryBatch];
    float t[kEntryBatch];
    int pq[kEntryBatch];
#pragma unroll
    for (int u = 0; u < kEntryBatch; ++u) {
        const uint32_t ee = (uint32_t)min(e + 256 * u, e1 - 1);  // (the pass has < 2^31 entries)
        x[u].q = (int)(ee / (uint32_t)ps.P);
        x[u].j = (int)(ee - (uint32_t)x[u].q * (uint32_t)ps.P);
        x[u].i = ee;
    }
#pragma unroll
    for (int u = 0; u < kEntryBatch; ++u) {
        x[u].l = probes[x[u].i];
        pq[u] = NPQ ? ps.npq[x[u].q] : ps.P;
        x[u].dc = CD ? pf.cd[x[u].i] : 0.0f;
        t[u] = CD ? pf.tau[x[u].q] : 0.0f;
    }
    float2 rb[kEntryBatch];
#pragma unroll
    for (int u = 0; u < kEntryBatch; ++u) rb[u] = CD ? pf.lbound[max(x[u].l, 0)] : make_float2(0.0f, 0.0f);
#pragma unroll
    for (int u = 0; u < kEntryBatch; ++u) {
        x[u].in = e + 256 * u < e1 && x[u].j < min(pq[u], ps.P) && x[u].l >= 0;
        x[u].pass = x[u].in;
        x[u].fr = 0.0f;
        if (CD) {
            const float dlo = sqrtf(fmaxf(x[u].dc, 0.0f) * (1.0f - 1e-4f));
            const float dhi = sqrtf(x[u].dc * (1.0f + 1e-4f));
            const float sq = sqrtf(t[u] * (1.0f + 1e-4f));
            x[u].pass = x[u].in && t[u] >= 0.0f && !(dlo - rb[u].y > sq) && !(rb[u].x - dhi > sq);
            x[u].fr = reach_frac(dlo, sq, rb[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < kEntryBatch; ++u)
        if (e + 256 * u < e1) f(x[u]);
}

template <bool NPQ, bool CD, class F>
__device__ inline void entry_batches(const PairSet& ps, const PairFilter& pf, const int32_t* __restrict__ probes,
                                     int64_t e0, int64_t e1, F&& f) {
    for (int64_t e = e0 + threadIdx.x; e < e1; e += 256 * kEntryBatch)
        entry_batch<NPQ, CD>(ps, pf, probes, e, e1, f);
}

// Every entry of [e0, e1) owned by this thread (stride 256), batched when
// the pass covers every query.
template <class F>
__device__ inline void for_entries(const PairSet& ps, const PairFilter& pf, const int32_t* __restrict__ probes,
                                   int64_t e0, int64_t e1, F&& f) {
    if (!ps.qlist && !ps.active) {
        if (ps.npq && pf.cd)
            entry_batches<true, true>(ps, pf, probes, e0, e1, f);
        else if (ps.npq)
            entry_batches<true, false>(ps, pf, probes, e0, e1, f);
        else if (pf.cd)
            entry_batches<false, true>(ps, pf, probes, e0, e1, f);
        else
            entry_batches<false, false>(ps, pf, probes, e0, e1, f);
    } else {
        for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) f(entry_at(ps, pf, probes, e));
    }
}

__global__ __launch_bounds__(256) void count_lds_kernel(const int32_t* __restrict__ probes, PairSet ps,
                                                        const int32_t* __restrict__ list_size, int nlist,
                                                        PairFilter pf, uint32_t* __restrict__ lcnt,
                                                        unsigned long long* cand) {
    __shared__ uint32_t hist[kLdsLists];
    __shared__ unsigned long long s_c;
    const int64_t e0 = (int64_t)blockIdx.x * kPlanEntries;
    const int64_t e1 = min(ps.size(), e0 + kPlanEntries);
    if (e0 >= e1) return;  // whole workgroup
    for (int b = threadIdx.x; b < nlist; b += 256) hist[b] = 0;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    unsigned long long c = 0;
    for_entries(ps, pf, probes, e0, e1, [&](const Entry& x) {
        if (cand && x.in) c += (unsigned long long)list_size[x.l];  // every probed row (the algorithmic work)
        if (x.pass) atomicAdd(&hist[x.l], 1u);
    });
    if (cand && c) atomicAdd(&s_c, c);
    __syncthreads();
    for (int b = threadIdx.x; b < nlist; b += 256)
        if (hist[b]) atomicAdd(&lcnt[b], hist[b]);
    if (cand && threadIdx.x == 0 && s_c) atomicAdd(cand, s_c);
}

// fill_kernel with a per-workgroup LDS histogram: each workgroup reserves one
// range per list with a global atomic, then places its entries with LDS
// atomics.
__global__ __launch_bounds__(256) void fill_lds_kernel(const int32_t* __restrict__ probes,
                                                       const float* __restrict__ cd, PairSet ps, int nlist,
                                                       PairFilter pf, uint32_t* __restrict__ lcur,
                                                       uint32_t* __restrict__ lend, int32_t* __restrict__ lq,
                                                       float* __restrict__ lqd) {
    __shared__ uint32_t hist[kLdsLists];
    const int64_t e0 = (int64_t)blockIdx.x * kPlanEntries;
    const int64_t e1 = min(ps.size(), e0 + kPlanEntries);
    if (e0 >= e1) return;  // whole workgroup
    for (int b = threadIdx.x; b < nlist; b += 256) hist[b] = 0;
    __syncthreads();
    // near entries (few) straight to the front of their list; far ones
    // counted here and placed from the back
    for_entries(ps, pf, probes, e0, e1, [&](const Entry& x) {
        if (x.pass) {
            if (x.j < kNearProbes) {
                const uint32_t pos = atomicAdd(&lcur[x.l], 1u);
                lq[pos] = x.q;
                lqd[pos] = cd[x.i];
            } else {
                atomicAdd(&hist[x.l], 1u);
            }
        }
    });
    __syncthreads();
    for (int b = threadIdx.x; b < nlist; b += 256)
        if (hist[b]) hist[b] = atomicSub(&lend[b], hist[b]) - hist[b];
    __syncthreads();
    for_entries(ps, pf, probes, e0, e1, [&](const Entry& x) {
        if (x.pass && x.j >= kNearProbes) {
            const uint32_t pos = atomicAdd(&hist[x.l], 1u);
            lq[pos] = x.q;
            lqd[pos] = cd[x.i];
        }
    });
}

// Planning without global atomics (C * nlist <= kWgSlots).  A list's
// queries are placed in C classes: class 0 the queries for which it is one
// of their kNearProbes nearest lists, classes 1 .. C-1 the others by where
// their reach begins in the list's radius interval (reach_frac, in C - 1
// equal steps).  Work items take consecutive queries, so the queries of an
// item reach similar parts of the list, whose rows the int8 shadow holds by
// increasing radius: the tiles below every query's reach are skipped.
// count_wg_kernel writes each workgroup's per-(class, list) counts,
// wg_scan_kernel turns them, per list, into each workgroup's offsets (class
// by class) and the list's total, and fill_wg_kernel places every entry at
// qoff[list] + its workgroup's offset + its rank among the workgroup's
// entries of that class and list (an LDS counter).  The workgroups evaluate
// the same entries with the same arithmetic, so the counts match.
constexpr int kWgSlots = 8192;

template <int C>
__device__ inline int entry_class(const Entry& x) {
    if (x.j < kNearProbes) return 0;
    return 1 + min(C - 2, (int)(x.fr * (float)(C - 1)));
}

template <int C>
__global__ __launch_bounds__(256) void count_wg_kernel(const int32_t* __restrict__ probes, PairSet ps,
                                                       const int32_t* __restrict__ list_size, int nlist,
                                                       PairFilter pf, uint32_t* __restrict__ hwg,
                                                       unsigned long long* cand) {
    __shared__ uint32_t hc[kWgSlots];  // [class][list]
    __shared__ unsigned long long s_c;
    const int64_t e0 = (int64_t)blockIdx.x * kPlanEntries;
    const int64_t e1 = min(ps.size(), e0 + kPlanEntries);
    const int nb = C * nlist;
    uint32_t* on = hwg + (size_t)blockIdx.x * nb;  // this workgroup's counts
    if (e0 >= e1) {  // whole workgroup: no entries, zero counts
        for (int b = threadIdx.x; b < nb; b += 256) on[b] = 0;
        return;
    }
    for (int b = threadIdx.x; b < nb; b += 256) hc[b] = 0;
    if (threadIdx.x == 0) s_c = 0;
    __syncthreads();
    unsigned long long c = 0;
    for_entries(ps, pf, probes, e0, e1, [&](const Entry& x) {
        if (cand && x.in) c += (unsigned long long)list_size[x.l];  // every probed row (the algorithmic work)
        if (x.pass) atomicAdd(&hc[entry_class<C>(x) * nlist + x.l], 1u);
    });
    if (cand && c) atomicAdd(&s_c, c);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += 256) on[b] = hc[b];
    if (cand && threadIdx.x == 0 && s_c) atomicAdd(cand, s_c);
}

// Exclusive scans over the workgroups' counts, per list and in place: class
// c's offsets start after all of the list's entries of lower classes; the
// list's total into lcnt.  A workgroup of 16 waves takes 64 lists, one per
// lane; wave w sums a contiguous run of planning workgroups, the waves'
// sums are combined through LDS, then each wave rewrites its run.
template <int C>
__global__ __launch_bounds__(1024) void wg_scan_kernel(uint32_t* __restrict__ hwg, int nwg, int nlist,
                                                       uint32_t* __restrict__ lcnt, uint32_t* __restrict__ lnear) {
    __shared__ uint32_t sw[C][16][64];
    __shared__ uint32_t st[C][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l = blockIdx.x * 64 + lane;
    const int S = (nwg + 15) / 16;
    const int g0 = min(nwg, w * S), g1 = min(nwg, g0 + S);
    const size_t nb = (size_t)C * nlist;
    // (the run is read kB workgroups at a time, every load issued before
    // any is used; a batch past the run's end re-reads its last workgroup
    // and drops the values)
    constexpr int kB = 32 / C;  // workgroups per batch (32 loads in flight)
    // A run of at most kRegS workgroups (10k queries x 64 probes: 157
    // planning workgroups, runs of 10) is read once, all its loads in flight
    // together, and kept in registers for the rewrite.
    constexpr int kRegS = 10;
    const bool in_regs = S <= kRegS;
    const int lc = min(l, nlist - 1);
    uint32_t a[C];
    uint32_t r[kRegS][C];
#pragma unroll
    for (int c = 0; c < C; ++c) a[c] = 0;
    if (in_regs) {
#pragma unroll
        for (int u = 0; u < kRegS; ++u) {
            const uint32_t* p = hwg + (size_t)min(g0 + u, max(g1 - 1, 0)) * nb + lc;
#pragma unroll
            for (int c = 0; c < C; ++c) r[u][c] = g0 + u < g1 ? p[(size_t)c * nlist] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kRegS; ++u)
#pragma unroll
            for (int c = 0; c < C; ++c) a[c] += r[u][c];
    }
    for (int g = g0; !in_regs && g < g1; g += kB) {
        uint32_t v[kB][C];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const uint32_t* p = hwg + (size_t)min(g + u, g1 - 1) * nb + lc;
#pragma unroll
            for (int c = 0; c < C; ++c) v[u][c] = p[(size_t)c * nlist];
        }
#pragma unroll
        for (int u = 0; u < kB; ++u)
#pragma unroll
            for (int c = 0; c < C; ++c) a[c] += g + u < g1 ? v[u][c] : 0u;
    }
#pragma unroll
    for (int c = 0; c < C; ++c) sw[c][w][lane] = a[c];
    __syncthreads();
    if (w < C) {  // wave c: class c's exclusive prefix over the waves, in place, and its total
        uint32_t run = 0;
        for (int v = 0; v < 16; ++v) {
            const uint32_t x = sw[w][v][lane];
            sw[w][v][lane] = run;
            run += x;
        }
        st[w][lane] = run;
    }
    __syncthreads();
    uint32_t base = 0;  // entries of the lower classes
#pragma unroll
    for (int c = 0; c < C; ++c) {
        a[c] = base + sw[c][w][lane];
        base += st[c][lane];
    }
    if (l >= nlist) return;
    if (w == 0) {
        lcnt[l] = base;
        lnear[l] = st[0][lane];  // class 0: the near queries
    }
    if (in_regs) {
#pragma unroll
        for (int u = 0; u < kRegS; ++u) {
            if (g0 + u >= g1) break;
            uint32_t* p = hwg + (size_t)(g0 + u) * nb + l;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                p[(size_t)c * nlist] = a[c];
                a[c] += r[u][c];
            }
        }
        return;
    }
    for (int g = g0; g < g1; g += kB) {
        uint32_t v[kB][C];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const uint32_t* p = hwg + (size_t)min(g + u, g1 - 1) * nb + l;
#pragma unroll
            for (int c = 0; c < C; ++c) v[u][c] = p[(size_t)c * nlist];
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            if (g + u >= g1) break;
            uint32_t* p = hwg + (size_t)(g + u) * nb + l;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                p[(size_t)c * nlist] = a[c];
                a[c] += v[u][c];
            }
        }
    }
}

// One thread per work item: find its list position by binary search over
// the item offsets (grouped order), then its (query chunk, tile chunk).
struct ItemArgs {
    const uint32_t* lcnt;
    const uint32_t* lnear;  // near queries per list (null: none)
    const uint32_t* qoff;
    const uint32_t* item_off;
    const uint32_t* meta;
    const int32_t* list_size;
    int nlist;
    uint32_t qc, rt;
    int64_t max_items;
    BatchItem* items;
};

// Within a group, the near-query items of its range come first (in list
// order), then the others (in list order): the costly items are claimed
// early, and the queue ends with cheap ones.  Item u of the list order
// belongs to the list l with item_off[l] <= u < item_off[l + 1] and is a
// near-query item when u - item_off[l] < (near items of l); the near items
// before u number near_off[l] + min(u - item_off[l], near items of l).
// io: item_off and near_off (2 (nlist + 1) words), gs: the group starts
// meta[kMetaOff ..] (kXcds + 1 words), read from LDS where the caller staged
// them (the binary searches below are chains of dependent reads).
__device__ void item_one(const ItemArgs& a, int64_t it, const uint32_t* io, const uint32_t* gs) {
    const int nl = a.nlist;
    if (it >= a.max_items || it >= (int64_t)io[nl]) return;
    const uint32_t* no = io + nl + 1;
    const uint32_t u = (uint32_t)it;
    int g = 0;  // the (non-empty) group holding queue position u
    while (g + 1 < kXcds && gs[g + 1] <= u) ++g;
    // the last list l with f(l) <= x (f non-decreasing, f(0) = 0 <= x)
    auto last_le = [&](auto f, uint32_t x) {
        int lo = 0, hi = nl;
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (f(mid) <= x) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    auto near_before = [&](uint32_t v) {  // near items among the first v of the list order
        if (v >= io[nl]) return no[nl];
        const int l = last_le([&](int m) { return io[m]; }, v);
        return no[l] + min(v - io[l], no[l + 1] - no[l]);
    };
    const uint32_t S = gs[g], E = gs[g + 1];
    const uint32_t N0 = near_before(S), NG = near_before(E) - N0;
    const uint32_t r = u - S;
    int l;
    uint32_t local;
    if (r < NG) {
        // the (N0 + r)-th near item of the list order
        const uint32_t x = N0 + r;
        l = last_le([&](int m) { return no[m]; }, x);
        while (no[l + 1] <= x) ++l;  // (lists with no near items share no[])
        local = x - no[l];
    } else {
        // the (S - N0 + r - NG)-th other item
        const uint32_t x = (S - N0) + (r - NG);
        l = last_le([&](int m) { return io[m] - no[m]; }, x);
        while (io[l + 1] - no[l + 1] <= x) ++l;
        local = (no[l + 1] - no[l]) + (x - (io[l] - no[l]));
    }
    const uint32_t t32 = (uint32_t)(a.list_size[l] + 31) / 32;
    const uint32_t nv = no[l + 1] - no[l];  // the list's near-query items
    uint32_t q, t0, ntl;
    if (local < nv) {  // a near-query item: rt / kNearSplit tiles
        ntl = near_tiles(a.rt);
        const uint32_t ntn = (t32 + ntl - 1) / ntl;
        q = local / ntn;
        t0 = (local % ntn) * ntl;
    } else {
        ntl = a.rt;
        const uint32_t ntc = (t32 + a.rt - 1) / a.rt;
        const uint32_t nch = nv ? near_chunks_of(a.lcnt[l], a.lnear[l], a.qc) : 0u;
        q = nch + (local - nv) / ntc;
        t0 = ((local - nv) % ntc) * a.rt;
    }
    BatchItem w;
    w.list = l;
    w.q0 = (int32_t)(a.qoff[l] + q * a.qc);
    w.nq = (int32_t)(min(a.qc, a.lcnt[l] - q * a.qc) | (ntl << 16));  // tiles of the item in the high half
    w.t0 = (int32_t)t0;
    a.items[it] = w;
}

__global__ void items_par_kernel(ItemArgs a) {
    item_one(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, a.item_off, a.meta + kMetaOff);
}

// The fill (workgroups 0 .. nfill - 1) and the work items (the rest, one
// thread per item) in one launch: both only read the plan.
template <int C>
__global__ __launch_bounds__(256) void fill_wg_kernel(const int32_t* __restrict__ probes,
                                                      const float* __restrict__ cd, PairSet ps, int nlist,
                                                      PairFilter pf, const uint32_t* __restrict__ hwg,
                                                      const uint32_t* __restrict__ qoff, int32_t* __restrict__ lq,
                                                      float* __restrict__ lqd, unsigned nfill, ItemArgs ia) {
    __shared__ uint32_t rc[kWgSlots];  // [class][list] next position; item offsets in the item workgroups
    if (blockIdx.x >= nfill) {  // (whole workgroup)
        const int64_t it = (int64_t)(blockIdx.x - nfill) * blockDim.x + threadIdx.x;
        const int nio = 2 * (ia.nlist + 1);
        if (nio + kXcds + 1 <= kWgSlots) {
            // the item and near-item offsets and the group starts staged in
            // LDS, 8 loads in flight per thread
            for (int i0 = 0; i0 < nio; i0 += 256 * 8) {
                uint32_t v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = ia.item_off[min(i0 + j * 256 + (int)threadIdx.x, nio - 1)];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (i0 + j * 256 + (int)threadIdx.x < nio) rc[i0 + j * 256 + threadIdx.x] = v[j];
            }
            if (threadIdx.x <= kXcds) rc[nio + threadIdx.x] = ia.meta[kMetaOff + threadIdx.x];
            __syncthreads();
            item_one(ia, it, rc, rc + nio);
        } else {
            item_one(ia, it, ia.item_off, ia.meta + kMetaOff);
        }
        return;
    }
    const int64_t e0 = (int64_t)blockIdx.x * kPlanEntries;
    const int64_t e1 = min(ps.size(), e0 + kPlanEntries);
    if (e0 >= e1) return;  // whole workgroup
    const int nb = C * nlist;
    const uint32_t* on = hwg + (size_t)blockIdx.x * nb;
    // (4 lists per thread at a time, every load issued before any is used)
    for (int l0 = threadIdx.x; l0 < nlist; l0 += 4 * 256) {
        uint32_t q0[4], v[4][C];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int l = min(l0 + u * 256, nlist - 1);
            q0[u] = qoff[l];
#pragma unroll
            for (int c = 0; c < C; ++c) v[u][c] = on[c * nlist + l];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int l = l0 + u * 256;
            if (l < nlist) {
#pragma unroll
                for (int c = 0; c < C; ++c) rc[c * nlist + l] = q0[u] + v[u][c];
            }
        }
    }
    __syncthreads();
    for_entries(ps, pf, probes, e0, e1, [&](const Entry& x) {
        if (x.pass) {
            const uint32_t pos = atomicAdd(&rc[entry_class<C>(x) * nlist + x.l], 1u);
            lq[pos] = x.q;
            lqd[pos] = cd[x.i];
        }
    });
}



struct BatchArgs {
    const BatchItem* items;
    uint32_t* meta;            // [0] items, [1] queue head
    const int32_t* lq;         // query ids per list
    const uint16_t* xbf;       // bf16 arena
    const uint16_t* qbf;       // [nq x dpad16]
    const uint16_t* qbf2;      // [nq x dpad16], -2 q
    const uint16_t* qaug;      // [nq x 8], norm chunk
    const float* qnorm;
    const int32_t* qint;       // query integer-valued in [0, 255]
    const float* norms;        // [rows] squared norms of list rows
    const int32_t* list_tile_off;
    const int32_t* list_size;
    const float* tau;
    int nkc;
    int aug;                   // the arena carries the norm chunk (stride nkc + 1)
    int dpad16;
    int exact_ok;              // lists integer in [0,255] and d <= 129
    float mcoef;               // margin coefficient for non-exact queries
    const uint32_t* all_int;   // every query of the batch integer-valued in [0, 255]
    uint32_t max_items;        // capacity of the item table
    uint32_t* scnt;            // [nq] survivors per query
    uint64_t* sbuf;            // [nq x cap]
    uint32_t cap;
    int item_tiles;
    // int8 form (integer lists and queries in [0, 255], d <= 256)
    int i8;                    // the int8 arena exists
    const i32x4* x8;           // [tiles32 x nkc8 x 64 lanes] 16 x int8 (x - 128)
    const int32_t* xn8;        // [rows] -sum (x - 128)^2
    const int8_t* q8;          // [nq x dpad32] q - 128
    const int32_t* qn8;        // [nq] sum (q - 128)^2
    int nkc8;
    int dpad32;
    // radius bound: skip a tile when no query of the item can have a row of
    // it within its bound (triangle inequality through the list's centroid)
    const float* lqd;          // squared coarse distance of each (list, query) entry
    const float2* rbound;      // [tiles32] radius interval of the tile's rows
    const float2* rbound_bf16; // the same in arena order (the BF16 form run by the int8 launch)
    unsigned long long* tstat; // [2] tiles of the items, tiles skipped (null: not counted)
};

// Bits r of a lane's 16 accumulator rows ((r / 4) * 8 + half * 4 + r % 4)
// that lie within the first `left` rows of the tile.
__device__ inline uint32_t tile_rows_ok(int left, int half) {
    if (left >= 32) return 0xFFFFu;
    uint32_t m = 0;
#pragma unroll
    for (int rq = 0; rq < 4; ++rq) {
        const int c = min(4, max(0, left - (rq * 8 + half * 4)));
        m |= ((1u << c) - 1u) << (4 * rq);
    }
    return m;
}

// Which scan kernel handles the batch: 2 = int8 (every query and list
// integer in [0, 255], d <= 256), 1 = BF16 with the norm chunk (same, d <=
// 129, no int8 arena), 0 = BF16 with the error margin.
__device__ inline int batch_kind(const BatchArgs& a) {
    const bool all_int = *a.all_int != 0;
    if (a.i8 && all_int) return 2;
    return (a.aug && a.exact_ok && all_int) ? 1 : 0;
}

// Persistent: each wave takes work items from the queue until it is empty.
// Fast form (every query of the item exact and the arena augmented): the
// MFMA chain over -2q and the norm chunk yields the exact distance, and the
// epilogue is one compare per distance.  General form: ||q||^2 + ||x||^2 -
// 2 q.x with the error margin.
template <int NKC, int QG, bool FAST>
__device__ __forceinline__ void batched_scan_body(const BatchArgs& a) {
    // exactly one of the scan kernels works on a batch; none when there is
    // no work (a rescue pass with nothing flagged)
    if (batch_kind(a) != (FAST ? 1 : 0) || a.meta[0] == 0) return;
    // one-query-group waves stage their survivors per query slot in LDS
    // (kStage keys per slot): the keys of many tiles go out with one counted
    // atomic per query, instead of one atomic round trip per tile.  A tile
    // adds at most 32 keys per slot, so a flush before it always makes room.
    constexpr bool kStaged = QG == 1;
    __shared__ uint64_t st_key[kStaged ? kNT / 64 : 1][kStaged ? 32 : 1][kStaged ? kStage : 1];
    const int wv = threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const int slot = lane & 31;
    const int stride = a.nkc + a.aug;
    // the items of this workgroup's XCD group first, then the other groups'
    // leftovers; a group is left once its queue is empty, so the loop ends
    // after at most kXcds empty claims
    // (the group's bounds are read once per group: loads that share a line
    // with a contended counter cost far more than the claims themselves)
    int grp = blockIdx.x % kXcds, switches = 0;
    uint32_t start = a.meta[kMetaOff + grp];
    uint32_t end = min(a.meta[kMetaOff + grp + 1], a.max_items);
    for (;;) {
        uint32_t it = 0;
        if (lane == 0) it = start + atomicAdd(&a.meta[kMetaHead + grp * kHeadStride], 1u);
        it = __shfl(it, 0);
        if (it >= end) {
            if (++switches == kXcds) break;
            grp = (grp + 1) % kXcds;
            start = a.meta[kMetaOff + grp];
            end = min(a.meta[kMetaOff + grp + 1], a.max_items);
            continue;
        }
        const BatchItem w = a.items[it];
        const int nrows = a.list_size[w.list];
        const int t32 = (nrows + 31) >> 5;
        const int te = min(t32, w.t0 + (w.nq >> 16));  // (the item's tiles in the high half of nq)
        const int64_t G0 = (int64_t)a.list_tile_off[w.list] * 2;

        int qid[QG];
        float tau[QG];
#pragma unroll
        for (int g = 0; g < QG; ++g) {
            const int s = g * 32 + slot;
            qid[g] = s < (w.nq & 0xFFFF) ? a.lq[w.q0 + s] : -1;
            tau[g] = qid[g] >= 0 ? a.tau[qid[g]] : -FLT_MAX;
        }
        // keys staged for this lane's query slot since the last flush (the
        // same count in both lanes of the slot)
        uint32_t mycnt = 0;
        auto flush = [&]() {
            uint32_t b0 = 0;
            if (half == 0 && mycnt > 0) b0 = atomicAdd(&a.scnt[qid[0]], mycnt);
            b0 = __shfl(b0, slot);
            for (uint32_t j = half; j < mycnt; j += 2) {
                const uint32_t dst = b0 + j;
                if (dst < a.cap) a.sbuf[(size_t)qid[0] * a.cap + dst] = st_key[wv][slot][j];
            }
            mycnt = 0;
            __builtin_amdgcn_wave_barrier();
        };
        constexpr bool fast = FAST;
        float qn[QG], marg[QG];
        bf16x8 b[QG][NKC];
        bf16x8 bx[QG];
#pragma unroll
        for (int g = 0; g < QG; ++g) {
            const bool live = qid[g] >= 0;
            qn[g] = (!FAST && live) ? a.qnorm[qid[g]] : 0.0f;
            marg[g] = (!FAST && live && !(a.exact_ok && a.qint[qid[g]])) ? a.mcoef : 0.0f;
            const uint16_t* base = fast ? a.qbf2 : a.qbf;
            const bf16x8* src = reinterpret_cast<const bf16x8*>(base + (size_t)max(qid[g], 0) * a.dpad16 + half * 8);
            bf16x8 z = {};
#pragma unroll
            for (int kc = 0; kc < NKC; ++kc) b[g][kc] = (live && kc < a.nkc) ? src[kc * 2] : z;
            bx[g] = (fast && live && half == 0) ? *reinterpret_cast<const bf16x8*>(a.qaug + (size_t)qid[g] * 8) : z;
        }
        // A fragments (list rows) of the first tile; the next tile's are
        // loaded while the current tile's epilogue runs
        bf16x8 av[NKC];
        bf16x8 ax;
        {
            const bf16x8* xa = reinterpret_cast<const bf16x8*>(a.xbf) + (G0 + w.t0) * stride * 64 + lane;
            bf16x8 z = {};
#pragma unroll
            for (int kc = 0; kc < NKC; ++kc) av[kc] = kc < a.nkc ? xa[kc * 64] : z;
            ax = fast ? xa[a.nkc * 64] : z;
        }
        for (int T = w.t0; T < te; ++T) {
            const int64_t G = G0 + T;
            floatx16 acc[QG];
#pragma unroll
            for (int g = 0; g < QG; ++g) acc[g] = floatx16{};
#pragma unroll
            for (int kc = 0; kc < NKC; ++kc) {
                if (kc < a.nkc) {
#pragma unroll
                    for (int g = 0; g < QG; ++g)
                        acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[kc], b[g][kc], acc[g], 0, 0, 0);
                }
            }
            if (fast) {
#pragma unroll
                for (int g = 0; g < QG; ++g)
                    acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ax, bx[g], acc[g], 0, 0, 0);
            }
            if (T + 1 < te) {
                const bf16x8* xa = reinterpret_cast<const bf16x8*>(a.xbf) + (G + 1) * stride * 64 + lane;
#pragma unroll
                for (int kc = 0; kc < NKC; ++kc)
                    if (kc < a.nkc) av[kc] = xa[kc * 64];
                if (fast) ax = xa[a.nkc * 64];
            }
            // rows of this lane that exist (only the list's last tile is partial)
            const uint32_t rows_ok = tile_rows_ok(nrows - T * 32, half);
            float xn[16];
            uint64_t m = 0;
            if constexpr (FAST) {
                // acc is the exact distance: reject the tile with one min per
                // distance when no query of the wave can keep a row
                bool any = false;
#pragma unroll
                for (int g = 0; g < QG; ++g) {
                    float mn = acc[g][0];
#pragma unroll
                    for (int r = 1; r < 16; ++r) mn = fminf(mn, acc[g][r]);
                    any = any || mn <= tau[g];
                }
                if (__ballot(any) == 0) continue;
#pragma unroll
                for (int g = 0; g < QG; ++g) {
                    uint32_t mg = 0;
#pragma unroll
                    for (int r = 0; r < 16; ++r) mg |= (acc[g][r] <= tau[g] ? 1u : 0u) << r;
                    m |= (uint64_t)(mg & rows_ok) << (16 * g);
                }
            } else {
#pragma unroll
                for (int rq = 0; rq < 4; ++rq) {
                    float4 v = *reinterpret_cast<const float4*>(a.norms + G * 32 + rq * 8 + half * 4);
                    xn[rq * 4 + 0] = v.x;
                    xn[rq * 4 + 1] = v.y;
                    xn[rq * 4 + 2] = v.z;
                    xn[rq * 4 + 3] = v.w;
                }
#pragma unroll
                for (int g = 0; g < QG; ++g) {
                    uint32_t mg = 0;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const float sn = qn[g] + xn[r];
                        const float dist = __builtin_fmaf(-2.0f, acc[g][r], sn);
                        const float lo = __builtin_fmaf(-marg[g], sn, dist);
                        mg |= (lo <= tau[g] ? 1u : 0u) << r;
                    }
                    m |= (uint64_t)(mg & rows_ok) << (16 * g);
                }
            }
            if (__ballot(m != 0) == 0) continue;
            if constexpr (kStaged) {
                const uint32_t cg = (uint32_t)__builtin_popcountll(m);
                const uint32_t oc = __shfl_xor(cg, 32);
                if (__ballot(mycnt + cg + oc > (uint32_t)kStage) != 0) flush();
                const uint32_t local = mycnt + (half ? oc : 0u);
                mycnt += cg + oc;
                uint64_t* sk = &st_key[wv][slot][local];
                const uint32_t m16 = (uint32_t)m;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    if ((m16 >> r) & 1u) {
                        const int row = (r >> 2) * 8 + half * 4 + (r & 3);
                        const float dist = fast ? acc[0][r] : __builtin_fmaf(-2.0f, acc[0][r], qn[0] + xn[r]);
                        sk[__builtin_popcount(m16 & ((1u << r) - 1u))] = make_key(dist, (uint32_t)(G * 32 + row));
                    }
                }
                continue;
            }
            // one counted atomic per query group, then the stores
            uint32_t base[QG];
#pragma unroll
            for (int g = 0; g < QG; ++g) {
                const uint32_t cg = __builtin_popcount((uint32_t)(m >> (16 * g)) & 0xFFFFu);
                base[g] = cg ? atomicAdd(&a.scnt[qid[g]], cg) : 0u;
            }
#pragma unroll
            for (int g = 0; g < QG; ++g) {
                uint32_t sl = base[g];
                uint64_t* dst = a.sbuf + (size_t)max(qid[g], 0) * a.cap;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    if ((m >> (16 * g + r)) & 1ull) {
                        if (sl < a.cap) {
                            const int row = (r >> 2) * 8 + half * 4 + (r & 3);
                            const float dist =
                                fast ? acc[g][r] : __builtin_fmaf(-2.0f, acc[g][r], qn[g] + xn[r]);
                            dst[sl] = make_key(dist, (uint32_t)(G * 32 + row));
                        }
                        ++sl;
                    }
                }
            }
        }
        if constexpr (kStaged) {
            if (__ballot(mycnt != 0) != 0) flush();
        }
    }
}

template <int NKC, int QG, bool FAST>
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(FAST && QG == 1 && NKC == 8 ? 4 : 1)))
void batched_scan_kernel(BatchArgs a) {
    batched_scan_body<NKC, QG, FAST>(a);
}

// Int8 form of the scan for integer data.  With x' = x - 128 and q' = q -
// 128 (int8), ||q - x||^2 = ||q'||^2 + ||x'||^2 - 2 q'.x' exactly in int32,
// and q'.x' comes from v_mfma_i32_32x32x32_i8 (32 dims per instruction, half
// the bytes of BF16 per row).  Every distance is an integer below 2^24
// (d <= 256), so it is exact and bit-identical to the FP32 chain; the bound
// compare is done on integers (dist <= tau  <=>  dist <= floor(tau)).
// One query group of 32 per wave, survivors staged in LDS as in the BF16
// form.
template <int NKC8, int QG>
constexpr bool kI8RunsBf16 = NKC8 == 4 && QG == 2;  // d <= 128, 64-query items

template <int NKC8, int QG>
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(NKC8 == 4 ? (QG == 1 ? 3 : 2) : 1)))
void batched_scan_i8_kernel(BatchArgs a) {
    if (a.meta[0] == 0) return;  // no work (a rescue pass with nothing flagged)
    if (batch_kind(a) != 2) {
        // Not every query integer: the 64-query form runs the BF16 form
        // itself (its arena order, its bound intervals), which saves the
        // separate launch that would otherwise follow every int8 scan and
        // find nothing to do.  An int8 index has no norm chunk, so the BF16
        // form here is always the general one.
        if constexpr (kI8RunsBf16<NKC8, QG>) {
            BatchArgs ab = a;
            ab.rbound = a.rbound_bf16;
            batched_scan_body<8, 2, false>(ab);
        }
        return;
    }
    unsigned long long n_tiles = 0, n_skipped = 0;  // tiles of this wave's items, and skipped
    // survivor staging: per wave, one slot of kStage keys per query of the
    // item, each slot padded by one key so that the 32 slots' rows start on
    // different LDS banks (the flush reads key j of every slot at once)
    __shared__ uint64_t st_key[kNT / 64][32 * QG][kStage + 1];
    const int wv = threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const int slot = lane & 31;
    int grp = blockIdx.x % kXcds, switches = 0;
    uint32_t start = a.meta[kMetaOff + grp];
    uint32_t end = min(a.meta[kMetaOff + grp + 1], a.max_items);
    for (;;) {
        uint32_t it = 0;
        if (lane == 0) it = start + atomicAdd(&a.meta[kMetaHead + grp * kHeadStride], 1u);
        it = __shfl(it, 0);
        if (it >= end) {
            if (++switches == kXcds) break;
            grp = (grp + 1) % kXcds;
            start = a.meta[kMetaOff + grp];
            end = min(a.meta[kMetaOff + grp + 1], a.max_items);
            continue;
        }
        const BatchItem w = a.items[it];
        const int nrows = a.list_size[w.list];
        const int t32 = (nrows + 31) >> 5;
        const int te = min(t32, w.t0 + (w.nq >> 16));  // (the item's tiles in the high half of nq)
        const int64_t G0 = (int64_t)a.list_tile_off[w.list] * 2;
        // Query group g of the item: queries g*32 + slot.  Radius bound: a
        // row x of the list (centroid c) is within the query's bound only if
        // | ||q - c|| - ||x - c|| | <= sqrt(tau); a tile is needed when that
        // holds for some query of the item and the tile's radius interval
        // (all widened for rounding).  Integer bound: dist <= tau <=> dist <=
        // floor(tau) (dist >= 0).  With the row norms stored negated,
        //   t = 2 acc - ||x'||^2,  dist = ||x'||^2 + ||q'||^2 - 2 q'.x' = qn - t <= tau_i  <=>  t >= qn - tau_i
        // (|t| < 2^24, tau_i >= -1, qn < 2^23: no overflow; an absent query
        // has b = 0 and qn = 0, so t = -||x'||^2 <= 0 < nthr = 1).
        int qid[QG], qn[QG], nthr[QG];
        float dlo[QG], dhi[QG], sq[QG], dcv[QG], tqv[QG];
        i32x4 b[QG][NKC8];
        uint32_t mycnt[QG];
        // Two rounds of loads, each issued whole before any of it is used:
        // the slots' query ids and coarse distances, then what the ids index
        // (bound, norm, fragments).  A slot past the item's queries loads
        // the last one's entry, then query 0's, and drops them.
        const int nqi = w.nq & 0xFFFF;
        // (and the first chunk's tile intervals, which need only the list's
        // tile offset: an item's first tile lies inside its list)
        const float2 rbl0 = a.rbound[G0 + w.t0 + min(lane, max(min(64, te - w.t0), 1) - 1)];
#pragma unroll
        for (int g = 0; g < QG; ++g) {
            const int e = w.q0 + min(g * 32 + slot, nqi - 1);
            qid[g] = a.lq[e];
            dcv[g] = a.lqd[e];
        }
#pragma unroll
        for (int g = 0; g < QG; ++g) {
            if (g * 32 + slot >= nqi) qid[g] = -1;
            const int qr = max(qid[g], 0);
            tqv[g] = a.tau[qr];
            qn[g] = a.qn8[qr];
            const i32x4* src = reinterpret_cast<const i32x4*>(a.q8 + (size_t)qr * a.dpad32 + half * 16);
#pragma unroll
            for (int kc = 0; kc < NKC8; ++kc) b[g][kc] = src[kc * 2];
        }
#pragma unroll
        for (int g = 0; g < QG; ++g) {
            // (computed for every slot and selected, so that the coarse
            // distance's load stays in the first round)
            const float tq = qid[g] >= 0 ? tqv[g] : -1.0f;
            const bool on = qid[g] >= 0 && tq >= 0.0f;
            const float dc = dcv[g];
            const float dl = sqrtf(fmaxf(dc, 0.0f) * (1.0f - 1e-4f)), dh = sqrtf(dc * (1.0f + 1e-4f)),
                        sv = sqrtf(tq * (1.0f + 1e-4f));  // (NaN off the selected slots: not used)
            dlo[g] = on ? dl : 0.0f;
            dhi[g] = on ? dh : 0.0f;
            sq[g] = on ? sv : -1.0f;
            int tau_i = -1;
            const i32x4 z = {};
            if (qid[g] >= 0) {
                tau_i = tq >= 2147483520.0f ? INT_MAX : (tq >= 0.0f ? (int)tq : -1);
            } else {
                qn[g] = 0;
#pragma unroll
                for (int kc = 0; kc < NKC8; ++kc) b[g][kc] = z;
            }
            nthr[g] = qn[g] - tau_i;
            mycnt[g] = 0;
        }
        // A slot's keys go out as a run of even length (an odd count is
        // padded with kNoKey, which every selection skips), so that runs
        // start at even positions of the query's buffer (cap is even) and
        // each lane stores two keys at once: half 0 keys j, j + 1, half 1
        // keys j + 2, j + 3.
        auto flush = [&](int g) {
            const uint32_t take = (mycnt[g] + 1u) & ~1u;
            uint32_t b0 = 0;
            if (half == 0 && mycnt[g] > 0) {
                b0 = atomicAdd(&a.scnt[qid[g]], take);
                if (take != mycnt[g]) st_key[wv][g * 32 + slot][mycnt[g]] = kNoKey;  // (row has kStage + 1 keys)
            }
            b0 = __shfl(b0, slot);
            __builtin_amdgcn_wave_barrier();
            for (uint32_t j = 2 * half; j < take; j += 4) {
                const uint32_t dst = b0 + j;
                const uint64_t k0 = st_key[wv][g * 32 + slot][j], k1 = st_key[wv][g * 32 + slot][j + 1];
                uint64_t* out = a.sbuf + (size_t)qid[g] * a.cap + dst;
                if (dst + 1 < a.cap)
                    *reinterpret_cast<ulonglong2*>(out) = make_ulonglong2(k0, k1);
                else if (dst < a.cap)
                    *out = k0;
            }
            mycnt[g] = 0;
            __builtin_amdgcn_wave_barrier();
        };
        // a tile's operands: its A fragments (4 KB) and its 32 row norms,
        // one per lane of each half (the epilogue spreads them with LDS
        // permutes: loading each lane's 16 norms directly would cost the
        // vector memory return path as much again as a quarter of the tile)
        auto load = [&](int T, i32x4 (&av)[NKC8], int& xw) {
            const i32x4* xa = a.x8 + (G0 + T) * NKC8 * 64 + lane;
#pragma unroll
            for (int kc = 0; kc < NKC8; ++kc) av[kc] = xa[kc * 64];
            xw = a.xn8[(G0 + T) * 32 + slot];
        };
        // the survivors of query group g in one tile: the common case (no
        // row passes for any query) costs one shift-add per distance, a max3
        // tree and one compare per lane
        auto epilogue = [&](int g, int T, const i32x16& acc, const int (&xn)[16]) {
            // acc[r]: row (r / 4) * 8 + half * 4 + r % 4 of the tile
            int t[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) t[r] = (acc[r] << 1) + xn[r];
            int mx = max(max(t[0], t[1]), t[2]);
#pragma unroll
            for (int r = 3; r < 15; r += 2) mx = max(max(mx, t[r]), t[r + 1]);
            mx = max(mx, t[15]);
            if (__ballot(mx >= nthr[g]) == 0) return;
            uint32_t m = 0;
#pragma unroll
            for (int r = 0; r < 16; ++r) m |= (t[r] >= nthr[g] ? 1u : 0u) << r;
            m &= tile_rows_ok(nrows - T * 32, half);
            if (__ballot(m != 0) == 0) return;
            const uint32_t cg = (uint32_t)__builtin_popcount(m);
            const uint32_t oc = __shfl_xor(cg, 32);
            if (__ballot(mycnt[g] + cg + oc > (uint32_t)kStage) != 0) flush(g);
            const uint32_t local = mycnt[g] + (half ? oc : 0u);
            mycnt[g] += cg + oc;
            uint64_t* sk = &st_key[wv][g * 32 + slot][local];
            const int64_t G = G0 + T;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if ((m >> r) & 1u) {
                    const int row = (r >> 2) * 8 + half * 4 + (r & 3);
                    sk[__builtin_popcount(m & ((1u << r) - 1u))] =
                        make_key((float)(qn[g] - t[r]), (uint32_t)(G * 32 + row));
                }
            }
        };
        // one tile: NKC8 MFMAs per query group (the groups' chains
        // interleaved), then each group's epilogue
        auto tile = [&](int T, const i32x4 (&av)[NKC8], const int xw) {
            int xn[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) xn[r] = __shfl(xw, (r >> 2) * 8 + half * 4 + (r & 3));
            i32x16 acc[QG];
#pragma unroll
            for (int g = 0; g < QG; ++g) acc[g] = i32x16{};
#pragma unroll
            for (int kc = 0; kc < NKC8; ++kc)
#pragma unroll
                for (int g = 0; g < QG; ++g)
                    acc[g] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av[kc], b[g][kc], acc[g], 0, 0, 0);
#pragma unroll
            for (int g = 0; g < QG; ++g) epilogue(g, T, acc[g], xn);
        };
        i32x4 avA[NKC8], avB[NKC8];
        int xnA, xnB;
        // The radius reach of each query group as one interval [rlo, rhi]:
        // rlo = min (dlo - sq), rhi = max (dhi + sq) over its queries, each
        // widened by 2^-18 (dlo + dhi + sq), far more than the rounding of
        // the per-query test !(dlo - ry > sq) && !(rx - dhi > sq).  So every
        // tile [rx, ry] that some query reaches has ry >= rlo and rx <= rhi
        // for its group; a tile that passes without being reached only costs
        // its MFMAs (each row is still tested against each query's bound).
        float rlo[QG], rhi[QG];
#pragma unroll
        for (int g = 0; g < QG; ++g) {
            const bool on = sq[g] >= 0.0f;
            const float m = 0x1p-18f * (dlo[g] + dhi[g] + sq[g]);
            rlo[g] = wave_min(on ? dlo[g] - sq[g] - m : INFINITY);
            rhi[g] = wave_max(on ? dhi[g] + sq[g] + m : -INFINITY);
        }
        // Tiles in chunks of 64: a chunk's needed tiles form one mask (a
        // ballot over the chunk's radius intervals, one per lane); the needed
        // tiles then run with the next one's operands loading meanwhile (two
        // register sets).
        for (int c0 = w.t0; c0 < te; c0 += 64) {
            const int nt = min(64, te - c0);
            const float2 rbl = c0 == w.t0 ? rbl0 : lane < nt ? a.rbound[G0 + c0 + lane] : make_float2(0.0f, 0.0f);
            // (a tile whose interval is not finite is always taken)
            bool nd = !(rbl.x >= -FLT_MAX && rbl.x <= FLT_MAX && rbl.y >= -FLT_MAX && rbl.y <= FLT_MAX);
#pragma unroll
            for (int g = 0; g < QG; ++g) nd = nd || (rbl.y >= rlo[g] && rbl.x <= rhi[g]);
            uint64_t need = __ballot(lane < nt && nd);
            n_skipped += nt - __popcll(need);
            if (need == 0) continue;
            // (the next operands are loaded unconditionally -- the current
            // tile again after the chunk's last -- so that every path has
            // the same loads in flight and the waits before a tile's MFMAs
            // cover only that tile's loads)
            int T = c0 + __builtin_ctzll(need);
            need &= need - 1;
            load(T, avA, xnA);
            for (;;) {
                int Tn = need ? c0 + __builtin_ctzll(need) : -1;
                need &= need - 1;
                load(Tn >= 0 ? Tn : T, avB, xnB);
                tile(T, avA, xnA);
                if (Tn < 0) break;
                T = Tn;
                Tn = need ? c0 + __builtin_ctzll(need) : -1;
                need &= need - 1;
                load(Tn >= 0 ? Tn : T, avA, xnA);
                tile(T, avB, xnB);
                if (Tn < 0) break;
                T = Tn;
            }
        }
#pragma unroll
        for (int g = 0; g < QG; ++g)
            if (__ballot(mycnt[g] != 0) != 0) flush(g);
        n_tiles += te - w.t0;
    }
    if (a.tstat && lane == 0) {
        atomicAdd(&a.tstat[0], n_tiles);
        atomicAdd(&a.tstat[1], n_skipped);
    }
}

// The grid of a persistent kernel capped at the workgroups resident at once
// (cus * the occupancy of kNT-thread workgroups), cached per variant.
template <typename K>
dim3 resident_grid(K kernel, int variant, int cus, dim3 want) {
    static int per_cu[3] = {0, 0, 0};
    if (per_cu[variant] == 0) {
        int n = 0;
        per_cu[variant] =
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kNT, 0) == hipSuccess && n > 0 ? n : -1;
    }
    if (per_cu[variant] < 0) return want;
    return dim3(std::min<unsigned>(want.x, (unsigned)std::max(1, cus * per_cu[variant])));
}

struct SelArgs {
    const int32_t* active;      // only these queries (null = all)
    const int32_t* qlist;       // or: only the queries of this list (count in *qcount)
    const uint32_t* qcount;
    uint32_t* scnt;             // survivor counts; cleared once a query is selected
    const uint64_t* sbuf;
    uint32_t cap;
    const int32_t* qint;
    const float* qnorm;         // [nq] squared query norms (as the BF16 scan used them)
    const float* norms;         // [rows] squared row norms
    float mcoef;                // BF16 margin coefficient (margin_coef)
    int exact_ok;               // BF16 keys exact for integer queries
    int i8;                     // the int8 scan ran when *all_int
    const uint32_t* all_int;
    int keys_exact;             // every key holds the canonical distance (the pair-parallel rescan)
    const float* tau;
    int check_tau;              // learned bound: need k exact survivors <= tau
    const uint32_t* ids;        // [rows]
    const uint32_t* ids8;       // [rows] in int8 shadow order (keys of the int8 scans)
    const float4* codes;        // FP32 tile arena (rescoring)
    const float4* q;            // padded queries
    int nquad;
    int k;
    float* out_d;
    int64_t* out_i;
    int32_t* fallback;          // [nq] flags
    float* rescan_tau;          // [nq] bound for the rescan of flagged queries
    unsigned long long* stats;  // [3] survivors, [fb_slot] flagged queries, [6] overflows of the first pass
    int fb_slot;
    int32_t* flist;             // flagged queries are appended here (count in *fcount)
    uint32_t* fcount;
    uint32_t* lcnt;             // per-list query counts, cleared for the next pass (workgroup 0)
    int nlist;
};

__device__ void select_one(const SelArgs& a, const int qi) {
    __shared__ LdsTopK<kNT, 2048> top;
    __shared__ uint32_t s_real;         // survivors other than padding keys
    __shared__ uint32_t s_np;           // positions waiting for their rescoring
    __shared__ uint32_t pend[2 * kNT];
    extern __shared__ float4 qs[];
    const int tid = threadIdx.x;
    const uint32_t total = a.scnt[qi];
    const bool overflow = total > a.cap;
    const uint32_t n = overflow ? a.cap : total;
    // the key holds the exact distance when the int8 scan ran (every query
    // integer) or the query is integer and the BF16 scan is exact for it
    const bool exact = a.keys_exact || (a.i8 && *a.all_int) || (a.exact_ok && a.qint[qi]);
    const uint32_t* ids = a.i8 && *a.all_int ? a.ids8 : a.ids;  // the keys' positions are shadow rows
    const uint64_t* sb = a.sbuf + (size_t)qi * a.cap;
    const float tau = a.tau[qi];
    if (tid == 0) s_real = s_np = 0;
    top.init();
    uint64_t thr = top.thr;
    uint32_t real = 0;
    const int rounds = (int)((n + kNT - 1) / kNT);
    if (exact) {
        for (int r = 0; r < rounds; ++r) {
            const uint32_t i = (uint32_t)r * kNT + tid;
            const uint64_t key = i < n ? sb[i] : kNoKey;
            if (key != kNoKey) {  // (padding of the int8 scan's runs)
                ++real;
                top.offer(make_key(key_rank(key), ids[(uint32_t)key]), thr);
            }
            __syncthreads();
            top.maybe_compact(a.k, kNT);
            thr = top.thr;
        }
    } else {
        // A BF16 key's distance v is within E = mcoef (||q||^2 + ||x||^2) of
        // the FP32 chain's (the margin the scan admitted the row with).
        // Pass 1: U = the k-th smallest upper bound v + E.  At least k rows
        // have chain distance <= U, so a row with v - E > U is beyond the
        // k-th and is not rescored.  Pass 2 rescores the others, gathered
        // in LDS so that every lane of a round runs a chain.
        const float qn = a.qnorm[qi];
        for (int r = 0; r < rounds; ++r) {
            const uint32_t i = (uint32_t)r * kNT + tid;
            const uint64_t key = i < n ? sb[i] : kNoKey;
            if (key != kNoKey) {
                ++real;
                const uint32_t pos = (uint32_t)key;
                const float sn = qn + a.norms[pos];
                top.offer(make_key(__builtin_fmaf(a.mcoef, sn, key_rank(key)), pos), thr);
            }
            __syncthreads();
            top.maybe_compact(a.k, kNT);
            thr = top.thr;
        }
        top.finish(a.k);
        const float U = top.cnt >= a.k ? key_rank(top.buf[a.k - 1]) : FLT_MAX;
        for (int j = tid; j < a.nquad; j += kNT) qs[j] = a.q[(size_t)qi * a.nquad + j];
        top.init();  // (its barrier orders the reads of buf above before any offer)
        thr = top.thr;
        for (int r = 0; r < rounds; ++r) {
            const uint32_t i = (uint32_t)r * kNT + tid;
            const uint64_t key = i < n ? sb[i] : kNoKey;
            if (key != kNoKey) {
                const uint32_t pos = (uint32_t)key;
                const float sn = qn + a.norms[pos];
                if (__builtin_fmaf(-a.mcoef, sn, key_rank(key)) <= U) pend[atomicAdd(&s_np, 1u)] = pos;
            }
            __syncthreads();
            const uint32_t c = s_np;  // < 2 kNT
            const uint32_t take = r == rounds - 1 ? c : (c >= (uint32_t)kNT ? (uint32_t)kNT : 0u);
            if (take) {
                for (uint32_t b = 0; b < take; b += kNT) {
                    const uint32_t j = b + tid;
                    if (j < take) {
                        const uint32_t pos = pend[j];
                        const float4* tp = a.codes + (size_t)(pos / 64) * a.nquad * 64 + (pos % 64);
                        float acc = 0.0f;
                        for (int jj = 0; jj < a.nquad; ++jj) {
                            const float4 x = tp[(size_t)jj * 64];
                            const float4 qq = qs[jj];
                            float t;
                            t = qq.x - x.x; acc = __builtin_fmaf(t, t, acc);
                            t = qq.y - x.y; acc = __builtin_fmaf(t, t, acc);
                            t = qq.z - x.z; acc = __builtin_fmaf(t, t, acc);
                            t = qq.w - x.w; acc = __builtin_fmaf(t, t, acc);
                        }
                        top.offer(make_key(acc, ids[pos]), thr);
                    }
                    __syncthreads();
                    top.maybe_compact(a.k, kNT);
                    thr = top.thr;
                }
                // keep the rest (fewer than kNT) at the front
                const uint32_t rest = c - take;
                const uint32_t v = (uint32_t)tid < rest ? pend[take + tid] : 0u;
                __syncthreads();
                if ((uint32_t)tid < rest) pend[tid] = v;
                if (tid == 0) s_np = rest;
            }
            __syncthreads();
        }
    }
    top.finish(a.k);
    if (real) atomicAdd(&s_real, real);
    __syncthreads();
    if (tid == 0 && a.stats) atomicAdd(&a.stats[3], (unsigned long long)s_real);  // (padding keys not counted)
    // Exact unless the buffer overflowed or a learned bound fell below the
    // k-th distance (fewer than k survivors within it).  Then the query is
    // rescanned; the k-th of the rows kept here (all real candidates)
    // bounds its k-th distance from above and seeds the rescan's pruning
    // bound.
    if (tid == 0) a.scnt[qi] = 0;  // read by every thread before the barriers above
    const bool tau_ok = top.cnt >= a.k && key_rank(top.buf[a.k - 1]) <= tau;
    if (overflow || (a.check_tau && !tau_ok)) {
        if (tid == 0) {
            a.fallback[qi] = 1;
            a.rescan_tau[qi] = top.cnt >= a.k ? key_rank(top.buf[a.k - 1]) : FLT_MAX;
            a.flist[atomicAdd(a.fcount, 1u)] = qi;
            if (a.stats) atomicAdd(&a.stats[a.fb_slot], 1ull);
            if (a.stats && overflow && a.fb_slot == 4) atomicAdd(&a.stats[6], 1ull);  // first pass
        }
        return;
    }
    const int cnt = top.cnt;
    for (int i = tid; i < a.k; i += kNT) {
        float dv = FLT_MAX;
        int64_t iv = -1;
        if (i < cnt) {
            dv = key_rank(top.buf[i]);
            iv = (int64_t)(uint32_t)top.buf[i];
        }
        a.out_d[(size_t)qi * a.k + i] = dv;
        a.out_i[(size_t)qi * a.k + i] = iv;
    }
}

// One workgroup per query (all, or those flagged in `active`), or, with a
// query list, a grid that walks the list.
__global__ __launch_bounds__(kNT) void select_kernel(SelArgs a) {
    if (a.qlist) {
        const uint32_t nl = *a.qcount;
        for (uint32_t i = blockIdx.x; i < nl; i += gridDim.x) {
            __syncthreads();  // the previous query's LDS reads are done
            select_one(a, a.qlist[i]);
        }
        return;
    }
    const int qi = blockIdx.x;
    if (a.active && !a.active[qi]) return;  // whole workgroup, before any barrier
    select_one(a, qi);
}

// One-wave form of the selection, for queries whose survivor keys already
// hold exact distances (every query of an int8 scan; integer queries of an
// exact BF16 scan): the wave finds a bound V >= the k-th smallest distance
// with two 256-bin linear histograms over the keys (the bins are monotone in
// the distance, so {distance <= V} is a prefix of the sorted keys that holds
// the top k), re-keys those candidates with their ids and sorts them in LDS.
// The keys are read once per pass (they are in L2 from the scan).  The
// result -- and the flags, rescan bounds and counts -- are those of
// select_one.  Every other query (not exact, more than kSelSort candidates,
// k above kSelSort / 2) goes to the workgroup form through hard[] / *nhard.
constexpr int kSelSort = 512;   // candidates one wave sorts in LDS
constexpr int kSelThin = 128;   // more admissible keys than this (and k): thinned by the histograms first

// REG: at most 64 * kSelRegSlots survivors, held in registers (one read of
// the keys); otherwise the keys are read once per pass.
constexpr int kSelRegSlots = 16;
// slots whose keys are loaded together with the survivor count, before it
// is known (positions past survivor_cap read the last slot instead), so the
// count's and the keys' latencies overlap; most queries keep fewer than 512
constexpr int kSelSpecSlots = 8;

// The first kSelSpecSlots slots of this lane's keys are in kr already
// (loaded before n was known); tau and ids8 (the keys' positions are int8
// shadow rows) were loaded with the survivor count.
template <bool REG>
__device__ void select_wave_one(const SelArgs& a, int qi, uint32_t n, bool overflow, uint32_t* hs, uint64_t* kk,
                                int lane, int32_t* hard, uint32_t* nhard, float tau, bool ids8,
                                uint64_t (&kr)[kSelRegSlots]) {
    auto hand_over = [&]() {
        if (lane == 0) hard[atomicAdd(nhard, 1u)] = qi;
    };
    const uint64_t* sb = a.sbuf + (size_t)qi * a.cap;
    if constexpr (REG) {
#pragma unroll
        for (int t = 0; t < kSelRegSlots; ++t) {
            const uint32_t i = (uint32_t)t * 64 + lane;
            if (t < kSelSpecSlots)
                kr[t] = i < n ? kr[t] : ~0ull;
            else
                kr[t] = i < n ? sb[i] : ~0ull;
        }
    }
    // f(key) over this lane's survivors
    auto each = [&](auto&& f) {
        if constexpr (REG) {
#pragma unroll
            for (int t = 0; t < kSelRegSlots; ++t)
                if ((uint32_t)t * 64 + lane < n) f(kr[t]);
        } else {
#pragma unroll 4
            for (uint32_t i = lane; i < n; i += 64) f(sb[i]);
        }
    };
    // distances <= tau (the learned bound's check), the range of the
    // admissible ones (rank < FLT_MAX: the heap's strict admission)
    float mn = INFINITY, mx = -INFINITY;
    uint32_t le = 0, fin = 0;
    each([&](uint64_t key) {
        const float r = key_rank(key);
        le += r <= tau;
        if (key < key_limit()) {
            mn = fminf(mn, r);
            mx = fmaxf(mx, r);
            ++fin;
        }
    });
    mn = wave_min(mn);
    mx = wave_max(mx);
    for (int off = 32; off > 0; off >>= 1) {
        le += __shfl_xor(le, off);
        fin += __shfl_xor(fin, off);
    }
    // candidate bins: all admissible keys when there are few, else the bins
    // up to the k-th (two histogram passes cost less than sorting a few
    // hundred keys more)
    int b1 = 255, sb2 = 255;
    float w1 = 0.0f, inv1 = 0.0f, lo2 = 0.0f, inv2 = 0.0f;
    const bool thin = fin > (uint32_t)max(kSelThin, a.k);
    if (REG && thin) {
        // (at most 1024 keys) one 1024-bin histogram, in the sort buffer
        // (free until the compaction): the k-th's bin holds few keys at
        // this resolution, and the candidates are the bins up to it
        uint32_t* h1k = reinterpret_cast<uint32_t*>(kk);
        wave_clear1024(h1k, lane);
        __builtin_amdgcn_wave_barrier();
        w1 = (mx - mn) * (1.0f / 1024.0f);
        inv1 = w1 > 0.0f ? 1.0f / w1 : 0.0f;
        each([&](uint64_t key) {
            if (key < key_limit()) atomicAdd(&h1k[lin_bin1024(key_rank(key), mn, inv1)], 1u);
        });
        __builtin_amdgcn_wave_barrier();
        b1 = wave_find_bin1024(h1k, (uint32_t)a.k, lane);
        __builtin_amdgcn_wave_barrier();  // the histogram's reads are done before the compaction's writes
    } else if (thin) {
        for (int i = lane; i < 256; i += 64) hs[i] = 0;
        __builtin_amdgcn_wave_barrier();
        w1 = (mx - mn) * (1.0f / 256.0f);
        inv1 = w1 > 0.0f ? 1.0f / w1 : 0.0f;
        each([&](uint64_t key) {
            if (key < key_limit()) atomicAdd(&hs[lin_bin(key_rank(key), mn, inv1)], 1u);
        });
        __builtin_amdgcn_wave_barrier();
        uint32_t need = (uint32_t)a.k;
        b1 = wave_find_bin(hs, &need, lane);
        // keys in the bins up to the k-th's; when they fit one small sort
        // the k-th's bin is not split (sb2 = 255 takes all of it)
        const uint32_t upto = (uint32_t)a.k - need + hs[b1];
        if (upto > (uint32_t)max(kSelThin, a.k)) {
            __builtin_amdgcn_wave_barrier();
            for (int i = lane; i < 256; i += 64) hs[i] = 0;
            __builtin_amdgcn_wave_barrier();
            lo2 = mn + (float)b1 * w1;
            inv2 = w1 > 0.0f ? 256.0f / w1 : 0.0f;
            each([&](uint64_t key) {
                if (key < key_limit()) {
                    const float r = key_rank(key);
                    if (lin_bin(r, mn, inv1) == b1) atomicAdd(&hs[lin_bin(r, lo2, inv2)], 1u);
                }
            });
            __builtin_amdgcn_wave_barrier();
            sb2 = wave_find_bin(hs, &need, lane);
        }
    }
    // compaction (ballots) of the candidates' keys, then re-keyed by
    // (distance, id)
    uint32_t nc = 0;
    auto group = [&](uint64_t key) {
        bool take = key < key_limit();
        if (take && thin) {
            const float r = key_rank(key);
            if (REG) {
                take = lin_bin1024(r, mn, inv1) <= b1;
            } else {
                const int bb = lin_bin(r, mn, inv1);
                take = bb < b1 || (bb == b1 && lin_bin(r, lo2, inv2) <= sb2);
            }
        }
        const uint64_t mk = __ballot(take);
        const uint32_t pos = nc + lanes_below(mk);
        nc += (uint32_t)__popcll(mk);
        if (take && pos < (uint32_t)kSelSort) kk[pos] = key;
    };
    if constexpr (REG) {
#pragma unroll
        for (int t = 0; t < kSelRegSlots; ++t)
            if ((uint32_t)t * 64 < n) group(kr[t]);
    } else {
        for (uint32_t base = 0; base < n; base += 64) {
            const uint32_t i = base + lane;
            group(i < n ? sb[i] : ~0ull);
        }
    }
    if (nc > (uint32_t)kSelSort) {
        hand_over();
        return;
    }
    if (lane == 0) a.scnt[qi] = 0;  // read by every lane at the start
    uint32_t n2 = 64;
    while (n2 < nc) n2 <<= 1;
    __builtin_amdgcn_wave_barrier();
    // the candidates' ids, gathered together after the compaction, every
    // load issued before any is used (one wait for memory)
    const uint32_t* ids = ids8 ? a.ids8 : a.ids;
    uint32_t idv[kSelSort / 64];
#pragma unroll
    for (int t = 0; t < kSelSort / 64; ++t) {
        const uint32_t i = (uint32_t)t * 64 + lane;
        if ((uint32_t)t * 64 < nc) idv[t] = ids[i < nc ? (uint32_t)kk[i] : 0u];
    }
#pragma unroll
    for (int t = 0; t < kSelSort / 64; ++t) {
        const uint32_t i = (uint32_t)t * 64 + lane;
        if (i < nc) kk[i] = (kk[i] & 0xffffffff00000000ull) | idv[t];
    }
    for (uint32_t i = nc + lane; i < n2; i += 64) kk[i] = ~0ull;
    __builtin_amdgcn_wave_barrier();
    wave_sort(kk, n2, lane);
    if (lane == 0 && a.stats) atomicAdd(&a.stats[3], (unsigned long long)fin);  // (padding keys not counted)
    if (overflow || (a.check_tau && le < (uint32_t)a.k)) {
        if (lane == 0) {
            a.fallback[qi] = 1;
            a.rescan_tau[qi] = nc >= (uint32_t)a.k ? key_rank(kk[a.k - 1]) : FLT_MAX;
            a.flist[atomicAdd(a.fcount, 1u)] = qi;
            if (a.stats) atomicAdd(&a.stats[a.fb_slot], 1ull);
            if (a.stats && overflow && a.fb_slot == 4) atomicAdd(&a.stats[6], 1ull);  // first pass
        }
        return;
    }
    for (int i = lane; i < a.k; i += 64) {
        float dv = FLT_MAX;
        int64_t iv = -1;
        if ((uint32_t)i < nc) {
            dv = key_rank(kk[i]);
            iv = (int64_t)(uint32_t)kk[i];
        }
        a.out_d[(size_t)qi * a.k + i] = dv;
        a.out_i[(size_t)qi * a.k + i] = iv;
    }
}

__device__ void select_wave_any(const SelArgs& a, int qi, uint32_t* hs, uint64_t* kk, int lane, int32_t* hard,
                                uint32_t* nhard) {
    // everything the selection reads first, loaded together: the count,
    // the flags, the bound and (when the buffer holds them) the first
    // kSelSpecSlots slots of keys, whatever the count turns out to be
    const uint32_t total = a.scnt[qi];
    const uint32_t all_int = *a.all_int;
    const int32_t qint = a.qint[qi];
    const float tau = a.tau[qi];
    const uint64_t* sb = a.sbuf + (size_t)qi * a.cap;
    uint64_t kr[kSelRegSlots];
#pragma unroll
    for (int t = 0; t < kSelSpecSlots; ++t) kr[t] = sb[min((uint32_t)t * 64 + lane, a.cap - 1)];
    const bool overflow = total > a.cap;
    const uint32_t n = overflow ? a.cap : total;
    const bool ids8 = a.i8 && all_int;
    const bool exact = a.keys_exact || ids8 || (a.exact_ok && qint);
    if (!exact || a.k > kSelSort / 2) {
        if (lane == 0) hard[atomicAdd(nhard, 1u)] = qi;
        return;
    }
    if (n <= 64u * kSelRegSlots)
        select_wave_one<true>(a, qi, n, overflow, hs, kk, lane, hard, nhard, tau, ids8, kr);
    else
        select_wave_one<false>(a, qi, n, overflow, hs, kk, lane, hard, nhard, tau, ids8, kr);
}

// One wave per query: all queries (nq, or those flagged in `active`), or a
// walk over the query list.
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(4))) void select_wave_kernel(SelArgs a, int nq, int32_t* hard, uint32_t* nhard) {
    constexpr int kW = kNT / 64;
    __shared__ uint32_t hist[kW][256];
    __shared__ __align__(16) uint64_t keys[kW][kSelSort];
    const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (blockIdx.x == 0)  // the pass's plan is done with them
        for (int l = threadIdx.x; l < a.nlist; l += kNT) a.lcnt[l] = 0;
    if (a.qlist) {
        const uint32_t nl = *a.qcount;
        for (uint32_t w = blockIdx.x * kW + wv; w < nl; w += gridDim.x * kW) {
            __builtin_amdgcn_wave_barrier();  // the previous query's LDS reads are done
            select_wave_any(a, a.qlist[w], hist[wv], keys[wv], lane, hard, nhard);
        }
        return;
    }
    const int qi = blockIdx.x * kW + wv;
    if (qi >= nq || (a.active && !a.active[qi])) return;
    select_wave_any(a, qi, hist[wv], keys[wv], lane, hard, nhard);
}


// Rescue pass over the listed queries.  Its bound: the k-th distance among
// the rows the first pass kept (an upper bound of D_k when there were k of
// them: a query that overflowed), else the learned bound widened by
// kRescueGrow (a query the learned bound missed; the selection then checks
// that k rows fall within it, and the exact per-query rescan takes the rest).
// Each (query, probed list) pair is a workgroup task (grid-stride), its
// tiles spread over the workgroup's 16 waves.  When the int8 scan serves the batch,
// a tile is 32 rows of the int8 shadow: two lanes per row, each summing its
// half of the dims with v_dot4_i32_i8 (exact int32, the same integer
// distance as the int8 scan); otherwise 64 rows of the FP32 arena, lane per
// row, with the canonical FP32 chain.  Rows within the bound are appended to
// the query's survivor buffer with one counted atomic per wave and tile.
// The keys hold canonical distances, so the one-wave selection applies as
// is.  Pairs that the list's radius interval excludes are skipped, as in the
// plan.
constexpr float kRescueGrow = 1.25f;
constexpr int kRescueNT = 1024;

struct PairScanArgs {
    DevLists lists;
    const float4* q;         // padded queries [nq x nquad]
    const int32_t* probes;   // [nq x P]
    int P;
    const int32_t* npq;      // per-query probe counts (null: P)
    const int32_t* qlist;    // the listed queries (count in *qcount)
    const uint32_t* qcount;
    const float* cd;         // [nq x P] coarse distances
    const float2* lbound;    // list radius intervals
    const float* kept_tau;   // k-th distance among the first pass's kept rows (FLT_MAX: fewer than k)
    const float* tau1;       // the first pass's bound
    int learned;
    float* tau_out;          // the bound used, for the selection
    uint32_t* scnt;
    uint64_t* sbuf;
    uint32_t cap;
    // the int8 form (i8 and *all_int)
    int i8;
    const uint32_t* all_int;
    const i32x4* x8;         // int8 shadow, [32-row tile][chunk][lane][16 int8]
    const int32_t* xn8;      // negated row norms
    const int8_t* q8;        // [nq x dpad32]
    const int32_t* qn8;
    int dpad32, nkc8;
};

__global__ __launch_bounds__(kRescueNT) void pair_scan_kernel(PairScanArgs a) {
    extern __shared__ float4 qs[];  // nquad
    constexpr int kW = kRescueNT / 64;
    const uint32_t ntask = *a.qcount * (uint32_t)a.P;
    const int wv = threadIdx.x / 64, lane = threadIdx.x & 63;
    const int nquad = a.lists.nquad;
    const bool i8 = a.i8 && *a.all_int;
    for (uint32_t e = blockIdx.x; e < ntask; e += gridDim.x) {
        // (every value below is uniform over the workgroup)
        const int w = (int)(e / (uint32_t)a.P), j = (int)(e % (uint32_t)a.P);
        const int qi = a.qlist[w];
        const float kt = a.kept_tau[qi];
        const float tau = kt < FLT_MAX || !a.learned ? kt : kRescueGrow * a.tau1[qi];
        if (j == 0 && threadIdx.x == 0) a.tau_out[qi] = tau;
        const int Pq = a.npq ? min(a.npq[qi], a.P) : a.P;
        const int64_t pi = (int64_t)qi * a.P + j;
        const int l = j < Pq ? a.probes[pi] : -1;
        if (l < 0) continue;
        const int t0 = a.lists.list_tile_off[l];
        const int nt = a.lists.list_tile_off[l + 1] - t0;
        // the list's radius bound (as the plan's pair filter)
        const float dc = a.cd[pi];
        const float dlo = sqrtf(fmaxf(dc, 0.0f) * (1.0f - 1e-4f));
        const float dhi = sqrtf(dc * (1.0f + 1e-4f));
        const float sq = sqrtf(tau * (1.0f + 1e-4f));
        const float2 rb = a.lbound[l];
        if (!(tau >= 0.0f) || dlo - rb.y > sq || rb.x - dhi > sq) continue;
        // (one survivor append per wave and tile)
        auto append = [&](bool pass, float dist, uint32_t pos) {
            const uint64_t mk = __ballot(pass);
            if (mk) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(&a.scnt[qi], (uint32_t)__popcll(mk));
                base = __shfl(base, 0);
                const uint32_t at = base + lanes_below(mk);
                if (pass && at < a.cap) a.sbuf[(size_t)qi * a.cap + at] = make_key(dist, pos);
            }
        };
        const int nrows = a.lists.list_size[l];
        if (i8) {
            // 32-row tiles; no barrier in this form
            for (int t = wv; t < 2 * nt; t += kW) {
                const int64_t G = (int64_t)t0 * 2 + t;
                const int half = lane >> 5, row = lane & 31;
                const i32x4* xa = a.x8 + G * a.nkc8 * 64 + lane;
                const i32x4* qa = reinterpret_cast<const i32x4*>(a.q8 + (size_t)qi * a.dpad32 + half * 16);
                int dot = 0;
                for (int kc = 0; kc < a.nkc8; ++kc) {
                    const i32x4 x = xa[kc * 64], y = qa[kc * 2];
                    dot = __builtin_amdgcn_sdot4(x[0], y[0], dot, false);
                    dot = __builtin_amdgcn_sdot4(x[1], y[1], dot, false);
                    dot = __builtin_amdgcn_sdot4(x[2], y[2], dot, false);
                    dot = __builtin_amdgcn_sdot4(x[3], y[3], dot, false);
                }
                dot += __shfl_xor(dot, 32);
                // dist = ||x'||^2 + ||q'||^2 - 2 q'.x' (exact; norms stored negated)
                const int dist = a.qn8[qi] - (2 * dot + a.xn8[G * 32 + row]);
                const float fd = (float)dist;  // < 2^24: exact
                append(half == 0 && t * 32 + row < nrows && fd <= tau, fd, (uint32_t)(G * 32 + row));
            }
            continue;
        }
        __syncthreads();  // the previous task's reads of qs are done
        for (int t = threadIdx.x; t < nquad; t += kRescueNT) qs[t] = a.q[(size_t)qi * nquad + t];
        __syncthreads();
        for (int t = wv; t < nt; t += kW) {  // 64-row tiles
            const float4* tp = a.lists.codes + (size_t)(t0 + t) * nquad * 64 + lane;
            float acc = 0.0f;
#pragma unroll 16
            for (int jq = 0; jq < nquad; ++jq) acc = accum<kL2>(acc, qs[jq], tp[(size_t)jq * 64]);
            append(t * 64 + lane < nrows && acc <= tau, acc, (uint32_t)((t0 + t) * 64 + lane));
        }
    }
}

// Bound on |fl(||q||^2 + ||x||^2 - 2 q~.x~) - chain(q, x)| for BF16 operands
// (unit roundoff u = 2^-8) with FP32 accumulation over d products
// (gamma_m = m 2^-24 / (1 - m 2^-24)), chain(q, x) being the FP32 FMA chain
// the results are rescored with:
//   |q.x - q~.x~| <= (2u + u^2) ||q|| ||x||,  accumulation adds
//   gamma_d (1 + u)^2 ||q|| ||x||  (taken twice, for any rounding inside the
//   matrix core's sums),  the FP32 norms and the epilogue add at most
//   4 * 2^-24 (||q||^2 + ||x||^2 + 2 ||q|| ||x||),  and the chain differs from
//   ||q - x||^2 by at most gamma_{d+2} ||q - x||^2 <= 2 gamma_{d+2} (...).
// With ||q|| ||x|| <= (||q||^2 + ||x||^2) / 2 the total is
//   E <= c (||q||^2 + ||x||^2),
//   c = (2u + u^2) + 2 gamma_d (1 + u)^2 + 8 * 2^-24 + 2 gamma_{d+2},
// scaled by 1.01 to cover the rounding of the check itself.
float margin_coef(int64_t d) {
    const double u = 1.0 / 256.0;
    const double e = 1.0 / 16777216.0;
    auto gamma = [&](double m) { return m * e / (1.0 - m * e); };
    return (float)(((2 * u + u * u) + 2 * gamma((double)d) * (1 + u) * (1 + u) + 8 * e + 2 * gamma(d + 2.0)) *
                   1.01);
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

void BatchedState::release() {
    if (bf16) hivf::dev_free(bf16);
    if (scratch) hivf::dev_free(scratch);
    if (x8) hivf::dev_free(x8);
    if (xn8) hivf::dev_free(xn8);
    if (rbound) hivf::dev_free(rbound);
    if (lbound) hivf::dev_free(lbound);
    if (rbound8) hivf::dev_free(rbound8);
    if (ids8) hivf::dev_free(ids8);
    rbound = lbound = rbound8 = ids8 = nullptr;
    bf16 = nullptr;
    scratch = nullptr;
    x8 = xn8 = nullptr;
    i8 = 0;
    bf16_bytes = scratch_bytes = 0;
    valid = false;
}

int run_batched(hivf_index* h, const float4* q, int64_t n, int64_t k, int P, const int32_t* npq,
                const hivf_search_params& p, float* D, int64_t* I, hivf_search_stats* st, hipStream_t s,
                bool timing) {
    BatchedState& B = h->batched;
    if (h->metric != HIVF_METRIC_L2) return h->fail(HIVF_EUNSUPPORTED, "batched scan: L2 only");
    const int64_t dpad16 = (h->d + 15) / 16 * 16;
    const int nkc = (int)(dpad16 / 16);
    if (nkc > 32) return h->fail(HIVF_EUNSUPPORTED, "batched scan: d <= 512");
    if (n * P >= INT32_MAX) return h->fail(HIVF_EUNSUPPORTED, "batched scan: queries x nprobe < 2^31");
    const int64_t tiles32 = h->tiles * 2;
    if (!B.valid) {
        B.release();
        B.nkc = nkc;
        // integer lists: the int8 arena (d <= 256, distances below 2^24);
        // otherwise the BF16 norm chunk where it is exact (d <= 129)
        B.i8 = (h->lists_uint8 && h->d <= 256) ? 1 : 0;
        B.aug = (!B.i8 && h->lists_uint8 && h->d <= 129) ? 1 : 0;
        const int stride = nkc + B.aug;
        B.bf16_bytes = (size_t)std::max<int64_t>(tiles32, 1) * stride * 64 * 8 * sizeof(uint16_t);
        HIP_TRY(h, hipMalloc(&B.bf16, B.bf16_bytes));
        int64_t total = tiles32 * stride * 64;
        if (total > 0)
            hipLaunchKernelGGL(build_bf16_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                               h->d_codes.p, h->d_norms.p, h->nquad, tiles32, nkc, stride,
                               static_cast<uint16_t*>(B.bf16));
        HIP_TRY(h, hipGetLastError());
        // tile radius intervals (list of every 32-row tile from the host's
        // tile offsets)
        HIP_TRY(h, hipMalloc(&B.rbound, (size_t)std::max<int64_t>(tiles32, 1) * sizeof(float2)));
        HIP_TRY(h, hipMalloc(&B.lbound, (size_t)std::max<int64_t>(h->nlist, 1) * sizeof(float2)));
        const int64_t rows = std::max<int64_t>(tiles32, 1) * 32;
        DevBuf<float> drr;  // per arena row: distance to its list's centroid
        if (B.i8 && tiles32 > 0) HIP_TRY(h, drr.ensure((size_t)rows));
        if (tiles32 > 0) {
            std::vector<int32_t> tl((size_t)tiles32, -1);
            for (int64_t l = 0; l < h->nlist; ++l)
                for (int64_t t = 2 * (int64_t)h->tile_off[l]; t < 2 * (int64_t)h->tile_off[l + 1]; ++t) tl[t] = (int32_t)l;
            DevBuf<int32_t> dtl;
            HIP_TRY(h, dtl.ensure(tl.size()));
            HIP_TRY(h, hipMemcpyAsync(dtl.p, tl.data(), tl.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(tile_radius_kernel, dim3((unsigned)((tiles32 * 32 + 255) / 256)), dim3(256), 0, s,
                               h->d_codes.p, h->d_cent.p, h->nquad, dtl.p, h->d_tile_off.p, h->d_size.p, tiles32,
                               static_cast<float2*>(B.rbound), drr.p);
            hipLaunchKernelGGL(list_radius_kernel, dim3((unsigned)((h->nlist + 255) / 256)), dim3(256), 0, s,
                               static_cast<const float2*>(B.rbound), h->d_tile_off.p, h->d_size.p, (int)h->nlist,
                               static_cast<float2*>(B.lbound));
            HIP_TRY(h, hipGetLastError());
            HIP_TRY(h, hipStreamSynchronize(s));  // tl and dtl are locals
        }
        if (B.i8) {
            // 4 or 8 chunks of 32 dims (zero past d): the kernel's chunk loop
            // has no runtime bound
            B.nkc8 = h->d <= 128 ? 4 : 8;
            // Shadow order: every list's rows by increasing distance to its
            // centroid (ties by arena position), so that a 32-row tile spans
            // a narrow radius interval and a query whose reach ends inside
            // the list skips the tiles beyond it.  Padding rows keep their
            // place.  The tile intervals are widened as tile_radius_kernel
            // widens them.  Computed on the device (build.hip).
            const int64_t t8 = std::max<int64_t>(tiles32, 1) * B.nkc8 * 64;
            HIP_TRY(h, hipMalloc(&B.x8, (size_t)t8 * 16));
            HIP_TRY(h, hipMalloc(&B.xn8, (size_t)rows * sizeof(int32_t)));
            HIP_TRY(h, hipMalloc(&B.ids8, (size_t)rows * sizeof(uint32_t)));
            HIP_TRY(h, hipMalloc(&B.rbound8, (size_t)std::max<int64_t>(tiles32, 1) * sizeof(float2)));
            if (tiles32 > 0) {
                DevBuf<int32_t> dperm;
                HIP_TRY(h, dperm.ensure((size_t)rows));
                int rc = shadow_order(h, drr.p, dperm.p, static_cast<float2*>(B.rbound8), s);
                if (rc) return rc;
                hipLaunchKernelGGL(build_i8_kernel, dim3((unsigned)((t8 + 255) / 256)), dim3(256), 0, s,
                                   h->d_codes.p, h->nquad, (int)h->d, tiles32, B.nkc8, dperm.p,
                                   static_cast<i32x4*>(B.x8));
                hipLaunchKernelGGL(norms_i8_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s,
                                   h->d_codes.p, h->nquad, (int)h->d, rows, dperm.p, static_cast<int32_t*>(B.xn8));
                hipLaunchKernelGGL(gather_ids_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s,
                                   h->d_ids.p, dperm.p, rows, static_cast<uint32_t*>(B.ids8));
                HIP_TRY(h, hipGetLastError());
                HIP_TRY(h, hipStreamSynchronize(s));  // dperm is a local
            }
        }
        hipDeviceProp_t prop;
        HIP_TRY(h, hipGetDeviceProperties(&prop, h->device));
        B.cus = prop.multiProcessorCount;
        B.valid = true;
    }
    // the int8 scan serves integer lists unless the caller asks for the BF16
    // shadow (HIVF_PREC_BF16); results are the same either way
    const int i8 = (B.i8 && p.precision != HIVF_PREC_BF16) ? 1 : 0;
    // (even: the int8 scan writes its survivor runs two keys at a time)
    const uint32_t cap = p.survivor_cap > 0 ? ((uint32_t)std::min<int64_t>(p.survivor_cap, 1 << 30) + 1u) & ~1u : 16384u;
    // tile shape: queries per item (32 x QG), 32-row tiles per item, waves per CU
    const int qg_max = nkc <= 8 ? 4 : nkc <= 16 ? 2 : 1;  // keeps the B fragments <= 128 VGPRs
    int qg = p.batch_queries > 0 ? std::max(1, std::min(qg_max, p.batch_queries / 32)) : h->tune.qg;
    // default: 64-query items for the int8 scan at d <= 128 (it shares each
    // tile between two query groups), 32 otherwise
    if (qg <= 0) qg = i8 && B.nkc8 == 4 ? 2 : 1;
    if (qg > qg_max) qg = 1;
    if (qg == 3) qg = 2;
    if (i8) qg = B.nkc8 == 4 ? std::min(qg, 2) : 1;  // the int8 kernel takes 32 or (d <= 128) 64 queries per item
    int rt = p.batch_rows > 0 ? std::max(1, p.batch_rows / 32) : h->tune.item_tiles;
    if (rt <= 0) rt = 32;
    int waves = p.batch_waves > 0 ? p.batch_waves : h->tune.waves_per_cu;
    if (waves <= 0) waves = 16;
    const uint32_t qc = 32u * qg;
    const int64_t nlist = h->nlist;
    // upper bound of work items: sum over lists of ceil(c/QC) * ceil(t32/RT)
    int64_t max_t32 = 1;
    for (int64_t l = 0; l < nlist; ++l)
        max_t32 = std::max<int64_t>(max_t32, 2 * (h->tile_off[l + 1] - h->tile_off[l]));
    // (near-query items hold rt / kNearSplit tiles: bound every item by that)
    const int64_t rtn = std::max(1, rt / (int)kNearSplit);
    const int64_t max_items = (n * P / qc + nlist) * ((max_t32 + rtn - 1) / rtn);

    // scratch layout
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align256(off + bytes);
        return o;
    };
    const size_t o_qbf = take((size_t)n * dpad16 * 2), o_qbf2 = take((size_t)n * dpad16 * 2),
                 o_qaug = take((size_t)n * 16), o_qn = take(n * 4), o_qint = take(n * 4),
                 o_lcnt = take(nlist * 4), o_lcur = take(nlist * 4), o_lend = take(nlist * 4),
                 o_qoff = take((nlist + 1) * 4), o_lqd = take((size_t)n * P * 4),
                 o_hwg = take(nlist * 2 <= kWgSlots  // the atomic-free planning's counts
                                  ? (size_t)((n * P + kPlanEntries - 1) / kPlanEntries + 1) * kWgSlots * 4 : 0),
                 o_ioff = take((nlist + 1) * 8), o_lnear = take(nlist * 4), o_meta = take(kMetaWords * 4), o_lq = take((size_t)n * P * 4),
                 o_items = take((size_t)max_items * sizeof(BatchItem)), o_scnt = take(n * 4),
                 o_sbuf = take((size_t)n * cap * 8), o_fb = take(n * 4), o_tau = take(n * 4),
                 o_rtau = take(n * 4), o_tau2 = take(n * 4), o_rtau2 = take(n * 4), o_fb2 = take(n * 4),
                 o_fbl = take(n * 4), o_fbl2 = take(n * 4), o_fbc = take(16), o_shard = take(n * 4),
                 o_allint = take(4), o_q8 = take(i8 ? (size_t)n * B.nkc8 * 32 : 0),
                 o_qn8 = take(i8 ? (size_t)n * 4 : 0);
    if (off > B.scratch_bytes) {
        if (B.scratch) hivf::dev_free(B.scratch);
        B.scratch = nullptr;
        B.scratch_bytes = 0;
        HIP_TRY(h, hipMalloc(&B.scratch, off));
        B.scratch_bytes = off;
    }
    char* base = static_cast<char*>(B.scratch);
    auto at = [&](size_t o) { return static_cast<void*>(base + o); };
    uint16_t* qbf = (uint16_t*)at(o_qbf);
    uint16_t* qbf2 = (uint16_t*)at(o_qbf2);
    uint16_t* qaug = (uint16_t*)at(o_qaug);
    float* qnorm = (float*)at(o_qn);
    int32_t* qint = (int32_t*)at(o_qint);
    uint32_t* lcnt = (uint32_t*)at(o_lcnt);
    uint32_t* lcur = (uint32_t*)at(o_lcur);
    uint32_t* lend = (uint32_t*)at(o_lend);
    float* lqd = (float*)at(o_lqd);
    uint32_t* qoff = (uint32_t*)at(o_qoff);
    uint32_t* hwg = (uint32_t*)at(o_hwg);  // per planning workgroup: [near | far] counts, then offsets
    uint32_t* ioff = (uint32_t*)at(o_ioff);  // item offsets, then near-query item offsets
    uint32_t* lnear = (uint32_t*)at(o_lnear);
    uint32_t* meta = (uint32_t*)at(o_meta);
    int32_t* lq = (int32_t*)at(o_lq);
    BatchItem* items = (BatchItem*)at(o_items);
    uint32_t* scnt = (uint32_t*)at(o_scnt);
    uint64_t* sbuf = (uint64_t*)at(o_sbuf);
    int32_t* fb = (int32_t*)at(o_fb);
    float* tau = (float*)at(o_tau);
    float* rtau = (float*)at(o_rtau);
    h->batched_qnorm = qnorm;

    uint32_t* all_int = (uint32_t*)at(o_allint);
    int8_t* q8 = (int8_t*)at(o_q8);
    int32_t* qn8 = (int32_t*)at(o_qn8);
    const int dpad32 = B.nkc8 * 32;
    float* tau2 = (float*)at(o_tau2);
    float* rtau2 = (float*)at(o_rtau2);
    int32_t* fb2 = (int32_t*)at(o_fb2);
    const int64_t np = n * P;
    const unsigned gnp = (unsigned)((np + 255) / 256);
    // the BF16 scans are exact for integer queries and lists when d <= 129
    const int exact_bf16 = (h->lists_uint8 && h->d <= 129) ? 1 : 0;
    // the queries each rescue pass handles, as lists the selections append
    // to (counts in fbc[0..1]), and the queries each pass's one-wave
    // selection hands to the workgroup form (one list, counts in fbc[2..3])
    int32_t* fbl = (int32_t*)at(o_fbl);
    int32_t* fbl2 = (int32_t*)at(o_fbl2);
    uint32_t* fbc = (uint32_t*)at(o_fbc);
    int32_t* shard = (int32_t*)at(o_shard);
    {
        // the coarse step, whose query-preparation launch also prepares the
        // queries for the scan (else prep_kernel does)
        PrepArgs pa{reinterpret_cast<const float*>(q), (int)h->d, (int)h->d_pad, (int)dpad16, dpad32, qbf, qbf2,
                    qaug, qnorm, qint, i8 ? q8 : nullptr, qn8, fb, fb2, scnt, fbc, lcnt, (int)nlist};
        bool prepped = false;
        const int rc0 = run_coarse(h, q, n, P, s, &pa, &prepped);
        if (rc0) return rc0;
        if (timing) HIP_TRY(h, hipEventRecord(h->ev[1], s));
        if (!prepped) hipLaunchKernelGGL(prep_kernel, dim3((unsigned)n), dim3(64), 0, s, pa);
    }

    // Warm-up bound (src/Index.cpp:999-1051) of all queries, or of the
    // flagged ones only.
    auto warmup = [&](float* out, const int32_t* qlist, const uint32_t* qcount, const float* cap_by) -> int {
        ScanArgs w{};
        w.lists = h->lists();
        w.q = q;
        w.nq = (int)n;
        w.probes = h->d_probes.p;
        w.probe_stride = P;
        w.nprobe_q = npq;
        w.nprobe = P;
        w.k = (int)k;
        w.block_dims = (int)h->d;
        w.nblocks = 1;
        w.row_limit = p.warmup_list_size > 0 ? p.warmup_list_size : (int)k;
        w.warm_lists = p.warmup_lists > 0 ? p.warmup_lists : 10;
        w.out_tau = out;
        w.tau_cap = cap_by;
        w.qlist = qlist;
        w.qcount = qcount;
        HIP_TRY(h, launch_scan(w, kScanWarmup, h->metric, s));
        return HIVF_OK;
    };

    // Top-k of every query's survivors (all queries, or the listed ones):
    // one wave per query where the keys are exact, the rest (listed by the
    // wave kernel) one workgroup per query.  Queries that cannot be finished
    // exactly are flagged in fb_out, with a rescan bound in rtau_out, and
    // appended to the pass's rescue list.
    auto select = [&](const int32_t* active, const int32_t* qlist, const uint32_t* qcount, const float* tau_in,
                      int32_t* fb_out, float* rtau_out, bool check_tau, int fb_slot, bool first,
                      bool keys_exact) -> hipError_t {
        SelArgs sa;
        sa.active = active;
        sa.qlist = qlist;
        sa.qcount = qcount;
        sa.scnt = scnt;
        sa.sbuf = sbuf;
        sa.cap = cap;
        sa.qint = qint;
        sa.qnorm = qnorm;
        sa.norms = h->d_norms.p;
        sa.mcoef = margin_coef(h->d);
        sa.exact_ok = exact_bf16;
        sa.i8 = i8;
        sa.all_int = all_int;
        sa.keys_exact = keys_exact ? 1 : 0;
        sa.tau = tau_in;
        sa.check_tau = check_tau ? 1 : 0;
        sa.ids = h->d_ids.p;
        sa.ids8 = static_cast<const uint32_t*>(B.ids8);
        sa.codes = h->d_codes.p;
        sa.q = q;
        sa.nquad = h->nquad;
        sa.k = (int)k;
        sa.out_d = D;
        sa.out_i = I;
        sa.fallback = fb_out;
        sa.rescan_tau = rtau_out;
        sa.stats = timing ? h->d_stats.p : nullptr;
        sa.fb_slot = fb_slot;
        sa.flist = first ? fbl : fbl2;
        sa.fcount = fbc + (first ? 0 : 1);
        sa.lcnt = lcnt;
        sa.nlist = (int)nlist;
        uint32_t* nhard = fbc + (first ? 2 : 3);
        const int64_t wgs = (n + 3) / 4;
        hipLaunchKernelGGL(select_wave_kernel, dim3((unsigned)(qlist ? std::min<int64_t>(wgs, 1024) : wgs)),
                           dim3(kNT), 0, s, sa, (int)n, shard, nhard);
        sa.active = nullptr;
        sa.qlist = shard;
        sa.qcount = nhard;
        hipLaunchKernelGGL(select_kernel, dim3((unsigned)std::min<int64_t>(n, 1024)), dim3(kNT),
                           h->nquad * sizeof(float4), s, sa);
        return hipGetLastError();
    };

    // One pass: invert the probes of the active queries, batched scan under
    // tau_in, select.  Queries it cannot finish exactly are flagged in
    // fb_out with a rescan bound in rtau_out.
    auto pass = [&](const int32_t* active, const int32_t* qlist, const uint32_t* qcount, const float* tau_in,
                    int32_t* fb_out, float* rtau_out, bool check_tau,
                    int fb_slot, bool first, int rt) -> int {
        unsigned long long* cand = (timing && first) ? h->d_stats.p : nullptr;
        PairFilter pf;
        pf.cd = h->d_cd.p;
        pf.tau = tau_in;
        pf.lbound = static_cast<const float2*>(B.lbound);
        const bool lds = nlist <= kLdsLists;
        const unsigned gpe = (unsigned)((np + kPlanEntries - 1) / kPlanEntries);
        PairSet ps{npq, active, qlist, qcount, (int)n, P};
        const int wgc = nlist * 8 <= kWgSlots ? 8 : nlist * 4 <= kWgSlots ? 4 : nlist * 2 <= kWgSlots ? 2 : 0;
        const ItemArgs ia{lcnt, wgc ? lnear : nullptr, qoff, ioff, meta, h->d_size.p, (int)nlist, qc, (uint32_t)rt,
                          max_items, items};
        const unsigned iwg = (unsigned)((max_items + 255) / 256);
        if (wgc) {
            const unsigned nwg = std::max(1u, gpe);
            const dim3 g1(nwg), gs((unsigned)((nlist + 63) / 64));
            if (wgc == 8) {
                hipLaunchKernelGGL(count_wg_kernel<8>, g1, dim3(256), 0, s, h->d_probes.p, ps, h->d_size.p,
                                   (int)nlist, pf, hwg, cand);
                hipLaunchKernelGGL(wg_scan_kernel<8>, gs, dim3(1024), 0, s, hwg, (int)nwg, (int)nlist, lcnt, lnear);
            } else if (wgc == 4) {
                hipLaunchKernelGGL(count_wg_kernel<4>, g1, dim3(256), 0, s, h->d_probes.p, ps, h->d_size.p,
                                   (int)nlist, pf, hwg, cand);
                hipLaunchKernelGGL(wg_scan_kernel<4>, gs, dim3(1024), 0, s, hwg, (int)nwg, (int)nlist, lcnt, lnear);
            } else {
                hipLaunchKernelGGL(count_wg_kernel<2>, g1, dim3(256), 0, s, h->d_probes.p, ps, h->d_size.p,
                                   (int)nlist, pf, hwg, cand);
                hipLaunchKernelGGL(wg_scan_kernel<2>, gs, dim3(1024), 0, s, hwg, (int)nwg, (int)nlist, lcnt, lnear);
            }
            hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(1024), 0, s, lcnt, lnear, h->d_size.p, (int)nlist, qc,
                               (uint32_t)rt, qoff, lcur, lend, ioff, meta, qint, (int)n, all_int);
            const dim3 gf(nwg + iwg);
            if (wgc == 8)
                hipLaunchKernelGGL(fill_wg_kernel<8>, gf, dim3(256), 0, s, h->d_probes.p, h->d_cd.p, ps, (int)nlist,
                                   pf, hwg, qoff, lq, lqd, nwg, ia);
            else if (wgc == 4)
                hipLaunchKernelGGL(fill_wg_kernel<4>, gf, dim3(256), 0, s, h->d_probes.p, h->d_cd.p, ps, (int)nlist,
                                   pf, hwg, qoff, lq, lqd, nwg, ia);
            else
                hipLaunchKernelGGL(fill_wg_kernel<2>, gf, dim3(256), 0, s, h->d_probes.p, h->d_cd.p, ps, (int)nlist,
                                   pf, hwg, qoff, lq, lqd, nwg, ia);
        } else {
        if (lds)
            hipLaunchKernelGGL(count_lds_kernel, dim3(std::max(1u, gpe)), dim3(256), 0, s, h->d_probes.p, ps,
                               h->d_size.p, (int)nlist, pf, lcnt, cand);
        else
            hipLaunchKernelGGL(count_kernel, dim3(gnp), dim3(256), 0, s, h->d_probes.p, (int)n, P, npq, active,
                               h->d_size.p, pf, lcnt, cand);
        hipLaunchKernelGGL(plan_kernel, dim3(1), dim3(1024), 0, s, lcnt, (const uint32_t*)nullptr, h->d_size.p,
                           (int)nlist, qc, (uint32_t)rt, qoff, lcur, lend, ioff, meta, qint, (int)n, all_int);
        if (lds)
            hipLaunchKernelGGL(fill_lds_kernel, dim3(std::max(1u, gpe)), dim3(256), 0, s, h->d_probes.p, h->d_cd.p,
                               ps, (int)nlist, pf, lcur, lend, lq, lqd);
        else
            hipLaunchKernelGGL(fill_kernel, dim3(gnp), dim3(256), 0, s, h->d_probes.p, h->d_cd.p, (int)n, P, npq,
                               active, pf, lcur, lend, lq, lqd);
        }
        if (!wgc) hipLaunchKernelGGL(items_par_kernel, dim3(iwg), dim3(256), 0, s, ia);
        HIP_TRY(h, hipGetLastError());
        if (timing && first) HIP_TRY(h, hipEventRecord(h->ev[6], s));
        BatchArgs ba;
        ba.items = items;
        ba.meta = meta;
        ba.lq = lq;
        ba.xbf = static_cast<const uint16_t*>(B.bf16);
        ba.qbf = qbf;
        ba.qbf2 = qbf2;
        ba.qaug = qaug;
        ba.qnorm = qnorm;
        ba.qint = qint;
        ba.norms = h->d_norms.p;
        ba.list_tile_off = h->d_tile_off.p;
        ba.list_size = h->d_size.p;
        ba.tau = tau_in;
        ba.nkc = nkc;
        ba.aug = B.aug;
        ba.dpad16 = (int)dpad16;
        ba.exact_ok = exact_bf16;
        ba.i8 = i8;
        ba.x8 = static_cast<const i32x4*>(B.x8);
        ba.xn8 = static_cast<const int32_t*>(B.xn8);
        ba.q8 = q8;
        ba.qn8 = qn8;
        ba.nkc8 = B.nkc8;
        ba.dpad32 = dpad32;
        ba.lqd = lqd;
        ba.tstat = cand ? h->d_stats.p + 1 : nullptr;  // blocks_total, blocks_skipped (first pass)
        ba.rbound = static_cast<const float2*>(B.rbound);
        ba.rbound_bf16 = ba.rbound;
        ba.mcoef = margin_coef(h->d);
        ba.all_int = all_int;
        ba.max_items = (uint32_t)std::min<int64_t>(max_items, UINT32_MAX);
        ba.scnt = scnt;
        ba.sbuf = sbuf;
        ba.cap = cap;
        ba.item_tiles = rt;
        // persistent grid: waves_per_cu waves on each CU (4 waves per workgroup)
        const dim3 grid((unsigned)std::max(1, B.cus * std::max(1, waves / 4)));
        // the fast instantiation exists for the shapes an integer index can
        // take (d <= 129); it returns at once unless every query is exact,
        // the general one returns at once when the fast one ran
        if (i8) {
            BatchArgs b8 = ba;  // tile bounds in shadow order
            b8.rbound = static_cast<const float2*>(B.rbound8);
            // no more workgroups than are resident at once: a later one would
            // only find the queues empty, after a resident one has ended
            if (B.nkc8 == 4 && qg == 2)
                hipLaunchKernelGGL((batched_scan_i8_kernel<4, 2>),
                                   resident_grid((batched_scan_i8_kernel<4, 2>), 0, B.cus, grid), dim3(kNT), 0, s, b8);
            else if (B.nkc8 == 4)
                hipLaunchKernelGGL((batched_scan_i8_kernel<4, 1>),
                                   resident_grid((batched_scan_i8_kernel<4, 1>), 1, B.cus, grid), dim3(kNT), 0, s, b8);
            else
                hipLaunchKernelGGL((batched_scan_i8_kernel<8, 1>),
                                   resident_grid((batched_scan_i8_kernel<8, 1>), 2, B.cus, grid), dim3(kNT), 0, s, b8);
        }
        if (nkc <= 8) {
            if (qg == 4) {
                if (B.aug) hipLaunchKernelGGL((batched_scan_kernel<8, 4, true>), grid, dim3(kNT), 0, s, ba);
                hipLaunchKernelGGL((batched_scan_kernel<8, 4, false>), grid, dim3(kNT), 0, s, ba);
            } else if (qg == 2) {
                if (B.aug) hipLaunchKernelGGL((batched_scan_kernel<8, 2, true>), grid, dim3(kNT), 0, s, ba);
                if (!(i8 && kI8RunsBf16<4, 2> && B.nkc8 == 4))  // else the int8 launch runs it
                    hipLaunchKernelGGL((batched_scan_kernel<8, 2, false>), grid, dim3(kNT), 0, s, ba);
            } else {
                if (B.aug) hipLaunchKernelGGL((batched_scan_kernel<8, 1, true>), grid, dim3(kNT), 0, s, ba);
                hipLaunchKernelGGL((batched_scan_kernel<8, 1, false>), grid, dim3(kNT), 0, s, ba);
            }
        } else if (nkc <= 16) {
            if (B.aug) {
                if (qg == 2)
                    hipLaunchKernelGGL((batched_scan_kernel<16, 2, true>), grid, dim3(kNT), 0, s, ba);
                else
                    hipLaunchKernelGGL((batched_scan_kernel<16, 1, true>), grid, dim3(kNT), 0, s, ba);
            }
            if (qg == 2)
                hipLaunchKernelGGL((batched_scan_kernel<16, 2, false>), grid, dim3(kNT), 0, s, ba);
            else
                hipLaunchKernelGGL((batched_scan_kernel<16, 1, false>), grid, dim3(kNT), 0, s, ba);
        } else {
            hipLaunchKernelGGL((batched_scan_kernel<32, 1, false>), grid, dim3(kNT), 0, s, ba);
        }
        HIP_TRY(h, hipGetLastError());
        if (timing && first) HIP_TRY(h, hipEventRecord(h->ev[3], s));
        HIP_TRY(h, select(active, qlist, qcount, tau_in, fb_out, rtau_out, check_tau, fb_slot, first, false));
        HIP_TRY(h, hipGetLastError());
        return HIVF_OK;
    };

    // pass 1: every query under the requested bound
    const bool learned = p.tau_source == HIVF_TAU_LEARNED;
    int rc = learned ? run_tau_model(h, q, n, P, npq, tau, s) : warmup(tau, nullptr, nullptr, nullptr);
    if (rc) return rc;
    if (timing) HIP_TRY(h, hipEventRecord(h->ev[2], s));
    rc = pass(nullptr, nullptr, nullptr, tau, fb, rtau, learned, 4, true, rt);
    if (rc) return rc;
    if (timing) HIP_TRY(h, hipEventRecord(h->ev[4], s));
    // Under the learned bound, flagged queries are few (the model is trained
    // and its offset tuned for a small miss rate): where one query's exact
    // rescan is small (at most 2^24 list elements at the mean list size:
    // 1M x 128 at nprobe 128) they go straight to the exact per-query rescan below, one launch, where the
    // pair-parallel pass and its selection would cost three launches in every
    // search that flags none.  Larger probe sets (a miss over 10M rows at
    // nprobe 256 would hold one workgroup for milliseconds), and every
    // warm-up-bound search (overflowing survivor buffers can flag many
    // queries), take pass 2 first.
    const double rescan_elems = (double)h->ntotal / (double)std::max<int64_t>(1, nlist) * (double)P * (double)h->d;
    const bool pair_rescue = !learned || rescan_elems > (double)(1 << 24);
    // pass 2: the flagged queries, each (query, probed list, tile group) a
    // workgroup task, bounded by the k-th of the rows pass 1 kept or, for a
    // learned-bound miss, the widened learned bound
    if (pair_rescue) {
        PairScanArgs ra;
        ra.lists = h->lists();
        ra.q = q;
        ra.probes = h->d_probes.p;
        ra.P = P;
        ra.npq = npq;
        ra.qlist = fbl;
        ra.qcount = fbc;
        ra.cd = h->d_cd.p;
        ra.lbound = static_cast<const float2*>(B.lbound);
        ra.kept_tau = rtau;
        ra.tau1 = tau;
        ra.learned = learned ? 1 : 0;
        ra.tau_out = tau2;
        ra.scnt = scnt;
        ra.sbuf = sbuf;
        ra.cap = cap;
        ra.i8 = i8 ? 1 : 0;
        ra.all_int = all_int;
        ra.x8 = static_cast<const i32x4*>(B.x8);
        ra.xn8 = static_cast<const int32_t*>(B.xn8);
        ra.q8 = q8;
        ra.qn8 = qn8;
        ra.dpad32 = dpad32;
        ra.nkc8 = B.nkc8;
        hipLaunchKernelGGL(pair_scan_kernel, dim3((unsigned)std::min<int64_t>(n * P, 2048)), dim3(kRescueNT),
                           h->nquad * sizeof(float4), s, ra);
        HIP_TRY(h, hipGetLastError());
        HIP_TRY(h, select(nullptr, fbl, fbc, tau2, fb2, rtau2, true, 5, false, true));
    }
    // last resort: exact per-query rescan of what pass 2 could not finish
    // (or of every flagged query, bounded by pass 1's rescan bound)
    ScanArgs f{};
    f.lists = h->lists();
    f.q = q;
    f.nq = (int)n;
    f.probes = h->d_probes.p;
    f.probe_stride = P;
    f.nprobe_q = npq;
    f.nprobe = P;
    f.k = (int)k;
    pruning_blocks(h->d, 0, &f.block_dims, &f.nblocks);
    f.tau = pair_rescue ? rtau2 : rtau;
    f.row_limit = INT_MAX;
    f.warm_lists = INT_MAX;
    f.out_d = D;
    f.out_i = I;
    f.qlist = pair_rescue ? fbl2 : fbl;
    f.qcount = pair_rescue ? fbc + 1 : fbc;
    HIP_TRY(h, launch_scan(f, kScanPruned, h->metric, s));
    if (timing) {
        HIP_TRY(h, hipEventRecord(h->ev[5], s));
        HIP_TRY(h, hipEventSynchronize(h->ev[5]));
        unsigned long long c[8];
        HIP_TRY(h, hipMemcpy(c, h->d_stats.p, sizeof(c), hipMemcpyDeviceToHost));
        std::memset(st, 0, sizeof(*st));
        st->coarse_ms = elapsed_ms(h->ev[0], h->ev[1]);
        st->tau_ms = elapsed_ms(h->ev[1], h->ev[2]);
        st->scan_ms = elapsed_ms(h->ev[6], h->ev[3]);
        st->select_ms = elapsed_ms(h->ev[3], h->ev[4]);
        st->fallback_ms = elapsed_ms(h->ev[4], h->ev[5]);
        st->total_ms = elapsed_ms(h->ev[0], h->ev[5]);
        st->survivors = (int64_t)c[3];
        st->fallback_queries = (int64_t)c[4];  // rescanned by the second batched pass
        st->scan_launches = 2;
        st->exact_rescans = (int64_t)(pair_rescue ? c[5] : c[4]);
        st->overflow_queries = (int64_t)c[6];
        st->batch_queries = qg * 32;
        st->batch_rows = rt * 32;
        st->batch_waves = waves;
        st->plan_ms = elapsed_ms(h->ev[2], h->ev[6]);
        uint32_t ai = 0;
        HIP_TRY(h, hipMemcpy(&ai, all_int, sizeof(ai), hipMemcpyDeviceToHost));
        st->scan_kind = (i8 && ai) ? 2 : ((B.aug && exact_bf16 && ai) ? 1 : 0);
        st->blocks_total = (int64_t)c[1];    // 32-row tiles of the first pass's work items
        st->blocks_skipped = (int64_t)c[2];  // of which skipped by the radius bound
        st->candidates = (int64_t)c[0];
    }
    return HIVF_OK;
}

}  // namespace hivf
