// Per-query kernels of the IVF-Flat query path on CDNA4:
//   coarse_kernel : query x centroid distances + top-nprobe
//                   (src/Index.cpp:967-995, src/IVFScan.hpp:108-138)
//   scan_kernel   : inverted-list scan with top-k, three modes
//       plain   : full distances (src/Index.cpp:1420-1554,
//                 src/IVFScan.hpp:247-272)
//       pruned  : dimension-block partial distances, a row is dropped once
//                 its running distance exceeds min(tau_q, running k-th)
//                 (src/node.cpp:606-678, 1056-1174)
//       warm-up : k-th distance over the first W rows of the first L probed
//                 lists (src/Index.cpp:999-1051)
//
// One workgroup of 256 threads per query.  Each wavefront owns one tile of
// 64 list rows per round, one row per lane, and streams it with coalesced
// 1 KiB loads (tile layout in kernels.hpp); the query sits in LDS and is
// read with broadcast loads.  A distance is one sequential FMA chain over
// the dims, the same arithmetic as the oracle, so results are bit-identical.
#include <algorithm>

#include "common.hpp"
#include "kernels.hpp"
#include "select.hpp"

namespace hivf {

namespace {

constexpr int kNT = 256;
constexpr int kNW = kNT / kWave;
constexpr int kBuf = 2048;
constexpr int kListGrid = 1024;  // workgroups of a query-list launch


__device__ inline void load_query(float4* qs, const float4* q, int nquad) {
    for (int j = threadIdx.x; j < nquad; j += kNT) qs[j] = q[j];
}

__device__ inline float quad_comp(float4 v, int c) {
    return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w;
}

// Dims [d0, d1) of the canonical chain of one row: a partial first quad,
// whole quads, a partial last quad (the partial steps only where a block
// boundary falls inside a quad).
template <int METRIC>
__device__ inline float accum_dims(float acc, const float4* qs, const float4* tp, int d0, int d1) {
    int j = d0 >> 2;
    const int c0 = d0 & 3;
    if (c0) {
        const float4 q = qs[j], x = tp[(size_t)j * kWave];
        const int c1 = min(4, d1 - 4 * j);
        for (int c = c0; c < c1; ++c) {
            const float qv = quad_comp(q, c), xv = quad_comp(x, c);
            if constexpr (METRIC == kL2) {
                const float t = qv - xv;
                acc = __builtin_fmaf(t, t, acc);
            } else {
                acc = __builtin_fmaf(qv, xv, acc);
            }
        }
        ++j;
    }
    const int jend = d1 >> 2;
#pragma unroll 8
    for (; j < jend; ++j) acc = accum<METRIC>(acc, qs[j], tp[(size_t)j * kWave]);
    const int r = d1 & 3;
    if (r && 4 * j >= d0 && 4 * j < d1) {
        const float4 q = qs[j], x = tp[(size_t)j * kWave];
        for (int c = 0; c < r; ++c) {
            const float qv = quad_comp(q, c), xv = quad_comp(x, c);
            if constexpr (METRIC == kL2) {
                const float t = qv - xv;
                acc = __builtin_fmaf(t, t, acc);
            } else {
                acc = __builtin_fmaf(qv, xv, acc);
            }
        }
    }
    return acc;
}

// Tiles of probe j of a query (rows capped at row_limit).
__device__ inline int probe_tiles(const ScanArgs& a, const int32_t* probes, int j) {
    int l = probes[j];
    if (l < 0) return 0;
    int rows = a.lists.list_size[l];
    if (rows > a.row_limit) rows = a.row_limit;
    return (rows + kTileRows - 1) / kTileRows;
}

// One query, by the whole workgroup.
template <int MODE, int METRIC>
__device__ void scan_one(const ScanArgs& a, const int qi) {
    __shared__ LdsTopK<kNT, kBuf> top;
    __shared__ int s_tiles;
    __shared__ unsigned long long s_stats[3];
    extern __shared__ float4 qs[];

    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int lane = tid % kWave;
    const int nquad = a.lists.nquad;
    const int32_t* probes = a.probes + (size_t)qi * a.probe_stride;
    int P = a.nprobe_q ? a.nprobe_q[qi] : a.nprobe;
    if (P > a.probe_stride) P = a.probe_stride;
    if (MODE == kScanWarmup && P > a.warm_lists) P = a.warm_lists;  // src/Index.cpp:1008-1011

    top.init();
    if (tid == 0) {
        s_tiles = 0;
        s_stats[0] = s_stats[1] = s_stats[2] = 0;
    }
    load_query(qs, a.q + (size_t)qi * nquad, nquad);
    __syncthreads();
    {
        int t = 0;
        for (int j = tid; j < P; j += kNT) t += probe_tiles(a, probes, j);
        if (t) atomicAdd(&s_tiles, t);
    }
    __syncthreads();
    const int total = s_tiles;
    const int rounds = (total + kNW - 1) / kNW;

    // wave cursor: probe j, tile t within that list, for global tile `wave`
    int j = 0, t = wave;
    int nt_j = P > 0 ? probe_tiles(a, probes, 0) : 0;
    while (j < P && t >= nt_j) {
        t -= nt_j;
        ++j;
        nt_j = j < P ? probe_tiles(a, probes, j) : 0;
    }

    const float tau_q = (MODE == kScanPruned && a.tau) ? a.tau[qi] : FLT_MAX;
    const int nblocks = MODE == kScanPruned ? a.nblocks : 1;
    unsigned long long my_cand = 0, my_skip = 0;
    uint64_t thr = top.thr;

    for (int r = 0; r < rounds; ++r) {
        if (j < P) {
            const int l = probes[j];
            int rows = a.lists.list_size[l];
            if (rows > a.row_limit) rows = a.row_limit;
            const int tile = a.lists.list_tile_off[l] + t;
            const bool valid = lane < rows - t * kTileRows;
            const float4* tp = a.lists.codes + (size_t)tile * nquad * kWave + lane;
            float acc = 0.0f;
            bool keep = valid;
            if constexpr (MODE == kScanPruned) {
                const float bound = fminf(tau_q, key_rank(thr));
                for (int b = 0; b < nblocks; ++b) {
                    if (__ballot(keep) == 0) break;
                    if (keep) {
                        const int d0 = b * a.block_dims;
                        const int d1 = b < nblocks - 1 ? d0 + a.block_dims : 4 * nquad;
                        acc = accum_dims<METRIC>(acc, qs, tp, d0, d1);
                        if (b < nblocks - 1 && acc > bound) {
                            keep = false;
                            my_skip += nblocks - 1 - b;
                        }
                    }
                }
            } else {
                if (valid) {
#pragma unroll 8
                    for (int jj = 0; jj < nquad; ++jj)
                        acc = accum<METRIC>(acc, qs[jj], tp[(size_t)jj * kWave]);
                }
            }
            if (valid) my_cand++;
            if (keep) {
                uint32_t id = a.lists.ids[(size_t)tile * kWave + lane];
                top.offer(make_key(rank_of<METRIC>(acc), id), thr);
            }
            // advance this wave's cursor by kNW tiles
            t += kNW;
            while (j < P && t >= nt_j) {
                t -= nt_j;
                ++j;
                nt_j = j < P ? probe_tiles(a, probes, j) : 0;
            }
        }
        __syncthreads();
        top.maybe_compact(a.k, kNT);
        thr = top.thr;
    }
    top.finish(a.k);

    if (a.stats) {
        if (my_cand) atomicAdd(&s_stats[0], my_cand);
        if (my_skip) atomicAdd(&s_stats[2], my_skip);
        __syncthreads();
        if (tid == 0) {
            atomicAdd(&a.stats[0], s_stats[0]);
            atomicAdd(&a.stats[1], s_stats[0] * (unsigned long long)nblocks);
            atomicAdd(&a.stats[2], s_stats[2]);
        }
    }

    if constexpr (MODE == kScanWarmup) {
        if (tid == 0) {
            float v = sentinel<METRIC>();
            if (top.cnt >= a.k) {
                float rk = key_rank(top.buf[a.k - 1]);
                v = dist_of_rank(METRIC, rk);
            }
            if (a.tau_cap) v = fminf(v, a.tau_cap[qi]);
            a.out_tau[qi] = v;
        }
    } else {
        const int cnt = top.cnt;
        for (int i = tid; i < a.k; i += kNT) {
            float dv = sentinel<METRIC>();
            int64_t iv = -1;
            if (i < cnt) {
                uint64_t key = top.buf[i];
                float rk = key_rank(key);
                dv = dist_of_rank(METRIC, rk);
                iv = (int64_t)(uint32_t)key;
            }
            a.out_d[(size_t)qi * a.k + i] = dv;
            a.out_i[(size_t)qi * a.k + i] = iv;
        }
    }
}

// One workgroup per query (all, or those flagged in `only`), or, with a
// query list, a grid that walks the list.
template <int MODE, int METRIC>
__global__ __launch_bounds__(kNT) void scan_kernel(ScanArgs a) {
    if (a.qlist) {
        const uint32_t nl = *a.qcount;
        for (uint32_t i = blockIdx.x; i < nl; i += gridDim.x) {
            __syncthreads();  // the previous query's LDS reads are done
            scan_one<MODE, METRIC>(a, a.qlist[i]);
        }
        return;
    }
    const int qi = blockIdx.x;
    if (a.only && !a.only[qi]) return;  // whole workgroup, before any barrier
    scan_one<MODE, METRIC>(a, qi);
}

template <int METRIC>
__global__ __launch_bounds__(kNT) void coarse_kernel(CoarseArgs a) {
    __shared__ LdsTopK<kNT, kBuf> top;
    extern __shared__ float4 qs[];
    const int qi = blockIdx.x;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int lane = tid % kWave;
    top.init();
    load_query(qs, a.q + (size_t)qi * a.nquad, a.nquad);
    __syncthreads();
    const int tiles = (a.nlist + kTileRows - 1) / kTileRows;
    const int rounds = (tiles + kNW - 1) / kNW;
    uint64_t thr = top.thr;
    for (int r = 0; r < rounds; ++r) {
        const int tile = r * kNW + wave;
        const int c = tile * kTileRows + lane;
        if (tile < tiles && c < a.nlist) {
            const float4* tp = a.cent + (size_t)tile * a.nquad * kWave + lane;
            float acc = 0.0f;
#pragma unroll 8
            for (int jj = 0; jj < a.nquad; ++jj)
                acc = accum<METRIC>(acc, qs[jj], tp[(size_t)jj * kWave]);
            top.offer(make_key(rank_of<METRIC>(acc), (uint32_t)c), thr);
        }
        __syncthreads();
        top.maybe_compact(a.P, kNT);
        thr = top.thr;
    }
    top.finish(a.P);
    const int cnt = top.cnt;
    for (int i = tid; i < a.P; i += kNT) {
        float dv = sentinel<METRIC>();
        int32_t lv = -1;
        if (i < cnt) {
            uint64_t key = top.buf[i];
            float rk = key_rank(key);
            dv = dist_of_rank(METRIC, rk);
            lv = (int32_t)(uint32_t)key;
        }
        a.out_d[(size_t)qi * a.P + i] = dv;
        a.out_l[(size_t)qi * a.P + i] = lv;
    }
}

template <int MODE>
hipError_t launch_scan_mode(const ScanArgs& a, int metric, hipStream_t s) {
    size_t lds = (size_t)a.lists.nquad * sizeof(float4);
    const dim3 grid((unsigned)(a.qlist ? std::min(a.nq, kListGrid) : a.nq));
    if (metric == kL2)
        hipLaunchKernelGGL((scan_kernel<MODE, kL2>), grid, dim3(kNT), lds, s, a);
    else
        hipLaunchKernelGGL((scan_kernel<MODE, kIP>), grid, dim3(kNT), lds, s, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_scan(const ScanArgs& a, int mode, int metric, hipStream_t s) {
    if (a.nq <= 0) return hipSuccess;
    switch (mode) {
        case kScanPlain: return launch_scan_mode<kScanPlain>(a, metric, s);
        case kScanPruned: return launch_scan_mode<kScanPruned>(a, metric, s);
        default: return launch_scan_mode<kScanWarmup>(a, metric, s);
    }
}

hipError_t launch_coarse(const CoarseArgs& a, int metric, hipStream_t s) {
    if (a.nq <= 0) return hipSuccess;
    size_t lds = (size_t)a.nquad * sizeof(float4);
    if (metric == kL2)
        hipLaunchKernelGGL((coarse_kernel<kL2>), dim3(a.nq), dim3(kNT), lds, s, a);
    else
        hipLaunchKernelGGL((coarse_kernel<kIP>), dim3(a.nq), dim3(kNT), lds, s, a);
    return hipGetLastError();
}

}  // namespace hivf
