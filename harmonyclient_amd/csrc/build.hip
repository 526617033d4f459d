// Index build on the GPU: the Lloyd update of k-means and the list fill of
// Index::add.
//
//   k-means (src/Clustering.cpp:93-152, src/Index.cpp:381-425): every round
//     assigns each training row to its nearest centroid (the coarse kernel
//     with one probe, lowest id on exact ties), groups the rows by centroid
//     in row order, and sets each centroid to the mean of its rows,
//     accumulated in FP64 in row order and rounded once to FP32 (a centroid
//     without rows keeps its value).  The fixed order makes the result
//     independent of scheduling, unlike the reference's `omp critical`
//     accumulation.
//   add (src/Index.cpp:581-698): assigns every row, orders each list by
//     distance to its centroid (ascending for L2, descending for IP; equal
//     distances by row id -- the canonical form of the reference's unstable
//     std::sort, src/Index.cpp:643-651), and writes the tiled list arena,
//     the ids and the FP64-accumulated squared norms.
//
// Grouping and ordering are stable least-significant-digit radix passes of
// 8 bits: each pass counts digits per 1024-element chunk (one wave per
// chunk, an LDS histogram), scans the counts of every digit over the chunks,
// and places each element at its digit's base + its chunk's offset + its
// rank among the chunk's earlier elements with the same digit.  Elements
// start in row order, so every pass keeps equal keys in row order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "index_impl.hpp"
#include "wave.hpp"

namespace hivf {

namespace {

constexpr int kChunk = 1024;  // elements per chunk of a radix pass
constexpr int kDigits = 256;  // 8-bit digits

// Digit counts of every chunk: counts[digit * nchunks + chunk].
__global__ __launch_bounds__(64) void rp_count_kernel(const uint32_t* __restrict__ key,
                                                      const int32_t* __restrict__ perm_in, int shift, int64_t n,
                                                      int nchunks, uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[kDigits];
    const int c = blockIdx.x, lane = threadIdx.x;
    for (int b = lane; b < kDigits; b += 64) hist[b] = 0;
    __syncthreads();
    const int64_t i0 = (int64_t)c * kChunk;
    const int64_t i1 = std::min<int64_t>(n, i0 + kChunk);
    for (int64_t i = i0 + lane; i < i1; i += 64) {
        const int32_t e = perm_in ? perm_in[i] : (int32_t)i;
        atomicAdd(&hist[(key[e] >> shift) & (kDigits - 1)], 1u);
    }
    __syncthreads();
    for (int b = lane; b < kDigits; b += 64) counts[(size_t)b * nchunks + c] = hist[b];
}

// Per digit (one workgroup each): exclusive scan of its chunk counts, in
// place, and the digit's total.
__global__ __launch_bounds__(256) void rp_scan_kernel(uint32_t* __restrict__ counts, int nchunks,
                                                      uint32_t* __restrict__ total) {
    __shared__ uint32_t part[256];
    uint32_t* col = counts + (size_t)blockIdx.x * nchunks;
    const int t = threadIdx.x;
    const int per = (nchunks + 255) / 256;
    const int c0 = t * per, c1 = std::min(nchunks, c0 + per);
    uint32_t s = 0;
    for (int c = c0; c < c1; ++c) s += col[c];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (int i = 0; i < 256; ++i) {
            const uint32_t v = part[i];
            part[i] = run;
            run += v;
        }
        total[blockIdx.x] = run;
    }
    __syncthreads();
    uint32_t run = part[t];
    for (int c = c0; c < c1; ++c) {
        const uint32_t v = col[c];
        col[c] = run;
        run += v;
    }
}

// Stable placement: one wave per chunk walks its elements 64 at a time in
// order; lanes with equal digits take consecutive positions by lane.
__global__ __launch_bounds__(64) void rp_fill_kernel(const uint32_t* __restrict__ key,
                                                     const int32_t* __restrict__ perm_in, int shift, int64_t n,
                                                     int nchunks, const uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ total,
                                                     int32_t* __restrict__ perm_out) {
    __shared__ uint32_t next[kDigits];
    const int c = blockIdx.x, lane = threadIdx.x;
    if (lane == 0) {
        uint32_t run = 0;
        for (int b = 0; b < kDigits; ++b) {
            next[b] = run + counts[(size_t)b * nchunks + c];
            run += total[b];
        }
    }
    __syncthreads();
    const int64_t i0 = (int64_t)c * kChunk;
    const int64_t i1 = std::min<int64_t>(n, i0 + kChunk);
    for (int64_t s = i0; s < i1; s += 64) {
        const int64_t i = s + lane;
        const bool valid = i < i1;
        const int32_t e = valid ? (perm_in ? perm_in[i] : (int32_t)i) : 0;
        const uint32_t dg = valid ? (key[e] >> shift) & (kDigits - 1) : 0u;
        uint64_t left = __ballot(valid);
        while (left) {
            const int leader = __builtin_ctzll(left);
            const uint32_t ld = __shfl(dg, leader);
            const uint64_t same = __ballot(valid && dg == ld) & left;
            const uint32_t start = next[ld];
            if (valid && dg == ld) perm_out[start + lanes_below(same)] = e;
            __syncthreads();  // every lane has read next[ld]
            if (lane == leader) next[ld] = start + (uint32_t)__builtin_popcountll(same);
            __syncthreads();
            left &= ~same;
        }
    }
}

// Sort key of a row's distance to its centroid: ascending for L2,
// descending for IP; -0 and +0 equal.
__global__ void dist_key_kernel(const float* __restrict__ dist, int64_t n, int l2, uint32_t* __restrict__ key) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = f2ord(dist[i]);
    key[i] = l2 ? o : ~o;
}

// Sort key of a row's label: the label, or nlist for a row the assignment
// left without a centroid (non-finite values), which then sorts after every
// list; counts per label in cnt[0 .. nlist) and those rows in cnt[nlist].
__global__ void label_key_kernel(const int32_t* __restrict__ label, int64_t n, int nlist,
                                 uint32_t* __restrict__ key, uint32_t* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t l = label[i];
    const uint32_t kv = (l >= 0 && l < nlist) ? (uint32_t)l : (uint32_t)nlist;
    key[i] = kv;
    atomicAdd(&cnt[kv], 1u);
}

__global__ void zero_u32_kernel(uint32_t* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Exclusive scan of cnt[0..m) into off[0..m] (one workgroup).
__global__ __launch_bounds__(1024) void exscan_kernel(const uint32_t* __restrict__ cnt, int64_t m,
                                                      int64_t* __restrict__ off) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (m + 1023) / 1024;
    const int64_t a = t * per, b = std::min<int64_t>(m, a + per);
    int64_t s = 0;
    for (int64_t i = a; i < b; ++i) s += cnt[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int i = 0; i < 1024; ++i) {
            const int64_t v = part[i];
            part[i] = run;
            run += v;
        }
        off[m] = run;
    }
    __syncthreads();
    int64_t run = part[t];
    for (int64_t i = a; i < b; ++i) {
        off[i] = run;
        run += cnt[i];
    }
}

// One workgroup per centroid: FP64 sums of its rows (members in row order)
// per dimension, then the mean rounded to FP32.
__global__ __launch_bounds__(256) void centroid_mean_kernel(const float* __restrict__ x, int64_t ldx, int d,
                                                            const int32_t* __restrict__ members,
                                                            const int64_t* __restrict__ off,
                                                            float* __restrict__ cent) {
    const int l = blockIdx.x;
    const int64_t m0 = off[l], m1 = off[l + 1];
    if (m1 == m0) return;  // keeps its value
    for (int j = threadIdx.x; j < d; j += blockDim.x) {
        double acc = 0.0;
        for (int64_t m = m0; m < m1; ++m) acc += (double)x[(int64_t)members[m] * ldx + j];
        cent[(size_t)l * d + j] = (float)(acc / (double)(m1 - m0));
    }
}

// The tiled list arena from rows in list order: one thread per (tile, lane).
__global__ __launch_bounds__(256) void fill_arena_kernel(const float4* __restrict__ x, int nquad, int d,
                                                         const int32_t* __restrict__ order,
                                                         const int64_t* __restrict__ list_off,
                                                         const int32_t* __restrict__ tile_off,
                                                         const int32_t* __restrict__ tile_list, int64_t tiles,
                                                         float4* __restrict__ codes, uint32_t* __restrict__ ids,
                                                         float* __restrict__ norms, uint32_t* __restrict__ not_u8) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= tiles * kTileRows) return;
    const int64_t T = g / kTileRows;
    const int r = (int)(g % kTileRows);
    const int l = tile_list[T];
    const int64_t rr = (T - tile_off[l]) * kTileRows + r;
    const int64_t n_l = list_off[l + 1] - list_off[l];
    float4* dst = codes + (size_t)T * nquad * kTileRows + r;
    if (rr >= n_l) {
        for (int j = 0; j < nquad; ++j) dst[(size_t)j * kTileRows] = make_float4(0.f, 0.f, 0.f, 0.f);
        ids[g] = kNoId;
        norms[g] = 0.0f;
        return;
    }
    const int32_t row = order[list_off[l] + rr];
    const float4* src = x + (size_t)row * nquad;
    double s = 0.0;
    bool u8 = true;
    for (int j = 0; j < nquad; ++j) {
        const float4 v = src[j];
        dst[(size_t)j * kTileRows] = v;
        const float c[4] = {v.x, v.y, v.z, v.w};
        for (int i = 0; i < 4; ++i) {
            if (4 * j + i >= d) break;
            s += (double)c[i] * (double)c[i];
            u8 = u8 && c[i] >= 0.0f && c[i] <= 255.0f && c[i] == rintf(c[i]);
        }
    }
    ids[g] = (uint32_t)row;
    norms[g] = (float)s;
    if (!u8) atomicOr(not_u8, 1u);
}

// Row i divided by its norm (the FP32 FMA chain over the d dims, then a
// correctly rounded square root and quotient); a zero row stays zero.
__global__ void normalize_rows_kernel(const float* __restrict__ x, int64_t n, int d, int ld,
                                      float* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = x + i * ld;
    float s = 0.0f;
    for (int j = 0; j < d; ++j) s = __builtin_fmaf(r[j], r[j], s);
    const float nr = __builtin_sqrtf(s);
    float* o = y + i * ld;
    for (int j = 0; j < ld; ++j) o[j] = nr > 0.0f ? r[j] / nr : r[j];
}

unsigned grid_of(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Stable radix passes over 8-bit digits [shift0, shift0 + 8 * passes) of
// key[], starting from the order perm (nullptr = row order); the result
// is left in *out (one of a, b).
hipError_t radix_passes(const uint32_t* key, int64_t n, int shift0, int passes, const int32_t* perm, int32_t* a,
                        int32_t* b, uint32_t* counts, uint32_t* total, hipStream_t s, const int32_t** out) {
    const int nchunks = (int)((n + kChunk - 1) / kChunk);
    const int32_t* in = perm;
    int32_t* bufs[2] = {a, b};
    for (int p = 0; p < passes; ++p) {
        int32_t* o = bufs[p & 1];
        if (o == in) o = bufs[(p + 1) & 1];
        const int shift = shift0 + 8 * p;
        hipLaunchKernelGGL(rp_count_kernel, dim3(nchunks), dim3(64), 0, s, key, in, shift, n, nchunks, counts);
        hipLaunchKernelGGL(rp_scan_kernel, dim3(kDigits), dim3(256), 0, s, counts, nchunks, total);
        hipLaunchKernelGGL(rp_fill_kernel, dim3(nchunks), dim3(64), 0, s, key, in, shift, n, nchunks, counts, total,
                           o);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        in = o;
    }
    *out = in;
    return hipSuccess;
}

// 8-bit passes covering the label keys 0 .. nlist
int label_passes(int64_t nlist) { return nlist < 256 ? 1 : nlist < 65536 ? 2 : 3; }

}  // namespace

// Rows grouped by label (stable: row order inside a label) and the label
// offsets, for n labels already on the device.  `key` may hold a second
// key (sorted first, inside each label) or be null.
struct Grouping {
    DevBuf<int32_t> pa, pb;
    DevBuf<uint32_t> counts, total, cnt, lkey;
    DevBuf<int64_t> off;
    const int32_t* order = nullptr;
};

static int group_rows(hivf_index* h, const int32_t* label, const uint32_t* key2, int64_t n, Grouping& G,
                      hipStream_t s) {
    const int nchunks = (int)((n + kChunk - 1) / kChunk);
    HIP_TRY(h, G.pa.ensure((size_t)n));
    HIP_TRY(h, G.pb.ensure((size_t)n));
    HIP_TRY(h, G.counts.ensure((size_t)nchunks * kDigits));
    HIP_TRY(h, G.total.ensure(kDigits));
    HIP_TRY(h, G.cnt.ensure((size_t)h->nlist + 1));
    HIP_TRY(h, G.lkey.ensure((size_t)n));
    HIP_TRY(h, G.off.ensure((size_t)h->nlist + 1));
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_of(h->nlist + 1, 256)), dim3(256), 0, s, G.cnt.p, h->nlist + 1);
    hipLaunchKernelGGL(label_key_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, label, n, (int)h->nlist, G.lkey.p,
                       G.cnt.p);
    const int32_t* order = nullptr;
    if (key2) HIP_TRY(h, radix_passes(key2, n, 0, 4, nullptr, G.pa.p, G.pb.p, G.counts.p, G.total.p, s, &order));
    // labels are the final (most significant) key
    HIP_TRY(h, radix_passes(G.lkey.p, n, 0, label_passes(h->nlist), order, G.pa.p, G.pb.p, G.counts.p, G.total.p, s,
                            &order));
    hipLaunchKernelGGL(exscan_kernel, dim3(1), dim3(1024), 0, s, G.cnt.p, h->nlist, G.off.p);
    HIP_TRY(h, hipGetLastError());
    G.order = order;
    return HIVF_OK;
}

int gpu_lloyd_round(hivf_index* h, const float4* xa, const float* xsum, int64_t ldx, int64_t n, float* cent_host,
                    DevBuf<float>& dcent) {
    hipStream_t s = h->stream;
    // assignment: nearest centroid of every row (coarse kernel, one probe)
    HIP_TRY(h, h->d_cd.ensure((size_t)n));
    HIP_TRY(h, h->d_probes.ensure((size_t)n));
    CoarseArgs c;
    c.q = xa;
    c.nq = (int)n;
    c.nquad = h->nquad;
    c.cent = h->d_cent.p;
    c.nlist = (int)h->nlist;
    c.P = 1;
    c.out_d = h->d_cd.p;
    c.out_l = h->d_probes.p;
    HIP_TRY(h, launch_coarse(c, h->metric, s));
    Grouping G;
    int rc = group_rows(h, h->d_probes.p, nullptr, n, G, s);
    if (rc) return rc;
    const size_t cb = (size_t)h->nlist * h->d * sizeof(float);
    HIP_TRY(h, dcent.ensure((size_t)h->nlist * h->d));
    HIP_TRY(h, hipMemcpyAsync(dcent.p, cent_host, cb, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(centroid_mean_kernel, dim3((unsigned)h->nlist), dim3(256), 0, s, xsum, ldx, (int)h->d,
                       G.order, G.off.p, dcent.p);
    HIP_TRY(h, hipGetLastError());
    HIP_TRY(h, hipMemcpyAsync(cent_host, dcent.p, cb, hipMemcpyDeviceToHost, s));
    HIP_TRY(h, hipStreamSynchronize(s));
    return HIVF_OK;
}

namespace {

// Element e of the CSR order (row r of list l) -> its arena row, its list
// and the order key of its radius.
__global__ void shadow_elems_kernel(const int32_t* __restrict__ tile_list, const int32_t* __restrict__ tile_off,
                                    const int32_t* __restrict__ size, const int64_t* __restrict__ list_off,
                                    const float* __restrict__ radius, int64_t arena_rows,
                                    int32_t* __restrict__ elem_row, int32_t* __restrict__ elem_list,
                                    uint32_t* __restrict__ key) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= arena_rows) return;
    const int l = tile_list[g / kTileRows];
    const int64_t r = g - (int64_t)tile_off[l] * kTileRows;
    if (r >= size[l]) return;  // padding
    const int64_t e = list_off[l] + r;
    elem_row[e] = (int32_t)g;
    elem_list[e] = l;
    key[e] = f2ord(radius[g]);
}

// perm[arena position] = arena row shown there: every list's real rows by
// (radius, arena row), padding rows in place.
__global__ void shadow_perm_kernel(const int32_t* __restrict__ tile_list, const int32_t* __restrict__ tile_off,
                                   const int32_t* __restrict__ size, const int64_t* __restrict__ list_off,
                                   const int32_t* __restrict__ order, const int32_t* __restrict__ elem_row,
                                   int64_t arena_rows, int32_t* __restrict__ perm) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= arena_rows) return;
    const int l = tile_list[g / kTileRows];
    const int64_t r = g - (int64_t)tile_off[l] * kTileRows;
    perm[g] = r < size[l] ? elem_row[order[list_off[l] + r]] : (int32_t)g;
}

// Radius interval of every 32-row tile in shadow order, widened by 10^-4
// (empty tiles: an empty interval).
__global__ void shadow_bounds_kernel(const int32_t* __restrict__ tile_list, const int32_t* __restrict__ tile_off,
                                     const int32_t* __restrict__ size, const int32_t* __restrict__ perm,
                                     const float* __restrict__ radius, int64_t tiles32, float2* __restrict__ rb) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles32) return;
    const int l = tile_list[t / 2];
    const int64_t r0 = (t - 2 * (int64_t)tile_off[l]) * 32;
    float lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < 32; ++i) {
        if (r0 + i >= size[l]) break;
        const float x = radius[perm[t * 32 + i]];
        lo = fminf(lo, x * (1.0f - 1e-4f));
        hi = fmaxf(hi, x * (1.0f + 1e-4f) + 1e-30f);
    }
    rb[t] = make_float2(lo, hi);
}

}  // namespace

int shadow_order(hivf_index* h, const float* radius, int32_t* perm, float2* rb8, hipStream_t s) {
    const int64_t nlist = h->nlist, tiles = h->tiles, arena_rows = tiles * kTileRows, n = h->ntotal;
    if (tiles == 0) return HIVF_OK;
    std::vector<int32_t> tile_list((size_t)tiles);
    for (int64_t l = 0; l < nlist; ++l)
        for (int32_t t = h->tile_off[l]; t < h->tile_off[l + 1]; ++t) tile_list[t] = (int32_t)l;
    DevBuf<int32_t> dtl, erow, elist;
    DevBuf<uint32_t> key;
    DevBuf<int64_t> loff;
    HIP_TRY(h, dtl.ensure((size_t)tiles));
    HIP_TRY(h, loff.ensure((size_t)nlist + 1));
    HIP_TRY(h, erow.ensure((size_t)std::max<int64_t>(n, 1)));
    HIP_TRY(h, elist.ensure((size_t)std::max<int64_t>(n, 1)));
    HIP_TRY(h, key.ensure((size_t)std::max<int64_t>(n, 1)));
    HIP_TRY(h, hipMemcpyAsync(dtl.p, tile_list.data(), tiles * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(h, hipMemcpyAsync(loff.p, h->list_off.data(), (nlist + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(shadow_elems_kernel, dim3(grid_of(arena_rows, 256)), dim3(256), 0, s, dtl.p, h->d_tile_off.p,
                       h->d_size.p, loff.p, radius, arena_rows, erow.p, elist.p, key.p);
    HIP_TRY(h, hipGetLastError());
    Grouping G;
    if (n > 0) {
        int rc = group_rows(h, elist.p, key.p, n, G, s);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(shadow_perm_kernel, dim3(grid_of(arena_rows, 256)), dim3(256), 0, s, dtl.p, h->d_tile_off.p,
                       h->d_size.p, loff.p, n > 0 ? G.order : erow.p, erow.p, arena_rows, perm);
    hipLaunchKernelGGL(shadow_bounds_kernel, dim3(grid_of(2 * tiles, 256)), dim3(256), 0, s, dtl.p, h->d_tile_off.p,
                       h->d_size.p, perm, radius, 2 * tiles, rb8);
    HIP_TRY(h, hipGetLastError());
    HIP_TRY(h, hipStreamSynchronize(s));  // locals
    return HIVF_OK;
}

int normalize_rows(hivf_index* h, const float4* x, int64_t n, float4* y) {
    hipLaunchKernelGGL(normalize_rows_kernel, dim3(grid_of(n, 256)), dim3(256), 0, h->stream,
                       reinterpret_cast<const float*>(x), n, (int)h->d, (int)h->d_pad, reinterpret_cast<float*>(y));
    HIP_TRY(h, hipGetLastError());
    return HIVF_OK;
}

int gpu_add(hivf_index* h, int64_t n) {
    // rows are staged in h->d_q (padded float4 rows)
    hipStream_t s = h->stream;
    const int64_t nlist = h->nlist;
    HIP_TRY(h, h->d_cd.ensure((size_t)std::max<int64_t>(n, 1)));
    HIP_TRY(h, h->d_probes.ensure((size_t)std::max<int64_t>(n, 1)));
    Grouping G;
    DevBuf<uint32_t> key;
    if (n > 0) {
        CoarseArgs c;
        c.q = h->d_q.p;
        c.nq = (int)n;
        c.nquad = h->nquad;
        c.cent = h->d_cent.p;
        c.nlist = (int)nlist;
        c.P = 1;
        c.out_d = h->d_cd.p;
        c.out_l = h->d_probes.p;
        HIP_TRY(h, launch_coarse(c, h->metric, s));
        HIP_TRY(h, key.ensure((size_t)n));
        hipLaunchKernelGGL(dist_key_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, h->d_cd.p, n,
                           h->metric == HIVF_METRIC_L2 ? 1 : 0, key.p);
        int rc = group_rows(h, h->d_probes.p, key.p, n, G, s);
        if (rc) return rc;
    }
    std::vector<int64_t> list_off(nlist + 1, 0);
    if (n > 0) {
        uint32_t orphans = 0;
        HIP_TRY(h, hipMemcpyAsync(list_off.data(), G.off.p, (nlist + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(h, hipMemcpyAsync(&orphans, G.cnt.p + nlist, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(h, hipStreamSynchronize(s));
        if (orphans) return h->fail(HIVF_EINVAL, "rows without a nearest centroid (non-finite values)");
    }
    std::vector<int32_t> tile_off(nlist + 1, 0), size(nlist, 0);
    int64_t tiles = 0;
    for (int64_t l = 0; l < nlist; ++l) {
        const int64_t n_l = list_off[l + 1] - list_off[l];
        if (n_l > INT32_MAX) return h->fail(HIVF_EINVAL, "list too long");
        size[l] = (int32_t)n_l;
        tile_off[l] = (int32_t)tiles;
        tiles += (n_l + kTileRows - 1) / kTileRows;
        if (tiles > INT32_MAX / kTileRows) return h->fail(HIVF_EINVAL, "index too large");
    }
    tile_off[nlist] = (int32_t)tiles;
    std::vector<int32_t> tile_list((size_t)std::max<int64_t>(tiles, 1), 0);
    for (int64_t l = 0; l < nlist; ++l)
        for (int32_t t = tile_off[l]; t < tile_off[l + 1]; ++t) tile_list[t] = (int32_t)l;
    const int nquad = h->nquad;
    HIP_TRY(h, h->d_codes.ensure((size_t)std::max<int64_t>(tiles, 1) * nquad * kTileRows));
    HIP_TRY(h, h->d_ids.ensure((size_t)std::max<int64_t>(tiles, 1) * kTileRows));
    HIP_TRY(h, h->d_norms.ensure((size_t)std::max<int64_t>(tiles, 1) * kTileRows));
    HIP_TRY(h, h->d_tile_off.ensure(nlist + 1));
    HIP_TRY(h, h->d_size.ensure(nlist));
    HIP_TRY(h, hipMemcpyAsync(h->d_tile_off.p, tile_off.data(), (nlist + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                              s));
    HIP_TRY(h, hipMemcpyAsync(h->d_size.p, size.data(), nlist * sizeof(int32_t), hipMemcpyHostToDevice, s));
    uint32_t not_u8 = 0;
    if (tiles > 0) {
        DevBuf<int32_t> dtl;
        DevBuf<uint32_t> flag;
        HIP_TRY(h, dtl.ensure((size_t)tiles));
        HIP_TRY(h, flag.ensure(1));
        HIP_TRY(h, hipMemcpyAsync(dtl.p, tile_list.data(), tiles * sizeof(int32_t), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, flag.p, 1);
        hipLaunchKernelGGL(fill_arena_kernel, dim3(grid_of(tiles * kTileRows, 256)), dim3(256), 0, s, h->d_q.p,
                           nquad, (int)h->d, G.order, G.off.p, h->d_tile_off.p, dtl.p, tiles, h->d_codes.p,
                           h->d_ids.p, h->d_norms.p, flag.p);
        HIP_TRY(h, hipGetLastError());
        HIP_TRY(h, hipMemcpyAsync(&not_u8, flag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(h, hipStreamSynchronize(s));  // host vectors and local buffers
    h->lists_uint8 = not_u8 == 0;
    h->list_off = list_off;
    h->tile_off = tile_off;
    h->ntotal = list_off[nlist];
    h->tiles = tiles;
    h->have_lists = true;
    h->batched.invalidate();
    h->drop_graph();
    return HIVF_OK;
}

}  // namespace hivf
