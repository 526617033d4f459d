// Coarse quantizer on MFMA (src/Index.cpp:967-995): each query's nprobe
// nearest centroids, sorted, with the exact FP32 distances of the
// canonical chain -- the same output as the per-query coarse kernel, bit for
// bit, at a fraction of its cost.
//
//   coarse_dot_kernel : approximate ranks of every (query, centroid) pair
//       from v_mfma_f32_32x32x16_bf16.  Centroids enter as three BF16 pieces
//       (c = c1 + c2 + c3 exactly), queries as two (q ~ qh + ql, residual
//       below 2^-16 |q|), so the contraction loses almost nothing; the rank
//       is ||q||^2 + ||c||^2 - 2 q.c (L2) or -q.c (IP).
//   coarse_pick_kernel: per query, E_c = coef (||q||^2 + ||c||^2) bounds
//       |approximate rank - exact chain rank| (coarse_coef below).  U, at
//       least the nprobe-th smallest finite upper bound rank + E, comes from
//       two linear histograms; every centroid whose lower bound rank - E is
//       <= U (or not finite) is rescored with the exact FP32 chain and the
//       exact top-nprobe is taken by (distance, id).  The nprobe-th smallest
//       finite upper bound is >= D_P, the nprobe-th smallest exact rank, so
//       every centroid of the exact top-nprobe is a candidate, and a centroid
//       that is not has exact rank > U >= D_P.  The selection is therefore
//       exact, ties included.
#include <climits>
#include <cmath>

#include "bf16.hpp"
#include "index_impl.hpp"
#include "prep.hpp"
#include "select.hpp"
#include "wave.hpp"

namespace hivf {
namespace {

constexpr int kNT = 256;
constexpr int kCandChunk = 2048;   // centroids examined per candidate round
constexpr int kTilesPerTask = 2;   // 32-centroid tiles per wave task
constexpr int64_t kMaxRankFloats = 64ll << 20;  // approximate-rank buffer
// absolute part of the error bound: BF16 / FP32 values below the normal
// range may be flushed to zero by the matrix cores
constexpr float kEabs = 1e-30f;

// Centroid shadow: [32-row tile T][piece p < 3][16-dim chunk kc][lane][8 bf16],
// lane l holding centroid row l % 32, dims kc*16 + (l / 32) * 8 .. +7 (the
// MFMA A-operand order).  Piece p of a value v is split3(v)[p].
__global__ void cent_split_kernel(const float4* __restrict__ cent, int nquad, int64_t tiles32, int nkc,
                                  uint16_t* __restrict__ out) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= tiles32 * 3 * nkc * 64) return;
    const int lane = (int)(gid % 64);
    const int kc = (int)((gid / 64) % nkc);
    const int p = (int)((gid / (64 * (int64_t)nkc)) % 3);
    const int64_t T = gid / (64 * (int64_t)nkc * 3);
    const int64_t t64 = T / 2;
    const int r64 = (int)(T % 2) * 32 + lane % 32;
    const int c0 = kc * 16 + (lane / 32) * 8;
    uint16_t v[8];
    for (int h = 0; h < 2; ++h) {
        const int qd = c0 / 4 + h;
        const float4 f = qd < nquad ? cent[(t64 * nquad + qd) * 64 + r64] : make_float4(0.f, 0.f, 0.f, 0.f);
        const float e[4] = {f.x, f.y, f.z, f.w};
        for (int i = 0; i < 4; ++i) {
            uint16_t s3[3];
            split3(e[i], s3);
            v[h * 4 + i] = s3[p];
        }
    }
    uint4 pk;
    pk.x = v[0] | ((uint32_t)v[1] << 16);
    pk.y = v[2] | ((uint32_t)v[3] << 16);
    pk.z = v[4] | ((uint32_t)v[5] << 16);
    pk.w = v[6] | ((uint32_t)v[7] << 16);
    reinterpret_cast<uint4*>(out)[gid] = pk;
}

// Per query (one wave): qh = bf16(q), ql = bf16(q - qh), row-major
// [n x dpad16], and the squared norm accumulated in FP64.
__global__ void coarse_prep_kernel(const float* __restrict__ q, int d, int ld, int dpad16, uint16_t* __restrict__ qh,
                                   uint16_t* __restrict__ ql, float* __restrict__ qn, PrepArgs pa, int with_prep) {
    const int64_t qi = blockIdx.x;
    const int lane = threadIdx.x;
    float qv[kQRegs];  // (one round of loads for both preparations)
    query_regs(q, qi, ld, d, lane, qv);
    if (with_prep) prep_query(pa, (int)qi, lane, qv);
    double s = 0.0;
    auto one = [&](int c, float v) {
        const uint16_t h = f2bf(v);
        qh[qi * dpad16 + c] = h;
        ql[qi * dpad16 + c] = f2bf(v - bf2f(h));
        s += (double)v * (double)v;
    };
    if (dpad16 <= 64 * kQRegs) {
#pragma unroll
        for (int j = 0; j < kQRegs; ++j)
            if (j * 64 + lane < dpad16) one(j * 64 + lane, qv[j]);
    } else {
        for (int c = lane; c < dpad16; c += 64) one(c, c < d ? q[qi * ld + c] : 0.0f);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) qn[qi] = (float)s;
}

struct CoarseDotArgs {
    const uint16_t* cbf;  // centroid shadow
    const uint16_t* qh;   // [nq x dpad16]
    const uint16_t* ql;
    const float* qn;      // [nq]
    const float* cn;      // [tiles32 * 32] centroid squared norms (0 on padding)
    float* out;           // [nq x ldo] approximate ranks
    int64_t ldo;          // tiles32 * 32
    int nq;
    int nkc;
    int dpad16;
    int tiles32;
    int groups;           // tasks per 32-query tile
    uint32_t* clear;      // set to zero by the first thread (the pick's hand-over count)
};

// One wave task: 32 queries x kTilesPerTask centroid tiles.
// NKC > 0: the chunk count at compile time (the query fragments stay in
// registers for the task and the chunk loop is unrolled, so a tile's
// fragment loads are all in flight together); NKC = 0: any chunk count.
template <int METRIC, int NKC>
__global__ __launch_bounds__(kNT) void coarse_dot_kernel(CoarseDotArgs a) {
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const int slot = lane & 31;
    const int64_t task = (int64_t)blockIdx.x * (kNT / 64) + threadIdx.x / 64;
    const int64_t qtiles = (a.nq + 31) / 32;
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.clear = 0u;
    if (task >= qtiles * a.groups) return;
    const int qt = (int)(task / a.groups), g = (int)(task % a.groups);
    const int q = qt * 32 + slot;
    const bool live = q < a.nq;
    const bf16x8* bh = reinterpret_cast<const bf16x8*>(a.qh + (size_t)(live ? q : 0) * a.dpad16 + half * 8);
    const bf16x8* bl = reinterpret_cast<const bf16x8*>(a.ql + (size_t)(live ? q : 0) * a.dpad16 + half * 8);
    const float qn = live ? a.qn[q] : 0.0f;
    const int T0 = g * kTilesPerTask, T1 = min(a.tiles32, T0 + kTilesPerTask);
    const bf16x8 z = {};
    const int nkc = NKC > 0 ? NKC : a.nkc;
    bf16x8 hr[NKC > 0 ? NKC : 1], lr[NKC > 0 ? NKC : 1];
    if constexpr (NKC > 0) {
#pragma unroll
        for (int kc = 0; kc < NKC; ++kc) {
            hr[kc] = live ? bh[kc * 2] : z;
            lr[kc] = live ? bl[kc * 2] : z;
        }
    }
    if constexpr (NKC > 0) {
        // A tile's 3 NKC A fragments are loaded before its first MFMA (the
        // chain then waits for memory once), and the next tile's as soon as
        // this tile's MFMAs are issued, so they are in flight during its
        // epilogue and stores; the 32 centroid norms are loaded while the
        // chain runs.
        bf16x8 av[3][NKC];
        auto load = [&](int T) {
            const bf16x8* A = reinterpret_cast<const bf16x8*>(a.cbf) + (size_t)T * 3 * NKC * 64 + lane;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int kc = 0; kc < NKC; ++kc) av[p][kc] = A[(size_t)(p * NKC + kc) * 64];
        };
        if (T0 < T1) load(T0);
        for (int T = T0; T < T1; ++T) {
            float4 cv[4];
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4)
                cv[r4] = *reinterpret_cast<const float4*>(a.cn + (size_t)T * 32 + r4 * 8 + half * 4);
            __builtin_amdgcn_sched_barrier(0);
            floatx16 acc = {};
#pragma unroll
            for (int kc = 0; kc < NKC; ++kc) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[2][kc], lr[kc], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[1][kc], lr[kc], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[0][kc], lr[kc], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[2][kc], hr[kc], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[1][kc], hr[kc], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[0][kc], hr[kc], acc, 0, 0, 0);
            }
            __builtin_amdgcn_sched_barrier(0);
            if (T + 1 < T1) load(T + 1);
            __builtin_amdgcn_sched_barrier(0);
            if (!live) continue;
            // acc[r]: centroid row (r / 4) * 8 + half * 4 + r % 4 of tile T
            float* o = a.out + (size_t)q * a.ldo + (size_t)T * 32;
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int row = r4 * 8 + half * 4;
                const float cn[4] = {cv[r4].x, cv[r4].y, cv[r4].z, cv[r4].w};
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float dot = acc[r4 * 4 + e];
                    v[e] = METRIC == kL2 ? (qn + cn[e]) - 2.0f * dot : -dot;
                }
                *reinterpret_cast<float4*>(o + row) = make_float4(v[0], v[1], v[2], v[3]);
            }
        }
        return;
    }
    for (int T = T0; T < T1; ++T) {
        const bf16x8* A = reinterpret_cast<const bf16x8*>(a.cbf) + (size_t)T * 3 * nkc * 64 + lane;
        floatx16 acc = {};
        auto chunk = [&](int kc, const bf16x8& h, const bf16x8& l) {
            const bf16x8 a0 = A[(size_t)kc * 64];
            const bf16x8 a1 = A[(size_t)(nkc + kc) * 64];
            const bf16x8 a2 = A[(size_t)(2 * nkc + kc) * 64];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, l, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, l, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, l, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, h, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, h, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, h, acc, 0, 0, 0);
        };
        if constexpr (NKC > 0) {
#pragma unroll
            for (int kc = 0; kc < NKC; ++kc) chunk(kc, hr[kc], lr[kc]);
        } else {
#pragma unroll 2
            for (int kc = 0; kc < nkc; ++kc) chunk(kc, live ? bh[kc * 2] : z, live ? bl[kc * 2] : z);
        }
        if (!live) continue;
        // acc[r]: centroid row (r / 4) * 8 + half * 4 + r % 4 of tile T
        float* o = a.out + (size_t)q * a.ldo + (size_t)T * 32;
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
            const int row = r4 * 8 + half * 4;
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float dot = acc[r4 * 4 + e];
                v[e] = METRIC == kL2 ? (qn + a.cn[(size_t)T * 32 + row + e]) - 2.0f * dot : -dot;
            }
            *reinterpret_cast<float4*>(o + row) = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
}

struct CoarsePickArgs {
    const float* rk;      // [nq x ldo] approximate ranks
    int64_t ldo;
    const float* qn;
    const float* cn;
    float coef;
    const float4* q;      // staged FP32 queries [nq x nquad]
    int nquad;
    const float4* cent;   // FP32 centroid tiles
    const float4* crow;   // FP32 centroids, row-major [nlist x nquad]
    int nlist;
    int P;
    float* out_d;         // [nq x P]
    int32_t* out_l;
    int32_t* hard;        // queries the one-wave form hands to the workgroup form
    uint32_t* nhard;
};

// Workgroup form of the pick for one query (all kNT threads).
template <int METRIC>
__device__ void pick_one(const CoarsePickArgs& a, int64_t qi) {
    constexpr int kW = kNT / 64;
    __shared__ LdsTopK<kNT, 2048> top;
    __shared__ uint32_t hist[kW][256];  // one histogram per wave (less LDS atomic contention)
    __shared__ float s_red[2][kW];
    __shared__ int s_bin, s_sub;
    __shared__ uint32_t s_need, s_fin;
    __shared__ int cand[kCandChunk];
    __shared__ int ncand;
    extern __shared__ float4 qs[];
    const int tid = threadIdx.x;
    const int wv = tid / 64, lane = tid % 64;
    for (int j = tid; j < a.nquad; j += kNT) qs[j] = a.q[qi * a.nquad + j];
    const float* rk = a.rk + qi * a.ldo;
    const float qn = a.qn[qi];
    auto ub_of = [&](int c) { return rk[c] + (a.coef * (qn + a.cn[c]) + kEabs); };
    // U: an upper bound of the P-th smallest upper bound, as the largest
    // element of a set of >= P bounds found by two linear histograms (the
    // first over [min, max] of the finite bounds, the second inside the bin
    // holding the P-th).  Binning is monotone, so the set is a prefix of the
    // sorted bounds.
    float mn = INFINITY, mx = -INFINITY;
    uint32_t fin = 0;
    for (int c = tid; c < a.nlist; c += kNT) {
        const float u = ub_of(c);
        if (u <= FLT_MAX && u >= -FLT_MAX) {
            mn = fminf(mn, u);
            mx = fmaxf(mx, u);
            ++fin;
        }
    }
    mn = wave_min(mn);
    mx = wave_max(mx);
    for (int off = 32; off > 0; off >>= 1) fin += __shfl_xor(fin, off);
    for (int i = lane; i < 256; i += 64) hist[wv][i] = 0;
    if (lane == 0) {
        s_red[0][wv] = mn;
        s_red[1][wv] = mx;
    }
    if (tid == 0) s_fin = 0;
    __syncthreads();
    if (lane == 0) atomicAdd(&s_fin, fin);
    mn = s_red[0][0];
    mx = s_red[1][0];
    for (int w = 1; w < kW; ++w) {
        mn = fminf(mn, s_red[0][w]);
        mx = fmaxf(mx, s_red[1][w]);
    }
    __syncthreads();
    float U = INFINITY;  // fewer than P finite bounds: every centroid is a candidate
    if (s_fin >= (uint32_t)a.P) {
        const float w1 = (mx - mn) * (1.0f / 256.0f);
        const float inv1 = w1 > 0.0f ? 1.0f / w1 : 0.0f;
        for (int c = tid; c < a.nlist; c += kNT) {
            const float u = ub_of(c);
            if (u <= FLT_MAX && u >= -FLT_MAX) atomicAdd(&hist[wv][lin_bin(u, mn, inv1)], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t need = (uint32_t)a.P;
            int b = 0;
            for (; b < 255; ++b) {
                const uint32_t h = hist[0][b] + hist[1][b] + hist[2][b] + hist[3][b];
                if (h >= need) break;
                need -= h;
            }
            s_bin = b;
            s_need = need;
        }
        __syncthreads();
        const int b1 = s_bin;
        for (int i = lane; i < 256; i += 64) hist[wv][i] = 0;
        __syncthreads();
        const float lo2 = mn + (float)b1 * w1;
        const float inv2 = w1 > 0.0f ? 256.0f / w1 : 0.0f;
        for (int c = tid; c < a.nlist; c += kNT) {
            const float u = ub_of(c);
            if (u <= FLT_MAX && u >= -FLT_MAX && lin_bin(u, mn, inv1) == b1)
                atomicAdd(&hist[wv][lin_bin(u, lo2, inv2)], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t need = s_need;
            int sb = 0;
            for (; sb < 255; ++sb) {
                const uint32_t h = hist[0][sb] + hist[1][sb] + hist[2][sb] + hist[3][sb];
                if (h >= need) break;
                need -= h;
            }
            s_sub = sb;
        }
        __syncthreads();
        const int sb = s_sub;
        float m = -INFINITY;
        for (int c = tid; c < a.nlist; c += kNT) {
            const float u = ub_of(c);
            if (u <= FLT_MAX && u >= -FLT_MAX) {
                const int bb = lin_bin(u, mn, inv1);
                if (bb < b1 || (bb == b1 && lin_bin(u, lo2, inv2) <= sb)) m = fmaxf(m, u);
            }
        }
        m = wave_max(m);
        if (lane == 0) s_red[0][wv] = m;
        __syncthreads();
        U = s_red[0][0];
        for (int w = 1; w < kW; ++w) U = fmaxf(U, s_red[0][w]);
    }
    // candidates, rescored with the exact chain, into the exact top-P
    top.init();
    uint64_t thr = top.thr;
    for (int base = 0; base < a.nlist; base += kCandChunk) {
        if (tid == 0) ncand = 0;
        __syncthreads();
        const int end = min(a.nlist, base + kCandChunk);
        for (int c = base + tid; c < end; c += kNT) {
            const float lb = rk[c] - (a.coef * (qn + a.cn[c]) + kEabs);
            // a bound that is not finite (overflow in the approximation)
            // says nothing: rescore that centroid too
            if (lb <= U || !(lb >= -FLT_MAX && lb <= FLT_MAX)) cand[atomicAdd(&ncand, 1)] = c;
        }
        __syncthreads();
        const int m = ncand;
        for (int i0 = 0; i0 < m; i0 += kNT) {
            if (i0 + tid < m) {
                const int c = cand[i0 + tid];
                const float4* cp = a.cent + (size_t)(c / kTileRows) * a.nquad * kWave + (c % kTileRows);
                float acc = 0.0f;
#pragma unroll 8
                for (int j = 0; j < a.nquad; ++j) acc = accum<METRIC>(acc, qs[j], cp[(size_t)j * kWave]);
                top.offer(make_key(rank_of<METRIC>(acc), (uint32_t)c), thr);
            }
            __syncthreads();
            top.maybe_compact(a.P, kNT);
            thr = top.thr;
        }
        __syncthreads();
    }
    top.finish(a.P);
    const int cnt = top.cnt;
    for (int i = tid; i < a.P; i += kNT) {
        float dv = sentinel<METRIC>();
        int32_t lv = -1;
        if (i < cnt) {
            const uint64_t key = top.buf[i];
            const float rk2 = key_rank(key);
            dv = dist_of_rank(METRIC, rk2);
            lv = (int32_t)(uint32_t)key;
        }
        a.out_d[qi * a.P + i] = dv;
        a.out_l[qi * a.P + i] = lv;
    }
}

// The queries the one-wave form handed over (a.hard[0 .. *a.nhard)), one
// per workgroup at a time.
template <int METRIC>
__global__ __launch_bounds__(kNT) void coarse_pick_kernel(CoarsePickArgs a) {
    const uint32_t nh = *a.nhard;
    for (uint32_t i = blockIdx.x; i < nh; i += gridDim.x) {
        __syncthreads();  // the previous query's LDS reads are done
        pick_one<METRIC>(a, a.hard[i]);
    }
}

constexpr int kWaveCand = 512;  // candidates one wave sorts in LDS
constexpr int kRankCache = 1024; // approximate ranks one wave keeps in LDS
constexpr int kRegSlots = 16;    // centroids per lane whose bounds stay in registers
constexpr int kBinSlack = 8;     // the P-th's bin is split when the bins up to it hold more than P + this


// One wave per query: the same bound U as the workgroup form (two linear
// histograms over the finite upper bounds, then the largest bound of the
// prefix set), the candidates compacted with ballots, rescored with the
// exact chain, sorted in LDS by (distance, id).  A query with more than
// kWaveCand / 2 candidates, or nprobe above that, goes to the workgroup form.
template <int METRIC, bool REG>
__global__ __launch_bounds__(kNT) void coarse_pick_wave_kernel(CoarsePickArgs a, int nq) {
    constexpr int kW = kNT / 64;
    __shared__ uint32_t hist[kW][256];
    __shared__ uint64_t keys[kW][kWaveCand];
    __shared__ __align__(16) float rkc[kW][kRankCache];  // the query's approximate ranks (without REG, when
                                                        // they fit); then the rescoring's staging area
    extern __shared__ float4 qsh[];  // kW x nquad
    const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int64_t qi = (int64_t)blockIdx.x * kW + wv;
    if (qi >= nq) return;  // whole wave; no workgroup barrier below
    float4* qs = qsh + wv * a.nquad;
    for (int j = lane; j < a.nquad; j += 64) qs[j] = a.q[qi * a.nquad + j];
    const float* rk = a.rk + qi * a.ldo;
    const float qn = a.qn[qi];
    // REG (nlist <= 64 kRegSlots): each lane keeps its centroids' bounds in
    // registers.  Otherwise the ranks are read from HBM once and kept in LDS
    // when they fit, and the bounds recomputed from them with the same
    // arithmetic (the norms stay in L1).
    const bool cached = !REG && a.nlist <= kRankCache;
    float* rc = rkc[wv];
    auto err_of = [&](int c) { return a.coef * (qn + a.cn[c]) + kEabs; };
    auto finite = [](float u) { return u <= FLT_MAX && u >= -FLT_MAX; };
    auto hand_over = [&]() {
        if (lane == 0) a.hard[atomicAdd(a.nhard, 1u)] = (int32_t)qi;
    };
    if (a.P > kWaveCand / 2) {
        hand_over();
        return;
    }
    float ubr[REG ? kRegSlots : 1], lbr[REG ? kRegSlots : 1];
    float mn = INFINITY, mx = -INFINITY;
    uint32_t fin = 0;
    if constexpr (REG) {
        // every slot's rank and centroid norm loaded before any is used (a
        // slot past nlist reads centroid 0's and is then ignored), so the
        // wave waits for memory once, not once per slot
        float rr[kRegSlots], cn[kRegSlots];
#pragma unroll
        for (int t = 0; t < kRegSlots; ++t) {
            const int c = t * 64 + lane < a.nlist ? t * 64 + lane : 0;
            rr[t] = rk[c];
            cn[t] = a.cn[c];
        }
#pragma unroll
        for (int t = 0; t < kRegSlots; ++t) {
            const int c = t * 64 + lane;
            ubr[t] = INFINITY;
            lbr[t] = INFINITY;
            if (c < a.nlist) {
                const float e = a.coef * (qn + cn[t]) + kEabs;  // err_of(c)
                ubr[t] = rr[t] + e;
                lbr[t] = rr[t] - e;
                if (finite(ubr[t])) {
                    mn = fminf(mn, ubr[t]);
                    mx = fmaxf(mx, ubr[t]);
                    ++fin;
                }
            }
        }
    } else {
#pragma unroll 4
        for (int c = lane; c < a.nlist; c += 64) {
            const float r = rk[c];
            if (cached) rc[c] = r;
            const float u = r + err_of(c);
            if (finite(u)) {
                mn = fminf(mn, u);
                mx = fmaxf(mx, u);
                ++fin;
            }
        }
    }
    // f(c, upper bound, lower bound) over this lane's centroids
    auto each = [&](auto&& f) {
        if constexpr (REG) {
#pragma unroll
            for (int t = 0; t < kRegSlots; ++t)
                if (t * 64 + lane < a.nlist) f(t * 64 + lane, ubr[t], lbr[t]);
        } else {
            for (int c = lane; c < a.nlist; c += 64) {
                const float r = cached ? rc[c] : rk[c], e = err_of(c);
                f(c, r + e, r - e);
            }
        }
    };
    mn = wave_min(mn);
    mx = wave_max(mx);
    for (int off = 32; off > 0; off >>= 1) fin += __shfl_xor(fin, off);
    float U = INFINITY;
    if (REG && fin >= (uint32_t)a.P) {
        // One 1024-bin histogram of the finite upper bounds, in the rank
        // cache (unused with REG until the rescoring), and U the upper edge
        // of the P-th's bin (widened as below).  The P-th's bin holds few
        // bounds at this resolution, so one pass does what two 256-bin
        // passes do otherwise.
        uint32_t* hs = reinterpret_cast<uint32_t*>(rc);
        wave_clear1024(hs, lane);
        __builtin_amdgcn_wave_barrier();
        const float w = (mx - mn) * (1.0f / 1024.0f);
        const float inv = w > 0.0f ? 1.0f / w : 0.0f;
        each([&](int, float u, float) {
            if (finite(u)) atomicAdd(&hs[lin_bin1024(u, mn, inv)], 1u);
        });
        __builtin_amdgcn_wave_barrier();
        const int b = wave_find_bin1024(hs, (uint32_t)a.P, lane);
        float edge = mx;
        if (inv > 0.0f && b < 1023) edge = mn + (float)(b + 1) * w;
        U = fminf(mx, edge + 0x1p-16f * (fabsf(mn) + fabsf(mx) + w));
        __builtin_amdgcn_wave_barrier();  // the histogram's reads are done before the staging reuses it
    } else if (fin >= (uint32_t)a.P) {
        uint32_t* hs = hist[wv];
        for (int i = lane; i < 256; i += 64) hs[i] = 0;
        __builtin_amdgcn_wave_barrier();
        const float w1 = (mx - mn) * (1.0f / 256.0f);
        const float inv1 = w1 > 0.0f ? 1.0f / w1 : 0.0f;
        each([&](int, float u, float) {
            if (finite(u)) atomicAdd(&hs[lin_bin(u, mn, inv1)], 1u);
        });
        __builtin_amdgcn_wave_barrier();
        uint32_t need = (uint32_t)a.P;
        const int b1 = wave_find_bin(hs, &need, lane);
        // bounds in the bins up to the P-th's: when they are few more than P
        // (the usual case), the P-th's bin is not split (sb = 255 keeps all
        // of it) -- a few more candidates, one pass fewer
        const uint32_t upto = (uint32_t)a.P - need + hs[b1];
        int sb = 255;
        float lo2 = 0.0f, inv2 = 0.0f;
        if (upto > (uint32_t)a.P + kBinSlack) {
            __builtin_amdgcn_wave_barrier();
            for (int i = lane; i < 256; i += 64) hs[i] = 0;
            __builtin_amdgcn_wave_barrier();
            lo2 = mn + (float)b1 * w1;
            inv2 = w1 > 0.0f ? 256.0f / w1 : 0.0f;
            each([&](int, float u, float) {
                if (finite(u) && lin_bin(u, mn, inv1) == b1) atomicAdd(&hs[lin_bin(u, lo2, inv2)], 1u);
            });
            __builtin_amdgcn_wave_barrier();
            sb = wave_find_bin(hs, &need, lane);
        }
        // U: the upper edge of the last bin taken (of the second histogram
        // when it ran), widened far beyond the rounding of lin_bin, or mx
        // when that bin is the last.  At least P finite upper bounds lie in
        // the bins taken, so U >= the P-th smallest of them >= the P-th
        // smallest exact rank; the few bounds between the largest one taken
        // and the edge only add candidates.  (An exact maximum over the
        // bins taken costs one more pass over the bounds.)
        float edge = mx;
        if (inv1 > 0.0f && b1 < 255) edge = mn + (float)(b1 + 1) * w1;
        if (inv2 > 0.0f && sb < 255) edge = fminf(edge, lo2 + (float)(sb + 1) * (w1 * (1.0f / 256.0f)));
        U = fminf(mx, edge + 0x1p-16f * (fabsf(mn) + fabsf(mx) + w1));
    }
    // candidates (ballot compaction of their ids), then rescored in
    // parallel, one candidate per lane, into keys[]
    uint64_t* kk = keys[wv];
    uint32_t* cid = reinterpret_cast<uint32_t*>(kk) + kWaveCand;  // second half of keys[]
    uint32_t ncand = 0;
    bool over = false;
    // one group of 64 centroids (c0 = group start; lb = this lane's lower bound)
    auto group = [&](int c0, float lb) {
        const int c = c0 + lane;
        const bool take = c < a.nlist && (lb <= U || !(lb >= -FLT_MAX && lb <= FLT_MAX));
        const uint64_t mk = __ballot(take);
        const uint32_t pos = ncand + lanes_below(mk);
        ncand += (uint32_t)__popcll(mk);
        over = ncand > (uint32_t)kWaveCand / 2;
        if (take && !over) cid[pos] = (uint32_t)c;
    };
    if constexpr (REG) {
#pragma unroll
        for (int t = 0; t < kRegSlots; ++t)
            if (t * 64 < a.nlist && !over) group(t * 64, lbr[t]);
    } else {
        for (int c0 = 0; c0 < a.nlist && !over; c0 += 64) {
            const int c = c0 + lane;
            group(c0, c < a.nlist ? (cached ? rc[c] : rk[c]) - err_of(c) : INFINITY);
        }
    }
    if (over) {
        hand_over();
        return;
    }
    __builtin_amdgcn_wave_barrier();
    // Rescoring, 64 candidates per round.  The candidates' rows are brought
    // in 4 quads (64 B of a row-major row) at a time, 16 rows per load
    // instruction, staged in LDS (the rank cache is free by now), and every
    // lane runs its candidate's chain from there.  kAhead chunks are in
    // flight (register buffers; the loads are issued unconditionally, the
    // last chunk again past the end, so every chunk's wait covers only its
    // own loads).  Element (i, jj) sits at i*4 + (jj ^ ((i >> 2) & 3)), so
    // both the staging writes and the chain's reads hit 16 distinct 16-byte
    // bank groups per 16 lanes.
    constexpr int kAhead = 4;
    float4* st = reinterpret_cast<float4*>(rc);
    const int nchunk = (a.nquad + 3) / 4;
    const int nround = (nchunk + kAhead - 1) / kAhead * kAhead;
    const int jj = lane & 3;
    for (uint32_t i0 = 0; i0 < ncand; i0 += 64) {
        const uint32_t nc = min(64u, ncand - i0);
        if (nc <= 16) {
            // a short round (typically the few candidates past the first
            // 64): each lane reads its candidate's row directly, every load
            // in flight at once, instead of the staged chunk chain
            if ((uint32_t)lane < nc) {
                const uint32_t c = cid[i0 + lane];
                const float4* rp = a.crow + (size_t)c * a.nquad;
                float acc = 0.0f;
#pragma unroll 16
                for (int j = 0; j < a.nquad; ++j) acc = accum<METRIC>(acc, qs[j], rp[j]);
                const float r = rank_of<METRIC>(acc);
                kk[i0 + lane] = r < FLT_MAX ? make_key(r, c) : ~0ull;
            }
            continue;
        }
        const float4* src[4];
        bool ok[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t i = (uint32_t)(lane >> 2) + 16u * t;
            ok[t] = i < nc;
            src[t] = a.crow + (size_t)(ok[t] ? cid[i0 + i] : 0u) * a.nquad + jj;
        }
        float4 v[kAhead][4];
        auto load = [&](int ch, float4 (&b)[4]) {
            ch = min(ch, nchunk - 1);
            const bool in = ch * 4 + jj < a.nquad;
#pragma unroll
            for (int t = 0; t < 4; ++t) b[t] = ok[t] && in ? src[t][ch * 4] : make_float4(0.f, 0.f, 0.f, 0.f);
        };
#pragma unroll
        for (int u = 0; u < kAhead; ++u) load(u, v[u]);
        float acc = 0.0f;
        for (int ch0 = 0; ch0 < nround; ch0 += kAhead) {
#pragma unroll
            for (int u = 0; u < kAhead; ++u) {
                const int ch = ch0 + u;
                __builtin_amdgcn_wave_barrier();  // the previous chunk's reads are done
                if (ch < nchunk) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int i = (lane >> 2) + 16 * t;
                        st[i * 4 + (jj ^ ((i >> 2) & 3))] = v[u][t];
                    }
                }
                __builtin_amdgcn_wave_barrier();
                load(ch + kAhead, v[u]);
                if (ch < nchunk) {
                    const int nj = min(4, a.nquad - ch * 4);
                    for (int j = 0; j < nj; ++j)
                        acc = accum<METRIC>(acc, qs[ch * 4 + j], st[lane * 4 + (j ^ ((lane >> 2) & 3))]);
                }
            }
        }
        if ((uint32_t)lane < nc) {
            const uint32_t c = cid[i0 + lane];
            // the heap's strict admission: inf and NaN never enter
            const float r = rank_of<METRIC>(acc);
            kk[i0 + lane] = r < FLT_MAX ? make_key(r, c) : ~0ull;
        }
    }
    __builtin_amdgcn_wave_barrier();
    // bitonic sort of the candidates (padded with ~0 to a power of two)
    uint32_t n2 = 64;
    while (n2 < ncand) n2 <<= 1;
    for (uint32_t i = ncand + lane; i < n2; i += 64) kk[i] = ~0ull;
    __builtin_amdgcn_wave_barrier();
    wave_sort(kk, n2, lane);
    for (int i = lane; i < a.P; i += 64) {
        const uint64_t key = kk[i];
        const bool ok = key < key_limit();
        a.out_d[qi * a.P + i] = ok ? dist_of_rank(METRIC, key_rank(key)) : sentinel<METRIC>();
        a.out_l[qi * a.P + i] = ok ? (int32_t)(uint32_t)key : -1;
    }
}

}  // namespace

// Bound on |approximate rank - exact chain rank| per unit of
// ||q||^2 + ||c||^2 (u = 2^-8, e = 2^-24, gamma_m = m e / (1 - m e)):
//   query residual |r| <= 2^-16 (1 + u) |q|:  2^-16 (1 + u) ||q|| ||c||
//   FP32 accumulation of 6 d exact products, whose magnitudes sum to at
//   most (1 + 2u)^2 ||q|| ||c||: gamma_{6d} (1 + 2u)^2 ||q|| ||c||, taken
//   twice to allow any rounding inside the matrix core's sums
//   FP32 norms and the two epilogue operations: 5 e (||q||^2 + ||c||^2)
//   the exact chain's own error:  gamma_{d+2} ||q - c||^2 <= 2 gamma_{d+2} (...)
// doubled for L2 (-2 q.c) and halved by ||q|| ||c|| <= (||q||^2 + ||c||^2) / 2;
// the IP rank (-q.c) is within the same bound.  Scaled by 1.01 for the
// rounding of rank +- E.
float coarse_coef(int64_t d) {
    const double u = 1.0 / 256.0, e = 1.0 / 16777216.0;
    auto gamma = [&](double m) { return m * e / (1.0 - m * e); };
    const double c = std::ldexp(1.0, -16) * (1 + u) + 2 * gamma(6.0 * d) * (1 + 2 * u) * (1 + 2 * u) + 5 * e +
                     2 * gamma(d + 2.0);
    return (float)(c * 1.01);
}

// nprobe nearest centroids of n staged queries into h->d_cd / h->d_probes.
// Returns HIVF_EUNSUPPORTED (nothing launched) where the MFMA form does not
// apply (d > 512), so the caller runs the per-query kernel.
int run_coarse_mfma(hivf_index* h, const float4* q, int64_t n, int64_t P, hipStream_t s, const PrepArgs* prep) {
    const int64_t dpad16 = (h->d + 15) / 16 * 16;
    const int nkc = (int)(dpad16 / 16);
    if (nkc > 32) return HIVF_EUNSUPPORTED;
    const int64_t tiles32 = 2 * ((h->nlist + kTileRows - 1) / kTileRows);
    const int64_t ldo = tiles32 * 32;
    CoarseState& C = h->coarse;
    if (!C.valid) {
        HIP_TRY(h, C.cbf.ensure((size_t)tiles32 * 3 * nkc * 64 * 8));
        const int64_t total = tiles32 * 3 * nkc * 64;
        hipLaunchKernelGGL(cent_split_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, h->d_cent.p,
                           h->nquad, tiles32, nkc, C.cbf.p);
        HIP_TRY(h, hipGetLastError());
        C.valid = true;
    }
    const int64_t chunk = std::max<int64_t>(32, std::min<int64_t>((n + 31) / 32 * 32, kMaxRankFloats / ldo / 32 * 32));
    HIP_TRY(h, C.qh.ensure((size_t)n * dpad16));
    HIP_TRY(h, C.ql.ensure((size_t)n * dpad16));
    HIP_TRY(h, C.qn.ensure((size_t)n));
    HIP_TRY(h, C.rank.ensure((size_t)chunk * ldo));
    HIP_TRY(h, C.hard.ensure((size_t)chunk));
    HIP_TRY(h, C.nhard.ensure(1));
    HIP_TRY(h, h->d_cd.ensure((size_t)n * P));
    HIP_TRY(h, h->d_probes.ensure((size_t)n * P));
    hipLaunchKernelGGL(coarse_prep_kernel, dim3((unsigned)n), dim3(64), 0, s, reinterpret_cast<const float*>(q),
                       (int)h->d, (int)h->d_pad, (int)dpad16, C.qh.p, C.ql.p, C.qn.p, prep ? *prep : PrepArgs{},
                       prep ? 1 : 0);
    const float coef = coarse_coef(h->d);
    const int groups = (int)((tiles32 + kTilesPerTask - 1) / kTilesPerTask);
    for (int64_t q0 = 0; q0 < n; q0 += chunk) {
        const int64_t m = std::min(chunk, n - q0);
        CoarseDotArgs da;
        da.cbf = C.cbf.p;
        da.qh = C.qh.p + q0 * dpad16;
        da.ql = C.ql.p + q0 * dpad16;
        da.qn = C.qn.p + q0;
        da.cn = h->d_cent_norm.p;
        da.out = C.rank.p;
        da.ldo = ldo;
        da.nq = (int)m;
        da.nkc = nkc;
        da.dpad16 = (int)dpad16;
        da.tiles32 = (int)tiles32;
        da.groups = groups;
        da.clear = C.nhard.p;
        const int64_t tasks = (m + 31) / 32 * groups;
        const dim3 grid((unsigned)((tasks + 3) / 4));
        CoarsePickArgs pa;
        pa.rk = C.rank.p;
        pa.ldo = ldo;
        pa.qn = C.qn.p + q0;
        pa.cn = h->d_cent_norm.p;
        pa.coef = coef;
        pa.q = q + q0 * h->nquad;
        pa.nquad = h->nquad;
        pa.cent = h->d_cent.p;
        pa.crow = h->d_cent_rows.p;
        pa.nlist = (int)h->nlist;
        pa.P = (int)P;
        pa.out_d = h->d_cd.p + q0 * P;
        pa.out_l = h->d_probes.p + q0 * P;
        pa.hard = C.hard.p;
        pa.nhard = C.nhard.p;
        const size_t lds = (size_t)h->nquad * sizeof(float4);
        const dim3 wgrid((unsigned)((m + 3) / 4));
        const dim3 hgrid((unsigned)std::min<int64_t>(m, 1024));
        if (h->metric == HIVF_METRIC_L2) {
            if (nkc == 8) hipLaunchKernelGGL((coarse_dot_kernel<kL2, 8>), grid, dim3(kNT), 0, s, da);
            else hipLaunchKernelGGL((coarse_dot_kernel<kL2, 0>), grid, dim3(kNT), 0, s, da);
            if (h->nlist <= 64 * kRegSlots)
                hipLaunchKernelGGL((coarse_pick_wave_kernel<kL2, true>), wgrid, dim3(kNT), 4 * lds, s, pa, (int)m);
            else
                hipLaunchKernelGGL((coarse_pick_wave_kernel<kL2, false>), wgrid, dim3(kNT), 4 * lds, s, pa, (int)m);
            hipLaunchKernelGGL((coarse_pick_kernel<kL2>), hgrid, dim3(kNT), lds, s, pa);
        } else {
            if (nkc == 8) hipLaunchKernelGGL((coarse_dot_kernel<kIP, 8>), grid, dim3(kNT), 0, s, da);
            else hipLaunchKernelGGL((coarse_dot_kernel<kIP, 0>), grid, dim3(kNT), 0, s, da);
            if (h->nlist <= 64 * kRegSlots)
                hipLaunchKernelGGL((coarse_pick_wave_kernel<kIP, true>), wgrid, dim3(kNT), 4 * lds, s, pa, (int)m);
            else
                hipLaunchKernelGGL((coarse_pick_wave_kernel<kIP, false>), wgrid, dim3(kNT), 4 * lds, s, pa, (int)m);
            hipLaunchKernelGGL((coarse_pick_kernel<kIP>), hgrid, dim3(kNT), lds, s, pa);
        }
        HIP_TRY(h, hipGetLastError());
    }
    return HIVF_OK;
}

}  // namespace hivf
