// oracle/_ref driver -- TEST INFRASTRUCTURE ONLY (checker and CPU baseline).
//
// Compiles the reference's own top-k heap (src/heap.hpp, with src/common.h)
// where it lies in the reference tree, under the reference's release flags
// (-Ofast, CMakeLists.txt:68), and drives it with the reference's plain
// search loop.  The rest of the reference's search path cannot be compiled
// here (src/utils.h needs <format>, mkl.h and Eigen), so the two pieces of
// it this driver needs are restated below:
//   - the distance loop of calculatedEuclideanDistance /
//     calculatedInnerProduct (src/utils.h:480-487, 512-519), same loop form,
//     left to the compiler's -Ofast vectorisation as in the reference;
//   - the query loop of Index::single_thread_search (src/Index.cpp:1420-1554)
//     with lite_scan_codes / scan_codes admission (src/IVFScan.hpp:108-172,
//     247-272) and the OpenMP query split of Index::search
//     (src/Index.cpp:1667-1699).
// Ties at equal distance are therefore resolved exactly as the reference's
// heap resolves them for this scan order.
//
// Two more CPU baselines for the same index, restated here under the same
// flags (they time the reference's other CPU algorithms; results are the
// same top-k distances):
//   - ref_search_base1: "Harmony base with one worker": the warm-up bound of
//     Index::warmUpSearch (src/Index.cpp:999-1051: the heap top after the
//     first min(W, n_l) rows of the first L probed lists), then every probed
//     list scanned with the four dimension checks of
//     BaseWorker::single_thread_search_simple (src/node.cpp:606-678: d / 4
//     dims per check, the last check taking the rest, a row dropped once its
//     running distance exceeds the bound), survivors admitted to the heap.
//   - ref_search_faiss: the algorithm of faiss IndexIVFFlat::search with
//     parallel_mode 0 (third_party/faiss/faiss/IndexIVF.cpp:294-388): query
//     slices per thread, the coarse quantizer in GEMM form ||q||^2 + ||c||^2
//     - 2 q.c (third_party/faiss/faiss/utils/distances.cpp:263-344) with a
//     top-nprobe heap, then IVFFlatScanner's scan (IndexIVFFlat.cpp:155-179):
//     strict admission into a max-heap ordered by (distance, id)
//     (third_party/faiss/faiss/utils/Heap.h:111-149).
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <omp.h>

#include "heap.hpp"

using harmony::idx_t;
using harmony::MetricType;

namespace {

inline float l2(const float* a, const float* b, size_t n) {
    float s = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float t = a[i] - b[i];
        s += t * t;
    }
    return s;
}

inline float ip(const float* a, const float* b, size_t n) {
    float s = 0.0f;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

template <MetricType M>
void scan_rows(size_t k, const float* q, const float* codes, const int64_t* ids,
               size_t n, size_t d, float* simi, idx_t* idxi) {
    for (size_t r = 0; r < n; ++r) {
        const float* x = codes + r * d;
        if constexpr (M == MetricType::METRIC_L2) {
            float dis = l2(q, x, d);
            if (dis < simi[0]) harmony::heap_replace_top<M>(k, simi, idxi, dis, ids[r]);
        } else {
            float dis = ip(q, x, d);
            if (dis > simi[0]) harmony::heap_replace_top<M>(k, simi, idxi, dis, ids[r]);
        }
    }
}

template <MetricType M>
void search_slice(int64_t nq, const float* q, int64_t d, int64_t nlist,
                  const float* cent, const int64_t* list_off,
                  const int64_t* ids, const float* codes, int64_t P, int64_t k,
                  float* D, int64_t* I, int64_t* probes_out) {
    std::unique_ptr<float[]> cd(new float[P]);
    std::unique_ptr<idx_t[]> cl(new idx_t[P]);
    std::unique_ptr<int64_t[]> cid(new int64_t[nlist]);
    for (int64_t c = 0; c < nlist; ++c) cid[c] = c;
    for (int64_t i = 0; i < nq; ++i) {
        const float* qv = q + i * d;
        harmony::heap_init<M>(P, cd.get(), cl.get());
        scan_rows<M>(P, qv, cent, cid.get(), nlist, d, cd.get(), cl.get());
        harmony::heap_sort<M>(P, cd.get(), cl.get());
        if (probes_out)
            for (int64_t j = 0; j < P; ++j) probes_out[i * P + j] = cl[j];
        float* simi = D + i * k;
        idx_t* idxi = I + i * k;
        harmony::heap_init<M>(k, simi, idxi);
        for (int64_t j = 0; j < P; ++j) {
            int64_t l = cl[j];
            if (l < 0) continue;
            int64_t lo = list_off[l], hi = list_off[l + 1];
            scan_rows<M>(k, qv, codes + lo * d, ids + lo, hi - lo, d, simi, idxi);
        }
        harmony::heap_sort<M>(k, simi, idxi);
    }
}

// Max-heap of (distance, id) ordered lexicographically (faiss CMax with
// cmp2): the root is the worst kept entry.
inline bool worse(float a, int64_t ia, float b, int64_t ib) { return a > b || (a == b && ia > ib); }

void fheap_replace_top(int64_t k, float* v, int64_t* id, float nv, int64_t nid) {
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1;
        if (l >= k) break;
        int64_t c = l;
        if (r < k && worse(v[r], id[r], v[l], id[l])) c = r;
        if (!worse(v[c], id[c], nv, nid)) break;
        v[i] = v[c];
        id[i] = id[c];
        i = c;
    }
    v[i] = nv;
    id[i] = nid;
}

void fheap_sort(int64_t k, float* v, int64_t* id) {
    // ascending (distance, id); sentinels (FLT_MAX, -1) stay last
    for (int64_t m = k; m > 1; --m) {
        float tv = v[0];
        int64_t ti = id[0];
        fheap_replace_top(m - 1, v, id, v[m - 1], id[m - 1]);
        v[m - 1] = tv;
        id[m - 1] = ti;
    }
}

void fheap_init(int64_t k, float* v, int64_t* id) {
    for (int64_t i = 0; i < k; ++i) {
        v[i] = 3.402823466e+38f;
        id[i] = -1;
    }
}

void base1_slice(int64_t nq, const float* q, int64_t d, int64_t nlist, const float* cent, const int64_t* list_off,
                 const int64_t* ids, const float* codes, int64_t P, int64_t k, int64_t L, int64_t W, float* D,
                 int64_t* I) {
    constexpr MetricType M = MetricType::METRIC_L2;
    std::unique_ptr<float[]> cd(new float[P]);
    std::unique_ptr<idx_t[]> cl(new idx_t[P]);
    std::unique_ptr<int64_t[]> cid(new int64_t[nlist]);
    for (int64_t c = 0; c < nlist; ++c) cid[c] = c;
    const int num_check = 4;
    const size_t check_dim = (size_t)d / num_check;  // src/node.cpp:640
    for (int64_t i = 0; i < nq; ++i) {
        const float* qv = q + i * d;
        harmony::heap_init<M>(P, cd.get(), cl.get());
        scan_rows<M>(P, qv, cent, cid.get(), nlist, d, cd.get(), cl.get());
        harmony::heap_sort<M>(P, cd.get(), cl.get());
        float* simi = D + i * k;
        idx_t* idxi = I + i * k;
        // warm-up bound
        harmony::heap_init<M>(k, simi, idxi);
        const int64_t Lw = L < P ? L : P;
        for (int64_t j = 0; j < Lw; ++j) {
            int64_t l = cl[j];
            if (l < 0) continue;
            int64_t lo = list_off[l], n = list_off[l + 1] - lo;
            if (n > W) n = W;
            scan_rows<M>(k, qv, codes + lo * d, ids + lo, n, d, simi, idxi);
        }
        const float tau = simi[0];
        harmony::heap_init<M>(k, simi, idxi);
        for (int64_t j = 0; j < P; ++j) {
            int64_t l = cl[j];
            if (l < 0) continue;
            for (int64_t r = list_off[l]; r < list_off[l + 1]; ++r) {
                const float* x = codes + r * d;
                float dis = 0.0f;
                for (int check = 0; check < num_check; ++check) {
                    if (check == num_check - 1) {
                        dis += l2(qv + check_dim * check, x + check_dim * check, d - check * check_dim);
                    } else {
                        dis += l2(qv + check_dim * check, x + check_dim * check, check_dim);
                        if (dis > tau) {
                            dis = 3.402823466e+38f;
                            break;
                        }
                    }
                }
                if (dis < simi[0]) harmony::heap_replace_top<M>(k, simi, idxi, dis, ids[r]);
            }
        }
        harmony::heap_sort<M>(k, simi, idxi);
    }
}

void faiss_slice(int64_t nq, const float* q, int64_t d, int64_t nlist, const float* cent, const float* cnorm,
                 const int64_t* list_off, const int64_t* ids, const float* codes, int64_t P, int64_t k, float* D,
                 int64_t* I) {
    std::unique_ptr<float[]> ipb(new float[nlist]);
    std::unique_ptr<float[]> cd(new float[P]);
    std::unique_ptr<int64_t[]> cl(new int64_t[P]);
    for (int64_t i = 0; i < nq; ++i) {
        const float* qv = q + i * d;
        const float qn = ip(qv, qv, d);
        for (int64_t c = 0; c < nlist; ++c) ipb[c] = ip(qv, cent + c * d, d);
        fheap_init(P, cd.get(), cl.get());
        for (int64_t c = 0; c < nlist; ++c) {
            float dis = qn + cnorm[c] - 2.0f * ipb[c];
            if (dis < 0.0f) dis = 0.0f;
            if (worse(cd[0], cl[0], dis, c)) fheap_replace_top(P, cd.get(), cl.get(), dis, c);
        }
        fheap_sort(P, cd.get(), cl.get());
        float* simi = D + i * k;
        int64_t* idxi = I + i * k;
        fheap_init(k, simi, idxi);
        for (int64_t j = 0; j < P; ++j) {
            const int64_t l = cl[j];
            if (l < 0) continue;
            for (int64_t r = list_off[l]; r < list_off[l + 1]; ++r) {
                const float dis = l2(qv, codes + r * d, d);
                if (dis < simi[0]) fheap_replace_top(k, simi, idxi, dis, ids[r]);
            }
        }
        fheap_sort(k, simi, idxi);
    }
}

template <class F>
void query_slices(int64_t nq, int threads, F f) {
    int nt = threads > 0 ? threads : omp_get_max_threads();
    if (nt > nq) nt = nq > 0 ? (int)nq : 1;
#pragma omp parallel for num_threads(nt)
    for (int t = 0; t < nt; ++t) {
        const int64_t s = nq * t / nt, e = nq * (t + 1) / nt;
        if (s < e) f(s, e);
    }
}

}  // namespace

extern "C" {

// "Harmony base, one worker" (L2): warm-up bound over the first L probed
// lists (W rows each), then the four-check pruned scan.  Returns P.
int64_t ref_search_base1(int64_t nq, const float* q, int64_t d, int64_t nlist, const float* cent,
                         const int64_t* list_off, const int64_t* ids, const float* codes, int64_t nprobe, int64_t k,
                         int64_t L, int64_t W, float* D, int64_t* I, int threads) {
    const int64_t P = (nprobe <= 0 || nprobe > nlist) ? nlist : nprobe;
    query_slices(nq, threads, [&](int64_t s, int64_t e) {
        base1_slice(e - s, q + s * d, d, nlist, cent, list_off, ids, codes, P, k, L, W, D + s * k, I + s * k);
    });
    return P;
}

// faiss IndexIVFFlat search algorithm (L2).  Returns P.
int64_t ref_search_faiss(int64_t nq, const float* q, int64_t d, int64_t nlist, const float* cent,
                         const int64_t* list_off, const int64_t* ids, const float* codes, int64_t nprobe, int64_t k,
                         float* D, int64_t* I, int threads) {
    const int64_t P = (nprobe <= 0 || nprobe > nlist) ? nlist : nprobe;
    std::unique_ptr<float[]> cnorm(new float[nlist]);
    for (int64_t c = 0; c < nlist; ++c) cnorm[c] = ip(cent + c * d, cent + c * d, d);
    query_slices(nq, threads, [&](int64_t s, int64_t e) {
        faiss_slice(e - s, q + s * d, d, nlist, cent, cnorm.get(), list_off, ids, codes, P, k, D + s * k, I + s * k);
    });
    return P;
}

// Reference heap as a streaming top-k: strict admission against the root,
// heap_replace_top, then heap_sort (src/heap.hpp:13-90, src/IVFScan.hpp:131).
void ref_topk(int metric, int64_t n, const float* val, const int64_t* ids,
              int64_t k, float* out_val, int64_t* out_ids) {
    if (metric == 1) {
        harmony::heap_init<MetricType::METRIC_L2>(k, out_val, out_ids);
        for (int64_t i = 0; i < n; ++i)
            if (val[i] < out_val[0])
                harmony::heap_replace_top<MetricType::METRIC_L2>(k, out_val, out_ids, val[i], ids[i]);
        harmony::heap_sort<MetricType::METRIC_L2>(k, out_val, out_ids);
    } else {
        harmony::heap_init<MetricType::METRIC_IP>(k, out_val, out_ids);
        for (int64_t i = 0; i < n; ++i)
            if (val[i] > out_val[0])
                harmony::heap_replace_top<MetricType::METRIC_IP>(k, out_val, out_ids, val[i], ids[i]);
        harmony::heap_sort<MetricType::METRIC_IP>(k, out_val, out_ids);
    }
}

// Plain IVF-Flat search of the reference's --mode original on `threads`
// OpenMP threads (0 = all), over a CSR index (list l = rows
// [list_off[l], list_off[l+1])).  probes_out (optional) receives the coarse
// quantizer's probe lists, [nq x P].  Returns P (nprobe clamped to nlist,
// src/Index.cpp:1646-1648).
int64_t ref_search(int metric, int64_t nq, const float* q, int64_t d,
                   int64_t nlist, const float* cent, const int64_t* list_off,
                   const int64_t* ids, const float* codes, int64_t nprobe,
                   int64_t k, float* D, int64_t* I, int64_t* probes_out,
                   int threads) {
    int64_t P = (nprobe <= 0 || nprobe > nlist) ? nlist : nprobe;
    int nt = threads > 0 ? threads : omp_get_max_threads();
    if (nt > nq) nt = nq > 0 ? (int)nq : 1;
    int64_t base = nq / nt, extra = nq % nt;
#pragma omp parallel for num_threads(nt)
    for (int t = 0; t < nt; ++t) {
        int64_t s = t < extra ? t * (base + 1) : t * base + extra;
        int64_t e = s + base + (t < extra ? 1 : 0);
        if (s >= e) continue;
        int64_t* po = probes_out ? probes_out + s * P : nullptr;
        if (metric == 1)
            search_slice<MetricType::METRIC_L2>(e - s, q + s * d, d, nlist, cent, list_off, ids,
                                                codes, P, k, D + s * k, I + s * k, po);
        else
            search_slice<MetricType::METRIC_IP>(e - s, q + s * d, d, nlist, cent, list_off, ids,
                                                codes, P, k, D + s * k, I + s * k, po);
    }
    return P;
}

}  // extern "C"
