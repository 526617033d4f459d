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

}  // namespace

extern "C" {

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
