// hivf::Index -- header-only C++ mirror of harmony::Index (src/Index.h:15-93)
// over the C ABI in hivf.h, so the reference's C++ callers (src/query.cpp,
// src/main.cpp) can switch to the MI355X path by changing the type.
//
// Method names, argument meaning and error behaviour (std::runtime_error)
// follow the reference; search() additionally accepts hivf_search_params.
#pragma once

#include <algorithm>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "hivf.h"

namespace hivf {

// Per-search counters (the reference's Stats, src/stats.h:14-152).
using Stats = hivf_search_stats;

class Index {
   public:
    // src/Index.h:15-25: Index(d, nlist, nprobe, metric, ...)
    Index(size_t d, size_t nlist, size_t nprobe = 1, int metric = HIVF_METRIC_L2)
        : d(d), nlist(nlist), nprobe(nprobe), metric(metric) {
        h_ = hivf_index_create((int64_t)d, (int64_t)nlist, metric);
        if (!h_) throw std::runtime_error(std::string("hivf: ") + hivf_last_error(nullptr));
    }
    Index(const Index&) = delete;
    Index& operator=(const Index&) = delete;
    Index(Index&& o) noexcept : d(o.d), nlist(o.nlist), nprobe(o.nprobe), metric(o.metric), h_(o.h_) {
        o.h_ = nullptr;
    }
    ~Index() { hivf_index_destroy(h_); }

    // src/Index.cpp:381-425: k-means with niter rounds (20 in the reference)
    // from nlist rows of a seeded sample of at most 256 * nlist rows
    // (src/Clustering.cpp:48-89).
    void train(size_t n, const float* codes, int niter = 20, uint64_t seed = 6666) {
        const size_t cap = 256 * nlist;
        std::vector<size_t> idx(n);
        for (size_t i = 0; i < n; ++i) idx[i] = i;
        std::mt19937_64 rng(seed);
        std::shuffle(idx.begin(), idx.end(), rng);
        const size_t m = n < cap ? n : cap;
        if (m < nlist) throw std::runtime_error("hivf: fewer training rows than lists");
        std::vector<float> sample(m * d), init(nlist * d);
        for (size_t i = 0; i < m; ++i) std::copy(codes + idx[i] * d, codes + idx[i] * d + d, &sample[i * d]);
        for (size_t c = 0; c < nlist; ++c) std::copy(&sample[c * d], &sample[c * d] + d, &init[c * d]);
        check(hivf_index_train(h_, (int64_t)m, sample.data(), init.data(), niter));
    }

    // src/Index.cpp:581-698: replaces all lists; ids are row numbers.
    void add(size_t n, const float* codes) { check(hivf_index_add(h_, (int64_t)n, codes)); }

    // src/Index.h:82.  `ratio` is accepted for signature compatibility and
    // ignored, as in the reference's L2 path.
    Stats search(size_t n, const float* queries, size_t k, float* distances, int64_t* labels, float ratio = 1,
                 const hivf_search_params* params = nullptr) {
        (void)ratio;
        hivf_search_params p;
        hivf_search_params_default(&p);
        if (params) p = *params;
        if (p.nprobe <= 0) p.nprobe = (int64_t)nprobe;
        Stats st{};
        check(hivf_index_search(h_, (int64_t)n, queries, (int64_t)k, distances, labels, &p, &st));
        return st;
    }

    // Same on device buffers (HBM) and a HIP stream (0 = the index's own);
    // with params->use_graph the search is captured once per shape and
    // replayed (no statistics then).
    void search_device(size_t n, const float* queries_dev, size_t k, float* distances_dev, int64_t* labels_dev,
                       const hivf_search_params* params = nullptr, void* stream = nullptr,
                       Stats* stats = nullptr) {
        hivf_search_params p;
        hivf_search_params_default(&p);
        if (params) p = *params;
        if (p.nprobe <= 0) p.nprobe = (int64_t)nprobe;
        check(hivf_index_search_device(h_, (int64_t)n, queries_dev, (int64_t)k, distances_dev, labels_dev, &p,
                                       stats, stream));
    }

    // src/Index.cpp:1708-1770 file format.
    void save_index(const std::string& path) const { check(hivf_index_save_harmony(h_, path.c_str())); }
    void load_index(const std::string& path) {
        int st = 0;
        hivf_index* h = hivf_index_load_harmony(path.c_str(), &st);
        if (!h) throw std::runtime_error(std::string("hivf: ") + hivf_last_error(nullptr));
        hivf_index_destroy(h_);
        h_ = h;
        int64_t dd = 0, nl = 0;
        int m = 0;
        hivf_index_info(h_, &dd, &nl, &m);
        d = (size_t)dd;
        nlist = (size_t)nl;
        metric = m;
    }

    hivf_index* handle() { return h_; }

    // public fields of the reference (src/Index.h:99-104)
    size_t d;
    size_t nlist;
    size_t nprobe;
    int metric;

   private:
    void check(int rc) const {
        if (rc != HIVF_OK) throw std::runtime_error(std::string("hivf: ") + hivf_last_error(h_));
    }
    hivf_index* h_ = nullptr;
};

// Top-k merge of per-partition results [parts][n][k_in] into [n][k]
// (src/Index.cpp:1327-1343), e.g. the rows of a list-sharded search.
inline void merge_topk(int metric, size_t n, size_t parts, size_t k_in, const float* D_parts,
                       const int64_t* I_parts, size_t k, float* D, int64_t* I) {
    if (hivf_merge_topk(metric, (int64_t)n, (int64_t)parts, (int64_t)k_in, D_parts, I_parts, (int64_t)k, D, I) !=
        HIVF_OK)
        throw std::runtime_error(std::string("hivf: ") + hivf_last_error(nullptr));
}

}  // namespace hivf
