// hivf::Index -- header-only C++ mirror of harmony::Index (src/Index.h:13-149)
// over the C ABI in hivf.h, so the reference's C++ callers (src/query.cpp,
// src/main.cpp, src/tests/index_test.cpp) switch to the MI355X path by
// changing the type.
//
// Constructor, move-assignment, method names, argument meaning, public
// fields and error behaviour (std::runtime_error) follow the reference;
// search() additionally accepts hivf_search_params.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "hivf.h"

namespace hivf {

// src/common.h:20-38
enum MetricType { METRIC_IP = HIVF_METRIC_IP, METRIC_L2 = HIVF_METRIC_L2 };
enum OptLevel {
    OPT_NONE = 0b000,
    OPT_TRIANGLE = 0b001,
    OPT_SUBNN_L2 = 0b010,
    OPT_SUBNN_IP = 0b100,
    OPT_TRI_SUBNN_L2 = 0b011,
    OPT_TRI_SUBNN_IP = 0b101,
    OPT_SUBNN_ONLY = 0b110,
    OPT_ALL = 0b111
};
enum EdgeDevice { EDGEDEVIVE_ENABLED, EDGEDEVIVE_DISABLED };

// Per-search counters (the reference's Stats, src/stats.h:14-152).
using Stats = hivf_search_stats;

class Index {
   public:
    // src/Index.h:15-25 / src/Index.cpp:24-43.  A default-constructed index
    // (d = 0) holds no device state until an index is move-assigned or
    // loaded into it, as `Index index; index = Index(...)` in
    // src/query.cpp:825-835.  opt_level, sub_k, sub_nlist and sub_nprobe
    // select the reference's Tribase pruning structures; they are kept in the
    // public fields and do not change results, so this path ignores them.
    // verbose and edge_device_enabled are kept likewise.
    explicit Index(size_t d = 0, size_t nlist = 0, size_t nprobe = 0, MetricType metric = METRIC_L2,
                   OptLevel opt_level = OPT_NONE, size_t sub_k = 0, size_t sub_nlist = 1, size_t sub_nprobe = 1,
                   bool verbose = false, EdgeDevice edge_device_enabled = EDGEDEVIVE_DISABLED)
        : d(d),
          nlist(nlist),
          nprobe(nprobe),
          metric(metric),
          opt_level(opt_level),
          added_opt_level(opt_level),
          sub_k(sub_k),
          sub_nlist(sub_nlist),
          sub_nprobe(sub_nprobe),
          verbose(verbose),
          edge_device_enabled(edge_device_enabled) {
        if (d == 0 && nlist == 0) return;
        h_ = hivf_index_create((int64_t)d, (int64_t)nlist, (int)metric);
        if (!h_) throw std::runtime_error(std::string("hivf: ") + hivf_last_error(nullptr));
    }
    Index(const Index&) = delete;
    Index& operator=(const Index&) = delete;
    Index(Index&& o) noexcept { *this = std::move(o); }
    // src/Index.cpp:363-379: takes the other index's fields and lists
    Index& operator=(Index&& o) noexcept {
        if (this == &o) return *this;
        hivf_index_destroy(h_);
        h_ = o.h_;
        o.h_ = nullptr;
        d = o.d;
        nlist = o.nlist;
        nprobe = o.nprobe;
        metric = o.metric;
        opt_level = o.opt_level;
        added_opt_level = o.added_opt_level;
        sub_k = o.sub_k;
        sub_nlist = o.sub_nlist;
        sub_nprobe = o.sub_nprobe;
        verbose = o.verbose;
        edge_device_enabled = o.edge_device_enabled;
        niter_ = o.niter_;
        seed_ = o.seed_;
        return *this;
    }
    ~Index() { hivf_index_destroy(h_); }

    // src/Index.cpp:381-425: k-means with the reference's recipe -- seed
    // 6666, at most 256 rows per centroid, 20 rounds (2 with `lite`),
    // sample and initial centroids drawn by std::shuffle under
    // std::default_random_engine(seed) (src/Clustering.cpp:48-89).  `faiss`
    // selects faiss's trainer in the reference; faiss is not linked here, so
    // both settings run this library's k-means.  set_train_params overrides
    // rounds and seed.
    void train(size_t n, const float* codes, bool faiss = false, bool lite = false) {
        (void)faiss;
        const int niter = niter_ > 0 ? niter_ : (lite ? 2 : 20);
        check(hivf_index_train_kmeans(handle(), (int64_t)n, codes, niter, seed_, 256));
    }
    void set_train_params(int niter, uint64_t seed) {
        niter_ = niter;
        seed_ = seed;
    }

    // src/Index.cpp:581-698: replaces all lists; ids are row numbers.
    void add(size_t n, const float* codes) { check(hivf_index_add(handle(), (int64_t)n, codes)); }

    // src/Index.h:82.  The public field `nprobe` is read at every call (0 =
    // every list, as src/query.cpp:849 sets it); `ratio` is accepted for
    // signature compatibility and ignored, as in the reference's L2 path.
    Stats search(size_t n, const float* queries, size_t k, float* distances, int64_t* labels, float ratio = 1,
                 const hivf_search_params* params = nullptr) {
        (void)ratio;
        hivf_search_params p = resolve(params);
        Stats st{};
        check(hivf_index_search(handle(), (int64_t)n, queries, (int64_t)k, distances, labels, &p, &st));
        return st;
    }

    // Same on device buffers (HBM) and a HIP stream (0 = the index's own);
    // with params->use_graph the search is captured once per shape and
    // replayed (no statistics then).
    void search_device(size_t n, const float* queries_dev, size_t k, float* distances_dev, int64_t* labels_dev,
                       const hivf_search_params* params = nullptr, void* stream = nullptr,
                       Stats* stats = nullptr) {
        hivf_search_params p = resolve(params);
        check(hivf_index_search_device(handle(), (int64_t)n, queries_dev, (int64_t)k, distances_dev, labels_dev,
                                       &p, stats, stream));
    }

    // src/Index.cpp:1708-1770 file format.
    void save_index(const std::string& path) const {
        if (!h_) throw std::runtime_error("hivf: empty index");
        check(hivf_index_save_harmony(h_, path.c_str()));
    }
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
        metric = (MetricType)m;
        opt_level = OPT_NONE;  // src/Index.cpp:1745
    }

    hivf_index* handle() const {
        if (!h_) throw std::runtime_error("hivf: empty index (constructed without d and nlist)");
        return h_;
    }

    // public fields of the reference (src/Index.h:99-116)
    size_t d;
    size_t nlist;
    size_t nprobe;
    MetricType metric;
    OptLevel opt_level;
    OptLevel added_opt_level;
    size_t sub_k;
    size_t sub_nlist;
    size_t sub_nprobe;
    bool verbose;
    EdgeDevice edge_device_enabled;

   private:
    hivf_search_params resolve(const hivf_search_params* params) const {
        hivf_search_params p;
        hivf_search_params_default(&p);
        if (params) p = *params;
        check(hivf_index_set_nprobe(handle(), (int64_t)nprobe));  // 0 -> nlist
        if (p.nprobe <= 0) p.nprobe = nprobe > 0 ? (int64_t)nprobe : (int64_t)nlist;
        return p;
    }
    void check(int rc) const {
        if (rc != HIVF_OK) throw std::runtime_error(std::string("hivf: ") + hivf_last_error(h_));
    }
    hivf_index* h_ = nullptr;
    int niter_ = 0;
    uint64_t seed_ = 6666;
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
