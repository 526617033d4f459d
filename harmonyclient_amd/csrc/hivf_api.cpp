// Host side of the hivf C ABI (include/hivf.h): index state, device arena,
// and the search pipeline that drives the kernels.
//
// Mirrors harmony::Index (src/Index.h:15-149): create = ctor, train, add,
// search, save/load.  The handle owns one HIP stream on the device that was
// current at creation; every entry point re-selects that device.
#include "hivf.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "harmony_format.hpp"
#include "index_impl.hpp"

using namespace hivf;

namespace {
std::mutex g_err_mu;
std::string g_err;
std::atomic<uint64_t> g_free_gen{0};
}  // namespace

void hivf::note_device_free() { g_free_gen.fetch_add(1, std::memory_order_relaxed); }
uint64_t hivf::device_free_generation() { return g_free_gen.load(std::memory_order_relaxed); }

int hivf::api_exception(hivf_index* h, const char* fn) noexcept {
    int code = HIVF_EINVAL;
    const char* what = "unknown exception";
    try {
        throw;
    } catch (const std::bad_alloc&) {
        code = HIVF_ENOMEM;
        what = "host allocation failed";
    } catch (const std::length_error&) {
        code = HIVF_ENOMEM;
        what = "size beyond what a host buffer can hold";
    } catch (const std::exception& e) {
        what = e.what();
    } catch (...) {
    }
    try {
        std::string m = std::string(fn) + ": " + what;
        if (h)
            h->err = m;
        else
            set_global_error(m);
    } catch (...) {
    }
    return code;
}

void hivf::set_global_error(const std::string& m) {
    std::lock_guard<std::mutex> lk(g_err_mu);
    g_err = m;
}

namespace hivf {

// Top-P centroids of n staged queries into d_cd / d_probes: the MFMA form
// for searches, the per-query kernel for row assignment (P = 1) and d > 512.
// With prep (the batched search's query preparation), the MFMA form runs it
// in its own query-preparation launch and sets *prep_done.
int run_coarse(hivf_index* h, const float4* q, int64_t n, int64_t P, hipStream_t s, const PrepArgs* prep,
               bool* prep_done) {
    if (prep_done) *prep_done = false;
    if (P > 1) {
        int rc = run_coarse_mfma(h, q, n, P, s, prep);
        if (rc != HIVF_EUNSUPPORTED) {
            if (prep_done && rc == HIVF_OK) *prep_done = prep != nullptr;
            return rc;
        }
    }
    HIP_TRY(h, h->d_cd.ensure((size_t)n * P));
    HIP_TRY(h, h->d_probes.ensure((size_t)n * P));
    CoarseArgs c;
    c.q = q;
    c.nq = (int)n;
    c.nquad = h->nquad;
    c.cent = h->d_cent.p;
    c.nlist = (int)h->nlist;
    c.P = (int)P;
    c.out_d = h->d_cd.p;
    c.out_l = h->d_probes.p;
    HIP_TRY(h, launch_coarse(c, h->metric, s));
    return HIVF_OK;
}

}  // namespace hivf

namespace {

int parallel_threads() {
    unsigned n = std::thread::hardware_concurrency();
    return n == 0 ? 4 : (int)std::min(n, 32u);
}

template <class F>
void parallel_for(int64_t n, F f) {
    int nt = (int)std::min<int64_t>(parallel_threads(), std::max<int64_t>(n, 1));
    if (nt <= 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int64_t i = t; i < n; i += nt) f(i);
        });
    for (auto& x : th) x.join();
}

__global__ void fill_u64_kernel(unsigned long long* p, int n, unsigned long long v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Row-major [n x d] -> [n x d_pad] with zero padding (device to device).
__global__ void pad_rows_kernel(const float* __restrict__ x, float* __restrict__ y, int64_t n, int d,
                                int d_pad) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * d_pad) return;
    int64_t r = i / d_pad;
    int c = (int)(i % d_pad);
    y[i] = c < d ? x[r * d + c] : 0.0f;
}

}  // namespace

namespace {

// Rows of a list in tile layout: see DevLists in kernels.hpp.
void tile_rows(const float* rows, int64_t nrows, int64_t d, int nquad,
               float* tile_base /* first tile of the list */) {
    for (int64_t r = 0; r < nrows; ++r) {
        int64_t t = r / kTileRows, lane = r % kTileRows;
        const float* src = rows + r * d;
        for (int j = 0; j < nquad; ++j) {
            float* dst = tile_base + ((t * nquad + j) * kTileRows + lane) * 4;
            for (int c = 0; c < 4; ++c) {
                int64_t dim = (int64_t)j * 4 + c;
                dst[c] = dim < d ? src[dim] : 0.0f;
            }
        }
    }
}

// Upload lists given as a row gatherer: row(l, r) -> pointer to d floats and
// its id.
template <class RowFn>
int upload_lists(hivf_index* h, const std::vector<int64_t>& list_off, RowFn row) {
    const int64_t nlist = h->nlist;
    std::vector<int32_t> tile_off(nlist + 1, 0);
    std::vector<int32_t> size(nlist, 0);
    int64_t tiles = 0;
    for (int64_t l = 0; l < nlist; ++l) {
        int64_t n_l = list_off[l + 1] - list_off[l];
        if (n_l > INT32_MAX) return h->fail(HIVF_EINVAL, "list too long");
        size[l] = (int32_t)n_l;
        tile_off[l] = (int32_t)tiles;
        tiles += (n_l + kTileRows - 1) / kTileRows;
        if (tiles > INT32_MAX / kTileRows) return h->fail(HIVF_EINVAL, "index too large");
    }
    tile_off[nlist] = (int32_t)tiles;
    const int nquad = h->nquad;
    const size_t tile_floats = (size_t)nquad * kTileRows * 4;
    std::vector<float> codes;
    std::vector<uint32_t> ids;
    std::vector<float> norms;
    try {
        codes.assign(std::max<size_t>(tiles * tile_floats, 4), 0.0f);
        ids.assign(std::max<int64_t>(tiles * kTileRows, 1), kNoId);
        norms.assign(std::max<int64_t>(tiles * kTileRows, 1), 0.0f);
    } catch (const std::bad_alloc&) {
        return h->fail(HIVF_ENOMEM, "host allocation for the list arena failed");
    }
    int bad = 0;
    std::vector<char> small(nlist, 1);
    parallel_for(nlist, [&](int64_t l) {
        int64_t n_l = size[l];
        float* base = codes.data() + (size_t)tile_off[l] * tile_floats;
        std::vector<float> tmp((size_t)std::max<int64_t>(n_l, 1) * h->d);
        for (int64_t r = 0; r < n_l; ++r) {
            int64_t id;
            const float* src = row(l, r, &id);
            if (id < 0 || id >= (int64_t)kNoId) bad = 1;
            std::memcpy(tmp.data() + r * h->d, src, sizeof(float) * h->d);
            size_t slot = (size_t)tile_off[l] * kTileRows + r;
            ids[slot] = (uint32_t)id;
            double s = 0.0;  // norms in FP64, rounded once
            for (int64_t c = 0; c < h->d; ++c) {
                s += (double)src[c] * (double)src[c];
                if (!(src[c] >= 0.0f && src[c] <= 255.0f && src[c] == std::rint(src[c]))) small[l] = 0;
            }
            norms[slot] = (float)s;
        }
        tile_rows(tmp.data(), n_l, h->d, nquad, base);
    });
    if (bad) return h->fail(HIVF_EINVAL, "ids must lie in [0, 2^32 - 2]");

    HIP_TRY(h, h->d_codes.ensure(codes.size() / 4));
    HIP_TRY(h, h->d_ids.ensure(ids.size()));
    HIP_TRY(h, h->d_norms.ensure(norms.size()));
    HIP_TRY(h, h->d_tile_off.ensure(nlist + 1));
    HIP_TRY(h, h->d_size.ensure(nlist));
    HIP_TRY(h, hipMemcpy(h->d_codes.p, codes.data(), codes.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(h, hipMemcpy(h->d_ids.p, ids.data(), ids.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(h, hipMemcpy(h->d_norms.p, norms.data(), norms.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(h, hipMemcpy(h->d_tile_off.p, tile_off.data(), (nlist + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(h, hipMemcpy(h->d_size.p, size.data(), nlist * sizeof(int32_t), hipMemcpyHostToDevice));
    h->lists_uint8 = std::all_of(small.begin(), small.end(), [](char c) { return c != 0; });
    h->list_off = list_off;
    h->tile_off = tile_off;
    h->ntotal = list_off[nlist];
    h->tiles = tiles;
    h->have_lists = true;
    h->batched.invalidate();
    h->drop_graph();
    return HIVF_OK;
}

int upload_centroids(hivf_index* h, const float* c) {
    h->cent.assign(c, c + h->nlist * h->d);
    int64_t tiles = (h->nlist + kTileRows - 1) / kTileRows;
    std::vector<float> t((size_t)tiles * h->nquad * kTileRows * 4, 0.0f);
    for (int64_t tt = 0; tt < tiles; ++tt) {
        int64_t rows = std::min<int64_t>(kTileRows, h->nlist - tt * kTileRows);
        tile_rows(c + tt * kTileRows * h->d, rows, h->d, h->nquad,
                  t.data() + (size_t)tt * h->nquad * kTileRows * 4);
    }
    HIP_TRY(h, h->d_cent.ensure(t.size() / 4));
    HIP_TRY(h, hipMemcpy(h->d_cent.p, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> rows((size_t)h->nlist * h->nquad * 4, 0.0f);
    for (int64_t i = 0; i < h->nlist; ++i)
        std::memcpy(rows.data() + (size_t)i * h->nquad * 4, c + i * h->d, sizeof(float) * h->d);
    HIP_TRY(h, h->d_cent_rows.ensure(rows.size() / 4));
    HIP_TRY(h, hipMemcpy(h->d_cent_rows.p, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> cn((size_t)tiles * kTileRows, 0.0f);
    for (int64_t i = 0; i < h->nlist; ++i) {
        double acc = 0.0;
        for (int64_t j = 0; j < h->d; ++j) acc += (double)c[i * h->d + j] * (double)c[i * h->d + j];
        cn[i] = (float)acc;
    }
    HIP_TRY(h, h->d_cent_norm.ensure(cn.size()));
    HIP_TRY(h, hipMemcpy(h->d_cent_norm.p, cn.data(), cn.size() * sizeof(float), hipMemcpyHostToDevice));
    h->coarse.valid = false;
    h->have_centroids = true;
    h->batched.invalidate();
    h->drop_graph();
    return HIVF_OK;
}

// Queries (host or device, row-major d floats) into the padded device
// buffer d_q.  Host rows are uploaded to a staging area first; the padding
// is written by a kernel, not a memset, so the device path stays safe under
// stream capture.
int stage_queries(hivf_index* h, int64_t n, const float* x, hipMemcpyKind kind, hipStream_t s) {
    HIP_TRY(h, h->d_q.ensure((size_t)n * h->nquad));
    if (h->d_pad == h->d) {
        HIP_TRY(h, hipMemcpyAsync(h->d_q.p, x, (size_t)n * h->d * sizeof(float), kind, s));
        return HIVF_OK;
    }
    const float* src = x;
    if (kind == hipMemcpyHostToDevice) {
        HIP_TRY(h, h->d_stage.ensure((size_t)n * h->d));
        HIP_TRY(h, hipMemcpyAsync(h->d_stage.p, x, (size_t)n * h->d * sizeof(float), kind, s));
        src = h->d_stage.p;
    }
    int64_t total = n * h->d_pad;
    hipLaunchKernelGGL(pad_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src,
                       reinterpret_cast<float*>(h->d_q.p), n, (int)h->d, (int)h->d_pad);
    HIP_TRY(h, hipGetLastError());
    return HIVF_OK;
}

int64_t clamp_nprobe(const hivf_index* h, int64_t p) {
    if (p <= 0 || p > h->nlist) return h->nlist;
    return p;
}

ScanArgs base_scan_args(hivf_index* h, const float4* q, int64_t n, int64_t P, int64_t k) {
    ScanArgs a{};
    a.lists = h->lists();
    a.q = q;
    a.nq = (int)n;
    a.probes = h->d_probes.p;
    a.probe_stride = (int)P;
    a.nprobe_q = nullptr;
    a.nprobe = (int)P;
    a.k = (int)k;
    a.block_dims = (int)h->d;
    a.nblocks = 1;
    a.tau = nullptr;
    a.row_limit = INT_MAX;
    a.warm_lists = INT_MAX;
    a.stats = nullptr;
    return a;
}

int check_params(hivf_index* h, int64_t k, const hivf_search_params& p) {
    if (k < 1 || k > kMaxK) return h->fail(HIVF_EUNSUPPORTED, "k must lie in [1, 1024]");
    if (p.algo < HIVF_ALGO_PLAIN || p.algo > HIVF_ALGO_BATCHED)
        return h->fail(HIVF_EINVAL, "unknown algo");
    if (p.tau_source < HIVF_TAU_NONE || p.tau_source > HIVF_TAU_LEARNED)
        return h->fail(HIVF_EINVAL, "unknown tau_source");
    if (p.algo == HIVF_ALGO_PRUNED && h->metric != HIVF_METRIC_L2)
        return h->fail(HIVF_EUNSUPPORTED, "dimension-block pruning is defined for L2 only");
    if (p.dim_block < 0) return h->fail(HIVF_EINVAL, "dim_block must be >= 0");
    if (p.precision != HIVF_PREC_FP32 && p.precision != HIVF_PREC_BF16)
        return h->fail(HIVF_EINVAL, "unknown precision");
    if (p.precision == HIVF_PREC_BF16 && p.algo != HIVF_ALGO_BATCHED)
        return h->fail(HIVF_EUNSUPPORTED, "the per-query scans compute in FP32; BF16 is a batched-scan option");
    return HIVF_OK;
}

void set_blocks(const hivf_index* h, const hivf_search_params& p, ScanArgs* a) {
    pruning_blocks(h->d, p.dim_block, &a->block_dims, &a->nblocks);
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0f;
    return ms;
}

// Core search on staged device queries q; results into device D / I.
// Probe count of the search and the device copy of per-query probe counts
// (a synchronous upload: done before any capture).
int resolve_probes(hivf_index* h, int64_t n, const hivf_search_params& p, hipStream_t s, int64_t* P_out,
                   const int32_t** npq_out) {
    int64_t P = clamp_nprobe(h, p.nprobe > 0 ? p.nprobe : h->nprobe);
    if (P > kMaxK) return h->fail(HIVF_EUNSUPPORTED, "nprobe must be <= 1024");
    *npq_out = nullptr;
    if (p.nprobe_per_query_dev) {
        if (p.nprobe_per_query) return h->fail(HIVF_EINVAL, "per-query nprobe given twice (host and device)");
        if (p.nprobe <= 0) return h->fail(HIVF_EINVAL, "device per-query nprobe needs nprobe as their bound");
        *npq_out = p.nprobe_per_query_dev;
    } else if (p.nprobe_per_query) {
        int64_t mx = 1;
        std::vector<int32_t> v(p.nprobe_per_query, p.nprobe_per_query + n);
        for (auto& x : v) {
            if (x <= 0 || x > h->nlist) x = (int32_t)h->nlist;
            mx = std::max<int64_t>(mx, x);
        }
        if (mx > kMaxK) return h->fail(HIVF_EUNSUPPORTED, "nprobe must be <= 1024");
        P = mx;
        HIP_TRY(h, h->d_npq.ensure(n));
        HIP_TRY(h, hipMemcpyAsync(h->d_npq.p, v.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
        HIP_TRY(h, hipStreamSynchronize(s));  // v is a local
        *npq_out = h->d_npq.p;
    }
    *P_out = P;
    return HIVF_OK;
}

int run_search_core(hivf_index* h, const float4* q, int64_t n, int64_t k, float* D, int64_t* I,
                    const hivf_search_params& p, int64_t P, const int32_t* npq, hivf_search_stats* st,
                    hipStream_t s);

int run_search(hivf_index* h, const float4* q, int64_t n, int64_t k, float* D, int64_t* I,
               const hivf_search_params& p, hivf_search_stats* st, hipStream_t s) {
    int rc = check_params(h, k, p);
    if (rc) return rc;
    if (!h->have_centroids || !h->have_lists)
        return h->fail(HIVF_ESTATE, "search before centroids and lists are set");
    int64_t P;
    const int32_t* npq;
    rc = resolve_probes(h, n, p, s, &P, &npq);
    if (rc) return rc;
    return run_search_core(h, q, n, k, D, I, p, P, npq, st, s);
}

int run_search_core(hivf_index* h, const float4* q, int64_t n, int64_t k, float* D, int64_t* I,
                    const hivf_search_params& p, int64_t P, const int32_t* npq, hivf_search_stats* st,
                    hipStream_t s) {
    int rc;
    const bool timing = st != nullptr;
    if (timing) {
        HIP_TRY(h, h->d_stats.ensure(8));
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, h->d_stats.p, 8, 0ull);
        HIP_TRY(h, hipEventRecord(h->ev[0], s));
    }
    // (the batched search runs the coarse step itself, with its own query
    // preparation in the same launch)
    if (p.algo == HIVF_ALGO_BATCHED) {
        return run_batched(h, q, n, k, (int)P, npq, p, D, I, st, s, timing);
    }
    rc = run_coarse(h, q, n, P, s);
    if (rc) return rc;
    if (timing) HIP_TRY(h, hipEventRecord(h->ev[1], s));

    ScanArgs a = base_scan_args(h, q, n, P, k);
    a.nprobe_q = npq;
    a.out_d = D;
    a.out_i = I;
    a.stats = timing ? h->d_stats.p : nullptr;
    const float* tau = nullptr;
    if (p.algo == HIVF_ALGO_PRUNED && p.tau_source == HIVF_TAU_WARMUP) {
        HIP_TRY(h, h->d_tau.ensure(n));
        ScanArgs w = base_scan_args(h, q, n, P, k);
        w.nprobe_q = npq;
        w.row_limit = p.warmup_list_size > 0 ? p.warmup_list_size : (int)k;  // src/query.cpp:847
        w.warm_lists = p.warmup_lists > 0 ? p.warmup_lists : 10;             // src/query.cpp:845
        w.out_tau = h->d_tau.p;
        HIP_TRY(h, launch_scan(w, kScanWarmup, h->metric, s));
        tau = h->d_tau.p;
    } else if (p.algo == HIVF_ALGO_PRUNED && p.tau_source == HIVF_TAU_LEARNED) {
        return h->fail(HIVF_EUNSUPPORTED, "the learned bound is used by the batched scan");
    }
    if (timing) HIP_TRY(h, hipEventRecord(h->ev[2], s));
    a.tau = tau;
    set_blocks(h, p, &a);
    HIP_TRY(h, launch_scan(a, p.algo == HIVF_ALGO_PRUNED ? kScanPruned : kScanPlain, h->metric, s));
    if (timing) {
        HIP_TRY(h, hipEventRecord(h->ev[3], s));
        HIP_TRY(h, hipEventSynchronize(h->ev[3]));
        unsigned long long c[8];
        HIP_TRY(h, hipMemcpy(c, h->d_stats.p, sizeof(c), hipMemcpyDeviceToHost));
        std::memset(st, 0, sizeof(*st));
        st->coarse_ms = elapsed(h->ev[0], h->ev[1]);
        st->tau_ms = elapsed(h->ev[1], h->ev[2]);
        st->scan_ms = elapsed(h->ev[2], h->ev[3]);
        st->total_ms = elapsed(h->ev[0], h->ev[3]);
        st->candidates = (int64_t)c[0];
        st->blocks_total = (int64_t)c[1];
        st->blocks_skipped = (int64_t)c[2];
        st->scan_launches = 1;
    }
    return HIVF_OK;
}

}  // namespace

extern "C" {

void hivf_search_params_default(hivf_search_params* p) {
    HIVF_API_BEGIN
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->algo = HIVF_ALGO_PLAIN;
    p->tau_source = HIVF_TAU_WARMUP;
    p->precision = HIVF_PREC_FP32;
    p->warmup_lists = 10;
    HIVF_API_END_VOID
}

hivf_index* hivf_index_create(int64_t d, int64_t nlist, int metric) {
    HIVF_API_BEGIN
    if (d < 1 || d > 16384 || nlist < 1 || nlist > (1 << 24) ||
        (metric != HIVF_METRIC_L2 && metric != HIVF_METRIC_IP)) {
        set_global_error("hivf_index_create: bad d, nlist or metric");
        return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_global_error("hivf_index_create: no HIP device");
        return nullptr;
    }
    hivf_index* h = new (std::nothrow) hivf_index();
    if (!h) {
        set_global_error("hivf_index_create: out of host memory");
        return nullptr;
    }
    h->d = d;
    h->d_pad = (d + 3) / 4 * 4;
    h->nquad = (int)(h->d_pad / 4);
    h->nlist = nlist;
    h->metric = metric;
    if (hipGetDevice(&h->device) != hipSuccess ||
        hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess ||
        hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        set_global_error("hivf_index_create: device query or stream creation failed");
        delete h;
        return nullptr;
    }
    {
        // the no-allocation search's scratch array on this device: its
        // address is taken here, once per device, with the device current
        hipError_t e = noalloc_setup(h->device, nullptr);
        if (e != hipSuccess) {
            set_global_error(std::string("hivf_index_create: hipGetSymbolAddress: ") + hipGetErrorString(e));
            hivf_index_destroy(h);
            return nullptr;
        }
    }
    for (auto& e : h->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            set_global_error("hivf_index_create: event creation failed");
            hivf_index_destroy(h);
            return nullptr;
        }
    }
    return h;
    HIVF_API_END_PTR(nullptr)
}

void hivf_index_destroy(hivf_index* h) {
    HIVF_API_BEGIN
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->drop_graph();
    for (auto& e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    HIVF_API_END_VOID
}

const char* hivf_last_error(const hivf_index* h) {
    if (h) return h->err.c_str();
    std::lock_guard<std::mutex> lk(g_err_mu);
    static thread_local std::string copy;
    try {
        copy = g_err;
    } catch (...) {
        return "";
    }
    return copy.c_str();
}

int hivf_index_set_nprobe(hivf_index* h, int64_t nprobe) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    h->nprobe = nprobe <= 0 ? h->nlist : nprobe;
    return HIVF_OK;
    HIVF_API_END(h)
}

int hivf_index_set_centroids(hivf_index* h, const float* c) {
    HIVF_API_BEGIN
    if (!h || !c) return h ? h->fail(HIVF_EINVAL, "null centroids") : HIVF_EINVAL;
    HIP_TRY(h, hipSetDevice(h->device));
    return upload_centroids(h, c);
    HIVF_API_END(h)
}

int hivf_index_get_centroids(const hivf_index* h, float* c) {
    HIVF_API_BEGIN
    if (!h || !c) return HIVF_EINVAL;
    if (!h->have_centroids) return HIVF_ESTATE;
    std::memcpy(c, h->cent.data(), sizeof(float) * h->cent.size());
    return HIVF_OK;
    HIVF_API_END(const_cast<hivf_index*>(h))
}

int hivf_index_train(hivf_index* h, int64_t n, const float* x, const float* init, int32_t niter) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (!x || !init || n < 1 || niter < 0) return h->fail(HIVF_EINVAL, "bad train arguments");
    if (n > INT32_MAX) return h->fail(HIVF_EINVAL, "too many training rows");
    HIP_TRY(h, hipSetDevice(h->device));
    const int64_t d = h->d, nlist = h->nlist;
    std::vector<float> cent(init, init + nlist * d);
    int rc = upload_centroids(h, cent.data());
    if (rc) return rc;
    // training rows on the device (padded); the means are taken over them
    DevBuf<float4> dx;
    HIP_TRY(h, dx.ensure((size_t)n * h->nquad));
    if (h->d_pad != d) HIP_TRY(h, hipMemset(dx.p, 0, (size_t)n * h->nquad * sizeof(float4)));
    HIP_TRY(h, hipMemcpy2D(dx.p, h->d_pad * sizeof(float), x, d * sizeof(float), d * sizeof(float), n,
                           hipMemcpyHostToDevice));  // training is never captured
    // the assignment of the IP metric compares unit-norm rows
    // (src/Clustering.cpp:125-131)
    DevBuf<float4> dxn;
    const float4* xa = dx.p;
    if (h->metric == HIVF_METRIC_IP) {
        HIP_TRY(h, dxn.ensure((size_t)n * h->nquad));
        rc = normalize_rows(h, dx.p, n, dxn.p);
        if (rc) return rc;
        xa = dxn.p;
    }
    DevBuf<float> dcent;
    for (int it = 0; it < niter; ++it) {
        rc = gpu_lloyd_round(h, xa, reinterpret_cast<const float*>(dx.p), h->d_pad, n, cent.data(), dcent);
        if (rc) return rc;
        rc = upload_centroids(h, cent.data());
        if (rc) return rc;
    }
    if (h->metric == HIVF_METRIC_IP) {  // src/Index.cpp:409-421
        for (int64_t c = 0; c < nlist; ++c) {
            float s = 0.0f;
            for (int64_t j = 0; j < d; ++j) s = std::fma(cent[c * d + j], cent[c * d + j], s);
            if (s != 0.0f) {
                float r = std::sqrt(s);
                for (int64_t j = 0; j < d; ++j) cent[c * d + j] /= r;
            }
        }
        rc = upload_centroids(h, cent.data());
        if (rc) return rc;
    }
    return HIVF_OK;
    HIVF_API_END(h)
}

int hivf_kmeans_sample(int64_t n, int64_t nlist, int32_t max_points_per_centroid, uint64_t seed,
                       int64_t* sample_rows, int64_t* n_sample, int64_t* init_rows) {
    HIVF_API_BEGIN
    if (n < 1 || nlist < 1 || nlist > n || max_points_per_centroid < 1 || !sample_rows || !n_sample || !init_rows) {
        set_global_error("hivf_kmeans_sample: bad arguments (need 1 <= nlist <= n)");
        return HIVF_EINVAL;
    }
    // Clustering::subsample_training_set (src/Clustering.cpp:48-68): above
    // nlist * max_points_per_centroid rows, the first that many of a
    // std::shuffle of 0..n-1 under default_random_engine(seed)
    const int64_t cap = nlist * (int64_t)max_points_per_centroid;
    const int seed_int = (int)seed;  // ClusteringParameters::seed is an int
    int64_t m = n;
    if (n > cap) {
        std::vector<size_t> idx(n);
        for (int64_t i = 0; i < n; ++i) idx[i] = (size_t)i;
        std::shuffle(idx.begin(), idx.end(), std::default_random_engine(seed_int));
        for (int64_t i = 0; i < cap; ++i) sample_rows[i] = (int64_t)idx[i];
        m = cap;
    } else {
        for (int64_t i = 0; i < n; ++i) sample_rows[i] = i;
    }
    // Clustering::initialize_centroids (src/Clustering.cpp:72-89): the first
    // nlist of a fresh shuffle of the sample's positions
    std::vector<size_t> pos(m);
    for (int64_t i = 0; i < m; ++i) pos[i] = (size_t)i;
    std::shuffle(pos.begin(), pos.end(), std::default_random_engine(seed_int));
    for (int64_t c = 0; c < nlist; ++c) init_rows[c] = (int64_t)pos[c];
    *n_sample = m;
    return HIVF_OK;
    HIVF_API_END(nullptr)
}

int hivf_index_train_kmeans(hivf_index* h, int64_t n, const float* x, int32_t niter, uint64_t seed,
                            int32_t max_points_per_centroid) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (!x || n < h->nlist || niter < 0) return h->fail(HIVF_EINVAL, "bad train arguments (need n >= nlist)");
    if (max_points_per_centroid <= 0) max_points_per_centroid = 256;
    const int64_t cap = h->nlist * (int64_t)max_points_per_centroid;
    const int64_t d = h->d;
    std::vector<int64_t> rows((size_t)std::min(n, cap)), init_rows(h->nlist);
    int64_t m = 0;
    int rc = hivf_kmeans_sample(n, h->nlist, max_points_per_centroid, seed, rows.data(), &m, init_rows.data());
    if (rc) return h->fail(rc, hivf_last_error(nullptr));
    std::vector<float> sample, init((size_t)h->nlist * d);
    const float* xs = x;
    if (m < n) {
        sample.resize((size_t)m * d);
        for (int64_t i = 0; i < m; ++i) std::memcpy(&sample[i * d], x + rows[i] * d, sizeof(float) * d);
        xs = sample.data();
    }
    for (int64_t c = 0; c < h->nlist; ++c) std::memcpy(&init[c * d], xs + init_rows[c] * d, sizeof(float) * d);
    return hivf_index_train(h, m, xs, init.data(), niter);
    HIVF_API_END(h)
}

int hivf_index_add(hivf_index* h, int64_t n, const float* x) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (n < 0 || (n > 0 && !x)) return h->fail(HIVF_EINVAL, "bad add arguments");
    if (!h->have_centroids) return h->fail(HIVF_ESTATE, "add before train / set_centroids");
    if (n >= (int64_t)kNoId || n > INT32_MAX) return h->fail(HIVF_EINVAL, "too many rows");
    HIP_TRY(h, hipSetDevice(h->device));
    if (n > 0) {
        int rc = stage_queries(h, n, x, hipMemcpyHostToDevice, h->stream);
        if (rc) return rc;
    }
    return gpu_add(h, n);
    HIVF_API_END(h)
}

int hivf_index_set_lists(hivf_index* h, const int64_t* list_off, const int64_t* ids, const float* codes) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (!list_off) return h->fail(HIVF_EINVAL, "null list_off");
    std::vector<int64_t> off(list_off, list_off + h->nlist + 1);
    if (off[0] != 0) return h->fail(HIVF_EINVAL, "list_off[0] must be 0");
    for (int64_t l = 0; l < h->nlist; ++l)
        if (off[l + 1] < off[l]) return h->fail(HIVF_EINVAL, "list_off must be non-decreasing");
    if (off[h->nlist] > 0 && (!ids || !codes)) return h->fail(HIVF_EINVAL, "null ids or codes");
    HIP_TRY(h, hipSetDevice(h->device));
    const int64_t d = h->d;
    return upload_lists(h, off, [&](int64_t l, int64_t r, int64_t* id) {
        int64_t g = off[l] + r;
        *id = ids[g];
        return codes + g * d;
    });
    HIVF_API_END(h)
}

int64_t hivf_index_ntotal(const hivf_index* h) { return h ? h->ntotal : 0; }

int hivf_index_list_size(const hivf_index* h, int64_t list, int64_t* size) {
    HIVF_API_BEGIN
    if (!h || !size || list < 0 || list >= h->nlist) return HIVF_EINVAL;
    *size = h->have_lists ? h->list_off[list + 1] - h->list_off[list] : 0;
    return HIVF_OK;
    HIVF_API_END(const_cast<hivf_index*>(h))
}

int hivf_index_get_list(const hivf_index* h_, int64_t list, int64_t* ids, float* codes) {
    HIVF_API_BEGIN
    hivf_index* h = const_cast<hivf_index*>(h_);
    if (!h || list < 0 || list >= h->nlist) return HIVF_EINVAL;
    if (!h->have_lists) return h->fail(HIVF_ESTATE, "no lists");
    HIP_TRY(h, hipSetDevice(h->device));
    int64_t n_l = h->list_off[list + 1] - h->list_off[list];
    int64_t t0 = h->tile_off[list], nt = h->tile_off[list + 1] - t0;
    if (nt == 0) return HIVF_OK;
    std::vector<float> t((size_t)nt * h->nquad * kTileRows * 4);
    std::vector<uint32_t> id((size_t)nt * kTileRows);
    HIP_TRY(h, hipMemcpy(t.data(), h->d_codes.p + (size_t)t0 * h->nquad * kTileRows, t.size() * sizeof(float),
                         hipMemcpyDeviceToHost));
    HIP_TRY(h, hipMemcpy(id.data(), h->d_ids.p + (size_t)t0 * kTileRows, id.size() * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < n_l; ++r) {
        int64_t tt = r / kTileRows, lane = r % kTileRows;
        if (ids) ids[r] = id[r];
        if (codes)
            for (int64_t c = 0; c < h->d; ++c)
                codes[r * h->d + c] = t[((tt * h->nquad + c / 4) * kTileRows + lane) * 4 + c % 4];
    }
    return HIVF_OK;
    HIVF_API_END(const_cast<hivf_index*>(h_))
}

int hivf_index_search(hivf_index* h, int64_t n, const float* x, int64_t k, float* D, int64_t* I,
                      const hivf_search_params* params, hivf_search_stats* stats) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (n == 0) return HIVF_OK;  // src/Index.cpp:1639-1641
    if (n < 0 || !x || !D || !I) return h->fail(HIVF_EINVAL, "bad search arguments");
    if (n > INT32_MAX) return h->fail(HIVF_EINVAL, "too many queries in one call");
    HIP_TRY(h, hipSetDevice(h->device));
    hivf_search_params p;
    hivf_search_params_default(&p);
    if (params) p = *params;
    int rc = check_params(h, k, p);
    if (rc) return rc;
    rc = stage_queries(h, n, x, hipMemcpyHostToDevice, h->stream);
    if (rc) return rc;
    HIP_TRY(h, h->d_D.ensure((size_t)n * k));
    HIP_TRY(h, h->d_I.ensure((size_t)n * k));
    rc = run_search(h, h->d_q.p, n, k, h->d_D.p, h->d_I.p, p, stats, h->stream);
    if (rc) return rc;
    HIP_TRY(h, hipMemcpyAsync(D, h->d_D.p, (size_t)n * k * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(I, h->d_I.p, (size_t)n * k * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return HIVF_OK;
    HIVF_API_END(h)
}

int hivf_index_search_device(hivf_index* h, int64_t n, const float* x_dev, int64_t k, float* D_dev,
                             int64_t* I_dev, const hivf_search_params* params, hivf_search_stats* stats,
                             void* stream) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (n == 0) return HIVF_OK;
    if (n < 0 || !x_dev || !D_dev || !I_dev) return h->fail(HIVF_EINVAL, "bad search arguments");
    if (n > INT32_MAX) return h->fail(HIVF_EINVAL, "too many queries in one call");
    HIP_TRY(h, hipSetDevice(h->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
    hivf_search_params p;
    hivf_search_params_default(&p);
    if (params) p = *params;
    int rc = check_params(h, k, p);
    if (rc) return rc;
    const bool staged = h->d_pad != h->d || (reinterpret_cast<uintptr_t>(x_dev) & 15u);
    if (!(p.use_graph && !stats)) {
        const float4* q = reinterpret_cast<const float4*>(x_dev);
        if (staged) {
            rc = stage_queries(h, n, x_dev, hipMemcpyDeviceToDevice, s);
            if (rc) return rc;
            q = h->d_q.p;
        }
        return run_search(h, q, n, k, D_dev, I_dev, p, stats, s);
    }
    // captured search: the key is everything the captured work depends on
    if (!h->have_centroids || !h->have_lists)
        return h->fail(HIVF_ESTATE, "search before centroids and lists are set");
    int64_t P;
    const int32_t* npq;
    rc = resolve_probes(h, n, p, s, &P, &npq);
    if (rc) return rc;
    hivf_search_params kp = p;
    kp.nprobe_per_query = nullptr;
    struct {
        int64_t n, k, P;
        const void *x, *D, *I, *npq;
    } shape{n, k, P, x_dev, D_dev, I_dev, npq};
    std::vector<char> key(sizeof(shape) + sizeof(kp));
    std::memcpy(key.data(), &shape, sizeof(shape));
    std::memcpy(key.data() + sizeof(shape), &kp, sizeof(kp));
    if (!(h->cap.valid && h->cap.key == key && h->cap.free_gen == device_free_generation())) {
        h->drop_graph();
        // eager run: allocates every buffer the capture will use, and is this
        // call's result
        const float4* q = reinterpret_cast<const float4*>(x_dev);
        if (staged) {
            rc = stage_queries(h, n, x_dev, hipMemcpyDeviceToDevice, s);
            if (rc) return rc;
            q = h->d_q.p;
        }
        rc = run_search_core(h, q, n, k, D_dev, I_dev, p, P, npq, nullptr, s);
        if (rc) return rc;
        HIP_TRY(h, hipStreamSynchronize(s));
        HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        if (staged) rc = stage_queries(h, n, x_dev, hipMemcpyDeviceToDevice, h->stream);
        if (!rc) rc = run_search_core(h, q, n, k, D_dev, I_dev, p, P, npq, nullptr, h->stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : h->hip_fail(e, "hipStreamEndCapture");
        }
        e = hipGraphInstantiate(&h->cap.exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(g);
            h->cap.exec = nullptr;
            return h->hip_fail(e, "hipGraphInstantiate");
        }
        h->cap.graph = g;
        h->cap.key = key;
        h->cap.free_gen = device_free_generation();
        h->cap.valid = true;
        return HIVF_OK;
    }
    HIP_TRY(h, hipGraphLaunch(h->cap.exec, s));
    return HIVF_OK;
    HIVF_API_END(h)
}

int hivf_index_coarse(hivf_index* h, int64_t n, const float* x, int64_t nprobe, float* dist, int64_t* lists) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (n == 0) return HIVF_OK;
    if (n < 0 || !x || !dist || !lists) return h->fail(HIVF_EINVAL, "bad coarse arguments");
    if (!h->have_centroids) return h->fail(HIVF_ESTATE, "no centroids");
    HIP_TRY(h, hipSetDevice(h->device));
    int64_t P = clamp_nprobe(h, nprobe);
    if (P > kMaxK) return h->fail(HIVF_EUNSUPPORTED, "nprobe must be <= 1024");
    int rc = stage_queries(h, n, x, hipMemcpyHostToDevice, h->stream);
    if (rc) return rc;
    rc = run_coarse(h, h->d_q.p, n, P, h->stream);
    if (rc) return rc;
    std::vector<int32_t> l((size_t)n * P);
    HIP_TRY(h, hipMemcpyAsync(dist, h->d_cd.p, (size_t)n * P * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(l.data(), h->d_probes.p, l.size() * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (size_t i = 0; i < l.size(); ++i) lists[i] = l[i];
    return HIVF_OK;
    HIVF_API_END(h)
}

static int upload_probes(hivf_index* h, int64_t n, const int64_t* probes, int64_t nprobe) {
    std::vector<int32_t> v((size_t)n * nprobe);
    for (size_t i = 0; i < v.size(); ++i) {
        if (probes[i] < -1 || probes[i] >= h->nlist) return h->fail(HIVF_EINVAL, "probe list id out of range");
        v[i] = (int32_t)probes[i];
    }
    HIP_TRY(h, h->d_probes.ensure(v.size()));
    HIP_TRY(h, hipMemcpy(h->d_probes.p, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    return HIVF_OK;
}

int hivf_index_warmup_tau(hivf_index* h, int64_t n, const float* x, const int64_t* probes, int64_t nprobe,
                          int64_t k, int64_t L, int64_t W, float* tau) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (n == 0) return HIVF_OK;
    if (n < 0 || !x || !probes || !tau || nprobe < 1 || L < 0 || W < 0)
        return h->fail(HIVF_EINVAL, "bad warm-up arguments");
    if (k < 1 || k > kMaxK) return h->fail(HIVF_EUNSUPPORTED, "k must lie in [1, 1024]");
    if (!h->have_lists) return h->fail(HIVF_ESTATE, "no lists");
    HIP_TRY(h, hipSetDevice(h->device));
    int rc = stage_queries(h, n, x, hipMemcpyHostToDevice, h->stream);
    if (rc) return rc;
    rc = upload_probes(h, n, probes, nprobe);
    if (rc) return rc;
    HIP_TRY(h, h->d_tau.ensure(n));
    ScanArgs w = base_scan_args(h, h->d_q.p, n, nprobe, k);
    w.row_limit = (int)std::min<int64_t>(W, INT_MAX);
    w.warm_lists = (int)std::min<int64_t>(L, INT_MAX);
    w.out_tau = h->d_tau.p;
    HIP_TRY(h, launch_scan(w, kScanWarmup, h->metric, h->stream));
    HIP_TRY(h, hipMemcpyAsync(tau, h->d_tau.p, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return HIVF_OK;
    HIVF_API_END(h)
}

int hivf_index_scan(hivf_index* h, int64_t n, const float* x, const int64_t* probes, int64_t nprobe, int64_t k,
                    const float* tau, const hivf_search_params* params, float* D, int64_t* I,
                    hivf_search_stats* stats) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (n == 0) return HIVF_OK;
    if (n < 0 || !x || !probes || !D || !I || nprobe < 1) return h->fail(HIVF_EINVAL, "bad scan arguments");
    if (!h->have_lists) return h->fail(HIVF_ESTATE, "no lists");
    HIP_TRY(h, hipSetDevice(h->device));
    hivf_search_params p;
    hivf_search_params_default(&p);
    if (params) p = *params;
    int rc = check_params(h, k, p);
    if (rc) return rc;
    if (p.algo == HIVF_ALGO_BATCHED) return h->fail(HIVF_EUNSUPPORTED, "scan alone: plain or pruned");
    rc = stage_queries(h, n, x, hipMemcpyHostToDevice, h->stream);
    if (rc) return rc;
    rc = upload_probes(h, n, probes, nprobe);
    if (rc) return rc;
    HIP_TRY(h, h->d_D.ensure((size_t)n * k));
    HIP_TRY(h, h->d_I.ensure((size_t)n * k));
    if (tau) {
        HIP_TRY(h, h->d_tau.ensure(n));
        HIP_TRY(h, hipMemcpy(h->d_tau.p, tau, n * sizeof(float), hipMemcpyHostToDevice));
    }
    ScanArgs a = base_scan_args(h, h->d_q.p, n, nprobe, k);
    a.out_d = h->d_D.p;
    a.out_i = h->d_I.p;
    a.tau = tau ? h->d_tau.p : nullptr;
    set_blocks(h, p, &a);
    if (stats) {
        HIP_TRY(h, h->d_stats.ensure(8));
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, h->stream, h->d_stats.p, 8, 0ull);
        a.stats = h->d_stats.p;
        HIP_TRY(h, hipEventRecord(h->ev[2], h->stream));
    }
    HIP_TRY(h, launch_scan(a, p.algo == HIVF_ALGO_PRUNED ? kScanPruned : kScanPlain, h->metric, h->stream));
    if (stats) HIP_TRY(h, hipEventRecord(h->ev[3], h->stream));
    HIP_TRY(h, hipMemcpyAsync(D, h->d_D.p, (size_t)n * k * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(I, h->d_I.p, (size_t)n * k * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    if (stats) {
        unsigned long long c[8];
        HIP_TRY(h, hipMemcpy(c, h->d_stats.p, sizeof(c), hipMemcpyDeviceToHost));
        stats->scan_ms = elapsed(h->ev[2], h->ev[3]);
        stats->total_ms = stats->scan_ms;
        stats->candidates = (int64_t)c[0];
        stats->blocks_total = (int64_t)c[1];
        stats->blocks_skipped = (int64_t)c[2];
        stats->scan_launches = 1;
    }
    return HIVF_OK;
    HIVF_API_END(h)
}

}  // extern "C"

// ---- persistence: the reference's index file format (harmony_format.cpp) ----
extern "C" {

int hivf_index_info(const hivf_index* h, int64_t* d, int64_t* nlist, int* metric) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (d) *d = h->d;
    if (nlist) *nlist = h->nlist;
    if (metric) *metric = h->metric;
    return HIVF_OK;
    HIVF_API_END(const_cast<hivf_index*>(h))
}

int hivf_index_save_harmony(const hivf_index* h_, const char* path) {
    HIVF_API_BEGIN
    hivf_index* h = const_cast<hivf_index*>(h_);
    if (!h || !path) return HIVF_EINVAL;
    if (!h->have_centroids) return h->fail(HIVF_ESTATE, "nothing to save");
    HarmonyFile f;
    f.d = (uint64_t)h->d;
    f.nlist = (uint64_t)h->nlist;
    f.metric = h->metric;
    f.centroids = h->cent;
    f.list_off.assign(h->nlist + 1, 0);
    if (h->have_lists) f.list_off = h->list_off;
    f.ids.resize(f.list_off[h->nlist]);
    f.codes.resize((size_t)f.list_off[h->nlist] * h->d);
    for (int64_t l = 0; h->have_lists && l < h->nlist; ++l) {
        if (f.list_off[l + 1] == f.list_off[l]) continue;
        int rc = hivf_index_get_list(h, l, f.ids.data() + f.list_off[l], f.codes.data() + f.list_off[l] * h->d);
        if (rc) return rc;
    }
    std::string m;
    int rc = write_harmony_file(path, f, &m);
    return rc ? h->fail(rc, m) : HIVF_OK;
    HIVF_API_END(const_cast<hivf_index*>(h_))
}

hivf_index* hivf_index_load_harmony(const char* path, int* status) {
    HIVF_API_BEGIN
    auto failed = [&](int code, const std::string& m) -> hivf_index* {
        set_global_error("hivf_index_load_harmony: " + m);
        if (status) *status = code;
        return nullptr;
    };
    if (status) *status = HIVF_OK;
    HarmonyFile f;
    std::string m;
    int rc = read_harmony_file(path, &f, &m);
    if (rc) return failed(rc, m);
    hivf_index* h = hivf_index_create((int64_t)f.d, (int64_t)f.nlist, f.metric);
    if (!h) return failed(HIVF_EHIP, "cannot create index");
    rc = hivf_index_set_centroids(h, f.centroids.data());
    if (rc == HIVF_OK) rc = hivf_index_set_lists(h, f.list_off.data(), f.ids.data(), f.codes.data());
    if (rc != HIVF_OK) {
        std::string e = h->err;
        hivf_index_destroy(h);
        return failed(rc, e);
    }
    return h;
    HIVF_API_END_PTR(status)
}

}  // extern "C"

// ---- tile-shape autotuner ------------------------------------------------
extern "C" {

int hivf_index_set_tune(hivf_index* h, int32_t bq, int32_t br, int32_t bw) {
    HIVF_API_BEGIN
    if (!h || bq < 0 || br < 0 || bw < 0) return HIVF_EINVAL;
    h->tune.qg = bq / 32;
    h->tune.item_tiles = br / 32;
    h->tune.waves_per_cu = bw;
    h->drop_graph();
    return HIVF_OK;
    HIVF_API_END(h)
}

int hivf_index_autotune(hivf_index* h, int64_t n, const float* x_dev, int64_t k, const hivf_search_params* base,
                        int32_t reps, hivf_tune_result* out) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (n < 1 || !x_dev || !out) return h->fail(HIVF_EINVAL, "bad autotune arguments");
    HIP_TRY(h, hipSetDevice(h->device));
    hivf_search_params p;
    hivf_search_params_default(&p);
    if (base) p = *base;
    p.algo = HIVF_ALGO_BATCHED;
    p.use_graph = 0;
    if (reps < 1) reps = 3;
    DevBuf<float> D;
    DevBuf<int64_t> I;
    HIP_TRY(h, D.ensure((size_t)n * k));
    HIP_TRY(h, I.ensure((size_t)n * k));
    const int64_t dpad16 = (h->d + 15) / 16 * 16;
    int qg_max = dpad16 <= 128 ? 4 : dpad16 <= 256 ? 2 : 1;
    auto time_one = [&](int bq, int br, int bw, double* ms) -> int {
        hivf_search_params t = p;
        t.batch_queries = bq;
        t.batch_rows = br;
        t.batch_waves = bw;
        int rc = hivf_index_search_device(h, n, x_dev, k, D.p, I.p, &t, nullptr, h->stream);  // warm-up
        if (rc) return rc;
        HIP_TRY(h, hipEventRecord(h->ev[6], h->stream));
        for (int r = 0; r < reps; ++r) {
            rc = hivf_index_search_device(h, n, x_dev, k, D.p, I.p, &t, nullptr, h->stream);
            if (rc) return rc;
        }
        HIP_TRY(h, hipEventRecord(h->ev[7], h->stream));
        HIP_TRY(h, hipEventSynchronize(h->ev[7]));
        *ms = elapsed_ms(h->ev[6], h->ev[7]) / reps;
        return HIVF_OK;
    };
    const auto saved = h->tune;
    h->tune = {};
    double dflt = 0.0;
    int rc = time_one(0, 0, 0, &dflt);
    if (rc) {
        h->tune = saved;
        return rc;
    }
    // integer data runs the int8 scan, whose items hold 32 queries, or 64
    // when d <= 128
    if (h->batched.valid && h->batched.i8) qg_max = h->batched.nkc8 == 4 ? 2 : 1;
    double best = 1e30;
    int bq = 0, br = 0, bw = 0, tried = 0;
    for (int q = 32; q <= 32 * qg_max; q *= 2)
        for (int rows : {256, 512, 1024, 2048})
            for (int w : {8, 16}) {
                double ms = 0.0;
                rc = time_one(q, rows, w, &ms);
                if (rc) {
                    h->tune = saved;
                    return rc;
                }
                ++tried;
                if (ms < best) {
                    best = ms;
                    bq = q;
                    br = rows;
                    bw = w;
                }
            }
    h->tune.qg = bq / 32;
    h->tune.item_tiles = br / 32;
    h->tune.waves_per_cu = bw;
    h->drop_graph();  // a captured search holds the old default shape
    std::memset(out, 0, sizeof(*out));
    out->batch_queries = bq;
    out->batch_rows = br;
    out->batch_waves = bw;
    out->tried = tried;
    out->ms = best;
    out->ms_default = dflt;
    return HIVF_OK;
    HIVF_API_END(h)
}

}  // extern "C"
