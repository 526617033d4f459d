// Search entry point that allocates no device memory
// (hivf_index_search_device_noalloc, include/hivf.h).
//
// The per-query search (coarse quantizer, optional warm-up bound, plain or
// pruned scan) needs scratch beyond the caller's inputs and outputs: the
// probe lists with their coarse distances, the warm-up bounds, and a padded
// copy of the queries when d is not a multiple of 4 or the query buffer is
// not 16-byte aligned.  Here that scratch is one file-scope device array of
// fixed size, so a call performs no hipMalloc / hipFree.  Its address is
// taken with hipGetSymbolAddress once per device, with that device current,
// and every kernel receives that address; the array's name is never used
// from host code.
//
// Every call on a device shares that device's array and leaves its last
// contents there: calls that overlap on one device (other streams, other
// threads) race.  Each call writes every scratch element it reads, so no
// clearing is needed.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>
#include <mutex>

#include "index_impl.hpp"

namespace hivf {

namespace {

// Limits of one call (n = queries, P = probes after clamping to nlist).
constexpr int64_t kNaMaxPairs = int64_t(1) << 20;       // n * P
constexpr int64_t kNaMaxQueries = int64_t(1) << 16;     // n
constexpr int64_t kNaMaxQueryFloats = int64_t(1) << 21; // n * d padded to a multiple of 4

// Carve-up of the array, in bytes: probe distances, probe lists, bounds,
// padded queries.
constexpr size_t kNaOffDist = 0;
constexpr size_t kNaOffList = kNaOffDist + sizeof(float) * kNaMaxPairs;
constexpr size_t kNaOffTau = kNaOffList + sizeof(int32_t) * kNaMaxPairs;
constexpr size_t kNaOffQ = kNaOffTau + sizeof(float) * kNaMaxQueries;
constexpr size_t kNaBytes = kNaOffQ + sizeof(float) * kNaMaxQueryFloats;  // 16.25 MiB

__device__ __align__(256) unsigned char scratch[kNaBytes];

constexpr int kMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kMaxDevices] = {};

__global__ void na_pad_rows_kernel(const float* __restrict__ x, float* __restrict__ y, int64_t n, int d,
                                   int d_pad) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * d_pad) return;
    int64_t r = i / d_pad;
    int c = (int)(i % d_pad);
    y[i] = c < d ? x[r * d + c] : 0.0f;
}

}  // namespace

// Address of this device's scratch array (the current device must be
// `device`); taken once per device.
hipError_t noalloc_setup(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_na_ptr[device] = p;
    }
    if (out) *out = g_na_ptr[device];
    return hipSuccess;
}

}  // namespace hivf

using namespace hivf;

extern "C" {

int hivf_index_search_device_noalloc(hivf_index* h, int64_t n, const float* x_dev, int64_t k, float* D_dev,
                                     int64_t* I_dev, const hivf_search_params* params, void* stream) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (n == 0) return HIVF_OK;
    if (n < 0 || !x_dev || !D_dev || !I_dev) return h->fail(HIVF_EINVAL, "bad search arguments");
    if (!h->have_centroids || !h->have_lists)
        return h->fail(HIVF_ESTATE, "search before centroids and lists are set");
    hivf_search_params p;
    hivf_search_params_default(&p);
    if (params) p = *params;
    if (p.algo != HIVF_ALGO_PLAIN && p.algo != HIVF_ALGO_PRUNED)
        return h->fail(HIVF_EUNSUPPORTED, "noalloc search: plain or pruned scan only");
    if (p.precision != HIVF_PREC_FP32) return h->fail(HIVF_EUNSUPPORTED, "noalloc search: FP32 only");
    if (p.nprobe_per_query)
        return h->fail(HIVF_EUNSUPPORTED, "noalloc search: per-query nprobe from device memory only");
    if (p.nprobe_per_query_dev && p.nprobe <= 0)
        return h->fail(HIVF_EINVAL, "device per-query nprobe needs nprobe as their bound");
    if (p.tau_source == HIVF_TAU_LEARNED) return h->fail(HIVF_EUNSUPPORTED, "noalloc search: no learned bound");
    if (p.tau_source < HIVF_TAU_NONE || p.tau_source > HIVF_TAU_LEARNED)
        return h->fail(HIVF_EINVAL, "unknown tau_source");
    if (p.algo == HIVF_ALGO_PRUNED && h->metric != HIVF_METRIC_L2)
        return h->fail(HIVF_EUNSUPPORTED, "dimension-block pruning is defined for L2 only");
    if (k < 1 || k > kMaxK) return h->fail(HIVF_EUNSUPPORTED, "k must lie in [1, 1024]");
    if (p.dim_block < 0) return h->fail(HIVF_EINVAL, "dim_block must be >= 0");
    int64_t P = p.nprobe > 0 ? p.nprobe : h->nprobe;
    if (P <= 0 || P > h->nlist) P = h->nlist;  // src/Index.cpp:1646-1648
    if (P > kMaxK) return h->fail(HIVF_EUNSUPPORTED, "nprobe must be <= 1024");
    if (n > kNaMaxQueries || n * P > kNaMaxPairs)
        return h->fail(HIVF_EINVAL, "noalloc search: at most 65536 queries and 2^20 (query, probe) pairs per call");
    const bool staged = h->d_pad != h->d || (reinterpret_cast<uintptr_t>(x_dev) & 15u);
    if (staged && n * h->d_pad > kNaMaxQueryFloats)
        return h->fail(HIVF_EINVAL, "noalloc search: at most 2^21 padded query floats per call");
    HIP_TRY(h, hipSetDevice(h->device));
    void* base = nullptr;
    HIP_TRY(h, noalloc_setup(h->device, &base));
    char* sb = static_cast<char*>(base);
    float* cd = reinterpret_cast<float*>(sb + kNaOffDist);
    int32_t* probes = reinterpret_cast<int32_t*>(sb + kNaOffList);
    float* tau = reinterpret_cast<float*>(sb + kNaOffTau);
    const float4* q = reinterpret_cast<const float4*>(x_dev);
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
    if (staged) {
        float* qp = reinterpret_cast<float*>(sb + kNaOffQ);
        const int64_t total = n * h->d_pad;
        hipLaunchKernelGGL(na_pad_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, x_dev, qp,
                           n, (int)h->d, (int)h->d_pad);
        HIP_TRY(h, hipGetLastError());
        q = reinterpret_cast<const float4*>(qp);
    }
    // coarse quantizer: the per-query kernel, which keeps its candidates in LDS
    CoarseArgs c;
    c.q = q;
    c.nq = (int)n;
    c.nquad = h->nquad;
    c.cent = h->d_cent.p;
    c.nlist = (int)h->nlist;
    c.P = (int)P;
    c.out_d = cd;
    c.out_l = probes;
    HIP_TRY(h, launch_coarse(c, h->metric, s));
    ScanArgs a{};
    a.lists = h->lists();
    a.q = q;
    a.nq = (int)n;
    a.probes = probes;
    a.probe_stride = (int)P;
    a.nprobe_q = p.nprobe_per_query_dev;
    a.nprobe = (int)P;
    a.k = (int)k;
    a.block_dims = (int)h->d;
    a.nblocks = 1;
    a.row_limit = INT_MAX;
    a.warm_lists = INT_MAX;
    a.out_d = D_dev;
    a.out_i = I_dev;
    if (p.algo == HIVF_ALGO_PRUNED && p.tau_source == HIVF_TAU_WARMUP) {
        ScanArgs w = a;
        w.row_limit = p.warmup_list_size > 0 ? p.warmup_list_size : (int)k;  // src/query.cpp:847
        w.warm_lists = p.warmup_lists > 0 ? p.warmup_lists : 10;             // src/query.cpp:845
        w.out_d = nullptr;
        w.out_i = nullptr;
        w.out_tau = tau;
        HIP_TRY(h, launch_scan(w, kScanWarmup, h->metric, s));
        a.tau = tau;
    }
    if (p.algo == HIVF_ALGO_PRUNED) pruning_blocks(h->d, p.dim_block, &a.block_dims, &a.nblocks);
    HIP_TRY(h, launch_scan(a, p.algo == HIVF_ALGO_PRUNED ? kScanPruned : kScanPlain, h->metric, s));
    return HIVF_OK;
    HIVF_API_END(h)
}

}  // extern "C"
