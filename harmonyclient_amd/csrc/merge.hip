// Top-k merge of per-partition results: the final step of a list-sharded
// search, where each partition (a GPU, or a worker in the reference's
// distributed modes) returns its own sorted top-k of every query and the
// merged row is the best k of their union (src/Index.cpp:1327-1343, which
// offers each partition's entries to a k-heap and sorts it).
//
// One 256-thread workgroup per query streams the parts x k_in entries of its
// row through the LDS top-k (threshold filter + bitonic compaction), so the
// result is the exact top-k in the (distance, id) order; entries with a
// negative id (a partial result's sentinel padding) are skipped.
#include <climits>
#include <cstring>
#include <string>

#include "index_impl.hpp"
#include "select.hpp"

namespace hivf {
namespace {

constexpr int kMergeNT = 256;

struct MergeArgs {
    const float* Dp;   // [parts][n][k_in]
    const int64_t* Ip;
    int64_t n;
    int parts;
    int k_in;
    int k;
    int metric;
    float* D;          // [n][k]
    int64_t* I;
};

__global__ __launch_bounds__(kMergeNT) void merge_kernel(MergeArgs a) {
    __shared__ LdsTopK<kMergeNT, 2048> tk;
    const int64_t q = blockIdx.x;
    tk.init();
    const int64_t total = (int64_t)a.parts * a.k_in;
    for (int64_t base = 0; base < total; base += kMergeNT) {
        const uint64_t thr = tk.thr;
        const int64_t e = base + threadIdx.x;
        if (e < total) {
            const int64_t p = e / a.k_in, i = e % a.k_in;
            const size_t off = ((size_t)p * a.n + q) * a.k_in + i;
            const int64_t id = a.Ip[off];
            const float v = a.Dp[off];
            const float r = a.metric == kL2 ? v : -v;
            // r < FLT_MAX: the heap's strict admission against its sentinel
            // (never inf or NaN)
            if (id >= 0 && r < FLT_MAX) tk.offer(make_key(r, (uint32_t)id), thr);
        }
        __syncthreads();
        tk.maybe_compact(a.k, kMergeNT);
    }
    tk.finish(a.k);
    const int have = min(tk.cnt, a.k);
    for (int i = threadIdx.x; i < a.k; i += kMergeNT) {
        float dv;
        int64_t iv;
        if (i < have) {
            const uint64_t key = tk.buf[i];
            const float r = key_rank(key);
            dv = dist_of_rank(a.metric, r);
            iv = (int64_t)(uint32_t)key;
        } else {
            dv = a.metric == kL2 ? FLT_MAX : -FLT_MAX;
            iv = -1;
        }
        a.D[(size_t)q * a.k + i] = dv;
        a.I[(size_t)q * a.k + i] = iv;
    }
}

int check_merge(int metric, int64_t n, int64_t parts, int64_t k_in, int64_t k) {
    if (metric != HIVF_METRIC_L2 && metric != HIVF_METRIC_IP) return HIVF_EINVAL;
    if (n < 0 || parts < 1 || k_in < 1 || k < 1) return HIVF_EINVAL;
    if (k > kMaxK) return HIVF_EUNSUPPORTED;
    if (n > INT32_MAX || parts * k_in > INT32_MAX) return HIVF_EUNSUPPORTED;
    return HIVF_OK;
}

hipError_t launch_merge(int metric, int64_t n, int64_t parts, int64_t k_in, const float* Dp, const int64_t* Ip,
                        int64_t k, float* D, int64_t* I, hipStream_t s) {
    if (n == 0) return hipSuccess;
    MergeArgs a{Dp, Ip, n, (int)parts, (int)k_in, (int)k, metric, D, I};
    hipLaunchKernelGGL(merge_kernel, dim3((unsigned)n), dim3(kMergeNT), 0, s, a);
    return hipGetLastError();
}

}  // namespace
}  // namespace hivf

using namespace hivf;

extern "C" {

int hivf_merge_topk_device(int metric, int64_t n, int64_t parts, int64_t k_in, const float* D_parts,
                           const int64_t* I_parts, int64_t k, float* D, int64_t* I, void* stream) {
    HIVF_API_BEGIN
    int rc = check_merge(metric, n, parts, k_in, k);
    if (rc) {
        set_global_error("hivf_merge_topk_device: bad metric, n, parts, k_in or k (k <= 1024)");
        return rc;
    }
    if (n == 0) return HIVF_OK;
    if (!D_parts || !I_parts || !D || !I) {
        set_global_error("hivf_merge_topk_device: null buffer");
        return HIVF_EINVAL;
    }
    hipError_t e = launch_merge(metric, n, parts, k_in, D_parts, I_parts, k, D, I, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        set_global_error(std::string("hivf_merge_topk_device: ") + hipGetErrorString(e));
        return HIVF_EHIP;
    }
    return HIVF_OK;
    HIVF_API_END(nullptr)
}

int hivf_merge_topk(int metric, int64_t n, int64_t parts, int64_t k_in, const float* D_parts,
                    const int64_t* I_parts, int64_t k, float* D, int64_t* I) {
    HIVF_API_BEGIN
    int rc = check_merge(metric, n, parts, k_in, k);
    if (rc) {
        set_global_error("hivf_merge_topk: bad metric, n, parts, k_in or k (k <= 1024)");
        return rc;
    }
    if (n == 0) return HIVF_OK;
    if (!D_parts || !I_parts || !D || !I) {
        set_global_error("hivf_merge_topk: null buffer");
        return HIVF_EINVAL;
    }
    const size_t in_el = (size_t)parts * n * k_in, out_el = (size_t)n * k;
    for (size_t i = 0; i < in_el; ++i)
        if (I_parts[i] >= (int64_t)kNoId) {
            set_global_error("hivf_merge_topk: ids must be below 2^32 - 1");
            return HIVF_EUNSUPPORTED;
        }
    DevBuf<float> dp, d;
    DevBuf<int64_t> ip, i;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = dp.ensure(in_el);
    if (e == hipSuccess) e = ip.ensure(in_el);
    if (e == hipSuccess) e = d.ensure(out_el);
    if (e == hipSuccess) e = i.ensure(out_el);
    if (e == hipSuccess) e = hipMemcpyAsync(dp.p, D_parts, in_el * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ip.p, I_parts, in_el * sizeof(int64_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_merge(metric, n, parts, k_in, dp.p, ip.p, k, d.p, i.p, s);
    if (e == hipSuccess) e = hipMemcpyAsync(D, d.p, out_el * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(I, i.p, out_el * sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (s) (void)hipStreamDestroy(s);
    if (e != hipSuccess) {
        set_global_error(std::string("hivf_merge_topk: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? HIVF_ENOMEM : HIVF_EHIP;
    }
    return HIVF_OK;
    HIVF_API_END(nullptr)
}

}  // extern "C"
