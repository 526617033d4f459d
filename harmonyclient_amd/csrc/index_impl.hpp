// Internal state behind the opaque hivf_index handle (include/hivf.h).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include <memory>

#include "batched.hpp"
#include "tau_model.hpp"
#include "common.hpp"
#include "hivf.h"
#include "kernels.hpp"

namespace hivf {

// Every device buffer the library frees goes through dev_free, which counts
// the frees process-wide.  A captured search (hipGraph) records the count
// at capture and is replayed only while it is unchanged: a freed buffer may
// have been one the graph's kernels read or write.
void note_device_free();
uint64_t device_free_generation();
inline void dev_free(void* p) {
    if (!p) return;
    (void)hipFree(p);
    note_device_free();
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        dev_free(p);
        p = nullptr;
        n = 0;
    }
    // Grow to at least cnt elements (contents not kept).
    hipError_t ensure(size_t cnt) {
        if (cnt <= n && p) return hipSuccess;
        release();
        size_t want = std::max<size_t>(cnt, 1);
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        n = want;
        return hipSuccess;
    }
};

// State of the MFMA coarse quantizer: the 3-piece BF16 centroid shadow and
// per-search scratch.
struct CoarseState {
    bool valid = false;  // shadow matches the current centroids
    DevBuf<uint16_t> cbf;
    DevBuf<uint16_t> qh, ql;
    DevBuf<float> qn, rank;
    DevBuf<int32_t> hard;   // queries the one-wave pick hands to the workgroup pick
    DevBuf<uint32_t> nhard;
};

}  // namespace hivf

struct hivf_index {
    int64_t d = 0, d_pad = 0, nlist = 0;
    int nquad = 0;
    int metric = HIVF_METRIC_L2;
    int64_t nprobe = 1;
    int device = 0;
    int cus = 0;  // compute units of the device
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    std::string err;

    std::vector<float> cent;  // [nlist x d]
    bool have_centroids = false;
    bool have_lists = false;
    bool lists_uint8 = false;  // every list value is an integer in [0, 255]

    // list arena (device) and its host-side shape
    std::vector<int64_t> list_off;  // [nlist + 1] rows
    std::vector<int32_t> tile_off;  // [nlist + 1] tiles
    int64_t ntotal = 0;
    int64_t tiles = 0;
    hivf::DevBuf<float4> d_cent;      // centroid tiles
    hivf::DevBuf<float4> d_cent_rows; // centroids row-major [nlist x nquad], zero-padded
    hivf::DevBuf<float> d_cent_norm;  // [tiles * 64] FP64-computed squared norms (0 on padding)
    hivf::CoarseState coarse;
    hivf::DevBuf<float4> d_codes;
    hivf::DevBuf<uint32_t> d_ids;
    hivf::DevBuf<float> d_norms;      // [tiles * 64] FP64-computed squared norms
    hivf::DevBuf<int32_t> d_tile_off;
    hivf::DevBuf<int32_t> d_size;

    // search scratch
    hivf::DevBuf<float4> d_q;
    hivf::DevBuf<float> d_stage;  // unpadded host queries on their way to d_q
    hivf::DevBuf<float> d_cd;
    hivf::DevBuf<int32_t> d_probes;
    hivf::DevBuf<int32_t> d_npq;
    hivf::DevBuf<float> d_tau;
    hivf::DevBuf<float> d_D;
    hivf::DevBuf<int64_t> d_I;
    hivf::DevBuf<unsigned long long> d_stats;

    hivf::BatchedState batched;
    float* batched_qnorm = nullptr;  // query norms of the last batched search (its scratch)
    std::unique_ptr<hivf::TauModel> tau_model;  // learned bound attached to this index
    hivf::DevBuf<float> d_feat;                 // bound-model features
    // captured search (hivf_search_params.use_graph): one shape per handle
    struct {
        bool valid = false;
        std::vector<char> key;
        uint64_t free_gen = 0;  // device_free_generation() when captured
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    } cap;
    void drop_graph() {
        if (cap.exec) (void)hipGraphExecDestroy(cap.exec);
        if (cap.graph) (void)hipGraphDestroy(cap.graph);
        cap.exec = nullptr;
        cap.graph = nullptr;
        cap.valid = false;
    }
    // tile shape of the batched scan picked by the autotuner (0 = default)
    struct {
        int qg = 0;
        int item_tiles = 0;
        int waves_per_cu = 0;
    } tune;

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
    int hip_fail(hipError_t e, const char* where) {
        err = std::string(where) + ": " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? HIVF_ENOMEM : HIVF_EHIP;
    }
    hivf::DevLists lists() const {
        hivf::DevLists L;
        L.codes = d_codes.p;
        L.ids = d_ids.p;
        L.list_tile_off = d_tile_off.p;
        L.list_size = d_size.p;
        L.nquad = nquad;
        return L;
    }
};

// Every extern "C" body runs between HIVF_API_BEGIN and one of the
// HIVF_API_END forms: a C++ exception (a host allocation that fails, a size
// beyond a container's limit) becomes a status code and an error message
// instead of crossing the C ABI.
namespace hivf {
// Maps the exception in flight to a HIVF_E* code and records its message on
// h (NULL: the global error).
int api_exception(hivf_index* h, const char* fn) noexcept;
}  // namespace hivf
#define HIVF_API_BEGIN try {
#define HIVF_API_END(h)                              \
    }                                                \
    catch (...) {                                    \
        return hivf::api_exception((h), __func__);   \
    }
#define HIVF_API_END_PTR(status_ptr)                              \
    }                                                             \
    catch (...) {                                                 \
        int code_ = hivf::api_exception(nullptr, __func__);       \
        int* st_ = (status_ptr);                                  \
        if (st_) *st_ = code_;                                    \
        return nullptr;                                           \
    }
#define HIVF_API_END_VOID                         \
    }                                             \
    catch (...) {                                 \
        (void)hivf::api_exception(nullptr, __func__); \
    }

#define HIP_TRY(h, expr)                                       \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return (h)->hip_fail(e_, #expr); \
    } while (0)

namespace hivf {
inline float elapsed_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0f;
    return ms;
}
// Learned bound: tau[q] from the attached model (mlp.hip).
int run_tau_model(hivf_index* h, const float4* q, int64_t n, int P, const int32_t* npq, float* tau,
                  hipStream_t s);
// Defined in batched.hip.
int run_batched(hivf_index* h, const float4* q, int64_t n, int64_t k, int P, const int32_t* npq,
                const hivf_search_params& p, float* D, int64_t* I, hivf_search_stats* st,
                hipStream_t s, bool timing);
struct PrepArgs;  // prep.hpp
// MFMA coarse quantizer (coarse.hip); HIVF_EUNSUPPORTED when d > 512.  With
// prep, its query-preparation launch also runs the batched search's.
int run_coarse_mfma(hivf_index* h, const float4* q, int64_t n, int64_t P, hipStream_t s,
                    const PrepArgs* prep = nullptr);
// Top-P centroids of n staged queries (hivf_api.cpp): the MFMA form, or the
// per-query kernel for P = 1 and d > 512.
int run_coarse(hivf_index* h, const float4* q, int64_t n, int64_t P, hipStream_t s,
               const PrepArgs* prep = nullptr, bool* prep_done = nullptr);
// Index build on the GPU (build.hip).  One Lloyd round: assign the rows xa
// (padded device rows), then set every centroid with rows to the FP64 mean
// of its rows of xsum (row stride ldx floats) in row order; cent_host holds
// the centroids before and after.
int gpu_lloyd_round(hivf_index* h, const float4* xa, const float* xsum, int64_t ldx, int64_t n, float* cent_host,
                    DevBuf<float>& dcent);
// Unit-norm copies of n padded device rows (zero rows stay zero).
int normalize_rows(hivf_index* h, const float4* x, int64_t n, float4* y);
// Index::add of the n rows staged in d_q: lists, their order, the arena.
int gpu_add(hivf_index* h, int64_t n);
// Order of the int8 shadow: perm[arena position] = the arena row shown
// there (every list's rows by increasing radius, ties by arena row, padding
// in place) and the 32-row tiles' widened radius intervals in that order;
// radius[arena row] = distance to the list's centroid.
int shadow_order(hivf_index* h, const float* radius, int32_t* perm, float2* rb8, hipStream_t s);
// Address of the no-allocation search's scratch array on `device` (the
// current device), taken once per device (noalloc.hip).
hipError_t noalloc_setup(int device, void** out);
// Error text of calls without a handle (hivf_last_error(NULL)).
void set_global_error(const std::string& m);
}  // namespace hivf
