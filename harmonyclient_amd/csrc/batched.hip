// List-major batched scan (placeholder until the MFMA path lands).
#include "index_impl.hpp"

namespace hivf {

int run_batched(hivf_index* h, const float4*, int64_t, int64_t, int, const int32_t*,
                const hivf_search_params&, float*, int64_t*, hivf_search_stats*, hipStream_t, bool) {
    return h->fail(HIVF_EUNSUPPORTED, "batched scan not built yet");
}

}  // namespace hivf
