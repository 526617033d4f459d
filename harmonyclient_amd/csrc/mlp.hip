// Learned pruning bound (placeholder until the model lands).
#include "index_impl.hpp"

namespace hivf {

int run_tau_model(hivf_index* h, const float4*, int64_t, int, const int32_t*, float*, hipStream_t) {
    return h->fail(HIVF_EUNSUPPORTED, "no bound model attached");
}

}  // namespace hivf
