// Per-index state of the list-major batched scan.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hivf {

struct BatchedState {
    bool valid = false;  // derived arena (BF16 shadow etc.) matches the lists
    void invalidate() { valid = false; }
};

}  // namespace hivf
