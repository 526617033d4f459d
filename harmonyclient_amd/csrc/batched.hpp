// Per-index state of the list-major batched scan (device buffers derived
// from the list arena, plus per-batch scratch).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hivf {

template <class T>
struct DevBuf;

// One unit of work of the batched scan: queries lq[q0 .. q0+nq) of list
// `list` against its 32-row tiles [t0, t0 + kItemTiles).
struct BatchItem {
    int32_t list;
    int32_t q0;
    int32_t nq;
    int32_t t0;
};

struct BatchedState {
    bool valid = false;     // bf16 arena matches the current lists
    int nkc = 0;            // 16-dim chunks of the padded dimension
    int aug = 0;            // the arena carries the norm chunk per row
    void* bf16 = nullptr;   // [tiles32 x (nkc + aug) x 64 lanes x 8] bf16
    size_t bf16_bytes = 0;

    // int8 arena of integer lists (d <= 256): operands and row norms
    int i8 = 0;
    int nkc8 = 0;
    void* x8 = nullptr;
    void* xn8 = nullptr;
    // the shadow holds every list's rows by increasing distance to the
    // centroid: the ids and 32-row tile radius intervals in that order
    void* ids8 = nullptr;
    void* rbound8 = nullptr;

    // radius interval of every 32-row tile (distance of its rows to their
    // list's centroid), for the scan's tile skipping
    void* rbound = nullptr;
    void* lbound = nullptr;  // the same per list

    // per-batch scratch, grown on demand
    void* scratch = nullptr;
    size_t scratch_bytes = 0;

    int cus = 0;            // compute units of the device

    void invalidate() { valid = false; }
    void release();
    ~BatchedState() { release(); }
};

}  // namespace hivf
