// The batched search's per-query preparation, shared by its own launch and
// the coarse quantizer's query preparation (one launch for both when the
// coarse step runs first with it).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bf16.hpp"

namespace hivf {

// Per query (one wave): the BF16 fragments (q and -2q), the FP64-summed
// squared norm with its 3-piece augmentation, whether the query is integer
// in [0, 255], and, for the int8 scan (q8 non-null), q - 128 as int8
// (row-major, dpad32; 0 past d) with sum (q - 128)^2.  It also clears the
// search's per-query flags and survivor counts, and (workgroup 0) the rescue
// list counts and the per-list query counts of the first pass.
struct PrepArgs {
    const float* q;
    int d, ld, dpad16, dpad32;
    uint16_t* qbf;
    uint16_t* qbf2;
    uint16_t* qaug;
    float* qnorm;
    int32_t* qint;
    int8_t* q8;      // null: no int8 scan
    int32_t* qn8;
    int32_t* fb;
    int32_t* fb2;
    uint32_t* scnt;
    uint32_t* fbc;   // [4]
    uint32_t* lcnt;  // [nlist]
    int nlist;
};

// A query's first kQRegs * 64 components (0 past d), one round of loads:
// lane l holds components l, l + 64, ...  Every preparation of the query
// reads them from here when its padded width fits (d <= 256).
constexpr int kQRegs = 4;

__device__ inline void query_regs(const float* q, int64_t qi, int ld, int d, int lane, float (&v)[kQRegs]) {
#pragma unroll
    for (int j = 0; j < kQRegs; ++j) {
        const int c = j * 64 + lane;
        v[j] = c < d ? q[qi * ld + c] : 0.0f;
    }
}

// Query qi's preparation by one 64-lane wave (lane = its lane); qv: the
// query's components from query_regs (used when dpad16 and dpad32 fit).
__device__ inline void prep_query(const PrepArgs& a, const int qi, const int lane, const float (&qv)[kQRegs]) {
    const bool regs = a.dpad16 <= 64 * kQRegs && a.dpad32 <= 64 * kQRegs;
    double s = 0.0;
    int ok = 1;
    auto one = [&](int c, float v) {
        a.qbf[(size_t)qi * a.dpad16 + c] = f2bf(v);
        a.qbf2[(size_t)qi * a.dpad16 + c] = f2bf(-2.0f * v);  // = -2 bf16(v) exactly
        s += (double)v * (double)v;
        ok &= (v == rintf(v)) && v >= 0.0f && v <= 255.0f;
    };
    if (regs) {
#pragma unroll
        for (int j = 0; j < kQRegs; ++j)
            if (j * 64 + lane < a.dpad16) one(j * 64 + lane, qv[j]);
    } else {
        for (int c = lane; c < a.dpad16; c += 64) one(c, c < a.d ? a.q[(size_t)qi * a.ld + c] : 0.0f);
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off);
        ok &= __shfl_down(ok, off);
    }
    if (lane == 0) {
        const float n = (float)s;
        a.qnorm[qi] = n;
        a.qint[qi] = ok;
        uint16_t v[8] = {kBf16One, kBf16One, kBf16One, 0, 0, 0, 0, 0};
        split3(n, v + 3);
        for (int e = 0; e < 8; ++e) a.qaug[(size_t)qi * 8 + e] = v[e];
        a.fb[qi] = 0;
        a.fb2[qi] = 0;
        a.scnt[qi] = 0;
    }
    if (a.q8) {
        int s8 = 0;
        auto one8 = [&](int c, float v) {
            const int iv = (v >= 0.0f && v <= 255.0f) ? (int)v - 128 : 0;
            a.q8[(size_t)qi * a.dpad32 + c] = (int8_t)iv;
            s8 += iv * iv;
        };
        if (regs) {
#pragma unroll
            for (int j = 0; j < kQRegs; ++j) {
                const int c = j * 64 + lane;
                if (c < a.dpad32) one8(c, c < a.d ? qv[j] : 128.0f);
            }
        } else {
            for (int c = lane; c < a.dpad32; c += 64) one8(c, c < a.d ? a.q[(size_t)qi * a.ld + c] : 128.0f);
        }
        for (int off = 32; off > 0; off >>= 1) s8 += __shfl_down(s8, off);
        if (lane == 0) a.qn8[qi] = s8;
    }
    if (qi == 0) {
        if (lane < 4) a.fbc[lane] = 0;
        for (int l = lane; l < a.nlist; l += 64) a.lcnt[l] = 0;
    }
}

}  // namespace hivf
