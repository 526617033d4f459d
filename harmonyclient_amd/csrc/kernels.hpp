// Launch interfaces of the hivf kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hivf {

// Tiled list arena.  Rows of one list occupy whole tiles of 64; inside a
// tile the layout is [quad j][lane][4 floats], so lane i of a wavefront
// reading quad j of tile t touches one contiguous 1 KiB segment:
//   codes[(t * nquad + j) * 64 + i] = dims 4j..4j+3 of the tile's row i.
// A dimension block of b quads is the contiguous range of quads it spans.
struct DevLists {
    const float4* codes;          // [tiles * nquad * 64]
    const uint32_t* ids;          // [tiles * 64], kNoId on padding
    const int32_t* list_tile_off; // [nlist + 1], in tiles
    const int32_t* list_size;     // [nlist]
    int nquad;                    // d padded to a multiple of 4, / 4
};

enum ScanMode : int { kScanPlain = 0, kScanPruned = 1, kScanWarmup = 2 };

struct ScanArgs {
    DevLists lists;
    const float4* q;          // [nq x nquad]
    int nq;
    const int32_t* probes;    // [nq x probe_stride], -1 = none
    int probe_stride;
    const int32_t* nprobe_q;  // optional per-query probe counts
    int nprobe;               // probes per query otherwise
    int k;
    int block_dims;           // pruned: dims per dimension block (any size)
    int nblocks;              // pruned: blocks; the last one ends at the padded dimension
    const float* tau;         // pruned: optional external bound [nq]
    int row_limit;            // warm-up: rows per list (W); else INT_MAX
    int warm_lists;           // warm-up: probed lists scanned (L)
    float* out_d;             // [nq x k]
    int64_t* out_i;           // [nq x k]
    float* out_tau;           // warm-up: [nq]
    const float* tau_cap;     // warm-up: optional [nq] upper bounds the output is capped by
    unsigned long long* stats;// optional: candidates, blocks_total, blocks_skipped
    const int32_t* only;      // optional [nq] flags: scan only the flagged queries
    const int32_t* qlist;     // optional: scan only these queries (count in *qcount)
    const uint32_t* qcount;
};

hipError_t launch_scan(const ScanArgs& a, int mode, int metric, hipStream_t s);

// Dimension blocks of the pruned scan for d dims: dim_block > 0 gives
// ceil(d / dim_block) blocks of dim_block dims (the last one shorter), as
// the worker's block_dim does (src/node.cpp:1056-1174); 0 gives the
// reference's four checks of d / 4 dims, the last block taking the
// remainder (src/node.cpp:639-653).
inline void pruning_blocks(int64_t d, int64_t dim_block, int* block_dims, int* nblocks) {
    if (dim_block > 0) {
        *block_dims = (int)(dim_block < d ? dim_block : d);
        *nblocks = (int)((d + *block_dims - 1) / *block_dims);
    } else if (d >= 4) {
        *block_dims = (int)(d / 4);
        *nblocks = 4;
    } else {
        *block_dims = (int)d;
        *nblocks = 1;
    }
}

struct CoarseArgs {
    const float4* q;        // [nq x nquad]
    int nq;
    int nquad;
    const float4* cent;     // centroid tiles, same layout as list tiles
    int nlist;
    int P;                  // probes to keep
    float* out_d;           // [nq x P]
    int32_t* out_l;         // [nq x P], -1 = none
};

hipError_t launch_coarse(const CoarseArgs& a, int metric, hipStream_t s);

}  // namespace hivf
