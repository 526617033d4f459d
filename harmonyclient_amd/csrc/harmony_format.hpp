// The reference's index file format, read and written on the host.
//
// Layout (src/Index.cpp:1708-1732 save_index, src/IVF.cpp:76-99 save_IVF):
//   size_t d, nlist; int32 metric; int32 opt level (OptLevel, src/common.h:23-32);
//   size_t sub_k, sub_nlist, sub_nprobe;
//   float centroids[nlist * d];
//   then, for every non-empty list, in any order:
//     size_t list_id; size_t n, d, sub_k; int32 opt;
//     size_t ids[n]; float codes[n * d]; float norms[n];
//     float candidate2centroid[n], sqrt_candidate2centroid[n]   if opt & (TRIANGLE | SUBNN_IP)
//     int64 sub_nearest_L2_id[n * sub_k]; float sub_nearest_L2_dis[n * sub_k]   if opt & SUBNN_L2
//     int64 ip_id[n * sub_k]; float ip_dis[n * sub_k];
//     int64 far_ip_id[n * sub_k]; float far_ip_dis[n * sub_k]                  if opt & SUBNN_IP
// A list record that repeats a list id replaces the earlier one, as the
// reference's load_index does (src/Index.cpp:1753-1762).
//
// Host C++ only (no HIP), so the parser also builds under the host
// sanitizers on its own.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hivf {

// OptLevel bits, src/common.h:23-32
constexpr int32_t kOptTriangle = 1, kOptSubnnL2 = 2, kOptSubnnIP = 4;

struct HarmonyFile {
    uint64_t d = 0, nlist = 0;
    int32_t metric = 1, opt = 0;
    uint64_t sub_k = 0, sub_nlist = 1, sub_nprobe = 1;
    std::vector<float> centroids;   // [nlist * d]
    std::vector<int64_t> list_off;  // [nlist + 1], CSR over ids / codes
    std::vector<int64_t> ids;       // [ntotal]
    std::vector<float> codes;       // [ntotal * d]
};

// Status codes are the HIVF_E* values of include/hivf.h.  Every size the
// file states is checked against the bytes that remain before anything is
// allocated; a malformed file gives HIVF_EIO and a message, never an
// exception or an abort.
int read_harmony_file(const char* path, HarmonyFile* out, std::string* err);

// Writes the lists of `f` (opt level 0, no extras); norms are the
// sequential FP32 FMA chain of each row (src/Index.cpp:689-696 stores
// calculatedInnerProduct(x, x, d)).
int write_harmony_file(const char* path, const HarmonyFile& f, std::string* err);

}  // namespace hivf
