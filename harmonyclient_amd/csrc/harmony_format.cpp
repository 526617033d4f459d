// Reader and writer of the reference's index file format (layout in
// harmony_format.hpp).  Host C++ only.
#include "harmony_format.hpp"

#include <cmath>
#include <cstdio>
#include <limits>
#include <new>
#include <stdexcept>

#include "hivf.h"

namespace hivf {
namespace {

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

template <class T>
bool get(FILE* f, T* v, size_t n = 1) {
    return n == 0 || fread(v, sizeof(T), n, f) == n;
}

template <class T>
bool put(FILE* f, const T* v, size_t n = 1) {
    return n == 0 || fwrite(v, sizeof(T), n, f) == n;
}

int fail(std::string* err, int code, const std::string& m) {
    if (err) *err = m;
    return code;
}

}  // namespace

int read_harmony_file(const char* path, HarmonyFile* out, std::string* err) {
    if (!path || !out) return fail(err, HIVF_EINVAL, "null path or output");
    File fc;
    fc.f = fopen(path, "rb");
    if (!fc.f) return fail(err, HIVF_EIO, std::string("cannot open ") + path);
    if (fseeko(fc.f, 0, SEEK_END) != 0) return fail(err, HIVF_EIO, "cannot seek");
    const off_t size = ftello(fc.f);
    if (size < 0 || fseeko(fc.f, 0, SEEK_SET) != 0) return fail(err, HIVF_EIO, "cannot seek");
    uint64_t left = (uint64_t)size;
    auto take = [&](uint64_t bytes) {  // true when `bytes` more remain
        if (bytes > left) return false;
        left -= bytes;
        return true;
    };

    HarmonyFile f;
    if (!take(48) || !(get(fc.f, &f.d) && get(fc.f, &f.nlist) && get(fc.f, &f.metric) && get(fc.f, &f.opt) &&
                       get(fc.f, &f.sub_k) && get(fc.f, &f.sub_nlist) && get(fc.f, &f.sub_nprobe)))
        return fail(err, HIVF_EIO, "truncated header");
    if (f.d < 1 || f.d > 16384 || f.nlist < 1 || f.nlist > (1u << 24) || (f.metric != 0 && f.metric != 1))
        return fail(err, HIVF_EIO, "implausible header (d, nlist or metric)");
    const uint64_t d = f.d, nlist = f.nlist;
    if (!take(nlist * d * sizeof(float))) return fail(err, HIVF_EIO, "truncated centroids");
    f.centroids.resize(nlist * d);
    if (!get(fc.f, f.centroids.data(), f.centroids.size())) return fail(err, HIVF_EIO, "read error (centroids)");

    std::vector<std::vector<int64_t>> ids(nlist);
    std::vector<std::vector<float>> codes(nlist);
    std::vector<uint64_t> u;
    std::vector<float> norms;
    for (;;) {
        // the reference stops at the first list id it cannot read whole
        // (src/Index.cpp:1755-1759)
        uint64_t lid;
        if (left < sizeof(lid)) break;
        take(sizeof(lid));
        if (!get(fc.f, &lid)) return fail(err, HIVF_EIO, "read error (list id)");
        uint64_t n, dd, sk;
        int32_t lopt;
        if (!take(28) || !(get(fc.f, &n) && get(fc.f, &dd) && get(fc.f, &sk) && get(fc.f, &lopt)))
            return fail(err, HIVF_EIO, "truncated list record header");
        if (lid >= nlist) return fail(err, HIVF_EIO, "list id beyond nlist");
        if (dd != d) return fail(err, HIVF_EIO, "list record of another dimension");
        if (lopt < 0 || lopt > 7) return fail(err, HIVF_EIO, "unknown opt level in a list record");
        // bytes per row: id, codes, norm, and the extras of the opt level
        const uint64_t sk_cap = left / 12 + 1;  // keeps the products below in range
        if ((lopt & (kOptSubnnL2 | kOptSubnnIP)) && sk > sk_cap)
            return fail(err, HIVF_EIO, "truncated list extras");
        uint64_t row = sizeof(uint64_t) + d * sizeof(float) + sizeof(float);
        if (lopt & (kOptTriangle | kOptSubnnIP)) row += 2 * sizeof(float);
        if (lopt & kOptSubnnL2) row += sk * (sizeof(int64_t) + sizeof(float));
        if (lopt & kOptSubnnIP) row += 2 * sk * (sizeof(int64_t) + sizeof(float));
        if (n > left / row) return fail(err, HIVF_EIO, "truncated list (its size exceeds the file)");
        take(n * row);
        u.resize(n);
        norms.resize(n);
        std::vector<float> c(n * d);
        if (!(get(fc.f, u.data(), n) && get(fc.f, c.data(), n * d) && get(fc.f, norms.data(), n)))
            return fail(err, HIVF_EIO, "read error (list)");
        const uint64_t extras = n * (row - (sizeof(uint64_t) + d * sizeof(float) + sizeof(float)));
        if (extras && fseeko(fc.f, (off_t)extras, SEEK_CUR) != 0) return fail(err, HIVF_EIO, "cannot skip extras");
        std::vector<int64_t> id(n);
        for (uint64_t r = 0; r < n; ++r) {
            if (u[r] > (uint64_t)std::numeric_limits<int64_t>::max())
                return fail(err, HIVF_EIO, "row id beyond the int64 range");
            id[r] = (int64_t)u[r];
        }
        ids[lid] = std::move(id);  // a repeated list id replaces the earlier record
        codes[lid] = std::move(c);
    }

    f.list_off.assign(nlist + 1, 0);
    for (uint64_t l = 0; l < nlist; ++l) f.list_off[l + 1] = f.list_off[l] + (int64_t)ids[l].size();
    const int64_t total = f.list_off[nlist];
    f.ids.resize(total);
    f.codes.resize((size_t)total * d);
    for (uint64_t l = 0; l < nlist; ++l) {
        std::copy(ids[l].begin(), ids[l].end(), f.ids.begin() + f.list_off[l]);
        std::copy(codes[l].begin(), codes[l].end(), f.codes.begin() + f.list_off[l] * d);
    }
    *out = std::move(f);
    return HIVF_OK;
}

int write_harmony_file(const char* path, const HarmonyFile& f, std::string* err) {
    if (!path) return fail(err, HIVF_EINVAL, "null path");
    const uint64_t d = f.d, nlist = f.nlist;
    if (f.centroids.size() != nlist * d || f.list_off.size() != nlist + 1 ||
        f.ids.size() != (size_t)f.list_off[nlist] || f.codes.size() != f.ids.size() * d)
        return fail(err, HIVF_EINVAL, "inconsistent index shape");
    File fc;
    fc.f = fopen(path, "wb");
    if (!fc.f) return fail(err, HIVF_EIO, std::string("cannot open ") + path);
    const int32_t opt = 0;
    bool ok = put(fc.f, &f.d) && put(fc.f, &f.nlist) && put(fc.f, &f.metric) && put(fc.f, &opt) &&
              put(fc.f, &f.sub_k) && put(fc.f, &f.sub_nlist) && put(fc.f, &f.sub_nprobe) &&
              put(fc.f, f.centroids.data(), f.centroids.size());
    std::vector<uint64_t> u;
    std::vector<float> norms;
    for (uint64_t l = 0; ok && l < nlist; ++l) {
        const uint64_t n = (uint64_t)(f.list_off[l + 1] - f.list_off[l]);
        if (n == 0) continue;  // empty lists are not written (src/Index.cpp:1725-1730)
        const int64_t* id = f.ids.data() + f.list_off[l];
        const float* x = f.codes.data() + (size_t)f.list_off[l] * d;
        u.assign(id, id + n);
        norms.resize(n);
        for (uint64_t r = 0; r < n; ++r) {
            float s = 0.0f;
            for (uint64_t c = 0; c < d; ++c) s = std::fma(x[r * d + c], x[r * d + c], s);
            norms[r] = s;
        }
        ok = put(fc.f, &l) && put(fc.f, &n) && put(fc.f, &f.d) && put(fc.f, &f.sub_k) && put(fc.f, &opt) &&
             put(fc.f, u.data(), n) && put(fc.f, x, n * d) && put(fc.f, norms.data(), n);
    }
    if (!ok) return fail(err, HIVF_EIO, std::string("write failed: ") + path);
    if (fflush(fc.f) != 0) return fail(err, HIVF_EIO, std::string("write failed: ") + path);
    return HIVF_OK;
}

}  // namespace hivf
