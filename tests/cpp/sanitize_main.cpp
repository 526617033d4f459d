// Host-only exercise of the index-file parser (harmony_format.cpp) and the
// CPU oracle (oracle.c), built under AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_format.py.  Exit status 0 and
// "sanitize ok" on success; a sanitizer report aborts the run.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "harmony_format.hpp"
#include "hivf.h"
#include "oracle.h"

static int failures = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            std::printf("check failed line %d: %s\n", __LINE__, #c); \
            ++failures;                                             \
        }                                                           \
    } while (0)

static void oracle_cases() {
    std::mt19937 rng(7);
    std::uniform_int_distribution<int> u(0, 15);
    const int64_t d = 13, n = 600, nlist = 9, nq = 17;
    std::vector<float> x(n * d), q(nq * d), cent(nlist * d);
    for (auto& v : x) v = (float)u(rng);
    for (auto& v : q) v = (float)u(rng);
    for (int64_t c = 0; c < nlist; ++c)
        for (int64_t j = 0; j < d; ++j) cent[c * d + j] = x[(c * 31) * d + j];
    orc_kmeans(ORC_METRIC_L2, n, x.data(), nlist, d, 3, cent.data());
    std::vector<int64_t> off(nlist + 1), order(n);
    orc_build_lists(ORC_METRIC_L2, n, x.data(), nlist, cent.data(), d, off.data(), order.data());
    std::vector<float> codes(n * d);
    for (int64_t i = 0; i < n; ++i) std::memcpy(&codes[i * d], &x[order[i] * d], d * sizeof(float));
    orc_index ix{ORC_METRIC_L2, d, nlist, cent.data(), off.data(), order.data(), codes.data()};
    for (int64_t k : {1, 10, 700}) {  // 700 > every candidate count: sentinel tails
        for (int64_t P : {1, 4, 50}) {
            std::vector<float> D(nq * k), D2(nq * k), tau(nq);
            std::vector<int64_t> I(nq * k), I2(nq * k), probes(nq * std::min(P, nlist));
            orc_stats st{};
            int64_t Pc = orc_search(&ix, nq, q.data(), P, k, D.data(), I.data(), 1, &st);
            CHECK(Pc == std::min(P, nlist));
            std::vector<float> cd(nq * Pc);
            orc_coarse(ORC_METRIC_L2, nq, q.data(), nlist, cent.data(), d, Pc, cd.data(), probes.data());
            orc_warmup_tau(&ix, nq, q.data(), probes.data(), Pc, k, 2, k, tau.data());
            for (int64_t blk : {1, 4, 13}) {
                orc_search_pruned(&ix, nq, q.data(), probes.data(), Pc, k, blk, tau.data(), 1, D2.data(),
                                  I2.data(), 1, &st);
                CHECK(I == I2);
                CHECK(std::memcmp(D.data(), D2.data(), D.size() * sizeof(float)) == 0);
            }
            // merging the result with an all-padding partition keeps it
            std::vector<float> Dp(2 * nq * k, 3.402823466e+38f), Dm(nq * k);
            std::vector<int64_t> Ip(2 * nq * k, -1), Im(nq * k);
            std::copy(D.begin(), D.end(), Dp.begin());
            std::copy(I.begin(), I.end(), Ip.begin());
            orc_merge(ORC_METRIC_L2, nq, 2, k, Dp.data(), Ip.data(), k, Dm.data(), Im.data());
            CHECK(Im == I);
            int64_t gt_k = k;
            double r = orc_recall(I.data(), I.data(), nq, k, gt_k);
            CHECK(r >= 0.0 && r <= 1.0);
        }
    }
    // IP metric coarse and search
    ix.metric = ORC_METRIC_IP;
    std::vector<float> D(nq * 5);
    std::vector<int64_t> I(nq * 5);
    orc_search(&ix, nq, q.data(), 3, 5, D.data(), I.data(), 1, nullptr);
    std::vector<float> v(40);
    std::vector<int64_t> id(40), oi(8);
    for (int i = 0; i < 40; ++i) {
        v[i] = (float)(i % 5);
        id[i] = 40 - i;
    }
    orc_topk(ORC_METRIC_L2, 40, v.data(), id.data(), 8, D.data(), oi.data());
    orc_topk(ORC_METRIC_IP, 0, v.data(), id.data(), 8, D.data(), oi.data());
    CHECK(oi[0] == -1);
}

static void format_cases(const std::string& dir) {
    hivf::HarmonyFile f;
    f.d = 6;
    f.nlist = 4;
    f.metric = 1;
    f.centroids.assign(24, 1.5f);
    f.list_off = {0, 3, 3, 5, 9};  // list 1 empty
    for (int i = 0; i < 9; ++i) f.ids.push_back(100 + i);
    for (int i = 0; i < 54; ++i) f.codes.push_back((float)(i % 7));
    const std::string path = dir + "/san.index";
    std::string err;
    CHECK(hivf::write_harmony_file(path.c_str(), f, &err) == HIVF_OK);
    hivf::HarmonyFile g;
    CHECK(hivf::read_harmony_file(path.c_str(), &g, &err) == HIVF_OK);
    CHECK(g.list_off == f.list_off && g.ids == f.ids && g.codes == f.codes && g.centroids == f.centroids);
    // every truncation parses or fails cleanly
    FILE* in = std::fopen(path.c_str(), "rb");
    std::vector<char> raw;
    int c;
    while ((c = std::fgetc(in)) != EOF) raw.push_back((char)c);
    std::fclose(in);
    const std::string cut = dir + "/san_cut.index";
    for (size_t len = 0; len <= raw.size(); ++len) {
        FILE* o = std::fopen(cut.c_str(), "wb");
        std::fwrite(raw.data(), 1, len, o);
        std::fclose(o);
        hivf::HarmonyFile h;
        int rc = hivf::read_harmony_file(cut.c_str(), &h, &err);
        CHECK(rc == HIVF_OK || rc == HIVF_EIO);
    }
    // corrupt every 8-byte word of the list records with huge values
    const size_t hdr = 48 + 24 * 4;
    for (size_t pos = hdr; pos + 8 <= raw.size(); pos += 4) {
        for (uint64_t val : {~0ull, 1ull << 40, 1ull << 62, 7ull}) {
            std::vector<char> b = raw;
            std::memcpy(&b[pos], &val, 8);
            FILE* o = std::fopen(cut.c_str(), "wb");
            std::fwrite(b.data(), 1, b.size(), o);
            std::fclose(o);
            hivf::HarmonyFile h;
            int rc = hivf::read_harmony_file(cut.c_str(), &h, &err);
            CHECK(rc == HIVF_OK || rc == HIVF_EIO);
        }
    }
    hivf::HarmonyFile h;
    CHECK(hivf::read_harmony_file((dir + "/none").c_str(), &h, &err) == HIVF_EIO);
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    oracle_cases();
    format_cases(dir);
    if (failures) {
        std::printf("%d checks failed\n", failures);
        return 1;
    }
    std::printf("sanitize ok\n");
    return 0;
}
