// C++ caller of the GPU path through include/hivf_index.hpp, shaped like the
// reference's src/tests/index_test.cpp: build an index from random data,
// search with nprobe = nlist, and require recall 1 against brute force
// (distances exact: integer data).  Exit status 0 on success.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "hivf_index.hpp"

int main() {
    const size_t d = 16, nb = 20000, nq = 100, nlist = 64, k = 10;
    std::mt19937 rng(123);
    std::uniform_int_distribution<int> u(0, 31);
    std::vector<float> xb(nb * d), xq(nq * d);
    for (auto& v : xb) v = (float)u(rng);
    for (auto& v : xq) v = (float)u(rng);
    hivf::Index index(d, nlist, nlist, HIVF_METRIC_L2);
    index.train(nb, xb.data());
    index.add(nb, xb.data());
    std::vector<float> D(nq * k);
    std::vector<int64_t> I(nq * k);
    for (int algo : {HIVF_ALGO_PLAIN, HIVF_ALGO_PRUNED, HIVF_ALGO_BATCHED}) {
        hivf_search_params p;
        hivf_search_params_default(&p);
        p.algo = algo;
        index.search(nq, xq.data(), k, D.data(), I.data(), 1, &p);
        size_t bad = 0;
        for (size_t q = 0; q < nq; ++q) {
            // brute-force k-th distance; every returned row must be within it
            std::vector<float> all(nb);
            for (size_t i = 0; i < nb; ++i) {
                float s = 0;
                for (size_t c = 0; c < d; ++c) {
                    float t = xq[q * d + c] - xb[i * d + c];
                    s += t * t;
                }
                all[i] = s;
            }
            std::nth_element(all.begin(), all.begin() + (k - 1), all.end());
            float kth = all[k - 1];
            for (size_t j = 0; j < k; ++j) {
                const float* x = &xb[(size_t)I[q * k + j] * d];
                float s = 0;
                for (size_t c = 0; c < d; ++c) {
                    float t = xq[q * d + c] - x[c];
                    s += t * t;
                }
                if (s != D[q * k + j] || s > kth) ++bad;
                if (j && D[q * k + j] < D[q * k + j - 1]) ++bad;
            }
            if (D[q * k + k - 1] != kth) ++bad;
        }
        std::printf("algo %d: %zu mismatches\n", algo, bad);
        if (bad) return 1;
    }
    std::printf("recall 1 for all algorithms\n");
    // merging these rows with an all-padding partition gives them back
    std::vector<float> Dp(2 * nq * k, 3.402823466e+38f), Dm(nq * k);
    std::vector<int64_t> Ip(2 * nq * k, -1), Im(nq * k);
    std::copy(D.begin(), D.end(), Dp.begin());
    std::copy(I.begin(), I.end(), Ip.begin());
    hivf::merge_topk(HIVF_METRIC_L2, nq, 2, k, Dp.data(), Ip.data(), k, Dm.data(), Im.data());
    if (Dm != D || Im != I) {
        std::printf("merge changed the rows\n");
        return 1;
    }
    std::printf("merge ok\n");
    return 0;
}
