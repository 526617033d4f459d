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
    // the call sequence of src/query.cpp:825-849: a default-constructed
    // index, move-assigned from the 9-argument constructor (OPT_ALL passed
    // as sub_k, as the reference does), trained with the default recipe,
    // filled, nprobe 0 meaning every list
    using namespace hivf;
    Index index;
    index = Index(d, nlist, 0, METRIC_L2, OPT_NONE, OPT_ALL, 1, 1, false);
    index.train(nb, xb.data());
    index.add(nb, xb.data());
    index.nprobe = 0;
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
    // save, then load into another index: the same results (src/Index.cpp:1708-1770)
    index.save_index("index_test.index");
    Index loaded;
    loaded.load_index("index_test.index");
    loaded.nprobe = 0;
    std::vector<float> D2(nq * k);
    std::vector<int64_t> I2(nq * k);
    hivf_search_params p;
    hivf_search_params_default(&p);
    p.algo = HIVF_ALGO_BATCHED;
    loaded.search(nq, xq.data(), k, D2.data(), I2.data(), 1, &p);
    if (D2 != D || I2 != I || loaded.d != d || loaded.nlist != nlist) {
        std::printf("the loaded index searches differently\n");
        return 1;
    }
    std::printf("save / load ok\n");
    // a small nprobe read from the public field at search time
    index.nprobe = 2;
    Stats st = index.search(nq, xq.data(), k, D2.data(), I2.data());
    if (st.candidates >= (int64_t)(nq * nb)) {
        std::printf("nprobe 2 scanned every list\n");
        return 1;
    }
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
