// Device state of the learned-bound MLP (hivf_tau_model in the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

struct hivf_tau_model {
    int in = 0, hid = 0;
    // flat parameter vector: W1[in x hid], b1[hid], W2[hid x hid], b2[hid],
    // w3[hid], b3, then the untrained mu[in], sd[in], scale, offset
    int64_t off_W1 = 0, off_b1 = 0, off_W2 = 0, off_b2 = 0, off_w3 = 0, off_b3 = 0, off_mu = 0, off_sd = 0,
            off_scale = 0, off_offset = 0, n_params = 0, n_trained = 0;
    float* params = nullptr;
    float* grads = nullptr;
    float* m1 = nullptr;  // Adam moments
    float* m2 = nullptr;
    float* work = nullptr;  // activations
    int64_t work_floats = 0;
    float host_offset = 0.0f;
    float host_scale = 1.0f;  // distance scale s of the target log1p(D_k / s)
    hipStream_t stream = nullptr;

    void layout();
    hipError_t ensure_work(int64_t rows);
    hipError_t forward(hipStream_t s, const float* Xs, int rows, float* H1, float* H2, float* yhat,
                       const float* P) const;
    hipError_t backward(hipStream_t s, const float* Xs, int rows, const float* H1, const float* H2, const float* dy,
                        float* dZ2, float* dZ1, float* G) const;
    hipError_t predict(hipStream_t s, const float* F, int64_t n, float* tau, const float* P);
    ~hivf_tau_model();
};

namespace hivf {
using TauModel = ::hivf_tau_model;
}
