// Learned pruning bound: a small MLP predicting each query's k-th neighbour
// distance from its coarse-quantizer distances.
//
// Features (per query): log1p of the 32 nearest centroid distances (the
// coarse quantizer's sorted output, src/Index.cpp:967-995) and log1p of the
// squared query norm, standardised with the training means / deviations.
// Network: 33 -> H -> H -> 1, ReLU, FP32.  It predicts log1p(D_k / s) for a
// distance scale s fixed at training (a hundredth of the training traces'
// median k-th distance, so the target is close to log D_k -- relative errors
// -- over the whole range of the workload's distances, small or large); the
// bound is tau = s max(expm1(prediction + offset), 0), the offset calibrated
// on held-out traces.  The bound only steers pruning: the batched scan rescans exactly
// any query with fewer than k survivors at or below it.
//
// Training runs on the GPU with the kernels below: a tiled FP32 GEMM with
// bias / ReLU / mask epilogues (forward and both backward products), column
// sums for bias gradients, the pinball-loss gradient and Adam.  Every
// reduction runs in a fixed order, so training is deterministic.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "index_impl.hpp"
#include "tau_model.hpp"

using namespace hivf;

namespace hivf {
namespace {

constexpr int kFeatCoarse = 32;

struct GemmArgs {
    const float* A;
    int lda;
    int transA;  // A'(m, k) = transA ? A[k * lda + m] : A[m * lda + k]
    const float* B;
    int ldb;
    int transB;  // B'(k, n) = transB ? B[n * ldb + k] : B[k * ldb + n]
    float* C;
    int ldc;
    int M, N, K;
    const float* bias;  // [N] or null
    int relu;
    const float* mask;  // C *= (mask[m * ldm + n] > 0) when non-null
    int ldm;
};

// C = epilogue(A' B'): 64 x 64 output tile per 256-thread workgroup, 4 x 4
// per thread, K in chunks of 16 staged in LDS; fixed summation order.
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
    __shared__ float As[16][65];
    __shared__ float Bs[16][65];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        for (int e = threadIdx.x; e < 1024; e += 256) {
            int kk = e / 64, r = e % 64;
            int m = m0 + r, k = k0 + kk, n = n0 + r;
            As[kk][r] = (m < g.M && k < g.K) ? (g.transA ? g.A[(size_t)k * g.lda + m] : g.A[(size_t)m * g.lda + k]) : 0.0f;
            Bs[kk][r] = (n < g.N && k < g.K) ? (g.transB ? g.B[(size_t)n * g.ldb + k] : g.B[(size_t)k * g.ldb + n]) : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int m = m0 + ty * 4 + i;
        if (m >= g.M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int n = n0 + tx * 4 + j;
            if (n >= g.N) continue;
            float v = acc[i][j];
            if (g.bias) v += g.bias[n];
            if (g.relu) v = fmaxf(v, 0.0f);
            if (g.mask && !(g.mask[(size_t)m * g.ldm + n] > 0.0f)) v = 0.0f;
            g.C[(size_t)m * g.ldc + n] = v;
        }
    }
}

// out[c] = sum_r X[r * ld + c], rows in a fixed order.
__global__ __launch_bounds__(256) void colsum_kernel(const float* __restrict__ X, int rows, int cols, int ld,
                                                     float* __restrict__ out) {
    __shared__ float part[4][64];
    const int c = blockIdx.x * 64 + threadIdx.x % 64;
    const int w = threadIdx.x / 64;
    float s = 0.0f;
    if (c < cols)
        for (int r = w; r < rows; r += 4) s += X[(size_t)r * ld + c];
    part[w][threadIdx.x % 64] = s;
    __syncthreads();
    if (w == 0 && c < cols) out[c] = (part[0][threadIdx.x] + part[1][threadIdx.x]) + (part[2][threadIdx.x] + part[3][threadIdx.x]);
}

// Raw features from the sorted coarse distances (stride P) and the query's
// squared norm.
__global__ void features_kernel(const float* __restrict__ cd, int P, const int32_t* __restrict__ npq,
                                const float* __restrict__ qnorm, int n, float* __restrict__ F) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    int Pq = npq ? min(npq[q], P) : P;
    if (Pq < 1) Pq = 1;
    for (int i = 0; i < kFeatCoarse; ++i) {
        float v = cd[(size_t)q * P + min(i, Pq - 1)];
        F[(size_t)q * (kFeatCoarse + 1) + i] = log1pf(fmaxf(v, 0.0f));
    }
    F[(size_t)q * (kFeatCoarse + 1) + kFeatCoarse] = log1pf(fmaxf(qnorm[q], 0.0f));
}

// Xs[i] = (X[idx[i]] - mu) / sd, ys[i] = y[idx[i]] (idx null: identity).
__global__ void gather_std_kernel(const float* __restrict__ X, const float* __restrict__ y,
                                  const int32_t* __restrict__ idx, int n, int in, const float* __restrict__ mu,
                                  const float* __restrict__ sd, float* __restrict__ Xs, float* __restrict__ ys) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * in) return;
    int i = e / in, c = e % in;
    int r = idx ? idx[i] : i;
    Xs[e] = (X[(size_t)r * in + c] - mu[c]) / sd[c];
    if (c == 0 && y && ys) ys[i] = y[r];
}

// Pinball loss at quantile qt on r = y - yhat; gradient w.r.t. yhat, /n.
__global__ void pinball_kernel(const float* __restrict__ yhat, const float* __restrict__ y, int n, float qt,
                               float* __restrict__ dy, float* __restrict__ loss) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float r = y[i] - yhat[i];
    loss[i] = r >= 0.0f ? qt * r : (qt - 1.0f) * r;
    dy[i] = (r > 0.0f ? -qt : (r < 0.0f ? 1.0f - qt : 0.0f)) / (float)n;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// tau = scale max(expm1(yhat + offset), 0)
__global__ void tau_out_kernel(const float* __restrict__ yhat, int n, float offset, float scale,
                               float* __restrict__ tau) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tau[i] = scale * fmaxf(expm1f(yhat[i] + offset), 0.0f);
}

unsigned blocks(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// The whole inference for a search in one launch: features, standardisation,
// the three layers and the bound, 8 R queries per 256-thread workgroup (R
// chosen at launch so that the workgroups spread over the compute units
// about two each).  W1,
// then each half of W2 in turn, sit in one LDS buffer; the next weights are
// loaded into registers while the current layer runs.  Every output is the
// same FMA chain as the GEMM path above (k ascending from zero, then bias,
// then ReLU), so the bounds are bit-identical to features_kernel +
// gather_std_kernel + the three gemm_kernel launches + tau_out_kernel.
// Hidden width: a multiple of 8, at most kFusedHid.
constexpr int kFusedHid = 128;
constexpr int kFusedMaxR = 8;
constexpr int kFeat = kFeatCoarse + 1;
constexpr int kHalfRegs = kFusedHid * kFusedHid / 2 / 256;  // W2 floats per thread per half

struct FusedArgs {
    const float* cd;      // [n x P] sorted coarse distances
    int P;
    const int32_t* npq;   // per-query probe counts (null: P)
    const float* qnorm;   // [n] squared query norms
    int n;
    const float* prm;     // model parameters (offsets of hivf_tau_model::layout)
    int hid;
    int64_t off_W1, off_b1, off_W2, off_b2, off_w3, off_b3, off_mu, off_sd;
    float offset;
    float scale;
    float* tau;
};

// LDS: weights max([33][H], [H/2][H]), features [8R][33], activations
// [8R][H + 1] (padded: the last layer reads one row per thread), w3 [H]
size_t fused_lds_bytes(int H, int R) {
    const size_t w = std::max<size_t>((size_t)kFeat * H, (size_t)(H / 2) * H);
    return sizeof(float) * (w + (size_t)8 * R * kFeat + (size_t)8 * R * (H + 1) + H);
}

template <int R>
__global__ __launch_bounds__(256) void tau_mlp_kernel(FusedArgs a) {
    constexpr int kRows = 8 * R;
    extern __shared__ float lds[];
    const int H = a.hid, ldh = H + 1, Hh = H / 2;
    float* Ws = lds;
    float* Xs = Ws + max(kFeat * H, Hh * H);   // [8R][33]
    float* Hs = Xs + kRows * kFeat;            // [8R][H + 1]: H1, then H2
    float* w3 = Hs + kRows * ldh;              // [H]
    const int tid = threadIdx.x;
    const int q0 = blockIdx.x * kRows;
    const int half_n = Hh * H;                 // floats of one half of W2
    float wr[kHalfRegs];
    auto fetch = [&](int hf) {
#pragma unroll
        for (int t = 0; t < kHalfRegs; ++t) {
            const int e = t * 256 + tid;
            wr[t] = e < half_n ? a.prm[a.off_W2 + (int64_t)hf * half_n + e] : 0.0f;
        }
    };
    auto put = [&]() {
#pragma unroll
        for (int t = 0; t < kHalfRegs; ++t) {
            const int e = t * 256 + tid;
            if (e < half_n) Ws[e] = wr[t];
        }
    };
    {
        // The features, W1, w3 and the first half of W2, every load of a
        // thread issued before any of them is used: the workgroup waits for
        // memory once (twice with per-query probe counts, which the coarse
        // distances' addresses need), not once per element.  Padding
        // elements read row n - 1 and are not stored.
        constexpr int kW1Regs = (kFeat * kFusedHid + 255) / 256;
        constexpr int kFE = (kRows * kFeat + 255) / 256;
        int pq[kFE];
        float mu[kFE], sd[kFE], fx[kFE];
#pragma unroll
        for (int u = 0; u < kFE; ++u) pq[u] = a.P;
        if (a.npq) {
#pragma unroll
            for (int u = 0; u < kFE; ++u) pq[u] = a.npq[min(q0 + (u * 256 + tid) / kFeat, a.n - 1)];
        }
#pragma unroll
        for (int u = 0; u < kFE; ++u) {
            const int e = u * 256 + tid, c = e % kFeat, q = min(q0 + e / kFeat, a.n - 1);
            const int Pq = max(1, min(pq[u], a.P));
            const float cv = a.cd[(size_t)q * a.P + min(min(c, kFeatCoarse - 1), Pq - 1)], qv = a.qnorm[q];
            fx[u] = c < kFeatCoarse ? cv : qv;  // (both loaded: no branch between the loads)
            mu[u] = a.prm[a.off_mu + c];
            sd[u] = a.prm[a.off_sd + c];
        }
        const int nw1 = kFeat * H;
        float w1r[kW1Regs];
#pragma unroll
        for (int t = 0; t < kW1Regs; ++t) w1r[t] = a.prm[a.off_W1 + min(t * 256 + tid, nw1 - 1)];
        const float w3v = a.prm[a.off_w3 + min(tid, H - 1)];
        fetch(0);
#pragma unroll
        for (int u = 0; u < kFE; ++u) {
            const int e = u * 256 + tid;
            if (e < kRows * kFeat) {
                float x = 0.0f;
                if (q0 + e / kFeat < a.n) x = (log1pf(fmaxf(fx[u], 0.0f)) - mu[u]) / sd[u];
                Xs[e] = x;
            }
        }
#pragma unroll
        for (int t = 0; t < kW1Regs; ++t)
            if (t * 256 + tid < nw1) Ws[t * 256 + tid] = w1r[t];
        if (tid < H) w3[tid] = w3v;
    }
    __syncthreads();
    // thread (ty, tx): rows ty*R .. ty*R + R - 1, columns tx*4 .. +3
    const int ty = tid / 32, tx = tid % 32;
    const bool on = tx * 4 < H;
    float o[R][4];
    auto part = [&](const float* In, int ldi, int k0, int k1, int kw) {
        if (!on) return;
#pragma unroll 4
        for (int k = k0; k < k1; ++k) {
            float x[R];
#pragma unroll
            for (int i = 0; i < R; ++i) x[i] = In[(ty * R + i) * ldi + k];
            const float4 u = *reinterpret_cast<const float4*>(Ws + (k - kw) * H + tx * 4);
            const float w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int i = 0; i < R; ++i) o[i][j] = __builtin_fmaf(x[i], w[j], o[i][j]);
        }
    };
    auto reset = [&]() {
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) o[i][j] = 0.0f;
    };
    auto finish_store = [&](const float* bias) {
        if (!on) return;
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = tx * 4 + j;
                Hs[(ty * R + i) * ldh + c] = fmaxf(o[i][j] + bias[c], 0.0f);
            }
    };
    reset();
    part(Xs, kFeat, 0, kFeat, 0);
    finish_store(a.prm + a.off_b1);
    reset();
    __syncthreads();  // H1 stored; W1 no longer read
    put();
    fetch(1);
    __syncthreads();
    part(Hs, ldh, 0, Hh, 0);
    __syncthreads();
    put();
    __syncthreads();
    part(Hs, ldh, Hh, H, Hh);
    __syncthreads();  // every read of H1 is done
    finish_store(a.prm + a.off_b2);
    __syncthreads();
    if (tid < kRows && q0 + tid < a.n) {
        float y = 0.0f;
#pragma unroll 8
        for (int k = 0; k < H; ++k) y = __builtin_fmaf(Hs[tid * ldh + k], w3[k], y);
        y += a.prm[a.off_b3];
        a.tau[q0 + tid] = a.scale * fmaxf(expm1f(y + a.offset), 0.0f);
    }
}

__global__ void fill_zero_kernel(float* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace

// ---- model --------------------------------------------------------------
}  // namespace hivf

hivf_tau_model::~hivf_tau_model() {
    for (float** p : {&params, &grads, &m1, &m2, &work})
        if (*p) hivf::dev_free(*p);
    if (stream) (void)hipStreamDestroy(stream);
}

hipError_t hivf_tau_model::ensure_work(int64_t rows) {
    // activations and gradients of one pass over `rows` rows
    int64_t need = rows * (2 * in + 6 * hid + 8) + 64;
    if (need <= work_floats && work) return hipSuccess;
    if (work) hivf::dev_free(work);
    work = nullptr;
    work_floats = 0;
    hipError_t e = hipMalloc(&work, need * sizeof(float));
    if (e == hipSuccess) work_floats = need;
    return e;
}

namespace hivf {
static hipError_t gemm(hipStream_t s, const float* A, int lda, int tA, const float* B, int ldb, int tB, float* C,
                       int ldc, int M, int N, int K, const float* bias, int relu, const float* mask, int ldm) {
    GemmArgs g{A, lda, tA, B, ldb, tB, C, ldc, M, N, K, bias, relu, mask, ldm};
    hipLaunchKernelGGL(gemm_kernel, dim3(blocks(M, 64), blocks(N, 64)), dim3(256), 0, s, g);
    return hipGetLastError();
}

}  // namespace hivf

// Forward of `rows` standardised feature rows Xs -> yhat; keeps H1, H2.
hipError_t hivf_tau_model::forward(hipStream_t s, const float* Xs, int rows, float* H1, float* H2, float* yhat,
                             const float* P) const {
    using namespace hivf;
    const float *W1 = P + off_W1, *b1 = P + off_b1, *W2 = P + off_W2, *b2 = P + off_b2, *w3 = P + off_w3,
                *b3 = P + off_b3;
    hipError_t e;
    if ((e = gemm(s, Xs, in, 0, W1, hid, 0, H1, hid, rows, hid, in, b1, 1, nullptr, 0)) != hipSuccess) return e;
    if ((e = gemm(s, H1, hid, 0, W2, hid, 0, H2, hid, rows, hid, hid, b2, 1, nullptr, 0)) != hipSuccess) return e;
    return gemm(s, H2, hid, 0, w3, 1, 0, yhat, 1, rows, 1, hid, b3, 0, nullptr, 0);
}

// Gradient of sum_i dy[i] yhat[i] w.r.t. the trained parameters into G.
hipError_t hivf_tau_model::backward(hipStream_t s, const float* Xs, int rows, const float* H1, const float* H2,
                              const float* dy, float* dZ2, float* dZ1, float* G) const {
    using namespace hivf;
    const float *W2 = params + off_W2, *w3 = params + off_w3;
    hipError_t e;
    // layer 3: dw3 = H2^T dy, db3 = sum dy
    if ((e = gemm(s, H2, hid, 1, dy, 1, 0, G + off_w3, 1, hid, 1, rows, nullptr, 0, nullptr, 0)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(256), 0, s, dy, rows, 1, 1, G + off_b3);
    // dZ2 = (dy w3^T) * [H2 > 0]
    if ((e = gemm(s, dy, 1, 0, w3, 1, 1, dZ2, hid, rows, hid, 1, nullptr, 0, H2, hid)) != hipSuccess) return e;
    if ((e = gemm(s, H1, hid, 1, dZ2, hid, 0, G + off_W2, hid, hid, hid, rows, nullptr, 0, nullptr, 0)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(colsum_kernel, dim3(blocks(hid, 64)), dim3(256), 0, s, dZ2, rows, hid, hid, G + off_b2);
    // dZ1 = (dZ2 W2^T) * [H1 > 0]
    if ((e = gemm(s, dZ2, hid, 0, W2, hid, 1, dZ1, hid, rows, hid, hid, nullptr, 0, H1, hid)) != hipSuccess) return e;
    if ((e = gemm(s, Xs, in, 1, dZ1, hid, 0, G + off_W1, hid, in, hid, rows, nullptr, 0, nullptr, 0)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(colsum_kernel, dim3(blocks(hid, 64)), dim3(256), 0, s, dZ1, rows, hid, hid, G + off_b1);
    return hipGetLastError();
}

void hivf_tau_model::layout() {
    off_W1 = 0;
    off_b1 = off_W1 + (int64_t)in * hid;
    off_W2 = off_b1 + hid;
    off_b2 = off_W2 + (int64_t)hid * hid;
    off_w3 = off_b2 + hid;
    off_b3 = off_w3 + hid;
    off_mu = off_b3 + 1;
    off_sd = off_mu + in;
    off_scale = off_sd + in;
    off_offset = off_scale + 1;
    n_params = off_offset + 1;
    n_trained = off_mu;
}

// Predict tau for n queries from device features F (raw, [n x in]).
hipError_t hivf_tau_model::predict(hipStream_t s, const float* F, int64_t n, float* tau, const float* P) {
    using namespace hivf;
    hipError_t e = ensure_work(n);
    if (e != hipSuccess) return e;
    float* Xs = work;
    float* H1 = Xs + n * in;
    float* H2 = H1 + n * hid;
    float* yh = H2 + n * hid;
    hipLaunchKernelGGL(gather_std_kernel, dim3(blocks(n * in)), dim3(256), 0, s, F, nullptr, nullptr, (int)n, in,
                       P + off_mu, P + off_sd, Xs, nullptr);
    if ((e = forward(s, Xs, (int)n, H1, H2, yh, P)) != hipSuccess) return e;
    hipLaunchKernelGGL(tau_out_kernel, dim3(blocks(n)), dim3(256), 0, s, yh, (int)n, host_offset, host_scale, tau);
    return hipGetLastError();
}

namespace hivf {
int run_tau_model(hivf_index* h, const float4* q, int64_t n, int P, const int32_t* npq, float* tau, hipStream_t s) {
    (void)q;
    if (!h->tau_model) return h->fail(HIVF_ESTATE, "learned bound requested but no model is attached");
    TauModel& m = *h->tau_model;
    if (m.in != kFeatCoarse + 1) return h->fail(HIVF_EUNSUPPORTED, "model input must be 33 features");
    if (m.hid <= kFusedHid && m.hid % 8 == 0) {
        FusedArgs fa{h->d_cd.p, P,        npq,      h->batched_qnorm, (int)n,   m.params, m.hid,
                     m.off_W1,  m.off_b1, m.off_W2, m.off_b2,         m.off_w3, m.off_b3, m.off_mu,
                     m.off_sd,  m.host_offset, m.host_scale, tau};
        // 8 R rows per workgroup, about two workgroups per compute unit (a
        // second wave per SIMD hides the layer loops' LDS waits: 22 us for
        // 10k queries at R = 3, 25 us at one workgroup per CU, R = 5)
        const int64_t cus = std::max(1, h->cus);
        int R = (int)std::min<int64_t>(kFusedMaxR, std::max<int64_t>(1, (n + 16 * cus - 1) / (16 * cus)));
        while (R > 1 && fused_lds_bytes(m.hid, R) > 65536) --R;  // 64 KiB per workgroup
        const dim3 grid(blocks(n, 8 * R));
        const size_t lds = fused_lds_bytes(m.hid, R);
        switch (R) {
            case 1: hipLaunchKernelGGL(tau_mlp_kernel<1>, grid, dim3(256), lds, s, fa); break;
            case 2: hipLaunchKernelGGL(tau_mlp_kernel<2>, grid, dim3(256), lds, s, fa); break;
            case 3: hipLaunchKernelGGL(tau_mlp_kernel<3>, grid, dim3(256), lds, s, fa); break;
            case 4: hipLaunchKernelGGL(tau_mlp_kernel<4>, grid, dim3(256), lds, s, fa); break;
            case 5: hipLaunchKernelGGL(tau_mlp_kernel<5>, grid, dim3(256), lds, s, fa); break;
            case 6: hipLaunchKernelGGL(tau_mlp_kernel<6>, grid, dim3(256), lds, s, fa); break;
            case 7: hipLaunchKernelGGL(tau_mlp_kernel<7>, grid, dim3(256), lds, s, fa); break;
            default: hipLaunchKernelGGL(tau_mlp_kernel<8>, grid, dim3(256), lds, s, fa); break;
        }
        HIP_TRY(h, hipGetLastError());
        return HIVF_OK;
    }
    HIP_TRY(h, h->d_feat.ensure((size_t)n * m.in));
    hipLaunchKernelGGL(features_kernel, dim3(blocks(n)), dim3(256), 0, s, h->d_cd.p, P, npq, h->batched_qnorm, (int)n,
                       h->d_feat.p);
    HIP_TRY(h, m.predict(s, h->d_feat.p, n, tau, m.params));
    return HIVF_OK;
}

}  // namespace hivf

// ---- C ABI ----------------------------------------------------------------
extern "C" {

hivf_tau_model* hivf_tau_model_create(int32_t in_dim, int32_t hidden, uint64_t seed) {
    HIVF_API_BEGIN
    if (in_dim < 1 || hidden < 1 || hidden > 4096) return nullptr;
    auto* m = new (std::nothrow) TauModel();
    if (!m) return nullptr;
    m->in = in_dim;
    m->hid = hidden;
    m->layout();
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&m->params, m->n_params * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->grads, m->n_params * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->m1, m->n_params * sizeof(float)) != hipSuccess ||
        hipMalloc(&m->m2, m->n_params * sizeof(float)) != hipSuccess) {
        delete m;
        return nullptr;
    }
    // He-uniform weights, zero biases, identity standardisation
    std::vector<float> p(m->n_params, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int64_t off, int64_t cnt, int fan_in) {
        std::uniform_real_distribution<float> u(-std::sqrt(6.0f / fan_in), std::sqrt(6.0f / fan_in));
        for (int64_t i = 0; i < cnt; ++i) p[off + i] = u(rng);
    };
    fill(m->off_W1, (int64_t)m->in * m->hid, m->in);
    fill(m->off_W2, (int64_t)m->hid * m->hid, m->hid);
    fill(m->off_w3, m->hid, m->hid);
    for (int i = 0; i < m->in; ++i) p[m->off_sd + i] = 1.0f;
    p[m->off_scale] = 1.0f;
    if (hivf_tau_model_set_params(m, p.data()) != HIVF_OK) {
        delete m;
        return nullptr;
    }
    return m;
    HIVF_API_END_PTR(nullptr)
}

void hivf_tau_model_destroy(hivf_tau_model* m) {
    HIVF_API_BEGIN
    delete m;
    HIVF_API_END_VOID
}

int64_t hivf_tau_model_num_params(const hivf_tau_model* m) { return m ? m->n_params : 0; }

int hivf_tau_model_get_params(const hivf_tau_model* m, float* out) {
    HIVF_API_BEGIN
    if (!m || !out) return HIVF_EINVAL;
    if (hipMemcpy(out, m->params, m->n_params * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return HIVF_EHIP;
    return HIVF_OK;
    HIVF_API_END(nullptr)
}

int hivf_tau_model_set_params(hivf_tau_model* m, const float* in) {
    HIVF_API_BEGIN
    if (!m || !in) return HIVF_EINVAL;
    if (hipMemcpy(m->params, in, m->n_params * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return HIVF_EHIP;
    if (!(in[m->off_scale] > 0.0f) || !(in[m->off_scale] <= FLT_MAX)) return HIVF_EINVAL;  // scale: finite, > 0
    m->host_offset = in[m->off_offset];
    m->host_scale = in[m->off_scale];
    return HIVF_OK;
    HIVF_API_END(nullptr)
}

int hivf_tau_model_forward(hivf_tau_model* m, int64_t n, const float* x, float* y) {
    HIVF_API_BEGIN
    if (!m || n < 0 || (n > 0 && (!x || !y))) return HIVF_EINVAL;
    if (n == 0) return HIVF_OK;
    float *dx = nullptr, *dy = nullptr;
    int rc = HIVF_OK;
    if (hipMalloc(&dx, n * m->in * sizeof(float)) != hipSuccess || hipMalloc(&dy, n * sizeof(float)) != hipSuccess ||
        hipMemcpy(dx, x, n * m->in * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        m->ensure_work(n) != hipSuccess) {
        rc = HIVF_EHIP;
    } else {
        float* Xs = m->work;
        float* H1 = Xs + n * m->in;
        float* H2 = H1 + n * m->hid;
        hipLaunchKernelGGL(gather_std_kernel, dim3(blocks(n * m->in)), dim3(256), 0, m->stream, dx, nullptr, nullptr,
                           (int)n, m->in, m->params + m->off_mu, m->params + m->off_sd, Xs, nullptr);
        if (m->forward(m->stream, Xs, (int)n, H1, H2, dy, m->params) != hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess ||
            hipMemcpy(y, dy, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = HIVF_EHIP;
    }
    if (dx) hivf::dev_free(dx);
    if (dy) hivf::dev_free(dy);
    return rc;
    HIVF_API_END(nullptr)
}

int hivf_tau_model_grad(hivf_tau_model* m, int64_t n, const float* x, const float* dyh, float* grad) {
    HIVF_API_BEGIN
    if (!m || n < 1 || !x || !dyh || !grad) return HIVF_EINVAL;
    float *dx = nullptr, *ddy = nullptr;
    int rc = HIVF_OK;
    if (hipMalloc(&dx, n * m->in * sizeof(float)) != hipSuccess || hipMalloc(&ddy, n * sizeof(float)) != hipSuccess ||
        hipMemcpy(dx, x, n * m->in * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ddy, dyh, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        m->ensure_work(n) != hipSuccess) {
        rc = HIVF_EHIP;
    } else {
        float* Xs = m->work;
        float* H1 = Xs + n * m->in;
        float* H2 = H1 + n * m->hid;
        float* yh = H2 + n * m->hid;
        float* dZ2 = yh + n;
        float* dZ1 = dZ2 + n * m->hid;
        hipLaunchKernelGGL(gather_std_kernel, dim3(blocks(n * m->in)), dim3(256), 0, m->stream, dx, nullptr, nullptr,
                           (int)n, m->in, m->params + m->off_mu, m->params + m->off_sd, Xs, nullptr);
        std::vector<float> zero(m->n_params, 0.0f);
        if (hipMemcpy(m->grads, zero.data(), m->n_params * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
            m->forward(m->stream, Xs, (int)n, H1, H2, yh, m->params) != hipSuccess ||
            m->backward(m->stream, Xs, (int)n, H1, H2, ddy, dZ2, dZ1, m->grads) != hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess ||
            hipMemcpy(grad, m->grads, m->n_params * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = HIVF_EHIP;
    }
    if (dx) hivf::dev_free(dx);
    if (ddy) hivf::dev_free(ddy);
    return rc;
    HIVF_API_END(nullptr)
}

int hivf_tau_model_train(hivf_tau_model* m, int64_t n, const float* x, const float* y,
                         const hivf_train_params* tp, hivf_train_stats* ts) {
    HIVF_API_BEGIN
    if (!m || n < 2 || !x || !y || !tp) return HIVF_EINVAL;
    const int in = m->in, hid = m->hid;
    // hold out a calibration split (every `holdout`-th trace)
    const int64_t hold = tp->holdout_every > 1 ? tp->holdout_every : 10;
    std::vector<int64_t> tr, ho;
    for (int64_t i = 0; i < n; ++i) (i % hold == hold - 1 ? ho : tr).push_back(i);
    if (ho.empty()) ho.push_back(n - 1);
    // standardisation from the training split
    std::vector<float> p(m->n_params);
    if (hivf_tau_model_get_params(m, p.data())) return HIVF_EHIP;
    for (int c = 0; c < in; ++c) {
        double s = 0.0, s2 = 0.0;
        for (int64_t i : tr) {
            s += x[i * in + c];
            s2 += (double)x[i * in + c] * x[i * in + c];
        }
        double mu = s / tr.size(), var = std::max(0.0, s2 / tr.size() - mu * mu);
        p[m->off_mu + c] = (float)mu;
        p[m->off_sd + c] = (float)std::max(std::sqrt(var), 1e-6);
    }
    p[m->off_offset] = 0.0f;
    p[m->off_scale] = tp->target_scale > 0.0f ? tp->target_scale : 1.0f;
    {  // start the output at the targets' mean: the network then learns the
       // per-query deviation, which is small next to log1p(k-th) itself
        double ym = 0.0;
        for (int64_t i : tr) ym += y[i];
        p[m->off_b3] = (float)(ym / tr.size());
    }
    if (hivf_tau_model_set_params(m, p.data())) return HIVF_EHIP;

    const int B = tp->batch > 0 ? tp->batch : 1024;
    float *dX = nullptr, *dY = nullptr, *dXb = nullptr, *dyb = nullptr, *dloss = nullptr, *ddy = nullptr;
    int32_t* didx = nullptr;
    hipStream_t s = m->stream;
    auto cleanup = [&]() {
        for (float* q : {dX, dY, dXb, dyb, dloss, ddy})
            if (q) hivf::dev_free(q);
        if (didx) hivf::dev_free(didx);
    };
    if (hipMalloc(&dX, n * in * sizeof(float)) || hipMalloc(&dY, n * sizeof(float)) ||
        hipMalloc(&dXb, (size_t)B * in * sizeof(float)) || hipMalloc(&dyb, B * sizeof(float)) ||
        hipMalloc(&dloss, B * sizeof(float)) || hipMalloc(&ddy, B * sizeof(float)) ||
        hipMalloc(&didx, std::max<int64_t>(n, B) * sizeof(int32_t)) ||
        hipMemcpy(dX, x, n * in * sizeof(float), hipMemcpyHostToDevice) ||
        hipMemcpy(dY, y, n * sizeof(float), hipMemcpyHostToDevice) || m->ensure_work(B)) {
        cleanup();
        return HIVF_EHIP;
    }
    std::vector<float> zero(m->n_params, 0.0f);
    if (hipMemcpy(m->m1, zero.data(), m->n_params * sizeof(float), hipMemcpyHostToDevice) ||
        hipMemcpy(m->m2, zero.data(), m->n_params * sizeof(float), hipMemcpyHostToDevice)) {
        cleanup();
        return HIVF_EHIP;
    }
    float* H1 = m->work;
    float* H2 = H1 + (size_t)B * hid;
    float* yh = H2 + (size_t)B * hid;
    float* dZ2 = yh + B;
    float* dZ1 = dZ2 + (size_t)B * hid;
    std::mt19937_64 rng(tp->seed);
    std::vector<int32_t> perm(tr.begin(), tr.end());
    int64_t step = 0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double last_loss = 0.0;
    std::vector<float> hl(B);
    for (int ep = 0; ep < tp->epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        if (hipMemcpy(didx, perm.data(), perm.size() * sizeof(int32_t), hipMemcpyHostToDevice)) {
            cleanup();
            return HIVF_EHIP;
        }
        double ep_loss = 0.0;
        for (size_t b0 = 0; b0 < perm.size(); b0 += B) {
            const int rows = (int)std::min<size_t>(B, perm.size() - b0);
            hipLaunchKernelGGL(gather_std_kernel, dim3(blocks((int64_t)rows * in)), dim3(256), 0, s, dX, dY,
                               didx + b0, rows, in, m->params + m->off_mu, m->params + m->off_sd, dXb, dyb);
            hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks(m->n_params)), dim3(256), 0, s, m->grads, m->n_params);
            if (m->forward(s, dXb, rows, H1, H2, yh, m->params) != hipSuccess) {
                cleanup();
                return HIVF_EHIP;
            }
            hipLaunchKernelGGL(pinball_kernel, dim3(blocks(rows)), dim3(256), 0, s, yh, dyb, rows, tp->quantile, ddy,
                               dloss);
            if (m->backward(s, dXb, rows, H1, H2, ddy, dZ2, dZ1, m->grads) != hipSuccess) {
                cleanup();
                return HIVF_EHIP;
            }
            ++step;
            const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
            hipLaunchKernelGGL(adam_kernel, dim3(blocks(m->n_trained)), dim3(256), 0, s, m->params, m->grads, m->m1,
                               m->m2, (int)m->n_trained, tp->lr, b1, b2, eps, c1, c2);
            if (b0 + B >= perm.size()) {  // last batch of the epoch: its mean loss
                if (hipMemcpyAsync(hl.data(), dloss, rows * sizeof(float), hipMemcpyDeviceToHost, s) ||
                    hipStreamSynchronize(s)) {
                    cleanup();
                    return HIVF_EHIP;
                }
                for (int i = 0; i < rows; ++i) ep_loss += hl[i];
                ep_loss /= rows;
            }
        }
        last_loss = ep_loss;
    }
    if (hipStreamSynchronize(s)) {
        cleanup();
        return HIVF_EHIP;
    }
    cleanup();
    // calibrate the offset on the held-out split: the smallest offset with
    // at most `miss_rate` of held-out bounds below their k-th distance
    std::vector<float> hx(ho.size() * in), hy(ho.size()), pred(ho.size());
    for (size_t i = 0; i < ho.size(); ++i) {
        std::memcpy(&hx[i * in], x + ho[i] * in, in * sizeof(float));
        hy[i] = y[ho[i]];
    }
    if (hivf_tau_model_forward(m, (int64_t)ho.size(), hx.data(), pred.data())) return HIVF_EHIP;
    std::vector<float> resid(ho.size());
    for (size_t i = 0; i < ho.size(); ++i) resid[i] = hy[i] - pred[i];
    std::sort(resid.begin(), resid.end());
    const double miss = tp->miss_rate > 0 ? tp->miss_rate : 0.01;
    size_t at = (size_t)std::min<double>(resid.size() - 1, std::ceil((1.0 - miss) * resid.size()) - 1);
    float offset = std::max(0.0f, resid[at]);
    if (hivf_tau_model_get_params(m, p.data())) return HIVF_EHIP;
    p[m->off_offset] = offset;
    if (hivf_tau_model_set_params(m, p.data())) return HIVF_EHIP;
    if (ts) {
        std::memset(ts, 0, sizeof(*ts));
        ts->final_loss = last_loss;
        ts->offset = offset;
        ts->steps = step;
        size_t over = 0;
        double err = 0.0;
        for (size_t i = 0; i < ho.size(); ++i) {
            over += (pred[i] + offset) < hy[i];
            err += std::fabs(pred[i] - hy[i]);
        }
        ts->holdout_miss_rate = (double)over / ho.size();
        ts->holdout_mean_abs_err = err / ho.size();
        ts->holdout = (int64_t)ho.size();
    }
    return HIVF_OK;
    HIVF_API_END(nullptr)
}

int hivf_index_set_tau_model(hivf_index* h, const hivf_tau_model* m) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    HIP_TRY(h, hipSetDevice(h->device));
    (void)hipStreamSynchronize(h->stream);
    h->drop_graph();  // a captured search reads the old model's weights
    if (!m) {
        h->tau_model.reset();
        return HIVF_OK;
    }
    auto c = std::make_unique<TauModel>();
    c->in = m->in;
    c->hid = m->hid;
    c->layout();
    HIP_TRY(h, hipMalloc(&c->params, c->n_params * sizeof(float)));
    HIP_TRY(h, hipMemcpy(c->params, m->params, c->n_params * sizeof(float), hipMemcpyDeviceToDevice));
    c->host_offset = m->host_offset;
    c->host_scale = m->host_scale;
    h->tau_model = std::move(c);
    return HIVF_OK;
    HIVF_API_END(h)
}

// Traces of the exact search: features of each query and its k-th distance
// (FLT_MAX when it has fewer than k candidates).
int hivf_index_tau_traces(hivf_index* h, int64_t n, const float* x, int64_t k, int64_t nprobe, float* features,
                          float* kth) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (n < 1 || !x || !features || !kth) return h->fail(HIVF_EINVAL, "bad trace arguments");
    std::vector<float> D((size_t)n * k);
    std::vector<int64_t> I((size_t)n * k);
    hivf_search_params p;
    hivf_search_params_default(&p);
    p.nprobe = nprobe;
    p.algo = HIVF_ALGO_BATCHED;
    p.tau_source = HIVF_TAU_WARMUP;
    int rc = hivf_index_search(h, n, x, k, D.data(), I.data(), &p, nullptr);
    if (rc) return rc;
    const int64_t P = std::min<int64_t>(nprobe > 0 ? nprobe : h->nprobe, h->nlist);
    if (P < 1) return h->fail(HIVF_EINVAL, "nprobe");
    // the search left the coarse distances and query norms on the device:
    // features by the same kernel the learned-bound search uses
    HIP_TRY(h, h->d_feat.ensure((size_t)n * (kFeatCoarse + 1)));
    hipLaunchKernelGGL(features_kernel, dim3(blocks(n)), dim3(256), 0, h->stream, h->d_cd.p, (int)P, nullptr,
                       h->batched_qnorm, (int)n, h->d_feat.p);
    HIP_TRY(h, hipGetLastError());
    HIP_TRY(h, hipMemcpyAsync(features, h->d_feat.p, (size_t)n * (kFeatCoarse + 1) * sizeof(float),
                              hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int64_t q = 0; q < n; ++q) kth[q] = D[q * k + k - 1];
    return HIVF_OK;
    HIVF_API_END(h)
}

int hivf_index_tau_bounds(hivf_index* h, int64_t n, const float* x, int64_t k, int64_t nprobe, float* tau) {
    HIVF_API_BEGIN
    if (!h) return HIVF_EINVAL;
    if (n < 1 || !x || !tau) return h->fail(HIVF_EINVAL, "bad bound arguments");
    if (!h->tau_model) return h->fail(HIVF_ESTATE, "no bound model attached");
    std::vector<float> D((size_t)n * k);
    std::vector<int64_t> I((size_t)n * k);
    hivf_search_params p;
    hivf_search_params_default(&p);
    p.nprobe = nprobe;
    p.algo = HIVF_ALGO_BATCHED;
    p.tau_source = HIVF_TAU_WARMUP;
    int rc = hivf_index_search(h, n, x, k, D.data(), I.data(), &p, nullptr);
    if (rc) return rc;
    const int64_t P = std::min<int64_t>(nprobe > 0 ? nprobe : h->nprobe, h->nlist);
    if (P < 1) return h->fail(HIVF_EINVAL, "nprobe");
    // the search left the coarse distances and query norms on the device
    HIP_TRY(h, h->d_tau.ensure((size_t)n));
    rc = run_tau_model(h, nullptr, n, (int)P, nullptr, h->d_tau.p, h->stream);
    if (rc) return rc;
    HIP_TRY(h, hipMemcpyAsync(tau, h->d_tau.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return HIVF_OK;
    HIVF_API_END(h)
}

}  // extern "C"
