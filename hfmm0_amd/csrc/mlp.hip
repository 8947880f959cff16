// Degree-predicting MLP: inference and training kernels for gfx950.
//
// Network: 6 inputs -> 3 hidden layers of 64 units with GELU (tanh form) -> 1 output.
// The forward pass uses BF16-rounded weights (round to nearest even) with FP32
// accumulation; training keeps FP32 master weights, accumulates the gradients of a
// minibatch with FP32 atomics and applies Adam.  One 64-lane wave handles one sample:
// lane j owns hidden unit j of every layer, the layer inputs sit in LDS.
//
// Parameter layout (float32, 8833 values): W1[64][6], b1[64], W2[64][64], b2[64],
// W3[64][64], b3[64], W4[64], b4.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/hfmm.h"
#include "common.h"

namespace hfmm {
namespace {

constexpr int kIn = HFMM_MLP_INPUTS, kH = 64, kP = HFMM_MLP_PARAMS;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kH;
static_assert(oB4 + 1 == kP, "parameter layout");

__device__ inline float bf16_round(float x) {
    unsigned u = __float_as_uint(x);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    return __uint_as_float(u);
}

__device__ inline float gelu(float z) {
    const float c = 0.7978845608028654f;   // sqrt(2/pi)
    return 0.5f * z * (1.0f + tanhf(c * (z + 0.044715f * z * z * z)));
}

__device__ inline float gelu_grad(float z) {
    const float c = 0.7978845608028654f;
    const float t = tanhf(c * (z + 0.044715f * z * z * z));
    const float du = c * (1.0f + 3.0f * 0.044715f * z * z);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

__device__ inline float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Forward pass of one sample on one wave; returns the output (all lanes) and leaves the
// pre-activations of lane j's units in z[0..2] and the activations in LDS.
__device__ inline float forward_one(const float* __restrict__ w, const float* x_in, float* sx, float* h1,
                                    float* h2, float* h3, float z[3]) {
    const int j = threadIdx.x;
    if (j < kIn) sx[j] = x_in[j];
    __syncthreads();
    float a = w[oB1 + j];
    for (int i = 0; i < kIn; ++i) a += w[oW1 + j * kIn + i] * sx[i];
    z[0] = a;
    h1[j] = gelu(a);
    __syncthreads();
    a = w[oB2 + j];
    for (int i = 0; i < kH; ++i) a += w[oW2 + j * kH + i] * h1[i];
    z[1] = a;
    h2[j] = gelu(a);
    __syncthreads();
    a = w[oB3 + j];
    for (int i = 0; i < kH; ++i) a += w[oW3 + j * kH + i] * h2[i];
    z[2] = a;
    h3[j] = gelu(a);
    __syncthreads();
    return wave_sum(w[oW4 + j] * h3[j]) + w[oB4];
}

__global__ __launch_bounds__(64) void mlp_forward_kernel(const float* wq, const float* X, int64_t n, float* y) {
    __shared__ float sx[kIn], h1[kH], h2[kH], h3[kH];
    const int64_t s = blockIdx.x;
    if (s >= n) return;
    float z[3];
    const float out = forward_one(wq, X + s * kIn, sx, h1, h2, h3, z);
    if (threadIdx.x == 0) y[s] = out;
}

// Forward + backward of one sample (index idx[blockIdx.x]); gradients of the mean
// squared error over a batch of nb samples are added into grad, the loss into loss[0].
__global__ __launch_bounds__(64) void mlp_grad_kernel(const float* wq, const float* X, const float* yt,
                                                      const int* idx, int nb, float* grad, float* loss) {
    __shared__ float sx[kIn], h1[kH], h2[kH], h3[kH], dz[kH];
    if ((int)blockIdx.x >= nb) return;
    const int64_t s = idx ? idx[blockIdx.x] : blockIdx.x;
    const int j = threadIdx.x;
    float z[3];
    const float out = forward_one(wq, X + s * kIn, sx, h1, h2, h3, z);
    const float r = out - yt[s];
    const float dy = 2.0f * r / nb;
    if (j == 0) {
        atomicAdd(loss, r * r / nb);
        atomicAdd(&grad[oB4], dy);
    }
    // layer 4
    atomicAdd(&grad[oW4 + j], dy * h3[j]);
    float d = dy * wq[oW4 + j] * gelu_grad(z[2]);
    dz[j] = d;
    atomicAdd(&grad[oB3 + j], d);
    for (int i = 0; i < kH; ++i) atomicAdd(&grad[oW3 + j * kH + i], d * h2[i]);
    __syncthreads();
    // layer 3 -> 2: lane j is input unit j of W3
    float g = 0.0f;
    for (int i = 0; i < kH; ++i) g += wq[oW3 + i * kH + j] * dz[i];
    d = g * gelu_grad(z[1]);
    __syncthreads();
    dz[j] = d;
    atomicAdd(&grad[oB2 + j], d);
    for (int i = 0; i < kH; ++i) atomicAdd(&grad[oW2 + j * kH + i], d * h1[i]);
    __syncthreads();
    g = 0.0f;
    for (int i = 0; i < kH; ++i) g += wq[oW2 + i * kH + j] * dz[i];
    d = g * gelu_grad(z[0]);
    atomicAdd(&grad[oB1 + j], d);
    for (int i = 0; i < kIn; ++i) atomicAdd(&grad[oW1 + j * kIn + i], d * sx[i]);
}

// Adam on the FP32 master weights; refreshes the BF16-rounded copy and clears grad.
__global__ void mlp_adam_kernel(float* w, float* wq, float* m, float* v, float* grad, float lr, float c1,
                                float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kP) return;
    const float g = grad[i];
    const float mi = 0.9f * m[i] + 0.1f * g;
    const float vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    w[i] = p;
    wq[i] = bf16_round(p);
    grad[i] = 0.0f;
}

__global__ void mlp_round_kernel(const float* w, float* wq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kP) wq[i] = bf16_round(w[i]);
}

}  // namespace
}  // namespace hfmm

using namespace hfmm;

struct hfmm_mlp {
    float* w = nullptr;    // FP32 master weights
    float* wq = nullptr;   // BF16-rounded copy used by the forward pass
    float* m = nullptr;
    float* v = nullptr;
    float* grad = nullptr;
    float* loss = nullptr;
    int step = 0;
    ~hfmm_mlp() {
        for (float* p : {w, wq, m, v, grad, loss})
            if (p) (void)hipFree(p);
    }
};

namespace {

#define MLPCHK(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) throw Error{HFMM_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

template <typename F>
int mlp_guard(F&& f) {
    try {
        f();
        return HFMM_OK;
    } catch (const Error& e) {
        set_last_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return HFMM_ERR_INTERNAL;
    }
}

struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) { MLPCHK(hipMalloc(&p, bytes ? bytes : 1)); }
    ~DevBuf() { (void)hipFree(p); }
    template <typename T>
    T* as() { return (T*)p; }
};

void check_kernel(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error{HFMM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)};
}

}  // namespace

extern "C" {

int hfmm_mlp_create(const float* params, hfmm_mlp** out) {
    return mlp_guard([&] {
        if (!params || !out) throw Error{HFMM_ERR_ARG, "null argument"};
        *out = nullptr;
        hfmm_mlp* M = new hfmm_mlp();
        try {
            for (float** p : {&M->w, &M->wq, &M->m, &M->v, &M->grad}) {
                MLPCHK(hipMalloc((void**)p, kP * sizeof(float)));
                MLPCHK(hipMemset(*p, 0, kP * sizeof(float)));
            }
            MLPCHK(hipMalloc((void**)&M->loss, sizeof(float)));
            MLPCHK(hipMemcpy(M->w, params, kP * sizeof(float), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(mlp_round_kernel, dim3((kP + 255) / 256), dim3(256), 0, 0, M->w, M->wq);
            check_kernel("mlp_round");
            MLPCHK(hipDeviceSynchronize());
        } catch (...) {
            delete M;
            throw;
        }
        *out = M;
    });
}

int hfmm_mlp_params(const hfmm_mlp* M, float* out) {
    return mlp_guard([&] {
        if (!M || !out) throw Error{HFMM_ERR_ARG, "null argument"};
        MLPCHK(hipMemcpy(out, M->w, kP * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int hfmm_mlp_forward(hfmm_mlp* M, const float* X, int64_t n, float* y) {
    return mlp_guard([&] {
        if (!M || !X || !y || n < 0) throw Error{HFMM_ERR_ARG, "bad argument"};
        if (n == 0) return;
        DevBuf dx(n * kIn * sizeof(float)), dy(n * sizeof(float));
        MLPCHK(hipMemcpy(dx.p, X, n * kIn * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)n), dim3(64), 0, 0, M->wq, dx.as<float>(), n,
                           dy.as<float>());
        check_kernel("mlp_forward");
        MLPCHK(hipMemcpy(y, dy.p, n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int hfmm_mlp_gradient(hfmm_mlp* M, const float* X, const float* y, int64_t n, float* grad, float* loss) {
    return mlp_guard([&] {
        if (!M || !X || !y || !grad || !loss || n <= 0 || n > (1 << 30)) throw Error{HFMM_ERR_ARG, "bad argument"};
        DevBuf dx(n * kIn * sizeof(float)), dyt(n * sizeof(float));
        MLPCHK(hipMemcpy(dx.p, X, n * kIn * sizeof(float), hipMemcpyHostToDevice));
        MLPCHK(hipMemcpy(dyt.p, y, n * sizeof(float), hipMemcpyHostToDevice));
        MLPCHK(hipMemset(M->grad, 0, kP * sizeof(float)));
        MLPCHK(hipMemset(M->loss, 0, sizeof(float)));
        hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)n), dim3(64), 0, 0, M->wq, dx.as<float>(), dyt.as<float>(),
                           (const int*)nullptr, (int)n, M->grad, M->loss);
        check_kernel("mlp_grad");
        MLPCHK(hipMemcpy(grad, M->grad, kP * sizeof(float), hipMemcpyDeviceToHost));
        MLPCHK(hipMemcpy(loss, M->loss, sizeof(float), hipMemcpyDeviceToHost));
        MLPCHK(hipMemset(M->grad, 0, kP * sizeof(float)));
    });
}

int hfmm_mlp_train(hfmm_mlp* M, const float* X, const float* y, int64_t n, int epochs, int batch, float lr,
                   int seed, float* loss_per_epoch) {
    return mlp_guard([&] {
        if (!M || !X || !y || n <= 0 || epochs < 1 || batch < 1 || !(lr > 0) || n > (1 << 30))
            throw Error{HFMM_ERR_ARG, "bad argument"};
        DevBuf dx(n * kIn * sizeof(float)), dyt(n * sizeof(float)), didx(n * sizeof(int));
        MLPCHK(hipMemcpy(dx.p, X, n * kIn * sizeof(float), hipMemcpyHostToDevice));
        MLPCHK(hipMemcpy(dyt.p, y, n * sizeof(float), hipMemcpyHostToDevice));
        MLPCHK(hipMemset(M->grad, 0, kP * sizeof(float)));
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::mt19937 rng((unsigned)seed);
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(order.begin(), order.end(), rng);
            MLPCHK(hipMemcpy(didx.p, order.data(), n * sizeof(int), hipMemcpyHostToDevice));
            MLPCHK(hipMemset(M->loss, 0, sizeof(float)));
            int nbatches = 0;
            for (int64_t b0 = 0; b0 < n; b0 += batch) {
                const int nb = (int)std::min<int64_t>(batch, n - b0);
                hipLaunchKernelGGL(mlp_grad_kernel, dim3(nb), dim3(64), 0, 0, M->wq, dx.as<float>(), dyt.as<float>(),
                                   didx.as<int>() + b0, nb, M->grad, M->loss);
                check_kernel("mlp_grad");
                ++M->step;
                const float c1 = 1.0f - std::pow(0.9f, (float)M->step);
                const float c2 = 1.0f - std::pow(0.999f, (float)M->step);
                hipLaunchKernelGGL(mlp_adam_kernel, dim3((kP + 255) / 256), dim3(256), 0, 0, M->w, M->wq, M->m, M->v,
                                   M->grad, lr, c1, c2);
                check_kernel("mlp_adam");
                ++nbatches;
            }
            float l = 0;
            MLPCHK(hipMemcpy(&l, M->loss, sizeof(float), hipMemcpyDeviceToHost));
            if (loss_per_epoch) loss_per_epoch[ep] = l / nbatches;
        }
        MLPCHK(hipDeviceSynchronize());
    });
}

int hfmm_mlp_destroy(hfmm_mlp* M) {
    return mlp_guard([&] { delete M; });
}

}  // extern "C"
