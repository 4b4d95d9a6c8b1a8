// hhg_mlp.hip — training and evaluation of the step-proposal network on the
// GPU (FP32 master weights, BF16 weights and activations in every forward
// pass, FP32 accumulation), and its C ABI.
//
// A training step on a batch of n samples:
//   fwd_bwd_kernel   one wave per sample, one lane per hidden unit: forward pass
//                    with the BF16-rounded weights (staged in LDS) and
//                    BF16-rounded activations, loss
//                    0.5 |y - t|^2 averaged over the batch, backward pass with
//                    the straight-through derivative of the BF16 rounding;
//                    activations and deltas of every sample go to a work array
//   grad_partial_kernel / grad_reduce_kernel
//                    the batch sum of delta x activation per parameter, in
//                    fixed 64-sample segments then in segment order
//                    (deterministic, no atomics)
//   adam_kernel      Adam update of the FP32 masters
//   refresh_kernel   BF16 roundings (FP32 copy) and the MFMA operand packing
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/hhg_amd.h"
#include "hhg_host.hpp"
#include "hhg_mlp.hpp"

using hhg::fail;

namespace hhg {
namespace {

constexpr int kRec = 400;  // floats per sample in the work array
constexpr int kA0 = 0, kA1 = 8, kA2 = 72, kA3 = 136, kD1 = 200, kD2 = 264, kD3 = 328, kDY = 392, kLoss = 394;

__device__ __forceinline__ float bf16r(float x) { return (float)(__bf16)x; }

__global__ void refresh_kernel(const float* __restrict__ params, float* __restrict__ params_bf,
                               __bf16* __restrict__ pack) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < kMlpParams) params_bf[idx] = bf16r(params[idx]);
    if (idx < kPackTotal) {
        const int tile = idx / kPackTile, rem = idx % kPackTile;
        const int lane = rem / 8, i = rem % 8;
        const int m = lane & 31, h = lane >> 5;
        float w = 0.0f;
        if (tile < kTileL2) {
            // layer 1 [nt][mt]: features in lane half nt, bias at k = 0 of the other half
            const int nt = tile / 2, row = 32 * (tile % 2) + m;
            w = h == nt ? params[kOffW1 + row * kMlpIn + i] : (i == 0 ? params[kOffB1 + row] : 0.0f);
        } else if (tile < kTileL4) {
            const int t = (tile - kTileL2) % 8, mt = t / 4, ks = t % 4;
            const int off = tile < kTileL3 ? kOffW2 : kOffW3;
            w = params[off + (32 * mt + m) * kMlpHidden + hidden_unit(ks, h, i)];
        } else {
            // output layer [nt][ks]: outputs on rows 4 nt and 4 nt + 1
            const int nt = (tile - kTileL4) / 4, ks = (tile - kTileL4) % 4;
            const int o = m - 4 * nt;
            if (o == 0 || o == 1) w = params[kOffW4 + o * kMlpHidden + hidden_unit(ks, h, i)];
        }
        pack[idx] = (__bf16)w;
    }
}

// LDS copy of the BF16-rounded parameters for the training passes; the rows
// of W1, W2 and W3 are padded by one float so that both the forward (lane =
// output unit) and the backward (lane = input unit) reads are free of bank
// conflicts.
constexpr int kS1 = kMlpIn + 1, kS2 = kMlpHidden + 1;
struct TrainLds {
    float W1[kMlpHidden * kS1], b1[kMlpHidden];
    float W2[kMlpHidden * kS2], b2[kMlpHidden];
    float W3[kMlpHidden * kS2], b3[kMlpHidden];
    float W4[kMlpOut * kMlpHidden], b4[kMlpOut];
    float act[4][4][kMlpHidden];  // per wave: a0 (first 8), a1, a2, a3 / deltas
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One wave per sample, lane u = hidden unit u; W: the BF16-rounded parameters.
// x, t raw (t may be null: forward only, y written de-normalised to y_out).
// Every sample's activations and deltas go to its work row for the gradient sums.
__global__ __launch_bounds__(256) void fwd_bwd_kernel(const float* __restrict__ Wg, const float* __restrict__ x,
                                                      const float* __restrict__ t, const int* __restrict__ idx,
                                                      int n, MlpDev norm, float inv_n, float* __restrict__ work,
                                                      float* __restrict__ y_out) {
    __shared__ TrainLds L;
    for (int p = threadIdx.x; p < kMlpHidden * kMlpIn; p += blockDim.x)
        L.W1[(p / kMlpIn) * kS1 + p % kMlpIn] = Wg[kOffW1 + p];
    for (int p = threadIdx.x; p < kMlpHidden * kMlpHidden; p += blockDim.x) {
        L.W2[(p / kMlpHidden) * kS2 + p % kMlpHidden] = Wg[kOffW2 + p];
        L.W3[(p / kMlpHidden) * kS2 + p % kMlpHidden] = Wg[kOffW3 + p];
    }
    for (int p = threadIdx.x; p < kMlpHidden; p += blockDim.x) {
        L.b1[p] = Wg[kOffB1 + p];
        L.b2[p] = Wg[kOffB2 + p];
        L.b3[p] = Wg[kOffB3 + p];
    }
    for (int p = threadIdx.x; p < kMlpOut * kMlpHidden; p += blockDim.x) L.W4[p] = Wg[kOffW4 + p];
    if (threadIdx.x < kMlpOut) L.b4[threadIdx.x] = Wg[kOffB4 + threadIdx.x];
    __syncthreads();

    const int wv = threadIdx.x / 64, u = threadIdx.x & 63;
    float(*A)[kMlpHidden] = L.act[wv];
    for (int b = blockIdx.x * 4 + wv; b < n; b += gridDim.x * 4) {
        const int s = idx ? idx[b] : b;
        float* R = work + (size_t)b * kRec;  // used only with targets
        if (u < kMlpIn) {
            const float a0 = bf16r(mlp_norm_input(x[(size_t)s * kMlpIn + u], norm.in_mean[u], norm.in_istd[u]));
            A[0][u] = a0;
            if (t) R[kA0 + u] = a0;
        }
        wave_lds_sync();
        // forward
        float z = L.b1[u];
        for (int i = 0; i < kMlpIn; ++i) z = fmaf(L.W1[u * kS1 + i], A[0][i], z);
        const float a1 = bf16r(z > 0.0f ? z : 0.0f);
        A[1][u] = a1;
        wave_lds_sync();
        z = L.b2[u];
        for (int v = 0; v < kMlpHidden; ++v) z = fmaf(L.W2[u * kS2 + v], A[1][v], z);
        const float a2 = bf16r(z > 0.0f ? z : 0.0f);
        A[2][u] = a2;
        wave_lds_sync();
        z = L.b3[u];
        for (int v = 0; v < kMlpHidden; ++v) z = fmaf(L.W3[u * kS2 + v], A[2][v], z);
        const float a3 = bf16r(z > 0.0f ? z : 0.0f);
        float y[kMlpOut];
        for (int o = 0; o < kMlpOut; ++o) y[o] = wave_sum_f32(L.W4[o * kMlpHidden + u] * a3) + L.b4[o];
        if (!t) {
            if (u < kMlpOut) y_out[(size_t)b * kMlpOut + u] = y[u] * norm.out_std[u] + norm.out_mean[u];
            wave_lds_sync();
            continue;
        }
        R[kA1 + u] = a1;
        R[kA2 + u] = a2;
        R[kA3 + u] = a3;
        // loss and backward
        float loss = 0.0f, dy[kMlpOut];
        for (int o = 0; o < kMlpOut; ++o) {
            const float tn = (t[(size_t)s * kMlpOut + o] - norm.out_mean[o]) / norm.out_std[o];
            const float e = y[o] - tn;
            loss += 0.5f * e * e;
            dy[o] = e * inv_n;
        }
        if (u < kMlpOut) R[kDY + u] = dy[u];
        if (u == 0) R[kLoss] = loss;
        float g = 0.0f;
        for (int o = 0; o < kMlpOut; ++o) g = fmaf(L.W4[o * kMlpHidden + u], dy[o], g);
        const float d3 = a3 > 0.0f ? g : 0.0f;
        R[kD3 + u] = d3;
        A[3][u] = d3;
        wave_lds_sync();
        g = 0.0f;
        for (int w = 0; w < kMlpHidden; ++w) g = fmaf(L.W3[w * kS2 + u], A[3][w], g);
        const float d2 = a2 > 0.0f ? g : 0.0f;
        R[kD2 + u] = d2;
        A[0][u] = d2;  // a0 is no longer needed
        wave_lds_sync();
        g = 0.0f;
        for (int w = 0; w < kMlpHidden; ++w) g = fmaf(L.W2[w * kS2 + u], A[0][w], g);
        R[kD1 + u] = a1 > 0.0f ? g : 0.0f;
        wave_lds_sync();
    }
}

// blocks of fwd_bwd_kernel for n samples (4 per block pass, grid-stride beyond)
inline unsigned fwd_bwd_blocks(int64_t n) { return (unsigned)std::min<int64_t>((n + 3) / 4, 4096); }

constexpr int kGradSeg = 64;  // samples per partial sum of the gradient

// thread per (parameter, segment of kGradSeg samples): the segment's sum in
// sample order; index kMlpParams carries the loss
__global__ void grad_partial_kernel(const float* __restrict__ work, int n, float* __restrict__ partial) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    const int b0 = blockIdx.y * kGradSeg, b1 = min(n, b0 + kGradSeg);
    float acc = 0.0f;
    if (p == kMlpParams) {
        for (int b = b0; b < b1; ++b) acc += work[(size_t)b * kRec + kLoss];
    } else {
        int dofs, aofs = -1;
        if (p < kOffB1) dofs = kD1 + (p - kOffW1) / kMlpIn, aofs = kA0 + (p - kOffW1) % kMlpIn;
        else if (p < kOffW2) dofs = kD1 + (p - kOffB1);
        else if (p < kOffB2) dofs = kD2 + (p - kOffW2) / kMlpHidden, aofs = kA1 + (p - kOffW2) % kMlpHidden;
        else if (p < kOffW3) dofs = kD2 + (p - kOffB2);
        else if (p < kOffB3) dofs = kD3 + (p - kOffW3) / kMlpHidden, aofs = kA2 + (p - kOffW3) % kMlpHidden;
        else if (p < kOffW4) dofs = kD3 + (p - kOffB3);
        else if (p < kOffB4) dofs = kDY + (p - kOffW4) / kMlpHidden, aofs = kA3 + (p - kOffW4) % kMlpHidden;
        else dofs = kDY + (p - kOffB4);
        for (int b = b0; b < b1; ++b) {
            const float* R = work + (size_t)b * kRec;
            acc = aofs >= 0 ? fmaf(R[dofs], R[aofs], acc) : acc + R[dofs];
        }
    }
    partial[(size_t)blockIdx.y * (kMlpParams + 1) + p] = acc;
}

// thread per parameter: the segment sums in segment order (deterministic for
// a given batch size); the loss goes to loss_out as the batch mean
__global__ void grad_reduce_kernel(const float* __restrict__ partial, int n_seg, int n, float* __restrict__ grad,
                                   float* __restrict__ loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    float acc = 0.0f;
    for (int q = 0; q < n_seg; ++q) acc += partial[(size_t)q * (kMlpParams + 1) + p];
    if (p == kMlpParams) *loss_out = acc / (float)n;
    else grad[p] = acc;
}

__global__ void adam_kernel(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = grad[p];
    m[p] = b1 * m[p] + (1.0f - b1) * g;
    v[p] = b2 * v[p] + (1.0f - b2) * g * g;
    params[p] -= lr * (m[p] * c1) / (sqrtf(v[p] * c2) + eps);
}

extern __shared__ __align__(16) unsigned char mfma_fwd_lds[];

// The solve kernel's wave-level MFMA inference on a feature array: one wave
// per 64 samples (padding lanes take zero features and write nothing).
__global__ __launch_bounds__(256) void mfma_forward_kernel(const MlpDev* M, const float* __restrict__ x, int64_t n,
                                                           float* __restrict__ y) {
    mlp_stage_lds(M, mfma_fwd_lds);
    __syncthreads();
    const int w = threadIdx.x / 64, lane = threadIdx.x & 63;
    const int64_t s = ((int64_t)blockIdx.x * 4 + w) * 64 + lane;
    float feat[kMlpIn], out[kMlpOut];
    for (int i = 0; i < kMlpIn; ++i) feat[i] = s < n ? x[s * kMlpIn + i] : 0.0f;
    mlp_infer_wave(M, mfma_fwd_lds, feat, out);
    if (s < n)
        for (int o = 0; o < kMlpOut; ++o) y[s * kMlpOut + o] = out[o];
}

}  // namespace
}  // namespace hhg

struct hhg_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *d_params = nullptr, *d_params_bf = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
    __bf16* d_pack = nullptr;
    hhg::MlpDev* d_dev = nullptr;
    hhg::MlpDev host_dev{};
    float* d_loss = nullptr;
    int64_t adam_t = 0;
    // training scratch
    float *d_x = nullptr, *d_t = nullptr, *d_work = nullptr, *d_y = nullptr, *d_partial = nullptr;
    float* d_loss_hist = nullptr;  // per-step losses of one epoch
    int64_t cap_hist = 0;
    int* d_idx = nullptr;
    int64_t cap_data = 0, cap_batch = 0;
};

namespace {

hhg_status chk(hipError_t e, const char* what) {
    if (e == hipSuccess) return HHG_SUCCESS;
    return fail(e == hipErrorOutOfMemory ? HHG_ERROR_OUT_OF_MEMORY : HHG_ERROR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define MLP_TRY(expr, what)                               \
    do {                                                  \
        const hhg_status _s = chk((expr), (what));        \
        if (_s != HHG_SUCCESS) return _s;                 \
    } while (0)

hhg_status refresh(hhg_mlp* m) {
    const int n = std::max(hhg::kMlpParams, hhg::kPackTotal);
    hipLaunchKernelGGL(hhg::refresh_kernel, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->d_params,
                       m->d_params_bf, m->d_pack);
    MLP_TRY(hipGetLastError(), "refresh");
    m->host_dev.pack = m->d_pack;
    m->host_dev.params = m->d_params_bf;
    MLP_TRY(hipMemcpyAsync(m->d_dev, &m->host_dev, sizeof(hhg::MlpDev), hipMemcpyHostToDevice, m->stream),
            "upload network record");
    MLP_TRY(hipStreamSynchronize(m->stream), "refresh sync");
    return HHG_SUCCESS;
}

hhg_status ensure_data(hhg_mlp* m, int64_t n) {
    if (n <= m->cap_data) return HHG_SUCCESS;
    hipFree(m->d_x);
    hipFree(m->d_t);
    hipFree(m->d_idx);
    hipFree(m->d_y);
    m->d_x = m->d_t = m->d_y = nullptr;
    m->d_idx = nullptr;
    m->cap_data = 0;
    MLP_TRY(hipMalloc(&m->d_x, sizeof(float) * hhg::kMlpIn * n), "x");
    MLP_TRY(hipMalloc(&m->d_t, sizeof(float) * hhg::kMlpOut * n), "t");
    MLP_TRY(hipMalloc(&m->d_y, sizeof(float) * hhg::kMlpOut * n), "y");
    MLP_TRY(hipMalloc(&m->d_idx, sizeof(int) * n), "index");
    m->cap_data = n;
    return HHG_SUCCESS;
}

hhg_status ensure_batch(hhg_mlp* m, int64_t n) {
    if (n <= m->cap_batch) return HHG_SUCCESS;
    hipFree(m->d_work);
    hipFree(m->d_partial);
    m->d_work = nullptr;
    m->d_partial = nullptr;
    m->cap_batch = 0;
    MLP_TRY(hipMalloc(&m->d_work, sizeof(float) * hhg::kRec * n), "work");
    const int64_t n_seg = (n + hhg::kGradSeg - 1) / hhg::kGradSeg;
    MLP_TRY(hipMalloc(&m->d_partial, sizeof(float) * (hhg::kMlpParams + 1) * n_seg), "gradient partial sums");
    m->cap_batch = n;
    return HHG_SUCCESS;
}

// gradient of the mean loss over samples idx[0..n) (or 0..n) of the uploaded data
// One forward/backward pass over a batch: gradient to d_grad, mean loss to *loss_dst.
hhg_status gradient_step(hhg_mlp* m, const int* d_idx, int n, float* loss_dst) {
    const hhg_status eb = ensure_batch(m, n);
    if (eb != HHG_SUCCESS) return eb;
    hipLaunchKernelGGL(hhg::fwd_bwd_kernel, dim3(hhg::fwd_bwd_blocks(n)), dim3(256), 0, m->stream, m->d_params_bf, m->d_x,
                       m->d_t, d_idx, n, m->host_dev, 1.0f / (float)n, m->d_work, nullptr);
    MLP_TRY(hipGetLastError(), "forward/backward");
    const int n_seg = (n + hhg::kGradSeg - 1) / hhg::kGradSeg;
    const int blocks = (hhg::kMlpParams + 1 + 255) / 256;
    hipLaunchKernelGGL(hhg::grad_partial_kernel, dim3(blocks, n_seg), dim3(256), 0, m->stream, m->d_work, n,
                       m->d_partial);
    MLP_TRY(hipGetLastError(), "gradient partial sums");
    hipLaunchKernelGGL(hhg::grad_reduce_kernel, dim3(blocks), dim3(256), 0, m->stream, m->d_partial, n_seg, n,
                       m->d_grad, loss_dst);
    MLP_TRY(hipGetLastError(), "gradient");
    return HHG_SUCCESS;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

namespace hhg {
const MlpDev* mlp_device_record(const hhg_mlp* m) { return m ? m->d_dev : nullptr; }
int mlp_device(const hhg_mlp* m) { return m ? m->device : -1; }
}  // namespace hhg

extern "C" {

int64_t hhg_mlp_param_count(void) { return hhg::kMlpParams; }

hhg_status hhg_mlp_create(int32_t device, uint64_t seed, hhg_mlp** out) {
    hhg::clear_error();
    if (!out) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_create: null argument");
    *out = nullptr;
    int n_dev = 0;
    if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev < 1) return fail(HHG_ERROR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= n_dev) return fail(HHG_ERROR_INVALID_ARGUMENT, "device ordinal out of range");
    hhg_mlp* m = new hhg_mlp();
    m->device = device;
    auto bail = [&](hhg_status st) {
        hhg_mlp_destroy(m);
        return st;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(fail(HHG_ERROR_HIP, "hipSetDevice"));
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(HHG_ERROR_HIP, "stream"));
    const size_t P = sizeof(float) * hhg::kMlpParams;
    hhg_status st = HHG_SUCCESS;
    for (float** p : {&m->d_params, &m->d_params_bf, &m->d_grad, &m->d_m, &m->d_v})
        if (st == HHG_SUCCESS) st = chk(hipMalloc(p, P), "parameters");
    if (st == HHG_SUCCESS) st = chk(hipMalloc(&m->d_pack, sizeof(__bf16) * hhg::kPackTotal), "packed weights");
    if (st == HHG_SUCCESS) st = chk(hipMalloc(&m->d_dev, sizeof(hhg::MlpDev)), "network record");
    if (st == HHG_SUCCESS) st = chk(hipMalloc(&m->d_loss, sizeof(float)), "loss");
    if (st != HHG_SUCCESS) return bail(st);
    // He-uniform initialisation from a splitmix64 stream; zero biases
    std::vector<float> w(hhg::kMlpParams, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int rows, int cols) {
        const float bound = std::sqrt(6.0f / (float)cols);
        for (int i = 0; i < rows * cols; ++i) {
            const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            w[off + i] = (float)((2.0 * u - 1.0) * bound);
        }
    };
    fill(hhg::kOffW1, hhg::kMlpHidden, hhg::kMlpIn);
    fill(hhg::kOffW2, hhg::kMlpHidden, hhg::kMlpHidden);
    fill(hhg::kOffW3, hhg::kMlpHidden, hhg::kMlpHidden);
    fill(hhg::kOffW4, hhg::kMlpOut, hhg::kMlpHidden);
    for (int i = 0; i < hhg::kMlpIn; ++i) m->host_dev.in_mean[i] = 0.0f, m->host_dev.in_istd[i] = 1.0f;
    for (int i = 0; i < hhg::kMlpOut; ++i) m->host_dev.out_mean[i] = 0.0f, m->host_dev.out_std[i] = 1.0f;
    std::vector<float> zeros(hhg::kMlpParams, 0.0f);
    if ((st = chk(hipMemcpy(m->d_params, w.data(), P, hipMemcpyHostToDevice), "init")) != HHG_SUCCESS ||
        (st = chk(hipMemcpy(m->d_m, zeros.data(), P, hipMemcpyHostToDevice), "init")) != HHG_SUCCESS ||
        (st = chk(hipMemcpy(m->d_v, zeros.data(), P, hipMemcpyHostToDevice), "init")) != HHG_SUCCESS)
        return bail(st);
    if ((st = refresh(m)) != HHG_SUCCESS) return bail(st);
    *out = m;
    return HHG_SUCCESS;
}

hhg_status hhg_mlp_destroy(hhg_mlp* m) {
    if (!m) return HHG_SUCCESS;
    hipSetDevice(m->device);
    for (void* p : {(void*)m->d_params, (void*)m->d_params_bf, (void*)m->d_grad, (void*)m->d_m, (void*)m->d_v,
                    (void*)m->d_pack, (void*)m->d_dev, (void*)m->d_loss, (void*)m->d_x, (void*)m->d_t,
                    (void*)m->d_work, (void*)m->d_y, (void*)m->d_idx, (void*)m->d_partial, (void*)m->d_loss_hist})
        hipFree(p);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
    return HHG_SUCCESS;
}

hhg_status hhg_mlp_set_normalization(hhg_mlp* m, const float* in_mean, const float* in_std, const float* out_mean,
                                     const float* out_std) {
    hhg::clear_error();
    if (!m || !in_mean || !in_std || !out_mean || !out_std)
        return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_set_normalization: null argument");
    for (int i = 0; i < hhg::kMlpIn; ++i) {
        if (!(in_std[i] > 0.0f)) return fail(HHG_ERROR_INVALID_ARGUMENT, "input std must be > 0");
        m->host_dev.in_mean[i] = in_mean[i];
        m->host_dev.in_istd[i] = 1.0f / in_std[i];
    }
    for (int i = 0; i < hhg::kMlpOut; ++i) {
        if (!(out_std[i] > 0.0f)) return fail(HHG_ERROR_INVALID_ARGUMENT, "output std must be > 0");
        m->host_dev.out_mean[i] = out_mean[i];
        m->host_dev.out_std[i] = out_std[i];
    }
    MLP_TRY(hipSetDevice(m->device), "hipSetDevice");
    return refresh(m);
}

hhg_status hhg_mlp_get_params(const hhg_mlp* m, float* params) {
    hhg::clear_error();
    if (!m || !params) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_get_params: null argument");
    MLP_TRY(hipSetDevice(m->device), "hipSetDevice");
    MLP_TRY(hipMemcpy(params, m->d_params, sizeof(float) * hhg::kMlpParams, hipMemcpyDeviceToHost), "download");
    return HHG_SUCCESS;
}

hhg_status hhg_mlp_set_params(hhg_mlp* m, const float* params) {
    hhg::clear_error();
    if (!m || !params) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_set_params: null argument");
    MLP_TRY(hipSetDevice(m->device), "hipSetDevice");
    MLP_TRY(hipMemcpy(m->d_params, params, sizeof(float) * hhg::kMlpParams, hipMemcpyHostToDevice), "upload");
    return refresh(m);
}

hhg_status hhg_mlp_forward(hhg_mlp* m, const float* x, int64_t n, float* y) {
    hhg::clear_error();
    if (!m || (n > 0 && (!x || !y))) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_forward: null argument");
    if (n <= 0) return HHG_SUCCESS;
    if (n > (1 << 26)) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_forward: n too large");
    MLP_TRY(hipSetDevice(m->device), "hipSetDevice");
    hhg_status st = ensure_data(m, n);
    if (st != HHG_SUCCESS) return st;
    MLP_TRY(hipMemcpy(m->d_x, x, sizeof(float) * hhg::kMlpIn * n, hipMemcpyHostToDevice), "upload x");
    hipLaunchKernelGGL(hhg::fwd_bwd_kernel, dim3(hhg::fwd_bwd_blocks(n)), dim3(256), 0, m->stream, m->d_params_bf,
                       m->d_x, nullptr, nullptr, (int)n, m->host_dev, 1.0f, nullptr, m->d_y);
    MLP_TRY(hipGetLastError(), "forward");
    MLP_TRY(hipStreamSynchronize(m->stream), "forward sync");
    MLP_TRY(hipMemcpy(y, m->d_y, sizeof(float) * hhg::kMlpOut * n, hipMemcpyDeviceToHost), "download y");
    return HHG_SUCCESS;
}

hhg_status hhg_mlp_forward_mfma(hhg_mlp* m, const float* x, int64_t n, float* y) {
    hhg::clear_error();
    if (!m || (n > 0 && (!x || !y))) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_forward_mfma: null argument");
    if (n <= 0) return HHG_SUCCESS;
    if (n > (1 << 26)) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_forward_mfma: n too large");
    MLP_TRY(hipSetDevice(m->device), "hipSetDevice");
    hhg_status st = ensure_data(m, n);
    if (st != HHG_SUCCESS) return st;
    MLP_TRY(hipMemcpy(m->d_x, x, sizeof(float) * hhg::kMlpIn * n, hipMemcpyHostToDevice), "upload x");
    const int blocks = (int)((n + 255) / 256);
    const size_t smem = hhg::kMlpLdsBytes;
    hipLaunchKernelGGL(hhg::mfma_forward_kernel, dim3(blocks), dim3(256), smem, m->stream, m->d_dev, m->d_x, n, m->d_y);
    MLP_TRY(hipGetLastError(), "mfma forward");
    MLP_TRY(hipStreamSynchronize(m->stream), "mfma forward sync");
    MLP_TRY(hipMemcpy(y, m->d_y, sizeof(float) * hhg::kMlpOut * n, hipMemcpyDeviceToHost), "download y");
    return HHG_SUCCESS;
}

hhg_status hhg_mlp_gradients(hhg_mlp* m, const float* x, const float* t, int64_t n, float* grad, float* loss) {
    hhg::clear_error();
    if (!m || n <= 0 || !x || !t || !grad) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_gradients: bad argument");
    if (n > (1 << 20)) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_gradients: n too large");
    MLP_TRY(hipSetDevice(m->device), "hipSetDevice");
    hhg_status st = ensure_data(m, n);
    if (st != HHG_SUCCESS) return st;
    MLP_TRY(hipMemcpy(m->d_x, x, sizeof(float) * hhg::kMlpIn * n, hipMemcpyHostToDevice), "upload x");
    MLP_TRY(hipMemcpy(m->d_t, t, sizeof(float) * hhg::kMlpOut * n, hipMemcpyHostToDevice), "upload t");
    if ((st = gradient_step(m, nullptr, (int)n, m->d_loss)) != HHG_SUCCESS) return st;
    MLP_TRY(hipStreamSynchronize(m->stream), "gradient sync");
    MLP_TRY(hipMemcpy(grad, m->d_grad, sizeof(float) * hhg::kMlpParams, hipMemcpyDeviceToHost), "download");
    if (loss) MLP_TRY(hipMemcpy(loss, m->d_loss, sizeof(float), hipMemcpyDeviceToHost), "download loss");
    return HHG_SUCCESS;
}

hhg_status hhg_mlp_train(hhg_mlp* m, const float* x, const float* t, int64_t n, int32_t epochs, int32_t batch,
                         float lr, uint64_t seed, float* epoch_loss) {
    hhg::clear_error();
    if (!m || n <= 0 || !x || !t || epochs < 1 || batch < 1 || !(lr > 0.0f))
        return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_train: bad argument");
    if (n > (1 << 26) || batch > (1 << 16)) return fail(HHG_ERROR_INVALID_ARGUMENT, "hhg_mlp_train: too large");
    MLP_TRY(hipSetDevice(m->device), "hipSetDevice");
    hhg_status st = ensure_data(m, n);
    if (st != HHG_SUCCESS) return st;
    MLP_TRY(hipMemcpy(m->d_x, x, sizeof(float) * hhg::kMlpIn * n, hipMemcpyHostToDevice), "upload x");
    MLP_TRY(hipMemcpy(m->d_t, t, sizeof(float) * hhg::kMlpOut * n, hipMemcpyHostToDevice), "upload t");
    const int64_t steps_per_epoch = (n + batch - 1) / batch;
    if (steps_per_epoch > m->cap_hist) {
        hipFree(m->d_loss_hist);
        m->d_loss_hist = nullptr;
        m->cap_hist = 0;
        MLP_TRY(hipMalloc(&m->d_loss_hist, sizeof(float) * steps_per_epoch), "loss history");
        m->cap_hist = steps_per_epoch;
    }
    std::vector<int> order(n);
    std::vector<float> hist(steps_per_epoch);
    uint64_t s = seed;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int nr = std::max(hhg::kMlpParams, hhg::kPackTotal);
    for (int e = 0; e < epochs; ++e) {
        std::iota(order.begin(), order.end(), 0);
        for (int64_t i = n - 1; i > 0; --i) std::swap(order[i], order[splitmix(s) % (uint64_t)(i + 1)]);
        MLP_TRY(hipMemcpyAsync(m->d_idx, order.data(), sizeof(int) * n, hipMemcpyHostToDevice, m->stream), "order");
        int64_t step = 0;
        for (int64_t b0 = 0; b0 < n; b0 += batch, ++step) {
            const int nb = (int)std::min<int64_t>(batch, n - b0);
            if ((st = gradient_step(m, m->d_idx + b0, nb, m->d_loss_hist + step)) != HHG_SUCCESS) return st;
            ++m->adam_t;
            const float c1 = 1.0f / (1.0f - std::pow(b1, (float)m->adam_t));
            const float c2 = 1.0f / (1.0f - std::pow(b2, (float)m->adam_t));
            hipLaunchKernelGGL(hhg::adam_kernel, dim3((hhg::kMlpParams + 255) / 256), dim3(256), 0, m->stream,
                               m->d_params, m->d_grad, m->d_m, m->d_v, lr, b1, b2, eps, c1, c2);
            MLP_TRY(hipGetLastError(), "adam");
            hipLaunchKernelGGL(hhg::refresh_kernel, dim3((nr + 255) / 256), dim3(256), 0, m->stream, m->d_params,
                               m->d_params_bf, m->d_pack);
            MLP_TRY(hipGetLastError(), "refresh");
        }
        // one sync per epoch; it also ends the upload of the host order vector before its reuse
        MLP_TRY(hipMemcpyAsync(hist.data(), m->d_loss_hist, sizeof(float) * steps_per_epoch, hipMemcpyDeviceToHost,
                               m->stream), "loss history");
        MLP_TRY(hipStreamSynchronize(m->stream), "train sync");
        double acc = 0.0;
        for (int64_t q = 0; q < steps_per_epoch; ++q) acc += hist[q];
        if (epoch_loss) epoch_loss[e] = (float)(acc / (double)steps_per_epoch);
    }
    return refresh(m);
}

}  // extern "C"
