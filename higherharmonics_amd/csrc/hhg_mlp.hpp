// hhg_mlp.hpp — the step-proposal network: layout shared by the training
// kernels (hhg_mlp.hip), the inference inside the solve kernel and the host.
//
// Network: 8 features -> 64 -> 64 -> 64 -> 2 outputs, ReLU hidden units.
// Parameters are FP32 master copies; every forward pass (training and
// inference) uses their BF16 roundings and BF16 activations between layers,
// with FP32 accumulation.  Inputs and outputs are normalised by fixed affine
// maps stored with the network; normalised inputs are clamped (kMlpInClamp).
//
// Features of one k-point at the start of a sampling interval t0:
//   0 kx, 1 ky, 2 |d(k - A(t0))|, 3 2|d(k)| (gap at k), 4 A(t0), 5 E(t0) = -dA/dt,
//   6 log(h_prev / Delta) (the controller's carried step), 7 log10 of the last error ratio clamped below at 5^-8
// Outputs: 0 log(h / Delta) of the first step to attempt (trained on the step the
// error model of the first accepted step allows), 1 log(attempts in the interval).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hhg {

constexpr int kMlpIn = 8;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 2;
// flat FP32 parameter vector: W1[64][8] b1 W2[64][64] b2 W3[64][64] b3 W4[2][64] b4
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHidden * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;
constexpr int kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpHidden * kMlpHidden;
constexpr int kOffW4 = kOffB3 + kMlpHidden;
constexpr int kOffB4 = kOffW4 + kMlpOut * kMlpHidden;
constexpr int kMlpParams = kOffB4 + kMlpOut;  // 9026
// Normalised inputs are clamped to +-kMlpInClamp standard deviations in every
// forward pass: a feature far outside the training distribution (a step size
// the logs never showed) cannot drive the output arbitrarily far.
constexpr float kMlpInClamp = 8.0f;

__host__ __device__ inline float mlp_norm_input(float x, float mean, float istd) {
    const float z = (x - mean) * istd;
    return fminf(fmaxf(z, -kMlpInClamp), kMlpInClamp);
}

// MFMA operand packing for v_mfma_f32_32x32x16_bf16 (A = weights, out x in;
// B = activations, in x samples; C = out x samples).  Operand layouts of the
// instruction: A lane l holds row l % 32, k = 8 (l / 32) + i in element i;
// B lane l holds column l % 32 with the same k; C lane l register j holds
// row (j / 4) * 8 + 4 (l / 32) + j % 4, column l % 32.
//
// Activations never leave registers.  A wave's 64 samples form two column
// tiles nt = 0, 1 (nt = 1: lanes 32..63).  The C registers of a hidden layer
// become the next layer's B operand unchanged, element i of k-step ks being
// c[8 (ks % 2) + i] of row tile ks / 2; the next layer's weight columns are
// stored in that order (hidden_unit).  Layer 1 folds its bias into the
// padding half of K (a constant-one input) and takes the features from the
// lanes of its own column tile; the output layer writes column tile nt's
// outputs to rows 4 nt, 4 nt + 1, so every lane finds its own sample's outputs
// in c[0], c[1] of tile l / 32.
//
// One A tile = 64 lanes x 8 bf16.  Tiles: layer 1 [nt][mt], layers 2 and 3
// [mt][ks], output layer [nt][ks].
constexpr int kPackTile = 64 * 8;
constexpr int kTileL1 = 0, kTileL2 = 4, kTileL3 = 12, kTileL4 = 20, kPackTiles = 28;
constexpr int kPackTotal = kPackTiles * kPackTile;  // bf16 elements

// input unit of element i of k-step ks in lane half h (hidden-layer B operand)
__host__ __device__ constexpr int hidden_unit(int ks, int h, int i) { return 16 * ks + 8 * (i / 4) + 4 * h + (i % 4); }

struct MlpDev {
    const __bf16* pack;  // kPackTotal bf16
    const float* params; // FP32 copy of the BF16-rounded parameters
    float in_mean[kMlpIn];
    float in_istd[kMlpIn];
    float out_mean[kMlpOut];
    float out_std[kMlpOut];
};

// Block-shared LDS of the inference: the packed weights, then the layer 2
// and 3 biases (FP32).
constexpr int kMlpLdsBytes = kPackTotal * 2 + 2 * kMlpHidden * 4;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Every thread of the block calls this, then the block synchronises.
__device__ __forceinline__ void mlp_stage_lds(const MlpDev* __restrict__ M, unsigned char* lds) {
    const uint4* src = reinterpret_cast<const uint4*>(M->pack);
    uint4* dst = reinterpret_cast<uint4*>(lds);
    for (int q = threadIdx.x; q < kPackTotal / 8; q += blockDim.x) dst[q] = src[q];
    float* bias = reinterpret_cast<float*>(lds + kPackTotal * 2);
    for (int q = threadIdx.x; q < 2 * kMlpHidden; q += blockDim.x)
        bias[q] = M->params[q < kMlpHidden ? kOffB2 + q : kOffB3 + q - kMlpHidden];
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// BF16 rounding (round to nearest even) of accumulator elements 8 G .. 8 G + 7,
// then ReLU as a signed 16-bit max of each packed pair (negative values, -0
// included, have a negative pattern, and rounding never changes a sign, so
// this equals rounding the ReLU): one conversion and one max per pair.
template <int G>
__device__ __forceinline__ bf16x8 relu_bf16(const floatx16& c) {
    union {
        bf16x2 h[4];
        unsigned u[4];
        bf16x8 v;
    } t;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        bf16x2 p;
        p[0] = (__bf16)c[8 * G + 2 * q];
        p[1] = (__bf16)c[8 * G + 2 * q + 1];
        t.h[q] = p;
        asm("v_pk_max_i16 %0, %1, 0" : "=v"(t.u[q]) : "v"(t.u[q]));
    }
    return t.v;
}

// Inference for the 64 lanes of a wave (batch = the wave's k-points).  Every
// lane of the wave must call it.  feat: this lane's raw features; out: this
// lane's de-normalised outputs.  lds: the block's staged weights.
__device__ __forceinline__ void mlp_infer_wave(const MlpDev* __restrict__ M, const unsigned char* lds,
                                               const float feat[kMlpIn], float out[kMlpOut]) {
    const int l = threadIdx.x & 63;
    const int h = l >> 5;
    const bf16x8* A = reinterpret_cast<const bf16x8*>(lds) + l;  // tile t: A[64 t]
    const float* bias = reinterpret_cast<const float*>(lds + kPackTotal * 2);
    bf16x8 x, one = {};
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) x[i] = (__bf16)mlp_norm_input(feat[i], M->in_mean[i], M->in_istd[i]);
    one[0] = (__bf16)1.0f;
    const bf16x8 b1[2] = {h == 0 ? x : one, h == 1 ? x : one};

    bf16x8 act[2][4];  // [nt][ks]
    const floatx16 zero = {};
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const floatx16 c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[64 * (kTileL1 + 2 * nt + mt)], b1[nt], zero,
                                                                       0, 0, 0);
            act[nt][2 * mt] = relu_bf16<0>(c);
            act[nt][2 * mt + 1] = relu_bf16<1>(c);
        }
    }
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const int t0 = layer == 0 ? kTileL2 : kTileL3;
        const float* bl = bias + layer * kMlpHidden;
        bf16x8 nxt[2][4];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            floatx16 c0, c1;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const floatx4 b = *reinterpret_cast<const floatx4*>(bl + 32 * mt + 8 * q + 4 * h);
#pragma unroll
                for (int r = 0; r < 4; ++r) c0[4 * q + r] = c1[4 * q + r] = b[r];
            }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const bf16x8 a = A[64 * (t0 + 4 * mt + ks)];
                c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, act[0][ks], c0, 0, 0, 0);
                c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, act[1][ks], c1, 0, 0, 0);
            }
            nxt[0][2 * mt] = relu_bf16<0>(c0);
            nxt[0][2 * mt + 1] = relu_bf16<1>(c0);
            nxt[1][2 * mt] = relu_bf16<0>(c1);
            nxt[1][2 * mt + 1] = relu_bf16<1>(c1);
        }
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) act[nt][ks] = nxt[nt][ks];
    }
    floatx16 o0 = {}, o1 = {};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        o0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[64 * (kTileL4 + ks)], act[0][ks], o0, 0, 0, 0);
        o1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[64 * (kTileL4 + 4 + ks)], act[1][ks], o1, 0, 0, 0);
    }
    const float* P = M->params;
    const float y0 = (h ? o1[0] : o0[0]) + P[kOffB4];
    const float y1 = (h ? o1[1] : o0[1]) + P[kOffB4 + 1];
    out[0] = y0 * M->out_std[0] + M->out_mean[0];
    out[1] = y1 * M->out_std[1] + M->out_mean[1];
}

}  // namespace hhg
