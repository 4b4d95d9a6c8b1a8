// hhg_mlp.hpp — the step-proposal network: layout shared by the training
// kernels (hhg_mlp.hip), the inference inside the solve kernel and the host.
//
// Network: 8 features -> 64 -> 64 -> 64 -> 2 outputs, ReLU hidden units.
// Parameters are FP32 master copies; every forward pass (training and
// inference) uses their BF16 roundings and BF16 activations between layers,
// with FP32 accumulation.  Inputs and outputs are normalised by fixed affine
// maps stored with the network.
//
// Features of one k-point at the start of a sampling interval t0:
//   0 kx, 1 ky, 2 |d(k - A(t0))|, 3 2|d(k)| (gap at k), 4 A(t0), 5 E(t0) = -dA/dt,
//   6 log(h_prev / Delta) (the controller's carried step), 7 log10 of the last error ratio
// Outputs: 0 log(h / Delta) of the first step to attempt, 1 log(attempts in the interval).
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

// MFMA operand packing of the hidden layers for v_mfma_f32_32x32x16_bf16:
// W (out x in) is the A operand; lane l of k-step ks in row tile mt holds
// W[32 mt + l % 32][16 ks + 8 (l / 32) + i], i = 0..7.
constexpr int kPackTile = 64 * 8;                 // bf16 per (mt, ks)
constexpr int kPackL1 = 2 * 1 * kPackTile;        // layer 1: K = 8 padded to one k-step of 16
constexpr int kPackL23 = 2 * 4 * kPackTile;       // layers 2, 3: K = 64
constexpr int kPackTotal = kPackL1 + 2 * kPackL23;

struct MlpDev {
    const __bf16* pack;  // kPackTotal bf16: layer 1, layer 2, layer 3
    const float* params; // FP32 copy of the BF16-rounded parameters (biases and W4 read here)
    float in_mean[kMlpIn];
    float in_istd[kMlpIn];
    float out_mean[kMlpOut];
    float out_std[kMlpOut];
};

// LDS per wave for inference: two activation buffers of 64 samples x 64 units.
constexpr int kMlpLdsPerWave = 2 * 64 * 64;  // bf16 elements

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Inference for the 64 lanes of a wave (batch = the wave's k-points).  Every
// lane of the wave must call it.  feat: this lane's raw features; out: this
// lane's de-normalised outputs.  lds: this wave's kMlpLdsPerWave bf16.
__device__ __forceinline__ void mlp_infer_wave(const MlpDev* __restrict__ M, const float feat[kMlpIn],
                                               float out[kMlpOut], __bf16* lds) {
    const int l = threadIdx.x & 63;
    const int half = l >> 5, col = l & 31;
    __bf16* buf0 = lds;
    __bf16* buf1 = lds + 64 * 64;
    // normalised features -> buf0[sample l][0..7], zeros in [8..15]
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) buf0[l * 64 + i] = (__bf16)((feat[i] - M->in_mean[i]) * M->in_istd[i]);
#pragma unroll
    for (int i = kMlpIn; i < 16; ++i) buf0[l * 64 + i] = (__bf16)0.0f;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();

    const bf16x8* pk = reinterpret_cast<const bf16x8*>(M->pack);
    const float* P = M->params;
    // one hidden layer: in -> out buffer, K = 16 * nks
    auto layer = [&](const bf16x8* wp, int nks, const float* bias, const __bf16* in, __bf16* outb) {
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                floatx16 c = {};
                for (int ks = 0; ks < nks; ++ks) {
                    const bf16x8 a = wp[(mt * nks + ks) * 64 + l];
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(in + (32 * nt + col) * 64 + 16 * ks + 8 * half);
                    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
                }
                const int sample = 32 * nt + col;
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int unit = 32 * mt + (j >> 2) * 8 + half * 4 + (j & 3);
                    const float v = c[j] + bias[unit];
                    outb[sample * 64 + unit] = (__bf16)(v > 0.0f ? v : 0.0f);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    };
    layer(pk, 1, P + kOffB1, buf0, buf1);
    layer(pk + kPackL1 / 8, 4, P + kOffB2, buf1, buf0);
    layer(pk + (kPackL1 + kPackL23) / 8, 4, P + kOffB3, buf0, buf1);
    // output layer on the VALU: this lane's own sample row
    float acc0 = P[kOffB4], acc1 = P[kOffB4 + 1];
    const bf16x8* row = reinterpret_cast<const bf16x8*>(buf1 + l * 64);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const bf16x8 v = row[q];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float x = (float)v[i];
            acc0 = fmaf(P[kOffW4 + 8 * q + i], x, acc0);
            acc1 = fmaf(P[kOffW4 + kMlpHidden + 8 * q + i], x, acc1);
        }
    }
    out[0] = acc0 * M->out_std[0] + M->out_mean[0];
    out[1] = acc1 * M->out_std[1] + M->out_mean[1];
    __builtin_amdgcn_wave_barrier();
}

}  // namespace hhg
