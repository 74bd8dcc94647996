// MFMA building blocks of the per-column MLP shared by the training kernels
// (mlp.hip) and the fused surrogate kernel of the dynamical core.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bf16.h"

namespace thor {
namespace mlp {

// K steps of an inference layer unrolled together (their operand loads are
// issued ahead of the MFMAs)
#ifndef THOR_MLP_KUNROLL
#define THOR_MLP_KUNROLL 8
#endif

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

// acc[ti][tj] += A[16 ti.., k] * W[k, n0 + 16 tj..] over k in [0, K), A in LDS
// (row stride lda), W in global memory in MFMA operand order (k_frag_bf16,
// ldw = the padded K of that copy).
template <int NT, int RT = 2>
__device__ __forceinline__ void infer_layer(const uint16_t *sA, int lda, int K, const uint16_t *Wt, int ldw, int n0,
                                            v4f acc[RT][NT]) {
    const int KT = ldw / 16;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = (lane >> 4) * 4;
#pragma unroll
    for (int ti = 0; ti < RT; ti++)
#pragma unroll
        for (int tj = 0; tj < NT; tj++) acc[ti][tj] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll THOR_MLP_KUNROLL
    for (int k0 = 0; k0 < K; k0 += 16) {
        v4s a[RT], b[NT];
#pragma unroll
        for (int ti = 0; ti < RT; ti++) a[ti] = *(const v4s *)&sA[(16 * ti + li) * lda + k0 + lk];
#pragma unroll
        for (int tj = 0; tj < NT; tj++)
            b[tj] = *(const v4s *)&Wt[(((size_t)(n0 / 16 + tj) * KT + k0 / 16) * 64 + lane) * 4];
#pragma unroll
        for (int ti = 0; ti < RT; ti++)
#pragma unroll
            for (int tj = 0; tj < NT; tj++)
                acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
    }
}

// Weights of the fused inference (NH = 256 hidden units, n_out <= 64,
// n_in <= 256): BF16 copies in MFMA operand order (k_frag_bf16) over
// [NH][kpad], [NH][NH], [64][NH] (n by k), zero-padded; FP32 biases.
struct InferWeights {
    const uint16_t *W1t, *W2t, *W3t;  // [NH][kpad], [NH][NH], [64][NH]
    const float *b1, *b2, *b3;
    int nin, nout, kpad;  // kpad = n_in rounded up to 16
};

// The three layers for 16 RT rows whose BF16 inputs are in bufA (row stride
// kpad + 8, zero beyond n_in): H1 = GELU(X W1 + b1) into bufB, H2 =
// GELU(H1 W2 + b2) into bufA (row stride NH + 8), and this lane's outputs of
// layer 3, y[ti][e] = (H2 W3 + b3)[16 ti + 4 (lane / 16) + e][16 wave +
// lane % 16] (valid where that column is < n_out). The K tiling and the
// epilogues are those of the three-GEMM forward pass, so the values are the
// same. Both buffers hold 16 RT x (NH + 8) BF16.
template <int NH, int RT = 2>
__device__ __forceinline__ void mlp3_rows(uint16_t *bufA, uint16_t *bufB, const InferWeights &w, float y[RT][4]) {
    constexpr int LDH = NH + 8;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = (lane >> 4) * 4;
    v4f acc[RT][4];
    infer_layer<4, RT>(bufA, w.kpad + 8, w.kpad, w.W1t, w.kpad, 64 * wave, acc);
#pragma unroll
    for (int ti = 0; ti < RT; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int row = 16 * ti + lk + e, col = 64 * wave + 16 * tj + li;
                bufB[row * LDH + col] = f2bf(gelu(acc[ti][tj][e] + w.b1[col]));
            }
    __syncthreads();
    infer_layer<4, RT>(bufB, LDH, NH, w.W2t, NH, 64 * wave, acc);
#pragma unroll
    for (int ti = 0; ti < RT; ti++)
#pragma unroll
        for (int tj = 0; tj < 4; tj++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int row = 16 * ti + lk + e, col = 64 * wave + 16 * tj + li;
                bufA[row * LDH + col] = f2bf(gelu(acc[ti][tj][e] + w.b2[col]));
            }
    __syncthreads();
    v4f acc3[RT][1];
    infer_layer<1, RT>(bufA, LDH, NH, w.W3t, NH, 16 * wave, acc3);
    const int col = 16 * wave + li;
    const float bias = col < w.nout ? w.b3[col] : 0.f;
#pragma unroll
    for (int ti = 0; ti < RT; ti++)
#pragma unroll
        for (int e = 0; e < 4; e++) y[ti][e] = acc3[ti][0][e] + bias;
}

} // namespace mlp
} // namespace thor
