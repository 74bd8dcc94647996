// MFMA building blocks of the per-column MLP shared by the training kernels
// (mlp.hip) and the surrogate kernel of the dynamical core.
//
// Matrix instruction: gfx950's v_mfma_f32_16x16x32_bf16 (BF16 inputs, FP32
// accumulation). Operand layout: lane l holds A[l % 16][8 (l / 16) + e] and
// B[8 (l / 16) + e][l % 16], e = 0..7; result element e of lane l is
// D[4 (l / 16) + e][l % 16], e = 0..3.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bf16.h"

namespace thor {
namespace mlp {

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma32(v8s a, v8s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0,
                                                   0, 0);
}

// GELU in its tanh form, 0.5 z (1 + tanh(c (z + 0.044715 z^3))) with
// c = sqrt(2 / pi), written as z sigma(2 c (z + 0.044715 z^3)) with the
// hardware exponential and reciprocal (about 10 instructions, against about
// 36 for the erf form; it is the activation of the whole network, training
// included). gelu_u is its argument u = 2 c z (1 + 0.044715 z^2).
__device__ __forceinline__ float gelu_u(float z) { return 1.5957691216057308f * z * (1.0f + 0.044715f * z * z); }
__device__ __forceinline__ float gelu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-gelu_u(z))); }

// Weights of the inference kernels (NH = 256 hidden units, n_out <= 64,
// n_in <= 128): BF16 copies in MFMA B-operand order (k_frag_bf16) of W1
// [kpad x 256], W2 [256 x 256], W3 [256 x 64], zero-padded; FP32 biases.
struct InferWeights {
    const uint16_t *W1t, *W2t, *W3t;
    const float *b1, *b2, *b3;
    int nin, nout, kpad;  // kpad = n_in rounded up to 32
};

// The inference kernels run 8 waves per block (two per SIMD, so one wave's
// GELU epilogue overlaps the other's MFMAs). A wave's weights are held in
// registers for the whole kernel: wave w computes the hidden units 32 w ..
// 32 w + 31 of layers 1 and 2 (two 16-unit tiles) and, for w < 4, the
// outputs 16 w .. 16 w + 15 of layer 3. KT1 = kpad / 32. The layers are
// formed transposed, H^T = W^T X^T: the weights are the MFMA's A operand (a
// B-operand fragment of W is the A-operand fragment of W^T) and the
// activations, stored [sample][unit] in LDS, its B operand; a lane's four
// results are four consecutive units of one sample, one 8-byte LDS store.
// Biases: this lane's units 16 tile + 4 (lane / 16) + e.
constexpr int kInferWaves = 8, kInferThreads = 64 * kInferWaves;
template <int KT1> struct WaveWeights {
    v8s w1[2][KT1], w2[2][8], w3[1][8];
    float b1[2][4], b2[2][4], b3[4];
};

// Fragment (n-tile nt, k-tile kt) of an operand stored by k_frag_bf16.
__device__ __forceinline__ v8s frag(const uint16_t *Wt, int KT, int nt, int kt, int lane) {
    return *(const v8s *)&Wt[(((size_t)nt * KT + kt) * 64 + lane) * 8];
}

template <int KT1> __device__ __forceinline__ void load_wave_weights(const InferWeights &w, WaveWeights<KT1> &r) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lq = (lane >> 4) * 4;
#pragma unroll
    for (int tj = 0; tj < 2; tj++) {
#pragma unroll
        for (int kt = 0; kt < KT1; kt++) r.w1[tj][kt] = frag(w.W1t, KT1, 2 * wave + tj, kt, lane);
#pragma unroll
        for (int kt = 0; kt < 8; kt++) r.w2[tj][kt] = frag(w.W2t, 8, 2 * wave + tj, kt, lane);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            r.b1[tj][e] = w.b1[32 * wave + 16 * tj + lq + e];
            r.b2[tj][e] = w.b2[32 * wave + 16 * tj + lq + e];
        }
    }
    const int w3t = wave < 4 ? wave : 0;  // waves 4..7 have no layer-3 tile
#pragma unroll
    for (int kt = 0; kt < 8; kt++) r.w3[0][kt] = frag(w.W3t, 8, w3t, kt, lane);
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int col = 16 * w3t + lq + e;
        r.b3[e] = col < w.nout ? w.b3[col] : 0.f;
    }
}

// acc[ti][tj] = sum over k-tiles of W^T[tile tj, 32 kt..] X^T[32 kt.., samples
// 16 ti..]: element e of lane l is unit 16 tj + 4 (l / 16) + e of sample
// 16 ti + l % 16. X in LDS as [sample][k] (row stride ldx, 16-byte aligned
// rows), W in registers.
template <int RT, int NT, int KT>
__device__ __forceinline__ void layer_regs(const uint16_t *sX, int ldx, const v8s (&w)[NT][KT], v4f acc[RT][NT]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = (lane >> 4) * 8;
#pragma unroll
    for (int ti = 0; ti < RT; ti++)
#pragma unroll
        for (int tj = 0; tj < NT; tj++) acc[ti][tj] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < KT; kt++) {
        v8s x[RT];
#pragma unroll
        for (int ti = 0; ti < RT; ti++) x[ti] = *(const v8s *)&sX[(16 * ti + li) * ldx + 32 * kt + lk];
#pragma unroll
        for (int ti = 0; ti < RT; ti++)
#pragma unroll
            for (int tj = 0; tj < NT; tj++) acc[ti][tj] = mfma32(w[tj][kt], x[ti], acc[ti][tj]);
    }
}

// GELU(acc + b) of a lane's four consecutive units, rounded to BF16, stored
// with one 8-byte write at sH[sample][unit0].
__device__ __forceinline__ void store_gelu4(uint16_t *sH, int ldh, int sample, int unit0, const v4f &a,
                                            const float (&b)[4]) {
    const uint32_t lo = f2bf2(gelu(a[0] + b[0]), gelu(a[1] + b[1]));
    const uint32_t hi = f2bf2(gelu(a[2] + b[2]), gelu(a[3] + b[3]));
    *(uint2 *)&sH[sample * ldh + unit0] = make_uint2(lo, hi);
}

// The three layers for 16 RT samples whose BF16 inputs are in bufA (row
// stride kpad + 8, zero beyond n_in): H1 = GELU(X W1 + b1) into bufB, H2 =
// GELU(H1 W2 + b2) into bufA (row stride NH + 8), and, for waves 0..3, this
// lane's outputs of layer 3, y[ti][e] = (H2 W3 + b3)[sample 16 ti +
// lane % 16][output 16 wave + 4 (lane / 16) + e] (valid where that output is
// < n_out). The products and epilogues are k_gemm's (training forward pass).
// Both buffers hold 16 RT x (NH + 8) BF16; bufB is free on return, and the
// caller synchronises before bufA is written again.
template <int KT1, int RT = 2>
__device__ __forceinline__ void mlp3_rows(uint16_t *bufA, uint16_t *bufB, int kpad, const WaveWeights<KT1> &r,
                                          float y[RT][4]) {
    constexpr int NH = 256, LDH = NH + 8;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lq = (lane >> 4) * 4;
    v4f acc[RT][2];
    layer_regs<RT, 2, KT1>(bufA, kpad + 8, r.w1, acc);
#pragma unroll
    for (int ti = 0; ti < RT; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++) store_gelu4(bufB, LDH, 16 * ti + li, 32 * wave + 16 * tj + lq, acc[ti][tj], r.b1[tj]);
    __syncthreads();
    // bufA's inputs were read by every wave before the barrier above
    layer_regs<RT, 2, 8>(bufB, LDH, r.w2, acc);
#pragma unroll
    for (int ti = 0; ti < RT; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++) store_gelu4(bufA, LDH, 16 * ti + li, 32 * wave + 16 * tj + lq, acc[ti][tj], r.b2[tj]);
    __syncthreads();
    if (wave < 4) {
        v4f acc3[RT][1];
        layer_regs<RT, 1, 8>(bufA, LDH, r.w3, acc3);
#pragma unroll
        for (int ti = 0; ti < RT; ti++)
#pragma unroll
            for (int e = 0; e < 4; e++) y[ti][e] = acc3[ti][0][e] + r.b3[e];
    }
}

} // namespace mlp
} // namespace thor
