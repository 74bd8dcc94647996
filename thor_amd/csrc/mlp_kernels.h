// Device code of the per-column MLP (see mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bf16.h"
#include "mlp_layer.h"

namespace thor {
namespace mlp {


enum { EPI_F32 = 0, EPI_BIAS_GELU = 1, EPI_BIAS_F32 = 2 };

// C[M x N] = A[M x K] * B[K x N] with element (i, k) of A at A[i*as_i + k*as_k]
// and (k, n) of B at B[k*bs_k + n*bs_n], so transposed operands need no copy.
// EPI_F32: C = product; EPI_BIAS_F32: C = product + bias; EPI_BIAS_GELU:
// C = product + bias (pre-activation, FP32) and H = bf16(GELU(C)).
struct GemmArgs {
    const uint16_t *A;
    long as_i, as_k;
    const uint16_t *B;
    long bs_k, bs_n;
    int M, N, K;
    float *C;
    long ldc;
    uint16_t *H;
    long ldh;
    const float *bias;
};

// d/dz of gelu: s + z s (1 - s) u'(z), s = sigma(u), u' = 2 c (1 + 3 0.044715 z^2).
__device__ inline float gelu_grad(float z) {
    const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-gelu_u(z)));
    return s + z * s * (1.0f - s) * 1.5957691216057308f * (1.0f + 3.0f * 0.044715f * z * z);
}

// 64 x 64 output tile per 256-thread block; each of the 4 waves owns a 32 x 32
// quarter made of 2 x 2 MFMA 16x16x32 tiles. K advances 32 at a time through
// LDS tiles of A (64 x 32) and B (32 x 64); out-of-range elements load as 0.
// MFMA 16x16x32 bf16 operand layout (gfx950): lane l holds A[l % 16][8 (l / 16)
// + e] and B[8 (l / 16) + e][l % 16], e = 0..7; the result element e of lane l
// is D[4 (l / 16) + e][l % 16] (e = 0..3).
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ __align__(16) uint16_t sA[64][32 + 8];
    __shared__ __align__(16) uint16_t sBt[64][32 + 8];  // B transposed: [n][k]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
    const int ar = tid >> 2, ak = (tid & 3) * 8;  // A staging: row, first k (8 per thread)
    const int bn = tid >> 2, bk = (tid & 3) * 8;  // B staging: column, first k
    const int li = lane & 15, lk = (lane >> 4) * 8;
    for (int k0 = 0; k0 < g.K; k0 += 32) {
        {
            const long i = m0 + ar;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const long k = k0 + ak + e;
                sA[ar][ak + e] = (i < g.M && k < g.K) ? g.A[i * g.as_i + k * g.as_k] : (uint16_t)0;
            }
        }
        {
            const long n = n0 + bn;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const long k = k0 + bk + e;
                sBt[bn][bk + e] = (k < g.K && n < g.N) ? g.B[k * g.bs_k + n * g.bs_n] : (uint16_t)0;
            }
        }
        __syncthreads();
        v8s a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            a[t] = *(const v8s *)&sA[wm + t * 16 + li][lk];
            b[t] = *(const v8s *)&sBt[wn + t * 16 + li][lk];
        }
#pragma unroll
        for (int ti = 0; ti < 2; ti++)
#pragma unroll
            for (int tj = 0; tj < 2; tj++) acc[ti][tj] = mfma32(a[ti], b[tj], acc[ti][tj]);
        __syncthreads();
    }
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const long i = m0 + wm + ti * 16 + (lane >> 4) * 4 + e;
                const long n = n0 + wn + tj * 16 + (lane & 15);
                if (i >= g.M || n >= g.N) continue;
                float v = acc[ti][tj][e];
                if (EPI != EPI_F32) v += g.bias[n];
                g.C[i * g.ldc + n] = v;
                if (EPI == EPI_BIAS_GELU) g.H[i * g.ldh + n] = f2bf_dev(gelu(v));
            }
}

// Inference over row groups of 32 (mlp3_rows), one 256-thread block per CU
// looping over the groups: each wave loads its weight slice into registers
// once, so the layers' B operands come from registers and only the
// activations pass through LDS. The group's input rows are one contiguous run
// of X.
struct InferArgs {
    const uint16_t *X;  // [rows][nin]
    long rows;
    InferWeights w;
    float *Y;  // [rows][nout]
};

template <int KT1> __global__ __launch_bounds__(kInferThreads, 1) void k_mlp_infer(InferArgs g) {
    constexpr int R = 32, LDH = 256 + 8;
    __shared__ __align__(16) uint16_t bufA[R * LDH], bufB[R * LDH];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lq = (lane >> 4) * 4;
    const int nin = g.w.nin, kpad = g.w.kpad, ldx = kpad + 8;
    WaveWeights<KT1> wr;
    load_wave_weights<KT1>(g.w, wr);
    const long ngroups = (g.rows + R - 1) / R;
    for (long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const long r0 = grp * R;
        {
            const int r = tid >> 4, part = tid & 15, kper = kpad / 16;  // 32 rows x 16 parts, kper <= 8
            const bool rv = r0 + r < g.rows;
            uint16_t v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = part * kper + j;
                v[j] = (j < kper && rv && k < nin) ? g.X[(r0 + r) * nin + k] : (uint16_t)0;
            }
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (j < kper) bufA[r * ldx + part * kper + j] = v[j];
        }
        __syncthreads();
        float y[2][4];
        mlp3_rows<KT1>(bufA, bufB, kpad, wr, y);
        if (wave < 4) {
#pragma unroll
            for (int ti = 0; ti < 2; ti++) {
                const long row = r0 + 16 * ti + li;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int col = 16 * wave + lq + e;
                    if (row < g.rows && col < g.w.nout) g.Y[row * g.w.nout + col] = y[ti][e];
                }
            }
        }
        __syncthreads();  // layer 3 has read bufA before the next group's inputs go there
    }
}

// BF16 copy of W [K][N] (FP32, row-major) in the order the inference kernels
// read their MFMA B operands: for each 32 x 16 tile (kt, nt) of W, lane l's
// eight values W[32 kt + 8 (l / 16) + e][16 nt + l % 16], e = 0..7, are at
// ((nt * Kpad / 32 + kt) * 64 + l) * 8 + e, so a wave's operand load is 1024
// contiguous bytes. Zero beyond K and N.
__global__ void k_frag_bf16(const float *src, int K, int N, uint16_t *dst, int Kpad, int Npad) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Npad * Kpad) return;
    const int KT = Kpad / 32;
    const int e = (int)(i & 7), l = (int)((i >> 3) & 63);
    const long tile = i >> 9;
    const int kt = (int)(tile % KT), nt = (int)(tile / KT);
    const int n = 16 * nt + (l & 15), k = 32 * kt + 8 * (l >> 4) + e;
    dst[i] = (n < N && k < K) ? f2bf(src[(size_t)k * N + n]) : (uint16_t)0;
}

__global__ void k_to_bf16(const float *src, uint16_t *dst, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = f2bf(src[i]);
}

// Mean-squared-error gradient dY = scale (Y - T) (FP32 and BF16) and the
// block's sum of squared differences in part[blockIdx.x].
__global__ __launch_bounds__(256) void k_mse(const float *Y, const float *T, long n, float scale, float *dY,
                                             uint16_t *dYb, double *part) {
    __shared__ double red[256];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double sq = 0.0;
    if (i < n) {
        const float d = Y[i] - T[i];
        dY[i] = scale * d;
        dYb[i] = f2bf(scale * d);
        sq = (double)d * d;
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dZ = dH * GELU'(Z), kept in FP32 in place of dH and in BF16 in dZb.
__global__ void k_gelu_back(float *dH, const float *Z, long n, uint16_t *dZb) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = dH[i] * gelu_grad(Z[i]);
    dH[i] = d;
    dZb[i] = f2bf(d);
}

// Column sums of a row-major M x N FP32 matrix: one block per column.
__global__ __launch_bounds__(256) void k_colsum(const float *X, int M, int N, float *out) {
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int r = threadIdx.x; r < M; r += 256) s += X[(long)r * N + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

__global__ void k_adam(float *w, float *m, float *v, const float *g, long n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// Mini-batch rows idx[0..rows) of the training set into X0 (BF16) and T.
__global__ void k_gather(const float *Xall, const float *Yall, const int *idx, long rows, int nin, int nout,
                         uint16_t *X0, float *T) {
    const int W = nin > nout ? nin : nout;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * W) return;
    const long r = t / W;
    const int j = (int)(t % W);
    const long src = idx[r];
    if (j < nin) X0[r * nin + j] = f2bf(Xall[src * nin + j]);
    if (j < nout) T[r * nout + j] = Yall[src * nout + j];
}

} // namespace mlp
} // namespace thor
