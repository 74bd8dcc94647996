// Per-column MLP of the radiative-heating surrogate: BF16 GEMMs on the gfx950
// matrix cores (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, GELU,
// mean-squared loss, backward pass and Adam, all as HIP kernels. Training is
// driven from the host one mini-batch at a time.
#include "../../include/thor_mlp.h"
#include "common.h"
#include "mlp_internal.h"
#include "mlp_kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace thor {
namespace mlp {

struct Net {
    int nin = 0, nh = 0, nout = 0;
    long np = 0;                       // parameter count
    float *w = nullptr;                // FP32 master parameters (flat layout of thor_mlp.h)
    float *gm = nullptr, *gv = nullptr;  // Adam moments
    float *grad = nullptr;             // gradients, same layout
    uint16_t *wb = nullptr;            // BF16 copies of W1, W2, W3 (biases stay FP32)
    // BF16 copies in MFMA operand order for the inference kernels (n_hidden =
    // 256, n_out <= 64, n_in <= 128 only): W1 [kpad][256], W2 [256][256],
    // W3 [256][64], zero-padded (k_frag_bf16).
    bool fused = false;
    int kpad = 0;
    int ncu = 0;  // compute units of the device: inference blocks
    uint16_t *wt1 = nullptr, *wt2 = nullptr, *wt3 = nullptr;
    long step = 0;
    // activations for up to cap rows
    long cap = 0;
    uint16_t *X0 = nullptr, *H1 = nullptr, *H2 = nullptr, *dYb = nullptr, *dZb = nullptr;
    float *Z1 = nullptr, *Z2 = nullptr, *Y = nullptr, *T = nullptr, *dY = nullptr, *dH = nullptr;
    double *loss_part = nullptr;
    int nloss_part = 0;
    // training set
    long ndata = 0;
    float *Xall = nullptr, *Yall = nullptr;
    int *idx = nullptr;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    // streams of dynamical cores that run this network's inference kernels
    // (thor_dyn_set_rt_surrogate): work that rewrites the weights or the
    // activation buffers first waits for what those streams have queued
    std::vector<hipStream_t> users;
    hipEvent_t ev_user = nullptr, ev_self = nullptr;

    long oW1() const { return 0; }
    long ob1() const { return (long)nin * nh; }
    long oW2() const { return ob1() + nh; }
    long ob2() const { return oW2() + (long)nh * nh; }
    long oW3() const { return ob2() + nh; }
    long ob3() const { return oW3() + (long)nh * nout; }
    // BF16 weight copies: W1, W2, W3 packed
    long bW1() const { return 0; }
    long bW2() const { return (long)nin * nh; }
    long bW3() const { return bW2() + (long)nh * nh; }
    long nwb() const { return bW3() + (long)nh * nout; }
};

template <class T> static hipError_t alloc(Net &n, T **p, size_t count) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) {
        n.allocs.push_back(*p);
        e = hipMemset(*p, 0, std::max<size_t>(count, 1) * sizeof(T));
        // the null-stream fill finishes before the network's own stream uses the buffer
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    }
    return e;
}

static hipError_t ensure_cap(Net &n, long rows) {
    if (rows <= n.cap) return hipSuccess;
    for (void *p : {(void *)n.X0, (void *)n.H1, (void *)n.H2, (void *)n.dYb, (void *)n.dZb, (void *)n.Z1,
                    (void *)n.Z2, (void *)n.Y, (void *)n.T, (void *)n.dY, (void *)n.dH, (void *)n.loss_part}) {
        if (!p) continue;
        auto it = std::find(n.allocs.begin(), n.allocs.end(), p);
        if (it != n.allocs.end()) n.allocs.erase(it);
        (void)hipFree(p);
    }
    const long r = (rows + 63) / 64 * 64;
    hipError_t e = hipSuccess;
    auto A = [&](auto **p, size_t c) {
        if (e == hipSuccess) e = alloc(n, p, c);
    };
    A(&n.X0, (size_t)r * n.nin);
    A(&n.H1, (size_t)r * n.nh);
    A(&n.H2, (size_t)r * n.nh);
    A(&n.dYb, (size_t)r * n.nout);
    A(&n.dZb, (size_t)r * n.nh);
    A(&n.Z1, (size_t)r * n.nh);
    A(&n.Z2, (size_t)r * n.nh);
    A(&n.Y, (size_t)r * n.nout);
    A(&n.T, (size_t)r * n.nout);
    A(&n.dY, (size_t)r * n.nout);
    A(&n.dH, (size_t)r * n.nh);
    n.nloss_part = (int)((r * n.nout + 255) / 256);
    A(&n.loss_part, (size_t)n.nloss_part);
    if (e == hipSuccess) n.cap = r;
    return e;
}

static void gemm(Net &n, int epi, const uint16_t *A, long as_i, long as_k, const uint16_t *B, long bs_k, long bs_n,
                 int M, int N, int K, float *C, long ldc, uint16_t *H, long ldh, const float *bias) {
    GemmArgs g{A, as_i, as_k, B, bs_k, bs_n, M, N, K, C, ldc, H, ldh, bias};
    const dim3 grid((N + 63) / 64, (M + 63) / 64);
    if (epi == EPI_F32) k_gemm<EPI_F32><<<grid, 256, 0, n.stream>>>(g);
    else if (epi == EPI_BIAS_GELU) k_gemm<EPI_BIAS_GELU><<<grid, 256, 0, n.stream>>>(g);
    else k_gemm<EPI_BIAS_F32><<<grid, 256, 0, n.stream>>>(g);
}

// Forward over rows of X0 (BF16, already in n.X0): Z1/H1, Z2/H2, Y.
static void forward(Net &n, long rows) {
    const int M = (int)rows;
    gemm(n, EPI_BIAS_GELU, n.X0, n.nin, 1, n.wb + n.bW1(), n.nh, 1, M, n.nh, n.nin, n.Z1, n.nh, n.H1, n.nh,
         n.w + n.ob1());
    gemm(n, EPI_BIAS_GELU, n.H1, n.nh, 1, n.wb + n.bW2(), n.nh, 1, M, n.nh, n.nh, n.Z2, n.nh, n.H2, n.nh,
         n.w + n.ob2());
    gemm(n, EPI_BIAS_F32, n.H2, n.nh, 1, n.wb + n.bW3(), n.nout, 1, M, n.nout, n.nh, n.Y, n.nout, nullptr, 0,
         n.w + n.ob3());
}

// Loss and gradients for rows of X0 with targets in n.T; writes n.grad.
static void backward(Net &n, long rows, int loss_slot) {
    const int M = (int)rows;
    const long nout = n.nout, nh = n.nh;
    const float scale = 2.0f / (float)(rows * nout);
    k_mse<<<(unsigned)((rows * nout + 255) / 256), 256, 0, n.stream>>>(n.Y, n.T, rows * nout, scale, n.dY, n.dYb,
                                                                        n.loss_part + loss_slot);
    // layer 3
    gemm(n, EPI_F32, n.H2, 1, nh, n.dYb, nout, 1, (int)nh, (int)nout, M, n.grad + n.oW3(), nout, nullptr, 0, nullptr);
    k_colsum<<<(unsigned)nout, 256, 0, n.stream>>>(n.dY, M, (int)nout, n.grad + n.ob3());
    gemm(n, EPI_F32, n.dYb, nout, 1, n.wb + n.bW3(), 1, nout, M, (int)nh, (int)nout, n.dH, nh, nullptr, 0, nullptr);
    k_gelu_back<<<(unsigned)((rows * nh + 255) / 256), 256, 0, n.stream>>>(n.dH, n.Z2, rows * nh, n.dZb);
    // layer 2
    gemm(n, EPI_F32, n.H1, 1, nh, n.dZb, nh, 1, (int)nh, (int)nh, M, n.grad + n.oW2(), nh, nullptr, 0, nullptr);
    k_colsum<<<(unsigned)nh, 256, 0, n.stream>>>(n.dH, M, (int)nh, n.grad + n.ob2());
    gemm(n, EPI_F32, n.dZb, nh, 1, n.wb + n.bW2(), 1, nh, M, (int)nh, (int)nh, n.dH, nh, nullptr, 0, nullptr);
    k_gelu_back<<<(unsigned)((rows * nh + 255) / 256), 256, 0, n.stream>>>(n.dH, n.Z1, rows * nh, n.dZb);
    // layer 1
    gemm(n, EPI_F32, n.X0, 1, n.nin, n.dZb, nh, 1, n.nin, (int)nh, M, n.grad + n.oW1(), nh, nullptr, 0, nullptr);
    k_colsum<<<(unsigned)nh, 256, 0, n.stream>>>(n.dH, M, (int)nh, n.grad + n.ob1());
}

static void refresh_bf16(Net &n) {
    const long nw[3] = {(long)n.nin * n.nh, (long)n.nh * n.nh, (long)n.nh * n.nout};
    const long src[3] = {n.oW1(), n.oW2(), n.oW3()}, dst[3] = {n.bW1(), n.bW2(), n.bW3()};
    for (int i = 0; i < 3; i++)
        k_to_bf16<<<(unsigned)((nw[i] + 255) / 256), 256, 0, n.stream>>>(n.w + src[i], n.wb + dst[i], nw[i]);
    if (n.fused) {
        const int K[3] = {n.nin, n.nh, n.nh}, N[3] = {n.nh, n.nh, n.nout};
        const int Kp[3] = {n.kpad, n.nh, n.nh}, Np[3] = {n.nh, n.nh, 64};
        uint16_t *wt[3] = {n.wt1, n.wt2, n.wt3};
        for (int i = 0; i < 3; i++) {
            const long cnt = (long)Np[i] * Kp[i];
            k_frag_bf16<<<(unsigned)((cnt + 255) / 256), 256, 0, n.stream>>>(n.w + src[i], K[i], N[i], wt[i], Kp[i],
                                                                             Np[i]);
        }
    }
}

static InferWeights infer_weights(const Net &n) {
    return InferWeights{n.wt1, n.wt2, n.wt3, n.w + n.ob1(), n.w + n.ob2(), n.w + n.ob3(), n.nin, n.nout, n.kpad};
}

// Inference over rows of X0 into Y: the fused kernel when the shape allows it,
// else the three GEMMs of the training forward pass (same results).
static void infer(Net &n, long rows) {
    if (!n.fused) {
        forward(n, rows);
        return;
    }
    InferArgs a{n.X0, rows, infer_weights(n), n.Y};
    const unsigned grid = (unsigned)std::max(1L, std::min((rows + 31) / 32, (long)n.ncu));
    switch (n.kpad / 32) {
    case 1: k_mlp_infer<1><<<grid, kInferThreads, 0, n.stream>>>(a); break;
    case 2: k_mlp_infer<2><<<grid, kInferThreads, 0, n.stream>>>(a); break;
    case 3: k_mlp_infer<3><<<grid, kInferThreads, 0, n.stream>>>(a); break;
    default: k_mlp_infer<4><<<grid, kInferThreads, 0, n.stream>>>(a); break;
    }
}

static double sum_loss(Net &n, int slots) {
    std::vector<double> part((size_t)slots);
    (void)hipMemcpyAsync(part.data(), n.loss_part, slots * sizeof(double), hipMemcpyDeviceToHost, n.stream);
    (void)hipStreamSynchronize(n.stream);
    double s = 0.0;
    for (double v : part) s += v;
    return s;
}

} // namespace mlp
} // namespace thor

struct thor_mlp : thor::mlp::Net {};

using namespace thor;
using namespace thor::mlp;

#define MLP_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess) return fail(THOR_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

namespace thor {
namespace mlp {
// Orders the network's own stream after everything its users have queued.
static hipError_t wait_users(Net &n) {
    for (hipStream_t s : n.users) {
        hipError_t e = hipEventRecord(n.ev_user, s);
        if (e == hipSuccess) e = hipStreamWaitEvent(n.stream, n.ev_user, 0);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
} // namespace mlp
} // namespace thor

extern "C" thor_status thor_mlp_create(int n_in, int n_hidden, int n_out, unsigned long long seed, thor_mlp **out) {
    if (!out || n_in < 1 || n_hidden < 1 || n_out < 1 || n_in > 4096 || n_hidden > 4096 || n_out > 4096)
        return fail(THOR_ERR_BAD_ARG, "thor_mlp_create: bad sizes");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(THOR_ERR_NO_DEVICE, "thor_mlp_create: no HIP device");
    thor_mlp *m = new thor_mlp();
    Net &n = *m;
    n.nin = n_in;
    n.nh = n_hidden;
    n.nout = n_out;
    n.np = n.ob3() + n_out;
    hipError_t e = hipStreamCreateWithFlags(&n.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&n.ev_user, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&n.ev_self, hipEventDisableTiming);
    if (e == hipSuccess) e = alloc(n, &n.w, n.np);
    if (e == hipSuccess) e = alloc(n, &n.gm, n.np);
    if (e == hipSuccess) e = alloc(n, &n.gv, n.np);
    if (e == hipSuccess) e = alloc(n, &n.grad, n.np);
    if (e == hipSuccess) e = alloc(n, &n.wb, n.nwb());
    n.fused = n_hidden == 256 && n_out <= 64 && n_in <= 128;
    if (e == hipSuccess) {
        int dev = 0;
        e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&n.ncu, hipDeviceAttributeMultiprocessorCount, dev);
    }
    if (n.fused) {
        n.kpad = (n_in + 31) / 32 * 32;
        if (e == hipSuccess) e = alloc(n, &n.wt1, (size_t)256 * n.kpad);
        if (e == hipSuccess) e = alloc(n, &n.wt2, (size_t)256 * 256);
        if (e == hipSuccess) e = alloc(n, &n.wt3, (size_t)64 * 256);
    }
    if (e != hipSuccess) {
        thor_mlp_destroy(m);
        return fail(THOR_ERR_HIP, std::string("thor_mlp_create: ") + hipGetErrorString(e));
    }
    // Kaiming-uniform weights (bound sqrt(6 / fan_in)), zero biases
    std::vector<float> w(n.np, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](long off, int fan_in, long count) {
        const double b = std::sqrt(6.0 / fan_in);
        for (long i = 0; i < count; i++) w[off + i] = (float)(b * (2.0 * ((rng() >> 11) * 0x1.0p-53) - 1.0));
    };
    fill(n.oW1(), n_in, (long)n_in * n_hidden);
    fill(n.oW2(), n_hidden, (long)n_hidden * n_hidden);
    fill(n.oW3(), n_hidden, (long)n_hidden * n_out);
    *out = m;
    return thor_mlp_set_params(m, w.data(), n.np);
}

extern "C" void thor_mlp_destroy(thor_mlp *m) {
    if (!m) return;
    for (hipStream_t s : m->users) (void)hipStreamSynchronize(s);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *p : m->allocs) (void)hipFree(p);
    if (m->ev_user) (void)hipEventDestroy(m->ev_user);
    if (m->ev_self) (void)hipEventDestroy(m->ev_self);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

extern "C" long thor_mlp_param_count(const thor_mlp *m) { return m ? m->np : -1; }

extern "C" thor_status thor_mlp_get_params(thor_mlp *m, float *dst, long n) {
    if (!m || !dst || n != m->np) return fail(THOR_ERR_BAD_ARG, "thor_mlp_get_params: bad argument");
    MLP_TRY(hipMemcpyAsync(dst, m->w, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    return THOR_OK;
}

extern "C" thor_status thor_mlp_set_params(thor_mlp *m, const float *src, long n) {
    if (!m || !src || n != m->np) return fail(THOR_ERR_BAD_ARG, "thor_mlp_set_params: bad argument");
    MLP_TRY(wait_users(*m));
    MLP_TRY(hipMemcpyAsync(m->w, src, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemsetAsync(m->gm, 0, n * sizeof(float), m->stream));
    MLP_TRY(hipMemsetAsync(m->gv, 0, n * sizeof(float), m->stream));
    m->step = 0;
    refresh_bf16(*m);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipStreamSynchronize(m->stream));
    return THOR_OK;
}

extern "C" thor_status thor_mlp_set_data(thor_mlp *m, const float *X, const float *Y, long n) {
    if (!m || !X || !Y || n < 1) return fail(THOR_ERR_BAD_ARG, "thor_mlp_set_data: bad argument");
    MLP_TRY(wait_users(*m));
    MLP_TRY(hipStreamSynchronize(m->stream));  // the old training set may still be read
    for (void *p : {(void *)m->Xall, (void *)m->Yall, (void *)m->idx}) {
        if (!p) continue;
        auto it = std::find(m->allocs.begin(), m->allocs.end(), p);
        if (it != m->allocs.end()) m->allocs.erase(it);
        (void)hipFree(p);
    }
    m->Xall = m->Yall = nullptr;
    m->idx = nullptr;
    MLP_TRY(alloc(*m, &m->Xall, (size_t)n * m->nin));
    MLP_TRY(alloc(*m, &m->Yall, (size_t)n * m->nout));
    MLP_TRY(alloc(*m, &m->idx, (size_t)n));
    MLP_TRY(hipMemcpyAsync(m->Xall, X, (size_t)n * m->nin * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemcpyAsync(m->Yall, Y, (size_t)n * m->nout * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    m->ndata = n;
    return THOR_OK;
}

extern "C" thor_status thor_mlp_train_epoch(thor_mlp *m, int batch, float lr, unsigned long long seed, double *loss) {
    if (!m || batch < 1 || !(lr > 0)) return fail(THOR_ERR_BAD_ARG, "thor_mlp_train_epoch: bad argument");
    if (m->ndata < 1) return fail(THOR_ERR_BAD_ARG, "thor_mlp_train_epoch: no training data");
    Net &n = *m;
    const long nb = (n.ndata + batch - 1) / batch;
    MLP_TRY(wait_users(n));
    MLP_TRY(ensure_cap(n, batch));
    // seeded shuffle on the host
    std::vector<int> perm((size_t)n.ndata);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(seed);
    for (long i = n.ndata - 1; i > 0; i--) std::swap(perm[i], perm[rng() % (unsigned long long)(i + 1)]);
    MLP_TRY(hipMemcpyAsync(n.idx, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, n.stream));
    double total = 0.0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (long b = 0; b < nb; b++) {
        const long r0 = b * batch, rows = std::min<long>(batch, n.ndata - r0);
        k_gather<<<(unsigned)((rows * std::max(n.nin, n.nout) + 255) / 256), 256, 0, n.stream>>>(
            n.Xall, n.Yall, n.idx + r0, rows, n.nin, n.nout, n.X0, n.T);
        forward(n, rows);
        MLP_TRY(hipMemsetAsync(n.loss_part, 0, n.nloss_part * sizeof(double), n.stream));
        backward(n, rows, 0);
        n.step++;
        const float bc1 = 1.0f - std::pow(b1, (float)n.step), bc2 = 1.0f - std::pow(b2, (float)n.step);
        k_adam<<<(unsigned)((n.np + 255) / 256), 256, 0, n.stream>>>(n.w, n.gm, n.gv, n.grad, n.np, lr, b1, b2, eps,
                                                                     bc1, bc2);
        refresh_bf16(n);
        MLP_TRY(hipGetLastError());
        total += sum_loss(n, (int)((rows * n.nout + 255) / 256));  // sum of squared errors of the batch
    }
    if (loss) *loss = total / (double)n.ndata / n.nout;
    return THOR_OK;
}

extern "C" thor_status thor_mlp_forward(thor_mlp *m, const float *X, float *Y, long n) {
    if (!m || !X || !Y || n < 1) return fail(THOR_ERR_BAD_ARG, "thor_mlp_forward: bad argument");
    Net &net = *m;
    MLP_TRY(wait_users(net));
    MLP_TRY(ensure_cap(net, n));
    float *dX = nullptr;
    MLP_TRY(hipMalloc((void **)&dX, (size_t)n * net.nin * sizeof(float)));
    hipError_t e = hipMemcpyAsync(dX, X, (size_t)n * net.nin * sizeof(float), hipMemcpyHostToDevice, net.stream);
    if (e == hipSuccess) {
        k_to_bf16<<<(unsigned)((n * net.nin + 255) / 256), 256, 0, net.stream>>>(dX, net.X0, n * net.nin);
        infer(net, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(Y, net.Y, (size_t)n * net.nout * sizeof(float), hipMemcpyDeviceToHost, net.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(net.stream);
    (void)hipFree(dX);
    if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("thor_mlp_forward: ") + hipGetErrorString(e));
    return THOR_OK;
}

extern "C" thor_status thor_mlp_gradients(thor_mlp *m, const float *X, const float *Y, long n, float *grad,
                                          double *loss) {
    if (!m || !X || !Y || !grad || n < 1 || n > 65536) return fail(THOR_ERR_BAD_ARG, "thor_mlp_gradients: bad argument");
    Net &net = *m;
    MLP_TRY(wait_users(net));
    MLP_TRY(ensure_cap(net, n));
    float *dX = nullptr;
    MLP_TRY(hipMalloc((void **)&dX, (size_t)n * net.nin * sizeof(float)));
    hipError_t e = hipMemcpyAsync(dX, X, (size_t)n * net.nin * sizeof(float), hipMemcpyHostToDevice, net.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(net.T, Y, (size_t)n * net.nout * sizeof(float), hipMemcpyHostToDevice, net.stream);
    if (e == hipSuccess) e = hipMemsetAsync(net.loss_part, 0, net.nloss_part * sizeof(double), net.stream);
    if (e == hipSuccess) {
        k_to_bf16<<<(unsigned)((n * net.nin + 255) / 256), 256, 0, net.stream>>>(dX, net.X0, n * net.nin);
        forward(net, n);
        backward(net, n, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(grad, net.grad, net.np * sizeof(float), hipMemcpyDeviceToHost, net.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(net.stream);
    (void)hipFree(dX);
    if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("thor_mlp_gradients: ") + hipGetErrorString(e));
    if (loss) *loss = sum_loss(net, (int)((n * net.nout + 255) / 256)) / (double)(n * net.nout);
    return THOR_OK;
}

extern "C" thor_status thor_mlp_time_forward(thor_mlp *m, long n, int repeats, double *ms) {
    if (!m || n < 1 || repeats < 1 || !ms) return fail(THOR_ERR_BAD_ARG, "thor_mlp_time_forward: bad argument");
    Net &net = *m;
    MLP_TRY(wait_users(net));
    MLP_TRY(ensure_cap(net, n));
    hipEvent_t a, b;
    MLP_TRY(hipEventCreate(&a));
    MLP_TRY(hipEventCreate(&b));
    infer(net, n);  // warm-up
    (void)hipEventRecord(a, net.stream);
    for (int i = 0; i < repeats; i++) infer(net, n);
    (void)hipEventRecord(b, net.stream);
    hipError_t e = hipEventSynchronize(b);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return fail(THOR_ERR_HIP, std::string("thor_mlp_time_forward: ") + hipGetErrorString(e));
    *ms = t / repeats;
    return THOR_OK;
}

// ------------------------------------------------------ surrogate helpers
namespace thor {
namespace mlp {
hipError_t prepare(thor_mlp *m, long rows) {
    // growing the buffers frees the old ones: nothing queued may still use them
    if (rows > m->cap) {
        for (hipStream_t s : m->users) {
            hipError_t e = hipStreamSynchronize(s);
            if (e != hipSuccess) return e;
        }
        hipError_t e = hipStreamSynchronize(m->stream);
        if (e != hipSuccess) return e;
    }
    return ensure_cap(*m, rows);
}
void attach(thor_mlp *m, hipStream_t s) {
    if (std::find(m->users.begin(), m->users.end(), s) == m->users.end()) m->users.push_back(s);
}
void detach(thor_mlp *m, hipStream_t s) {
    auto it = std::find(m->users.begin(), m->users.end(), s);
    if (it != m->users.end()) m->users.erase(it);
}
hipError_t order_on(thor_mlp *m, hipStream_t s) {
    hipError_t e = hipEventRecord(m->ev_self, m->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, m->ev_self, 0);
    return e;
}
uint16_t *input(thor_mlp *m) { return m->X0; }
const float *output(thor_mlp *m) { return m->Y; }
void forward_on(thor_mlp *m, long rows, hipStream_t s) {
    Net &n = *m;
    const hipStream_t keep = n.stream;
    n.stream = s;
    infer(n, rows);
    n.stream = keep;
}
bool fused_weights(const thor_mlp *m, InferWeights *w) {
    if (!m->fused) return false;
    *w = infer_weights(*m);
    return true;
}
void dims(const thor_mlp *m, int *nin, int *nout) {
    *nin = m->nin;
    *nout = m->nout;
}
} // namespace mlp
} // namespace thor
