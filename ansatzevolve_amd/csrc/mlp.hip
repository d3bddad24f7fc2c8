// Learned warm start: a small fully connected network (FP32) that maps
// integral features to initial TUPS angles, with hand-written HIP kernels for
// inference and training (forward, backward, Adam).  No reference
// counterpart; it sits beside the statevector engine.
//
// Network: x -> tanh(x W0 + b0) -> tanh(. W1 + b1) -> . W2 + b2 (linear
// output).  Weights are row-major [in][out].  Loss: masked mean squared
// error  L = sum_{b,j} m_bj (y_bj - t_bj)^2 / sum m.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/ansatzsynth.h"
#include "common.hpp"

namespace ae {
namespace {

constexpr int TB = 16;  // tile edge of the small GEMM kernels

// C[M][N] = A[M][K] * B[K][N] (+ bias[N]) with optional tanh; A, B row-major.
// TA: use A^T (A stored [K][M]); TB_: use B^T (B stored [N][K]).
template <bool TA, bool TBT>
__global__ void __launch_bounds__(TB *TB)
    k_gemm(const float *__restrict__ A, const float *__restrict__ B, const float *__restrict__ bias,
           float *__restrict__ C, float *__restrict__ Cpre, int M, int N, int K, int act, float beta)
{
    __shared__ float sa[TB][TB + 1];
    __shared__ float sb[TB][TB + 1];
    const int tx = threadIdx.x % TB, ty = threadIdx.x / TB;
    const int row = blockIdx.y * TB + ty, col = blockIdx.x * TB + tx;
    float acc = 0.f;
    for (int k0 = 0; k0 < K; k0 += TB) {
        const int ka = k0 + tx, kb = k0 + ty;
        float av = 0.f, bv = 0.f;
        if (row < M && ka < K) av = TA ? A[(size_t)ka * M + row] : A[(size_t)row * K + ka];
        if (kb < K && col < N) bv = TBT ? B[(size_t)col * K + kb] : B[(size_t)kb * N + col];
        sa[ty][tx] = av;
        sb[ty][tx] = bv;
        __syncthreads();
#pragma unroll
        for (int t = 0; t < TB; t++) acc += sa[ty][t] * sb[t][tx];
        __syncthreads();
    }
    if (row < M && col < N) {
        float v = acc + (bias ? bias[col] : 0.f);
        if (Cpre) Cpre[(size_t)row * N + col] = v;
        if (act) v = tanhf(v);
        const size_t o = (size_t)row * N + col;
        C[o] = beta != 0.f ? beta * C[o] + v : v;
    }
}

// dY = 2 m (y - t) / msum; loss partial per block.
__global__ void __launch_bounds__(256)
    k_loss_grad(const float *__restrict__ y, const float *__restrict__ t, const float *__restrict__ m, int n,
                float inv_msum, float *__restrict__ dy, float *__restrict__ part)
{
    __shared__ float red[256];
    float acc = 0.f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float d = (y[i] - t[i]) * m[i];
        dy[i] = 2.f * d * inv_msum;
        acc += d * (y[i] - t[i]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv_msum;
}

// dZ = dA * (1 - tanh(z)^2) with a = tanh(z) stored
__global__ void __launch_bounds__(256) k_tanh_back(const float *__restrict__ a, float *__restrict__ d, int n)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) d[i] *= 1.f - a[i] * a[i];
}

// db[j] = sum_b dZ[b][j]
__global__ void __launch_bounds__(256) k_colsum(const float *__restrict__ d, int B, int N, float *__restrict__ out)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int b = 0; b < B; b++) s += d[(size_t)b * N + j];
    out[j] = s;
}

__global__ void __launch_bounds__(256)
    k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1, float *__restrict__ m2, int n,
           float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    }
}

struct Mlp {
    int in = 0, hid = 0, out = 0;
    size_t nparam = 0;
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
    long long step = 0;
    hipStream_t stream = nullptr;
    // offsets of W0 b0 W1 b1 W2 b2 in params
    size_t oW0, ob0, oW1, ob1, oW2, ob2;
    // activations for a batch
    int cap = 0;
    float *X = nullptr, *A1 = nullptr, *A2 = nullptr, *Y = nullptr, *T = nullptr, *Mk = nullptr, *D2 = nullptr,
          *D1 = nullptr, *DY = nullptr, *part = nullptr;

    ~Mlp()
    {
        for (float *p : {params, grads, m1, m2, X, A1, A2, Y, T, Mk, D2, D1, DY, part})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    bool ensure_batch(int B)
    {
        if (B <= cap) return true;
        for (float **p : {&X, &A1, &A2, &Y, &T, &Mk, &D2, &D1, &DY})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
        auto al = [](float **p, size_t n) { return hipMalloc(p, n * sizeof(float)) == hipSuccess; };
        if (!al(&X, (size_t)B * in) || !al(&A1, (size_t)B * hid) || !al(&A2, (size_t)B * hid) ||
            !al(&Y, (size_t)B * out) || !al(&T, (size_t)B * out) || !al(&Mk, (size_t)B * out) ||
            !al(&D2, (size_t)B * hid) || !al(&D1, (size_t)B * hid) || !al(&DY, (size_t)B * out))
            return false;
        cap = B;
        return true;
    }
    dim3 grid(int M, int N) const { return dim3((N + TB - 1) / TB, (M + TB - 1) / TB); }

    void forward(int B)
    {
        hipLaunchKernelGGL((k_gemm<false, false>), grid(B, hid), dim3(TB * TB), 0, stream, X, params + oW0,
                           params + ob0, A1, nullptr, B, hid, in, 1, 0.f);
        hipLaunchKernelGGL((k_gemm<false, false>), grid(B, hid), dim3(TB * TB), 0, stream, A1, params + oW1,
                           params + ob1, A2, nullptr, B, hid, hid, 1, 0.f);
        hipLaunchKernelGGL((k_gemm<false, false>), grid(B, out), dim3(TB * TB), 0, stream, A2, params + oW2,
                           params + ob2, Y, nullptr, B, out, hid, 0, 0.f);
    }

    // gradients of the masked MSE into `grads`; returns via part the loss
    void backward(int B, float inv_msum)
    {
        const int ny = B * out;
        hipLaunchKernelGGL(k_loss_grad, dim3(64), dim3(256), 0, stream, Y, T, Mk, ny, inv_msum, DY, part);
        // layer 2: dW2 = A2^T DY, db2 = colsum DY, D2 = DY W2^T
        hipLaunchKernelGGL((k_gemm<true, false>), grid(hid, out), dim3(TB * TB), 0, stream, A2, DY, nullptr,
                           grads + oW2, nullptr, hid, out, B, 0, 0.f);
        hipLaunchKernelGGL(k_colsum, dim3((out + 255) / 256), dim3(256), 0, stream, DY, B, out, grads + ob2);
        hipLaunchKernelGGL((k_gemm<false, true>), grid(B, hid), dim3(TB * TB), 0, stream, DY, params + oW2, nullptr,
                           D2, nullptr, B, hid, out, 0, 0.f);
        hipLaunchKernelGGL(k_tanh_back, dim3(64), dim3(256), 0, stream, A2, D2, B * hid);
        // layer 1
        hipLaunchKernelGGL((k_gemm<true, false>), grid(hid, hid), dim3(TB * TB), 0, stream, A1, D2, nullptr,
                           grads + oW1, nullptr, hid, hid, B, 0, 0.f);
        hipLaunchKernelGGL(k_colsum, dim3((hid + 255) / 256), dim3(256), 0, stream, D2, B, hid, grads + ob1);
        hipLaunchKernelGGL((k_gemm<false, true>), grid(B, hid), dim3(TB * TB), 0, stream, D2, params + oW1, nullptr,
                           D1, nullptr, B, hid, hid, 0, 0.f);
        hipLaunchKernelGGL(k_tanh_back, dim3(64), dim3(256), 0, stream, A1, D1, B * hid);
        // layer 0
        hipLaunchKernelGGL((k_gemm<true, false>), grid(in, hid), dim3(TB * TB), 0, stream, X, D1, nullptr,
                           grads + oW0, nullptr, in, hid, B, 0, 0.f);
        hipLaunchKernelGGL(k_colsum, dim3((hid + 255) / 256), dim3(256), 0, stream, D1, B, hid, grads + ob0);
    }
};

} // namespace
} // namespace ae

using ae::Mlp;

extern "C" {

void *mlpCreate(int nin, int nhidden, int nout, unsigned long long seed)
{
    if (nin < 1 || nhidden < 1 || nout < 1) return nullptr;
    auto *m = new (std::nothrow) Mlp;
    if (!m) return nullptr;
    m->in = nin;
    m->hid = nhidden;
    m->out = nout;
    m->oW0 = 0;
    m->ob0 = m->oW0 + (size_t)nin * nhidden;
    m->oW1 = m->ob0 + nhidden;
    m->ob1 = m->oW1 + (size_t)nhidden * nhidden;
    m->oW2 = m->ob1 + nhidden;
    m->ob2 = m->oW2 + (size_t)nhidden * nout;
    m->nparam = m->ob2 + nout;
    std::vector<float> p(m->nparam, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, int fan_in, size_t cnt) {
        std::normal_distribution<float> nd(0.f, 1.f / std::sqrt((float)fan_in));
        for (size_t i = 0; i < cnt; i++) p[off + i] = nd(rng);
    };
    fill(m->oW0, nin, (size_t)nin * nhidden);
    fill(m->oW1, nhidden, (size_t)nhidden * nhidden);
    fill(m->oW2, nhidden, (size_t)nhidden * nout);
    for (size_t i = 0; i < (size_t)nhidden * nout; i++) p[m->oW2 + i] *= 0.1f;  // start near zero output
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&m->params, m->nparam * 4) != hipSuccess || hipMalloc(&m->grads, m->nparam * 4) != hipSuccess ||
        hipMalloc(&m->m1, m->nparam * 4) != hipSuccess || hipMalloc(&m->m2, m->nparam * 4) != hipSuccess ||
        hipMalloc(&m->part, 64 * 4) != hipSuccess ||
        hipMemcpy(m->params, p.data(), m->nparam * 4, hipMemcpyHostToDevice) != hipSuccess) {
        delete m;
        return nullptr;
    }
    std::vector<float> z(m->nparam, 0.f);
    (void)hipMemcpy(m->m1, z.data(), m->nparam * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(m->m2, z.data(), m->nparam * 4, hipMemcpyHostToDevice);
    return m;
}

int mlpDestroy(void *h)
{
    if (!h) return 1;
    delete static_cast<Mlp *>(h);
    return 0;
}

long long mlpParameterCount(void *h) { return h ? (long long)static_cast<Mlp *>(h)->nparam : -1; }

int mlpGetParameters(void *h, float *out)
{
    auto *m = static_cast<Mlp *>(h);
    if (!m || !out) return 1;
    return hipMemcpy(out, m->params, m->nparam * 4, hipMemcpyDeviceToHost) == hipSuccess ? 0 : 5;
}

int mlpSetParameters(void *h, const float *in)
{
    auto *m = static_cast<Mlp *>(h);
    if (!m || !in) return 1;
    return hipMemcpy(m->params, in, m->nparam * 4, hipMemcpyHostToDevice) == hipSuccess ? 0 : 5;
}

int mlpForward(void *h, int batch, const float *x, float *y)
{
    auto *m = static_cast<Mlp *>(h);
    if (!m || !x || !y) return 1;
    if (batch < 1) return 3;
    if (!m->ensure_batch(batch)) return 5;
    if (hipMemcpyAsync(m->X, x, (size_t)batch * m->in * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess) return 5;
    m->forward(batch);
    if (hipMemcpyAsync(y, m->Y, (size_t)batch * m->out * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess) return 5;
    return hipStreamSynchronize(m->stream) == hipSuccess ? 0 : 5;
}

// Loss and parameter gradients (same layout as the parameters) without an update.
int mlpGradient(void *h, int batch, const float *x, const float *target, const float *mask, float *loss,
                float *grads)
{
    auto *m = static_cast<Mlp *>(h);
    if (!m || !x || !target || !mask || !loss) return 1;
    if (batch < 1) return 3;
    if (!m->ensure_batch(batch)) return 5;
    const size_t ny = (size_t)batch * m->out;
    double msum = 0;
    for (size_t i = 0; i < ny; i++) msum += mask[i];
    if (msum <= 0) return 3;
    if (hipMemcpyAsync(m->X, x, (size_t)batch * m->in * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->T, target, ny * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess ||
        hipMemcpyAsync(m->Mk, mask, ny * 4, hipMemcpyHostToDevice, m->stream) != hipSuccess)
        return 5;
    m->forward(batch);
    m->backward(batch, (float)(1.0 / msum));
    float part[64];
    if (hipMemcpyAsync(part, m->part, sizeof(part), hipMemcpyDeviceToHost, m->stream) != hipSuccess) return 5;
    if (grads && hipMemcpyAsync(grads, m->grads, m->nparam * 4, hipMemcpyDeviceToHost, m->stream) != hipSuccess)
        return 5;
    if (hipStreamSynchronize(m->stream) != hipSuccess) return 5;
    double s = 0;
    for (float v : part) s += v;
    *loss = (float)s;
    return 0;
}

// One Adam step on a mini-batch; *loss receives the pre-step loss.
int mlpTrainStep(void *h, int batch, const float *x, const float *target, const float *mask, float lr,
                 float *loss)
{
    auto *m = static_cast<Mlp *>(h);
    if (!m) return 1;
    const int rc = mlpGradient(h, batch, x, target, mask, loss, nullptr);
    if (rc) return rc;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
    hipLaunchKernelGGL(ae::k_adam, dim3(256), dim3(256), 0, m->stream, m->params, m->grads, m->m1, m->m2,
                       (int)m->nparam, lr, b1, b2, eps, c1, c2);
    return hipStreamSynchronize(m->stream) == hipSuccess ? 0 : 5;
}

} // extern "C"
