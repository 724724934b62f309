// knc_mlp.hip -- training (forward / backward / Adam) and inference kernels
// of the learned c2p pressure guess, and its part of the C ABI.
//
// Training is FP32 on the vector ALUs with FP32 master weights: one thread
// per sample runs the forward and backward pass, the workgroup reduces its
// samples' weight gradients through LDS into a per-workgroup partial, and
// k_mlp_adam sums the partials in a fixed order (deterministic) before the
// Adam update.  Inference (k_mlp_predict and the stage kernel's learned
// guess) runs the BF16 MFMA path of knc_mlp.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/knc_gpu.h"
#include "knc_mlp.hpp"
#include "knc_mlp_host.hpp"

namespace knc {

constexpr int TB = 128;   // samples per training workgroup

__global__ void __launch_bounds__(TB) k_mlp_grad(const float* P, const float* x, const float* t, int64_t n,
                                                 float inv_n, float* partial, float* lossp)
{
    __shared__ float sW1[MLP_H * MLP_IN], sb1[MLP_H], sW2[MLP_H * MLP_H], sb2[MLP_H], sW3[MLP_H];
    __shared__ float sA[TB * (MLP_H + 1)];   // h1, later d1
    __shared__ float sB[TB * (MLP_H + 1)];   // d2
    __shared__ float sC[TB * (MLP_H + 1)];   // h2
    __shared__ float sX[TB * MLP_IN];
    __shared__ float sDy[TB];
    __shared__ float sL[TB];
    const int tid = threadIdx.x;
    for (int p = tid; p < MLP_H * MLP_IN; p += TB) sW1[p] = P[MLP_OFF_W1 + p];
    for (int p = tid; p < MLP_H * MLP_H; p += TB) sW2[p] = P[MLP_OFF_W2 + p];
    for (int p = tid; p < MLP_H; p += TB) {
        sb1[p] = P[MLP_OFF_B1 + p];
        sb2[p] = P[MLP_OFF_B2 + p];
        sW3[p] = P[MLP_OFF_W3 + p];
    }
    const float b3 = P[MLP_OFF_B3];
    __syncthreads();

    const int64_t s = (int64_t)blockIdx.x * TB + tid;
    const bool valid = s < n;
    float xs[MLP_IN];
    for (int m = 0; m < MLP_IN; ++m) xs[m] = valid ? x[4 * s + m] : 0.0f;
    const float target = valid ? t[s] : 0.0f;

    float h1[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) {
        float z = sb1[i];
#pragma unroll
        for (int m = 0; m < MLP_IN; ++m) z += sW1[i * MLP_IN + m] * xs[m];
        h1[i] = fast_tanh(z);
        sA[tid * (MLP_H + 1) + i] = h1[i];
    }
    float y = b3;
    for (int k = 0; k < MLP_H; ++k) {
        float z = sb2[k];
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) z += sW2[k * MLP_H + i] * h1[i];
        const float h2 = fast_tanh(z);
        sC[tid * (MLP_H + 1) + k] = h2;
        y += sW3[k] * h2;
    }
    const float e = valid ? y - target : 0.0f;
    const float dy = 2.0f * e * inv_n;
    sL[tid] = e * e;
    sDy[tid] = dy;
    for (int m = 0; m < MLP_IN; ++m) sX[tid * MLP_IN + m] = xs[m];
    // d2[k] = dy W3[k] (1 - h2^2), d1[i] = (1 - h1^2) sum_k d2[k] W2[k][i]
    float d1[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) d1[i] = 0.0f;
    for (int k = 0; k < MLP_H; ++k) {
        const float h2 = sC[tid * (MLP_H + 1) + k];
        const float d2 = dy * sW3[k] * (1.0f - h2 * h2);
        sB[tid * (MLP_H + 1) + k] = d2;
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) d1[i] += d2 * sW2[k * MLP_H + i];
    }
    __syncthreads();

    float* out = partial + (int64_t)blockIdx.x * MLP_NPARAM;
    // dW2[k][i] = sum_s d2[s][k] h1[s][i]; db2; dW3[k] = sum_s dy[s] h2[s][k]
    for (int e2 = tid; e2 < MLP_H * MLP_H; e2 += TB) {
        const int k = e2 / MLP_H, i = e2 % MLP_H;
        float acc = 0.0f;
        for (int q = 0; q < TB; ++q) acc += sB[q * (MLP_H + 1) + k] * sA[q * (MLP_H + 1) + i];
        out[MLP_OFF_W2 + e2] = acc;
    }
    for (int k = tid; k < MLP_H; k += TB) {
        float a2 = 0.0f, a3 = 0.0f;
        for (int q = 0; q < TB; ++q) {
            a2 += sB[q * (MLP_H + 1) + k];
            a3 += sDy[q] * sC[q * (MLP_H + 1) + k];
        }
        out[MLP_OFF_B2 + k] = a2;
        out[MLP_OFF_W3 + k] = a3;
    }
    if (tid == 0) {
        float a = 0.0f, l = 0.0f;
        for (int q = 0; q < TB; ++q) {
            a += sDy[q];
            l += sL[q];
        }
        out[MLP_OFF_B3] = a;
        lossp[blockIdx.x] = l;
    }
    __syncthreads();
    // d1 replaces h1 in sA
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) {
        const float hv = sA[tid * (MLP_H + 1) + i];
        d1[i] *= 1.0f - hv * hv;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) sA[tid * (MLP_H + 1) + i] = d1[i];
    __syncthreads();
    for (int e1 = tid; e1 < MLP_H * MLP_IN; e1 += TB) {
        const int i = e1 / MLP_IN, m = e1 % MLP_IN;
        float acc = 0.0f;
        for (int q = 0; q < TB; ++q) acc += sA[q * (MLP_H + 1) + i] * sX[q * MLP_IN + m];
        out[MLP_OFF_W1 + e1] = acc;
    }
    for (int i = tid; i < MLP_H; i += TB) {
        float acc = 0.0f;
        for (int q = 0; q < TB; ++q) acc += sA[q * (MLP_H + 1) + i];
        out[MLP_OFF_B1 + i] = acc;
    }
}

// Sums the partial gradients in workgroup order; optional Adam update.
__global__ void k_mlp_adam(const float* partial, const float* lossp, int nparts, float* P, float* m, float* v,
                           float lr, float b1t, float b2t, int adam, float* grad_out, float* loss_out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss_out) {
        float l = 0.0f;
        for (int w = 0; w < nparts; ++w) l += lossp[w];
        *loss_out = l;
    }
    if (p >= MLP_NPARAM) return;
    float g = 0.0f;
    for (int w = 0; w < nparts; ++w) g += partial[(int64_t)w * MLP_NPARAM + p];
    if (grad_out) grad_out[p] = g;
    if (!adam) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float mm = beta1 * m[p] + (1.0f - beta1) * g;
    const float vv = beta2 * v[p] + (1.0f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - b1t);
    const float vh = vv / (1.0f - b2t);
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Raw features of conserved states (the device definition, for building
// training sets that match inference exactly).
__global__ void k_mlp_features(int64_t n, const double* cons, const double* gamma, float* f)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const double D = cons[4 * k], S2 = cons[4 * k + 1] * cons[4 * k + 1] + cons[4 * k + 2] * cons[4 * k + 2];
        float ff[4];
        mlp_raw_features(D, S2, cons[4 * k + 3], gamma[k], ff);
        for (int m = 0; m < 4; ++m) f[4 * k + m] = ff[m];
    }
}

// MFMA inference over an array: normalised network output y and the guess.
__global__ void __launch_bounds__(256) k_mlp_predict(MlpDev M, int64_t n, const double* cons, const double* gamma,
                                                     float* y, double* guess)
{
    __shared__ MlpShared sh;
    mlp_shared_load(&sh, M);
    const int64_t base = (int64_t)blockIdx.x * blockDim.x;
    for (int64_t b0 = base; b0 < n; b0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = b0 + threadIdx.x;
        const bool valid = k < n;
        const int64_t kk = valid ? k : n - 1;
        const double D = cons[4 * kk], S2 = cons[4 * kk + 1] * cons[4 * kk + 1] + cons[4 * kk + 2] * cons[4 * kk + 2];
        float f[4], x[4];
        mlp_raw_features(D, S2, cons[4 * kk + 3], gamma[kk], f);
        for (int m = 0; m < 4; ++m) x[m] = (f[m] - M.norm.shift[m]) * M.norm.scale[m];
        const float out = mlp_wave(M, &sh, x);
        if (valid) {
            if (y) y[k] = out;
            if (guess) guess[k] = D * (double)__builtin_amdgcn_exp2f(out / M.norm.tscale + M.norm.tshift);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
int MlpHost::upload_fragments()
{
    // The MlpShared layout: a1 | a2b | a2 | a3 (bf16x8 per lane) | tanh table.
    std::vector<__bf16> img(MLP_IMAGE_FRAGS * 8, (__bf16)0.0f);
    __bf16* a1 = img.data();
    __bf16* a2b = a1 + 2 * 64 * 8;
    __bf16* a2 = a2b + 2 * 64 * 8;
    __bf16* a3 = a2 + 2 * 4 * 64 * 8;
    const float* W1 = params.data() + MLP_OFF_W1;
    const float* B1 = params.data() + MLP_OFF_B1;
    const float* W2 = params.data() + MLP_OFF_W2;
    const float* B2 = params.data() + MLP_OFF_B2;
    const float* W3 = params.data() + MLP_OFF_W3;
    for (int mt = 0; mt < 2; ++mt)
        for (int lane = 0; lane < 64; ++lane) {
            const int g = lane / 32, row = 32 * mt + lane % 32;
            for (int i = 0; i < 8; ++i) {
                const int k = 8 * g + i;
                a1[(mt * 64 + lane) * 8 + i] =
                    (__bf16)(k < MLP_IN ? W1[row * MLP_IN + k] : (k == MLP_IN ? B1[row] : 0.0f));
                a2b[(mt * 64 + lane) * 8 + i] = (__bf16)(k == 0 ? B2[row] : 0.0f);
                for (int c = 0; c < 4; ++c)
                    a2[((mt * 4 + c) * 64 + lane) * 8 + i] = (__bf16)W2[row * MLP_H + mlp_perm(c, g, i)];
            }
        }
    for (int c = 0; c < 4; ++c)
        for (int lane = 0; lane < 64; ++lane)
            for (int i = 0; i < 8; ++i)
                a3[(c * 64 + lane) * 8 + i] = (__bf16)(lane % 32 == 0 ? W3[mlp_perm(c, lane / 32, i)] : 0.0f);
    std::vector<uint16_t> tab(TANH_N);
    for (int m = 0; m < TANH_N; ++m) {
        const __bf16 t = (__bf16)(float)std::tanh((m - TANH_MID) / (double)TANH_RH);
        std::memcpy(&tab[m], &t, 2);
    }
    if (!d_frag) {
        if (hipMalloc(&d_frag, frag_bytes()) != hipSuccess) return KNC_ERR_HIP;
    }
    if (hipMemcpy(d_frag, img.data(), img.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy((char*)d_frag + img.size() * 2, tab.data(), tab.size() * 2, hipMemcpyHostToDevice) != hipSuccess)
        return KNC_ERR_HIP;
    dev.image = (const uint4*)d_frag;
    dev.b3 = params[MLP_OFF_B3];
    std::memcpy(dev.norm.shift, norm.data(), 4 * sizeof(float));
    std::memcpy(dev.norm.scale, norm.data() + 4, 4 * sizeof(float));
    dev.norm.tshift = norm[8];
    dev.norm.tscale = norm[9];
    return KNC_OK;
}

size_t MlpHost::frag_bytes() { return MLP_IMAGE_BYTES; }

}  // namespace knc

using namespace knc;

namespace {
thread_local std::string g_mlp_err;
int mfail(int code, const char* msg)
{
    g_mlp_err = msg;
    return code;
}
struct Buf {
    void* p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n ? n : 8); }
};
}  // namespace

extern "C" {

const char* knc_mlp_last_error_message(void) { return g_mlp_err.c_str(); }
int knc_mlp_num_params(void) { return MLP_NPARAM; }

int knc_mlp_create(int32_t device, knc_mlp** out)
{
    if (!out) return mfail(KNC_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mfail(KNC_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return mfail(KNC_ERR_INVALID_ARGUMENT, "device out of range");
    if (hipSetDevice(device) != hipSuccess) return mfail(KNC_ERR_HIP, "hipSetDevice failed");
    MlpHost* m = new MlpHost();
    m->device = device;
    m->params.assign(MLP_NPARAM, 0.0f);
    m->norm.assign(MLP_NNORM, 0.0f);
    for (int k = 4; k < 8; ++k) m->norm[k] = 1.0f;
    m->norm[9] = 1.0f;
    if (m->upload_fragments() != KNC_OK) {
        delete m;
        return mfail(KNC_ERR_HIP, "fragment upload failed");
    }
    *out = reinterpret_cast<knc_mlp*>(m);
    return KNC_OK;
}

void knc_mlp_destroy(knc_mlp* mm)
{
    MlpHost* m = reinterpret_cast<MlpHost*>(mm);
    if (!m) return;
    if (m->d_frag) (void)hipFree(m->d_frag);
    delete m;
}

int knc_mlp_set_params(knc_mlp* mm, const float* params, const float* norm)
{
    MlpHost* m = reinterpret_cast<MlpHost*>(mm);
    if (!m || !params || !norm) return mfail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    m->params.assign(params, params + MLP_NPARAM);
    m->norm.assign(norm, norm + MLP_NNORM);
    (void)hipSetDevice(m->device);
    return m->upload_fragments();
}

int knc_mlp_get_params(knc_mlp* mm, float* params, float* norm)
{
    MlpHost* m = reinterpret_cast<MlpHost*>(mm);
    if (!m) return mfail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    if (params) std::memcpy(params, m->params.data(), MLP_NPARAM * sizeof(float));
    if (norm) std::memcpy(norm, m->norm.data(), MLP_NNORM * sizeof(float));
    return KNC_OK;
}

// Glorot-uniform weights, zero biases.
int knc_mlp_init(knc_mlp* mm, uint64_t seed)
{
    MlpHost* m = reinterpret_cast<MlpHost*>(mm);
    if (!m) return mfail(KNC_ERR_INVALID_ARGUMENT, "NULL argument");
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int fan_out, int count) {
        const float a = std::sqrt(6.0f / (float)(fan_in + fan_out));
        std::uniform_real_distribution<float> u(-a, a);
        for (int k = 0; k < count; ++k) m->params[off + k] = u(rng);
    };
    std::fill(m->params.begin(), m->params.end(), 0.0f);
    fill(MLP_OFF_W1, MLP_IN, MLP_H, MLP_H * MLP_IN);
    fill(MLP_OFF_W2, MLP_H, MLP_H, MLP_H * MLP_H);
    fill(MLP_OFF_W3, MLP_H, 1, MLP_H);
    (void)hipSetDevice(m->device);
    return m->upload_fragments();
}

// Loss (mean squared error) and its gradient over n normalised samples.
int knc_mlp_loss_grad(knc_mlp* mm, int64_t n, const float* x, const float* t, float* loss, float* grad)
{
    MlpHost* m = reinterpret_cast<MlpHost*>(mm);
    if (!m || n <= 0 || !x || !t) return mfail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    (void)hipSetDevice(m->device);
    const int nb = (int)((n + TB - 1) / TB);
    Buf dP, dx, dt, dpart, dloss, dg, dl;
    if (dP.alloc(MLP_NPARAM * 4) || dx.alloc(n * 16) || dt.alloc(n * 4) || dpart.alloc((size_t)nb * MLP_NPARAM * 4) ||
        dloss.alloc(nb * 4) || dg.alloc(MLP_NPARAM * 4) || dl.alloc(4))
        return mfail(KNC_ERR_HIP, "hipMalloc failed");
    if (hipMemcpy(dP.p, m->params.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice) ||
        hipMemcpy(dx.p, x, n * 16, hipMemcpyHostToDevice) || hipMemcpy(dt.p, t, n * 4, hipMemcpyHostToDevice))
        return mfail(KNC_ERR_HIP, "hipMemcpy failed");
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(TB), 0, nullptr, (const float*)dP.p, (const float*)dx.p,
                       (const float*)dt.p, n, 1.0f / (float)n, (float*)dpart.p, (float*)dloss.p);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, nullptr, (const float*)dpart.p,
                       (const float*)dloss.p, nb, (float*)dP.p, nullptr, nullptr, 0.0f, 0.0f, 0.0f, 0, (float*)dg.p,
                       (float*)dl.p);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return mfail(KNC_ERR_HIP, "kernel failed");
    float L = 0.0f;
    if (hipMemcpy(&L, dl.p, 4, hipMemcpyDeviceToHost) ||
        (grad && hipMemcpy(grad, dg.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost)))
        return mfail(KNC_ERR_HIP, "hipMemcpy failed");
    if (loss) *loss = L / (float)n;
    return KNC_OK;
}

// Adam over `epochs` passes of the n samples in shuffled mini-batches; the
// learning rate halves after every epoch.  epoch_loss (may be NULL)
// receives the mean training loss of each epoch.
int knc_mlp_train(knc_mlp* mm, int64_t n, const float* x, const float* t, int32_t epochs, int32_t batch, float lr,
                  uint64_t seed, float* epoch_loss)
{
    MlpHost* m = reinterpret_cast<MlpHost*>(mm);
    if (!m || n <= 0 || !x || !t || epochs < 1 || batch < TB) return mfail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    (void)hipSetDevice(m->device);
    batch = (batch / TB) * TB;
    const int nb = batch / TB;
    Buf dP, dm, dv, dx, dt, dpart, dloss, dl;
    if (dP.alloc(MLP_NPARAM * 4) || dm.alloc(MLP_NPARAM * 4) || dv.alloc(MLP_NPARAM * 4) || dx.alloc(n * 16) ||
        dt.alloc(n * 4) || dpart.alloc((size_t)nb * MLP_NPARAM * 4) || dloss.alloc(nb * 4) || dl.alloc(4))
        return mfail(KNC_ERR_HIP, "hipMalloc failed");
    if (hipMemcpy(dP.p, m->params.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice) ||
        hipMemset(dm.p, 0, MLP_NPARAM * 4) || hipMemset(dv.p, 0, MLP_NPARAM * 4))
        return mfail(KNC_ERR_HIP, "init failed");
    std::mt19937_64 rng(seed);
    std::vector<int64_t> order((size_t)n);
    std::vector<float> xs((size_t)n * 4), ts((size_t)n);
    int64_t step = 0;
    float rate = lr;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int64_t k = 0; k < n; ++k) order[k] = k;
        std::shuffle(order.begin(), order.end(), rng);
        for (int64_t k = 0; k < n; ++k) {
            std::memcpy(&xs[4 * k], &x[4 * order[k]], 16);
            ts[k] = t[order[k]];
        }
        if (hipMemcpy(dx.p, xs.data(), n * 16, hipMemcpyHostToDevice) ||
            hipMemcpy(dt.p, ts.data(), n * 4, hipMemcpyHostToDevice))
            return mfail(KNC_ERR_HIP, "hipMemcpy failed");
        double sum = 0.0;
        int64_t cnt = 0;
        for (int64_t b0 = 0; b0 + batch <= n; b0 += batch) {
            ++step;
            const float b1t = std::pow(0.9f, (float)step), b2t = std::pow(0.999f, (float)step);
            hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(TB), 0, nullptr, (const float*)dP.p,
                               (const float*)dx.p + 4 * b0, (const float*)dt.p + b0, (int64_t)batch,
                               1.0f / (float)batch, (float*)dpart.p, (float*)dloss.p);
            hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, nullptr,
                               (const float*)dpart.p, (const float*)dloss.p, nb, (float*)dP.p, (float*)dm.p,
                               (float*)dv.p, rate, b1t, b2t, 1, nullptr, (float*)dl.p);
            if ((step & 63) == 0 || b0 + 2 * batch > n) {
                float L = 0.0f;
                if (hipMemcpy(&L, dl.p, 4, hipMemcpyDeviceToHost)) return mfail(KNC_ERR_HIP, "hipMemcpy failed");
                sum += L / (float)batch;
                ++cnt;
            }
        }
        if (hipGetLastError() != hipSuccess) return mfail(KNC_ERR_HIP, "training kernel failed");
        if (epoch_loss) epoch_loss[ep] = cnt ? (float)(sum / (double)cnt) : 0.0f;
        rate *= 0.5f;
    }
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(m->params.data(), dP.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost))
        return mfail(KNC_ERR_HIP, "readback failed");
    return m->upload_fragments();
}

int knc_mlp_features(int64_t n, const double* cons, const double* gamma, float* features)
{
    if (n <= 0 || !cons || !gamma || !features) return mfail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    Buf dc, dg, df;
    if (dc.alloc(n * 32) || dg.alloc(n * 8) || df.alloc(n * 16)) return mfail(KNC_ERR_HIP, "hipMalloc failed");
    if (hipMemcpy(dc.p, cons, n * 32, hipMemcpyHostToDevice) || hipMemcpy(dg.p, gamma, n * 8, hipMemcpyHostToDevice))
        return mfail(KNC_ERR_HIP, "hipMemcpy failed");
    int64_t g = std::min<int64_t>(65536, (n + 255) / 256);
    hipLaunchKernelGGL(k_mlp_features, dim3((unsigned)g), dim3(256), 0, nullptr, n, (const double*)dc.p,
                       (const double*)dg.p, (float*)df.p);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(features, df.p, n * 16, hipMemcpyDeviceToHost))
        return mfail(KNC_ERR_HIP, "kernel failed");
    return KNC_OK;
}

int knc_mlp_predict(knc_mlp* mm, int64_t n, const double* cons, const double* gamma, float* y, double* guess)
{
    MlpHost* m = reinterpret_cast<MlpHost*>(mm);
    if (!m || n <= 0 || !cons || !gamma) return mfail(KNC_ERR_INVALID_ARGUMENT, "bad arguments");
    (void)hipSetDevice(m->device);
    Buf dc, dg, dy, dq;
    if (dc.alloc(n * 32) || dg.alloc(n * 8) || dy.alloc(n * 4) || dq.alloc(n * 8))
        return mfail(KNC_ERR_HIP, "hipMalloc failed");
    if (hipMemcpy(dc.p, cons, n * 32, hipMemcpyHostToDevice) || hipMemcpy(dg.p, gamma, n * 8, hipMemcpyHostToDevice))
        return mfail(KNC_ERR_HIP, "hipMemcpy failed");
    int64_t g = std::min<int64_t>(65536, (n + 255) / 256);
    hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)g), dim3(256), 0, nullptr, m->dev, n, (const double*)dc.p,
                       (const double*)dg.p, (float*)dy.p, (double*)dq.p);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return mfail(KNC_ERR_HIP, "kernel failed");
    if ((y && hipMemcpy(y, dy.p, n * 4, hipMemcpyDeviceToHost)) ||
        (guess && hipMemcpy(guess, dq.p, n * 8, hipMemcpyDeviceToHost)))
        return mfail(KNC_ERR_HIP, "hipMemcpy failed");
    return KNC_OK;
}

}  // extern "C"
