// knc_mlp.hpp -- the learned initial pressure guess of the SRHD c2p.
//
// Network: 4 -> 64 -> 64 -> 1, tanh hidden layers, linear output, mapping
// log2-scaled features of a conserved state to log2(p/D).  Features (before
// the affine normalisation stored with the weights):
//   f0 = log2 D
//   f1 = log2(|S|/D)
//   f2 = log2((tau + D - sqrt(D^2 + S^2))/D)   thermal part of tau/D: the
//        rest-mass + kinetic part sqrt(D^2+S^2) is removed in FP64 first,
//        since for cold gas at Gamma ~ 100 it hides p in the last digits
//   f3 = gamma
// Inference runs per wavefront on BF16 MFMA (v_mfma_f32_32x32x16_bf16) with
// FP32 accumulation: 64 zones (one per lane) are the N dimension, the hidden
// units the M dimension.  Parameter vector (4545 FP32): W1[64][4] | b1[64] |
// W2[64][64] | b2[64] | W3[64] | b3.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace knc {

constexpr int MLP_IN = 4;
constexpr int MLP_H = 64;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + 1;   // 4545
constexpr int MLP_NNORM = 10;                  // shift[4], scale[4], tshift, tscale

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

struct MlpNorm {
    float shift[4];
    float scale[4];
    float tshift;
    float tscale;
};

// Weights laid out as MFMA operand fragments (built on the host):
//   a1[mt][lane]      layer-1 A operand, rows = hidden 32mt + lane%32,
//                     k = 8(lane/32) + i (features 0..3, zero padded)
//   a2[mt][c][lane]   layer-2 A operand, rows = output 32mt + lane%32,
//                     k-chunk c with the hidden units permuted so that the
//                     layer-1 accumulator registers are the B operand as is
//   c1/c2/w3[mt][g]   per-lane-group fragments (16 floats) of b1, b2, W3 in
//                     the accumulator row order 8(v/4) + 4g + v%4
struct MlpDev {
    const bf16x8* a1;
    const bf16x8* a2;
    const floatx16* c1;
    const floatx16* c2;
    const floatx16* w3;
    float b3;
    MlpNorm norm;
};

// tanh(x) = 1 - 2/(2^(2x/ln 2) + 1), from the hardware exp2 and reciprocal.
__device__ inline float fast_tanh(float x)
{
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

// Inference tanh from a table in LDS: TANH_N intervals over [-8, 8], each
// entry (tanh(x_i), tanh(x_{i+1}) - tanh(x_i)), linear interpolation.  The
// interpolation error is at most h^2/8 max|tanh''| = 2.4e-5 (h = 1/64), far
// below the BF16 resolution of the next layer's operands, and it costs no
// transcendental instruction (exp2 and rcp issue at a quarter of the rate).
constexpr int TANH_N = 1024;
constexpr float TANH_X0 = -8.0f;
constexpr float TANH_RH = TANH_N / 16.0f;   // intervals per unit of x

// Fills the table; every thread of the workgroup must call it, and it ends
// with a barrier.
__device__ inline void tanh_table_init(float2* tab)
{
    for (int i = threadIdx.x; i < TANH_N; i += blockDim.x) {
        const float a = tanhf(TANH_X0 + i / TANH_RH);
        const float b = tanhf(TANH_X0 + (i + 1) / TANH_RH);
        tab[i] = make_float2(a, b - a);
    }
    __syncthreads();
}

__device__ inline float table_tanh(const float2* tab, float x)
{
    const float t = fminf(fmaxf(fmaf(x, TANH_RH, -TANH_X0 * TANH_RH), 0.0f), TANH_N - 0.5f);
    const float2 e = tab[(int)t];
    return fmaf(__builtin_amdgcn_fractf(t), e.y, e.x);
}

__device__ inline float safe_log2(float x)
{
    return __builtin_amdgcn_logf(x > 1e-30f ? x : 1e-30f);
}

// Raw (unnormalised) features of a per-volume conserved state.
__device__ inline void mlp_raw_features(double D, double S2, double tau, double gamma, float f[4])
{
    const double q = sqrt(D * D + S2);
    const double th = (tau + D) - q;
    const float l2d = safe_log2((float)D);
    f[0] = l2d;
    f[1] = 0.5f * safe_log2((float)S2) - l2d;
    f[2] = safe_log2((float)th) - l2d;
    f[3] = (float)gamma;
}

// Accumulator row held by lane group g in element v of a 32x32 tile.
__host__ __device__ inline int mfma_row(int g, int v) { return 8 * (v / 4) + 4 * g + (v % 4); }
// Hidden unit feeding k index kappa = 16c + 8g + i of layer 2.
__host__ __device__ inline int mlp_perm(int c, int g, int i) { return 16 * c + 8 * (i / 4) + 4 * g + (i % 4); }

// Wave-level inference: every lane passes its zone's normalised features
// and receives y = MLP(x) for its own zone.  Must be called by all 64 lanes;
// tab is the workgroup's tanh table (tanh_table_init).
__device__ inline float mlp_wave(const MlpDev& M, const float2* tab, const float x[4])
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;
    // Layer 1 B operand: lanes 0..31 carry features (k = 0..3) of zone
    // 32nt + lane; lanes 32..63 carry the zero padding k = 8..15.
    float xp[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) xp[m] = __shfl_xor(x[m], 32, 64);
    bf16x8 bx0, bx1;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        bx0[i] = (__bf16)((g == 0 && i < 4) ? x[i & 3] : 0.0f);
        bx1[i] = (__bf16)((g == 0 && i < 4) ? xp[i & 3] : 0.0f);
    }
    floatx16 h[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const bf16x8 a = M.a1[mt * 64 + lane];
        const floatx16 c = M.c1[mt * 2 + g];
        h[mt][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bx0, c, 0, 0, 0);
        h[mt][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bx1, c, 0, 0, 0);
    }
    // tanh, then the accumulators become layer 2's B operand (bf16).
    bf16x8 b2[4][2];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int i = 0; i < 8; ++i) b2[c][nt][i] = (__bf16)table_tanh(tab, h[c >> 1][nt][8 * (c & 1) + i]);
    floatx16 o[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const floatx16 cb = M.c2[mt * 2 + g];
        o[mt][0] = cb;
        o[mt][1] = cb;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const bf16x8 a = M.a2[(mt * 4 + c) * 64 + lane];
            o[mt][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b2[c][0], o[mt][0], 0, 0, 0);
            o[mt][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b2[c][1], o[mt][1], 0, 0, 0);
        }
    }
    // Layer 3: each lane sums its 32 hidden rows, the partner lane the rest.
    float part[2] = {0.0f, 0.0f};
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        const floatx16 w = M.w3[mt * 2 + g];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            part[0] = fmaf(w[v], table_tanh(tab, o[mt][0][v]), part[0]);
            part[1] = fmaf(w[v], table_tanh(tab, o[mt][1][v]), part[1]);
        }
    }
    part[0] += __shfl_xor(part[0], 32, 64);
    part[1] += __shfl_xor(part[1], 32, 64);
    return M.b3 + (g == 0 ? part[0] : part[1]);
}

// log2(p/D) guessed for one zone by the wave inference: y/tscale + tshift.
__device__ inline float mlp_log2_guess_wave(const MlpDev& M, const float2* tab, double D, double S2, double tau,
                                           double gamma)
{
    float f[4], x[4];
    mlp_raw_features(D, S2, tau, gamma, f);
#pragma unroll
    for (int m = 0; m < 4; ++m) x[m] = (f[m] - M.norm.shift[m]) * M.norm.scale[m];
    const float y = mlp_wave(M, tab, x);
    return y / M.norm.tscale + M.norm.tshift;
}

// Pressure guess from its log2 form: p = D 2^l2.
__device__ inline double mlp_guess_pressure(double D, float l2)
{
    return D * (double)__builtin_amdgcn_exp2f(l2);
}

}  // namespace knc
