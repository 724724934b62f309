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

#include "knc_zone.hpp"

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

// Inference image: the weights as MFMA operand fragments plus the tanh
// table, built on the host and copied whole into each workgroup's LDS
// (MlpShared).  Fragments (one bf16x8 per lane):
//   a1[mt][lane]      layer-1 A operand, rows = hidden 32mt + lane%32,
//                     k = 8(lane/32) + i: W1 for k = 0..3, b1 for k = 4
//                     (the B operand carries 1 there), zero beyond
//   a2b[mt][lane]     layer-2 bias as an A operand: b2 of row 32mt + lane%32
//                     at k = 0, zero elsewhere (B operand: 1 at k = 0)
//   a2[mt][c][lane]   layer-2 A operand, rows = output 32mt + lane%32,
//                     k-chunk c with the hidden units permuted so that the
//                     layer-1 accumulator registers are the B operand as is
//   a3[c][lane]       layer-3 A operand: row 0 (lanes 0 and 32) holds W3 in
//                     the same permuted k order, rows 1..31 are zero
//   tanh              the TANH_N BF16 tanh entries
constexpr int MLP_IMAGE_FRAGS = 2 * 64 + 2 * 64 + 2 * 4 * 64 + 4 * 64;
constexpr int TANH_N = 2048;
constexpr int MLP_IMAGE_BYTES = MLP_IMAGE_FRAGS * 16 + TANH_N * 2;   // 20 KiB

struct MlpShared {
    bf16x8 a1[2 * 64];
    bf16x8 a2b[2 * 64];
    bf16x8 a2[2 * 4 * 64];
    bf16x8 a3[4 * 64];
    uint16_t tanh[TANH_N];
};
static_assert(sizeof(MlpShared) == MLP_IMAGE_BYTES, "inference image layout");

struct MlpDev {
    const uint4* image;   // MLP_IMAGE_BYTES, the MlpShared layout
    float b3;
    MlpNorm norm;
};

// tanh(x) = 1 - 2/(2^(2x/ln 2) + 1), from the hardware exp2 and reciprocal.
__device__ inline float fast_tanh(float x)
{
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

// Inference tanh: a table of TANH_N BF16 values in LDS, entry m holding
// tanh((m - 1024)/256) rounded to BF16, read at the nearest entry.  The
// index is formed two zones at a time in packed FP16: m 2^-24 =
// x 2^-16 + 2^-14, clamped to [0, 2047 2^-24].  Below 2^-13 the FP16 spacing
// is 2^-24, so the bit pattern of the rounded value is the integer m
// (subnormals included, which the kernels keep: FP16 denormal mode is
// IEEE).  For |x| < 4, |x - x_m| < 2^-8 (FP16 truncation of x, below 2^-9,
// plus half an entry), so the entry is within 2^-8 of tanh(x) before its
// BF16 rounding; beyond, tanh(x) rounds to +-1 in BF16 like the end
// entries.  The entries are part of the inference image.
constexpr int TANH_MID = 1024;       // entry of x = 0
constexpr float TANH_RH = 256.0f;    // entries per unit of x

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef short short2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Copies the inference image into the workgroup's LDS; every thread of the
// workgroup must call it, and it ends with a barrier.
__device__ inline void mlp_shared_load(MlpShared* sh, const MlpDev& M)
{
    uint4* dst = reinterpret_cast<uint4*>(sh);
    for (int i = threadIdx.x; i < MLP_IMAGE_BYTES / 16; i += blockDim.x) dst[i] = M.image[i];
    __syncthreads();
}

// Table indexes of two FP32 values, packed as two 16-bit halves (a low).
__device__ inline uint32_t tanh_index_pair(float a, float b)
{
    const half2_t x = __builtin_bit_cast(half2_t, __builtin_amdgcn_cvt_pkrtz(a, b));
    half2_t m = __builtin_elementwise_fma(x, (half2_t)(_Float16)0x1p-16f, (half2_t)(_Float16)0x1p-14f);
    m = __builtin_elementwise_max(m, (half2_t)(_Float16)0.0f);
    m = __builtin_elementwise_min(m, (half2_t)(_Float16)(2047 * 0x1p-24f));
    return __builtin_bit_cast(uint32_t, m);
}

// Eight accumulator values -> an MFMA operand of their tanh.  The four
// index pairs are formed first and the eight table reads issued together,
// so that one LDS round trip covers all eight.
__device__ inline bf16x8 tanh_operand(const uint16_t* tab, const float* v)
{
    uint32_t bits[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bits[q] = tanh_index_pair(v[2 * q], v[2 * q + 1]);
    uint16_t lo[4], hi[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        lo[q] = tab[bits[q] & 0xffffu];
        hi[q] = tab[bits[q] >> 16];
    }
    u32x4 w;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        short2_t r;
        r.x = (short)lo[q];
        r.y = (short)hi[q];
        w[q] = __builtin_bit_cast(uint32_t, r);
    }
    return __builtin_bit_cast(bf16x8, w);
}

__device__ inline float safe_log2(float x)
{
    return __builtin_amdgcn_logf(x > 1e-30f ? x : 1e-30f);
}

// Raw (unnormalised) features of a per-volume conserved state.
__device__ inline void mlp_raw_features(double D, double S2, double tau, double gamma, float f[4])
{
    const double q = sqrt_rn(D * D + S2);
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
// sh is the workgroup's inference image (mlp_shared_load).  The 64 zones go
// through the network as two tiles of 32 (the MFMA N dimension), one after
// the other, which halves the live accumulators and lets more wavefronts
// share a SIMD.
__device__ inline float mlp_wave(const MlpDev& M, const MlpShared* sh, const float x[4])
{
    const int lane = threadIdx.x & 63;
    const int g = lane >> 5;
    const floatx16 zero = {};
    // B operand with 1 at k = 0 (layer-2 bias) for lane group 0
    bf16x8 one;
#pragma unroll
    for (int i = 0; i < 8; ++i) one[i] = (__bf16)((g == 0 && i == 0) ? 1.0f : 0.0f);
    float y0 = 0.0f, y1 = 0.0f;
#pragma unroll 1
    for (int nt = 0; nt < 2; ++nt) {
        // The weight fragments are read from LDS for every tile: an opaque
        // copy of the lane index keeps the compiler from hoisting the reads
        // out of the loop into 64 long-lived VGPRs, which would cost spills
        // at four wavefronts per SIMD.
        int ln = lane;
        asm volatile("" : "+v"(ln));
        // Layer 1 B operand: lanes 0..31 carry the features (k = 0..3) of
        // zone 32nt + lane and 1 at k = 4; lanes 32..63 carry k = 8..15 = 0.
        float xs[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const float o = __shfl_xor(x[m], 32, 64);
            xs[m] = nt == 0 ? x[m] : o;
        }
        bf16x8 bx;
#pragma unroll
        for (int i = 0; i < 8; ++i) bx[i] = (__bf16)(g == 0 ? (i < 4 ? xs[i & 3] : (i == 4 ? 1.0f : 0.0f)) : 0.0f);
        floatx16 h[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) h[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sh->a1[mt * 64 + ln], bx, zero, 0, 0, 0);
        // tanh; the accumulators become layer 2's B operand (bf16).
        bf16x8 b2[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = h[c >> 1][8 * (c & 1) + i];
            b2[c] = tanh_operand(sh->tanh, v);
        }
        floatx16 o[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            o[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sh->a2b[mt * 64 + ln], one, zero, 0, 0, 0);
#pragma unroll
            for (int c = 0; c < 4; ++c)
                o[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sh->a2[(mt * 4 + c) * 64 + ln], b2[c], o[mt], 0, 0, 0);
        }
        // Layer 3 on the same MFMA, with W3 as row 0 of the A operand:
        // element 0 of lanes 0..31 receives W3 . tanh(o) of zone 32nt + lane.
        floatx16 acc = zero;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = o[c >> 1][8 * (c & 1) + i];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(sh->a3[c * 64 + ln], tanh_operand(sh->tanh, v), acc, 0, 0, 0);
        }
        if (nt == 0)
            y0 = acc[0];
        else
            y1 = acc[0];
    }
    const float hi = __shfl_xor(y1, 32, 64);
    return (g == 0 ? y0 : hi) + M.b3;
}

// log2(p/D) guessed for one zone by the wave inference: y/tscale + tshift.
__device__ inline float mlp_log2_guess_wave(const MlpDev& M, const MlpShared* sh, double D, double S2, double tau,
                                           double gamma)
{
    float f[4], x[4];
    mlp_raw_features(D, S2, tau, gamma, f);
#pragma unroll
    for (int m = 0; m < 4; ++m) x[m] = (f[m] - M.norm.shift[m]) * M.norm.scale[m];
    const float y = mlp_wave(M, sh, x);
    return y / M.norm.tscale + M.norm.tshift;
}

// Pressure guess from its log2 form: p = D 2^l2.  An entry of -inf (a zone
// with no guess yet) gives the plain guess p = 0; a NaN (a failed network)
// fails Newton, which then restarts from the plain guess and counts it.
__device__ inline double mlp_guess_pressure(double D, float l2)
{
    return D * (double)__builtin_amdgcn_exp2f(l2);
}

// The log2 form of a recovered pressure, log2(p/D) in FP32 (about 1e-7
// relative: a Newton starting point, not a result).
__device__ inline float guess_log2(double p, double D)
{
    return __builtin_amdgcn_logf((float)(p * __builtin_amdgcn_rcp(D)));
}

}  // namespace knc
