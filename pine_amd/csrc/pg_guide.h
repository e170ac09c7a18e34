// Learned path-guiding network, device side: a 225-64-64-64-41 MLP whose
// output is an 8-lobe von Mises-Fisher mixture plus the share of directions
// drawn from it (the rest come from the BSDF; one-sample MIS over the two).
// Matrix products run on the matrix cores (v_mfma_f32_16x16x32_bf16: BF16
// operands, FP32 accumulation).
//
// MFMA 16x16x32 operand layout (lane l, r = l % 16, g = l / 16):
//   A (16 x 32): A[r][8g .. 8g+7];  B (32 x 16): B[8g .. 8g+7][r];
//   C (16 x 16): C[4g .. 4g+3][r].
// The forward pass computes Z^T = W X^T per layer, with a wave's 16 queries as
// the N dimension, so every lane works on one query (r) and the layer's C
// registers -- units 16t + 4g + i of tile t -- are exactly the B operand of
// the next layer's k-step s when that layer reads its inputs in the order
// k = 32s + 8g + i' -> unit 16(2s + i'/4) + 4g + i'%4 (PHI).  Activations
// never leave the registers between layers.
//
// Parameter vector (FP32 masters, 27,888 floats, padded):
//   W0 [64][256] | b0 [64] | W1 [64][64] | b1 | W2 [64][64] | b2 | W3 [48][64] | b3 [48]
// Live: W0 columns < 225, W3 / b3 rows < 41.  Device copies in BF16:
//   wf: W0 | W1 | W2 | W3 row-major (forward A operands);
//   wt: W1^T | W2^T | W3^T padded to [64][64] (backward A operands).
#pragma once

#include "pg_vec.h"
#include "../../include/pinehip.h"

#include <hip/hip_runtime.h>

namespace pg {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

namespace guide {

constexpr int kIn = 9, kFreq = 12, kFeat = 256, kLiveFeat = 9 + 2 * kFreq * 9;  // 225
constexpr int kH = 64, kOut = 41, kOutPad = 48, kLobes = 8;
constexpr int OFF_W0 = 0, OFF_B0 = OFF_W0 + kH * kFeat, OFF_W1 = OFF_B0 + kH;
constexpr int OFF_B1 = OFF_W1 + kH * kH, OFF_W2 = OFF_B1 + kH, OFF_B2 = OFF_W2 + kH * kH;
constexpr int OFF_W3 = OFF_B2 + kH, OFF_B3 = OFF_W3 + kOutPad * kH;
constexpr int kParams = OFF_B3 + kOutPad;  // 27888
// BF16 weight copies
constexpr int WF0 = 0, WF1 = WF0 + kH * kFeat, WF2 = WF1 + kH * kH, WF3 = WF2 + kH * kH;
constexpr int kWf = WF3 + kOutPad * kH;  // 27648
constexpr int WT1 = 0, WT2 = kH * kH, WT3 = 2 * kH * kH, kWt = 3 * kH * kH;
// alpha <= 0.5: a direction the guide misses keeps at least half the BSDF's
// density, so one vertex scales the throughput by at most 2x the albedo (a
// larger share made the estimator's error tail heavier than it gained).
constexpr float kAlphaMax = 0.5f, kKrMin = -4.0f, kKrMax = 7.0f;
// Output unit of lobe j = 2g + h, component c (mu.xyz, log kappa, logit):
// slot s = 5h + c of lane group g's twelve output registers (units 16t + 4g + i,
// s = 4t + i), so each lane holds its two lobes.  Alpha's logit takes lane
// group 0's slot 10.  The other units (s = 10, 11 elsewhere) are padding.
__host__ __device__ constexpr int out_unit(int j, int c) {
    return 16 * ((5 * (j & 1) + c) >> 2) + 4 * (j >> 1) + ((5 * (j & 1) + c) & 3);
}
constexpr int kAlphaUnit = 34;
__host__ __device__ constexpr bool live_out(int u) {
    const int s = 4 * (u >> 4) + (u & 3), g = (u >> 2) & 3;
    return u < kOutPad && (s < 10 || (g == 0 && s == 10));
}
constexpr float kQDefensive = 0.1f;  // uniform share in the guide density's own objective
constexpr float kInv4PiF = 0.0795774715459476679f;
constexpr float kTwoPiF = 6.28318530717958647692f;

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Encoding: column f = 8b + i of W0.  Blocks b < 24 hold sin (b even) or cos
// (b odd) of pi 2^(b/2 + oct) x_i for inputs i < 8; blocks 24..26 the 24
// (frequency, sin/cos) rows rho = 8(b - 24) + i of input 8; block 27 inputs
// 0..7 raw; column 224 input 8 raw; columns 225.. are zero.  A lane's B
// operand of k-step s is block 4s + g, so its inputs are compile-time
// registers and only the octave and the phase depend on the lane.
//
// sin / cos(pi 2^e' x) = sin(2 pi (frac(2^(e'-1) x) + 0 / 0.25)): exact range
// reduction, then the hardware sine, which takes revolutions.
__device__ __forceinline__ float periodic(float x, int e, float phase) {
    return __builtin_amdgcn_sinf(__builtin_amdgcn_fractf(ldexpf(x, e)) + phase);
}

// Where feature f comes from: kind 0 zero, 1 raw input c, 2 periodic of input
// c with exponent e and phase.
struct FeatureSource {
    int kind, c, e;
    float phase;
};
__device__ __forceinline__ FeatureSource feature_source(int f, int oct) {
    const int b = f >> 3, i = f & 7;
    if (b < 24)
        return {2, i, (b >> 1) + oct - 1, (b & 1) ? 0.25f : 0.0f};
    if (b < 27) {
        const int rho = 8 * (b - 24) + i;
        return {2, 8, (rho >> 1) + oct - 1, (rho & 1) ? 0.25f : 0.0f};
    }
    if (b == 27)
        return {1, i, 0, 0.0f};
    if (b == 28 && i == 0)
        return {1, 8, 0, 0.0f};
    return {0, 0, 0, 0.0f};
}
__device__ __forceinline__ float feature_value(const FeatureSource& fs, const float* x) {
    if (fs.kind == 2)
        return periodic(x[fs.c], fs.e, fs.phase);
    return fs.kind == 1 ? x[fs.c] : 0.0f;
}

// The B operand of layer 0's k-step s for this lane's query (inputs x[9] in
// registers): features 32s + 8g + i.
template <int s>
__device__ __forceinline__ bf16x8 encode8(const float* x, int g, int oct) {
    bf16x8 b;
    if constexpr (s < 6) {
        const int blk = 4 * s + g, e = (blk >> 1) + oct - 1;
        const float ph = (blk & 1) ? 0.25f : 0.0f;
#pragma unroll
        for (int i = 0; i < 8; i++)
            b[i] = (__bf16)periodic(x[i], e, ph);
    } else if constexpr (s == 6) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float v = periodic(x[8], 4 * g + (i >> 1) + oct - 1, (i & 1) ? 0.25f : 0.0f);
            b[i] = (__bf16)(g < 3 ? v : x[i]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++)
            b[i] = (__bf16)((g == 0 && i == 0) ? x[8] : 0.0f);
    }
    return b;
}

// Input k of a hidden layer -> unit of the previous layer (PHI), and back.
__host__ __device__ __forceinline__ int phi(int k) {
    const int s = k >> 5, g = (k >> 3) & 3, i = k & 7;
    return 16 * (2 * s + (i >> 2)) + 4 * g + (i & 3);
}
__host__ __device__ __forceinline__ int phi_inv(int m) {
    const int t = m >> 4, g = (m >> 2) & 3, i = m & 3;
    return 32 * (t >> 1) + 8 * g + 4 * (t & 1) + i;
}

__device__ __forceinline__ bf16x8 load8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

// Bias + ReLU of a hidden layer's four C tiles, packed as the next layer's two
// B operands (k-steps 0 and 1).
__device__ __forceinline__ void relu_pack(const f32x4 acc[4], const float* bias, int g, float h[16],
                                          bf16x8 bop[2]) {
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float z = acc[t][i] + bias[16 * t + 4 * g + i];
            h[4 * t + i] = z > 0.0f ? z : 0.0f;
        }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int i = 0; i < 8; i++)
            bop[s][i] = (__bf16)h[4 * (2 * s + (i >> 2)) + (i & 3)];
}

// Hidden layer with K = 64: four 16-unit tiles from the two B operands.
__device__ __forceinline__ void hidden64(const __bf16* W, int ldw, const bf16x8 bop[2], int r,
                                         int g, f32x4 acc[4]) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
        acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; s++)
            acc[t] = mfma(load8(W + (16 * t + r) * ldw + 32 * s + 8 * g), bop[s], acc[t]);
    }
}

// Forward pass of a wave's 16 queries (query r of this lane, inputs x[9]).
// W* point at row-major BF16 weights with row strides ld0 (W0) and ld (others);
// params holds the FP32 biases.  On return out[t][i] is raw output
// 16t + 4g + i (t < 3), and h1..h3 (if non-null) the post-ReLU units
// 16t + 4g + i of each hidden layer.
__device__ __forceinline__ void forward16(const float* x, const __bf16* W0, int ld0,
                                          const __bf16* W1, const __bf16* W2, const __bf16* W3,
                                          int ld, const float* params, int oct, f32x4 out[3],
                                          float* h1, float* h2, float* h3) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++)
        acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    auto step = [&](const bf16x8 b, int s) {
#pragma unroll
        for (int t = 0; t < 4; t++)
            acc[t] = mfma(load8(W0 + (16 * t + r) * ld0 + 32 * s + 8 * g), b, acc[t]);
    };
    step(encode8<0>(x, g, oct), 0);
    step(encode8<1>(x, g, oct), 1);
    step(encode8<2>(x, g, oct), 2);
    step(encode8<3>(x, g, oct), 3);
    step(encode8<4>(x, g, oct), 4);
    step(encode8<5>(x, g, oct), 5);
    step(encode8<6>(x, g, oct), 6);
    step(encode8<7>(x, g, oct), 7);
    float h[16];
    bf16x8 bop[2];
    relu_pack(acc, params + OFF_B0, g, h, bop);
    if (h1)
        for (int k = 0; k < 16; k++)
            h1[k] = h[k];
    hidden64(W1, ld, bop, r, g, acc);
    relu_pack(acc, params + OFF_B1, g, h, bop);
    if (h2)
        for (int k = 0; k < 16; k++)
            h2[k] = h[k];
    hidden64(W2, ld, bop, r, g, acc);
    relu_pack(acc, params + OFF_B2, g, h, bop);
    if (h3)
        for (int k = 0; k < 16; k++)
            h3[k] = h[k];
#pragma unroll
    for (int t = 0; t < 3; t++) {
        f32x4 a = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < 2; s++)
            a = mfma(load8(W3 + (16 * t + r) * ld + 32 * s + 8 * g), bop[s], a);
#pragma unroll
        for (int i = 0; i < 4; i++)
            a[i] += params[OFF_B3 + 16 * t + 4 * g + i];
        out[t] = a;
    }
}

// Mixture parameters of one query from its raw outputs o[48]:
// p[6j..6j+5] = (mu.xyz, kappa, pi, vMF normaliser kappa/(2 pi (1 - e^-2kappa))),
// p[48] = alpha (guide share); records are 52 floats (16-byte aligned).
constexpr int kParamStride = 52, kAlphaAt = 48;
__device__ __forceinline__ void head(const float* o, float* p) {
    float lmax = -3.0e38f;
    for (int j = 0; j < kLobes; j++)
        lmax = fmaxf(lmax, o[out_unit(j, 4)]);
    float esum = 0.0f, e[kLobes];
    for (int j = 0; j < kLobes; j++) {
        e[j] = expf(o[out_unit(j, 4)] - lmax);
        esum += e[j];
    }
    for (int j = 0; j < kLobes; j++) {
        float mx = o[out_unit(j, 0)], my = o[out_unit(j, 1)], mz = o[out_unit(j, 2)];
        const float n = fmaxf(sqrtf(mx * mx + my * my + mz * mz), 1e-6f);
        const float kap = expf(fminf(fmaxf(o[out_unit(j, 3)], kKrMin), kKrMax));
        p[6 * j + 0] = mx / n;
        p[6 * j + 1] = my / n;
        p[6 * j + 2] = mz / n;
        p[6 * j + 3] = kap;
        p[6 * j + 4] = e[j] / esum;
        p[6 * j + 5] = kap / (kTwoPiF * -expm1f(-2.0f * kap));
    }
    p[kAlphaAt] = kAlphaMax / (1.0f + expf(-o[kAlphaUnit]));
}

// The same parameters computed by the four lanes of a query straight from the
// output registers of forward16 (lane group g holds lobes 2g and 2g + 1 in
// slots 0..9 and, for g = 0, alpha in slot 10; the softmax reduces across
// lanes l, l ^ 16, l ^ 32), written to the 16-byte aligned record p.  Every
// lane of the wave must call it (cross-lane reductions); `store` selects the
// lanes that write.
__device__ __forceinline__ void head4(const f32x4 out[3], int g, bool store, float* p) {
    float v[12];
#pragma unroll
    for (int sl = 0; sl < 12; sl++)
        v[sl] = out[sl >> 2][sl & 3];
    float m = fmaxf(v[4], v[9]);
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    const float e0 = expf(v[4] - m), e1 = expf(v[9] - m);
    float es = e0 + e1;
    es += __shfl_xor(es, 16);
    es += __shfl_xor(es, 32);
    const float inv = 1.0f / es;
    float r[12];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const float mx = v[5 * h], my = v[5 * h + 1], mz = v[5 * h + 2];
        const float n = fmaxf(sqrtf(mx * mx + my * my + mz * mz), 1e-6f);
        const float kap = expf(fminf(fmaxf(v[5 * h + 3], kKrMin), kKrMax));
        r[6 * h + 0] = mx / n;
        r[6 * h + 1] = my / n;
        r[6 * h + 2] = mz / n;
        r[6 * h + 3] = kap;
        r[6 * h + 4] = (h ? e1 : e0) * inv;
        r[6 * h + 5] = kap / (kTwoPiF * -expm1f(-2.0f * kap));
    }
    if (store) {
        float4* d = reinterpret_cast<float4*>(p + 12 * g);
        d[0] = make_float4(r[0], r[1], r[2], r[3]);
        d[1] = make_float4(r[4], r[5], r[6], r[7]);
        d[2] = make_float4(r[8], r[9], r[10], r[11]);
        if (g == 0)
            p[kAlphaAt] = kAlphaMax / (1.0f + expf(-v[10]));
    }
}

// Guide density q(w) of a direction under mixture parameters p.
__device__ __forceinline__ float mix_pdf(const float* p, f3 w) {
    float q = 0.0f;
#pragma unroll
    for (int j = 0; j < kLobes; j++) {
        const float c = p[6 * j] * w.x + p[6 * j + 1] * w.y + p[6 * j + 2] * w.z;
        q += p[6 * j + 4] * p[6 * j + 5] * expf(p[6 * j + 3] * (c - 1.0f));
    }
    return q;
}

// Draws a direction from the mixture: lobe by u0, then vMF(mu, kappa) by
// (u1, u2) with cos = 1 + log(u1 + (1 - u1) e^-2kappa) / kappa.
__device__ __forceinline__ f3 mix_sample(const float* p, float u0, float u1, float u2) {
    int j = 0;
    float acc = p[4];
    while (j < kLobes - 1 && u0 >= acc) {
        j++;
        acc += p[6 * j + 4];
    }
    const f3 mu = mk(p[6 * j], p[6 * j + 1], p[6 * j + 2]);
    const float kap = p[6 * j + 3];
    float cw = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * kap)) / kap;
    cw = fminf(fmaxf(cw, -1.0f), 1.0f);
    const float sw = sqrtf(fmaxf(0.0f, 1.0f - cw * cw));
    const float ph = kTwoPiF * u2;
    // Orthonormal frame around mu (Duff et al. 2017).
    const float sign = mu.z >= 0.0f ? 1.0f : -1.0f;
    const float a = -1.0f / (sign + mu.z), b = mu.x * mu.y * a;
    const f3 t = mk(1.0f + sign * mu.x * mu.x * a, sign * b, -sign * mu.x);
    const f3 bt = mk(b, sign + mu.y * mu.y * a, -mu.y);
    return t * (sw * __cosf(ph)) + bt * (sw * __sinf(ph)) + mu * cw;
}

}  // namespace guide

// Host interface of the network for the path tracer (guide.hip).
const float* guide_params_dev(const pg_guide* G);
int guide_freq_offset(const pg_guide* G);
const __bf16* guide_wf_dev(const pg_guide* G);
// One Adam step on sequence positions [first, first + count) of n slots, slot
// of position s = (s * perm_a) mod n (perm_a coprime to n).
int guide_train_step(pg_guide* G, const float* x, const float* dir, const float* weight,
                     const float* pb, int64_t n, int64_t first, int64_t count, uint64_t perm_a,
                     hipStream_t st);
}  // namespace pg
