// Learned path-guiding network: training and inference kernels and the C ABI
// (pg_guide_*).  pine has no counterpart; the network is defined in
// pg_guide.h, and its float64 restatement (oracle/guide.py) is the checker.
//
// Training step over a batch of n radiance samples (x, w, weight, p_bsdf):
//   k_guide_grad    one workgroup per 64 samples: forward (activations kept in
//                   LDS, transposed), head gradient in FP32, backward through
//                   the layers as MFMA tile products over LDS-staged operands;
//                   writes its partial gradient (27,888 floats) to its own row
//   k_guide_adam    sums the partial rows in a fixed order (deterministic),
//                   applies Adam to the FP32 masters and refreshes the BF16
//                   forward and transposed weight copies
#include "pg_api.h"
#include "pg_guide.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace pg;
using namespace pg::guide;

#define PG_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return pg_fail(PG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

struct pg_guide {
    float* params = nullptr;    // FP32 masters
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float* grad = nullptr;      // last reduced gradient
    float* partials = nullptr;  // [max_blocks][kParams]
    float* loss_part = nullptr; // [max_blocks]
    __bf16* wf = nullptr;
    __bf16* wt = nullptr;
    int max_blocks = 0;
    int last_blocks = 0;  // workgroups of the last gradient launch
    int64_t step = 0;
    float lr = 1e-3f;
    uint64_t seed = 1;
    int oct = 0;  // encoding octave offset

    void free_all() {
        for (void* p : {(void*)params, (void*)adam_m, (void*)adam_v, (void*)grad, (void*)partials,
                        (void*)loss_part, (void*)wf, (void*)wt})
            if (p)
                (void)hipFree(p);
    }
};

namespace {

constexpr int kQ = 64;          // samples per training workgroup
constexpr int kLdq = kQ + 8;    // LDS row stride of [unit][sample] tiles (BF16)
constexpr int kLd0 = kFeat + 8; // LDS row strides of staged weights (BF16)
constexpr int kLd = kH + 8;

__device__ __forceinline__ bool live_param(int p) {
    if (p >= OFF_W0 && p < OFF_B0)
        return (p - OFF_W0) % kFeat < kLiveFeat;
    if (p >= OFF_W3 && p < OFF_B3)
        return live_out((p - OFF_W3) / kH);
    if (p >= OFF_B3)
        return live_out(p - OFF_B3);
    return true;
}

// Writes the BF16 copies of parameter p (forward row-major and transposed).
__device__ __forceinline__ void refresh_copy(const float v, int p, __bf16* wf, __bf16* wt) {
    const __bf16 b = (__bf16)v;
    if (p < OFF_B0) {
        wf[WF0 + (p - OFF_W0)] = b;
    } else if (p >= OFF_W1 && p < OFF_B1) {
        const int m = (p - OFF_W1) / kH, k = (p - OFF_W1) % kH;
        wf[WF1 + m * kH + k] = b;
        wt[WT1 + k * kH + m] = b;
    } else if (p >= OFF_W2 && p < OFF_B2) {
        const int m = (p - OFF_W2) / kH, k = (p - OFF_W2) % kH;
        wf[WF2 + m * kH + k] = b;
        wt[WT2 + k * kH + m] = b;
    } else if (p >= OFF_W3 && p < OFF_B3) {
        const int m = (p - OFF_W3) / kH, k = (p - OFF_W3) % kH;
        wf[WF3 + m * kH + k] = b;
        wt[WT3 + k * kH + m] = b;
    }
}

__global__ void k_guide_refresh(const float* params, __bf16* wf, __bf16* wt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kParams)
        refresh_copy(params[p], p, wf, wt);
}

// C tile (16x16) = A (16 x K) . Bt (16 x K)^T, both BF16 row-major with strides
// lda / ldb (elements).  Lane holds C[4g + i][r].
__device__ __forceinline__ f32x4 tile(const __bf16* A, int lda, const __bf16* Bt, int ldb, int K) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 32)
        acc = mfma(load8(A + r * lda + k0 + 8 * g), load8(Bt + r * ldb + k0 + 8 * g), acc);
    return acc;
}

struct Batch {
    const float* x;       // n x 9
    const float* dir;     // n x 3
    const float* weight;  // n
    const float* pbsdf;   // n
    int64_t n;            // slots in the arrays
    int64_t first, end;   // positions [first, end) of the sample sequence in this launch
    float inv_n;          // 1 / samples in the loss mean
    uint64_t perm_a;      // sample s of the sequence is slot (s * perm_a) mod n (1 = in order)
    int oct;              // encoding octave offset
};

__device__ __forceinline__ int64_t slot_of(const Batch& B, int64_t s) {
    return B.perm_a == 1 ? s : (int64_t)(((unsigned __int128)s * B.perm_a) % (uint64_t)B.n);
}

// Head gradient of one sample: d(-weight log q_mix)/d o, scaled by inv_n
// (oracle/guide.py loss_and_grad).
__device__ float head_grad(const float* o, f3 w, float weight, float pb, float inv_n, float* d) {
    float p[kParamStride];
    head(o, p);
    float v[kLobes], q = 0.0f;
    for (int j = 0; j < kLobes; j++) {
        const float c = p[6 * j] * w.x + p[6 * j + 1] * w.y + p[6 * j + 2] * w.z;
        v[j] = p[6 * j + 5] * expf(p[6 * j + 3] * (c - 1.0f));
        q += p[6 * j + 4] * v[j];
    }
    for (int k = 0; k < kOutPad; k++)
        d[k] = 0.0f;
    if (!(weight > 0.0f))
        return 0.0f;  // a sample that carried no radiance teaches nothing
    // Guide density by its own cross-entropy (defensive uniform share bounds the
    // gradient); alpha by the mixture's, the guide density held constant.
    const float sig = p[kAlphaAt] / kAlphaMax, alpha = p[kAlphaAt];
    static_assert(out_unit(0, 0) == 0 && out_unit(7, 4) == 45 && live_out(kAlphaUnit), "output units");
    const float qm = fmaxf(alpha * q + (1.0f - alpha) * pb, 1e-30f);
    const float qu = (1.0f - kQDefensive) * q + kQDefensive * kInv4PiF;
    const float g_q = -weight * (1.0f - kQDefensive) / qu * inv_n;
    d[kAlphaUnit] = -weight / qm * inv_n * (q - pb) * kAlphaMax * sig * (1.0f - sig);
    for (int j = 0; j < kLobes; j++) {
        const float mx = o[out_unit(j, 0)], my = o[out_unit(j, 1)], mz = o[out_unit(j, 2)];
        const float rn = sqrtf(mx * mx + my * my + mz * mz), nrm = fmaxf(rn, 1e-6f);
        const float kap = p[6 * j + 3], pi = p[6 * j + 4];
        const f3 mu = mk(p[6 * j], p[6 * j + 1], p[6 * j + 2]);
        const float c = mu.x * w.x + mu.y * w.y + mu.z * w.z;
        const float kr = o[out_unit(j, 3)];
        d[out_unit(j, 4)] = g_q * pi * (v[j] - q);
        const float em = expf(-2.0f * kap);
        const float dlogv = 1.0f / kap - 2.0f * em / (-expm1f(-2.0f * kap)) + (c - 1.0f);
        d[out_unit(j, 3)] = (kr > kKrMin && kr < kKrMax) ? g_q * pi * v[j] * dlogv * kap : 0.0f;
        const float s = g_q * pi * v[j] * kap;  // dq/dmu = s * w
        const f3 dmu = w * s;
        const float proj = dmu.x * mu.x + dmu.y * mu.y + dmu.z * mu.z;
        const float inv = rn > 1e-6f ? 1.0f / nrm : 0.0f;
        d[out_unit(j, 0)] = (dmu.x - proj * mu.x) * inv;
        d[out_unit(j, 1)] = (dmu.y - proj * mu.y) * inv;
        d[out_unit(j, 2)] = (dmu.z - proj * mu.z) * inv;
    }
    return -weight * (logf(qu) + logf(qm)) * inv_n;
}

// One workgroup (4 waves) per 64 samples; partial gradient row per workgroup.
__global__ __launch_bounds__(256) void k_guide_grad(const float* __restrict__ params,
                                                    const __bf16* __restrict__ wf,
                                                    const __bf16* __restrict__ wt, Batch B,
                                                    float* __restrict__ partials,
                                                    float* __restrict__ loss_part) {
    __shared__ float sx[kQ * kIn];
    __shared__ __attribute__((aligned(16))) __bf16 ut[3][kH * kLdq];  // layer inputs u_L transposed: [k][q]
    __shared__ float so[kQ * (kOutPad + 1)];
    // delta, sample-major [q][unit] (units 48..63 zero for the head) and unit-major [unit][q]
    __shared__ __attribute__((aligned(16))) __bf16 dq[kQ * kLd];
    __shared__ __attribute__((aligned(16))) __bf16 dt[kH * kLdq];
    __shared__ float spart[4][kH];  // per-wave partial bias sums (summed in wave order)
    __shared__ float sl[kQ];        // per-sample losses

    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, r = l & 15, g = l >> 4;
    const int64_t q0 = B.first + (int64_t)blockIdx.x * kQ;
    float* part = partials + (size_t)blockIdx.x * kParams;
    for (int k = tid; k < kQ * kIn; k += 256) {
        const int64_t q = q0 + k / kIn;
        sx[k] = q < B.end ? B.x[slot_of(B, q) * kIn + k % kIn] : 0.0f;
    }
    __syncthreads();

    // Forward: wave w owns samples 16w .. 16w+15; lane r works on sample 16w + r.
    {
        const int q = 16 * wave + r;
        f32x4 out[3];
        float h[3][16];
        forward16(sx + q * kIn, wf + WF0, kFeat, wf + WF1, wf + WF2, wf + WF3, kH, params, B.oct, out,
                  h[0], h[1], h[2]);
        for (int L = 0; L < 3; L++)
            for (int t = 0; t < 4; t++)
                for (int i = 0; i < 4; i++)
                    ut[L][phi_inv(16 * t + 4 * g + i) * kLdq + q] = (__bf16)h[L][4 * t + i];
        for (int t = 0; t < 3; t++)
            for (int i = 0; i < 4; i++)
                so[q * (kOutPad + 1) + 16 * t + 4 * g + i] = out[t][i];
    }
    __syncthreads();

    // Head: one thread per sample; delta rows into dq / dt (BF16) and, in FP32,
    // over the sample's own row of `so` for the bias sums.
    if (tid < kQ) {
        const int q = tid;
        const int64_t gq = q0 + q;
        float d[kOutPad];
        float lo = 0.0f;
        if (gq < B.end) {
            const int64_t sq = slot_of(B, gq);
            const f3 w = mk(B.dir[3 * sq], B.dir[3 * sq + 1], B.dir[3 * sq + 2]);
            lo = head_grad(so + q * (kOutPad + 1), w, B.weight[sq], B.pbsdf[sq], B.inv_n, d);
        } else {
            for (int k = 0; k < kOutPad; k++)
                d[k] = 0.0f;
        }
        for (int k = 0; k < kH; k++) {
            const float v = k < kOutPad ? d[k] : 0.0f;
            dq[q * kLd + k] = (__bf16)v;
            dt[k * kLdq + q] = (__bf16)v;
            if (k < kOutPad)
                so[q * (kOutPad + 1) + k] = v;
        }
        sl[q] = lo;
    }
    __syncthreads();
    // Sums in a fixed order: the result does not depend on scheduling.
    if (tid < kOutPad) {
        float b = 0.0f;
        for (int q = 0; q < kQ; q++)
            b += so[q * (kOutPad + 1) + tid];
        part[OFF_B3 + tid] = b;
    }
    if (tid == kOutPad) {
        float lsum = 0.0f;
        for (int q = 0; q < kQ; q++)
            lsum += sl[q];
        loss_part[blockIdx.x] = lsum;
    }

    // dW3 [48][64] = D3t . U3t^T: 12 tiles.
    for (int tI = wave; tI < 12; tI += 4) {
        const int mt = tI / 4, nt = tI % 4;
        const f32x4 c = tile(dt + 16 * mt * kLdq, kLdq, ut[2] + 16 * nt * kLdq, kLdq, kQ);
        for (int i = 0; i < 4; i++)
            part[OFF_W3 + (16 * mt + 4 * g + i) * kH + 16 * nt + r] = c[i];
    }

    // Backward through the hidden layers L = 2, 1, 0 (weights W_{L+1}).
    for (int L = 2; L >= 0; L--) {
        const __bf16* Wt = wt + (L == 2 ? WT3 : (L == 1 ? WT2 : WT1));
        // du[k][q] = sum_o Wt[k][o] * delta[q][o]: 16 tiles, K = 64.
        f32x4 c[4];
        for (int j = 0; j < 4; j++) {
            const int tI = wave + 4 * j, kt = tI / 4, qt = tI % 4;
            c[j] = tile(Wt + 16 * kt * kH, kH, dq + 16 * qt * kLd, kLd, kH);
        }
        // Mask by the ReLU (input k of layer L+1 is unit phi(k) of layer L).
        for (int j = 0; j < 4; j++) {
            const int q = 16 * wave + r;
            for (int i = 0; i < 4; i++) {
                const int k = 16 * j + 4 * g + i;
                c[j][i] = (float)ut[L][k * kLdq + q] > 0.0f ? c[j][i] : 0.0f;
            }
        }
        __syncthreads();  // every read of the previous delta is done
        for (int j = 0; j < 4; j++) {
            const int q = 16 * wave + r;
            for (int i = 0; i < 4; i++) {
                const int k = 16 * j + 4 * g + i, m = phi(k);
                const float v = c[j][i];
                dt[m * kLdq + q] = (__bf16)v;
                dq[q * kLd + m] = (__bf16)v;
                // Sum over the wave's 16 samples (lanes of equal g), fixed butterfly.
                float sum = v;
                for (int off = 8; off > 0; off >>= 1)
                    sum += __shfl_xor(sum, off, 16);
                if (r == 0)
                    spart[wave][m] = sum;
            }
        }
        __syncthreads();
        const int offb = L == 2 ? OFF_B2 : (L == 1 ? OFF_B1 : OFF_B0);
        if (tid < kH)
            part[offb + tid] = ((spart[0][tid] + spart[1][tid]) + spart[2][tid]) + spart[3][tid];
        if (L > 0) {
            // dW_L [64][64] = Dt . U_{L-1}t^T
            const int offw = L == 2 ? OFF_W2 : OFF_W1;
            for (int tI = wave; tI < 16; tI += 4) {
                const int mt = tI / 4, nt = tI % 4;
                const f32x4 cc = tile(dt + 16 * mt * kLdq, kLdq, ut[L - 1] + 16 * nt * kLdq, kLdq, kQ);
                for (int i = 0; i < 4; i++)
                    part[offw + (16 * mt + 4 * g + i) * kH + 16 * nt + r] = cc[i];
            }
        } else {
            // dW0 [64][256] = Dt . X^T, the features of each sample recomputed:
            // lane needs Xt[f = 16 ft + r][q = k0 + 8g + i].
            for (int tI = wave; tI < 64; tI += 4) {
                const int mt = tI / 16, ft = tI % 16, f = 16 * ft + r;
                const FeatureSource fs = feature_source(f, B.oct);
                f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                for (int k0 = 0; k0 < kQ; k0 += 32) {
                    bf16x8 b;
                    for (int i = 0; i < 8; i++)
                        b[i] = (__bf16)feature_value(fs, sx + (k0 + 8 * g + i) * kIn);
                    acc = mfma(load8(dt + (16 * mt + r) * kLdq + k0 + 8 * g), b, acc);
                }
                for (int i = 0; i < 4; i++)
                    part[OFF_W0 + (16 * mt + 4 * g + i) * kFeat + f] = acc[i];
            }
        }
        __syncthreads();  // dq / dt are overwritten by the next layer
    }
}

// Sums the partial rows (fixed order), then Adam on live parameters.
// Sums the per-workgroup partial gradients and takes one Adam step.  A block
// holds 16 parameters x 16 row groups: group g sums rows g, g + 16, ... of its
// parameter (reads of 16 consecutive parameters per row), then the 16 group
// sums are added in group order, so the result is the same on every run.
__global__ __launch_bounds__(256) void k_guide_adam(float* params, float* m, float* v, float* grad,
                                                    const float* partials, int nblocks, float lr, float bc1,
                                                    float bc2, int apply, __bf16* wf, __bf16* wt) {
    __shared__ float part[16][17];
    const int pl = threadIdx.x & 15, rg = threadIdx.x >> 4;
    const int p = blockIdx.x * 16 + pl;
    float s = 0.0f;
    if (p < kParams)
        for (int b = rg; b < nblocks; b += 16)
            s += partials[(size_t)b * kParams + p];
    part[rg][pl] = s;
    __syncthreads();
    if (rg != 0 || p >= kParams)
        return;
    float gsum = 0.0f;
    for (int g = 0; g < 16; g++)
        gsum += part[g][pl];
    if (!live_param(p))
        gsum = 0.0f;
    grad[p] = gsum;
    if (!apply || !live_param(p))
        return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.0f - b1) * gsum;
    const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    const float val = params[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    params[p] = val;
    refresh_copy(val, p, wf, wt);
}

// Inference: grid-stride over tiles of 64 queries (16 per wave); W1..W3 staged
// in LDS with padded rows, W0 (32 KB) read through the caches.
template <int kMode>  // 0 raw outputs, 1 pdf (q, alpha, q_mix), 2 sample (dir, q)
__global__ __launch_bounds__(256) void k_guide_apply(const float* __restrict__ params,
                                                     const __bf16* __restrict__ wf,
                                                     const float* __restrict__ x, int64_t n,
                                                     const float* __restrict__ aux,
                                                     const float* __restrict__ aux2,
                                                     float* __restrict__ out, int oct) {
    __shared__ __attribute__((aligned(16))) __bf16 sw[(2 * kH + kOutPad) * kLd];
    __shared__ float so[4][16 * (kOutPad + 1)];
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63, r = l & 15, g = l >> 4;
    for (int k = tid; k < 2 * kH * kH + kOutPad * kH; k += 256)
        sw[(k / kH) * kLd + k % kH] = wf[WF1 + k];
    __syncthreads();
    // Every wave of the block runs the same number of iterations (barriers).
    for (int64_t tile0 = (int64_t)blockIdx.x * 64; tile0 < n; tile0 += (int64_t)gridDim.x * 64) {
        const int64_t q = tile0 + 16 * wave + r;
        float xin[kIn];
        for (int c = 0; c < kIn; c++)
            xin[c] = q < n ? x[q * kIn + c] : 0.0f;
        f32x4 o[3];
        forward16(xin, wf + WF0, kFeat, sw, sw + kH * kLd, sw + 2 * kH * kLd, kLd, params, oct, o,
                  nullptr, nullptr, nullptr);
        float* s = so[wave];
        for (int t = 0; t < 3; t++)
            for (int i = 0; i < 4; i++)
                s[r * (kOutPad + 1) + 16 * t + 4 * g + i] = o[t][i];
        __syncthreads();
        if (g == 0 && q < n) {
            const float* oq = s + r * (kOutPad + 1);
            if (kMode == 0) {
                for (int k = 0; k < kOutPad; k++)
                    out[q * kOutPad + k] = oq[k];
            } else {
                float p[kParamStride];
                head(oq, p);
                if (kMode == 1) {
                    const f3 w = mk(aux[3 * q], aux[3 * q + 1], aux[3 * q + 2]);
                    const float qq = mix_pdf(p, w);
                    out[3 * q] = qq;
                    out[3 * q + 1] = p[kAlphaAt];
                    out[3 * q + 2] = p[kAlphaAt] * qq + (1.0f - p[kAlphaAt]) * aux2[q];
                } else {
                    const f3 w = mix_sample(p, aux[3 * q], aux[3 * q + 1], aux[3 * q + 2]);
                    out[4 * q] = w.x;
                    out[4 * q + 1] = w.y;
                    out[4 * q + 2] = w.z;
                    out[4 * q + 3] = mix_pdf(p, w);
                }
            }
        }
        __syncthreads();
    }
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// He-uniform hidden layers, 0.1x output layer, zero biases, padding zero.
std::vector<float> init_params(uint64_t seed) {
    std::vector<float> p(kParams, 0.0f);
    uint64_t s = seed;
    auto uni = [&]() { return (float)((splitmix(s) >> 40) * (1.0 / 16777216.0)) * 2.0f - 1.0f; };
    auto fill = [&](int off, int rows, int cols, bool out_rows, int live_cols, float scale) {
        const float lim = std::sqrt(6.0f / (float)cols) * scale;
        for (int m = 0; m < rows; m++)
            for (int k = 0; k < cols; k++) {
                const float u = uni();
                if ((!out_rows || live_out(m)) && k < live_cols)
                    p[off + m * cols + k] = lim * u;
            }
    };
    fill(OFF_W0, kH, kFeat, false, kLiveFeat, 1.0f);
    fill(OFF_W1, kH, kH, false, kH, 1.0f);
    fill(OFF_W2, kH, kH, false, kH, 1.0f);
    fill(OFF_W3, kOutPad, kH, true, kH, 0.1f);
    return p;
}

int upload_params(pg_guide* G, const float* h) {
    PG_HIP(hipMemcpy(G->params, h, kParams * sizeof(float), hipMemcpyHostToDevice));
    PG_HIP(hipMemset(G->adam_m, 0, kParams * sizeof(float)));
    PG_HIP(hipMemset(G->adam_v, 0, kParams * sizeof(float)));
    PG_HIP(hipMemset(G->wt, 0, kWt * sizeof(__bf16)));
    hipLaunchKernelGGL(k_guide_refresh, dim3((kParams + 255) / 256), dim3(256), 0, 0, G->params,
                       G->wf, G->wt);
    PG_HIP(hipGetLastError());
    PG_HIP(hipDeviceSynchronize());
    G->step = 0;
    return PG_OK;
}

int ensure_partials(pg_guide* G, int blocks) {
    if (blocks <= G->max_blocks)
        return PG_OK;
    if (G->partials)
        (void)hipFree(G->partials);
    if (G->loss_part)
        (void)hipFree(G->loss_part);
    G->partials = nullptr;
    G->loss_part = nullptr;
    PG_HIP(hipMalloc(&G->partials, (size_t)blocks * kParams * sizeof(float)));
    PG_HIP(hipMalloc(&G->loss_part, (size_t)blocks * sizeof(float)));
    G->max_blocks = blocks;
    return PG_OK;
}

int check_n(pg_guide* G, int64_t n, const void* a, const char* fn) {
    if (!G)
        return pg_fail(PG_E_ARG, std::string(fn) + ": null guide");
    if (n < 0 || (n > 0 && !a))
        return pg_fail(PG_E_ARG, std::string(fn) + ": bad buffers");
    return PG_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Internal interface for the path tracer.

namespace pg {
const float* guide_params_dev(const pg_guide* G) { return G->params; }
int guide_freq_offset(const pg_guide* G) { return G->oct; }
const __bf16* guide_wf_dev(const pg_guide* G) { return G->wf; }

// One Adam step on samples [first, first + count) of the arrays, with the
// partial-gradient rows and per-workgroup losses in `partials` / `loss_part`
// (at least (count + 63) / 64 rows).  Stream-ordered; allocates nothing.
static int enqueue_step(pg_guide* G, const float* x, const float* dir, const float* weight,
                        const float* pb, int64_t n, int64_t first, int64_t count, uint64_t perm_a,
                        float* partials, float* loss_part, hipStream_t st) {
    const int blocks = (int)((count + kQ - 1) / kQ);
    Batch B{x, dir, weight, pb, n, first, first + count, 1.0f / (float)count, perm_a, G->oct};
    hipLaunchKernelGGL(k_guide_grad, dim3(blocks), dim3(256), 0, st, G->params, G->wf, G->wt, B,
                       partials, loss_part);
    G->last_blocks = blocks;
    G->step++;
    const float bc1 = 1.0f - std::pow(0.9f, (float)G->step);
    const float bc2 = 1.0f - std::pow(0.999f, (float)G->step);
    hipLaunchKernelGGL(k_guide_adam, dim3((kParams + 15) / 16), dim3(256), 0, st, G->params,
                       G->adam_m, G->adam_v, G->grad, partials, blocks, G->lr, bc1, bc2, 1,
                       G->wf, G->wt);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

// One Adam step on samples [first, first + count) of the arrays.
int guide_train_step(pg_guide* G, const float* x, const float* dir, const float* weight,
                     const float* pb, int64_t n, int64_t first, int64_t count, uint64_t perm_a,
                     hipStream_t st) {
    if (int rc = ensure_partials(G, (int)((count + kQ - 1) / kQ)))
        return rc;
    return enqueue_step(G, x, dir, weight, pb, n, first, count, perm_a, G->partials, G->loss_part, st);
}
}  // namespace pg

// ---------------------------------------------------------------------------
// Training without allocation (pg_guide_train_static): the partial-gradient
// rows and per-workgroup losses of one step live in a file-scope device array,
// one copy per GPU, held until the process ends.  Its device address is taken
// once per device with hipGetSymbolAddress; host code uses only that address.

namespace {

constexpr int kStaticBlocks = PG_GUIDE_STATIC_MAX_BATCH / kQ;
static_assert(PG_GUIDE_STATIC_MAX_BATCH % kQ == 0, "static batch limit: whole workgroups");
constexpr size_t kScratchFloats = (size_t)kStaticBlocks * kParams + kStaticBlocks;

__device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
float* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of this GPU's scratch array (the current device's).
int scratch_of_current_device(float** out) {
    int dev = 0;
    PG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return pg_fail(PG_E_ARG, "pg_guide_train_static: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        PG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return pg_fail(PG_E_HIP, "pg_guide_train_static: no device address for the scratch array");
        g_scratch[dev] = static_cast<float*>(p);
    }
    *out = g_scratch[dev];
    return PG_OK;
}

// The mean loss of a step: its per-workgroup losses summed in workgroup order
// in double precision, as pg_guide_train sums them on the host.
__global__ void k_loss_sum(const float* loss_part, int blocks, float* out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < blocks; b++)
            s += loss_part[b];
        *out = (float)s;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" int pg_guide_create(const pg_guide_config* cfg, pg_guide** out) {
    if (!out)
        return pg_fail(PG_E_ARG, "pg_guide_create: out is null");
    pg_guide* G = new (std::nothrow) pg_guide();
    if (!G)
        return pg_fail(PG_E_NOMEM, "pg_guide_create: out of host memory");
    G->seed = cfg ? cfg->seed : 1;
    G->lr = cfg && cfg->learning_rate > 0.0f ? cfg->learning_rate : 1e-3f;
    G->oct = cfg ? cfg->freq_offset : 0;
    if (G->oct < -16 || G->oct > 8) {
        delete G;
        return pg_fail(PG_E_ARG, "pg_guide_create: freq_offset out of range");
    }
    hipError_t e = hipSuccess;
    for (float** p : {&G->params, &G->adam_m, &G->adam_v, &G->grad})
        if (e == hipSuccess)
            e = hipMalloc(p, kParams * sizeof(float));
    if (e == hipSuccess)
        e = hipMalloc(&G->wf, kWf * sizeof(__bf16));
    if (e == hipSuccess)
        e = hipMalloc(&G->wt, kWt * sizeof(__bf16));
    if (e != hipSuccess) {
        G->free_all();
        delete G;
        return pg_fail(PG_E_HIP, std::string("pg_guide_create: ") + hipGetErrorString(e));
    }
    std::vector<float> p = init_params(G->seed);
    if (int rc = upload_params(G, p.data())) {
        G->free_all();
        delete G;
        return rc;
    }
    *out = G;
    return PG_OK;
}

extern "C" int pg_guide_set_params(pg_guide* G, const float* h_params) {
    if (!G || !h_params)
        return pg_fail(PG_E_ARG, "pg_guide_set_params: null argument");
    return upload_params(G, h_params);
}

extern "C" int pg_guide_get_params(const pg_guide* G, float* h_params) {
    if (!G || !h_params)
        return pg_fail(PG_E_ARG, "pg_guide_get_params: null argument");
    PG_HIP(hipMemcpy(h_params, G->params, kParams * sizeof(float), hipMemcpyDeviceToHost));
    return PG_OK;
}

static int apply_grid(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 63) / 64, 1024)); }

extern "C" int pg_guide_forward(pg_guide* G, const float* d_inputs, int64_t n, float* d_out,
                                void* stream) {
    if (int rc = check_n(G, n, d_inputs && d_out ? d_inputs : nullptr, "pg_guide_forward"))
        return rc;
    if (n == 0)
        return PG_OK;
    hipLaunchKernelGGL(k_guide_apply<0>, dim3(apply_grid(n)), dim3(256), 0, (hipStream_t)stream,
                       G->params, G->wf, d_inputs, n, nullptr, nullptr, d_out, G->oct);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

extern "C" int pg_guide_pdf(pg_guide* G, const float* d_inputs, const float* d_dirs,
                            const float* d_pbsdf, int64_t n, float* d_out, void* stream) {
    const bool ok = d_inputs && d_dirs && d_pbsdf && d_out;
    if (int rc = check_n(G, n, ok ? d_inputs : nullptr, "pg_guide_pdf"))
        return rc;
    if (n == 0)
        return PG_OK;
    hipLaunchKernelGGL(k_guide_apply<1>, dim3(apply_grid(n)), dim3(256), 0, (hipStream_t)stream,
                       G->params, G->wf, d_inputs, n, d_dirs, d_pbsdf, d_out, G->oct);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

extern "C" int pg_guide_sample(pg_guide* G, const float* d_inputs, const float* d_u, int64_t n,
                               float* d_out, void* stream) {
    const bool ok = d_inputs && d_u && d_out;
    if (int rc = check_n(G, n, ok ? d_inputs : nullptr, "pg_guide_sample"))
        return rc;
    if (n == 0)
        return PG_OK;
    hipLaunchKernelGGL(k_guide_apply<2>, dim3(apply_grid(n)), dim3(256), 0, (hipStream_t)stream,
                       G->params, G->wf, d_inputs, n, d_u, nullptr, d_out, G->oct);
    PG_HIP(hipGetLastError());
    return PG_OK;
}

extern "C" int pg_guide_gradient(pg_guide* G, const float* d_inputs, const float* d_dirs,
                                 const float* d_weight, const float* d_pbsdf, int64_t n,
                                 float* h_grad, float* h_loss) {
    const bool ok = d_inputs && d_dirs && d_weight && d_pbsdf && h_grad;
    if (int rc = check_n(G, n, ok ? d_inputs : nullptr, "pg_guide_gradient"))
        return rc;
    if (n == 0)
        return pg_fail(PG_E_ARG, "pg_guide_gradient: empty batch");
    const int blocks = (int)((n + kQ - 1) / kQ);
    if (int rc = ensure_partials(G, blocks))
        return rc;
    Batch B{d_inputs, d_dirs, d_weight, d_pbsdf, n, 0, n, 1.0f / (float)n, 1, G->oct};
    hipLaunchKernelGGL(k_guide_grad, dim3(blocks), dim3(256), 0, 0, G->params, G->wf, G->wt, B,
                       G->partials, G->loss_part);
    hipLaunchKernelGGL(k_guide_adam, dim3((kParams + 15) / 16), dim3(256), 0, 0, G->params,
                       G->adam_m, G->adam_v, G->grad, G->partials, blocks, 0.0f, 1.0f, 1.0f, 0,
                       G->wf, G->wt);
    PG_HIP(hipGetLastError());
    PG_HIP(hipMemcpy(h_grad, G->grad, kParams * sizeof(float), hipMemcpyDeviceToHost));
    if (h_loss) {
        std::vector<float> lp(blocks);
        PG_HIP(hipMemcpy(lp.data(), G->loss_part, blocks * sizeof(float), hipMemcpyDeviceToHost));
        double s = 0.0;
        for (float v : lp)
            s += v;
        *h_loss = (float)s;
    }
    return PG_OK;
}

extern "C" int pg_guide_train(pg_guide* G, const float* d_inputs, const float* d_dirs,
                              const float* d_weight, const float* d_pbsdf, int64_t n, int32_t steps,
                              int32_t batch, float* h_loss) {
    const bool ok = d_inputs && d_dirs && d_weight && d_pbsdf;
    if (int rc = check_n(G, n, ok ? d_inputs : nullptr, "pg_guide_train"))
        return rc;
    if (n == 0 || steps < 0 || batch <= 0)
        return pg_fail(PG_E_ARG, "pg_guide_train: empty batch or bad step count");
    int64_t first = 0;
    for (int s = 0; s < steps; s++) {
        const int64_t count = std::min<int64_t>(batch, n - first);
        if (int rc = guide_train_step(G, d_inputs, d_dirs, d_weight, d_pbsdf, n, first, count, 1, 0))
            return rc;
        first += count;
        if (first >= n)
            first = 0;
    }
    PG_HIP(hipDeviceSynchronize());
    if (h_loss) {
        const int blocks = G->last_blocks;
        std::vector<float> lp(blocks);
        PG_HIP(hipMemcpy(lp.data(), G->loss_part, blocks * sizeof(float), hipMemcpyDeviceToHost));
        double s = 0.0;
        for (float v : lp)
            s += v;
        *h_loss = (float)s;
    }
    return PG_OK;
}

extern "C" int pg_guide_train_static(pg_guide* G, const float* d_inputs, const float* d_dirs,
                                     const float* d_weight, const float* d_pbsdf, int64_t n,
                                     int32_t steps, int32_t batch, float* d_loss, void* stream) {
    const bool ok = d_inputs && d_dirs && d_weight && d_pbsdf;
    if (int rc = check_n(G, n, ok ? d_inputs : nullptr, "pg_guide_train_static"))
        return rc;
    if (n == 0 || steps < 0 || batch <= 0)
        return pg_fail(PG_E_ARG, "pg_guide_train_static: empty batch or bad step count");
    if (batch > PG_GUIDE_STATIC_MAX_BATCH)
        return pg_fail(PG_E_ARG, "pg_guide_train_static: batch " + std::to_string(batch) +
                                     " exceeds PG_GUIDE_STATIC_MAX_BATCH (" +
                                     std::to_string(PG_GUIDE_STATIC_MAX_BATCH) + ")");
    float* sc = nullptr;
    if (int rc = scratch_of_current_device(&sc))
        return rc;
    float* partials = sc;
    float* loss_part = sc + (size_t)kStaticBlocks * kParams;
    const hipStream_t st = (hipStream_t)stream;
    int64_t first = 0;
    for (int s = 0; s < steps; s++) {
        const int64_t count = std::min<int64_t>(batch, n - first);
        if (int rc = enqueue_step(G, d_inputs, d_dirs, d_weight, d_pbsdf, n, first, count, 1, partials,
                                  loss_part, st))
            return rc;
        first += count;
        if (first >= n)
            first = 0;
    }
    if (d_loss && steps > 0) {
        hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(64), 0, st, loss_part, G->last_blocks, d_loss);
        PG_HIP(hipGetLastError());
    }
    return PG_OK;
}

extern "C" int pg_guide_release(pg_guide* G) {
    if (!G)
        return PG_OK;
    G->free_all();
    delete G;
    return PG_OK;
}
