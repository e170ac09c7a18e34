// Shading library of the path tracer, host+device: pine's Halton and Sobol samplers,
// camera, layered BSDFs, lights and Path-integrator helpers, with pine's
// evaluation order.  The rules pine's -ffast-math build applies and that this
// file follows (each was seen bit for bit in pine's golden vectors):
//   - a vector divided by a scalar is multiplied by the scalar's reciprocal;
//   - a division by a constant is a multiplication by the folded reciprocal;
//   - other divisions are true IEEE divisions; no FMA contraction anywhere.
#pragma once

#include "pg_pstd.h"
#include "pg_sobol.h"

namespace pg {

static constexpr float kPi = 3.14159265358979323846f;
static constexpr float kInvPi = 1.0f / 3.14159265358979323846f;
static constexpr float kOneMinusEps = 0x1.fffffep-1f;
static constexpr float kEpsilon = 1.19209290e-07f;

// ---- small vector helpers -----------------------------------------------------
PG_HD float max0(float a, float b) { return a > b ? a : b; }
PG_HD f3 face_same_hemisphere(f3 v, f3 ref) { return dot(v, ref) < 0 ? -v : v; }  // vecmath.h:1252
PG_HD float abs_dot(f3 a, f3 b) { return fabsp(dot(a, b)); }

// 3x3 matrix stored by columns x, y, z (vecmath.h:543-640).
struct M3 {
    f3 x, y, z;
};
PG_HD f3 mul(const M3& m, f3 v) { return m.x * v.x + m.y * v.y + m.z * v.z; }  // vecmath.h:765

// CoordinateSystem(n) as a matrix (vecmath.h:1231-1236): z = n.
PG_HD M3 frame_from(f3 n) {
    M3 m;
    m.z = n;
    coordinate_system(n, m.x, m.y);
    return m;
}

// Inverse(mat3) (vecmath.h:1123-1145); m[c][r] = column c, row r.
PG_HD M3 inverse(const M3& m) {
    const float m00 = m.x.x, m01 = m.x.y, m02 = m.x.z;
    const float m10 = m.y.x, m11 = m.y.y, m12 = m.y.z;
    const float m20 = m.z.x, m21 = m.z.y, m22 = m.z.z;
    float det = m00 * (m11 * m22 - m21 * m12) + m10 * (m21 * m02 - m01 * m22) +
                m20 * (m01 * m12 - m11 * m02);
    M3 r;
    if (det == 0.0f) {
        r.x = mk(1, 0, 0);
        r.y = mk(0, 1, 0);
        r.z = mk(0, 0, 1);
        return r;
    }
    r.x = mk(m11 * m22 - m21 * m12, m21 * m02 - m01 * m22, m01 * m12 - m11 * m02);
    r.y = mk(m12 * m20 - m22 * m10, m22 * m00 - m02 * m20, m02 * m10 - m12 * m00);
    r.z = mk(m10 * m21 - m20 * m11, m20 * m01 - m00 * m21, m00 * m11 - m10 * m01);
    // mat3 / float: nine divisions by one value -> reciprocal multiply.
    const float inv = 1.0f / det;
    r.x = r.x * inv;
    r.y = r.y * inv;
    r.z = r.z * inv;
    return r;
}

// ---- Halton sampler (sampler.cpp:32-96, lowdiscrepancy.h:28-53) ----------------
struct HaltonDev {
    int baseScales[2], baseExponents[2], multInverse[2];
    int stride;
    const uint16_t* perms;  // ComputeRadicalInversePermutations
    const int* primes;      // first 1000 primes
    const int* primeSums;   // prefix sums of primes
    // Per dimension, three 16-byte records: prime base, offset of its
    // permutation in perms, 1/base and the series term perm[0] / (base + 1),
    // as floats' bits (the two quotients computed once on the host, IEEE as
    // the device would); then the base's division magic m and post-shift
    // l - 1 (see udiv_magic), the chunk length k (0: no chunk tables) and the
    // offset of the dimension's words in chunk32; then B = base^k, B's magic
    // and post-shift, and the offset of its tables in chunk16.
    const int4* dims;
    // Digit-chunk tables of the bases with base^2 <= kChunkEntries (see
    // scrambled_radical_inverse32): per dimension, in chunk16, the natural
    // table (B entries: the permuted digits of x < B folded as the digit loop
    // folds them, with the digit count in the top 4 bits) and the full table
    // (B entries: all k digits of x, leading zeros included); in chunk32,
    // base^n for n <= k, then the bits of invBase^n for n <= 32 as the loop
    // forms them (one rounded multiplication per digit).
    const uint16_t* chunk16;
    const uint32_t* chunk32;
};

constexpr int kChunkEntries = 4096, kMaxFullChunks = 3;

// Unsigned 32-bit division by an invariant b >= 2 (Granlund and Montgomery,
// "Division by invariant integers using multiplication", 1994): with
// l = ceil(log2 b) and m = floor(2^32 (2^l - b) / b) + 1,
// a / b = (t + ((a - t) >> 1)) >> (l - 1), t = the high word of m * a, for
// every 32-bit a.  A multiply-high, a subtract, an add and two shifts instead
// of the generic division sequence.
PG_HD uint32_t umulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}
PG_HD uint32_t udiv_magic(uint32_t a, uint32_t m, uint32_t post) {
    const uint32_t t = umulhi32(m, a);
    return (t + ((a - t) >> 1)) >> post;
}

// InverseRadicalInverse (sampler.cpp:32-41) for the pixel offsets of
// StartPixel, whose digits fit 32 bits: base 2 reverses the low nDigits bits,
// base 3 takes nDigits digits least significant first.
PG_HD uint32_t inverse_radical_inverse2(uint32_t inverse, int nDigits) {
    if (nDigits <= 0)
        return 0u;
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bitreverse32(inverse) >> (32 - nDigits);
#else
    uint32_t index = 0;
    for (int i = 0; i < nDigits; i++, inverse >>= 1)
        index = (index << 1) | (inverse & 1u);
    return index;
#endif
}
PG_HD uint32_t inverse_radical_inverse3(uint32_t inverse, int nDigits) {
    uint32_t index = 0;
    for (int i = 0; i < nDigits; i++) {
        const uint32_t next = inverse / 3u;
        index = index * 3u + (inverse - 3u * next);
        inverse = next;
    }
    return index;
}

// HaltonSampler::StartPixel (sampler.cpp:57-70): the Halton index of sample
// `sampleIndex` of pixel p.  The per-pixel offset is formed in 32 bits, which
// holds every intermediate of pine's 64-bit sum: the scales are at most 128
// (base 2) and 243 (base 3), each offset is below its own scale and each
// multiplicative inverse below the other scale, so each term is below
// 243 * 128 * 243 < 2^23 and the sum below 2^24.
PG_HD uint64_t halton_index(const HaltonDev& h, int px, int py, int sampleIndex) {
    uint32_t hi = 0;
    if (h.stride > 1) {
        const uint32_t o0 = inverse_radical_inverse2((uint32_t)(px % 128), h.baseExponents[0]);
        const uint32_t o1 = inverse_radical_inverse3((uint32_t)(py % 128), h.baseExponents[1]);
        hi = (o0 * (uint32_t)h.baseScales[1] * (uint32_t)h.multInverse[1] +
              o1 * (uint32_t)h.baseScales[0] * (uint32_t)h.multInverse[0]) %
             (uint32_t)h.stride;
    }
    return (uint64_t)((int64_t)hi + (int64_t)sampleIndex * h.stride);
}

// ScrambledRadicalInverse for an index below 2^32: the digit loop in 32 bits,
// four digits a round so that their permutation loads are in flight together
// (the digits are combined in the same order as one at a time).
//
// Bases with chunk tables (base^k = B <= kChunkEntries, k >= 2) take the
// index k digits at a time: each full chunk x = a mod B (the rest still
// nonzero, so all k digits count, zeros included) contributes
// reversed = reversed * B + full[x], and the last part (a < B, only its own
// digits) reversed * base^nd + natural[a]: the same integer as digit by
// digit, and invBase^(digits) from the table, the same float.  At most
// kMaxFullChunks full chunks for a 32-bit index (the host checks).
PG_HD float scrambled_radical_inverse32(const HaltonDev& h, int dim, uint32_t a) {
    const int4 t = h.dims[3 * dim], dm = h.dims[3 * dim + 1];
    if (dm.z > 0) {
        const int4 c = h.dims[3 * dim + 2];
        const uint32_t B = (uint32_t)c.x, k = (uint32_t)dm.z;
        const uint16_t* natural = h.chunk16 + c.w;
        const uint16_t* full = natural + B;
        const uint32_t* words = h.chunk32 + dm.w;
        uint32_t xs[kMaxFullChunks];
        int nf = 0;
#pragma unroll
        for (int j = 0; j < kMaxFullChunks; j++) {
            xs[j] = 0u;
            if (a >= B) {
                const uint32_t next = udiv_magic(a, (uint32_t)c.y, (uint32_t)c.z);
                xs[j] = a - next * B;
                a = next;
                nf++;
            }
        }
        uint32_t fv[kMaxFullChunks];
#pragma unroll
        for (int j = 0; j < kMaxFullChunks; j++)
            fv[j] = j < nf ? (uint32_t)full[xs[j]] : 0u;
        const uint32_t e = natural[a];
        uint64_t reversed = 0;
#pragma unroll
        for (int j = 0; j < kMaxFullChunks; j++)
            if (j < nf)
                reversed = reversed * (uint64_t)B + fv[j];
        const uint32_t nd = e >> 12;
        reversed = reversed * (uint64_t)words[nd] + (e & 0xFFFu);
        const float invBaseN = __builtin_bit_cast(float, words[k + 1 + (uint32_t)nf * k + nd]);
        const float v = ((float)reversed + __builtin_bit_cast(float, t.w)) * invBaseN;
        return v < kOneMinusEps ? v : kOneMinusEps;
    }
    const uint32_t b = (uint32_t)t.x, m = (uint32_t)dm.x, post = (uint32_t)dm.y;
    const uint16_t* perm = h.perms + t.y;
    const float invBase = __builtin_bit_cast(float, t.z);
    float invBaseN = 1.0f;
    uint64_t reversed = 0;
    while (a) {
        uint32_t dg[4];
        int nd = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t next = udiv_magic(a, m, post);
            dg[j] = a - next * b;
            nd += a != 0u ? 1 : 0;
            a = next;
        }
        uint32_t pv[4];
#pragma unroll
        for (int j = 0; j < 4; j++)
            pv[j] = j < nd ? (uint32_t)perm[dg[j]] : 0u;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (j < nd) {
                reversed = reversed * (uint64_t)b + pv[j];
                invBaseN *= invBase;
            }
    }
    const float series = __builtin_bit_cast(float, t.w);
    const float v = ((float)reversed + series) * invBaseN;
    return v < kOneMinusEps ? v : kOneMinusEps;
}

// ScrambledRadicalInverse.  Indices stay below 2^32 for every film up to
// 2^32/stride samples per pixel, so the digit loop runs in 32 bits.
PG_HD float scrambled_radical_inverse(const HaltonDev& h, int dim, uint64_t a64) {
    if (!(a64 >> 32))
        return scrambled_radical_inverse32(h, dim, (uint32_t)a64);
    const int base = h.primes[dim];
    const uint16_t* perm = h.perms + h.primeSums[dim];
    const float invBase = 1.0f / (float)base;
    float invBaseN = 1.0f;
    uint64_t reversed = 0;
    while (a64) {
        uint64_t next = a64 / (uint64_t)base;
        uint64_t digit = a64 - next * (uint64_t)base;
        reversed = reversed * (uint64_t)base + perm[digit];
        invBaseN *= invBase;
        a64 = next;
    }
    const float series = (float)perm[0] / ((float)base + 1.0f);
    const float v = ((float)reversed + series) * invBaseN;
    return v < kOneMinusEps ? v : kOneMinusEps;
}

// ---- the scene's sampler: Halton or Sobol (sampler.h:250-255) -------------------
enum SamplerKind : int32_t { kSamplerHalton = 0, kSamplerSobol = 1 };
struct SamplerDev {
    int32_t kind;
    HaltonDev halton;
    SobolDev sobol;
};

// Per-path sampler state: the sample's index (Halton index, or the 32-bit Sobol
// index) and the next dimension.
struct SamplerState {
    uint64_t index;
    int dim;
};

// StartPixel(p, sampleIndex): HaltonSampler::StartPixel (sampler.cpp:57-70) or
// SobolSampler::StartPixel (sampler.cpp:136-141); both restart at dimension 2.
PG_HD uint64_t sampler_start(const SamplerDev& h, int px, int py, int sampleIndex) {
    if (h.kind == kSamplerSobol)
        return sobol_interval_to_index(h.sobol, (uint32_t)sampleIndex, (uint32_t)px, (uint32_t)py);
    return halton_index(h.halton, px, py, sampleIndex);
}

PG_HD float get_1d(const SamplerDev& h, SamplerState& s) {
    if (h.kind == kSamplerSobol) {  // sampler.cpp:147-151
        if (s.dim >= kSobolDimensions)
            s.dim = 2;
        return sobol_sample(h.sobol, s.dim++, (uint32_t)s.index);
    }
    if (s.dim >= 1000)  // sampler.cpp:72-76
        s.dim = 2;
    return scrambled_radical_inverse(h.halton, s.dim++, s.index);
}
// get_1d compiled for one sampler (kernels instantiated per sampler keep only
// that sampler's code and registers): kS 0 = the scene's kind at run time,
// 1 = Halton with sample indices below 2^32, 2 = Sobol.  Same values as get_1d.
enum SamplerVariant : int { kSamplerAny = 0, kSamplerHalton32 = 1, kSamplerSobolOnly = 2 };
template <int kS>
PG_HD float get_1d_t(const SamplerDev& h, SamplerState& s) {
    if constexpr (kS == kSamplerSobolOnly) {
        if (s.dim >= kSobolDimensions)
            s.dim = 2;
        return sobol_sample(h.sobol, s.dim++, (uint32_t)s.index);
    } else if constexpr (kS == kSamplerHalton32) {
        if (s.dim >= 1000)
            s.dim = 2;
        return scrambled_radical_inverse32(h.halton, s.dim++, (uint32_t)s.index);
    } else {
        return get_1d(h, s);
    }
}

// Uniform numbers of this library's brick-majorant trackers (its own
// estimator; pine's global-majorant tracking keeps the sampler's values):
// each number takes the next sampler dimension with get_1d_t's bookkeeping
// (the same wrap points), so the dimensions a path uses after tracking are
// the ones it would use with the sampler's values, but the value is a hash of
// (sample index, dimension) instead of the sampler's radical inverse or Sobol
// point, which cost a digit loop or a matrix walk and dependent table reads
// per number inside the trackers' inner loops.  Distinct (index, dimension)
// pairs of a path give independent numbers, so the trackers stay unbiased.
// The hash: Wellons' lowbias32 mixer, applied to the dimension and the
// index's high word, then to that XOR the index's low word (bijective in the
// dimension for a fixed index); the top 24 bits give a float in [0, 1).
PG_HD uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
template <int kS>
PG_HD float track_uniform(const SamplerDev& h, SamplerState& s) {
    const bool sobol = kS == kSamplerSobolOnly || (kS == kSamplerAny && h.kind == kSamplerSobol);
    if (s.dim >= (sobol ? kSobolDimensions : 1000))
        s.dim = 2;
    const uint32_t d = (uint32_t)s.dim++;
    const uint32_t inner = lowbias32((uint32_t)(s.index >> 32) * 1024u + d);
    return (float)(lowbias32((uint32_t)s.index ^ inner) >> 8) * 0x1p-24f;
}

PG_HD void get_2d(const SamplerDev& h, SamplerState& s, float& u0, float& u1) {
    if (h.kind == kSamplerSobol) {
        // sampler.cpp:152-158 wraps only when the first dimension is past the
        // table, so a pair starting at 1023 reads one column block past pine's
        // table (undefined there); this reads dimension 2 for the second value.
        if (s.dim >= kSobolDimensions)
            s.dim = 2;
        const int d = s.dim;
        s.dim += 2;
        u0 = sobol_sample(h.sobol, d, (uint32_t)s.index);
        u1 = sobol_sample(h.sobol, d + 1 < kSobolDimensions ? d + 1 : 2, (uint32_t)s.index);
        return;
    }
    if (s.dim + 1 >= 1000)  // sampler.cpp:77-83
        s.dim = 2;
    int d = s.dim;
    s.dim += 2;
    u0 = scrambled_radical_inverse(h.halton, d, s.index);
    u1 = scrambled_radical_inverse(h.halton, d + 1, s.index);
}

// ---- sampling.h ----------------------------------------------------------------
PG_HD void sample_disk_polar(float u0, float u1, float& x, float& y) {  // sampling.h:16-20
    float r = pstd::sqrt(u0);
    float theta = 2.0f * kPi * u1;
    x = r * pstd::cos(theta);
    y = r * pstd::sin(theta);
}
PG_HD void sample_disk_concentric(float u0, float u1, float& x, float& y) {  // sampling.h:22-33
    u0 = u0 * 2.0f - 1.0f;
    u1 = u1 * 2.0f - 1.0f;
    float theta, r;
    if (fabsp(u0) > fabsp(u1)) {
        r = u0;
        theta = kPi / 4.0f * u1 / u0;
    } else {
        r = u1;
        theta = kPi / 2.0f - kPi / 4.0f * (u0 / u1);
    }
    x = r * pstd::cos(theta);
    y = r * pstd::sin(theta);
}
PG_HD f3 cosine_weighted(float u0, float u1) {  // sampling.h:35-39
    float dx, dy;
    sample_disk_concentric(u0, u1, dx, dy);
    float z = pstd::sqrt(max0(1.0f - dx * dx - dy * dy, 0.0f));
    return mk(dx, dy, z);
}
PG_HD float power_heuristic(float f, float g) {  // sampling.h:60-64 (nF = nG = 1)
    float ff = 1.0f * f, gg = 1.0f * g;
    return (ff * ff) / (ff * ff + gg * gg);
}

// ---- scattering.h ------------------------------------------------------------------
PG_HD float cos_theta(f3 w) { return w.z; }
PG_HD float cos2_theta(f3 w) { return w.z * w.z; }
PG_HD float abs_cos_theta(f3 w) { return fabsp(w.z); }
PG_HD float sin2_theta(f3 w) { return max0(1.0f - cos2_theta(w), 0.0f); }
PG_HD float sin_theta(f3 w) { return pstd::sqrt(sin2_theta(w)); }
PG_HD float tan2_theta(f3 w) { return sin2_theta(w) / cos2_theta(w); }
PG_HD float cos_phi(f3 w) {
    float s = sin_theta(w);
    return s == 0 ? 1.0f : pstd::clamp(w.x / s, -1.0f, 1.0f);
}
PG_HD float sin_phi(f3 w) {
    float s = sin_theta(w);
    return s == 0 ? 1.0f : pstd::clamp(w.y / s, -1.0f, 1.0f);
}
PG_HD bool same_hemisphere(f3 a, f3 b) { return a.z * b.z > 0.0f; }
PG_HD f3 face_forward(f3 v, f3 n) { return dot(v, n) < 0.0f ? -v : v; }
PG_HD f3 reflect(f3 wi, f3 n) { return 2.0f * dot(wi, n) * n - wi; }
PG_HD bool refract(f3 wi, f3 n, float eta, f3& wt) {  // scattering.h:72-91
    float cosThetaI = dot(n, wi);
    if (cosThetaI < 0) {
        eta = 1.0f / eta;
        cosThetaI = -cosThetaI;
        n = -n;
    }
    float sin2ThetaI = max0(0.0f, 1.0f - cosThetaI * cosThetaI);
    float sin2ThetaT = sin2ThetaI / (eta * eta);
    if (sin2ThetaT >= 1)
        return false;
    float cosThetaT = pstd::sqrt(1.0f - sin2ThetaT);
    wt = -wi / eta + (cosThetaI / eta - cosThetaT) * n;
    return true;
}
PG_HD float fr_dielectric(float cosThetaI, float eta) {  // scattering.h:93-109
    cosThetaI = pstd::clamp(cosThetaI, -1.0f, 1.0f);
    if (cosThetaI < 0) {
        eta = 1 / eta;
        cosThetaI = -cosThetaI;
    }
    float sin2ThetaI = 1.0f - cosThetaI * cosThetaI;
    float sin2ThetaT = sin2ThetaI / (eta * eta);
    if (sin2ThetaT >= 1.0f)
        return 1.0f;
    float cosThetaT = pstd::sqrt(1.0f - sin2ThetaT);
    float rParl = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
    float rPerp = (cosThetaI - eta * cosThetaT) / (cosThetaI + eta * cosThetaT);
    return (rParl * rParl + rPerp * rPerp) * 0.5f;
}
PG_HD f3 fr_schlick(f3 F0, float cosTheta) {  // scattering.h:111-113
    return F0 + (mk(1, 1, 1) - F0) * pstd::pow(1.0f - cosTheta, 5.0f);
}

// TrowbridgeReitzDistribution (scattering.h:115-170), isotropic alpha.
struct GGX {
    float ax, ay;
    PG_HD float D(f3 wm) const {
        float tan2 = tan2_theta(wm);
        float cos4 = cos2_theta(wm) * cos2_theta(wm);
        if (cos4 < 1e-6f)
            return 0.0f;
        float cp = cos_phi(wm) / ax, sp = sin_phi(wm) / ay;
        float e = tan2 * (cp * cp + sp * sp);
        float ope = 1 + e;
        return 1.0f / (kPi * ax * ay * cos4 * (ope * ope));
    }
    PG_HD float Lambda(f3 w) const {
        float tan2 = tan2_theta(w);
        float a = cos_phi(w) * ax, b = sin_phi(w) * ay;
        float alpha2 = a * a + b * b;
        return (pstd::sqrt(1.0f + alpha2 * tan2) - 1.0f) * 0.5f;
    }
    PG_HD float G1(f3 w) const { return 1.0f / (1.0f + Lambda(w)); }
    PG_HD float G(f3 wi, f3 wo) const { return 1.0f / (1.0f + Lambda(wi) + Lambda(wo)); }
    PG_HD float Dv(f3 w, f3 wm) const { return G1(w) / abs_cos_theta(w) * D(wm) * abs_dot(w, wm); }
    PG_HD float PDF(f3 w, f3 wm) const { return max0(Dv(w, wm), kEpsilon); }
    PG_HD f3 SampleWm(f3 w, float u0, float u1) const {
        f3 wh = normalize(mk(ax * w.x, ay * w.y, w.z));
        if (wh.z < 0.0f)
            wh = -wh;
        f3 T1 = (wh.z < 0.99999f) ? normalize(cross(mk(0, 0, 1), wh)) : mk(1, 0, 0);
        f3 T2 = cross(wh, T1);
        float px, py;
        sample_disk_polar(u0, u1, px, py);
        float h = pstd::sqrt(1.0f - px * px);
        float t = (1.0f + wh.z) * 0.5f;
        py = h * (1.0f - t) + py * t;  // pstd::lerp(t, h, p.y)
        float pz = pstd::sqrt(max0(0.0f, 1.0f - (px * px + py * py)));
        f3 nh = px * T1 + py * T2 + pz * wh;
        return normalize(mk(ax * nh.x, ay * nh.y, max0(1e-6f, nh.z)));
    }
};

PG_HD float ggx_alpha(float roughness) {
    return pstd::clamp(roughness * roughness, 0.001f, 1.0f);
}

// ---- layered materials (material.cpp, bxdf.cpp) --------------------------------
enum LayerType : int32_t { kDiffuse = 0, kConductor = 1, kDielectric = 2 };

struct Layer {
    int32_t type;
    f3 albedo;
    float roughness;
    float eta;
    // Node-graph programs for the inputs (offsets into the scene's program
    // table, -1: the constant above).
    int32_t prog_albedo, prog_roughness, prog_eta;
};

enum MaterialKind : int32_t { kMatNone = -1, kMatLayered = 0, kMatEmissive = 1 };

constexpr int kMaxLayers = 4;
struct MaterialDev {
    int32_t kind;
    int32_t nlayers;
    Layer layers[kMaxLayers];  // evaluation order: descending layer index
    f3 emission;
    int32_t has_nodes;  // some input is a node-graph program
    int32_t prog_bump;  // bumpMap program, -1 = none
};

struct BSDFSample {
    f3 wo;
    f3 f;
    float pdf;
};

// DiffuseBSDF::Sample/F/PDF (bxdf.cpp:7-31)
PG_HD bool diffuse_sample(const Layer& l, f3 wi, float u0, float u1, BSDFSample& bs) {
    f3 wo = cosine_weighted(u0, u1);
    if (cos_theta(wi) < 0)
        wo = wo * -1.0f;
    bs.wo = wo;
    bs.pdf = abs_cos_theta(wo) * kInvPi;
    bs.f = l.albedo * kInvPi;
    return true;
}
PG_HD f3 diffuse_f(const Layer& l, f3 wi, f3 wo) {
    if (!same_hemisphere(wi, wo))
        return mk(0, 0, 0);
    return l.albedo * kInvPi;
}
PG_HD float diffuse_pdf(f3 wi, f3 wo) {
    if (!same_hemisphere(wi, wo))
        return kEpsilon;
    return abs_cos_theta(wo) * kInvPi;
}

// ConductorBSDF (bxdf.cpp:33-78)
PG_HD bool conductor_sample(const Layer& l, f3 wi, float u0, float u1, BSDFSample& bs) {
    float alpha = ggx_alpha(l.roughness);
    GGX d{alpha, alpha};
    f3 wm = d.SampleWm(wi, u0, u1);
    f3 wo = reflect(wi, wm);
    if (!same_hemisphere(wi, wo))
        return false;
    f3 fr = fr_schlick(l.albedo, abs_cos_theta(wm));
    bs.wo = wo;
    bs.pdf = d.PDF(wi, wm) / (4.0f * abs_dot(wi, wm));
    bs.f = fr * d.D(wm) * d.G(wo, wi) / (4.0f * cos_theta(wi) * cos_theta(wo));
    return true;
}
PG_HD f3 conductor_f(const Layer& l, f3 wi, f3 wo) {
    if (!same_hemisphere(wi, wo))
        return mk(0, 0, 0);
    float alpha = ggx_alpha(l.roughness);
    GGX d{alpha, alpha};
    f3 wh = normalize(wi + wo);
    f3 fr = fr_schlick(l.albedo, abs_cos_theta(wh));
    return fr * d.D(wh) * d.G(wo, wi) / (4.0f * abs_cos_theta(wo) * abs_cos_theta(wi));
}
PG_HD float conductor_pdf(const Layer& l, f3 wi, f3 wo) {
    if (!same_hemisphere(wi, wo))
        return 0.0f;
    float alpha = ggx_alpha(l.roughness);
    GGX d{alpha, alpha};
    f3 wh = normalize(wi + wo);
    return d.PDF(wi, wh) / (4.0f * abs_dot(wi, wh));
}

// DielectricBSDF (bxdf.cpp:80-165)
PG_HD bool dielectric_sample(const Layer& l, f3 wi, float u1, float u2a, float u2b,
                             BSDFSample& bs) {
    float etap = l.eta;
    if (cos_theta(wi) < 0)
        etap = 1.0f / etap;
    float fr = fr_dielectric(abs_cos_theta(wi), etap);
    float alpha = ggx_alpha(l.roughness);
    GGX d{alpha, alpha};
    f3 wm = d.SampleWm(wi, u2a, u2b);
    if (u1 < fr) {
        f3 wo = reflect(wi, wm);
        if (!same_hemisphere(wi, wo))
            return false;
        bs.wo = wo;
        bs.pdf = fr * d.PDF(wi, wm) / (4.0f * abs_dot(wi, wm));
        bs.f = l.albedo * fr * d.D(wm) * d.G(wo, wi) / (4.0f * cos_theta(wi) * cos_theta(wo));
    } else {
        f3 wo;
        if (!refract(wi, wm, etap, wo))
            return false;
        float cosThetaO = cos_theta(wo), cosThetaI = cos_theta(wi);
        bs.wo = wo;
        float dd = dot(wo, wm) + dot(wi, wm) / etap;
        float denom = dd * dd;
        float dwm_dwo = abs_dot(wo, wm) / denom;
        bs.pdf = (1.0f - fr) * d.PDF(wi, wm) * dwm_dwo;
        bs.f = l.albedo * (1.0f - fr) * d.D(wm) * d.G(wi, wo) *
               fabsp(dot(wo, wm) * dot(wi, wm) / denom / cosThetaI / cosThetaO);
    }
    return true;
}
PG_HD f3 dielectric_f(const Layer& l, f3 wi, f3 wo) {
    float alpha = ggx_alpha(l.roughness);
    GGX d{alpha, alpha};
    float cosThetaO = cos_theta(wo), cosThetaI = cos_theta(wi);
    bool refl = cosThetaI * cosThetaO > 0;
    float etap = l.eta;
    if (!refl)
        etap = 1.0f / etap;
    f3 wm = face_forward(normalize(wi * etap + wo), mk(0, 0, 1));
    if (dot(wm, wo) * cosThetaI < 0.0f || dot(wm, wi) * cosThetaO < 0.0f)
        return mk(0, 0, 0);
    float fr = fr_dielectric(abs_cos_theta(wi), l.eta);
    if (refl)
        return l.albedo * fr * d.D(wm) * d.G(wo, wi) / (4.0f * cosThetaI * cosThetaO);
    float dd = dot(wo, wm) + dot(wi, wm) / etap;
    float denom = dd * dd * cosThetaI * cosThetaO;
    return l.albedo * (1.0f - fr) * d.D(wm) * d.G(wi, wo) *
           fabsp(dot(wo, wm) * dot(wi, wm) / denom);
}
PG_HD float dielectric_pdf(const Layer& l, f3 wi, f3 wo) {
    float alpha = ggx_alpha(l.roughness);
    GGX d{alpha, alpha};
    float cosThetaO = cos_theta(wo), cosThetaI = cos_theta(wi);
    bool refl = cosThetaI * cosThetaO > 0;
    float etap = l.eta;
    if (!refl)
        etap = 1.0f / etap;
    f3 wm = face_forward(normalize(wi * etap + wo), mk(0, 0, 1));
    if (dot(wm, wo) * cosThetaI < 0.0f || dot(wm, wi) * cosThetaO < 0.0f)
        return 0.0f;
    float fr = fr_dielectric(abs_cos_theta(wi), l.eta);
    if (refl)
        return fr * d.PDF(wi, wm) / (4.0f * abs_dot(wi, wm));
    float dd = dot(wo, wm) + dot(wi, wm) / etap;
    float denom = dd * dd;
    float dwm_dwo = abs_dot(wo, wm) / denom;
    return (1.0f - fr) * d.PDF(wi, wm) * dwm_dwo;
}

// LayeredMaterial::Sample (material.cpp:9-18) + Material::Sample (material.h:76-87):
// layers in order, the first whose sample stays in wi's hemisphere wins,
// otherwise the last sample made is returned.  Returns false for "no sample".
PG_HD bool material_sample(const MaterialDev& m, f3 wi, float u1, float u2a, float u2b,
                           BSDFSample& out) {
    if (m.kind != kMatLayered)
        return false;
    bool have = false;
    for (int i = 0; i < m.nlayers; i++) {
        const Layer& l = m.layers[i];
        BSDFSample bs;
        bool ok;
        if (l.type == kDiffuse)
            ok = diffuse_sample(l, wi, u2a, u2b, bs);
        else if (l.type == kConductor)
            ok = conductor_sample(l, wi, u2a, u2b, bs);
        else
            ok = dielectric_sample(l, wi, u1, u2a, u2b, bs);
        have = ok;
        if (ok)
            out = bs;
        if (ok && same_hemisphere(wi, bs.wo))
            break;
    }
    return have;
}
PG_HD f3 material_f(const MaterialDev& m, f3 wi, f3 wo) {
    f3 f = mk(0, 0, 0);
    if (m.kind != kMatLayered)
        return f;
    for (int i = 0; i < m.nlayers; i++) {
        const Layer& l = m.layers[i];
        if (l.type == kDiffuse)
            f = f + diffuse_f(l, wi, wo);
        else if (l.type == kConductor)
            f = f + conductor_f(l, wi, wo);
        else
            f = f + dielectric_f(l, wi, wo);
    }
    return f;
}
PG_HD float material_pdf(const MaterialDev& m, f3 wi, f3 wo) {
    float pdf = 0.0f;
    if (m.kind != kMatLayered)
        return pdf;
    for (int i = 0; i < m.nlayers; i++) {
        const Layer& l = m.layers[i];
        if (l.type == kDiffuse)
            pdf += diffuse_pdf(wi, wo);
        else if (l.type == kConductor)
            pdf += conductor_pdf(l, wi, wo);
        else
            pdf += dielectric_pdf(l, wi, wo);
    }
    return pdf;
}

// RGBSpectrum::y (spectrum.h:271-274)
PG_HD float luminance(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }
PG_HD bool is_black(f3 c) { return c.x == 0.0f && c.y == 0.0f && c.z == 0.0f; }

}  // namespace pg
