// Pine's Sobol sampler, host+device (sampler.cpp:129-220, lowdiscrepancy.h:128-158).
//
// A sample is the Sobol point of a per-pixel index: SobolIntervalToIndex picks,
// for sample `frame` of pixel p, the index whose first two Sobol dimensions fall
// in p's cell of a 2^m x 2^m grid (m = log2 of the film's larger side rounded up
// to a power of two).  Dimension d of that point is the XOR of the generator
// matrix columns selected by the index's bits, scrambled per dimension with a
// seed Hash(d) and mapped to [0, 1).
//
// Layout: the generator matrices of all 1024 dimensions live in device memory
// dimension-major, 32 columns of 32 bits each (128 KiB: pine's index is a
// 32-bit value, so columns 32..51 of its table are never reached).  The pixel
// -> index tables for the scene's m are 2 x 52 64-bit words.
#pragma once
#include <cstdint>

#include "pg_vec.h"

namespace pg {

static constexpr int kSobolDimensions = 1024;  // NSobolDimensions (sobolmetrices.h:10)
static constexpr int kSobolColumns = 32;

// SobolSampler::RandomizeStrategy (sampler.h:143).
enum SobolRandomize : int32_t { kSobolNone = 0, kSobolBinaryPermutate = 1, kSobolFastOwen = 2, kSobolOwen = 3 };

struct SobolDev {
    const uint32_t* C;        // [1024][32] generator matrix columns
    const uint64_t* vdc;      // [52] VdCSobolMatrices[m - 1]: frame bit -> pixel-code delta
    const uint64_t* vdc_inv;  // [52] VdCSobolMatricesInv[m - 1]: pixel-code bit -> index bits
    int32_t log2Scale;        // m
    int32_t randomize;        // SobolRandomize
};

// Hash64u (rng.h:10-15) of one 32-bit value zero-extended: Hash(int) (rng.h:17-29).
PG_HD uint32_t sobol_dim_seed(int dim) {
    uint64_t s = (uint64_t)(uint32_t)dim + 0x9E3779B97f4A7C15ULL;
    s = (s ^ (s >> 30)) * 0xBF58476D1CE4E5B9ULL;
    s = (s ^ (s >> 27)) * 0x94D049BB133111EBULL;
    return (uint32_t)(s ^ (s >> 31));
}

PG_HD uint32_t reverse_bits32(uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_bitreverse32(x);
#else
    x = (x & 0x55555555u) << 1 | (x & 0xaaaaaaaau) >> 1;
    x = (x & 0x33333333u) << 2 | (x & 0xccccccccu) >> 2;
    x = (x & 0x0f0f0f0fu) << 4 | (x & 0xf0f0f0f0u) >> 4;
    x = (x & 0x00ff00ffu) << 8 | (x & 0xff00ff00u) >> 8;
    return (x << 16) | (x >> 16);
#endif
}

// MixBits (rng.h:43-50).
PG_HD uint64_t mix_bits(uint64_t v) {
    v ^= (v >> 31);
    v *= 0x7fb5d329728ea185ULL;
    v ^= (v >> 27);
    v *= 0x81dadef4bc2dd44dULL;
    v ^= (v >> 33);
    return v;
}

// The four scramblers of sampler.cpp:159-208.
PG_HD uint32_t sobol_scramble(uint32_t v, uint32_t seed, int strategy) {
    if (strategy == kSobolBinaryPermutate)
        return v ^ seed;
    if (strategy == kSobolFastOwen) {
        v = reverse_bits32(v);
        v ^= v * 0x3d20adeau;
        v += seed;
        v *= (seed >> 16) | 1u;
        v ^= v * 0x05526c56u;
        v ^= v * 0x53a22864u;
        return reverse_bits32(v);
    }
    if (strategy == kSobolOwen) {
        if (seed & 1u)
            v ^= 1u << 31;
        for (int b = 1; b < 32; b++) {
            const uint32_t mask = (~0u) << (32 - b);
            if ((uint32_t)mix_bits((uint64_t)((v & mask) ^ seed)) & (1u << b))
                v ^= 1u << (31 - b);
        }
        return v;
    }
    return v;
}

// SobolSample (lowdiscrepancy.h:128-136) through SampleDimension (sampler.cpp:210-220).
PG_HD float sobol_sample(const SobolDev& s, int dim, uint32_t index) {
    const uint32_t* c = s.C + (int64_t)dim * kSobolColumns;
    uint32_t v = 0;
    while (index) {
        const int b = __builtin_ctz(index);
        v ^= c[b];
        index &= index - 1u;
    }
    v = sobol_scramble(v, sobol_dim_seed(dim), s.randomize);
    const float r = (float)v * 0x1p-32f;
    return r < 0x1.fffffep-1f ? r : 0x1.fffffep-1f;  // min(v * 0x1p-32f, OneMinusEpsilon)
}

// SobolIntervalToIndex (lowdiscrepancy.h:139-158).  Pine keeps the result in a
// 32-bit variable, so the returned index always fits in 32 bits.
PG_HD uint32_t sobol_interval_to_index(const SobolDev& s, uint32_t frame, uint32_t px, uint32_t py) {
    const uint32_t m = (uint32_t)s.log2Scale;
    if (m == 0)
        return frame;
    uint32_t index = (uint32_t)((uint64_t)frame << (2 * m));
    uint64_t delta = 0;
    for (int c = 0; frame; c++, frame >>= 1)
        if (frame & 1u)
            delta ^= s.vdc[c];
    uint64_t b = (((uint64_t)px << m) | (uint64_t)py) ^ delta;
    for (int c = 0; b; c++, b >>= 1)
        if (b & 1u)
            index ^= (uint32_t)s.vdc_inv[c];
    return index;
}

}  // namespace pg
