// Host check of the library's StartPixel (halton_index in pg_shade.h), built
// by hipcc as host code only: its 32-bit per-pixel offset against pine's 64-bit
// sum (HaltonSampler::StartPixel, sampler.cpp:57-70, with InverseRadicalInverse,
// sampler.cpp:32-41, and the scales and inverses of the constructor,
// sampler.cpp:9-30,43-55), restated here in 64-bit arithmetic, for every
// pixel of films from 1 to 1000 pixels a side (first 300 rows and columns)
// and sample indices up to 2e9.  Prints the count checked and exits nonzero
// on a mismatch.
#include <algorithm>
#include <cstdio>

#include "pg_shade.h"

using namespace pg;

static uint64_t pine_inverse_radical_inverse(uint64_t inverse, int base, int nDigits) {
    uint64_t index = 0;
    for (int i = 0; i < nDigits; i++) {
        const uint64_t digit = inverse % (uint64_t)base;
        inverse /= (uint64_t)base;
        index = index * (uint64_t)base + digit;
    }
    return index;
}

static uint64_t pine_start_pixel(const HaltonDev& h, int px, int py, int sampleIndex) {
    int64_t hi = 0;
    if (h.stride > 1) {
        const int pm[2] = {px % 128, py % 128};
        for (int i = 0; i < 2; i++) {
            const uint64_t off = pine_inverse_radical_inverse((uint64_t)pm[i], i == 0 ? 2 : 3, h.baseExponents[i]);
            hi += (int64_t)(off * (uint64_t)h.baseScales[1 - i] * (uint64_t)h.multInverse[1 - i]);
        }
        hi %= h.stride;
    }
    return (uint64_t)(hi + (int64_t)sampleIndex * h.stride);
}

static int64_t mul_inverse(int64_t a, int64_t n) {
    int64_t r0 = n, r1 = a % n, t0 = 0, t1 = 1;
    while (r1) {
        const int64_t q = r0 / r1;
        const int64_t r2 = r0 - q * r1, t2 = t0 - q * t1;
        r0 = r1, r1 = r2, t0 = t1, t1 = t2;
    }
    return ((t0 % n) + n) % n;
}

int main() {
    const int sizes[] = {1, 2, 3, 5, 8, 9, 27, 64, 81, 100, 127, 128, 129, 200, 243, 256, 512, 1000};
    const int samples[] = {0, 1, 7, 63, 4095, 100000, 2000000000};
    long checked = 0, bad = 0;
    for (int w : sizes)
        for (int ht : sizes) {
            HaltonDev h{};
            const int film[2] = {w, ht};
            for (int i = 0; i < 2; i++) {
                int base = i == 0 ? 2 : 3, scale = 1, exp = 0;
                while (scale < std::min(film[i], 128)) {
                    scale *= base;
                    ++exp;
                }
                h.baseScales[i] = scale;
                h.baseExponents[i] = exp;
            }
            h.multInverse[0] = (int)mul_inverse(h.baseScales[0], h.baseScales[1]);
            h.multInverse[1] = (int)mul_inverse(h.baseScales[1], h.baseScales[0]);
            h.stride = h.baseScales[0] * h.baseScales[1];
            for (int py = 0; py < std::min(ht, 300); py++)
                for (int px = 0; px < std::min(w, 300); px++)
                    for (int si : samples) {
                        checked++;
                        if (pine_start_pixel(h, px, py, si) != halton_index(h, px, py, si) && bad++ < 5)
                            std::printf("mismatch: film %dx%d pixel (%d, %d) sample %d\n", w, ht, px, py, si);
                    }
        }
    std::printf("%ld checked, %ld mismatches\n", checked, bad);
    return bad != 0;
}
