// Spectral transport, four wavelengths per path (hero-wavelength sampling):
// an extension beside pine's RGB pipeline (pine's RGBSpectrum,
// src/core/spectrum.h:230-276, has no spectral counterpart).
//
//   wavelengths  l_j = 360 + 470 frac(u + j/4) nm, u uniform per path, so each
//                l_j is uniform on [360, 830] (pdf 1/470);
//   RGB -> spectrum (reflectances, emission, sky): s(l) = b(l)^T Q rgb, with
//                b(l) a smooth partition of unity (blue, green, red bands with
//                logistic edges at 490 and 590 nm) and Q chosen so that the
//                projection of the spectrum of rgb is rgb again (primaries
//                round-trip exactly, gray maps to a flat spectrum);
//   spectrum -> RGB (film): rgb = sum_j W(l_j) L_j, W(l) = 470 / (4 Ybar) M cmf(l),
//                the Monte Carlo estimate of M * XYZ / Ybar over the four
//                wavelengths, M = linear sRGB matrix white-balanced so that the
//                equal-energy spectrum maps to (1, 1, 1);
//   cmf          CIE 1931 2-degree colour matching functions, the multi-lobe
//                piecewise-Gaussian fit of Wyman, Sloan and Shirley (JCGT 2(2),
//                2013).
// A scene of flat (gray) spectra is therefore rendered exactly as the RGB
// pipeline renders it, in expectation.
#pragma once

#include "pg_vec.h"

#include <cmath>

namespace pg {
namespace spec {

constexpr float kLamMin = 360.0f, kLamRange = 470.0f;

PG_HD inline float lobe(float x, float mu, float s1, float s2) {
    const float t = (x - mu) / (x < mu ? s1 : s2);
    return expf(-0.5f * t * t);
}

PG_HD inline f3 cmf(float l) {
    return mk(1.056f * lobe(l, 599.8f, 37.9f, 31.0f) + 0.362f * lobe(l, 442.0f, 16.0f, 26.7f) -
                  0.065f * lobe(l, 501.1f, 20.4f, 26.2f),
              0.821f * lobe(l, 568.8f, 46.9f, 40.5f) + 0.286f * lobe(l, 530.9f, 16.3f, 31.1f),
              1.217f * lobe(l, 437.0f, 11.8f, 36.0f) + 0.681f * lobe(l, 459.0f, 26.0f, 13.8f));
}

// (red, green, blue) band weights at wavelength l; they sum to 1.
PG_HD inline f3 basis(float l) {
    const float a = 1.0f / (1.0f + expf(-(l - 490.0f) / 12.0f));
    const float c = 1.0f / (1.0f + expf(-(l - 590.0f) / 12.0f));
    return mk(c, a - c, 1.0f - a);
}

// Tables the kernels use (computed on the host by spectral_tables).
struct Tables {
    float Q[9];      // RGB -> basis coefficients, row-major
    float M[9];      // XYZ -> RGB (white-balanced linear sRGB), row-major
    float w_scale;   // 470 / (4 * integral of ybar)
};

PG_HD inline f3 mul33(const float* A, f3 v) {
    return mk(A[0] * v.x + A[1] * v.y + A[2] * v.z, A[3] * v.x + A[4] * v.y + A[5] * v.z,
              A[6] * v.x + A[7] * v.y + A[8] * v.z);
}

// a(l) = Q^T b(l): the spectrum of rgb at l is max(0, a(l) . rgb).
PG_HD inline f3 upsample_row(const Tables& t, float l) {
    const f3 b = basis(l);
    const float* Q = t.Q;
    return mk(Q[0] * b.x + Q[3] * b.y + Q[6] * b.z, Q[1] * b.x + Q[4] * b.y + Q[7] * b.z,
              Q[2] * b.x + Q[5] * b.y + Q[8] * b.z);
}

// W(l): the RGB weight of a radiance sample at l.
PG_HD inline f3 film_weight(const Tables& t, float l) { return mul33(t.M, cmf(l)) * t.w_scale; }

PG_HD inline float wavelength(float u, int j) {
    float v = u + 0.25f * (float)j;
    v = v >= 1.0f ? v - 1.0f : v;
    return kLamMin + kLamRange * v;
}

// Host: integrals over [360, 830] nm at 1 nm (midpoint rule) and the tables.
inline void spectral_tables(Tables& t) {
    double Y = 0.0, flat[3] = {0, 0, 0}, proj[3][3] = {{0}};
    for (int k = 0; k < 470; k++) {
        const float l = kLamMin + (float)k + 0.5f;
        const f3 c = cmf(l), b = basis(l);
        const double cc[3] = {c.x, c.y, c.z}, bb[3] = {b.x, b.y, b.z};
        Y += cc[1];
        for (int i = 0; i < 3; i++) {
            flat[i] += cc[i];
            for (int j = 0; j < 3; j++)
                proj[i][j] += cc[i] * bb[j];  // XYZ of band j, row i
        }
    }
    const double S[9] = {3.2404542, -1.5371385, -0.4985314, -0.9692660, 1.8760108,
                         0.0415560, 0.0556434, -0.2040259, 1.0572252};  // XYZ -> linear sRGB
    double white[3], M[9];
    for (int i = 0; i < 3; i++)
        white[i] = (S[3 * i] * flat[0] + S[3 * i + 1] * flat[1] + S[3 * i + 2] * flat[2]) / Y;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            M[3 * i + j] = S[3 * i + j] / white[i];
    // P[i][j] = RGB channel i of band j's spectrum.
    double P[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            P[3 * i + j] = (M[3 * i] * proj[0][j] + M[3 * i + 1] * proj[1][j] + M[3 * i + 2] * proj[2][j]) / Y;
    const double det = P[0] * (P[4] * P[8] - P[5] * P[7]) - P[1] * (P[3] * P[8] - P[5] * P[6]) +
                       P[2] * (P[3] * P[7] - P[4] * P[6]);
    const double inv[9] = {(P[4] * P[8] - P[5] * P[7]) / det, (P[2] * P[7] - P[1] * P[8]) / det,
                           (P[1] * P[5] - P[2] * P[4]) / det, (P[5] * P[6] - P[3] * P[8]) / det,
                           (P[0] * P[8] - P[2] * P[6]) / det, (P[2] * P[3] - P[0] * P[5]) / det,
                           (P[3] * P[7] - P[4] * P[6]) / det, (P[1] * P[6] - P[0] * P[7]) / det,
                           (P[0] * P[4] - P[1] * P[3]) / det};
    for (int k = 0; k < 9; k++) {
        t.Q[k] = (float)inv[k];
        t.M[k] = (float)M[k];
    }
    t.w_scale = (float)(kLamRange / (4.0 * Y));
}

}  // namespace spec
}  // namespace pg
