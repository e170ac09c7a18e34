// pine's own libm (src/pstd/math.h:110-351), restated for host and device.
// Pine replaces every transcendental with short iterations (Newton square root,
// Taylor exp and cos, bisection log2, Newton acos/atan); matching pine's images
// means computing the same approximations, not the correctly rounded functions.
// Pine is compiled with -ffast-math, which lets g++ turn division by a constant
// into multiplication by its reciprocal and fold x*x inside unrolled loops; the
// forms below are the ones that reproduce the reference build's bits on the
// golden vectors (each function names its source lines).
#pragma once

#include "pg_vec.h"

namespace pg {
namespace pstd {

static constexpr float Pi = 3.14159265358979323846f;
static constexpr float Ln2 = 0.69314718055994530942f;
static constexpr float OneMinusEpsilon = 0x1.fffffep-1f;
static constexpr float Epsilon = 1.19209290e-07f;

// math.h:50-56
PG_HD int ieeeexp(float x) { return (int)(0xff & (f2u(x) >> 23)) - 127; }
PG_HD float exp2i(int e) { return u2f((uint32_t)(e + 127) << 23); }

// math.h:111-117 (cast-based floor)
PG_HD float floor(float v) {
    int32_t i = (int32_t)v;
    if (v < 0 && (float)i != v)
        i -= 1;
    return (float)i;
}
PG_HD float ceil(float v) {
    int32_t i = (int32_t)v;
    if (v > 0 && (float)i != v)
        i += 1;
    return (float)i;
}

// math.h:128-133 (cast-based fract; v - trunc(v) moved into [0, 1))
PG_HD float fract(float v) {
    v = v - (float)(int32_t)v;
    v = v + 1.0f;
    return v - (float)(int32_t)v;
}

// math.h:141-151: four Newton steps from 2^(e/2).
PG_HD float sqrt(float y) {
    float x = exp2i(ieeeexp(y) / 2);
    x = x * 0.5f + y / (2.0f * x);
    x = x * 0.5f + y / (2.0f * x);
    x = x * 0.5f + y / (2.0f * x);
    x = x * 0.5f + y / (2.0f * x);
    return x;
}

// math.h:153-165
PG_HD float powi(float x, int e) {
    float y = 1.0f;
    while (e) {
        if (e & 1)
            y *= x;
        x *= x;
        e >>= 1;
    }
    return y;
}

// math.h:167-185
PG_HD float pow(float x, float e) {
    int ei = (int)pstd::floor(e);
    e -= (float)ei;
    float y = ei > 0 ? powi(x, ei) : 1.0f / powi(x, -ei);
    for (int i = 0; i < 8; ++i) {
        if (e > 1.0f) {
            y *= x;
            e -= 1.0f;
        }
        e *= 2.0f;
        x = pstd::sqrt(x);
    }
    return y;
}

// math.h:187-212: eight bisection steps on the mantissa.
PG_HD float log2(float y) {
    int lg = ieeeexp(y);
    float logl = 0.0f, logr = 1.0f;
    float expl = exp2i(0), expr = exp2i(1);
    y = y / exp2i(lg);
    for (int i = 0; i < 8; ++i) {
        float logm = (logl + logr) * 0.5f;
        float expm = pstd::sqrt(expl * expr);
        if (expm < y) {
            logl = logm;
            expl = expm;
        } else {
            logr = logm;
            expr = expm;
        }
    }
    return (float)lg + (logl + logr) * 0.5f;
}
// math.h:214-217
PG_HD float log(float y) { return Ln2 * pstd::log2(y); }

// math.h:229-239: 8-term Horner Taylor series, no range reduction.  Pine's
// build evaluates x*y/i as y*(x*(1/i)) with 1/i folded to a float constant.
PG_HD float exp(float x) {
    float y = 0.0f;
#pragma unroll
    for (int i = 8; i > 0; --i)
        y = 1.0f + y * (x * (1.0f / (float)i));
    return y;
}

// math.h:251-260, evaluated by pine's build as y*((x*x)*(1/(i*(i-1)))).
PG_HD float cos_0_pi_2(float x) {
    float y = 1.0f;
    const float x2 = x * x;
#pragma unroll
    for (int i = 8; i > 0; i -= 2)
        y = 1.0f - y * (x2 * (1.0f / (float)(i * (i - 1))));
    return y;
}
// math.h:262-276.  Pine's build divides by 2*Pi as a multiply by the folded
// reciprocal and, in the x > Pi branch, evaluates 2*Pi - (x - y*2*Pi) as
// (2*Pi - x) + y*2*Pi.
PG_HD float cos(float x) {
    if (x < 0)
        x = -x;
    uint32_t y = (uint32_t)(int64_t)(x * (1.0f / (Pi * 2)));
    float yp = (float)y * (Pi * 2);
    float xr = x - yp;
    if (xr > Pi)
        xr = (Pi * 2 - x) + yp;
    if (xr < Pi / 2)
        return cos_0_pi_2(xr);
    else
        return -cos_0_pi_2(Pi - xr);
}
// math.h:278-281
PG_HD float sin(float x) { return pstd::cos(x - Pi / 2); }
// math.h:283-287
PG_HD float tan(float x) { return pstd::sin(x) / pstd::cos(x); }

// math.h:289-306
PG_HD float acos(float y) {
    if (y < -1 || y > 1)
        return 0.0f;
    float x = (1 - y) * Pi / 2;
    for (int i = 0; i < 8; ++i) {
        float y_ = pstd::cos(x) - y;
        float dydx = -pstd::sin(x);
        x -= y_ / dydx;
    }
    return x;
}

// math.h:327-341.  Pine's build folds y_ / (1 / cos^2) into y_ * cos^2.
// atan_steps runs the first seven Newton steps and returns the last correction
// in `d` (atan = x - d); atan2 needs the two parts.
PG_HD float atan_steps(float y, float& d) {
    float x = (y > 0 ? 1 - 1 / (1 + y) : -1 + 1 / (1 - y)) * Pi / 2;
    for (int i = 0; i < 8; ++i) {
        float y_ = pstd::tan(x) - y;
        float c = pstd::cos(x);
        d = y_ * (c * c);
        if (i == 7)
            break;
        x -= d;
    }
    return x;
}
PG_HD float atan(float y) {
    float d;
    float x = atan_steps(y, d);
    return x - d;
}
// math.h:343-351.  For x <= 0 pine's build reassociates +-Pi - (x7 - d) into
// (+-Pi - x7) + d.
PG_HD float atan2(float y, float x) {
    if (x > 0)
        return pstd::atan(y / x);
    float d;
    float x7 = atan_steps(-y / x, d);
    if (y > 0)
        return (Pi - x7) + d;
    return (-Pi - x7) + d;
}

PG_HD float clamp(float v, float a, float b) { return fminp(fmaxp(v, a), b); }
PG_HD float sqr(float v) { return v * v; }

}  // namespace pstd

// vecmath.h:834-860 with pstd::sqrt.
PG_HD float length(f3 v) { return pstd::sqrt(dot(v, v)); }
PG_HD f3 normalize(f3 v) { return v / length(v); }
PG_HD f3 normalize(f3 v, float& len) {
    len = length(v);
    return v / len;
}

// vecmath.h:1223-1236
PG_HD void coordinate_system(f3 n, f3& t, f3& b) {
    if (fabsp(n.x) > fabsp(n.y))
        t = normalize(cross(n, mk(0, 1, 0)));
    else
        t = normalize(cross(n, mk(1, 0, 0)));
    b = cross(n, t);
}

}  // namespace pg
