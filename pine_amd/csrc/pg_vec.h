// Small float3 toolkit shared by host and device code.  Every operation keeps
// pine's evaluation order (src/core/vecmath.h:817-895) so that, compiled with
// -ffp-contract=off, results round exactly as pine's do on x86-64 (no FMA).
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PG_HD __host__ __device__ __forceinline__
#else
#define PG_HD inline
#endif

namespace pg {

struct f3 {
    float x, y, z;
};

PG_HD f3 mk(float x, float y, float z) {
    f3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}
PG_HD f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PG_HD f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PG_HD f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
PG_HD f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
PG_HD f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
PG_HD f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
// Pine's -ffast-math build turns the three divisions of a vector by one scalar
// into one reciprocal and three multiplies (its golden vectors agree only with
// this form), so vector / scalar is written that way here.
PG_HD f3 operator/(f3 a, float s) {
    const float r = 1.0f / s;
    return mk(a.x * r, a.y * r, a.z * r);
}
PG_HD f3 operator/(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }

// vecmath.h:874-876: x*x + y*y + z*z, left to right.
PG_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// vecmath.h:892-895
PG_HD f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
PG_HD float fminp(float a, float b) { return a < b ? a : b; }  // pstd::min (math.h:19-21)
PG_HD float fmaxp(float a, float b) { return a > b ? a : b; }  // pstd::max (math.h:23-25)
PG_HD float fabsp(float x) { return x > 0 ? x : -x; }        // pstd::abs (math.h:37-39)
PG_HD f3 fmin3(f3 a, f3 b) { return mk(fminp(a.x, b.x), fminp(a.y, b.y), fminp(a.z, b.z)); }
PG_HD f3 fmax3(f3 a, f3 b) { return mk(fmaxp(a.x, b.x), fmaxp(a.y, b.y), fmaxp(a.z, b.z)); }

PG_HD uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
PG_HD float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static constexpr float kFloatMax = 3.40282346638528859812e+38f;

}  // namespace pg

namespace pg {

// pine's OffsetRayOrigin (src/core/ray.h:34-45): move p off the surface by 256*n
// units in the last place (integer bit offset), or by n/65536 near the origin.
PG_HD f3 offset_ray_origin(f3 p, f3 n) {
    const float origin = 1.0f / 32.0f;
    const float floatScale = 1.0f / 65536.0f;
    const float intScale = 256.0f;
    int32_t ox = (int32_t)(intScale * n.x), oy = (int32_t)(intScale * n.y),
            oz = (int32_t)(intScale * n.z);
    float px = u2f((uint32_t)((int32_t)f2u(p.x) + (p.x < 0 ? -ox : ox)));
    float py = u2f((uint32_t)((int32_t)f2u(p.y) + (p.y < 0 ? -oy : oy)));
    float pz = u2f((uint32_t)((int32_t)f2u(p.z) + (p.z < 0 ? -oz : oz)));
    return mk(fabsp(p.x) < origin ? p.x + n.x * floatScale : px,
              fabsp(p.y) < origin ? p.y + n.y * floatScale : py,
              fabsp(p.z) < origin ? p.z + n.z * floatScale : pz);
}

}  // namespace pg
