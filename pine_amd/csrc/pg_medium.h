// Participating media of pine's Path integrator: PhaseFunction,
// HomogeneousMedium and GridMedium (src/core/medium.h:13-83,
// src/core/medium.cpp:9-183), device side.
//
// Two tracking modes for a grid:
//  - pine's (brick == 0): delta / ratio tracking against the single global
//    majorant 1/invMaxDensity, consuming sampler dimensions exactly as
//    GridMedium::Sample/Tr do, with the arithmetic of pine's -ffast-math build
//    (the Ln2 of pstd::log folded into the majorant, trilinear taps summed in
//    double with a float rounding after each tap).
//  - bricks (brick > 0): the same estimators against a piecewise-constant
//    majorant, one value per brick of `brick`^3 voxels, walked brick by brick
//    (the exponential step restarts at each brick face).  Unbiased and far
//    fewer null collisions in sparse grids; consumes a different number of
//    sampler dimensions, so images agree with pine's statistically, not
//    sample for sample, and it uses exact logarithms, float densities and
//    hashed uniform numbers (track_uniform) rather than pine's arithmetic and
//    sampler values.  The brick majorants of a launch are staged in LDS.
#pragma once

#include "pg_shade.h"

#include <cmath>

namespace pg {

enum MediumType : int32_t { kMediumHomogeneous = 0, kMediumGrid = 1 };
enum GridMethod : int32_t { kDeltaTracking = 0, kRayMarching = 1 };

static constexpr float kInv4Pi = 0x1.45f306p-4f;  // float(1/(4 pi)), folded by pine's build
static constexpr float kTwoPi = 0x1.921fb6p+2f;

struct MediumDev {
    int32_t type;         // MediumType
    int32_t method;       // grid: GridMethod
    int32_t interpolate;  // grid: trilinear density
    int32_t brick;        // grid: majorant brick edge in voxels, 0 = pine's global majorant
    f3 sigma_a, sigma_s;
    f3 sigma_t3;          // homogeneous: sigma_a + sigma_s
    float sigma_t;        // grid: channel 0 of sigma_a + sigma_s (medium.cpp:168)
    float g;              // HG asymmetry
    float inv_max_density;
    float ln2_inv_max;    // Ln2 * invMaxDensity (pstd::log's Ln2 folded into the majorant)
    float march_step;     // rayMarchingStepSize / max(size) (medium.cpp:172)
    int32_t size[3];
    int32_t nbricks[3];
    f3 w2m_s, w2m_t;      // world -> unit cube: s * p + t (Translate * Scale, medium.cpp:180)
    float m2w[12];        // Inverse(w2m) columns c0, c1, c2, c3 (xyz of each)
    int32_t dblocks[2];     // density blocks along x and y (see density_blocked)
    const float* density;   // pine's densities in blocks of 4 x 4 x 2 voxels (density_blocked)
    const float* majorant;  // nbricks.x * nbricks.y * nbricks.z density bounds
};

// Density layout in HBM: blocks of 4 (x) x 4 (y) x 2 (z) voxels, 32 floats =
// one 128-byte line each, blocks x-fastest, then y, then z; inside a block x,
// then y, then z.  The 8 taps of a trilinear lookup touch 2.3 lines on
// average ((1 + 1/4)(1 + 1/4)(1 + 1/2)) instead of the 4 of pine's x-fastest
// layout.  The values are pine's; only their addresses change.
PG_HD int32_t dens_off_x(int x) { return (x >> 2) * 32 + (x & 3); }
PG_HD int32_t dens_off_y(const MediumDev& m, int y) { return (y >> 2) * (m.dblocks[0] * 32) + (y & 3) * 4; }
PG_HD int32_t dens_off_z(const MediumDev& m, int z) {
    return (z >> 1) * (m.dblocks[0] * m.dblocks[1] * 32) + (z & 1) * 16;
}
// Blocks per axis and padded length (floats) of a grid of size[3] voxels.
inline void density_blocks(const int32_t size[3], int32_t blocks[3], int64_t& floats) {
    blocks[0] = (size[0] + 3) / 4;
    blocks[1] = (size[1] + 3) / 4;
    blocks[2] = (size[2] + 1) / 2;
    floats = (int64_t)blocks[0] * blocks[1] * blocks[2] * 32;
}

// PhaseHG (medium.cpp:9-12) as pine's build evaluates it:
// ((1 - g^2) / (4 pi)) / (sqrt(d) * d), d = (2 cos g + g^2) + 1.
PG_HD float phase_hg(float cos_theta, float g) {
    const float denom = ((cos_theta + cos_theta) * g + g * g) + 1.0f;
    return ((1.0f - g * g) * kInv4Pi) / (pstd::sqrt(denom) * denom);
}

// PhaseFunction::P (medium.cpp:14-16): evaluated at |cos|, as pine does.
PG_HD float phase_p(float g, f3 wi, f3 wo) { return phase_hg(fabsp(dot(wi, wo)), g); }

// PhaseFunction::Sample (medium.cpp:17-31); returns the pdf (= the phase value).
PG_HD float phase_sample(float g, f3 wi, float u0, float u1, f3& wo) {
    float cos_theta;
    if (fabsp(g) < 1e-3f) {
        cos_theta = 1.0f - (u0 + u0);
    } else {
        const float sqr_term = (1.0f - g * g) / ((g + 1.0f) - (u0 + u0) * g);
        cos_theta = ((sqr_term * sqr_term + (-1.0f - g * g)) * 0.5f) / g;
    }
    const float sin_theta = pstd::sqrt(1.0f - cos_theta * cos_theta);
    const float phi = kTwoPi * u1;
    const M3 m = frame_from(wi);
    wo = mul(m, mk(sin_theta * pstd::cos(phi), sin_theta * pstd::sin(phi), cos_theta));
    return phase_hg(cos_theta, g);
}

// --- grid ------------------------------------------------------------------

struct MRay {
    f3 o, d;
    float tmin, tmax;
};

// TransformRayOD(w2m, ray) (ray.h:62-66); w2m is a scale plus a translation, so
// its zero entries add nothing and the product is s*p + t per axis.
PG_HD MRay to_model(const MediumDev& m, f3 o, f3 d, float tmin, float tmax) {
    MRay r;
    r.o = mk(m.w2m_s.x * o.x + m.w2m_t.x, m.w2m_s.y * o.y + m.w2m_t.y,
             m.w2m_s.z * o.z + m.w2m_t.z);
    r.d = normalize(mk(m.w2m_s.x * d.x, m.w2m_s.y * d.y, m.w2m_s.z * d.z));
    r.tmin = tmin;
    r.tmax = tmax;
    return r;
}

// m2w * vec4(p, 1) (vecmath.h:770-772) summed as pine's build does:
// (c1*y + c2*z) + (c0*x + c3).
PG_HD f3 to_world(const MediumDev& m, f3 p) {
    const float* c = m.m2w;
    return mk((c[3] * p.y + c[6] * p.z) + (c[0] * p.x + c[9]),
              (c[4] * p.y + c[7] * p.z) + (c[1] * p.x + c[10]),
              (c[5] * p.y + c[8] * p.z) + (c[2] * p.x + c[11]));
}

// AABB(0, 1).Hit(ray, tmin, tmax) (geometry.cpp:67-87), with the max/min
// operand order of pine's build (SSE maxss/minss keep the running value on ties
// and NaN).
PG_HD bool slab_axis(float o, float d, float& t0, float& t1) {
    const float inv = 1.0f / d;
    float tn = (0.0f - o) * inv, tf = (1.0f - o) * inv;
    if (d < 0.0f) {
        const float tmp = tn;
        tn = tf;
        tf = tmp;
    }
    t0 = t0 > tn ? t0 : tn;
    t1 = t1 < tf ? t1 : tf;
    return !(t0 > t1);
}
PG_HD bool unit_box_hit(const MRay& r, float& t0, float& t1) {
    t0 = r.tmin;
    t1 = r.tmax;
    if (t0 > t1)
        return false;
    return slab_axis(r.o.x, r.d.x, t0, t1) && slab_axis(r.o.y, r.d.y, t0, t1) &&
           slab_axis(r.o.z, r.d.z, t0, t1);
}

PG_HD f3 ray_at(const MRay& r, float t) { return mk(t * r.d.x + r.o.x, t * r.d.y + r.o.y, t * r.d.z + r.o.z); }

// GridMedium::D (medium.cpp:156-160): zero outside the grid.
PG_HD float grid_voxel(const MediumDev& m, int x, int y, int z) {
    if (x < 0 || x >= m.size[0] || y < 0 || y >= m.size[1] || z < 0 || z >= m.size[2])
        return 0.0f;
    return m.density[dens_off_x(x) + dens_off_y(m, y) + dens_off_z(m, z)];
}

// The 8 voxels of a trilinear lookup at (ix, iy, iz): each axis's two
// offsets and in-grid flags found once, every tap one add of three offsets.
struct GridTaps {
    int32_t ox[2], oy[2], oz[2];
    bool vx[2], vy[2], vz[2];
};
PG_HD GridTaps grid_taps(const MediumDev& m, int ix, int iy, int iz) {
    GridTaps T;
    for (int k = 0; k < 2; k++) {
        const int x = ix + k, y = iy + k, z = iz + k;
        T.vx[k] = x >= 0 && x < m.size[0];
        T.vy[k] = y >= 0 && y < m.size[1];
        T.vz[k] = z >= 0 && z < m.size[2];
        T.ox[k] = dens_off_x(x);
        T.oy[k] = dens_off_y(m, y);
        T.oz[k] = dens_off_z(m, z);
    }
    return T;
}
PG_HD float grid_tap(const MediumDev& m, const GridTaps& T, int a, int b, int c) {
    return (T.vx[a] && T.vy[b] && T.vz[c]) ? m.density[T.ox[a] + T.oy[b] + T.oz[c]] : 0.0f;
}

// GridMedium::Density (medium.cpp:139-155).  AABB(0,1).Offset(p) is p itself;
// TrilinearInterp (vecmath.h:1052-1061) runs in double (its weights are
// `1.0 - uvw`), each tap product ((wx*wy)*wz)*c added to the float result and
// rounded back to float, taps in x, y, z nested order.
PG_HD float grid_density(const MediumDev& m, f3 p) {
    const float px = p.x * (float)m.size[0], py = p.y * (float)m.size[1],
                pz = p.z * (float)m.size[2];
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    if (!m.interpolate)
        return grid_voxel(m, ix, iy, iz);
    const double X1 = (double)(px - (float)ix), Y1 = (double)(py - (float)iy),
                 Z1 = (double)(pz - (float)iz);
    const double X0 = 1.0 - X1, Y0 = 1.0 - Y1, Z0 = 1.0 - Z1;
    const GridTaps T = grid_taps(m, ix, iy, iz);
    const float c000 = grid_tap(m, T, 0, 0, 0), c001 = grid_tap(m, T, 0, 0, 1);
    const float c010 = grid_tap(m, T, 0, 1, 0), c011 = grid_tap(m, T, 0, 1, 1);
    const float c100 = grid_tap(m, T, 1, 0, 0), c101 = grid_tap(m, T, 1, 0, 1);
    const float c110 = grid_tap(m, T, 1, 1, 0), c111 = grid_tap(m, T, 1, 1, 1);
    const double x0y0 = X0 * Y0, x0y1 = X0 * Y1, x1y0 = X1 * Y0, x1y1 = X1 * Y1;
    float ret = (float)(x0y0 * Z0 * (double)c000);
    ret = (float)((double)ret + x0y0 * Z1 * (double)c001);
    ret = (float)((double)ret + x0y1 * Z0 * (double)c010);
    ret = (float)((double)ret + x0y1 * Z1 * (double)c011);
    ret = (float)((double)ret + x1y0 * Z0 * (double)c100);
    ret = (float)((double)ret + x1y0 * Z1 * (double)c101);
    ret = (float)((double)ret + x1y1 * Z0 * (double)c110);
    ret = (float)((double)ret + x1y1 * Z1 * (double)c111);
    return ret;
}

// --- brick-majorant tracking: this library's own estimator -------------------
// Tracking against brick majorants does not replay pine's sample sequence, so
// it need not reproduce pine's arithmetic: free flights use the hardware log2
// (pine's bisection log2 costs 32 IEEE divisions and is off by up to 2^-9,
// which skews the exponential distribution), densities a float trilinear
// interpolation, and the uniform numbers of flights, collisions and Russian
// roulette come from track_uniform (a hash of the sample index and the
// dimension, the dimensions advancing as the sampler's would).

// log2 of x > 0: v_log_f32 on the device.
PG_HD float fast_log2(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_logf(x);
#else
    return std::log2(x);
#endif
}

// A free flight from t at extinction `rate`: t - ln(1 - u) / rate, with the
// hardware log2 and scale = ln 2 / rate (brick_flight_scale), divided once per
// brick instead of once per flight.
PG_HD float brick_flight_scale(float rate) { return 0.693147182f / rate; }
PG_HD float brick_flight(float t, float u, float scale) {
    return t - fast_log2(1.0f - u) * scale;
}

// GridMedium::Density (medium.cpp:139-155) in float: trilinear weights as
// lerps over x, then y, then z; nearest voxel when interpolate is off.
PG_HD float grid_density_fast(const MediumDev& m, f3 p) {
    const float px = p.x * (float)m.size[0], py = p.y * (float)m.size[1],
                pz = p.z * (float)m.size[2];
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    if (!m.interpolate)
        return grid_voxel(m, ix, iy, iz);
    const float fx = px - (float)ix, fy = py - (float)iy, fz = pz - (float)iz;
    const GridTaps T = grid_taps(m, ix, iy, iz);
    const float c000 = grid_tap(m, T, 0, 0, 0), c001 = grid_tap(m, T, 0, 0, 1);
    const float c010 = grid_tap(m, T, 0, 1, 0), c011 = grid_tap(m, T, 0, 1, 1);
    const float c100 = grid_tap(m, T, 1, 0, 0), c101 = grid_tap(m, T, 1, 0, 1);
    const float c110 = grid_tap(m, T, 1, 1, 0), c111 = grid_tap(m, T, 1, 1, 1);
    const float c00 = fmaf(fx, c100 - c000, c000), c01 = fmaf(fx, c101 - c001, c001);
    const float c10 = fmaf(fx, c110 - c010, c010), c11 = fmaf(fx, c111 - c011, c011);
    const float c0 = fmaf(fy, c10 - c00, c00), c1 = fmaf(fy, c11 - c01, c01);
    return fmaf(fz, c1 - c0, c0);
}

// One free-flight step of pine's trackers: t -= log(1 - u) * invMaxDensity /
// sigma_t, with pstd::log's Ln2 folded into the majorant as pine's build does.
PG_HD float track_step(const MediumDev& m, float t, float u) {
    return t - (pstd::log2(1.0f - u) * m.ln2_inv_max) / m.sigma_t;
}

// Brick walker over the majorant grid (Amanatides-Woo in brick units): the
// current brick, and the ray distance at which it is left through each axis.
struct BrickWalk {
    int cx, cy, cz;
    int sx, sy, sz;
    float nx, ny, nz;  // distance of the next brick face per axis
    float dx, dy, dz;  // distance between brick faces per axis
};

PG_HD void walk_axis(float o, float d, float scale, int nb, float t0, int& c, int& step, float& next,
                     float& delta) {
    const float cf = o * scale;
    int ci = (int)floorf(cf);
    c = ci < 0 ? 0 : (ci >= nb ? nb - 1 : ci);
    const float dir = d * scale;  // brick faces crossed per unit t
    if (dir > 0.0f) {
        step = 1;
        delta = 1.0f / dir;
        next = t0 + ((float)(c + 1) - cf) / dir;
    } else if (dir < 0.0f) {
        step = -1;
        delta = -1.0f / dir;
        next = t0 + ((float)c - cf) / dir;
    } else {
        step = 0;
        delta = kFloatMax;
        next = kFloatMax;
    }
}

PG_HD void brick_walk_init(const MediumDev& m, const MRay& r, float t0, BrickWalk& w) {
    const f3 p = ray_at(r, t0);
    walk_axis(p.x, r.d.x, (float)m.size[0] / (float)m.brick, m.nbricks[0], t0, w.cx, w.sx, w.nx, w.dx);
    walk_axis(p.y, r.d.y, (float)m.size[1] / (float)m.brick, m.nbricks[1], t0, w.cy, w.sy, w.ny, w.dy);
    walk_axis(p.z, r.d.z, (float)m.size[2] / (float)m.brick, m.nbricks[2], t0, w.cz, w.sz, w.nz, w.dz);
}

// Distance at which the walk leaves the current brick.
PG_HD float brick_exit(const BrickWalk& w) { return fminf(w.nx, fminf(w.ny, w.nz)); }

// Steps into the next brick; false when the walk leaves the grid.
PG_HD bool brick_advance(const MediumDev& m, BrickWalk& w) {
    if (w.nx <= w.ny && w.nx <= w.nz) {
        w.cx += w.sx;
        w.nx += w.dx;
        return w.cx >= 0 && w.cx < m.nbricks[0];
    }
    if (w.ny <= w.nz) {
        w.cy += w.sy;
        w.ny += w.dy;
        return w.cy >= 0 && w.cy < m.nbricks[1];
    }
    w.cz += w.sz;
    w.nz += w.dz;
    return w.cz >= 0 && w.cz < m.nbricks[2];
}

PG_HD int brick_index(const MediumDev& m, const BrickWalk& w) {
    return w.cx + m.nbricks[0] * (w.cy + m.nbricks[1] * w.cz);
}

// --- medium interfaces used by the integrator --------------------------------

struct MediumEvent {
    f3 weight;   // multiplies the path throughput
    f3 p;        // world-space scattering point when scattered
    bool scattered;
};

// Media a tracking kernel is compiled for: every kind (kM 0), only grids
// tracked with brick majorants by delta / ratio tracking (kM 1), whose kernels
// then carry no code or registers for the other methods, or a scene whose only
// medium is such a grid (kM 2: the medium's fields are uniform across the wave
// and stay in scalar registers).
enum MediumVariant : int { kMediaAny = 0, kMediaBricks = 1, kMediaOneBrickGrid = 2 };

// Medium::Sample (medium.h:70-73) for a ray of the path; `maj` holds the brick
// majorants of grid media in brick mode (LDS copy or global).  kS: the
// sampler variant (get_1d_t), kM: the media variant.
template <int kS = kSamplerAny, int kM = kMediaAny>
PG_HD MediumEvent medium_sample(const MediumDev& m, const float* maj, const SamplerDev& hs,
                                SamplerState& ss, f3 o, f3 d, float tmax, int& lookups) {
    MediumEvent ev;
    ev.scattered = false;
    ev.weight = mk(1.0f, 1.0f, 1.0f);
    if (kM == kMediaAny && m.type == kMediumHomogeneous) {
        // HomogeneousMedium::Sample (medium.cpp:37-53)
        int ch = (int)(get_1d_t<kS>(hs, ss) * 3.0f);
        ch = ch < 0 ? 0 : (ch > 2 ? 2 : ch);
        const float st = ch == 0 ? m.sigma_t3.x : (ch == 1 ? m.sigma_t3.y : m.sigma_t3.z);
        const float dist = -pstd::log(1.0f - get_1d_t<kS>(hs, ss)) / st;
        const bool sampled = dist < tmax;
        const float t = dist < tmax ? dist : tmax;
        const f3 tr = mk(pstd::exp(-m.sigma_t3.x * t), pstd::exp(-m.sigma_t3.y * t),
                         pstd::exp(-m.sigma_t3.z * t));
        if (sampled) {
            const f3 st_tr = m.sigma_t3 * tr;
            const float pdf = (st_tr.x + st_tr.y + st_tr.z) / 3.0f;
            ev.p = o + t * d;
            ev.scattered = true;
            ev.weight = m.sigma_s * tr / pdf;
        } else {
            const float pdf = (tr.x + tr.y + tr.z) / 3.0f;
            ev.weight = tr / pdf;
        }
        return ev;
    }
    // GridMedium::Sample (medium.cpp:96-137)
    const MRay r = to_model(m, o, d, 0.0f, tmax);
    float t0, t1;
    if (!unit_box_hit(r, t0, t1))
        return ev;
    const f3 albedo = m.sigma_s / m.sigma_t;
    if (kM == kMediaAny && m.method == kRayMarching) {
        const float lnXi = -pstd::log(1.0f - get_1d_t<kS>(hs, ss));
        const float stepSize = m.march_step;
        const float jitter = stepSize * (get_1d_t<kS>(hs, ss) - 0.5f);
        float t = t0, sum = 0.0f;
        while (t < t1) {
            lookups++;
            sum += grid_density(m, ray_at(r, t)) * m.sigma_t * stepSize;
            if (sum >= lnXi) {
                ev.p = to_world(m, ray_at(r, t + jitter));
                ev.scattered = true;
                ev.weight = albedo;
                return ev;
            }
            t += stepSize;
        }
        return ev;
    }
    if (kM == kMediaAny && m.brick == 0) {
        float t = t0;
        while (true) {
            t = track_step(m, t, get_1d_t<kS>(hs, ss));
            if (!(t < t1))  // also ends on NaN, where pine would loop forever
                return ev;
            const f3 pm = ray_at(r, t);
            lookups++;
            if (grid_density(m, pm) * m.inv_max_density > get_1d_t<kS>(hs, ss)) {
                ev.p = to_world(m, pm);
                ev.scattered = true;
                ev.weight = albedo;
                return ev;
            }
        }
    }
    // Brick majorants: per brick, free flights against sigma_t * maj[brick].
    BrickWalk w;
    brick_walk_init(m, r, t0, w);
    float t = t0;
    while (true) {
        const float texit = fminf(brick_exit(w), t1);
        const float mj = maj[brick_index(m, w)];
        const float rate = m.sigma_t * mj;
        if (mj > 0.0f && rate > 0.0f) {
            const float inv_mj = 1.0f / mj, scale = brick_flight_scale(rate);
            while (true) {
                t = brick_flight(t, track_uniform<kS>(hs, ss), scale);
                if (!(t < texit))
                    break;
                const f3 pm = ray_at(r, t);
                lookups++;
                if (grid_density_fast(m, pm) * inv_mj > track_uniform<kS>(hs, ss)) {
                    ev.p = to_world(m, pm);
                    ev.scattered = true;
                    ev.weight = albedo;
                    return ev;
                }
            }
        }
        if (texit >= t1)
            return ev;
        t = texit;
        if (!brick_advance(m, w))
            return ev;
    }
}

// Medium::Tr (medium.h:66-69): transmittance over [0, tmax] of the ray.
template <int kS = kSamplerAny, int kM = kMediaAny>
PG_HD f3 medium_tr(const MediumDev& m, const float* maj, const SamplerDev& hs, SamplerState& ss,
                   f3 o, f3 d, float tmax, int& lookups) {
    if (kM == kMediaAny && m.type == kMediumHomogeneous) {
        // HomogeneousMedium::Tr (medium.cpp:33-35): Exp(-ray.tmax * sigma_t)
        return mk(pstd::exp(-tmax * m.sigma_t3.x), pstd::exp(-tmax * m.sigma_t3.y),
                  pstd::exp(-tmax * m.sigma_t3.z));
    }
    // GridMedium::Tr (medium.cpp:55-94)
    const MRay r = to_model(m, o, d, 0.0f, tmax);
    float t0, t1;
    if (!unit_box_hit(r, t0, t1))
        return mk(1.0f, 1.0f, 1.0f);
    if (kM == kMediaAny && m.method == kRayMarching) {
        float tr = 0.0f;
        const float stepSize = m.march_step;
        const float jitter = stepSize * (get_1d_t<kS>(hs, ss) - 0.5f);
        float t = t0;
        while (t < t1) {
            lookups++;
            tr += grid_density(m, ray_at(r, t + jitter)) * m.sigma_t * stepSize;
            t += stepSize;
        }
        const float e = pstd::exp(-tr);
        return mk(e, e, e);
    }
    float tr = 1.0f;
    const float rrThreshold = 0.1f;
    if (kM == kMediaAny && m.brick == 0) {
        float t = t0;
        while (true) {
            t = track_step(m, t, get_1d_t<kS>(hs, ss));
            if (!(t < t1))
                break;
            lookups++;
            const float density = grid_density(m, ray_at(r, t));
            tr *= 1.0f - max0(density * m.inv_max_density, 0.0f);
            if (tr < rrThreshold) {
                const float q = max0(1.0f - tr, 0.05f);
                if (get_1d_t<kS>(hs, ss) < q)
                    return mk(0.0f, 0.0f, 0.0f);
                tr /= 1.0f - q;
            }
        }
        return mk(tr, tr, tr);
    }
    BrickWalk w;
    brick_walk_init(m, r, t0, w);
    float t = t0;
    while (true) {
        const float texit = fminf(brick_exit(w), t1);
        const float mj = maj[brick_index(m, w)];
        const float rate = m.sigma_t * mj;
        if (mj > 0.0f && rate > 0.0f) {
            const float inv_mj = 1.0f / mj, scale = brick_flight_scale(rate);
            while (true) {
                t = brick_flight(t, track_uniform<kS>(hs, ss), scale);
                if (!(t < texit))
                    break;
                lookups++;
                const float density = grid_density_fast(m, ray_at(r, t));
                tr *= 1.0f - max0(density * inv_mj, 0.0f);
                if (tr < rrThreshold) {
                    const float q = max0(1.0f - tr, 0.05f);
                    if (track_uniform<kS>(hs, ss) < q)
                        return mk(0.0f, 0.0f, 0.0f);
                    tr /= 1.0f - q;
                }
            }
        }
        if (texit >= t1)
            break;
        t = texit;
        if (!brick_advance(m, w))
            break;
    }
    return mk(tr, tr, tr);
}

}  // namespace pg
