// Per-primitive ray tests with pine's arithmetic, shared by the trace kernels
// and the path tracer.  Compiled with -ffp-contract=off: every product and sum
// rounds separately, as pine's x86-64 build (no FMA) does.
#pragma once

#include "pg_pstd.h"

namespace pg {

// Device record layouts --------------------------------------------------------
// Triangle (48 B): {v0.xyz, prim}, {E1 = v1-v0, 0}, {E2 = v2-v0, 0}.
// Analytic (64 B): {a.xyz, prim}, {b.xyz, type}, {c.xyz, f0}, {d.xyz, f1}
//   sphere: a = centre, b.x = radius
//   rect  : a = position, b = ex/|ex|, c = ey/|ey| (f0 = |ex|), d = n (f1 = |ey|)
//   plane : a = position, b = n, c = u, d = v
enum AnalyticType : int32_t { kSphere = 1, kRect = 2, kPlane = 3 };

struct Ray3 {
    f3 o, d;
    float tmin, tmax;
};

// Triangle::Intersect (geometry.h:271-294).  The three divisions by D stay true,
// correctly rounded divisions, as in pine's build: its golden hits match these
// bit for bit, a reciprocal-multiply form does not.  Accepts when t is inside
// the open interval (tmin, best_t) or ties best_t with a smaller prim id.
PG_HD bool tri_closest(f3 o, f3 d, float tmin, f3 v0, f3 E1, f3 E2, int32_t prim, float& best_t,
                       int32_t& best_prim, float& bu, float& bv) {
    f3 T = o - v0;
    f3 P = cross(d, E2);
    f3 Q = cross(T, E1);
    float D = dot(P, E1);
    if (D == 0.0f)
        return false;
    float t = dot(Q, E2) / D;
    if (t <= tmin || !(t <= best_t))
        return false;
    if (t == best_t && !(best_prim >= 0 && prim < best_prim))
        return false;
    float u = dot(P, T) / D;
    if (u < 0.0f || u > 1.0f)
        return false;
    float v = dot(Q, d) / D;
    if (v < 0.0f || v > 1.0f)
        return false;
    if (u + v > 1.0f)
        return false;
    best_t = t;
    best_prim = prim;
    bu = u;
    bv = v;
    return true;
}

// Triangle::Hit (geometry.h:251-270): closed t interval, strict u+v < 1.
PG_HD bool tri_any(f3 o, f3 d, float tmin, float tmax, f3 v0, f3 E1, f3 E2) {
    f3 T = o - v0;
    f3 P = cross(d, E2);
    f3 Q = cross(T, E1);
    float D = dot(P, E1);
    if (D == 0.0f)
        return false;
    float t = dot(Q, E2) / D;
    if (t < tmin || t > tmax)
        return false;
    float u = dot(P, T) / D;
    if (u < 0.0f || u > 1.0f)
        return false;
    float v = dot(Q, d) / D;
    if (v < 0.0f || v > 1.0f)
        return false;
    return u + v < 1.0f;
}

// Sphere::ComputeT (geometry.cpp:114-126).  Pine's -ffast-math build evaluates
// c = Dot(ro,ro) + Dot(p,p) - 2*Dot(ro,p) - r*r as
// ((py^2 + pz^2) - (2*Dot(ro,p) + r^2)) + ((px^2 + roz^2) + (rox^2 + roy^2));
// a, b and the roots keep the source's rounding.
PG_HD float sphere_t(f3 ro, f3 rd, float tmin, f3 p, float r) {
    float a = dot(rd, rd);
    float b = 2.0f * dot(ro - p, rd);
    float c = ((p.y * p.y + p.z * p.z) - (2.0f * dot(ro, p) + r * r)) +
              ((p.x * p.x + ro.z * ro.z) + (ro.x * ro.x + ro.y * ro.y));
    float disc = b * b - 4.0f * a * c;
    if (disc <= 0.0f)
        return -1.0f;
    disc = pstd::sqrt(disc);
    float t = (-b - disc) / (2.0f * a);
    if (t < tmin)
        t = (-b + disc) / (2.0f * a);
    return t;
}

// Plane/Rect distance (geometry.cpp:90, 96, 161, 176) = (Dot(pos,n) - Dot(o,n))
// / Dot(d,n).  Pine's build orders the numerator differently in the Intersect
// and the Hit members; both orders are kept.
PG_HD float plane_t_intersect(f3 o, f3 d, f3 pos, f3 n) {
    float num = ((pos.z - o.z) * n.z - (o.x * n.x + o.y * n.y)) + (pos.x * n.x + pos.y * n.y);
    return num / dot(d, n);
}
PG_HD float plane_t_hit(f3 o, f3 d, f3 pos, f3 n) {
    float num = ((pos.x * n.x + pos.y * n.y) - (o.x * n.x + o.y * n.y)) + (pos.z - o.z) * n.z;
    return num / dot(d, n);
}
// Rect (u, v) coordinate: Dot(o + t*d - pos, e) / l, with o + t*d - pos
// evaluated as t*d + (o - pos) as in pine's build.
PG_HD float rect_coord(f3 o, f3 d, float t, f3 pos, f3 e, float l) {
    f3 q = mk(d.x * t + (o.x - pos.x), d.y * t + (o.y - pos.y), d.z * t + (o.z - pos.z));
    return dot(q, e) / l;
}

// Closest-hit test of one analytic record; same acceptance rule as tri_closest
// on top of each shape's own interval test.
PG_HD bool analytic_closest(const float* rec, f3 o, f3 d, float tmin, float& best_t,
                            int32_t& best_prim, float& bu, float& bv) {
    f3 a = mk(rec[0], rec[1], rec[2]);
    int32_t prim = (int32_t)f2u(rec[3]);
    f3 b = mk(rec[4], rec[5], rec[6]);
    int32_t type = (int32_t)f2u(rec[7]);
    float t, u = 0.0f, v = 0.0f;
    if (type == kSphere) {
        // Sphere::Intersect (geometry.cpp:131-137): rejects t < tmin and t > tmax.
        t = sphere_t(o, d, tmin, a, b.x);
        if (t < tmin || t > best_t)
            return false;
    } else if (type == kRect) {
        // Rect::Intersect (geometry.cpp:175-195)
        f3 ex = b, ey = mk(rec[8], rec[9], rec[10]), n = mk(rec[12], rec[13], rec[14]);
        float lx = rec[11], ly = rec[15];
        t = plane_t_intersect(o, d, a, n);
        if (t < tmin || !(t <= best_t))
            return false;
        u = rect_coord(o, d, t, a, ex, lx);
        if (u < -0.5f || u > 0.5f)
            return false;
        v = rect_coord(o, d, t, a, ey, ly);
        if (v < -0.5f || v > 0.5f)
            return false;
        u = u + 0.5f;
        v = v + 0.5f;
    } else {
        // Plane::Intersect (geometry.cpp:95-100)
        t = plane_t_intersect(o, d, a, b);
        if (t < tmin || !(t <= best_t))
            return false;
    }
    if (t == best_t && !(best_prim >= 0 && prim < best_prim))
        return false;
    if (!(t == t))
        return false;
    best_t = t;
    best_prim = prim;
    bu = u;
    bv = v;
    return true;
}

// Any-hit test of one analytic record (Sphere/Rect/Plane::Hit).
PG_HD bool analytic_any(const float* rec, f3 o, f3 d, float tmin, float tmax) {
    f3 a = mk(rec[0], rec[1], rec[2]);
    f3 b = mk(rec[4], rec[5], rec[6]);
    int32_t type = (int32_t)f2u(rec[7]);
    if (type == kSphere) {
        float t = sphere_t(o, d, tmin, a, b.x);
        return t > tmin && t < tmax;
    }
    if (type == kRect) {
        f3 ex = b, ey = mk(rec[8], rec[9], rec[10]), n = mk(rec[12], rec[13], rec[14]);
        float lx = rec[11], ly = rec[15];
        float t = plane_t_hit(o, d, a, n);
        if (t < tmin || t >= tmax)
            return false;
        float u = rect_coord(o, d, t, a, ex, lx);
        if (u < -0.5f || u > 0.5f)
            return false;
        float v = rect_coord(o, d, t, a, ey, ly);
        return !(v < -0.5f || v > 0.5f);
    }
    float t = plane_t_hit(o, d, a, b);
    if (t <= tmin)
        return false;
    return t < tmax;
}

}  // namespace pg
