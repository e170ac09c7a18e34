// pinecpu: CPU restatement of pine's path-tracing hot path, used ONLY as test
// infrastructure (the checker in tests/, smoke() and bench.py's cpu_baseline).
// It is never linked into or called by the shipped library.
//
// Every function cites the reference lines it restates.  Pins (how each part is
// checked against something other than itself) are listed in tests/:
//   - po_pstd_*            : golden vectors written by oracle/_ref/ref_harness
//                            (pine's own compiled pstd), tests/test_oracle.py
//   - po_trace_brute       : pine's BVH::Intersect/Hit via the harness on the same
//                            rays, plus closed forms (sphere/plane distances)
//   - po_bvh_*             : same answers as po_trace_brute (any valid tree gives
//                            the same closest t)
//   - po_halton_values     : golden vectors of pine's HaltonSampler (harness)
//   - po_noise             : pine's PerlinNoise / Turbulence / Turbulence3D at 2,000
//                            points (harness golden); closed form: gradient noise
//                            is 1/2 at every lattice point
//   - po_sobol_values      : golden vectors of pine's SobolSampler (harness), all
//                            four scramblers; the per-pixel index is found by
//                            search (closed form: its first two Sobol dimensions
//                            fall in the pixel), not from pine's index tables
//
// Built by oracle/Makefile into oracle/libpinecpu.so with g++ -O2
// -ffp-contract=off (no fast-math): x86-64 has no FMA by default, so each
// product and sum rounds exactly as in pine's build.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

// Joe & Kuo direction numbers (data only): {s, a, m_1..m_s} for dimensions 1..1023.
#include "../pine_amd/csrc/pg_sobol_dirnums.h"

namespace {

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
// vec3 / float: pine's -ffast-math build computes one reciprocal and three
// multiplies (a plane normal from the reference matches only this form).
inline V3 operator/(V3 a, float s) {
    const float r = 1.0f / s;
    return v3(a.x * r, a.y * r, a.z * r);
}
// vecmath.h:874-876 / 892-895
inline float Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 Cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

uint32_t bits(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
float fbits(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// ---- pstd libm (pstd/math.h) ------------------------------------------------
const float kPi = 3.14159265358979323846f;

int IeeeExp(float x) { return (int)(0xff & (bits(x) >> 23)) - 127; }  // math.h:50-52
float Exp2i(int e) { return fbits((uint32_t)(e + 127) << 23); }     // math.h:54-56

float Sqrt(float y) {  // math.h:141-151
    float x = Exp2i(IeeeExp(y) / 2);
    for (int i = 0; i < 4; i++)
        x = x / 2 + y / (2 * x);
    return x;
}
float Floor(float v) {  // math.h:111-117
    int32_t i = (int32_t)v;
    if (v < 0 && (float)i != v)
        i -= 1;
    return (float)i;
}
float Powi(float x, int e) {  // math.h:153-165
    float y = 1;
    while (e) {
        if (e & 1)
            y *= x;
        x *= x;
        e >>= 1;
    }
    return y;
}
float Pow(float x, float e) {  // math.h:167-185
    int ei = (int)Floor(e);
    e -= ei;
    float y = ei > 0 ? Powi(x, ei) : 1 / Powi(x, -ei);
    for (int i = 0; i < 8; ++i) {
        if (e > 1.0f) {
            y *= x;
            e -= 1.0f;
        }
        e *= 2;
        x = Sqrt(x);
    }
    return y;
}
float Log2(float y) {  // math.h:187-212
    int lg = IeeeExp(y);
    float logl = 0, logr = 1, expl = Exp2i(0), expr = Exp2i(1);
    y /= Exp2i(lg);
    for (int i = 0; i < 8; ++i) {
        float logm = (logl + logr) / 2;
        float expm = Sqrt(expl * expr);
        if (expm < y) {
            logl = logm;
            expl = expm;
        } else {
            logr = logm;
            expr = expm;
        }
    }
    return lg + (logl + logr) / 2;
}
float Log(float y) { return 0.69314718055994530942f * Log2(y); }  // math.h:214-217
// math.h:229-239.  Pine's -ffast-math build evaluates x*y/i as y*(x*(1/i)) with
// 1/i folded to a float constant (the only form matching pine's vectors).
float Exp(float x) {
    float y = 0;
    for (int i = 8; i > 0; --i)
        y = 1 + y * (x * (1.0f / (float)i));
    return y;
}
// math.h:251-260, compiled by pine's build as y*((x*x)*(1/(i*(i-1)))).
float Cos0(float x) {
    float y = 1;
    const float x2 = x * x;
    for (int i = 8; i > 0; i -= 2)
        y = 1 - y * (x2 * (1.0f / (float)(i * (i - 1))));
    return y;
}
// math.h:262-276.  Pine's build divides by 2*Pi as a multiply by the folded
// reciprocal and, in the x > Pi branch, reassociates 2*Pi - (x - y*2*Pi) into
// (2*Pi - x) + y*2*Pi; both forms are reproduced here.
float Cos(float x) {
    if (x < 0)
        x = -x;
    uint32_t y = (uint32_t)(int64_t)(x * (1.0f / (kPi * 2)));
    float yp = (float)y * (kPi * 2);
    float xr = x - yp;
    if (xr > kPi)
        xr = (kPi * 2 - x) + yp;
    if (xr < kPi / 2)
        return Cos0(xr);
    return -Cos0(kPi - xr);
}
float Sin(float x) { return Cos(x - kPi / 2); }  // math.h:278-281
float Tan(float x) { return Sin(x) / Cos(x); }   // math.h:283-287
float Acos(float y) {                            // math.h:289-306
    if (y < -1 || y > 1)
        return 0;
    float x = (1 - y) * kPi / 2;
    for (int i = 0; i < 8; ++i) {
        float y_ = Cos(x) - y;
        float dydx = -Sin(x);
        x -= y_ / dydx;
    }
    return x;
}
// math.h:327-341.  Pine's build folds y_ / (1 / cos^2) into y_ * cos^2.
// AtanSteps runs the first seven Newton steps and returns the last correction
// in `d`, so atan = x - d; atan2 needs the two parts (see below).
float AtanSteps(float y, float& d) {
    float x = (y > 0 ? 1 - 1 / (1 + y) : -1 + 1 / (1 - y)) * kPi / 2;
    for (int i = 0; i < 8; ++i) {
        float y_ = Tan(x) - y;
        float c = Cos(x);
        d = y_ * (c * c);
        if (i == 7)
            break;
        x -= d;
    }
    return x;
}
float Atan(float y) {
    float d;
    float x = AtanSteps(y, d);
    return x - d;
}
// math.h:343-351.  For x <= 0 pine's build reassociates +-Pi - (x7 - d) into
// (+-Pi - x7) + d.
float Atan2(float y, float x) {
    if (x > 0)
        return Atan(y / x);
    float d;
    float x7 = AtanSteps(-y / x, d);
    if (y > 0)
        return (kPi - x7) + d;
    return (-kPi - x7) + d;
}

float Length(V3 v) { return Sqrt(Dot(v, v)); }        // vecmath.h:834-836
V3 Normalize(V3 v) { return v / Length(v); }          // vecmath.h:856-858
void CoordSys(V3 n, V3& t, V3& b) {                   // vecmath.h:1223-1229
    if ((n.x > 0 ? n.x : -n.x) > (n.y > 0 ? n.y : -n.y))
        t = Normalize(Cross(n, v3(0, 1, 0)));
    else
        t = Normalize(Cross(n, v3(1, 0, 0)));
    b = Cross(n, t);
}

// ---- shapes -----------------------------------------------------------------
struct Ray {
    V3 o, d;
    float tmin, tmax;
};

// Triangle::Intersect (geometry.h:271-294); the divisions by D are true
// divisions in pine's build (checked bit for bit against its golden hits).
bool TriIntersect(const Ray& r, float tmax, V3 v0, V3 v1, V3 v2, float& t, float& u, float& v) {
    V3 E1 = v1 - v0, E2 = v2 - v0, T = r.o - v0;
    V3 P = Cross(r.d, E2), Q = Cross(T, E1);
    float D = Dot(P, E1);
    if (D == 0.0f)
        return false;
    t = Dot(Q, E2) / D;
    if (t <= r.tmin || t >= tmax)
        return false;
    u = Dot(P, T) / D;
    if (u < 0.0f || u > 1.0f)
        return false;
    v = Dot(Q, r.d) / D;
    if (v < 0.0f || v > 1.0f)
        return false;
    return !(u + v > 1.0f);
}
// Triangle::Hit (geometry.h:251-270)
bool TriHit(const Ray& r, V3 v0, V3 v1, V3 v2) {
    V3 E1 = v1 - v0, E2 = v2 - v0, T = r.o - v0;
    V3 P = Cross(r.d, E2), Q = Cross(T, E1);
    float D = Dot(P, E1);
    if (D == 0.0f)
        return false;
    float t = Dot(Q, E2) / D;
    if (t < r.tmin || t > r.tmax)
        return false;
    float u = Dot(P, T) / D;
    if (u < 0.0f || u > 1.0f)
        return false;
    float v = Dot(Q, r.d) / D;
    if (v < 0.0f || v > 1.0f)
        return false;
    return u + v < 1.0f;
}

// Analytic shape in scene terms: type 1 sphere {c, r}, 2 rect {pos, ex, ey},
// 3 plane {pos, normal}; derived fields follow the pine constructors.
struct Analytic {
    int type;
    V3 a, b, c;  // raw scene vectors
    // derived
    float r;
    V3 ex, ey, n, u, v;
    float lx, ly;
};

Analytic MakeAnalytic(const float* f) {
    Analytic s{};
    s.type = (int)f[0];
    s.a = v3(f[1], f[2], f[3]);
    s.b = v3(f[4], f[5], f[6]);
    s.c = v3(f[7], f[8], f[9]);
    if (s.type == 1) {
        s.r = f[4];
    } else if (s.type == 2) {  // Rect ctor, geometry.h:337-343
        s.ex = Normalize(s.b);
        s.ey = Normalize(s.c);
        s.n = Normalize(Cross(s.b, s.c));
        s.lx = Length(s.b);
        s.ly = Length(s.c);
    } else {  // Plane::Create + ctor, geometry.cpp:340-342, geometry.h:121-125
        s.n = Normalize(s.b);
        CoordSys(s.n, s.u, s.v);
    }
    return s;
}

// Sphere::ComputeT (geometry.cpp:114-126).  Pine's -ffast-math build
// reassociates c = Dot(ro,ro) + Dot(p,p) - 2*Dot(ro,p) - r*r into
// ((py^2 + pz^2) - (2*Dot(ro,p) + r^2)) + ((px^2 + roz^2) + (rox^2 + roy^2))
// (read from the reference object built here); a and b keep source order.
float SphereT(V3 ro, V3 rd, float tmin, V3 p, float r) {
    float a = Dot(rd, rd);
    float b = 2 * Dot(ro - p, rd);
    float c = ((p.y * p.y + p.z * p.z) - (2 * Dot(ro, p) + r * r)) +
              ((p.x * p.x + ro.z * ro.z) + (ro.x * ro.x + ro.y * ro.y));
    float d = b * b - 4 * a * c;
    if (d <= 0.0f)
        return -1.0f;
    d = Sqrt(d);
    float t = (-b - d) / (2 * a);
    if (t < tmin)
        t = (-b + d) / (2 * a);
    return t;
}

// Plane/Rect distance (geometry.cpp:90, 96, 161, 176): (Dot(pos,n) - Dot(o,n)) /
// Dot(d,n).  Pine's build evaluates the numerator in a different order in the
// Intersect and the Hit members; both are reproduced.
float PlaneNumIntersect(V3 o, V3 pos, V3 n) {
    return ((pos.z - o.z) * n.z - (o.x * n.x + o.y * n.y)) + (pos.x * n.x + pos.y * n.y);
}
float PlaneNumHit(V3 o, V3 pos, V3 n) {
    return ((pos.x * n.x + pos.y * n.y) - (o.x * n.x + o.y * n.y)) + (pos.z - o.z) * n.z;
}
// Rect (u, v): p = o + t*d, then Dot(p - pos, e) / l, evaluated by pine's build
// with p - pos = t*d + (o - pos).
float RectCoord(V3 o, V3 d, float t, V3 pos, V3 e, float l) {
    V3 q = v3(d.x * t + (o.x - pos.x), d.y * t + (o.y - pos.y), d.z * t + (o.z - pos.z));
    return Dot(q, e) / l;
}

// Closest test; returns true and sets t/u/v if accepted against tmax per the
// shape's own interval rule (geometry.cpp:95-195).
bool AnaIntersect(const Analytic& s, const Ray& r, float tmax, bool tie_ok, float& t, float& u,
                  float& v) {
    u = v = 0.0f;
    if (s.type == 1) {
        t = SphereT(r.o, r.d, r.tmin, s.a, s.r);
        if (t < r.tmin)
            return false;
        return tie_ok ? !(t > tmax) : t < tmax;
    }
    t = PlaneNumIntersect(r.o, s.a, s.n) / Dot(r.d, s.n);
    if (t < r.tmin)
        return false;
    if (tie_ok ? t > tmax : t >= tmax)
        return false;
    if (s.type == 2) {
        u = RectCoord(r.o, r.d, t, s.a, s.ex, s.lx);
        if (u < -0.5f || u > 0.5f)
            return false;
        v = RectCoord(r.o, r.d, t, s.a, s.ey, s.ly);
        if (v < -0.5f || v > 0.5f)
            return false;
        u += 0.5f;
        v += 0.5f;
    }
    return t == t;
}
bool AnaHit(const Analytic& s, const Ray& r) {  // Sphere/Rect/Plane::Hit
    if (s.type == 1) {
        float t = SphereT(r.o, r.d, r.tmin, s.a, s.r);
        return t > r.tmin && t < r.tmax;
    }
    float t = PlaneNumHit(r.o, s.a, s.n) / Dot(r.d, s.n);
    if (s.type == 3)
        return !(t <= r.tmin) && t < r.tmax;
    if (t < r.tmin || t >= r.tmax)
        return false;
    float u = RectCoord(r.o, r.d, t, s.a, s.ex, s.lx);
    if (u < -0.5f || u > 0.5f)
        return false;
    float v = RectCoord(r.o, r.d, t, s.a, s.ey, s.ly);
    return !(v < -0.5f || v > 0.5f);
}


// ---- RNG + Halton sampler ------------------------------------------------------
// rng.h:39-44 SplitMix64, rng.h:62-105 RNG (xoroshiro128+ step).
uint64_t SplitMix64(uint64_t& s) {
    uint64_t r = s += 0x9E3779B97f4A7C15ULL;
    r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ULL;
    r = (r ^ (r >> 27)) * 0x94D049BB133111EBULL;
    return r ^ (r >> 31);
}
struct Rng {
    uint64_t s[2];
    explicit Rng(uint64_t seed = 0) {
        s[0] = SplitMix64(seed);
        s[1] = SplitMix64(seed);
    }
    uint64_t U64() {
        const uint64_t s0 = s[0];
        uint64_t s1 = s[1];
        const uint64_t result = s0 + s1;
        s1 ^= s0;
        s[0] = ((s0 << 24) | (s0 >> 40)) ^ s1 ^ (s1 << 16);
        s[1] = (s1 << 37) | (s1 >> 27);
        return result;
    }
    uint32_t U32() {
        uint64_t h = U64();
        return (uint32_t)(h ^ (h >> 32));
    }
};

struct HaltonTables {
    std::vector<int> primes, sums;
    std::vector<uint16_t> perms;
    HaltonTables() {
        // First 1000 primes (util/primes.cpp) and their prefix sums.
        for (int n = 2; (int)primes.size() < 1000; n++) {
            bool prime = true;
            for (int p : primes) {
                if (p * p > n)
                    break;
                if (n % p == 0) {
                    prime = false;
                    break;
                }
            }
            if (prime)
                primes.push_back(n);
        }
        int acc = 0;
        for (int p : primes) {
            sums.push_back(acc);
            acc += p;
        }
        // ComputeRadicalInversePermutations (lowdiscrepancy.cpp:5-21) with
        // Shuffle (lowdiscrepancy.h:55-62) driven by RNG(0).
        perms.resize(acc);
        Rng rng(0);
        uint16_t* p = perms.data();
        for (int i = 0; i < 1000; i++) {
            for (int j = 0; j < primes[i]; j++)
                p[j] = (uint16_t)j;
            for (int k = 0; k < primes[i]; k++) {
                int other = k + (int)(rng.U32() % (uint32_t)(primes[i] - k));
                std::swap(p[k], p[other]);
            }
            p += primes[i];
        }
    }
};
const HaltonTables& Tables() {
    static HaltonTables t;
    return t;
}

// HaltonSampler (sampler.cpp:32-96).
struct Halton {
    int baseScales[2], baseExponents[2], multInverse[2];
    int stride;
    uint64_t index = 0;
    int dim = 0;

    static void Egcd(uint64_t a, uint64_t b, int64_t& g, int64_t& x, int64_t& y) {
        int d = (int)(a / b);
        int r = (int)(a - (uint64_t)d * b);
        if (r == 0) {
            g = (int64_t)b;
            x = 0;
            y = 1;
            return;
        }
        int64_t nx, ny;
        Egcd(b, (uint64_t)r, g, nx, ny);
        x = ny;
        y = nx - d * ny;
    }
    static uint64_t MulInv(int64_t a, int64_t n) {
        int64_t g, x, y;
        Egcd((uint64_t)a, (uint64_t)n, g, x, y);
        int64_t r = x - (x / n) * n;
        return (uint64_t)(r < 0 ? r + n : r);
    }
    Halton(int w, int h) {
        int film[2] = {w, h};
        for (int i = 0; i < 2; i++) {
            int base = i == 0 ? 2 : 3, scale = 1, exp = 0;
            while (scale < std::min(film[i], 128)) {
                scale *= base;
                ++exp;
            }
            baseScales[i] = scale;
            baseExponents[i] = exp;
        }
        multInverse[0] = (int)MulInv(baseScales[0], baseScales[1]);
        multInverse[1] = (int)MulInv(baseScales[1], baseScales[0]);
        stride = baseScales[0] * baseScales[1];
    }
    static uint64_t InverseRadicalInverse(uint64_t inverse, int base, int nDigits) {
        uint64_t index = 0;
        for (int i = 0; i < nDigits; i++) {
            uint64_t digit = inverse % base;
            inverse /= base;
            index = index * base + digit;
        }
        return index;
    }
    void StartPixel(int px, int py, int sampleIndex) {
        int64_t hi = 0;
        if (stride > 1) {
            int pm[2] = {px % 128, py % 128};
            for (int i = 0; i < 2; i++) {
                uint64_t off = InverseRadicalInverse((uint64_t)pm[i], i == 0 ? 2 : 3, baseExponents[i]);
                hi += (int64_t)(off * (uint64_t)baseScales[1 - i] * (uint64_t)multInverse[1 - i]);
            }
            hi %= stride;
        }
        hi += (int64_t)sampleIndex * stride;
        index = (uint64_t)hi;
        dim = 2;
    }
    void NextSample() {
        index += (uint64_t)stride;
        dim = 2;
    }
    // ScrambledRadicalInverse (lowdiscrepancy.h:28-42)
    static float Scrambled(int baseIndex, uint64_t a) {
        const HaltonTables& T = Tables();
        int base = T.primes[baseIndex];
        const uint16_t* perm = &T.perms[T.sums[baseIndex]];
        float invBase = 1.0f / base, invBaseN = 1;
        uint64_t reversed = 0;
        while (a) {
            uint64_t next = a / base;
            uint64_t digit = a - next * base;
            reversed = reversed * base + perm[digit];
            invBaseN *= invBase;
            a = next;
        }
        float series = perm[0] / (base + 1.0f);
        float v = ((float)reversed + series) * invBaseN;
        return v < 0x1.fffffep-1f ? v : 0x1.fffffep-1f;
    }
    float Get1D() {
        if (dim >= 1000)
            dim = 2;
        return Scrambled(dim++, index);
    }
};

Ray LoadRay(const float* f) {
    Ray r;
    r.o = v3(f[0], f[1], f[2]);
    r.tmin = f[3];
    r.d = v3(f[4], f[5], f[6]);
    r.tmax = f[7];
    return r;
}

V3 TriV(const float* tri9, int64_t k, int c) {
    const float* p = tri9 + 9 * k + 3 * c;
    return v3(p[0], p[1], p[2]);
}

template <typename F>
void ParallelFor(int64_t n, int threads, F&& f) {
    if (threads <= 0)
        threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&]() {
            while (true) {
                int64_t b = next.fetch_add(64);
                if (b >= n)
                    return;
                int64_t e = std::min(n, b + 64);
                for (int64_t i = b; i < e; i++)
                    f(i);
            }
        });
    for (auto& th : pool)
        th.join();
}

// ---- binned SAH BVH2 + pine-order traversal (CPU baseline) -------------------
struct BNode {
    float lo[2][3], hi[2][3];
    int32_t child[2];  // >=0 inner; <0: ~leaf index
};
struct Leaf {
    int32_t start, count;
};
struct CpuBVH {
    std::vector<BNode> nodes;
    std::vector<Leaf> leaves;
    std::vector<int32_t> order;
    std::vector<float> tri9;  // copy, prim order
    int32_t root_leaf = -1;   // whole tree is one leaf
};

struct PrimBox {
    float lo[3], hi[3], c[3];
    int32_t id;
};

int32_t BuildRec(CpuBVH& b, std::vector<PrimBox>& p, int32_t s, int32_t e, float* lo, float* hi) {
    // Returns child code for range [s, e) and writes its bounds.
    for (int k = 0; k < 3; k++) {
        lo[k] = 3.4e38f;
        hi[k] = -3.4e38f;
    }
    float clo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, chi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (int32_t i = s; i < e; i++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], p[i].lo[k]);
            hi[k] = std::max(hi[k], p[i].hi[k]);
            clo[k] = std::min(clo[k], p[i].c[k]);
            chi[k] = std::max(chi[k], p[i].c[k]);
        }
    int32_t n = e - s;
    auto leaf = [&]() {
        b.leaves.push_back({s, n});
        return ~(int32_t)(b.leaves.size() - 1);
    };
    if (n <= 4)
        return leaf();
    const int B = 16;
    float best = 3.4e38f;
    int bax = -1, bb = -1;
    for (int ax = 0; ax < 3; ax++) {
        float ext = chi[ax] - clo[ax];
        if (!(ext > 0))
            continue;
        int cnt[B] = {};
        float blo[B][3], bhi[B][3];
        for (int k = 0; k < B; k++)
            for (int j = 0; j < 3; j++)
                blo[k][j] = 3.4e38f, bhi[k][j] = -3.4e38f;
        for (int32_t i = s; i < e; i++) {
            int k = std::min(B - 1, (int)((p[i].c[ax] - clo[ax]) / ext * B));
            cnt[k]++;
            for (int j = 0; j < 3; j++) {
                blo[k][j] = std::min(blo[k][j], p[i].lo[j]);
                bhi[k][j] = std::max(bhi[k][j], p[i].hi[j]);
            }
        }
        for (int split = 0; split < B - 1; split++) {
            float l0[3] = {3.4e38f, 3.4e38f, 3.4e38f}, l1[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
            float r0[3] = {3.4e38f, 3.4e38f, 3.4e38f}, r1[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
            int nl = 0, nr = 0;
            for (int k = 0; k < B; k++) {
                float* a0 = k <= split ? l0 : r0;
                float* a1 = k <= split ? l1 : r1;
                (k <= split ? nl : nr) += cnt[k];
                if (cnt[k])
                    for (int j = 0; j < 3; j++) {
                        a0[j] = std::min(a0[j], blo[k][j]);
                        a1[j] = std::max(a1[j], bhi[k][j]);
                    }
            }
            if (!nl || !nr)
                continue;
            auto area = [](float* a, float* c) {
                float dx = c[0] - a[0], dy = c[1] - a[1], dz = c[2] - a[2];
                return dx * dy + dy * dz + dz * dx;
            };
            float cost = area(l0, l1) * nl + area(r0, r1) * nr;
            if (cost < best) {
                best = cost;
                bax = ax;
                bb = split;
            }
        }
    }
    int32_t mid;
    if (bax < 0) {
        mid = s + n / 2;
    } else {
        float ext = chi[bax] - clo[bax];
        mid = (int32_t)(std::partition(p.begin() + s, p.begin() + e,
                                       [&](const PrimBox& q) {
                                           int k = std::min(B - 1, (int)((q.c[bax] - clo[bax]) / ext * B));
                                           return k <= bb;
                                       }) -
                        p.begin());
        if (mid == s || mid == e)
            mid = s + n / 2;
    }
    int32_t id = (int32_t)b.nodes.size();
    b.nodes.push_back(BNode{});
    float l0[3], l1[3], r0[3], r1[3];
    int32_t c0 = BuildRec(b, p, s, mid, l0, l1);
    int32_t c1 = BuildRec(b, p, mid, e, r0, r1);
    BNode& nd = b.nodes[id];
    for (int k = 0; k < 3; k++) {
        nd.lo[0][k] = l0[k];
        nd.hi[0][k] = l1[k];
        nd.lo[1][k] = r0[k];
        nd.hi[1][k] = r1[k];
    }
    nd.child[0] = c0;
    nd.child[1] = c1;
    return id;
}

// AABB::Hit(RayOctant, tmin, *tmax) (geometry.h:93-106): exit distance to *tmax.
bool BoxHit(const float* lo, const float* hi, const V3& inv, const V3& nod, const int* oct,
            float tmin, float* tmax) {
    const float p[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
    const float iv[3] = {inv.x, inv.y, inv.z}, no[3] = {nod.x, nod.y, nod.z};
    float t0 = p[0 + oct[0]] * iv[0] + no[0];
    float t1 = p[1 + oct[1]] * iv[1] + no[1];
    float t2 = p[2 + oct[2]] * iv[2] + no[2];
    float u0 = p[3 - oct[0]] * iv[0] + no[0];
    float u1 = p[4 - oct[1]] * iv[1] + no[1];
    float u2 = p[5 - oct[2]] * iv[2] + no[2];
    float tn = std::max(std::max(std::max(t0, t1), t2), tmin);
    float tf = std::min(std::min(std::min(u0, u1), u2), *tmax);
    *tmax = tf;
    return tn <= tf;
}

}  // namespace

// ---- Sobol sampler (sampler.cpp:129-220, lowdiscrepancy.h:128-158) ---------------
// Generator matrices: column k of dimension d holds direction number v_{k+1}
// as a 32-bit fraction; dimension 0 is van der Corput, the others follow the
// Bratley-Fox recurrence v_k = v_{k-s} ^ (v_{k-s} >> s) ^ XOR_j a_j v_{k-j}.
struct SobolTables {
    std::vector<uint32_t> C;  // [1024][32]
    SobolTables() : C(1024 * 32) {
        for (int k = 0; k < 32; k++)
            C[k] = 1u << (31 - k);
        const uint16_t* q = kSobolDirnums;
        for (int d = 1; d < 1024; d++) {
            const int s = *q++, a = *q++;
            uint32_t* c = &C[d * 32];
            for (int k = 1; k <= 32; k++) {
                if (k <= s) {
                    c[k - 1] = (uint32_t)q[k - 1] << (32 - k);
                } else {
                    uint32_t v = c[k - s - 1] ^ (c[k - s - 1] >> s);
                    for (int j = 1; j < s; j++)
                        if ((a >> (s - 1 - j)) & 1)
                            v ^= c[k - j - 1];
                    c[k - 1] = v;
                }
            }
            q += s;
        }
    }
    // SobolSample before scrambling: the matrix-vector product over GF(2).
    uint32_t Raw(int dim, uint64_t a) const {
        uint32_t v = 0;
        for (int i = 0; a && i < 32; i++, a >>= 1)
            if (a & 1)
                v ^= C[dim * 32 + i];
        return v;
    }
};

uint32_t ReverseBits32(uint32_t x) {  // math.h:19-26
    uint32_t r = 0;
    for (int i = 0; i < 32; i++)
        r |= ((x >> i) & 1u) << (31 - i);
    return r;
}
uint64_t Hash64u(uint64_t s) {  // rng.h:10-15
    s += 0x9E3779B97f4A7C15ULL;
    s = (s ^ (s >> 30)) * 0xBF58476D1CE4E5B9ULL;
    s = (s ^ (s >> 27)) * 0x94D049BB133111EBULL;
    return s ^ (s >> 31);
}
uint64_t MixBits(uint64_t v) {  // rng.h:43-50
    v ^= (v >> 31);
    v *= 0x7fb5d329728ea185ULL;
    v ^= (v >> 27);
    v *= 0x81dadef4bc2dd44dULL;
    v ^= (v >> 33);
    return v;
}

struct Sobol {
    const SobolTables& T;
    int m = 0, strategy, dim = 2;
    uint32_t px = 0, py = 0;
    uint64_t index = 0;
    Sobol(const SobolTables& T, int w, int h, int strategy) : T(T), strategy(strategy) {
        int scale = 1;  // roundup2(max(w, h)), log2int
        while (scale < std::max(w, h))
            scale <<= 1;
        while ((1 << m) < scale)
            m++;
    }
    // SobolIntervalToIndex: the index frame * 4^m + j whose dimension 0 and 1
    // values, scaled by 2^m, floor to (px, py).  Dimension 0 is the bit reversal
    // of the index, so the low m bits of j are px reversed; the high m bits are
    // searched.  Pine keeps the index in 32 bits.
    void Start(uint64_t frame) {
        dim = 2;
        if (m == 0) {
            index = frame;
            return;
        }
        uint64_t lo = 0;
        for (int b = 0; b < m; b++)
            if ((px >> (m - 1 - b)) & 1)
                lo |= 1ull << b;
        for (uint64_t hi = 0; hi < (1ull << m); hi++) {
            const uint64_t i = (frame << (2 * m)) | (hi << m) | lo;
            if ((T.Raw(0, i) >> (32 - m)) == px && (T.Raw(1, i) >> (32 - m)) == py) {
                index = (uint32_t)i;
                return;
            }
        }
        index = ~0ull;  // unreachable: every cell holds exactly one index
    }
    float Sample(int d) const {  // SampleDimension (sampler.cpp:210-220)
        uint32_t v = T.Raw(d, index);
        const uint32_t seed = (uint32_t)Hash64u((uint64_t)(uint32_t)d);  // Hash(int)
        if (strategy == 1) {
            v ^= seed;
        } else if (strategy == 2) {  // FastOwenScramber (sampler.cpp:173-188)
            v = ReverseBits32(v);
            v ^= v * 0x3d20adeau;
            v += seed;
            v *= (seed >> 16) | 1;
            v ^= v * 0x05526c56u;
            v ^= v * 0x53a22864u;
            v = ReverseBits32(v);
        } else if (strategy == 3) {  // OwenScramber (sampler.cpp:190-208)
            if (seed & 1)
                v ^= 1u << 31;
            for (int b = 1; b < 32; b++) {
                uint32_t mask = (~0u) << (32 - b);
                if ((uint32_t)MixBits((v & mask) ^ seed) & (1u << b))
                    v ^= 1u << (31 - b);
            }
        }
        return std::min(v * 0x1p-32f, 0x1.fffffep-1f);
    }
    float Get1D() {  // sampler.cpp:147-151
        if (dim >= 1024)
            dim = 2;
        return Sample(dim++);
    }
};

// ---- Perlin noise (noise.cpp:7-52) -----------------------------------------------
// Corner gradients come from pine's RNG (rng.h:62-105) seeded with Hash(vec3i
// corner, seed) (rng.h:17-29): the corner's three ints as two little-endian
// 64-bit words, then the seed's own hash.
struct PineRng {
    uint64_t a, b;
    explicit PineRng(uint64_t seed) {
        a = SplitMix(seed);
        b = SplitMix(seed);
    }
    static uint64_t SplitMix(uint64_t& s) {  // rng.h:36-41
        uint64_t r = s += 0x9E3779B97f4A7C15ULL;
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ULL;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBULL;
        return r ^ (r >> 31);
    }
    uint64_t Next() {
        const uint64_t s0 = a;
        uint64_t s1 = b;
        const uint64_t r = s0 + s1;
        s1 ^= s0;
        a = ((s0 << 24) | (s0 >> 40)) ^ s1 ^ (s1 << 16);
        b = (s1 << 37) | (s1 >> 27);
        return r;
    }
    float Uniformf() {
        const uint64_t u = Next();
        const uint32_t w = (uint32_t)u + (uint32_t)(u >> 32);
        return std::min(w * 0x1p-32f, 0x1.fffffep-1f);
    }
};

float PerlinNoise(V3 q, float frequency, int seed) {
    q = q * frequency;
    const float fx = Floor(q.x), fy = Floor(q.y), fz = Floor(q.z);
    const float w[3] = {q.x - fx, q.y - fy, q.z - fz};
    float s3[3];
    for (int k = 0; k < 3; k++)
        s3[k] = w[k] * w[k] * (3.0f - 2 * w[k]);
    float ret = 0.0f;
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
            for (int z = 0; z < 2; z++) {
                const int32_t cx = (int32_t)(fx + x), cy = (int32_t)(fy + y), cz = (int32_t)(fz + z);
                uint64_t h = Hash64u(0 ^ ((uint64_t)(uint32_t)cx | (uint64_t)(uint32_t)cy << 32));
                h = Hash64u(h ^ (uint64_t)(uint32_t)cz);
                h = Hash64u(h ^ Hash64u((uint64_t)(uint32_t)seed));
                PineRng rng(h);
                const float u0 = rng.Uniformf(), u1 = rng.Uniformf();
                const float phi = u0 * kPi * 2, theta = u1 * kPi;
                const V3 g = v3(Sin(theta) * Cos(phi), Sin(theta) * Sin(phi), Cos(theta));
                const V3 d = v3(s3[0] - x, s3[1] - y, s3[2] - z);
                // PerlinInterp (vecmath.h:1063-1075): its 1.0 literals make the
                // weights double
                const double wt = (x * s3[0] + (1 - x) * (1.0 - s3[0])) * (y * s3[1] + (1 - y) * (1.0 - s3[1])) *
                                  (z * s3[2] + (1 - z) * (1.0 - s3[2]));
                ret = (float)(ret + wt * Dot(g, d));
            }
    return 0.5f * (1.0f + ret);
}

float Turbulence(V3 q, float frequency, int octaves) {
    float acc = 0.0f, wgt = 1.0f;
    for (int i = 0; i < octaves; i++) {
        acc += wgt * PerlinNoise(q, frequency, 0);
        wgt *= 0.5f;
        q = q * 2.0f;
    }
    const float r = acc / (2.0f - wgt * 2);
    return r * r;
}

V3 Turbulence3D(V3 q, float frequency, int octaves) {
    V3 acc = v3(0, 0, 0);
    float wgt = 1.0f;
    for (int i = 0; i < octaves; i++) {
        acc = acc + v3(PerlinNoise(q, frequency, 0), PerlinNoise(q, frequency, 1), PerlinNoise(q, frequency, 2)) * wgt;
        wgt *= 0.5f;
        q = q * 2.0f;
    }
    const V3 r = acc / (2.0f - wgt * 2);
    return v3(r.x * r.x, r.y * r.y, r.z * r.z);
}

extern "C" {

// Noise at points: out[8 i + 0..7] = perlin seeds 0..2, turbulence, turbulence3d xyz.
void po_noise(const float* pts, const float* freq, const float* octaves, int64_t n, float* out) {
    for (int64_t i = 0; i < n; i++) {
        const V3 q = v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        const V3 t3 = Turbulence3D(q, freq[i], (int)octaves[i]);
        const float v[8] = {PerlinNoise(q, freq[i], 0), PerlinNoise(q, freq[i], 1), PerlinNoise(q, freq[i], 2),
                            Turbulence(q, freq[i], (int)octaves[i]), t3.x, t3.y, t3.z};
        for (int k = 0; k < 7; k++)
            out[8 * i + k] = v[k];
        out[8 * i + 7] = 0.0f;
    }
}

// Sobol values in the harness's layout (as po_halton_values).
void po_sobol_values(int w, int h, int spp, int ndims, int px0, int py0, int npix, int strategy,
                     float* out) {
    static const SobolTables T;
    Sobol s(T, w, h, strategy);
    for (int k = 0; k < npix; k++) {
        int idx = py0 * w + px0 + k * 977;
        s.px = (uint32_t)(idx % w);
        s.py = (uint32_t)((idx / w) % h);
        for (int si = 0; si < spp; si++) {
            s.Start((uint64_t)si);
            for (int d = 0; d < ndims; d++)
                *out++ = s.Get1D();
        }
    }
}

// Halton values in the harness's layout: npix pixels spread as k*977 from
// (px0, py0), spp samples each, ndims successive Get1D values per sample.
void po_halton_values(int w, int h, int spp, int ndims, int px0, int py0, int npix, float* out) {
    Halton s(w, h);
    for (int k = 0; k < npix; k++) {
        int idx = py0 * w + px0 + k * 977;
        int px = idx % w, py = (idx / w) % h;
        s.StartPixel(px, py, 0);
        for (int si = 0; si < spp; si++) {
            for (int d = 0; d < ndims; d++)
                *out++ = s.Get1D();
            s.NextSample();
        }
    }
}

// pstd libm entry points for the golden-vector tests.
float po_pstd_sqrt(float x) { return Sqrt(x); }
float po_pstd_exp(float x) { return Exp(x); }
float po_pstd_log(float x) { return Log(x); }
float po_pstd_cos(float x) { return Cos(x); }
float po_pstd_sin(float x) { return Sin(x); }
float po_pstd_acos(float x) { return Acos(x); }
float po_pstd_pow(float x, float e) { return Pow(x, e); }
float po_pstd_atan2(float y, float x) { return Atan2(y, x); }

// Closest and any hit by testing every primitive.  Prim ids: triangles
// 0..ntri-1, then analytic shapes ntri..ntri+nana-1.  Closest = smallest t,
// exact ties to the smaller prim id.
void po_trace_brute(const float* tri9, int64_t ntri, const float* ana10, int64_t nana,
                    const float* rays8, int64_t n, int32_t* prim, float* tout, float* uout,
                    float* vout, int32_t* occluded, int threads) {
    std::vector<Analytic> ana;
    for (int64_t k = 0; k < nana; k++)
        ana.push_back(MakeAnalytic(ana10 + 10 * k));
    ParallelFor(n, threads, [&](int64_t i) {
        Ray r = LoadRay(rays8 + 8 * i);
        float best = r.tmax, bu = 0, bv = 0;
        int32_t bp = -1;
        bool occ = false;
        for (int64_t k = 0; k < ntri; k++) {
            V3 a = TriV(tri9, k, 0), b = TriV(tri9, k, 1), c = TriV(tri9, k, 2);
            float t, u, v;
            // A tie with the current best is examined against the smaller id rule.
            if (TriIntersect(r, bp >= 0 ? std::nextafter(best, __builtin_inff()) : best, a, b, c, t, u, v)) {
                if (t < best || bp < 0) {
                    best = t, bu = u, bv = v, bp = (int32_t)k;
                }
            }
            if (!occ && occluded && TriHit(r, a, b, c))
                occ = true;
        }
        for (int64_t k = 0; k < nana; k++) {
            float t, u, v;
            if (AnaIntersect(ana[k], r, best, bp >= 0, t, u, v)) {
                if (t < best || (bp < 0) || (t == best && ntri + k < bp))
                    best = t, bu = u, bv = v, bp = (int32_t)(ntri + k);
            }
            if (!occ && occluded && AnaHit(ana[k], r))
                occ = true;
        }
        prim[i] = bp;
        tout[i] = bp >= 0 ? best : r.tmax;
        if (uout)
            uout[i] = bu;
        if (vout)
            vout[i] = bv;
        if (occluded)
            occluded[i] = occ ? 1 : 0;
    });
}

// CPU baseline: binned-SAH BVH2 over triangles, traversed in pine's order
// (bvh.cpp:534-610: both child boxes per node, leaf children tested on the
// spot, the child with the larger exit distance pushed), first found wins ties.
void* po_bvh_build(const float* tri9, int64_t ntri) {
    CpuBVH* b = new CpuBVH();
    std::vector<PrimBox> p((size_t)ntri);
    for (int64_t k = 0; k < ntri; k++) {
        for (int j = 0; j < 3; j++) {
            float a = tri9[9 * k + j], bb = tri9[9 * k + 3 + j], c = tri9[9 * k + 6 + j];
            p[k].lo[j] = std::min(a, std::min(bb, c));
            p[k].hi[j] = std::max(a, std::max(bb, c));
            p[k].c[j] = (p[k].lo[j] + p[k].hi[j]) * 0.5f;
        }
        p[k].id = (int32_t)k;
    }
    float lo[3], hi[3];
    int32_t root = BuildRec(*b, p, 0, (int32_t)ntri, lo, hi);
    if (root < 0)
        b->root_leaf = ~root;
    b->order.resize((size_t)ntri);
    b->tri9.resize((size_t)ntri * 9);
    for (int64_t k = 0; k < ntri; k++) {
        b->order[k] = p[k].id;
        memcpy(&b->tri9[9 * k], tri9 + 9 * (int64_t)p[k].id, 36);
    }
    return b;
}

void po_bvh_free(void* h) { delete (CpuBVH*)h; }

// Returns wall seconds of the traversal loop.
double po_bvh_trace(void* h, const float* rays8, int64_t n, int32_t* prim, float* tout,
                    int threads) {
    const CpuBVH& b = *(const CpuBVH*)h;
    auto t0 = std::chrono::steady_clock::now();
    ParallelFor(n, threads, [&](int64_t i) {
        Ray r = LoadRay(rays8 + 8 * i);
        int oct[3] = {(r.d.x < 0) * 3, (r.d.y < 0) * 3, (r.d.z < 0) * 3};
        V3 inv = v3(r.d.x == 0 ? 1e20f : 1 / r.d.x, r.d.y == 0 ? 1e20f : 1 / r.d.y,
                    r.d.z == 0 ? 1e20f : 1 / r.d.z);
        V3 nod = v3(-r.o.x * inv.x, -r.o.y * inv.y, -r.o.z * inv.z);
        float best = r.tmax;
        int32_t bp = -1;
        auto leaf = [&](int32_t li) {
            const Leaf& L = b.leaves[li];
            for (int32_t k = L.start; k < L.start + L.count; k++) {
                float t, u, v;
                if (TriIntersect(r, best, TriV(b.tri9.data(), k, 0), TriV(b.tri9.data(), k, 1),
                                 TriV(b.tri9.data(), k, 2), t, u, v)) {
                    best = t;
                    bp = b.order[k];
                }
            }
        };
        if (b.root_leaf >= 0) {
            leaf(b.root_leaf);
        } else {
            int32_t stack[64];
            int sp = 0;
            int32_t next = 0;
            while (true) {
                const BNode& nd = b.nodes[next];
                int32_t l = -1, rr = -1;
                float t0 = best, t1 = best;
                if (BoxHit(nd.lo[0], nd.hi[0], inv, nod, oct, r.tmin, &t0)) {
                    if (nd.child[0] >= 0)
                        l = nd.child[0];
                    else
                        leaf(~nd.child[0]);
                }
                if (BoxHit(nd.lo[1], nd.hi[1], inv, nod, oct, r.tmin, &t1)) {
                    if (nd.child[1] >= 0)
                        rr = nd.child[1];
                    else
                        leaf(~nd.child[1]);
                }
                if (l != -1 && rr != -1) {
                    if (t0 > t1) {
                        stack[sp++] = l;
                        next = rr;
                    } else {
                        stack[sp++] = rr;
                        next = l;
                    }
                } else if (l != -1) {
                    next = l;
                } else if (rr != -1) {
                    next = rr;
                } else {
                    if (sp == 0)
                        break;
                    next = stack[--sp];
                }
            }
        }
        prim[i] = bp;
        tout[i] = bp >= 0 ? best : r.tmax;
    });
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
