// Pine's material node graph (src/core/node.h, node.cpp:32-97), host+device.
//
// A material input (albedo, roughness, eta, a bump map) is a tree of nodes
// evaluated per shading point for a float or a vec3 (pine's EvalFloat and
// EvalVec3, whose meanings differ per node).  The host compiles each tree for
// the mode its consumer asks for into a postfix program; the device runs it on
// a small value stack.  Every value is a float3; float-mode values are splats.
//
// Program layout: an header Inst {kProgram, n} followed by n instructions.
#pragma once

#include "pg_pstd.h"

namespace pg {
namespace nodes {

enum Op : int32_t {
    kProgram = 0,  // header: arg = instruction count
    kConst,        // push (x, y, z)
    kPosition,     // push p           (Position::EvalVec3)
    kNormal,       // push n           (Normal::EvalVec3)
    kTexCoord,     // push (u, v, 0)   (TexCoord::EvalVec3)
    kComponent,    // v -> splat(v[arg])                      (Decompose)
    kCompose,      // x, y, z -> (x.x, y.x, z.x)              (Composite::EvalVec3)
    kAdd, kSub, kMul, kDiv,  // a, b -> a op b, per component
    kMulAdd,       // a, b, c -> a * b + c                    (MultiplyAdd)
    kLength,       // v -> splat(Length(v))
    kSqr, kSqrt, kSin, kCos, kTan, kFract, kInvert,  // per component
    kPow,          // v, e -> pow(v_i, e.x)                   (Pow: exponent is EvalFloat)
    kChecker,      // p, f -> splat(checkerboard(p * f.x))
    kNoise,        // p, freq, oct -> splat(Turbulence(p, (int)freq, (int)oct))
    kNoise3D,      // same; arg 1: Turbulence3D (EvalVec3), arg 0: Turbulence
};

struct Inst {
    int32_t op, arg;
    float x, y, z;
};

// NodeEvalCtx (node.h:15-20).
struct Ctx {
    f3 p, n;
    float u, v;
};

constexpr int kStack = 16;

// ---- pine's RNG, Hash and Perlin noise (rng.h:10-105, noise.cpp:7-48) --------
PG_HD uint64_t hash64u(uint64_t s) {  // Hash64u
    s += 0x9E3779B97f4A7C15ULL;
    s = (s ^ (s >> 30)) * 0xBF58476D1CE4E5B9ULL;
    s = (s ^ (s >> 27)) * 0x94D049BB133111EBULL;
    return s ^ (s >> 31);
}
// Hash(vec3i, int): the vec3i's 12 bytes as two zero-padded 64-bit words, then
// the seed's hash mixed in.
PG_HD uint64_t hash_cell(int32_t x, int32_t y, int32_t z, int32_t seed) {
    uint64_t h = hash64u(0 ^ ((uint64_t)(uint32_t)x | ((uint64_t)(uint32_t)y << 32)));
    h = hash64u(h ^ (uint64_t)(uint32_t)z);
    return hash64u(h ^ hash64u((uint64_t)(uint32_t)seed));
}
PG_HD uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
struct Rng {  // RNG (rng.h:62-105): two SplitMix64 draws seed the state
    uint64_t s0, s1;
    PG_HD explicit Rng(uint64_t seed) {
        seed += 0x9E3779B97f4A7C15ULL;
        s0 = mix(seed);
        seed += 0x9E3779B97f4A7C15ULL;
        s1 = mix(seed);
    }
    PG_HD static uint64_t mix(uint64_t r) {
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ULL;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBULL;
        return r ^ (r >> 31);
    }
    PG_HD uint64_t next() {
        const uint64_t a = s0;
        uint64_t b = s1;
        const uint64_t r = a + b;
        b ^= a;
        s0 = rotl64(a, 24) ^ b ^ (b << 16);
        s1 = rotl64(b, 37);
        return r;
    }
    PG_HD float uniform() {
        const uint64_t u = next();
        const float f = (float)(uint32_t)((uint32_t)u + (uint32_t)(u >> 32)) * 0x1p-32f;
        return f < 0x1.fffffep-1f ? f : 0x1.fffffep-1f;
    }
};

// PerlinNoise (noise.cpp:7-21): random unit gradients at the 8 lattice corners
// (RNG seeded by the corner's hash), smoothstep weights, PerlinInterp
// (vecmath.h:1063-1075, whose 1 - w terms are double: 1.0 is a double literal).
PG_HD float perlin(f3 np, float frequency, int seed) {
    np = np * frequency;
    const f3 fl = mk(pstd::floor(np.x), pstd::floor(np.y), pstd::floor(np.z));
    f3 w = np - fl;
    w = w * w * (mk(3.0f, 3.0f, 3.0f) - 2.0f * w);
    float ret = 0.0f;
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
            for (int z = 0; z < 2; z++) {
                Rng r(hash_cell((int32_t)(fl.x + (float)x), (int32_t)(fl.y + (float)y),
                                (int32_t)(fl.z + (float)z), seed));
                const float u0 = r.uniform(), u1 = r.uniform();
                const float phi = u0 * pstd::Pi * 2.0f, theta = u1 * pstd::Pi;
                const float st = pstd::sin(theta);
                const f3 g = mk(st * pstd::cos(phi), st * pstd::sin(phi), pstd::cos(theta));
                const f3 d = w - mk((float)x, (float)y, (float)z);
                const double wx = (double)((float)x * w.x) + (double)(1 - x) * (1.0 - (double)w.x);
                const double wy = (double)((float)y * w.y) + (double)(1 - y) * (1.0 - (double)w.y);
                const double wz = (double)((float)z * w.z) + (double)(1 - z) * (1.0 - (double)w.z);
                ret = (float)((double)ret + wx * wy * wz * (double)dot(g, d));
            }
    return 0.5f * (1.0f + ret);
}

// Turbulence / Turbulence3D (noise.cpp:28-52): octaves of halving weight and
// doubling frequency, normalised, squared.
PG_HD float turbulence(f3 np, float frequency, int octaves) {
    float accum = 0.0f, weight = 1.0f;
    for (int i = 0; i < octaves; i++) {
        accum += weight * perlin(np, frequency, 0);
        weight *= 0.5f;
        np = np * 2.0f;
    }
    const float a = accum / (2.0f - weight * 2.0f);
    return a * a;
}
PG_HD f3 turbulence3d(f3 np, float frequency, int octaves) {
    f3 accum = mk(0.0f, 0.0f, 0.0f);
    float weight = 1.0f;
    for (int i = 0; i < octaves; i++) {
        accum = accum + weight * mk(perlin(np, frequency, 0), perlin(np, frequency, 1), perlin(np, frequency, 2));
        weight *= 0.5f;
        np = np * 2.0f;
    }
    const f3 a = accum / (2.0f - weight * 2.0f);
    return a * a;
}

PG_HD f3 splat(float v) { return mk(v, v, v); }

// Runs one program on the value stack `st` (kStack entries of the caller's
// storage: the shading kernels pass per-slot device memory, so that the
// interpreter adds no scratch memory to them).  A program the host compiled
// never needs more than kStack values (the compiler checks), so the stack
// accesses stay in bounds.
PG_HD f3 eval(const Inst* prog, const Ctx& c, f3* st) {
    int sp = 0;
    const int n = prog[0].arg;
    for (int k = 1; k <= n; k++) {
        const Inst in = prog[k];
        switch (in.op) {
        case kConst: st[sp++] = mk(in.x, in.y, in.z); break;
        case kPosition: st[sp++] = c.p; break;
        case kNormal: st[sp++] = c.n; break;
        case kTexCoord: st[sp++] = mk(c.u, c.v, 0.0f); break;
        case kComponent: {
            const f3 v = st[sp - 1];
            st[sp - 1] = splat(in.arg == 0 ? v.x : in.arg == 1 ? v.y : v.z);
            break;
        }
        case kCompose: st[sp - 3] = mk(st[sp - 3].x, st[sp - 2].x, st[sp - 1].x); sp -= 2; break;
        case kAdd: st[sp - 2] = st[sp - 2] + st[sp - 1]; sp--; break;
        case kSub: st[sp - 2] = st[sp - 2] - st[sp - 1]; sp--; break;
        case kMul: st[sp - 2] = st[sp - 2] * st[sp - 1]; sp--; break;
        case kDiv: {
            const f3 a = st[sp - 2], b = st[sp - 1];
            st[sp - 2] = mk(a.x / b.x, a.y / b.y, a.z / b.z);
            sp--;
            break;
        }
        case kMulAdd: st[sp - 3] = st[sp - 3] * st[sp - 2] + st[sp - 1]; sp -= 2; break;
        case kLength: st[sp - 1] = splat(length(st[sp - 1])); break;
        case kSqr: st[sp - 1] = st[sp - 1] * st[sp - 1]; break;
        case kSqrt: {
            const f3 v = st[sp - 1];
            st[sp - 1] = mk(pstd::sqrt(v.x), pstd::sqrt(v.y), pstd::sqrt(v.z));
            break;
        }
        case kSin: {
            const f3 v = st[sp - 1];
            st[sp - 1] = mk(pstd::sin(v.x), pstd::sin(v.y), pstd::sin(v.z));
            break;
        }
        case kCos: {
            const f3 v = st[sp - 1];
            st[sp - 1] = mk(pstd::cos(v.x), pstd::cos(v.y), pstd::cos(v.z));
            break;
        }
        case kTan: {
            const f3 v = st[sp - 1];
            st[sp - 1] = mk(pstd::tan(v.x), pstd::tan(v.y), pstd::tan(v.z));
            break;
        }
        case kFract: {
            const f3 v = st[sp - 1];
            st[sp - 1] = mk(pstd::fract(v.x), pstd::fract(v.y), pstd::fract(v.z));
            break;
        }
        case kInvert: st[sp - 1] = mk(1.0f, 1.0f, 1.0f) - st[sp - 1]; break;
        case kPow: {
            const f3 v = st[sp - 2];
            const float e = st[sp - 1].x;
            st[sp - 2] = mk(pstd::pow(v.x, e), pstd::pow(v.y, e), pstd::pow(v.z, e));
            sp--;
            break;
        }
        case kChecker: {
            // Checkerboard::EvalFloat (node.h:311-317)
            const f3 q = st[sp - 2] * st[sp - 1].x;
            const float r = (pstd::fract(q.x) - 0.5f) * (pstd::fract(q.y) - 0.5f) * (pstd::fract(q.z) - 0.5f);
            st[sp - 2] = splat(r > 0.0f ? 1.0f : 0.0f);
            sp--;
            break;
        }
        case kNoise:
        case kNoise3D: {
            // Noise / Noise3D (node.h:325-358): Turbulence(p, (int)freq, oct),
            // the octave count converted to int by the call.
            const f3 q = st[sp - 3];
            const float freq = (float)(int32_t)st[sp - 2].x;
            const int oct = (int32_t)st[sp - 1].x;
            st[sp - 3] = in.op == kNoise3D && in.arg ? turbulence3d(q, freq, oct) : splat(turbulence(q, freq, oct));
            sp -= 2;
            break;
        }
        default: break;
        }
    }
    return st[0];
}

// Material::BumpNormal (material.cpp:49-63): finite differences of the bump
// map along u and v (step 0.01, applied to uv and to p along dpdu / dpdv), the
// perturbed tangents' cross product turned to the side of n.
PG_HD f3 bump_normal(const Inst* prog, const Ctx& c, f3 dpdu, f3 dpdv, f3* st) {
    const float delta = 0.01f;
    Ctx c0 = c, c1 = c;
    c0.u = c.u + delta;
    c1.v = c.v + delta;
    c0.p = c.p + dpdu * delta;
    c1.p = c.p + dpdv * delta;
    const float b = eval(prog, c, st).x;
    const float dddu = (eval(prog, c0, st).x - b) * (1.0f / delta);
    const float dddv = (eval(prog, c1, st).x - b) * (1.0f / delta);
    const f3 du = dpdu + dddu * c.n, dv = dpdv + dddv * c.n;
    const f3 n = normalize(cross(du, dv));
    return dot(n, c.n) < 0 ? -n : n;
}

}  // namespace nodes
}  // namespace pg
