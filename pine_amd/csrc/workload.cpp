// Deterministic benchmark workloads: a displaced geodesic icosphere and the ray
// stream of the traversal benchmark.  Only IEEE +,-,*,/ and sqrt are used (the
// library is built with -ffp-contract=off), so any x86-64 host produces the same
// bits, which lets golden hit files refer to rays by index instead of storing them.
#include "pg_api.h"

#include <algorithm>
#include <cstdio>
#include <string>
#include <cmath>
#include <thread>
#include <unordered_map>
#include <vector>

namespace pg {
namespace {

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
float unit24(uint64_t& s) { return (float)(splitmix(s) >> 40) * (1.0f / 16777216.0f); }

uint32_t hash3(int32_t x, int32_t y, int32_t z, uint64_t seed) {
    uint64_t s = seed ^ ((uint64_t)(uint32_t)x * 0x8DA6B343ULL) ^
                 ((uint64_t)(uint32_t)y * 0xD8163841ULL) ^ ((uint64_t)(uint32_t)z * 0xCB1AB31FULL);
    return (uint32_t)(splitmix(s) >> 32);
}

// Value noise in [-1, 1] with smoothstep-weighted trilinear interpolation.
float value_noise(float px, float py, float pz, uint64_t seed) {
    float fx = std::floor(px), fy = std::floor(py), fz = std::floor(pz);
    int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
    float tx = px - fx, ty = py - fy, tz = pz - fz;
    float sx = tx * tx * (3.0f - 2.0f * tx);
    float sy = ty * ty * (3.0f - 2.0f * ty);
    float sz = tz * tz * (3.0f - 2.0f * tz);
    float c[8];
    for (int k = 0; k < 8; k++) {
        uint32_t h = hash3(ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2), seed);
        c[k] = (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
    }
    float x00 = c[0] + sx * (c[1] - c[0]);
    float x10 = c[2] + sx * (c[3] - c[2]);
    float x01 = c[4] + sx * (c[5] - c[4]);
    float x11 = c[6] + sx * (c[7] - c[6]);
    float y0 = x00 + sy * (x10 - x00);
    float y1 = x01 + sy * (x11 - x01);
    return y0 + sz * (y1 - y0);
}

float fbm5(f3 p, uint64_t seed) {
    float sum = 0.0f, amp = 0.5f, freq = 3.0f, norm = 0.0f;
    for (int o = 0; o < 5; o++) {
        sum += amp * value_noise(p.x * freq, p.y * freq, p.z * freq, seed + (uint64_t)o * 7919u);
        norm += amp;
        amp *= 0.5f;
        freq *= 2.0f;
    }
    return sum / norm;
}

// fBm of value noise with `octaves` octaves, amplitudes 1/2, 1/4, ... (so the
// sum stays inside (-1, 1)) and lacunarity 2.
float fbm(float x, float y, float z, int octaves, float freq, uint64_t seed) {
    float sum = 0.0f, amp = 0.5f;
    for (int o = 0; o < octaves; o++) {
        sum += amp * value_noise(x * freq, y * freq, z * freq, seed + (uint64_t)o * 7919u);
        amp *= 0.5f;
        freq *= 2.0f;
    }
    return sum;
}

f3 normalize(f3 v) {
    float l = std::sqrt(dot(v, v));
    return v / l;
}

}  // namespace
}  // namespace pg

using namespace pg;

extern "C" int pg_gen_icosphere(int32_t freq, float amp, uint64_t seed, int64_t max_triangles,
                                float* h_vertices, int64_t* n_vertices, uint32_t* h_indices,
                                int64_t* n_triangles) {
    if (freq < 1 || freq > 2048 || !n_vertices || !n_triangles)
        return pg_fail(PG_E_ARG, "pg_gen_icosphere: freq must be 1..2048, size outputs non-null");
    const int64_t f = freq;
    int64_t nt_full = 20 * f * f;
    int64_t nt = (max_triangles > 0 && max_triangles < nt_full) ? max_triangles : nt_full;
    int64_t nv = 10 * f * f + 2;
    if (!h_vertices || !h_indices) {
        *n_vertices = nv;
        *n_triangles = nt;
        return PG_OK;
    }
    // Icosahedron: (0, +-1, +-phi) and cyclic permutations; outward-wound faces.
    const float phi = (1.0f + std::sqrt(5.0f)) / 2.0f;
    const f3 base[12] = {mk(-1, phi, 0), mk(1, phi, 0),  mk(-1, -phi, 0), mk(1, -phi, 0),
                         mk(0, -1, phi), mk(0, 1, phi),  mk(0, -1, -phi), mk(0, 1, -phi),
                         mk(phi, 0, -1), mk(phi, 0, 1),  mk(-phi, 0, -1), mk(-phi, 0, 1)};
    static const int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                     {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                     {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                     {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    std::unordered_map<uint64_t, uint32_t> map;
    map.reserve((size_t)nv * 2);
    int64_t vcount = 0;
    // A lattice point is identified by its (corner id, weight) pairs sorted by id,
    // and its position is always evaluated in that order, so points shared by
    // neighbouring faces get identical bits.
    auto vertex = [&](const int ids[3], const int64_t w[3]) -> uint32_t {
        int order[3] = {0, 1, 2};
        std::sort(order, order + 3, [&](int a, int b) { return ids[a] < ids[b]; });
        uint64_t key = 0;
        f3 acc = mk(0, 0, 0);
        for (int k = 0; k < 3; k++) {
            int o = order[k];
            if (w[o] == 0)
                continue;
            key = (key << 21) | ((uint64_t)ids[o] << 12) | (uint64_t)w[o];
            acc = acc + base[ids[o]] * (float)w[o];
        }
        auto it = map.find(key);
        if (it != map.end())
            return it->second;
        f3 n = normalize(acc / (float)f);
        float r = 1.0f + amp * fbm5(n, seed);
        f3 p = n * r;
        uint32_t id = (uint32_t)vcount++;
        h_vertices[3 * id + 0] = p.x;
        h_vertices[3 * id + 1] = p.y;
        h_vertices[3 * id + 2] = p.z;
        map.emplace(key, id);
        return id;
    };
    int64_t t = 0;
    for (int fi = 0; fi < 20 && t < nt; fi++) {
        const int* ids = faces[fi];
        auto P = [&](int64_t i, int64_t j) {
            int64_t w[3] = {f - i - j, i, j};
            return vertex(ids, w);
        };
        for (int64_t i = 0; i < f && t < nt; i++) {
            for (int64_t j = 0; j < f - i && t < nt; j++) {
                uint32_t a = P(i, j), b = P(i + 1, j), c = P(i, j + 1);
                h_indices[3 * t + 0] = a;
                h_indices[3 * t + 1] = b;
                h_indices[3 * t + 2] = c;
                t++;
                if (j < f - i - 1 && t < nt) {
                    uint32_t d = P(i + 1, j + 1);
                    h_indices[3 * t + 0] = b;
                    h_indices[3 * t + 1] = d;
                    h_indices[3 * t + 2] = c;
                    t++;
                }
            }
        }
    }
    *n_vertices = vcount;
    *n_triangles = t;
    return PG_OK;
}

extern "C" int pg_gen_rays(const float* h_vertices, int64_t n_vertices, const uint32_t* h_indices,
                           int64_t n_triangles, int32_t res, const float eye[3], float fov_deg,
                           int64_t n_bounce, uint64_t seed, pg_ray* h_rays) {
    if (!h_rays || res < 0 || n_bounce < 0 || !eye)
        return pg_fail(PG_E_ARG, "pg_gen_rays: bad arguments");
    if (n_bounce > 0 && (!h_vertices || !h_indices || n_triangles <= 0 || n_vertices <= 0))
        return pg_fail(PG_E_ARG, "pg_gen_rays: bounce rays need a mesh");
    const int64_t n_primary = (int64_t)res * res;
    const f3 e = mk(eye[0], eye[1], eye[2]);
    const f3 fwd = normalize(-e);
    f3 right = normalize(cross(fwd, mk(0, 1, 0)));
    f3 up = cross(right, fwd);
    const float tanh = (float)std::tan((double)fov_deg * 3.14159265358979323846 / 360.0);
    const int64_t total = n_primary + n_bounce;
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned th = 0; th < nth; th++) {
        pool.emplace_back([&, th]() {
            for (int64_t k = th; k < total; k += nth) {
                pg_ray r;
                if (k < n_primary) {
                    int64_t x = k % res, y = k / res;
                    float u = ((float)x + 0.5f) / (float)res * 2.0f - 1.0f;
                    float v = 1.0f - ((float)y + 0.5f) / (float)res * 2.0f;
                    f3 d = normalize(fwd + right * (u * tanh) + up * (v * tanh));
                    r.o[0] = e.x, r.o[1] = e.y, r.o[2] = e.z;
                    r.d[0] = d.x, r.d[1] = d.y, r.d[2] = d.z;
                } else {
                    uint64_t s = seed ^ ((uint64_t)(k - n_primary) * 0xD1B54A32D192ED03ULL);
                    splitmix(s);
                    int64_t tri = (int64_t)(unit24(s) * (float)n_triangles);
                    if (tri >= n_triangles)
                        tri = n_triangles - 1;
                    float a = unit24(s), b = unit24(s);
                    if (a + b > 1.0f) {
                        a = 1.0f - a;
                        b = 1.0f - b;
                    }
                    const float* V = h_vertices;
                    const uint32_t* I = h_indices + 3 * tri;
                    f3 v0 = mk(V[3 * I[0]], V[3 * I[0] + 1], V[3 * I[0] + 2]);
                    f3 v1 = mk(V[3 * I[1]], V[3 * I[1] + 1], V[3 * I[1] + 2]);
                    f3 v2 = mk(V[3 * I[2]], V[3 * I[2] + 1], V[3 * I[2] + 2]);
                    f3 p = v0 + (v1 - v0) * a + (v2 - v0) * b;
                    f3 n = normalize(cross(v1 - v0, v2 - v0));
                    if (dot(n, p) < 0.0f)
                        n = -n;
                    f3 o = offset_ray_origin(p, n);
                    // Malley's method: uniform disk point by rejection, lifted to
                    // the hemisphere, gives a cosine-distributed direction.
                    float dx, dy;
                    do {
                        dx = unit24(s) * 2.0f - 1.0f;
                        dy = unit24(s) * 2.0f - 1.0f;
                    } while (dx * dx + dy * dy >= 1.0f);
                    float dz = std::sqrt(1.0f - dx * dx - dy * dy);
                    f3 t1 = fabsp(n.x) > fabsp(n.y) ? normalize(cross(n, mk(0, 1, 0)))
                                                    : normalize(cross(n, mk(1, 0, 0)));
                    f3 t2 = cross(n, t1);
                    f3 d = normalize(t1 * dx + t2 * dy + n * dz);
                    r.o[0] = o.x, r.o[1] = o.y, r.o[2] = o.z;
                    r.d[0] = d.x, r.d[1] = d.y, r.d[2] = d.z;
                }
                r.tmin = 0.0f;
                r.tmax = kFloatMax;
                h_rays[k] = r;
            }
        });
    }
    for (auto& t : pool)
        t.join();
    return PG_OK;
}

// Heterogeneous-medium density grid: voxel (x, y, z) holds
// max(0, fbm6(centre) - threshold) * scale, x fastest, the order pine's raw
// volume files store (fileio.cpp:194-203).
extern "C" int pg_gen_density_grid(int32_t nx, int32_t ny, int32_t nz, float base_freq,
                                   float threshold, float scale, uint64_t seed, float* h_density) {
    if (nx < 1 || ny < 1 || nz < 1 || (int64_t)nx * ny * nz > (int64_t)1 << 31 || !h_density)
        return pg_fail(PG_E_ARG, "pg_gen_density_grid: bad size or null output");
    const int64_t plane = (int64_t)nx * ny;
    const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++)
        pool.emplace_back([=]() {
            for (int32_t z = t; z < nz; z += nthreads)
                for (int32_t y = 0; y < ny; y++)
                    for (int32_t x = 0; x < nx; x++) {
                        float px = ((float)x + 0.5f) / (float)nx;
                        float py = ((float)y + 0.5f) / (float)ny;
                        float pz = ((float)z + 0.5f) / (float)nz;
                        float d = fbm(px, py, pz, 6, base_freq, seed) - threshold;
                        h_density[z * plane + (int64_t)y * nx + x] = d > 0.0f ? d * scale : 0.0f;
                    }
        });
    for (auto& th : pool)
        th.join();
    return PG_OK;
}

namespace {

// Fixed-point text of x with six decimals (rounded half away from zero).
char* put_fixed6(char* p, float x) {
    double d = x;
    if (d < 0) {
        *p++ = '-';
        d = -d;
    }
    uint64_t q = (uint64_t)(d * 1e6 + 0.5);
    uint64_t ip = q / 1000000, fp = q % 1000000;
    char tmp[24];
    int n = 0;
    do {
        tmp[n++] = (char)('0' + ip % 10);
        ip /= 10;
    } while (ip);
    while (n)
        *p++ = tmp[--n];
    *p++ = '.';
    for (int k = 5; k >= 0; k--) {
        p[k] = (char)('0' + fp % 10);
        fp /= 10;
    }
    return p + 6;
}

char* put_uint(char* p, uint64_t v) {
    char tmp[24];
    int n = 0;
    do {
        tmp[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (n)
        *p++ = tmp[--n];
    return p;
}

}  // namespace

extern "C" int pg_write_obj(const char* path, const float* v, int64_t nv, const uint32_t* idx, int64_t nt) {
    if (!path || nv < 0 || nt < 0 || (nv && !v) || (nt && !idx))
        return pg_fail(PG_E_ARG, "pg_write_obj: bad arguments");
    for (int64_t k = 0; k < 3 * nt; k++)
        if (idx[k] >= (uint64_t)nv)
            return pg_fail(PG_E_ARG, "pg_write_obj: index out of range");
    for (int64_t k = 0; k < 3 * nv; k++)
        if (!(std::fabs(v[k]) < 1e12f))
            return pg_fail(PG_E_ARG, "pg_write_obj: vertex coordinate not finite or too large");
    const int nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int64_t lines = nv + nt;
    std::vector<std::string> parts(nth);
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; t++)
        pool.emplace_back([&, t]() {
            const int64_t a = lines * t / nth, b = lines * (t + 1) / nth;
            std::string& out = parts[t];
            out.resize((size_t)(b - a) * 80);  // longest line: "v" + 3 x (" " + 21 chars) + "\n"
            char* p = &out[0];
            for (int64_t l = a; l < b; l++) {
                if (l < nv) {
                    *p++ = 'v';
                    for (int c = 0; c < 3; c++) {
                        *p++ = ' ';
                        p = put_fixed6(p, v[3 * l + c]);
                    }
                } else {
                    const int64_t f = l - nv;
                    *p++ = 'f';
                    for (int c = 0; c < 3; c++) {
                        *p++ = ' ';
                        p = put_uint(p, (uint64_t)idx[3 * f + c] + 1);
                    }
                }
                *p++ = '\n';
            }
            out.resize((size_t)(p - &out[0]));
        });
    for (auto& th : pool)
        th.join();
    FILE* fp = std::fopen(path, "wb");
    if (!fp)
        return pg_fail(PG_E_ARG, std::string("pg_write_obj: cannot open ") + path);
    bool ok = true;
    for (auto& s : parts)
        ok = ok && std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? PG_OK : pg_fail(PG_E_ARG, std::string("pg_write_obj: write failed: ") + path);
}
