// Oracle harness: drives the reference pine library (built from its own sources
// by oracle/Makefile into oracle/_ref/) through its public API and writes golden
// vectors.  This file is test infrastructure, never part of the shipped path.
//
// Modes
//   trace  <scene.bin> <rays.bin> <hits.bin> [threads]
//       Builds pine's BVH (BVH::Initialize, bvh.cpp:612-650) over the shapes in
//       scene.bin and answers every ray with the real BVH::Intersect
//       (bvh.cpp:669-709) and BVH::Hit (bvh.cpp:652-667).  Pine's Interaction
//       does not expose the hit triangle's index, so the harness also replays
//       pine's traversal order (bvh.cpp:534-610) over the public node arrays
//       with pine's own inline AABB::Hit / Triangle::Intersect, and checks per
//       ray that the replay reproduces the real call's t and (u,v) bit for bit
//       before it trusts the replay's triangle index.
//   time   <scene.bin> <rays.bin> [threads]
//       Times the real BVH::Intersect over the ray file (CPU Mrays/s).
//   (time and render wait for a line on stdin before their timed region when
//   PG_HARNESS_GATE is set; see Gate().)
//   math   <out.bin>
//       pstd libm outputs (pstd/math.h:141-351) on a fixed input grid.
//   halton <out.bin> <filmW> <filmH> <spp> <ndims> <px0> <py0> <npix>
//       HaltonSampler values (sampler.cpp:32-96) for a run of pixels.
//   sobol  <out.bin> <filmW> <filmH> <spp> <ndims> <px0> <py0> <npix> <strategy>
//       SobolSampler values (sampler.cpp:129-220) for a run of pixels; strategy
//       0 none, 1 binaryPermutate, 2 fastOwen, 3 Owen.
//   atmosphere <out.bin> <sx> <sy> <sz> <r> <g> <b> <sizeX> <sizeY>
//       The Atmosphere environment light's colour table and sun colour as its
//       constructor computes them (light.cpp:81-92): sizeX*sizeY*3 floats, then 3.
//   noise  <out.bin> <n>
//       PerlinNoise (seed 0..2), Turbulence and Turbulence3D (noise.cpp:7-52)
//       at n seeded points in [-8, 8)^3 with frequencies 1..7 and 1..5 octaves:
//       per point x, y, z, frequency, octaves, perlin0..2, turbulence, t3d xyz.
//   render <scene.txt> <out.bin> [threads] [sample0] [atmosphere.bin]
//       With atmosphere.bin (sizeX*sizeY*3 floats, then 3), every copy pine
//       holds of the scene's Atmosphere light (scene->envLight, scene->lights,
//       the integrator's light sampler) gets that colour table and sun colour
//       in place of its own before rendering.
//       Loads a scene with pine's own LoadScene (fileio.cpp:251-281) and renders
//       it with the scene's integrator (PathIntegrator::Li, path.cpp:6-87) in
//       pine's per-pixel sample order (integrator.cpp:121-158), accumulating
//       each sample with the film's filter-table weight (film.h:26-45) into a
//       linear float image; pine itself only writes tone-mapped 8-bit BMPs.
//       out.bin: i32 w, i32 h, then h*w*4 floats (sum w*r, sum w*g, sum w*b,
//       sum w) in pine's pixel order (row 0 = top, film.h:60-62).
//   finalize <film.bin> <out.bmp> <toneMapping 0|1>
//       pine's own image output for a film of render's layout: a Film with
//       applyToneMapping as given whose pixels (Film::GetPixel, film.h:59-61)
//       hold the film's sums, then Film::Finalize(1) (film.cpp:45-100: divide
//       by the weight, Uncharted2 filmic curve, gamma 1/2.2 with pstd::pow,
//       SaveImage -> WriteImageBMP, fileio.cpp:88-184).
//
// File formats (little endian):
//   scene.bin: u32 'PGS1', u32 nshapes, then per shape u32 type and payload:
//       0 mesh   : u32 nv, u32 nt, f32 v[3nv], u32 idx[3nt]
//       1 sphere : f32 c[3], f32 r
//       2 rect   : f32 position[3], ex[3], ey[3]
//       3 plane  : f32 position[3], normal[3]
//   rays.bin : u64 n, then n x {f32 o[3], tmin, d[3], tmax}
//   hits.bin : n x {i32 shape, i32 prim, f32 t, f32 u, f32 v, i32 occluded}
#include <core/scene.h>
#include <core/integrator.h>
#include <core/filter.h>
#include <core/film.h>
#include <util/fileio.h>
#include <core/sampler.h>
#include <core/noise.h>
#include <core/lowdiscrepancy.h>
#include <impl/accel/bvh.h>
#include <util/parameters.h>

#include <atomic>
#include <cmath>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

using namespace pine;

namespace {

struct RayRec {
    float o[3], tmin, d[3], tmax;
};
struct HitRec {
    int32_t shape, prim;
    float t, u, v;
    int32_t occluded;
};

std::vector<unsigned char> ReadAll(const char* path) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(2);
    }
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<unsigned char> buf(n);
    if (n && fread(buf.data(), 1, n, f) != (size_t)n) {
        fprintf(stderr, "short read %s\n", path);
        exit(2);
    }
    fclose(f);
    return buf;
}

struct Reader {
    const unsigned char* p;
    template <typename T>
    T Get() {
        T v;
        memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    void GetN(void* dst, size_t bytes) {
        memcpy(dst, p, bytes);
        p += bytes;
    }
};

vec3 V3(Reader& r) {
    vec3 v;
    v.x = r.Get<float>();
    v.y = r.Get<float>();
    v.z = r.Get<float>();
    return v;
}

void LoadSceneBin(const char* path, Scene& scene) {
    auto buf = ReadAll(path);
    Reader r{buf.data()};
    uint32_t magic = r.Get<uint32_t>();
    if (magic != 0x31534750u) {
        fprintf(stderr, "bad scene magic\n");
        exit(2);
    }
    uint32_t n = r.Get<uint32_t>();
    for (uint32_t i = 0; i < n; i++) {
        uint32_t type = r.Get<uint32_t>();
        Shape shape;
        if (type == 0) {
            uint32_t nv = r.Get<uint32_t>(), nt = r.Get<uint32_t>();
            pstd::vector<vec3> verts(nv);
            pstd::vector<uint32_t> idx(nt * 3);
            for (uint32_t k = 0; k < nv; k++)
                verts[k] = V3(r);
            r.GetN(&idx[0], sizeof(uint32_t) * idx.size());
            shape = TriangleMesh(pstd::move(verts), pstd::move(idx));
        } else if (type == 1) {
            vec3 c = V3(r);
            float rad = r.Get<float>();
            shape = Sphere(c, rad);
        } else if (type == 2) {
            vec3 p = V3(r), ex = V3(r), ey = V3(r);
            shape = Rect(p, ex, ey);
        } else if (type == 3) {
            vec3 p = V3(r), nn = V3(r);
            shape = Plane(p, Normalize(nn));
        } else {
            fprintf(stderr, "bad shape type %u\n", type);
            exit(2);
        }
        // Same as CreateShape (geometry.cpp:384).
        shape.aabb = shape.Dispatch([](auto&& x) { return x.GetAABB(); });
        scene.shapes.push_back(shape);
    }
}

std::vector<RayRec> LoadRays(const char* path) {
    auto buf = ReadAll(path);
    Reader r{buf.data()};
    uint64_t n = r.Get<uint64_t>();
    std::vector<RayRec> rays(n);
    r.GetN(rays.data(), n * sizeof(RayRec));
    return rays;
}

Ray ToRay(const RayRec& rr) {
    return Ray(vec3(rr.o[0], rr.o[1], rr.o[2]), vec3(rr.d[0], rr.d[1], rr.d[2]), rr.tmin, rr.tmax);
}

// Replays BVHImpl::Intersect (bvh.cpp:534-610) on a BVHImpl's public nodes, with
// `leaf(ray, index)` standing in for the closure pine passes.  Returns the index
// handed to the last accepted leaf call, or -1.
template <typename Leaf>
int ReplayIntersect(const BVHImpl& b, Ray& ray, Leaf&& leaf) {
    RayOctant oct(ray);
    const BVHImpl::Node* nodes = b.nodes.data();
    int closest = -1;
    int stack[64];
    int ptr = 0;
    int next = b.rootIndex;
    if (nodes[next].primitiveIndices.size()) {
        for (int index : nodes[next].primitiveIndices)
            if (leaf(ray, index))
                closest = index;
        return closest;
    }
    while (true) {
        const BVHImpl::Node& node = nodes[next];
        int l = -1, rr = -1;
        float t0 = ray.tmax, t1 = ray.tmax;
        if (node.aabbs[0].Hit(oct, ray.tmin, &t0)) {
            const BVHImpl::Node& c = nodes[node.children[0]];
            if (!c.primitiveIndices.size())
                l = node.children[0];
            else
                for (int index : c.primitiveIndices)
                    if (leaf(ray, index))
                        closest = index;
        }
        if (node.aabbs[1].Hit(oct, ray.tmin, &t1)) {
            const BVHImpl::Node& c = nodes[node.children[1]];
            if (!c.primitiveIndices.size())
                rr = node.children[1];
            else
                for (int index : c.primitiveIndices)
                    if (leaf(ray, index))
                        closest = index;
        }
        if (l != -1) {
            if (rr != -1) {
                if (t0 > t1) {
                    stack[ptr++] = l;
                    next = rr;
                } else {
                    stack[ptr++] = rr;
                    next = l;
                }
            } else {
                next = l;
            }
        } else if (rr != -1) {
            next = rr;
        } else {
            if (ptr == 0)
                break;
            next = stack[--ptr];
        }
    }
    return closest;
}

int Threads(int argc, char** argv, int pos) {
    if (argc > pos)
        return atoi(argv[pos]);
    return (int)std::thread::hardware_concurrency();
}

template <typename F>
void ParallelRange(int64_t n, int threads, F&& f) {
    std::atomic<int64_t> next{0};
    const int64_t chunk = 256;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&]() {
            while (true) {
                int64_t b = next.fetch_add(chunk);
                if (b >= n)
                    break;
                int64_t e = b + chunk < n ? b + chunk : n;
                for (int64_t i = b; i < e; i++)
                    f(i);
            }
        });
    for (auto& th : pool)
        th.join();
}

// With PG_HARNESS_GATE set in the environment, the timed modes print "ready"
// once their scene and BVH are built and wait for a line on stdin before the
// timed region, so a driver can build several in parallel and time them one
// at a time.
void Gate() {
    if (!getenv("PG_HARNESS_GATE"))
        return;
    printf("ready\n");
    fflush(stdout);
    char buf[16];
    if (!fgets(buf, sizeof(buf), stdin))
        exit(5);
}

double Now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

int ModeTrace(int argc, char** argv) {
    if (argc < 5)
        return 2;
    Scene scene;
    LoadSceneBin(argv[2], scene);
    auto rays = LoadRays(argv[3]);
    int threads = Threads(argc, argv, 5);

    BVH bvh{Parameters()};
    double t0 = Now();
    bvh.Initialize(&scene);
    double buildMs = (Now() - t0) * 1e3;

    std::vector<HitRec> hits(rays.size());
    std::atomic<int64_t> replayMismatch{0};
    ParallelRange((int64_t)rays.size(), threads, [&](int64_t i) {
        HitRec h{-1, -1, 0.0f, 0.0f, 0.0f, 0};
        Ray ray = ToRay(rays[i]);
        Interaction it;
        bool hit = bvh.Intersect(ray, it);
        h.occluded = bvh.Hit(ToRay(rays[i])) ? 1 : 0;
        if (hit) {
            h.shape = int32_t(it.shape - &scene.shapes[0]);
            h.t = ray.tmax;
            h.u = it.uv.x;
            h.v = it.uv.y;
            // Replay to recover the triangle index (pine keeps it local).
            Ray r2 = ToRay(rays[i]);
            Interaction it2;
            int triIndex = -1;
            int lbvhIndex = ReplayIntersect(bvh.tbvh, r2, [&](Ray& ray2, int li) {
                const Shape& shape = scene.shapes[bvh.indices[li]];
                if (li < (int)bvh.lbvh.size()) {
                    const TriangleMesh& mesh = shape.Be<TriangleMesh>();
                    int found = ReplayIntersect(bvh.lbvh[li], ray2, [&](Ray& ray3, int ti) {
                        return mesh.GetTriangle(ti).Intersect(ray3, it2);
                    });
                    if (found != -1)
                        triIndex = found;
                    return found != -1;
                }
                return shape.Intersect(ray2, it2);
            });
            int shapeIdx = lbvhIndex >= 0 ? bvh.indices[lbvhIndex] : -1;
            bool isMesh = lbvhIndex >= 0 && lbvhIndex < (int)bvh.lbvh.size();
            h.prim = isMesh ? triIndex : 0;
            if (shapeIdx != h.shape || memcmp(&r2.tmax, &ray.tmax, 4) != 0 ||
                memcmp(&it2.uv.x, &it.uv.x, 4) != 0 || memcmp(&it2.uv.y, &it.uv.y, 4) != 0)
                replayMismatch++;
        }
        hits[i] = h;
    });

    FILE* f = fopen(argv[4], "wb");
    fwrite(hits.data(), sizeof(HitRec), hits.size(), f);
    fclose(f);
    printf("{\"build_ms\": %.3f, \"rays\": %zu, \"replay_mismatch\": %lld, \"threads\": %d}\n",
           buildMs, rays.size(), (long long)replayMismatch.load(), threads);
    return replayMismatch.load() == 0 ? 0 : 3;
}

int ModeTime(int argc, char** argv) {
    if (argc < 4)
        return 2;
    Scene scene;
    LoadSceneBin(argv[2], scene);
    auto rays = LoadRays(argv[3]);
    int threads = Threads(argc, argv, 4);
    BVH bvh{Parameters()};
    double t0 = Now();
    bvh.Initialize(&scene);
    double buildMs = (Now() - t0) * 1e3;
    std::atomic<int64_t> nhit{0};
    Gate();
    double t1 = Now();
    ParallelRange((int64_t)rays.size(), threads, [&](int64_t i) {
        Ray ray = ToRay(rays[i]);
        Interaction it;
        if (bvh.Intersect(ray, it))
            nhit++;
    });
    double secs = Now() - t1;
    printf("{\"build_ms\": %.3f, \"rays\": %zu, \"hits\": %lld, \"trace_s\": %.6f, "
           "\"mrays_per_s\": %.4f, \"threads\": %d}\n",
           buildMs, rays.size(), (long long)nhit.load(), secs, rays.size() / secs * 1e-6, threads);
    return 0;
}

// pstd libm called through out-of-line wrappers on inputs the compiler cannot
// see, so each function is compiled as a standalone call: a constant-folded
// loop would let -ffast-math merge the input arithmetic into the function.
__attribute__((noinline)) float WSqrt(float x) { return pstd::sqrt(x); }
__attribute__((noinline)) float WExp(float x) { return pstd::exp(x); }
__attribute__((noinline)) float WLog(float x) { return pstd::log(x); }
__attribute__((noinline)) float WCos(float x) { return pstd::cos(x); }
__attribute__((noinline)) float WSin(float x) { return pstd::sin(x); }
__attribute__((noinline)) float WAcos(float x) { return pstd::acos(x); }
__attribute__((noinline)) float WPow(float x, float e) { return pstd::pow(x, e); }
__attribute__((noinline)) float WAtan2(float y, float x) { return pstd::atan2(y, x); }

int ModeMath(int argc, char** argv) {
    if (argc < 3)
        return 2;
    const int N = 4096;
    std::vector<float> in(N);
    volatile float sink = 0.0f;
    std::vector<float> out;
    auto run = [&](float (*fn)(float)) {
        for (int i = 0; i < N; i++) {
            float x = in[i];
            asm volatile("" : "+x"(x));  // opaque input
            float y = fn(x);
            sink = y;
            out.push_back(in[i]);
            out.push_back(y);
        }
    };
    // Layout: for each function, N (x, f(x)) pairs, in this order.
    for (int i = 0; i < N; i++)  // sqrt on [1e-6, 1e4] log-spaced
        in[i] = 1e-6f * (float)std::pow(1e10, (i + 0.5) / N);
    run(WSqrt);
    for (int i = 0; i < N; i++)  // exp on [-6, 3]
        in[i] = -6.0f + 9.0f * (i + 0.5f) / N;
    run(WExp);
    for (int i = 0; i < N; i++)  // log on (0, 1]
        in[i] = (i + 0.5f) / N;
    run(WLog);
    for (int i = 0; i < N; i++)  // cos on [-10, 10]
        in[i] = -10.0f + 20.0f * (i + 0.5f) / N;
    run(WCos);
    run(WSin);  // sin on the same grid
    for (int i = 0; i < N; i++)  // acos on [-1, 1]
        in[i] = -1.0f + 2.0f * (i + 0.5f) / N;
    run(WAcos);
    volatile float e5 = 5.0f, eg = 1.0f / 2.2f;
    for (int i = 0; i < N; i++) {  // pow(x, 5) on [0, 1]  (FrSchlick)
        float x = (i + 0.5f) / N;
        asm volatile("" : "+x"(x));
        out.push_back(x);
        out.push_back(WPow(x, e5));
    }
    for (int i = 0; i < N; i++) {  // pow(x, 1/2.2) on [0, 4] (gamma)
        float x = 4.0f * (i + 0.5f) / N;
        asm volatile("" : "+x"(x));
        out.push_back(x);
        out.push_back(WPow(x, eg));
    }
    for (int i = 0; i < N; i++) {  // atan2 on points of the unit circle
        float a = -3.1f + 6.2f * (i + 0.5f) / N;
        float y = (float)std::sin((double)a), x = (float)std::cos((double)a);
        asm volatile("" : "+x"(x), "+x"(y));
        out.push_back(y);
        out.push_back(x);
        out.push_back(WAtan2(y, x));
    }
    (void)sink;
    FILE* f = fopen(argv[2], "wb");
    fwrite(out.data(), sizeof(float), out.size(), f);
    fclose(f);
    printf("{\"floats\": %zu}\n", out.size());
    return 0;
}

int ModeHalton(int argc, char** argv) {
    if (argc < 10)
        return 2;
    vec2i film(atoi(argv[3]), atoi(argv[4]));
    int spp = atoi(argv[5]), ndims = atoi(argv[6]);
    int px0 = atoi(argv[7]), py0 = atoi(argv[8]), npix = atoi(argv[9]);
    HaltonSampler s(spp, film);
    std::vector<float> out;
    for (int k = 0; k < npix; k++) {
        int idx = py0 * film.x + px0 + k * 977;  // a spread of pixels
        vec2i p(idx % film.x, (idx / film.x) % film.y);
        s.StartPixel(p, 0);
        for (int si = 0; si < spp; si++) {
            for (int d = 0; d < ndims; d++)
                out.push_back(s.Get1D());
            s.StartNextSample();
        }
    }
    FILE* f = fopen(argv[2], "wb");
    fwrite(out.data(), sizeof(float), out.size(), f);
    fclose(f);
    printf("{\"floats\": %zu, \"stride\": %d}\n", out.size(), s.sampleStride);
    return 0;
}

int ModeSobol(int argc, char** argv) {
    if (argc < 11)
        return 2;
    vec2i film(atoi(argv[3]), atoi(argv[4]));
    int spp = atoi(argv[5]), ndims = atoi(argv[6]);
    int px0 = atoi(argv[7]), py0 = atoi(argv[8]), npix = atoi(argv[9]);
    const SobolSampler::RandomizeStrategy strategies[4] = {
        SobolSampler::RandomizeStrategy::None, SobolSampler::RandomizeStrategy::BinaryPermutate,
        SobolSampler::RandomizeStrategy::FastOwen, SobolSampler::RandomizeStrategy::Owen};
    SobolSampler s(spp, film, strategies[atoi(argv[10]) & 3]);
    std::vector<float> out;
    for (int k = 0; k < npix; k++) {
        int idx = py0 * film.x + px0 + k * 977;  // a spread of pixels
        vec2i p(idx % film.x, (idx / film.x) % film.y);
        s.StartPixel(p, 0);
        for (int si = 0; si < spp; si++) {
            for (int d = 0; d < ndims; d++)
                out.push_back(s.Get1D());
            s.StartNextSample();
        }
    }
    FILE* f = fopen(argv[2], "wb");
    fwrite(out.data(), sizeof(float), out.size(), f);
    fclose(f);
    printf("{\"floats\": %zu, \"log2scale\": %d}\n", out.size(), (int)s.log2Scale);
    return 0;
}

int ModeNoise(int argc, char** argv) {
    if (argc < 4)
        return 2;
    const int n = atoi(argv[3]);
    RNG rng(12345);
    std::vector<float> out;
    for (int i = 0; i < n; i++) {
        vec3 p = (rng.Uniform3f() * 2.0f - vec3(1.0f)) * 8.0f;
        const float freq = (float)(1 + i % 7);
        const int oct = 1 + (i / 7) % 5;
        vec3 t3 = Turbulence3D(p, freq, oct);
        const float vals[12] = {p.x, p.y, p.z, freq, (float)oct, PerlinNoise(p, freq, 0),
                                PerlinNoise(p, freq, 1), PerlinNoise(p, freq, 2), Turbulence(p, freq, oct),
                                t3.x, t3.y, t3.z};
        out.insert(out.end(), vals, vals + 12);
    }
    FILE* f = fopen(argv[2], "wb");
    fwrite(out.data(), sizeof(float), out.size(), f);
    fclose(f);
    printf("{\"floats\": %zu}\n", out.size());
    return 0;
}

int ModeAtmosphere(int argc, char** argv) {
    if (argc < 11)
        return 2;
    vec3 sun(atof(argv[3]), atof(argv[4]), atof(argv[5]));
    vec3 col(atof(argv[6]), atof(argv[7]), atof(argv[8]));
    vec2i size(atoi(argv[9]), atoi(argv[10]));
    Atmosphere a(sun, Spectrum(col), size, true);
    std::vector<float> out;
    for (const vec3& c : a.colors)
        for (int j = 0; j < 3; j++)
            out.push_back(c[j]);
    vec3 sc = a.sunSampledColor.ToRGB();
    for (int j = 0; j < 3; j++)
        out.push_back(sc[j]);
    FILE* f = fopen(argv[2], "wb");
    fwrite(out.data(), sizeof(float), out.size(), f);
    fclose(f);
    printf("{\"floats\": %zu}\n", out.size());
    return 0;
}

int ModeRender(int argc, char** argv) {
    if (argc < 4)
        return 2;
    int threads = Threads(argc, argv, 4);
    // Optional first sample index per pixel (disjoint sample ranges).
    const int64_t sample0 = argc > 5 ? atoll(argv[5]) : 0;
    auto scene = pstd::make_shared<Scene>();
    Parameters params = LoadScene(argv[2], scene.get());
    auto* ri = dynamic_cast<RadianceIntegrator*>(scene->integrator.get());
    if (!ri) {
        fprintf(stderr, "integrator is not a RadianceIntegrator\n");
        return 3;
    }
    if (argc > 6) {
        std::vector<float> tab;
        FILE* tf = fopen(argv[6], "rb");
        if (!tf)
            return 4;
        float v;
        while (fread(&v, 4, 1, tf) == 1)
            tab.push_back(v);
        fclose(tf);
        auto patch = [&](Atmosphere& a) {
            if (tab.size() != a.colors.size() * 3 + 3) {
                fprintf(stderr, "atmosphere table size mismatch\n");
                exit(5);
            }
            for (size_t i = 0; i < a.colors.size(); i++)
                a.colors[i] = vec3(tab[3 * i], tab[3 * i + 1], tab[3 * i + 2]);
            size_t n = a.colors.size() * 3;
            a.sunSampledColor = Spectrum(vec3(tab[n], tab[n + 1], tab[n + 2]));
        };
        auto patch_lights = [&](pstd::vector<Light>& ls) {
            for (auto& l : ls)
                if (l.Is<EnvironmentLight>() && l.Be<EnvironmentLight>().Is<Atmosphere>())
                    patch(l.Be<EnvironmentLight>().Be<Atmosphere>());
        };
        if (scene->envLight && scene->envLight->Is<Atmosphere>())
            patch(scene->envLight->Be<Atmosphere>());
        patch_lights(scene->lights);
        if (ri->lightSampler.Is<PowerLightSampler>())
            patch_lights(ri->lightSampler.Be<PowerLightSampler>().lights);
        else
            patch_lights(ri->lightSampler.Be<UniformLightSampler>().lights);
    }
    Film& film = scene->camera.GetFilm();
    vec2i size = film.Size();
    int spp = ri->samplesPerPixel;
    // Filter table exactly as Film's constructor builds it (film.cpp:22-28).
    Filter filter = CreateFilter(params["Camera"]["film"]["filter"]);
    const int TW = 16;
    float table[TW * TW];
    for (int y = 0, o = 0; y < TW; y++)
        for (int x = 0; x < TW; x++)
            table[o++] = filter.Evaluate(
                {(x + 0.5f) / TW * filter.Radius().x, (y + 0.5f) / TW * filter.Radius().y});
    std::vector<float> acc((size_t)size.x * size.y * 4, 0.0f);
    std::vector<std::mutex> locks(1024);
    Gate();
    double t0 = Now();
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int th = 0; th < threads; th++)
        pool.emplace_back([&]() {
            Sampler sampler = ri->samplers[0].Clone();
            while (true) {
                int64_t index = next.fetch_add(1);
                if (index >= (int64_t)size.x * size.y)
                    break;
                vec2i p = {int(index % size.x), int(index / size.x)};
                sampler.StartPixel(p, sample0);
                for (int s = 0; s < spp; s++) {
                    // RadianceIntegrator::Compute (integrator.cpp:152-158)
                    vec2 pFilm = (p + sampler.Get2D()) / film.Size();
                    Ray ray = scene->camera.GenRay(pFilm, sampler.Get2D());
                    Spectrum L = ri->Li(ray, sampler);
                    // Film::AddSample (film.h:26-45).  The barrier keeps g++ from
                    // folding (x / size) * size back to x here; pine's own Compute
                    // keeps the divide and the multiply (integrator.o).
                    asm volatile("" : "+m"(pFilm));
                    vec2 pf = pFilm * size;
                    pf -= vec2(0.5f);
                    vec2i p0 = Ceil(pf - filter.Radius());
                    vec2i p1 = Floor(pf + filter.Radius());
                    p0 = Max(p0, vec2i(0));
                    p1 = Min(p1, size - vec2i(1));
                    vec3 rgb = L.ToRGB();
                    for (int y = p0.y; y <= p1.y; y++)
                        for (int x = p0.x; x <= p1.x; x++) {
                            vec2 d = vec2(x, y) - pf;
                            vec2i pi = TW * Min(Abs(d) / filter.Radius(), vec2(OneMinusEpsilon));
                            float w = table[pi.y * TW + pi.x];
                            size_t pix = (size_t)(size.y - 1 - y) * size.x + x;
                            std::lock_guard<std::mutex> lk(locks[pix % locks.size()]);
                            acc[pix * 4 + 0] += rgb[0] * w;
                            acc[pix * 4 + 1] += rgb[1] * w;
                            acc[pix * 4 + 2] += rgb[2] * w;
                            acc[pix * 4 + 3] += w;
                        }
                    sampler.StartNextSample();
                }
            }
        });
    for (auto& t : pool)
        t.join();
    double secs = Now() - t0;
    FILE* f = fopen(argv[3], "wb");
    int32_t hdr[2] = {size.x, size.y};
    fwrite(hdr, 4, 2, f);
    fwrite(acc.data(), sizeof(float), acc.size(), f);
    fclose(f);
    float lightPdf = ri->lightSampler.SampleLight(0.5f).pdf;
    printf("{\"w\": %d, \"h\": %d, \"spp\": %d, \"render_s\": %.3f, \"msamples_per_s\": %.4f, "
           "\"threads\": %d, \"light_select_pdf\": %.9g, \"n_lights\": %zu}\n",
           size.x, size.y, spp, secs, (double)size.x * size.y * spp / secs * 1e-6, threads,
           lightPdf, scene->lights.size());
    return 0;
}

int ModeFinalize(int argc, char** argv) {
    if (argc < 5)
        return 2;
    FILE* f = fopen(argv[2], "rb");
    if (!f)
        return 1;
    int32_t hdr[2];
    if (fread(hdr, 4, 2, f) != 2)
        return 1;
    const vec2i size(hdr[0], hdr[1]);
    std::vector<float> acc((size_t)size.x * size.y * 4);
    if (fread(acc.data(), sizeof(float), acc.size(), f) != acc.size())
        return 1;
    fclose(f);
    Film film(size, CreateFilter(Parameters()), argv[3], atoi(argv[4]) != 0, false);
    film.Clear();
    for (int row = 0; row < size.y; row++)
        for (int x = 0; x < size.x; x++) {
            // Row 0 of the film file is the top: GetPixel's y = size.y - 1 - row.
            const float* a = &acc[((size_t)row * size.x + x) * 4];
            Pixel& px = film.GetPixel(vec2i(x, size.y - 1 - row));
            px.rgb[0] = a[0];
            px.rgb[1] = a[1];
            px.rgb[2] = a[2];
            px.weight = a[3];
        }
    film.Finalize(1.0f);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: ref_harness trace|time|math|halton|sobol|atmosphere|render ...\n");
        return 2;
    }
    std::string mode = argv[1];
    if (mode == "trace")
        return ModeTrace(argc, argv);
    if (mode == "time")
        return ModeTime(argc, argv);
    if (mode == "math")
        return ModeMath(argc, argv);
    if (mode == "halton")
        return ModeHalton(argc, argv);
    if (mode == "sobol")
        return ModeSobol(argc, argv);
    if (mode == "atmosphere")
        return ModeAtmosphere(argc, argv);
    if (mode == "noise")
        return ModeNoise(argc, argv);
    if (mode == "render")
        return ModeRender(argc, argv);
    if (mode == "finalize")
        return ModeFinalize(argc, argv);
    fprintf(stderr, "unknown mode\n");
    return 2;
}
