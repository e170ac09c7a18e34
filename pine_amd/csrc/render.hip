// Wavefront path tracer behind pine's `Integrator: Path` (PathIntegrator::Li,
// src/impl/integrator/path.cpp:6-87; PixelIntegrator::Render and
// RadianceIntegrator::Compute, src/core/integrator.cpp:121-158) plus the scene
// loader behind pine's LoadScene (src/util/fileio.cpp:251-281).
//
// One batch = every pixel times `spb` samples, one path per lane.  Each bounce
// is a sequence of kernels over compacted queues in HBM:
//   extend   closest hit for the active rays            (accel kernels)
//   shade    emission + MIS, light sample -> shadow ray, BSDF sample -> next
//            ray, Russian roulette; survivors appended to the next queue
//   shadow   occlusion of the shadow rays               (accel kernels)
//   resolve  adds the unoccluded light contributions to the paths' radiance
// and after the last bounce `film` adds each path to its pixel with pine's
// filter-table weight, sample by sample, so the film sum is deterministic.
// Dimension bookkeeping of the Halton sampler follows pine's call order, so a
// path sees exactly the sample values pine's would.
#include "bvh_build.h"
#include "pg_api.h"
#include "pg_guide.h"
#include "pg_isect.h"
#include "pg_medium.h"
#include "pg_nodes.h"
#include "pg_spectral.h"
#include "pg_shade.h"
#include "pg_sobol_dirnums.h"
#include "scene_parse.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

using namespace pg;

#define PG_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return pg_fail(PG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

namespace pg {

enum ShapeType : int32_t { kShapeMesh = 0, kShapeSphere = 1, kShapeRect = 2, kShapePlane = 3 };

struct ShapeDev {
    int32_t type;
    int32_t material;   // index into materials, -1 = none
    int32_t first_prim; // global prim id of the shape's first prim
    int32_t tri_offset; // meshes: first triangle in tri_verts
    f3 p0, e0, e1, n;   // sphere: p0 = c; rect: p0 = pos, e0/e1 = ex/ey unit, n; plane: p0, n, e0 = u, e1 = v
    float lx, ly, r, area;
    int32_t mi_inside, mi_outside;  // MediumInterface (geometry.cpp:392-409), -1 = none
};

enum LightType : int32_t {
    kLightArea = 0, kLightPoint = 1, kLightDirectional = 2, kLightSky = 3, kLightAtmosphere = 4,
    kLightConstantEnv = 5  // this library's extension: radiance `color` from every direction
};
enum EnvKind : int32_t { kEnvSky = 0, kEnvAtmosphere = 1, kEnvConstant = 2 };
// VizIntegrator::Type (viz.h:14)
enum VizType : int32_t { kVizPosition = 0, kVizNormal = 1, kVizTexcoord = 2, kVizDpDu = 3, kVizDpDv = 4, kVizBvh = 5 };
struct LightDev {
    int32_t type;
    int32_t shape;
    f3 v;      // point: position; directional: unit direction
    f3 color;
};

struct CameraDev {
    f3 cx, cy, cz, cw;  // LookAt columns (vecmath.h:1212-1221)
    float fov2d[2];
    float lensRadius, focalDistance;
};

struct SceneDev {
    const int32_t* prim_shape;
    const float* tri_verts;  // 9 floats per mesh triangle (v0, v1, v2)
    const ShapeDev* shapes;
    const MaterialDev* materials;
    const LightDev* lights;
    const float* light_func;  // Distribution1D func (lightsampler.cpp:20-25)
    const float* light_cdf;   // n+1 entries
    int32_t n_lights;
    int32_t light_uniform;    // UniformLightSampler
    float light_func_int;
    CameraDev cam;
    SamplerDev sampler;
    float filter_table[256];
    float filter_radius[2];
    int32_t width, height;
    int32_t max_depth;
    // Media (medium.cpp) and the environment light (scene->envLight).
    const MediumDev* media;
    int32_t n_media;
    int32_t camera_medium;  // ThinLenCamera medium (camera.cpp:98-105), -1 = none
    int32_t has_env;        // an environment light is present (scene->envLight)
    int32_t env_kind;       // EnvKind
    f3 sky_dir, sky_color;  // sunDirection (normalised) and sunColor
    // Atmosphere (light.cpp:81-139): colour table [size.y][size.x] of RGB, the
    // colour along the sun, bilinear lookups when `interpolate`.
    const float* atmo_colors;
    int32_t atmo_w, atmo_h, atmo_interp;
    f3 atmo_sun_le;
    // Learned path guiding (general mode only).
    int32_t guided;
    int32_t g_depth;  // vertices of depth < g_depth sample from the guide
    int32_t rr_depth;  // Russian roulette after this depth (pine's: 2, path.cpp:76)
    spec::Tables spec;  // spectral mode: RGB <-> spectrum tables
    int32_t viz_type;   // Viz integrator: VizType
    const nodes::Inst* node_prog;  // material node-graph programs (MaterialDev::has_nodes)
    const float* g_params;  // FP32 parameters (biases used in FP32)
    const __bf16* g_wf;     // BF16 forward weights
    f3 g_bmin, g_binv;      // scene bounds: network position input = (p - bmin) * binv
    int32_t g_oct;          // encoding octave offset
};

// Path state, structure of arrays (one entry per path of the batch).
struct PathsDev {
    float* L;       // 3 per path
    float* beta;    // 3 per path
    float* bsdf_pdf;
    uint64_t* sidx; // sampler index (Halton or Sobol)
    int32_t* sdim;  // next sampler dimension
    int32_t* depth;
    float* pfilm;   // 2 per path
    // queues
    int32_t* queue_in;
    int32_t* queue_out;
    pg_ray* rays_in;
    pg_ray* rays_out;
    pg_hit* hits;
    pg_hit_uv* uvs;
    int32_t* counters;  // [0] next-queue length, [1] shadow-queue length, [2] direction queue,
                        // [3] next shadow segments
    pg_ray* shadow_rays;
    int32_t* shadow_path;
    float* shadow_contrib;  // 3 per shadow ray
    // Spectral mode: hero-wavelength variable per path and 4-wavelength
    // radiance, throughput and shadow contributions.
    float* lam;
    float* L4;
    float* beta4;
    float* sh4;
    int32_t* occluded;
    // Film contributions a sample makes to a neighbouring pixel (a sample on a
    // pixel border): (pixel, w*r, w*g, w*b, w), applied after the main pass.
    int2* border;     // per pixel: first and last of its samples that also feed a neighbour, -1 if none (k_film)
    // General mode (media or material-less shapes): the bounce is split at the
    // shadow ray, because GridMedium::Tr consumes sampler dimensions between
    // the light sample and the BSDF / phase sample.
    int32_t* medium;       // medium of the path's current ray, -1 = none
    float* vtx;            // 9 per path: vertex p, n, wi
    int32_t* vinfo;        // 4 per path: kind (0 surface, 1 medium), material, medium in, out
    int32_t* queue_b;      // paths that sample a direction after their shadow ray
    float* sh_rec;         // 16 per shadow record: f.xyz, w, Le.xyz, pdf, p2.xyz, delta, wo.xyz, p_bsdf
    float* sh_tr;          // 3 per shadow record: transmittance so far
    pg_ray* seg_rays[2];   // shadow segments (IntersectTr loop), ping-pong
    int32_t* seg_rec[2];
    int32_t* seg_med[2];
    pg_hit* seg_hits;
    pg_hit_uv* seg_uvs;
    unsigned long long* lookups;  // density lookups of the tracking kernels (8 taps each)
    float4* mev;                  // 2 per queue slot: Medium::Sample's event (weight, scattered), (p, -)
    float* seg_trf;               // per shadow segment: its medium's transmittance (lane tracking)
    // Guiding: mixture of each path's current vertex, and training records
    // (slot = path * rec_depth + depth; rec_pdf 0 = empty slot).
    float* gparams;   // guide::kParamStride per vertex: per path (split bounce) or per queue slot (fused)
    int32_t* gflag;   // 1 if the vertex samples from the guide (same indexing)
    float* gx;        // fused bounce: network input of queue slot k (9 floats)
    int32_t* shadow_rec;  // fused bounce: record slot of each shadow ray's vertex, -1 if none
    float* rec_x;     // 9 per slot: network input
    float* rec_dir;   // 3 per slot: sampled direction
    float* rec_w;     // radiance / pdf, filled after the batch
    float* rec_pb;    // BSDF pdf of the direction
    float* rec_pdf;   // mixture pdf it was drawn with
    float* rec_lb;    // 3 per slot: path radiance before the direction
    float* rec_beta;  // 3 per slot: throughput after the direction
    float* rec_acc;   // 1: mean nonzero weight of a batch
    // Nonzero-weight records compacted in slot order (the training set of a
    // batch) and the per-256-slot counts, sums and offsets that place them.
    float* c_x;
    float* c_dir;
    float* c_w;
    float* c_pb;
    int32_t* rec_bcnt;
    float* rec_bsum;
    int32_t* rec_boff;
    int32_t* rec_n;
    int32_t rec_depth;
    // Scenes with material node graphs: each shading slot's copy of its
    // material with the node-graph inputs evaluated, and the interpreter's
    // value stack (nodes::kStack entries per slot).
    MaterialDev* node_mat;
    f3* node_stack;
};

}  // namespace pg

// ----------------------------------------------------------------------------
// Device code

namespace {

// ThinLenCamera::GenRay (camera.cpp:31-41) for pFilm in [0,1]^2.
__device__ __forceinline__ void gen_ray(const CameraDev& c, float fx, float fy, float u0, float u1, f3& o, f3& d) {
    float sx = (fx - 0.5f) * 2.0f * c.fov2d[0];
    float sy = (fy - 0.5f) * 2.0f * c.fov2d[1];
    f3 dir = normalize(mk(sx, sy, 1.0f));
    f3 pFocus = c.focalDistance * dir / dir.z;
    float lx, ly;
    sample_disk_polar(u0, u1, lx, ly);
    f3 pLen = c.lensRadius * mk(lx, ly, 0.0f);
    // c2w * vec4(pLen, 1) and (mat3)c2w * Normalize(pFocus - pLen)
    o = c.cx * pLen.x + c.cy * pLen.y + c.cz * pLen.z + c.cw * 1.0f;
    f3 dl = normalize(pFocus - pLen);
    d = c.cx * dl.x + c.cy * dl.y + c.cz * dl.z;
}

__device__ __forceinline__ void store_ray(pg_ray* r, f3 o, f3 d, float tmin, float tmax) {
    float4* p = reinterpret_cast<float4*>(r);
    p[0] = make_float4(o.x, o.y, o.z, tmin);
    p[1] = make_float4(d.x, d.y, d.z, tmax);
}

// d_np (optional): receives n_paths, the first bounce's path count for a
// bounce loop that reads its counts from device memory.
__global__ __launch_bounds__(256) void k_raygen(SceneDev s, PathsDev P, int32_t n_paths, int32_t spb, int32_t sample0,
                                                int32_t* d_np) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && d_np)
        *d_np = n_paths;
    if (i >= n_paths)
        return;
    const uint32_t pixel = (uint32_t)i / (uint32_t)spb, sl = (uint32_t)i - pixel * (uint32_t)spb;
    const int py = (int)(pixel / (uint32_t)s.width), px = (int)pixel - py * s.width;
    SamplerState ss;
    ss.index = sampler_start(s.sampler, px, py, sample0 + (int)sl);
    ss.dim = 2;
    float u0, u1, l0, l1;
    get_2d(s.sampler, ss, u0, u1);
    // (p + Get2D()) / film->Size()  (integrator.cpp:153)
    float fx = ((float)px + u0) / (float)s.width;
    float fy = ((float)py + u1) / (float)s.height;
    get_2d(s.sampler, ss, l0, l1);
    f3 o, d;
    gen_ray(s.cam, fx, fy, l0, l1, o, d);
    P.L[3 * i + 0] = P.L[3 * i + 1] = P.L[3 * i + 2] = 0.0f;
    P.beta[3 * i + 0] = P.beta[3 * i + 1] = P.beta[3 * i + 2] = 1.0f;
    P.bsdf_pdf[i] = 0.0f;
    P.sidx[i] = ss.index;
    P.sdim[i] = ss.dim;
    P.depth[i] = 0;
    P.pfilm[2 * i + 0] = fx;
    P.pfilm[2 * i + 1] = fy;
    P.queue_in[i] = i;
    store_ray(&P.rays_in[i], o, d, 0.0f, kFloatMax);
    if (P.medium)
        P.medium[i] = s.camera_medium;  // camera.h:61 ray.medium = medium
    if (P.lam) {
        // Spectral mode: the hero-wavelength variable from a hash of (pixel,
        // sample index), so a render is deterministic and batch-independent.
        uint32_t h = pixel * 0x9E3779B1u ^ ((uint32_t)sample0 + sl + 0x7F4A7C15u) * 0x85EBCA77u;
        h ^= h >> 15;
        h *= 0x2C1B3C6Du;
        h ^= h >> 12;
        h *= 0x297A2D39u;
        h ^= h >> 15;
        P.lam[i] = (float)(h >> 8) * (1.0f / 16777216.0f);
        for (int j = 0; j < 4; j++) {
            P.L4[4 * (int64_t)i + j] = 0.0f;
            P.beta4[4 * (int64_t)i + j] = 1.0f;
        }
    }
}

// Sky::Color / SkyColor (light.cpp:74-76, color.cpp:91-99): a gradient in the
// angle from +y (SkyColor swaps y and z, then takes acos of z), 8 x sunColor
// exactly along the sun direction.
__device__ __forceinline__ f3 sky_color(const SceneDev& s, f3 d) {
    if (d.x == s.sky_dir.x && d.y == s.sky_dir.y && d.z == s.sky_dir.z)
        return 8.0f * s.sky_color;
    const float theta = pstd::acos(d.y) * kInvPi;
    const f3 lo = mk(0.01f, 0.03f, 0.3f), hi = mk(0.3f, 0.5f, 0.8f);
    const f3 l = lo * (1.0f - theta) + hi * theta;  // pstd::lerp (math.h:96-98)
    return s.sky_color * (l * l);
}

// Atmosphere::Color (light.cpp:106-135): the direction's (phi / 2pi,
// (1 - cos theta) / 2) scaled to the table, nearest texel or a bilinear blend
// of the four around it (cast-based floor/ceil, indices reduced by size with
// C's signed remainder, the blend weights taken against those indices).
// Pine's approximate atan2 diverges for x -> 0+ with y < 0 (phi near -1e5
// instead of 3pi/2), so its x index can be negative there.  Pine then reads
// colors[y * size.x + x]: another row of its table, or memory outside it (its
// nearest-texel mode crashes on such scenes).  The flat index is formed the
// same way here and wrapped into the table, so every read pine makes inside
// its table is the same texel, and none leaves the table.
__device__ __forceinline__ f3 atmo_texel(const SceneDev& s, int x, int y) {
    const int64_t n = (int64_t)s.atmo_w * s.atmo_h;
    int64_t flat = ((int64_t)y * s.atmo_w + x) % n;
    flat = flat < 0 ? flat + n : flat;
    const float* c = s.atmo_colors + 3 * flat;
    return mk(c[0], c[1], c[2]);
}
__device__ __forceinline__ f3 atmosphere_color(const SceneDev& s, f3 d) {
    // InverseUniformSphereMampling (sampling.h:44-47) via CartesianToSpherical
    // (vecmath.h:1243-1250).
    float phi = pstd::atan2(d.y, d.x);
    phi = phi < 0.0f ? kPi * 2.0f + phi : phi;
    const float theta = pstd::acos(d.z);
    const float ux = phi * (1.0f / (kPi * 2.0f)), uy = (1.0f - pstd::cos(theta)) * 0.5f;
    const float fx = ux * (float)s.atmo_w, fy = uy * (float)s.atmo_h;
    if (!s.atmo_interp) {
        const float mx = (float)s.atmo_w - 1.0f, my = (float)s.atmo_h - 1.0f;
        return atmo_texel(s, (int)(fx < mx ? fx : mx), (int)(fy < my ? fy : my));
    }
    const int x0 = (int)pstd::floor(fx) % s.atmo_w, x1 = (int)pstd::ceil(fx) % s.atmo_w;
    const int y0 = (int)pstd::floor(fy) % s.atmo_h, y1 = (int)pstd::ceil(fy) % s.atmo_h;
    const float px = fx - (float)x0, py = fy - (float)y0;
    const f3 c00 = atmo_texel(s, x0, y0), c01 = atmo_texel(s, x1, y0);
    const f3 c10 = atmo_texel(s, x0, y1), c11 = atmo_texel(s, x1, y1);
    const f3 c0 = c00 * (1.0f - px) + c01 * px;  // pstd::lerp (math.h:96-98)
    const f3 c1 = c10 * (1.0f - px) + c11 * px;
    return c0 * (1.0f - py) + c1 * py;
}

// EnvironmentLight::Color and ::Pdf (light.h:150-160): Sky's pdf is 0,
// Atmosphere's 1 / 4pi (its sampling pdf is half that; pine uses this value in
// the BSDF-side MIS weight).
// The Constant environment (an extension; pine has no such light) is sampled
// uniformly over the sphere, so its pdf is 1 / 4pi as well.
__device__ __forceinline__ f3 env_color(const SceneDev& s, f3 d) {
    if (s.env_kind == kEnvConstant)
        return s.sky_color;
    return s.env_kind == kEnvAtmosphere ? atmosphere_color(s, d) : sky_color(s, d);
}
__device__ __forceinline__ float env_pdf(const SceneDev& s) {
    return s.env_kind == kEnvSky ? 0.0f : 1.0f / (kPi * 4.0f);
}

struct Hit {
    f3 p, n;
    int32_t shape;
};

// Surface data after a closest hit, as BVH::Intersect's post-hit closure and the
// shapes' Intersect members fill the Interaction (bvh.cpp:690-708,
// geometry.cpp:95-195).
__device__ __forceinline__ Hit surface(const SceneDev& s, int32_t prim, float t, float u, float v, f3 o, f3 d) {
    Hit h;
    h.shape = s.prim_shape[prim];
    const ShapeDev& sh = s.shapes[h.shape];
    if (sh.type == kShapeMesh) {
        const float* tv = s.tri_verts + 9 * (int64_t)(sh.tri_offset + (prim - sh.first_prim));
        f3 v0 = mk(tv[0], tv[1], tv[2]), v1 = mk(tv[3], tv[4], tv[5]), v2 = mk(tv[6], tv[7], tv[8]);
        h.p = (1.0f - u - v) * v0 + u * v1 + v * v2;  // InterpolatePosition (geometry.h:308-310)
        h.n = normalize(cross(v0 - v1, v0 - v2));
    } else if (sh.type == kShapeSphere) {
        // Normalize(ray(t) - c) with ray(t) - c evaluated as t*d + (o - c).
        f3 q = mk(d.x * t + (o.x - sh.p0.x), d.y * t + (o.y - sh.p0.y), d.z * t + (o.z - sh.p0.z));
        h.n = normalize(q);
        h.p = sh.p0 + h.n * sh.r;
    } else if (sh.type == kShapeRect) {
        float ur = u - 0.5f, vr = v - 0.5f;
        h.p = sh.p0 + sh.lx * sh.e0 * ur + sh.ly * sh.e1 * vr;
        h.n = sh.n;
    } else {
        f3 p = o + t * d;
        float pu = dot(p - sh.p0, sh.e0), pv = dot(p - sh.p0, sh.e1);
        h.p = sh.p0 + pu * sh.e0 + pv * sh.e1;
        h.n = sh.n;
    }
    return h;
}

// Surface parameterisation of a hit, as the shapes' Intersect members fill
// Interaction::uv, dpdu, dpdv: triangles (geometry.h:292, 311-314 via
// bvh.cpp:695-697) barycentric (u, v) and the edges v1 - v0, v2 - v0;
// spheres (geometry.cpp:140-145) (phi, theta) of the normal and the spherical
// tangents; rects (geometry.cpp:191-193) (u, v) + 1/2 and the unit edges;
// planes (geometry.cpp:104-107) the hit point's coordinates along u, v.
__device__ __forceinline__ void surface_param(const SceneDev& s, int32_t prim, const Hit& h, float t, float u, float v,
                              f3 o, f3 d, f3& uv, f3& dpdu, f3& dpdv) {
    const ShapeDev& sh = s.shapes[h.shape];
    if (sh.type == kShapeMesh) {
        const float* tv = s.tri_verts + 9 * (int64_t)(sh.tri_offset + (prim - sh.first_prim));
        const f3 v0 = mk(tv[0], tv[1], tv[2]), v1 = mk(tv[3], tv[4], tv[5]), v2 = mk(tv[6], tv[7], tv[8]);
        uv = mk(u, v, 0.0f);
        dpdu = v1 - v0;
        dpdv = v2 - v0;
    } else if (sh.type == kShapeSphere) {
        // CartesianToSpherical (vecmath.h:1243-1250)
        float phi = pstd::atan2(h.n.y, h.n.x);
        phi = phi < 0.0f ? kPi * 2.0f + phi : phi;
        const float theta = pstd::acos(h.n.z);
        const float st = pstd::sin(theta), ct = pstd::cos(theta);
        const float sp = pstd::sin(phi), cp = pstd::cos(phi);
        dpdu = mk(st * -sp, st * cp, 0.0f);
        dpdv = mk(ct * cp, ct * sp, -st);
        uv = mk(phi, theta, 0.0f);
    } else if (sh.type == kShapeRect) {
        uv = mk(u, v, 0.0f);
        dpdu = sh.e0;
        dpdv = sh.e1;
    } else {
        const f3 p = o + t * d;
        uv = mk(dot(p - sh.p0, sh.e0), dot(p - sh.p0, sh.e1), 0.0f);
        dpdu = sh.e0;
        dpdv = sh.e1;
    }
}

// A material with node-graph inputs, made concrete at a surface point: the
// bump map first replaces the shading normal (Material::BumpNormal,
// material.cpp:49-63, as PathIntegrator::Li does before anything else,
// path.cpp:48), then each input program is evaluated with NodeEvalCtx (p, the
// bumped n, uv) (material.h:14-33) into the layer's constant.
// The material `src` with its node-graph inputs evaluated at this shading
// point, written to the slot's own copy in device memory (P.node_mat; the
// interpreter's stack is P.node_stack).  The copy and the stack stay out of
// registers: a local copy, indexed by layer at run time, would put the
// material and the stack in scratch memory in every scene.
__device__ __forceinline__ const MaterialDev* apply_nodes(const SceneDev& s, const PathsDev& P, int64_t slot,
                                          const MaterialDev* src, Hit& h, int32_t prim, float t, float u,
                                          float v, f3 o, f3 d) {
    MaterialDev* m = P.node_mat + slot;
    f3* st = P.node_stack + nodes::kStack * slot;
    *m = *src;
    f3 uv, dpdu, dpdv;
    surface_param(s, prim, h, t, u, v, o, d, uv, dpdu, dpdv);
    nodes::Ctx c{h.p, h.n, uv.x, uv.y};
    if (m->prog_bump >= 0) {
        h.n = nodes::bump_normal(s.node_prog + m->prog_bump, c, dpdu, dpdv, st);
        c.n = h.n;
    }
    for (int i = 0; i < m->nlayers; i++) {
        Layer& l = m->layers[i];
        if (l.prog_albedo >= 0)
            l.albedo = nodes::eval(s.node_prog + l.prog_albedo, c, st);
        if (l.prog_roughness >= 0)
            l.roughness = nodes::eval(s.node_prog + l.prog_roughness, c, st).x;
        if (l.prog_eta >= 0)
            l.eta = nodes::eval(s.node_prog + l.prog_eta, c, st).x;
    }
    return m;
}

// EmissiveMaterial::Le through MaterialEvalCtx (material.h:14-33, 60-65): the
// emission when wi, expressed in the frame of n, points above the surface.
__device__ __forceinline__ f3 emitted(const MaterialDev& m, f3 n, f3 wi) {
    M3 w2n = inverse(frame_from(n));
    f3 wl = mul(w2n, wi);
    return cos_theta(wl) > 0.0f ? m.emission : mk(0, 0, 0);
}

// PowerLightSampler::SampleLight / UniformLightSampler::SampleLight
// (lightsampler.cpp:6-34) with Distribution1D::SampleDiscrete
// (distribution.h:27-52): pdf is func[i]/n, not normalised, as in pine.
__device__ __forceinline__ int sample_light(const SceneDev& s, float u, float& pdf) {
    const int n = s.n_lights;
    if (s.light_uniform) {
        int idx = (int)(u * (float)n);
        idx = idx < n - 1 ? idx : n - 1;
        pdf = 1.0f / (float)n;
        return idx;
    }
    float cont, fpdf;
    if (s.light_func_int == 0.0f) {
        fpdf = 1.0f;
        cont = u;
    } else {
        // FindInterval(size, cdf[i] <= u) (math.h FindInterval), clamped.
        int first = 0, len = n + 1;
        while (len > 0) {
            int half = len >> 1, middle = first + half;
            if (s.light_cdf[middle] <= u) {
                first = middle + 1;
                len -= half + 1;
            } else {
                len = half;
            }
        }
        int offset = first - 1;
        offset = offset < 0 ? 0 : (offset > n - 1 ? n - 1 : offset);
        float du = u - s.light_cdf[offset];
        if (s.light_cdf[offset + 1] - s.light_cdf[offset] > 0)
            du /= s.light_cdf[offset + 1] - s.light_cdf[offset];
        fpdf = s.light_func[offset];
        cont = ((float)offset + du) / (float)n;
        cont = cont < kOneMinusEps ? cont : kOneMinusEps;
    }
    int idx = (int)(cont * (float)n);
    pdf = fpdf * (1.0f / (float)n);
    return idx;
}

struct LightSampleD {
    f3 Le, wo;
    float distance, pdf;
    bool delta;
};

// Light::Sample (light.cpp:9-63) and Shape::Sample (geometry.h:435-442).
__device__ __forceinline__ LightSampleD sample_light_point(const SceneDev& s, const LightDev& L, f3 p, float u0,
                                           float u1) {
    LightSampleD ls;
    ls.delta = false;
    if (L.type == kLightPoint) {
        ls.wo = normalize(L.v - p, ls.distance);
        ls.pdf = ls.distance * ls.distance;
        ls.Le = L.color;
        ls.delta = true;
        return ls;
    }
    if (L.type == kLightSky) {
        // Sky::Sample (light.cpp:65-73): the sun as a delta light.
        ls.distance = 1e10f;
        ls.delta = true;
        ls.wo = s.sky_dir;
        ls.Le = 8.0f * s.sky_color;
        ls.pdf = 1.0f;
        return ls;
    }
    if (L.type == kLightAtmosphere) {
        // Atmosphere::Sample (light.cpp:94-111): half the time a uniform sphere
        // direction with the nearest table colour, else the sun as a delta light.
        ls.distance = 1e10f;
        if (u0 < 0.5f) {
            u0 *= 2.0f;
            const float phi = u0 * kPi * 2.0f, theta = pstd::acos(1.0f - 2.0f * u1);
            const float st = pstd::sin(theta);
            ls.wo = mk(st * pstd::cos(phi), st * pstd::sin(phi), pstd::cos(theta));
            ls.pdf = 1.0f / (kPi * 4.0f);
            const float fx = u0 * (float)s.atmo_w, fy = u1 * (float)s.atmo_h;
            const float mx = (float)s.atmo_w - 1.0f, my = (float)s.atmo_h - 1.0f;
            ls.Le = atmo_texel(s, (int)(fx < mx ? fx : mx), (int)(fy < my ? fy : my));
        } else {
            ls.wo = s.sky_dir;
            ls.pdf = 1.0f;
            ls.Le = s.atmo_sun_le;
            ls.delta = true;
        }
        ls.pdf = ls.pdf * 0.5f;
        return ls;
    }
    if (L.type == kLightConstantEnv) {
        // A direction uniform on the sphere (UniformSphereSampling,
        // sampling.h:41-43) and the constant radiance.
        ls.distance = 1e10f;
        const float phi = u0 * kPi * 2.0f, theta = pstd::acos(1.0f - 2.0f * u1);
        const float st = pstd::sin(theta);
        ls.wo = mk(st * pstd::cos(phi), st * pstd::sin(phi), pstd::cos(theta));
        ls.pdf = 1.0f / (kPi * 4.0f);
        ls.Le = L.color;
        return ls;
    }
    if (L.type == kLightDirectional) {
        ls.Le = L.color;
        ls.wo = L.v;
        ls.distance = 1e10f;
        ls.pdf = 1.0f;
        ls.delta = true;
        return ls;
    }
    const ShapeDev& sh = s.shapes[L.shape];
    f3 sp, sn;
    if (sh.type == kShapeRect) {
        sp = sh.p0 + sh.e0 * sh.lx * (u0 - 0.5f) + sh.e1 * sh.ly * (u1 - 0.5f);
        sn = sh.n;
        sp = offset_ray_origin(sp, sn);
    } else if (sh.type == kShapeSphere) {
        // UniformSphereSampling (sampling.h:41-43) via SphericalToCartesian.
        float phi = u0 * kPi * 2.0f, theta = pstd::acos(1.0f - 2.0f * u1);
        float st = pstd::sin(theta);
        sn = mk(st * pstd::cos(phi), st * pstd::sin(phi), pstd::cos(theta));
        sp = sh.p0 + sh.r * sn;
        sp = offset_ray_origin(sp, sn);
    } else {
        // Single-triangle emitters (meshes cannot be sampled in pine).
        const float* tv = s.tri_verts + 9 * (int64_t)sh.tri_offset;
        f3 v0 = mk(tv[0], tv[1], tv[2]), v1 = mk(tv[3], tv[4], tv[5]), v2 = mk(tv[6], tv[7], tv[8]);
        if (u0 + u1 > 1.0f) {
            u0 = 1.0f - u0;
            u1 = 1.0f - u1;
        }
        sp = (1.0f - u0 - u1) * v0 + u0 * v1 + u1 * v2;
        sn = normalize(cross(v0 - v1, v0 - v2));
        sp = offset_ray_origin(sp, sn);
    }
    ls.wo = normalize(sp - p, ls.distance);
    const MaterialDev& m = s.materials[sh.material];
    ls.Le = emitted(m, sn, -ls.wo);
    ls.pdf = ls.distance * ls.distance / (abs_dot(-ls.wo, sn) * sh.area);
    return ls;
}

// Network input of a surface vertex: position in the scene bounds, normal,
// outgoing direction.
__device__ __forceinline__ void guide_input(const SceneDev& s, f3 p, f3 n, f3 wo, float* x) {
    x[0] = (p.x - s.g_bmin.x) * s.g_binv.x;
    x[1] = (p.y - s.g_bmin.y) * s.g_binv.y;
    x[2] = (p.z - s.g_bmin.z) * s.g_binv.z;
    x[3] = n.x, x[4] = n.y, x[5] = n.z;
    x[6] = wo.x, x[7] = wo.y, x[8] = wo.z;
}

// Vertices the guide serves: single-layer diffuse surfaces, whose BSDF
// sampling density is exactly material_pdf (the mixture pdf needs it).
__device__ __forceinline__ bool guidable(const SceneDev& s, int4 info) {
    if (info.x != 0)
        return false;
    const MaterialDev& m = s.materials[info.y];
    return m.kind == kMatLayered && m.nlayers == 1 && m.layers[0].type == kDiffuse;
}

// Direction sampling at a guided vertex: one-sample MIS of the guide (share
// alpha) and the BSDF; the estimator divides by the mixture pdf, whichever
// strategy drew wo, so it stays unbiased for any guide.  With training on,
// the vertex is recorded (network input, direction, pdfs, path radiance so far
// and throughput after the direction).  False when the sample carries nothing.
__device__ __forceinline__ bool guided_direction(const SceneDev& s, PathsDev& P, SamplerState& ss, int32_t i, int depth,
                                 const float* gp, const MaterialDev& m, f3 p, f3 n, f3 wi, f3 L_before,
                                 f3& beta, f3& wo_out) {
    const M3 n2w = frame_from(n);
    const M3 w2n = inverse(n2w);
    const f3 wiL = mul(w2n, wi);
    const float alpha = gp[guide::kAlphaAt];
    const float usel = get_1d(s.sampler, ss);
    const float u1 = get_1d(s.sampler, ss);
    float u2a, u2b;
    get_2d(s.sampler, ss, u2a, u2b);
    f3 wo = mk(0.0f, 0.0f, 1.0f), f = mk(0.0f, 0.0f, 0.0f);
    float pb = 0.0f;
    if (usel < alpha) {
        wo = guide::mix_sample(gp, u1, u2a, u2b);
        const f3 woL = mul(w2n, wo);
        f = material_f(m, wiL, woL);
        pb = material_pdf(m, wiL, woL);
    } else {
        BSDFSample bs;
        if (material_sample(m, wiL, u1, u2a, u2b, bs)) {
            wo = mul(n2w, bs.wo);
            f = bs.f;
            pb = bs.pdf;
        }
    }
    const float pmix = alpha * guide::mix_pdf(gp, wo) + (1.0f - alpha) * pb;
    if (is_black(f) || !(pmix > 0.0f))
        return false;
    beta = beta * (abs_dot(wo, n) * f / pmix);
    P.bsdf_pdf[i] = pmix;
    wo_out = wo;
    if (depth < P.rec_depth) {
        const int64_t slot = (int64_t)i * P.rec_depth + depth;
        float x[guide::kIn];
        guide_input(s, p, n, wi, x);
        for (int c = 0; c < guide::kIn; c++)
            P.rec_x[guide::kIn * slot + c] = x[c];
        P.rec_dir[3 * slot + 0] = wo.x;
        P.rec_dir[3 * slot + 1] = wo.y;
        P.rec_dir[3 * slot + 2] = wo.z;
        P.rec_pb[slot] = pb;
        P.rec_pdf[slot] = pmix;
        P.rec_lb[3 * slot + 0] = L_before.x;
        P.rec_lb[3 * slot + 1] = L_before.y;
        P.rec_lb[3 * slot + 2] = L_before.z;
        P.rec_beta[3 * slot + 0] = beta.x;
        P.rec_beta[3 * slot + 1] = beta.y;
        P.rec_beta[3 * slot + 2] = beta.z;
    }
    return true;
}

// kGuided: directions at the vertices the guide serves (gflag, set for the
// bounce by k_guide_select / k_guide_infer) come from the guide/BSDF mixture,
// and the light sample's MIS weight uses the mixture pdf.
// d_active (optional): the active count in device memory, read in place of
// n_active (launches sized for the queue's capacity).
// --- queue slots claimed per workgroup -----------------------------------------
// Device-scope atomics on one address serialise: on MI355X about 11 ns each,
// so a 4M-path bounce whose waves each claim a queue slot (the compiler already
// merges a wave's lanes into one atomic) spends ~0.75 ms per counter on them.
// These helpers claim a whole workgroup's slots with one atomic: the waves'
// counts go through LDS, one lane adds the workgroup's total, every lane gets
// the counter's old value plus its rank among the requesting lanes (waves in
// order, lanes in order).  Every lane of the workgroup must call them (no early
// return before); they begin with a barrier so that they can follow each other.
constexpr int kSlotBlock = 256, kSlotWaves = kSlotBlock / 64;
// k_shade's workgroup: the fused bounce launches it for the batch's capacity,
// so a near-empty bounce dispatches a workgroup per kShadeBlock slots.
constexpr int kShadeBlock = 512;

// Two adjacent counters ctr[0], ctr[1] (8-byte aligned) in one 64-bit atomic:
// the low counter never carries into the high one (counts stay below 2^31).
template <int kWaves = kSlotWaves>
__device__ __forceinline__ void block_slots2(bool want_a, bool want_b, int32_t* ctr, int32_t& slot_a,
                                             int32_t& slot_b) {
    __shared__ uint32_t cnt[2][kWaves];
    __shared__ uint32_t base[2];
    __syncthreads();
    const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
    const uint64_t ma = __ballot(want_a), mb = __ballot(want_b);
    if (lane == 0) {
        cnt[0][wave] = (uint32_t)__popcll(ma);
        cnt[1][wave] = (uint32_t)__popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ta = 0, tb = 0;
        for (int w = 0; w < kWaves; w++) {
            ta += cnt[0][w];
            tb += cnt[1][w];
        }
        unsigned long long old = 0;
        if (ta | tb)
            old = atomicAdd(reinterpret_cast<unsigned long long*>(ctr),
                            (unsigned long long)ta | ((unsigned long long)tb << 32));
        base[0] = (uint32_t)old;
        base[1] = (uint32_t)(old >> 32);
    }
    __syncthreads();
    uint32_t pa = base[0], pb = base[1];
    for (int w = 0; w < wave; w++) {
        pa += cnt[0][w];
        pb += cnt[1][w];
    }
    const uint64_t lt = (1ull << lane) - 1ull;
    slot_a = (int32_t)(pa + (uint32_t)__popcll(ma & lt));
    slot_b = (int32_t)(pb + (uint32_t)__popcll(mb & lt));
}

// One counter.
__device__ __forceinline__ int32_t block_slot(bool want, int32_t* ctr) {
    __shared__ uint32_t cnt[kSlotWaves];
    __shared__ uint32_t base;
    __syncthreads();
    const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
    const uint64_t m = __ballot(want);
    if (lane == 0)
        cnt[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kSlotWaves; w++)
            t += cnt[w];
        base = t ? (uint32_t)atomicAdd(ctr, (int32_t)t) : 0u;
    }
    __syncthreads();
    uint32_t p = base;
    for (int w = 0; w < wave; w++)
        p += cnt[w];
    return (int32_t)(p + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)));
}

template <bool kGuided>
__global__ __launch_bounds__(kShadeBlock) void k_shade(SceneDev s, PathsDev P, int32_t n_active, const int32_t* d_active) {
    const int32_t n = d_active ? *d_active : n_active;
    if ((int64_t)blockIdx.x * blockDim.x >= n)
        return;  // the whole workgroup
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    // The shadow ray and the next ray are written after the workgroup has
    // claimed its queue slots (block_slots2); lanes past the count take part in
    // the claim only.
    bool alive = false, shadow = false;
    f3 nextO, nextD, sh_o, sh_d, sh_c;
    float sh_tmax = 0.0f;
    int32_t sh_rec = -1, i = 0;
    if (k < n) {
    i = P.queue_in[k];
    const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
    const float4 ra = r4[0], rb = r4[1];
    const f3 ro = mk(ra.x, ra.y, ra.z), rd = mk(rb.x, rb.y, rb.z);
    const pg_hit hit = P.hits[k];
    int depth = P.depth[i];
    f3 L = mk(P.L[3 * i], P.L[3 * i + 1], P.L[3 * i + 2]);
    f3 beta = mk(P.beta[3 * i], P.beta[3 * i + 1], P.beta[3 * i + 2]);
    const float bsdfPdf = P.bsdf_pdf[i];
    SamplerState ss;
    ss.index = P.sidx[i];
    ss.dim = P.sdim[i];

    if (hit.prim < 0) {
        if (s.has_env) {
            // path.cpp:29-39: the BSDF-side MIS weight against the environment's Pdf.
            const f3 le = env_color(s, rd);
            L = depth == 0 ? L + beta * le : L + beta * le * power_heuristic(bsdfPdf, env_pdf(s));
        }
    } else {
        const pg_hit_uv uv = P.uvs[k];
        Hit h = surface(s, hit.prim, hit.t, uv.u, uv.v, ro, rd);
        const ShapeDev& sh = s.shapes[h.shape];
        if (sh.material < 0) {
            // Surface without a material: continue straight through (path.cpp:42-46).
            nextD = rd;
            nextO = offset_ray_origin(h.p, face_same_hemisphere(h.n, rd));
            alive = true;
        } else {
            const MaterialDev* mp = s.materials + sh.material;
            if (mp->has_nodes)
                mp = apply_nodes(s, P, k, mp, h, hit.prim, hit.t, uv.u, uv.v, ro, rd);
            const MaterialDev& m = *mp;
            const f3 wi = -rd;
            if (m.kind == kMatEmissive) {
                f3 le = emitted(m, h.n, wi);
                if (depth == 0) {
                    L = L + beta * le;
                } else {
                    // Shape::Pdf (geometry.h:432-434); BSDF-hit MIS (path.cpp:51-60)
                    float lightPdf = hit.t * hit.t / (abs_dot(-rd, h.n) * sh.area);
                    L = L + beta * le * power_heuristic(bsdfPdf, lightPdf);
                }
            } else if (depth + 1 < s.max_depth) {
                const M3 n2w = frame_from(h.n);
                const M3 w2n = inverse(n2w);
                const f3 wiL = mul(w2n, wi);
                // The fused bounce's mixture records are in queue order.
                const bool guided = kGuided && P.gflag[k];
                const float* gp = P.gparams + guide::kParamStride * (int64_t)k;
                // EstimateDirect (integrator.cpp:88-119)
                float lpdf;
                const float ul = get_1d(s.sampler, ss);
                float la, lb;
                if (s.n_lights > 0) {
                    int li = sample_light(s, ul, lpdf);
                    get_2d(s.sampler, ss, la, lb);
                    LightSampleD ls = sample_light_point(s, s.lights[li], h.p, la, lb);
                    ls.pdf *= lpdf;
                    const f3 woL = mul(w2n, ls.wo);
                    f3 f = material_f(m, wiL, woL) * abs_dot(ls.wo, h.n);
                    float scatPdf = 0.0f;
                    if (!ls.delta) {
                        scatPdf = material_pdf(m, wiL, woL);
                        if (guided) {
                            const float alpha = gp[guide::kAlphaAt];
                            scatPdf = alpha * guide::mix_pdf(gp, ls.wo) + (1.0f - alpha) * scatPdf;
                        }
                    }
                    float w = power_heuristic(ls.pdf, scatPdf);
                    f3 ld = ls.delta ? f * ls.Le / ls.pdf : f * w * ls.Le / ls.pdf;
                    if (!is_black(ld)) {
                        shadow = true;
                        sh_c = beta * ld;
                        sh_o = offset_ray_origin(h.p, face_same_hemisphere(h.n, ls.wo));
                        sh_d = ls.wo;
                        sh_tmax = ls.distance * (1.0f - 1e-3f);
                        if (kGuided)
                            sh_rec = guided && depth < P.rec_depth ? i * P.rec_depth + depth : -1;
                    }
                }
                if (guided) {
                    f3 wo;
                    if (guided_direction(s, P, ss, i, depth, gp, m, h.p, h.n, wi, L, beta, wo)) {
                        nextD = wo;
                        nextO = offset_ray_origin(h.p, face_same_hemisphere(h.n, nextD));
                        alive = true;
                        if (depth > s.rr_depth) {
                            float q = pstd::clamp(1.0f - luminance(beta), 0.05f, 1.0f);
                            if (get_1d(s.sampler, ss) < q)
                                alive = false;
                            else
                                beta = beta / (1.0f - q);
                        }
                    }
                } else {
                    // BSDF sample (path.cpp:66-86)
                    const float u1 = get_1d(s.sampler, ss);
                    float u2a, u2b;
                    get_2d(s.sampler, ss, u2a, u2b);
                    BSDFSample bs;
                    if (material_sample(m, wiL, u1, u2a, u2b, bs)) {
                        bs.wo = mul(n2w, bs.wo);
                        if (!is_black(bs.f) && bs.pdf != 0.0f) {
                            beta = beta * (abs_dot(bs.wo, h.n) * bs.f / bs.pdf);
                            P.bsdf_pdf[i] = bs.pdf;
                            nextD = bs.wo;
                            nextO = offset_ray_origin(h.p, face_same_hemisphere(h.n, nextD));
                            alive = true;
                            if (depth > s.rr_depth) {
                                float q = pstd::clamp(1.0f - luminance(beta), 0.05f, 1.0f);
                                if (get_1d(s.sampler, ss) < q)
                                    alive = false;
                                else
                                    beta = beta / (1.0f - q);
                            }
                        }
                    }
                }
                depth++;
            }
        }
    }
    P.L[3 * i + 0] = L.x;
    P.L[3 * i + 1] = L.y;
    P.L[3 * i + 2] = L.z;
    P.beta[3 * i + 0] = beta.x;
    P.beta[3 * i + 1] = beta.y;
    P.beta[3 * i + 2] = beta.z;
    P.sdim[i] = ss.dim;
    P.depth[i] = depth;
    }
    int32_t slot = 0, sslot = 0;
    block_slots2<kShadeBlock / 64>(alive, shadow, P.counters, slot, sslot);
    if (alive) {
        P.queue_out[slot] = i;
        store_ray(&P.rays_out[slot], nextO, nextD, 0.0f, kFloatMax);
    }
    if (shadow) {
        store_ray(&P.shadow_rays[sslot], sh_o, sh_d, 0.0f, sh_tmax);
        P.shadow_path[sslot] = i;
        P.shadow_contrib[3 * sslot + 0] = sh_c.x;
        P.shadow_contrib[3 * sslot + 1] = sh_c.y;
        P.shadow_contrib[3 * sslot + 2] = sh_c.z;
        if (kGuided && P.rec_depth > 0)
            P.shadow_rec[sslot] = sh_rec;
    }
}

__global__ __launch_bounds__(256) void k_resolve_shadow(PathsDev P, int32_t n_shadow, const int32_t* d_shadow) {
    const int32_t n = d_shadow ? *d_shadow : n_shadow;
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (P.occluded[k])
            continue;
        const int32_t i = P.shadow_path[k];
        // One shadow ray per path per bounce: no two lanes share a path here.
        P.L[3 * i + 0] += P.shadow_contrib[3 * k + 0];
        P.L[3 * i + 1] += P.shadow_contrib[3 * k + 1];
        P.L[3 * i + 2] += P.shadow_contrib[3 * k + 2];
        // A recorded guided vertex's own light sample is not radiance along its
        // sampled direction: it joins the record's path radiance before it.
        if (P.rec_depth > 0 && P.shadow_rec) {
            const int32_t r = P.shadow_rec[k];
            if (r >= 0) {
                P.rec_lb[3 * (int64_t)r + 0] += P.shadow_contrib[3 * k + 0];
                P.rec_lb[3 * (int64_t)r + 1] += P.shadow_contrib[3 * k + 1];
                P.rec_lb[3 * (int64_t)r + 2] += P.shadow_contrib[3 * k + 2];
            }
        }
    }
}


// --- pine's AO, Viz and RandomWalk integrators --------------------------------

// AOIntegrator::Li (ao.cpp:9-19): at the camera ray's hit, one direction
// uniform over the sphere (UniformHemisphereMampling, sampling.h:49-51, with
// pine's acos/sin/cos) turned into the hemisphere of the face-forward normal;
// radiance 1 when BVH::Hit finds nothing along it.  Misses give 0.
__global__ __launch_bounds__(kSlotBlock) void k_shade_ao(SceneDev s, PathsDev P, int32_t n_active) {
    if ((int64_t)blockIdx.x * blockDim.x >= n_active)
        return;  // the whole workgroup
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    // The occlusion ray is written after the workgroup's slot claim.
    bool shadow = false;
    f3 sh_o, sh_d;
    int32_t i = 0;
    if (k < n_active) {
    i = P.queue_in[k];
    P.L[3 * i + 0] = P.L[3 * i + 1] = P.L[3 * i + 2] = 0.0f;
    const pg_hit hit = P.hits[k];
    if (hit.prim >= 0) {
    const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
    const float4 ra = r4[0], rb = r4[1];
    const f3 ro = mk(ra.x, ra.y, ra.z), rd = mk(rb.x, rb.y, rb.z);
    const pg_hit_uv uv = P.uvs[k];
    const Hit h = surface(s, hit.prim, hit.t, uv.u, uv.v, ro, rd);
    const f3 n = face_same_hemisphere(h.n, -rd);
    SamplerState ss;
    ss.index = P.sidx[i];
    ss.dim = P.sdim[i];
    float u0, u1;
    get_2d(s.sampler, ss, u0, u1);
    const float phi = u0 * pstd::Pi * 2.0f, theta = pstd::acos(u1);
    const float st = pstd::sin(theta);
    const f3 w = face_same_hemisphere(mk(st * pstd::cos(phi), st * pstd::sin(phi), pstd::cos(theta)), n);
    shadow = true;
    sh_o = offset_ray_origin(h.p, face_same_hemisphere(n, w));
    sh_d = w;
    }
    }
    const int32_t slot = block_slot(shadow, &P.counters[1]);
    if (shadow) {
        store_ray(&P.shadow_rays[slot], sh_o, sh_d, 0.0f, kFloatMax);
        P.shadow_path[slot] = i;
    }
}

__global__ __launch_bounds__(256) void k_resolve_ao(PathsDev P, int32_t n_shadow) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_shadow || P.occluded[k])
        return;
    const int32_t i = P.shadow_path[k];
    P.L[3 * i + 0] = P.L[3 * i + 1] = P.L[3 * i + 2] = 1.0f;
}

// VizIntegrator::Li (viz.cpp:27-41), Position and Normal: the camera ray's
// hit point or shading normal as the colour, 0 for a miss.
// ColorMap (color.cpp:25-35): black -> dark blue -> green -> red.
__device__ __forceinline__ f3 color_map(float v) {
    const f3 k0 = mk(0.0f, 0.0f, 0.0f), k1 = mk(0.0f, 0.0f, 0.5f), k2 = mk(0.0f, 1.0f, 0.0f), k3 = mk(1.0f, 0.0f, 0.0f);
    auto lerp = [](float t, f3 a, f3 b) { return (1.0f - t) * a + t * b; };  // pstd::lerp
    if (v < 1.0f / 3.0f)
        return lerp(v * 3.0f, k0, k1);
    if (v < 2.0f / 3.0f)
        return lerp((v - 1.0f / 3.0f) * 3.0f, k1, k2);
    if (v < 3.0f / 3.0f)
        return lerp((v - 2.0f / 3.0f) * 3.0f, k2, k3);
    return k3;
}

// VizIntegrator::Li (viz.cpp:24-41).  Bvh: pine adds 2 per hierarchy node its
// traversal visits (bvh.cpp:553) and maps count / 200 through ColorMap, hits
// and misses alike; here the count is of this library's hierarchy (`counts`,
// nodes fetched per ray), so the heat map shows this tree, not pine's.
__global__ __launch_bounds__(256) void k_shade_viz(SceneDev s, PathsDev P, int32_t n_active,
                                                   const int32_t* counts) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_active)
        return;
    const int32_t i = P.queue_in[k];
    f3 c = mk(0.0f, 0.0f, 0.0f);
    const pg_hit hit = P.hits[k];
    if (s.viz_type == kVizBvh) {
        c = color_map((2.0f * (float)counts[2 * k]) / 200.0f);
    } else if (hit.prim >= 0) {
        const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
        const float4 ra = r4[0], rb = r4[1];
        const f3 o = mk(ra.x, ra.y, ra.z), d = mk(rb.x, rb.y, rb.z);
        const pg_hit_uv uv = P.uvs[k];
        const Hit h = surface(s, hit.prim, hit.t, uv.u, uv.v, o, d);
        if (s.viz_type == kVizPosition) {
            c = h.p;
        } else if (s.viz_type == kVizNormal) {
            c = h.n;
        } else {
            f3 tc, du, dv;
            surface_param(s, hit.prim, h, hit.t, uv.u, uv.v, o, d, tc, du, dv);
            c = s.viz_type == kVizTexcoord ? tc : s.viz_type == kVizDpDu ? du : dv;
            if (s.viz_type != kVizTexcoord)  // Abs(dpdu), Abs(dpdv)
                c = mk(fabsf(c.x), fabsf(c.y), fabsf(c.z));
        }
    }
    P.L[3 * i + 0] = c.x;
    P.L[3 * i + 1] = c.y;
    P.L[3 * i + 2] = c.z;
}

// RandomWalkIntegrator::Li (randomwalk.cpp:6-62) without media: emission and
// the environment added with throughput and no light sampling; Material::Sample
// per vertex (Get1D, Get2D); no Russian roulette; material-less surfaces
// continue straight without counting depth.
__global__ __launch_bounds__(kSlotBlock) void k_shade_rw(SceneDev s, PathsDev P, int32_t n_active) {
    if ((int64_t)blockIdx.x * blockDim.x >= n_active)
        return;  // the whole workgroup
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    // The next ray is written after the workgroup's slot claim.
    bool alive = false;
    f3 nextO, nextD;
    int32_t i = 0;
    if (k < n_active) {
    i = P.queue_in[k];
    const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
    const float4 ra = r4[0], rb = r4[1];
    const f3 ro = mk(ra.x, ra.y, ra.z), rd = mk(rb.x, rb.y, rb.z);
    const pg_hit hit = P.hits[k];
    int depth = P.depth[i];
    f3 L = mk(P.L[3 * i], P.L[3 * i + 1], P.L[3 * i + 2]);
    f3 beta = mk(P.beta[3 * i], P.beta[3 * i + 1], P.beta[3 * i + 2]);
    SamplerState ss;
    ss.index = P.sidx[i];
    ss.dim = P.sdim[i];
    if (hit.prim < 0) {
        if (s.has_env)
            L = L + beta * env_color(s, rd);
    } else {
        const pg_hit_uv uv = P.uvs[k];
        Hit h = surface(s, hit.prim, hit.t, uv.u, uv.v, ro, rd);
        const ShapeDev& sh = s.shapes[h.shape];
        if (sh.material < 0) {
            nextD = rd;
            nextO = offset_ray_origin(h.p, face_same_hemisphere(h.n, rd));
            alive = true;
        } else {
            const MaterialDev* mp = s.materials + sh.material;
            if (mp->has_nodes)
                mp = apply_nodes(s, P, k, mp, h, hit.prim, hit.t, uv.u, uv.v, ro, rd);
            const MaterialDev& m = *mp;
            const f3 wi = -rd;
            if (m.kind == kMatEmissive) {
                L = L + beta * emitted(m, h.n, wi);
            } else if (depth + 1 < s.max_depth) {
                const M3 n2w = frame_from(h.n);
                const M3 w2n = inverse(n2w);
                const float u1 = get_1d(s.sampler, ss);
                float u2a, u2b;
                get_2d(s.sampler, ss, u2a, u2b);
                BSDFSample bs;
                if (material_sample(m, mul(w2n, wi), u1, u2a, u2b, bs)) {
                    bs.wo = mul(n2w, bs.wo);
                    beta = beta * (abs_dot(bs.wo, h.n) * bs.f / bs.pdf);
                    nextD = bs.wo;
                    nextO = offset_ray_origin(h.p, face_same_hemisphere(h.n, nextD));
                    alive = true;
                }
                depth++;
            }
        }
    }
    P.L[3 * i + 0] = L.x;
    P.L[3 * i + 1] = L.y;
    P.L[3 * i + 2] = L.z;
    P.beta[3 * i + 0] = beta.x;
    P.beta[3 * i + 1] = beta.y;
    P.beta[3 * i + 2] = beta.z;
    P.sdim[i] = ss.dim;
    P.depth[i] = depth;
    }
    const int32_t slot = block_slot(alive, &P.counters[0]);
    if (alive) {
        P.queue_out[slot] = i;
        store_ray(&P.rays_out[slot], nextO, nextD, 0.0f, kFloatMax);
    }
}

// --- spectral mode (fused bounce, no guide) -----------------------------------

struct Spec4 {
    float v[4];
};
__device__ __forceinline__ Spec4 spec_load(const float* p, int64_t i) {
    const float4 a = reinterpret_cast<const float4*>(p)[i];
    return {{a.x, a.y, a.z, a.w}};
}
__device__ __forceinline__ void spec_store(float* p, int64_t i, const Spec4& s) {
    reinterpret_cast<float4*>(p)[i] = make_float4(s.v[0], s.v[1], s.v[2], s.v[3]);
}

// k_shade's fused bounce with every colour a 4-wavelength spectrum: RGB
// quantities from the scene (albedo-driven BSDF values, emission, sky) are
// upsampled at the path's wavelengths (spec::upsample_row), products are taken
// per wavelength, Russian roulette uses the mean throughput.  Geometry,
// sampling and the sampler's dimension order are those of k_shade.
__global__ __launch_bounds__(kSlotBlock) void k_shade_spec(SceneDev s, PathsDev P, int32_t n_active) {
    if ((int64_t)blockIdx.x * blockDim.x >= n_active)
        return;  // the whole workgroup
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    // Queue writes wait for the workgroup's slot claim (block_slots2).
    bool alive = false, shadow = false;
    f3 nextO, nextD, sh_o, sh_d;
    float sh_tmax = 0.0f;
    Spec4 sh_c;
    int32_t i = 0;
    if (k < n_active) {
    i = P.queue_in[k];
    const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
    const float4 ra = r4[0], rb = r4[1];
    const f3 ro = mk(ra.x, ra.y, ra.z), rd = mk(rb.x, rb.y, rb.z);
    const pg_hit hit = P.hits[k];
    int depth = P.depth[i];
    Spec4 L = spec_load(P.L4, i), beta = spec_load(P.beta4, i);
    f3 up[4];
    {
        const float u = P.lam[i];
        for (int j = 0; j < 4; j++)
            up[j] = spec::upsample_row(s.spec, spec::wavelength(u, j));
    }
    auto lift = [&](f3 c, float scale) {
        Spec4 r;
        for (int j = 0; j < 4; j++)
            r.v[j] = fmaxf(0.0f, dot(up[j], c)) * scale;
        return r;
    };
    const float bsdfPdf = P.bsdf_pdf[i];
    SamplerState ss;
    ss.index = P.sidx[i];
    ss.dim = P.sdim[i];
    if (hit.prim < 0) {
        if (s.has_env) {
            const float w = depth == 0 ? 1.0f : power_heuristic(bsdfPdf, env_pdf(s));
            const Spec4 le = lift(env_color(s, rd), w);
            for (int j = 0; j < 4; j++)
                L.v[j] += beta.v[j] * le.v[j];
        }
    } else {
        const pg_hit_uv uv = P.uvs[k];
        Hit h = surface(s, hit.prim, hit.t, uv.u, uv.v, ro, rd);
        const ShapeDev& sh = s.shapes[h.shape];
        const MaterialDev* mp = s.materials + sh.material;
        if (mp->has_nodes)
            mp = apply_nodes(s, P, k, mp, h, hit.prim, hit.t, uv.u, uv.v, ro, rd);
        const MaterialDev& m = *mp;
        const f3 wi = -rd;
        if (m.kind == kMatEmissive) {
            const float w = depth == 0 ? 1.0f
                                       : power_heuristic(bsdfPdf, hit.t * hit.t / (abs_dot(-rd, h.n) * sh.area));
            const Spec4 le = lift(emitted(m, h.n, wi), w);
            for (int j = 0; j < 4; j++)
                L.v[j] += beta.v[j] * le.v[j];
        } else if (depth + 1 < s.max_depth) {
            const M3 n2w = frame_from(h.n);
            const M3 w2n = inverse(n2w);
            const f3 wiL = mul(w2n, wi);
            float lpdf;
            const float ul = get_1d(s.sampler, ss);
            float la, lb;
            if (s.n_lights > 0) {
                int li = sample_light(s, ul, lpdf);
                get_2d(s.sampler, ss, la, lb);
                LightSampleD ls = sample_light_point(s, s.lights[li], h.p, la, lb);
                ls.pdf *= lpdf;
                const f3 woL = mul(w2n, ls.wo);
                const f3 f = material_f(m, wiL, woL) * abs_dot(ls.wo, h.n);
                const float scatPdf = ls.delta ? 0.0f : material_pdf(m, wiL, woL);
                const float w = ls.delta ? 1.0f : power_heuristic(ls.pdf, scatPdf);
                const Spec4 fs = lift(f, 1.0f), le = lift(ls.Le, w / ls.pdf);
                Spec4 contrib;
                bool any = false;
                for (int j = 0; j < 4; j++) {
                    contrib.v[j] = beta.v[j] * fs.v[j] * le.v[j];
                    any = any || contrib.v[j] != 0.0f;
                }
                if (any) {
                    shadow = true;
                    sh_o = offset_ray_origin(h.p, face_same_hemisphere(h.n, ls.wo));
                    sh_d = ls.wo;
                    sh_tmax = ls.distance * (1.0f - 1e-3f);
                    sh_c = contrib;
                }
            }
            const float u1 = get_1d(s.sampler, ss);
            float u2a, u2b;
            get_2d(s.sampler, ss, u2a, u2b);
            BSDFSample bs;
            if (material_sample(m, wiL, u1, u2a, u2b, bs)) {
                bs.wo = mul(n2w, bs.wo);
                if (!is_black(bs.f) && bs.pdf != 0.0f) {
                    const Spec4 fs = lift(bs.f, abs_dot(bs.wo, h.n) / bs.pdf);
                    for (int j = 0; j < 4; j++)
                        beta.v[j] *= fs.v[j];
                    P.bsdf_pdf[i] = bs.pdf;
                    nextD = bs.wo;
                    nextO = offset_ray_origin(h.p, face_same_hemisphere(h.n, nextD));
                    alive = true;
                    if (depth > s.rr_depth) {
                        const float mean = 0.25f * (beta.v[0] + beta.v[1] + beta.v[2] + beta.v[3]);
                        const float q = pstd::clamp(1.0f - mean, 0.05f, 1.0f);
                        if (get_1d(s.sampler, ss) < q) {
                            alive = false;
                        } else {
                            for (int j = 0; j < 4; j++)
                                beta.v[j] /= (1.0f - q);
                        }
                    }
                }
            }
            depth++;
        }
    }
    spec_store(P.L4, i, L);
    spec_store(P.beta4, i, beta);
    P.sdim[i] = ss.dim;
    P.depth[i] = depth;
    }
    int32_t slot = 0, sslot = 0;
    block_slots2(alive, shadow, P.counters, slot, sslot);
    if (alive) {
        P.queue_out[slot] = i;
        store_ray(&P.rays_out[slot], nextO, nextD, 0.0f, kFloatMax);
    }
    if (shadow) {
        store_ray(&P.shadow_rays[sslot], sh_o, sh_d, 0.0f, sh_tmax);
        P.shadow_path[sslot] = i;
        spec_store(P.sh4, sslot, sh_c);
    }
}

__global__ __launch_bounds__(256) void k_resolve_shadow_spec(PathsDev P, int32_t n_shadow) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_shadow || P.occluded[k])
        return;
    const int32_t i = P.shadow_path[k];
    Spec4 L = spec_load(P.L4, i);
    const Spec4 c = spec_load(P.sh4, k);
    for (int j = 0; j < 4; j++)
        L.v[j] += c.v[j];
    spec_store(P.L4, i, L);
}

// Each path's 4-wavelength radiance to the film's RGB (spec::film_weight).
__global__ __launch_bounds__(256) void k_spec_to_rgb(SceneDev s, PathsDev P, int32_t n_paths) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_paths)
        return;
    const Spec4 L = spec_load(P.L4, i);
    const float u = P.lam[i];
    f3 rgb = mk(0.0f, 0.0f, 0.0f);
    for (int j = 0; j < 4; j++)
        rgb = rgb + spec::film_weight(s.spec, spec::wavelength(u, j)) * L.v[j];
    P.L[3 * i + 0] = rgb.x;
    P.L[3 * i + 1] = rgb.y;
    P.L[3 * i + 2] = rgb.z;
}

// ---------------------------------------------------------------------------
// General mode: media and material-less shapes.  One bounce of
// PathIntegrator::Li (path.cpp:6-87) becomes
//   k_shade_gen   medium sample, env / emission, light sample -> shadow record
//   shadow loop   trace_closest + k_shadow_tr per segment (IntersectTr,
//                 integrator.cpp:68-87), consuming Tr's sampler dimensions
//   k_resolve_gen L += beta * tr * f * w * Le / pdf
//   k_shade_gen_b BSDF or phase sample, Russian roulette -> next ray

// Adds a thread's density-lookup count to the global counter: a cross-lane
// sum per wave, the waves' sums through LDS, one atomic per workgroup (see
// block_slots2 on same-address atomics).  Every lane of the block reaches
// this call.
__device__ __forceinline__ void add_lookups(unsigned long long* counter, int n) {
    __shared__ int wsum[16];
    for (int off = 32; off > 0; off >>= 1)
        n += __shfl_down(n, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
        wsum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; w++)
            t += (unsigned long long)wsum[w];
        if (t > 0)
            atomicAdd(counter, t);
    }
}

// Brick majorants of medium m: the LDS copy when this launch staged it.
__device__ __forceinline__ const float* majorants(const SceneDev& s, int m, int lds_medium,
                                                  const float* lds) {
    return m == lds_medium ? lds : s.media[m].majorant;
}

// Loads the brick majorants of `lds_medium` into LDS (once per workgroup of a
// grid-stride kernel).
__device__ __forceinline__ const float* stage_majorants(const SceneDev& s, int lds_medium, float* lds) {
    if (lds_medium < 0)
        return nullptr;
    const MediumDev& m = s.media[lds_medium];
    const int nb = m.nbricks[0] * m.nbricks[1] * m.nbricks[2];
    for (int k = threadIdx.x; k < nb; k += blockDim.x)
        lds[k] = m.majorant[k];
    __syncthreads();
    return lds;
}

__device__ __forceinline__ int medium_for(f3 w, f3 n, int inside, int outside) {
    return dot(w, n) > 0.0f ? outside : inside;  // Interaction::GetMedium (interaction.h:54-56)
}

// EstimateDirect's light sample (integrator.cpp:88-119) up to the shadow ray,
// for a surface (its material, node-graph inputs evaluated) or medium (mat
// null, phase g) vertex.  The record is
// written even when the contribution is black: IntersectTr still runs, and in
// media it consumes sampler dimensions.
// EstimateDirect's light sample at a split-bounce vertex (integrator.cpp:
// 88-100): light choice, point, BSDF or phase value, MIS weight and the first
// shadow segment, held in registers until the workgroup has claimed the
// record's slot (light_record_write).
struct LightRecord {
    float4 r0, r1, r2, r3;  // {f, w}, {Le, pdf}, {target p2, delta}, {wo, pb}
    f3 so;                  // first segment: origin, direction r3.xyz, range stmax
    float stmax;
    int32_t seg_med;
    bool valid;
};

__device__ __forceinline__ void light_record(const SceneDev& s, SamplerState& ss, f3 p, f3 n, f3 wi,
                                             const MaterialDev* mat, float g, int med_in, int med_out,
                                             LightRecord& out) {
    const float ul = get_1d(s.sampler, ss);  // SampleLight's Get1D runs even with no lights
    if (s.n_lights <= 0)
        return;
    float lpdf, la, lb;
    const int li = sample_light(s, ul, lpdf);
    get_2d(s.sampler, ss, la, lb);
    LightSampleD ls = sample_light_point(s, s.lights[li], p, la, lb);
    ls.pdf *= lpdf;
    f3 f;
    float scatPdf = 0.0f, pb = 0.0f;
    if (mat) {
        const MaterialDev& m = *mat;
        const M3 w2n = inverse(frame_from(n));
        f = material_f(m, mul(w2n, wi), mul(w2n, ls.wo)) * abs_dot(ls.wo, n);
        pb = material_pdf(m, mul(w2n, wi), mul(w2n, ls.wo));
        if (!ls.delta)
            scatPdf = pb;
    } else {
        const float ph = phase_p(g, wi, ls.wo);
        f = mk(ph, ph, ph);
        scatPdf = ph;
    }
    const float w = power_heuristic(ls.pdf, scatPdf);
    // it.SpawnRay(ls.wo, ls.distance) (interaction.h:36-44)
    out.so = offset_ray_origin(p, face_same_hemisphere(n, ls.wo));
    out.stmax = ls.distance * (1.0f - 1e-3f);
    const f3 p2 = out.so + out.stmax * ls.wo;  // IntersectTr's target ray(ray.tmax)
    out.r0 = make_float4(f.x, f.y, f.z, w);
    out.r1 = make_float4(ls.Le.x, ls.Le.y, ls.Le.z, ls.pdf);
    out.r2 = make_float4(p2.x, p2.y, p2.z, ls.delta ? 1.0f : 0.0f);
    out.r3 = make_float4(ls.wo.x, ls.wo.y, ls.wo.z, pb);
    out.seg_med = medium_for(ls.wo, n, med_in, med_out);
    out.valid = true;
}

__device__ __forceinline__ void light_record_write(PathsDev& P, int32_t r, int32_t i, const LightRecord& L) {
    float4* rec = reinterpret_cast<float4*>(P.sh_rec + 16 * (int64_t)r);
    rec[0] = L.r0;
    rec[1] = L.r1;
    rec[2] = L.r2;
    rec[3] = L.r3;
    P.sh_tr[3 * r + 0] = P.sh_tr[3 * r + 1] = P.sh_tr[3 * r + 2] = 1.0f;
    P.shadow_path[r] = i;
    store_ray(&P.seg_rays[0][r], L.so, mk(L.r3.x, L.r3.y, L.r3.z), 0.0f, L.stmax);
    P.seg_rec[0][r] = r;
    P.seg_med[0][r] = L.seg_med;
}

// The vertex a direction will be sampled at (k_shade_gen_b), stored per path;
// the path joins the direction queue once the workgroup has claimed slots.
__device__ __forceinline__ void save_vertex(PathsDev& P, int32_t i, int kind, f3 p, f3 n, f3 wi, int mat,
                            int med_in, int med_out) {
    float* v = P.vtx + 9 * (int64_t)i;
    v[0] = p.x, v[1] = p.y, v[2] = p.z;
    v[3] = n.x, v[4] = n.y, v[5] = n.z;
    v[6] = wi.x, v[7] = wi.y, v[8] = wi.z;
    int4 info = make_int4(kind, mat, med_in, med_out);
    *reinterpret_cast<int4*>(P.vinfo + 4 * (int64_t)i) = info;
}

// Medium::Sample (medium.h:70-73) for every queued ray whose path is in a
// medium, ahead of k_shade_gen: the tracking loop (free flights and their
// dependent density lookups) runs in this lean kernel at several times the
// shading kernel's occupancy.  The event goes to the ray's queue slot and the
// sampler dimensions it consumed to the path, so k_shade_gen continues the
// path's sample sequence exactly where Sample left it.
template <int kS, int kM>
__global__ __launch_bounds__(256) void k_medium_track(SceneDev s, PathsDev P, int32_t n_active, int32_t lds_medium) {
    extern __shared__ float lds_maj[];
    const float* lds = stage_majorants(s, lds_medium, lds_maj);
    int lookups = 0;
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n_active;
         k += gridDim.x * blockDim.x) {
        const int32_t i = P.queue_in[k];
        const int med = P.medium[i];
        if (med < 0)
            continue;
        const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
        const float4 ra = r4[0], rb = r4[1];
        const pg_hit hit = P.hits[k];
        // Intersect leaves ray.tmax at the hit distance (bvh.cpp:652-709).
        const float tmax = hit.prim >= 0 ? hit.t : kFloatMax;
        SamplerState ss;
        ss.index = P.sidx[i];
        ss.dim = P.sdim[i];
        const int mi = kM == kMediaOneBrickGrid ? 0 : med;
        const MediumEvent ev = medium_sample<kS, kM>(s.media[mi], majorants(s, mi, lds_medium, lds), s.sampler,
                                                     ss, mk(ra.x, ra.y, ra.z), mk(rb.x, rb.y, rb.z), tmax, lookups);
        P.sdim[i] = ss.dim;
        P.mev[2 * (int64_t)k] = make_float4(ev.weight.x, ev.weight.y, ev.weight.z, ev.scattered ? 1.0f : 0.0f);
        P.mev[2 * (int64_t)k + 1] = make_float4(ev.p.x, ev.p.y, ev.p.z, 0.0f);
    }
    add_lookups(P.lookups, lookups);
}

// --- lane tracking ------------------------------------------------------------
// Delta and ratio tracking against brick majorants as persistent lanes
// (PG_RENDER_LANE_TRACKING): each lane takes the next ray of its workgroup's
// share whenever its own is finished, so a wave is not held by its longest
// path.  Measured on the 256^3 cloud it is 8% slower than one ray per thread
// (the branch for every phase runs predicated each step), so it is opt-in.  Each call of the step functions below does
// one piece of the brick walk of medium_sample / medium_tr (enter a brick, one
// free flight, leave a brick) with the same arithmetic and sampler draws, so
// the results are those of the per-path loops bit for bit.

// Gives each idle lane of the wave the next index of a work counter (one
// atomic per wave).  Every lane of the wave calls it; busy lanes get -1.  The
// counter is the workgroup's own, in LDS, over its contiguous share of the
// queue: one global counter for every wave of the chip serialises its atomics
// (measured: slower than no refilling at all).
__device__ __forceinline__ int32_t lanes_acquire(bool idle, int32_t* ctr) {
    const uint64_t need = __ballot(idle);
    if (need == 0)
        return -1;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)need) - 1;
    int32_t base = 0;
    if (lane == leader)
        base = atomicAdd(ctr, __popcll(need));
    base = __shfl(base, leader);
    return idle ? base + __popcll(need & ((1ull << lane) - 1ull)) : -1;
}

enum TrackPhase : int { kEnter = 0, kFlight = 1, kLeave = 2 };

struct LaneTrack {
    MRay r;
    BrickWalk w;
    float t, t1, texit, rate, inv_mj, scale;
    int phase;
};

// One step of the brick walk of a grid medium: returns 1 when the walk ends
// (left the grid or the ray's range), 2 when a free flight lands at r(t)
// (the caller tests that collision), 0 otherwise.
template <int kS>
__device__ __forceinline__ int track_step_bricks(const MediumDev& m, const float* maj, const SamplerDev& hs,
                                                 SamplerState& ss, LaneTrack& L) {
    if (L.phase == kEnter) {
        L.texit = fminf(brick_exit(L.w), L.t1);
        const float mj = maj[brick_index(m, L.w)];
        L.rate = m.sigma_t * mj;
        if (mj > 0.0f && L.rate > 0.0f) {
            L.inv_mj = 1.0f / mj;
            L.scale = brick_flight_scale(L.rate);
            L.phase = kFlight;
        } else {
            L.phase = kLeave;
        }
    }
    if (L.phase == kFlight) {
        L.t = brick_flight(L.t, track_uniform<kS>(hs, ss), L.scale);
        if (L.t < L.texit)
            return 2;
        L.phase = kLeave;
    }
    // kLeave
    if (L.texit >= L.t1)
        return 1;
    L.t = L.texit;
    if (!brick_advance(m, L.w))
        return 1;
    L.phase = kEnter;
    return 0;
}

// Medium::Sample of brick-tracked grid media as persistent lanes (see
// k_medium_track for what is read and written).
template <int kS, int kM>
__global__ __launch_bounds__(256) void k_medium_track_lanes(SceneDev s, PathsDev P, int32_t n_active,
                                                            int32_t lds_medium) {
    extern __shared__ float lds_maj[];
    const float* lds = stage_majorants(s, lds_medium, lds_maj);
    int lookups = 0;
    int32_t k = -1, i = 0, med = 0;
    SamplerState ss;
    LaneTrack L;
    // This workgroup's share of the queue, handed out through an LDS counter.
    __shared__ int32_t next;
    const int32_t share = (n_active + gridDim.x - 1) / gridDim.x;
    const int32_t end = min(n_active, (int32_t)blockIdx.x * share + share);
    if (threadIdx.x == 0)
        next = (int32_t)blockIdx.x * share;
    __syncthreads();
    bool exhausted = false;
    while (true) {
        if (!exhausted) {
            const bool idle = k < 0;
            const int32_t got = lanes_acquire(idle, &next);
            if (idle && got < end) {
                k = got;
                i = P.queue_in[k];
                med = P.medium[i];
                if (med < 0) {
                    k = -1;
                } else {
                    const int mi = kM == kMediaOneBrickGrid ? 0 : med;
                    const MediumDev& m = s.media[mi];
                    const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
                    const float4 ra = r4[0], rb = r4[1];
                    const pg_hit hit = P.hits[k];
                    // Intersect leaves ray.tmax at the hit distance (bvh.cpp:652-709).
                    const float tmax = hit.prim >= 0 ? hit.t : kFloatMax;
                    ss.index = P.sidx[i];
                    ss.dim = P.sdim[i];
                    L.r = to_model(m, mk(ra.x, ra.y, ra.z), mk(rb.x, rb.y, rb.z), 0.0f, tmax);
                    float t0, t1;
                    if (!unit_box_hit(L.r, t0, t1)) {
                        P.mev[2 * (int64_t)k] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
                        k = -1;
                    } else {
                        brick_walk_init(m, L.r, t0, L.w);
                        L.t = t0;
                        L.t1 = t1;
                        L.phase = kEnter;
                    }
                }
            }
            exhausted = __any(idle && got >= end);
        }
        if (!__any(k >= 0)) {
            if (exhausted)
                break;
            continue;
        }
        if (k < 0)
            continue;
        const int mi = kM == kMediaOneBrickGrid ? 0 : med;
        const MediumDev& m = s.media[mi];
        const int st = track_step_bricks<kS>(m, majorants(s, mi, lds_medium, lds), s.sampler, ss, L);
        bool scattered = false;
        f3 pm;
        if (st == 2) {
            pm = ray_at(L.r, L.t);
            lookups++;
            scattered = grid_density_fast(m, pm) * L.inv_mj > track_uniform<kS>(s.sampler, ss);
        }
        if (scattered) {
            const f3 albedo = m.sigma_s / m.sigma_t;
            const f3 p = to_world(m, pm);
            P.mev[2 * (int64_t)k] = make_float4(albedo.x, albedo.y, albedo.z, 1.0f);
            P.mev[2 * (int64_t)k + 1] = make_float4(p.x, p.y, p.z, 0.0f);
        } else if (st == 1) {
            P.mev[2 * (int64_t)k] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        }
        if (scattered || st == 1) {
            P.sdim[i] = ss.dim;
            k = -1;
        }
    }
    add_lookups(P.lookups, lookups);
}

// Medium::Tr of the shadow segments in brick-tracked grid media (ratio
// tracking with pine's Russian roulette, medium.cpp:55-94) as persistent
// lanes: the transmittance of segment k goes to seg_trf[k] (1 outside media),
// the sampler dimensions it used to its path.
template <int kS, int kM>
__global__ __launch_bounds__(256) void k_shadow_track_lanes(SceneDev s, PathsDev P, int32_t n_seg, int32_t src,
                                                            int32_t lds_medium) {
    extern __shared__ float lds_maj[];
    const float* lds = stage_majorants(s, lds_medium, lds_maj);
    int lookups = 0;
    int32_t k = -1, i = 0, med = 0;
    float tr = 1.0f;
    SamplerState ss;
    LaneTrack L;
    // This workgroup's share of the queue, handed out through an LDS counter.
    __shared__ int32_t next;
    const int32_t share = (n_seg + gridDim.x - 1) / gridDim.x;
    const int32_t end = min(n_seg, (int32_t)blockIdx.x * share + share);
    if (threadIdx.x == 0)
        next = (int32_t)blockIdx.x * share;
    __syncthreads();
    bool exhausted = false;
    while (true) {
        if (!exhausted) {
            const bool idle = k < 0;
            const int32_t got = lanes_acquire(idle, &next);
            if (idle && got < end) {
                k = got;
                med = P.seg_med[src][k];
                if (med < 0) {
                    P.seg_trf[k] = 1.0f;
                    k = -1;
                } else {
                    const int mi = kM == kMediaOneBrickGrid ? 0 : med;
                    const MediumDev& m = s.media[mi];
                    i = P.shadow_path[P.seg_rec[src][k]];
                    const float4* r4 = reinterpret_cast<const float4*>(&P.seg_rays[src][k]);
                    const float4 ra = r4[0], rb = r4[1];
                    const pg_hit hit = P.seg_hits[k];
                    const float tmax = hit.prim >= 0 ? hit.t : rb.w;
                    ss.index = P.sidx[i];
                    ss.dim = P.sdim[i];
                    L.r = to_model(m, mk(ra.x, ra.y, ra.z), mk(rb.x, rb.y, rb.z), 0.0f, tmax);
                    float t0, t1;
                    if (!unit_box_hit(L.r, t0, t1)) {
                        P.seg_trf[k] = 1.0f;
                        k = -1;
                    } else {
                        brick_walk_init(m, L.r, t0, L.w);
                        L.t = t0;
                        L.t1 = t1;
                        L.phase = kEnter;
                        tr = 1.0f;
                    }
                }
            }
            exhausted = __any(idle && got >= end);
        }
        if (!__any(k >= 0)) {
            if (exhausted)
                break;
            continue;
        }
        if (k < 0)
            continue;
        const int mi = kM == kMediaOneBrickGrid ? 0 : med;
        const MediumDev& m = s.media[mi];
        const int st = track_step_bricks<kS>(m, majorants(s, mi, lds_medium, lds), s.sampler, ss, L);
        bool done = st == 1;
        if (st == 2) {
            lookups++;
            const float density = grid_density_fast(m, ray_at(L.r, L.t));
            tr *= 1.0f - max0(density * L.inv_mj, 0.0f);
            if (tr < 0.1f) {
                const float q = max0(1.0f - tr, 0.05f);
                if (track_uniform<kS>(s.sampler, ss) < q) {
                    tr = 0.0f;
                    done = true;
                } else {
                    tr /= 1.0f - q;
                }
            }
        }
        if (done) {
            P.seg_trf[k] = tr;
            P.sdim[i] = ss.dim;
            k = -1;
        }
    }
    add_lookups(P.lookups, lookups);
}

// The four wavelengths' upsampling rows of path i (spectral mode): an RGB
// quantity of the scene becomes the path's 4-wavelength spectrum.
struct Lift4 {
    f3 up[4];
    __device__ __forceinline__ Lift4(const SceneDev& s, const PathsDev& P, int32_t i) {
        const float u = P.lam[i];
        for (int j = 0; j < 4; j++)
            up[j] = spec::upsample_row(s.spec, spec::wavelength(u, j));
    }
    __device__ __forceinline__ Spec4 operator()(f3 c, float scale) const {
        Spec4 r;
        for (int j = 0; j < 4; j++)
            r.v[j] = fmaxf(0.0f, dot(up[j], c)) * scale;
        return r;
    }
};

// kSpec: spectral mode (PG_RENDER_SPECTRAL), every colour of the path a
// 4-wavelength spectrum (L4, beta4), the scene's RGB factors upsampled; the
// geometry, sampling and sampler dimensions are those of the RGB bounce.
// kNodes: the scene has node-graph materials (the interpreter's code and
// registers only in that instantiation).
template <bool kSpec, bool kNodes>
__global__ __launch_bounds__(kSlotBlock) void k_shade_gen(SceneDev s, PathsDev P, int32_t n_active) {
    if ((int64_t)blockIdx.x * blockDim.x >= n_active)
        return;  // the whole workgroup
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    // One path per lane; the queue writes wait for the workgroup's slot claim.
    bool alive = false, vertex = false;
    f3 nextO, nextD;
    int32_t i = 0;
    int next_med = -1;
    LightRecord lrec;
    lrec.valid = false;
    if (k < n_active) {
        i = P.queue_in[k];
        const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
        const float4 ra = r4[0], rb = r4[1];
        const f3 ro = mk(ra.x, ra.y, ra.z), rd = mk(rb.x, rb.y, rb.z);
        const pg_hit hit = P.hits[k];
        const int depth = P.depth[i];
        f3 L, beta;
        Spec4 L4, beta4;
        if constexpr (kSpec) {
            L4 = spec_load(P.L4, i);
            beta4 = spec_load(P.beta4, i);
        } else {
            L = mk(P.L[3 * i], P.L[3 * i + 1], P.L[3 * i + 2]);
            beta = mk(P.beta[3 * i], P.beta[3 * i + 1], P.beta[3 * i + 2]);
        }
        // L += beta * c * w, in RGB as pine computes it, or per wavelength.
        auto add_light = [&](f3 c, float w, bool weighted) {
            if constexpr (kSpec) {
                const Spec4 cs = Lift4(s, P, i)(c, weighted ? w : 1.0f);
                for (int j = 0; j < 4; j++)
                    L4.v[j] += beta4.v[j] * cs.v[j];
            } else {
                L = weighted ? L + beta * c * w : L + beta * c;
            }
        };
        const float bsdfPdf = P.bsdf_pdf[i];
        SamplerState ss;
        ss.index = P.sidx[i];
        ss.dim = P.sdim[i];
        const int med = P.medium[i];
        next_med = med;
        MediumEvent ev;
        ev.scattered = false;
        if (med >= 0) {
            // k_medium_track's event for this slot (sdim above already follows it).
            const float4 e0 = P.mev[2 * (int64_t)k], e1 = P.mev[2 * (int64_t)k + 1];
            ev.weight = mk(e0.x, e0.y, e0.z);
            ev.scattered = e0.w != 0.0f;
            ev.p = mk(e1.x, e1.y, e1.z);
            if constexpr (kSpec) {
                const Spec4 ws = Lift4(s, P, i)(ev.weight, 1.0f);
                for (int j = 0; j < 4; j++)
                    beta4.v[j] *= ws.v[j];
            } else {
                beta = beta * ev.weight;
            }
        }
        if (ev.scattered) {
            if (depth + 1 != s.max_depth) {
                const f3 zero = mk(0.0f, 0.0f, 0.0f);
                light_record(s, ss, ev.p, zero, -rd, nullptr, s.media[med].g, med, med, lrec);
                save_vertex(P, i, 1, ev.p, zero, -rd, -1, med, med);
                vertex = true;
            }
        } else if (hit.prim < 0) {
            if (s.has_env) {
                const f3 le = env_color(s, rd);
                if (depth == 0)
                    add_light(le, 1.0f, false);
                else
                    add_light(le, power_heuristic(bsdfPdf, env_pdf(s)), true);
            }
        } else {
            const pg_hit_uv uv = P.uvs[k];
            Hit h = surface(s, hit.prim, hit.t, uv.u, uv.v, ro, rd);
            const ShapeDev& sh = s.shapes[h.shape];
            // bvh.cpp:702-707: a shape's own interface if it is a transition,
            // else the ray's medium on both sides.
            const bool transition = sh.mi_inside != sh.mi_outside;
            const int min_ = transition ? sh.mi_inside : med;
            const int mout = transition ? sh.mi_outside : med;
            if (sh.material < 0) {
                nextD = rd;
                nextO = offset_ray_origin(h.p, face_same_hemisphere(h.n, rd));
                next_med = medium_for(rd, h.n, min_, mout);
                alive = true;
            } else {
                // Node-graph inputs evaluated at this point (and the bump map
                // applied to h.n, path.cpp:48) into the path's own material
                // copy, which k_shade_gen_b samples after the shadow ray.
                const MaterialDev* mp = s.materials + sh.material;
                if (kNodes && mp->has_nodes)
                    mp = apply_nodes(s, P, i, mp, h, hit.prim, hit.t, uv.u, uv.v, ro, rd);
                const MaterialDev& m = *mp;
                if (m.kind == kMatEmissive) {
                    const f3 le = emitted(m, h.n, -rd);
                    if (depth == 0) {
                        add_light(le, 1.0f, false);
                    } else {
                        const float lightPdf = hit.t * hit.t / (abs_dot(-rd, h.n) * sh.area);
                        add_light(le, power_heuristic(bsdfPdf, lightPdf), true);
                    }
                } else if (depth + 1 != s.max_depth) {
                    light_record(s, ss, h.p, h.n, -rd, &m, 0.0f, min_, mout, lrec);
                    save_vertex(P, i, 0, h.p, h.n, -rd, sh.material, min_, mout);
                    vertex = true;
                }
            }
        }
        if constexpr (kSpec) {
            spec_store(P.L4, i, L4);
            spec_store(P.beta4, i, beta4);
        } else {
            P.L[3 * i + 0] = L.x;
            P.L[3 * i + 1] = L.y;
            P.L[3 * i + 2] = L.z;
            P.beta[3 * i + 0] = beta.x;
            P.beta[3 * i + 1] = beta.y;
            P.beta[3 * i + 2] = beta.z;
        }
        P.sdim[i] = ss.dim;
    }
    int32_t slot = 0, r = 0;
    block_slots2(alive, lrec.valid, P.counters, slot, r);
    const int32_t b = block_slot(vertex, &P.counters[2]);
    if (alive) {
        P.queue_out[slot] = i;
        P.medium[i] = next_med;
        store_ray(&P.rays_out[slot], nextO, nextD, 0.0f, kFloatMax);
    }
    if (lrec.valid)
        light_record_write(P, r, i, lrec);
    if (vertex)
        P.queue_b[b] = i;
}

// RandomWalkIntegrator::Li (randomwalk.cpp:6-62) with participating media:
// the ray's medium samples an event first (Medium::Sample, as in the Path
// integrator's split bounce; its weight multiplies beta either way); at a
// medium event the phase function picks the next direction (Get2D) unless the
// path is at maxDepth; emission and the environment are added with the
// throughput alone; material-less surfaces pass the ray into the medium on
// their far side without counting depth.
__global__ __launch_bounds__(kSlotBlock) void k_shade_rw_gen(SceneDev s, PathsDev P, int32_t n_active, int32_t lds_medium) {
    extern __shared__ float lds_maj[];
    const float* lds = stage_majorants(s, lds_medium, lds_maj);
    int lookups = 0;
    // The workgroup strides through the queue together, so that each round
    // claims its next-ray slots once (block_slot).
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n_active; base += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = (int32_t)(base + threadIdx.x);
        bool alive = false;
        f3 nextO, nextD;
        int32_t i = 0;
        int next_med = -1;
        if (k < n_active) {
        i = P.queue_in[k];
        const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
        const float4 ra = r4[0], rb = r4[1];
        const f3 ro = mk(ra.x, ra.y, ra.z), rd = mk(rb.x, rb.y, rb.z);
        const pg_hit hit = P.hits[k];
        int depth = P.depth[i];
        f3 L = mk(P.L[3 * i], P.L[3 * i + 1], P.L[3 * i + 2]);
        f3 beta = mk(P.beta[3 * i], P.beta[3 * i + 1], P.beta[3 * i + 2]);
        SamplerState ss;
        ss.index = P.sidx[i];
        ss.dim = P.sdim[i];
        const int med = P.medium[i];
        next_med = med;
        const float tmax = hit.prim >= 0 ? hit.t : kFloatMax;
        MediumEvent ev;
        ev.scattered = false;
        if (med >= 0) {
            ev = medium_sample(s.media[med], majorants(s, med, lds_medium, lds), s.sampler, ss, ro, rd, tmax,
                               lookups);
            beta = beta * ev.weight;
        }
        if (ev.scattered) {
            if (depth + 1 != s.max_depth) {
                float u0, u1;
                get_2d(s.sampler, ss, u0, u1);
                f3 wo;
                phase_sample(s.media[med].g, -rd, u0, u1, wo);
                // MediumInteraction::SpawnRay: the normal is zero, no offset.
                const f3 zero = mk(0.0f, 0.0f, 0.0f);
                nextD = wo;
                nextO = offset_ray_origin(ev.p, face_same_hemisphere(zero, wo));
                next_med = med;
                alive = true;
                depth++;
            }
        } else if (hit.prim < 0) {
            if (s.has_env)
                L = L + beta * env_color(s, rd);
        } else {
            const pg_hit_uv uv = P.uvs[k];
            Hit h = surface(s, hit.prim, hit.t, uv.u, uv.v, ro, rd);
            const ShapeDev& sh = s.shapes[h.shape];
            const bool transition = sh.mi_inside != sh.mi_outside;
            const int min_ = transition ? sh.mi_inside : med;
            const int mout = transition ? sh.mi_outside : med;
            if (sh.material < 0) {
                nextD = rd;
                nextO = offset_ray_origin(h.p, face_same_hemisphere(h.n, rd));
                next_med = medium_for(rd, h.n, min_, mout);
                alive = true;
            } else {
                const MaterialDev* mp = s.materials + sh.material;
                if (mp->has_nodes)
                    mp = apply_nodes(s, P, k, mp, h, hit.prim, hit.t, uv.u, uv.v, ro, rd);
                const MaterialDev& m = *mp;
                const f3 wi = -rd;
                if (m.kind == kMatEmissive) {
                    L = L + beta * emitted(m, h.n, wi);
                } else if (depth + 1 < s.max_depth) {
                    const M3 n2w = frame_from(h.n);
                    const M3 w2n = inverse(n2w);
                    const float u1 = get_1d(s.sampler, ss);
                    float u2a, u2b;
                    get_2d(s.sampler, ss, u2a, u2b);
                    BSDFSample bs;
                    if (material_sample(m, mul(w2n, wi), u1, u2a, u2b, bs)) {
                        bs.wo = mul(n2w, bs.wo);
                        beta = beta * (abs_dot(bs.wo, h.n) * bs.f / bs.pdf);
                        nextD = bs.wo;
                        nextO = offset_ray_origin(h.p, face_same_hemisphere(h.n, nextD));
                        next_med = medium_for(nextD, h.n, min_, mout);
                        alive = true;
                    }
                    depth++;
                }
            }
        }
        P.L[3 * i + 0] = L.x;
        P.L[3 * i + 1] = L.y;
        P.L[3 * i + 2] = L.z;
        P.beta[3 * i + 0] = beta.x;
        P.beta[3 * i + 1] = beta.y;
        P.beta[3 * i + 2] = beta.z;
        P.sdim[i] = ss.dim;
        P.depth[i] = depth;
        }
        const int32_t slot = block_slot(alive, &P.counters[0]);
        if (alive) {
            P.queue_out[slot] = i;
            P.medium[i] = next_med;
            store_ray(&P.rays_out[slot], nextO, nextD, 0.0f, kFloatMax);
        }
    }
    add_lookups(P.lookups, lookups);
}

// One segment of IntersectTr (integrator.cpp:68-87): transmittance of the
// segment's medium up to the closest hit, then stop at a surface with a
// material, or continue towards p2 through a material-less one.
// kPre: the segments' transmittance was tracked by k_shadow_track_lanes
// (seg_trf); otherwise it is tracked here, one segment per thread.
// d_nseg: the pass's segment count in device memory (NULL: n_seg); the next
// pass's segments are counted at next_ctr (NULL: P.counters[3]).
template <int kS, int kM, bool kPre>
__global__ __launch_bounds__(256) void k_shadow_tr(SceneDev s, PathsDev P, int32_t n_seg, int32_t src, int32_t lds_medium,
                                                   const int32_t* d_nseg, int32_t* next_ctr) {
    extern __shared__ float lds_maj[];
    const float* lds = kPre ? nullptr : stage_majorants(s, lds_medium, lds_maj);
    const int dst = 1 - src;
    const int32_t nseg = d_nseg ? *d_nseg : n_seg;
    int32_t* const ctr = next_ctr ? next_ctr : &P.counters[3];
    int lookups = 0;
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nseg; k += gridDim.x * blockDim.x) {
        const int32_t r = P.seg_rec[src][k];
        const int32_t i = P.shadow_path[r];
        const float4* r4 = reinterpret_cast<const float4*>(&P.seg_rays[src][k]);
        const float4 ra = r4[0], rb = r4[1];
        const f3 ro = mk(ra.x, ra.y, ra.z), rd = mk(rb.x, rb.y, rb.z);
        const pg_hit hit = P.seg_hits[k];
        const int med = P.seg_med[src][k];
        f3 tr = mk(P.sh_tr[3 * r], P.sh_tr[3 * r + 1], P.sh_tr[3 * r + 2]);
        if (kPre && med >= 0) {
            const float f = P.seg_trf[k];
            tr = tr * mk(f, f, f);
        } else if (med >= 0) {
            SamplerState ss;
            ss.index = P.sidx[i];
            ss.dim = P.sdim[i];
            const float tmax = hit.prim >= 0 ? hit.t : rb.w;
            const int mi = kM == kMediaOneBrickGrid ? 0 : med;
            tr = tr * medium_tr<kS, kM>(s.media[mi], majorants(s, mi, lds_medium, lds), s.sampler, ss, ro,
                                        rd, tmax, lookups);
            P.sdim[i] = ss.dim;
        }
        if (hit.prim >= 0) {
            const int shape = s.prim_shape[hit.prim];
            const ShapeDev& sh = s.shapes[shape];
            if (sh.material >= 0) {
                tr = mk(0.0f, 0.0f, 0.0f);
            } else {
                const pg_hit_uv uv = P.seg_uvs[k];
                const Hit h = surface(s, hit.prim, hit.t, uv.u, uv.v, ro, rd);
                const bool transition = sh.mi_inside != sh.mi_outside;
                const int min_ = transition ? sh.mi_inside : med;
                const int mout = transition ? sh.mi_outside : med;
                // Interaction::SpawnRayTo(p2) (interaction.h:27-35)
                const float* rec = P.sh_rec + 16 * (int64_t)r;
                const f3 p2 = mk(rec[8], rec[9], rec[10]);
                float len;
                const f3 d2 = normalize(p2 - h.p, len);
                const f3 o2 = offset_ray_origin(h.p, face_same_hemisphere(h.n, d2));
                const int slot = atomicAdd(ctr, 1);
                store_ray(&P.seg_rays[dst][slot], o2, d2, 0.0f, len * (1.0f - 1e-3f));
                P.seg_rec[dst][slot] = r;
                P.seg_med[dst][slot] = medium_for(d2, h.n, min_, mout);
            }
        }
        P.sh_tr[3 * r + 0] = tr.x;
        P.sh_tr[3 * r + 1] = tr.y;
        P.sh_tr[3 * r + 2] = tr.z;
    }
    add_lookups(P.lookups, lookups);
}

// EstimateDirect's tail (integrator.cpp:101-118) and `L += beta * ...`.
// At a guided vertex the directional strategy is the guide/BSDF mixture, so
// the light sample's MIS weight uses the mixture pdf of the light direction.
// d_n: the record count in device memory (NULL: n_rec).  ovf: when it holds
// a nonzero count (shadow segments left after the device-driven passes), the
// kernel does nothing; the host finishes the segments and launches it again.
template <bool kSpec>
__global__ __launch_bounds__(256) void k_resolve_gen(SceneDev s, PathsDev P, int32_t n_rec, const int32_t* d_n,
                                                     const int32_t* ovf) {
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (d_n ? *d_n : n_rec) || (ovf && *ovf != 0))
        return;
    const f3 tr = mk(P.sh_tr[3 * r], P.sh_tr[3 * r + 1], P.sh_tr[3 * r + 2]);
    if (is_black(tr))
        return;
    const int32_t i = P.shadow_path[r];
    const float4* rec = reinterpret_cast<const float4*>(P.sh_rec + 16 * (int64_t)r);
    const float4 a = rec[0], b = rec[1], c = rec[2];
    const f3 f = mk(a.x, a.y, a.z), le = mk(b.x, b.y, b.z);
    const float pdf = b.w;
    float w = a.w;
    if (s.guided && P.gflag[i] && c.w == 0.0f) {
        const float4 d = rec[3];
        const float* gp = P.gparams + guide::kParamStride * (int64_t)i;
        const float alpha = gp[guide::kAlphaAt];
        const float pmix = alpha * guide::mix_pdf(gp, mk(d.x, d.y, d.z)) + (1.0f - alpha) * d.w;
        w = power_heuristic(pdf, pmix);
    }
    if constexpr (kSpec) {
        // Transmittance, BSDF or phase value and emission each upsampled at the
        // path's wavelengths, multiplied per wavelength.
        const Lift4 lift(s, P, i);
        const Spec4 ts = lift(tr, 1.0f), fs = lift(f, 1.0f), ls = lift(le, (c.w != 0.0f ? 1.0f : w) / pdf);
        const Spec4 beta4 = spec_load(P.beta4, i);
        Spec4 L4 = spec_load(P.L4, i);
        for (int j = 0; j < 4; j++)
            L4.v[j] += beta4.v[j] * ts.v[j] * fs.v[j] * ls.v[j];
        spec_store(P.L4, i, L4);
    } else {
        const f3 ld = c.w != 0.0f ? tr * f * le / pdf : tr * f * w * le / pdf;
        const f3 beta = mk(P.beta[3 * i], P.beta[3 * i + 1], P.beta[3 * i + 2]);
        const f3 add = beta * ld;
        P.L[3 * i + 0] += add.x;
        P.L[3 * i + 1] += add.y;
        P.L[3 * i + 2] += add.z;
    }
}

// Direction sampling after the shadow ray: PhaseFunction::Sample at a medium
// vertex (path.cpp:19-25), the material's Sample + Russian roulette at a
// surface (path.cpp:66-86).
// d_n, ovf: as for k_resolve_gen (the direction queue's count).
template <bool kSpec>
__global__ __launch_bounds__(kSlotBlock) void k_shade_gen_b(SceneDev s, PathsDev P, int32_t n_b, const int32_t* d_n,
                                                            const int32_t* ovf) {
    const int32_t n = d_n ? *d_n : n_b;
    if ((int64_t)blockIdx.x * blockDim.x >= n || (ovf && *ovf != 0))
        return;  // the whole workgroup
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    f3 nextO, nextD;
    int next_med = -1, i = 0;
    if (k < n) {
    i = P.queue_b[k];
    const float* v = P.vtx + 9 * (int64_t)i;
    const f3 p = mk(v[0], v[1], v[2]), n = mk(v[3], v[4], v[5]), wi = mk(v[6], v[7], v[8]);
    const int4 info = *reinterpret_cast<const int4*>(P.vinfo + 4 * (int64_t)i);
    f3 beta;
    Spec4 beta4;
    if constexpr (kSpec)
        beta4 = spec_load(P.beta4, i);
    else
        beta = mk(P.beta[3 * i], P.beta[3 * i + 1], P.beta[3 * i + 2]);
    SamplerState ss;
    ss.index = P.sidx[i];
    ss.dim = P.sdim[i];
    int depth = P.depth[i];
    if (info.x == 1) {
        float u0, u1;
        get_2d(s.sampler, ss, u0, u1);
        f3 wo;
        P.bsdf_pdf[i] = phase_sample(s.media[info.z].g, wi, u0, u1, wo);
        nextD = wo;
        nextO = offset_ray_origin(p, face_same_hemisphere(n, wo));
        next_med = info.z;
        alive = true;
    } else if (!kSpec && s.guided && P.gflag[i]) {
        f3 wo;
        if (guided_direction(s, P, ss, i, depth, P.gparams + guide::kParamStride * (int64_t)i,
                             s.materials[info.y].has_nodes ? P.node_mat[i] : s.materials[info.y], p, n, wi,
                             mk(P.L[3 * i], P.L[3 * i + 1], P.L[3 * i + 2]), beta, wo)) {
            nextD = wo;
            nextO = offset_ray_origin(p, face_same_hemisphere(n, nextD));
            next_med = medium_for(nextD, n, info.z, info.w);
            alive = true;
            if (depth > s.rr_depth) {
                const float q = pstd::clamp(1.0f - luminance(beta), 0.05f, 1.0f);
                if (get_1d(s.sampler, ss) < q)
                    alive = false;
                else
                    beta = beta / (1.0f - q);
            }
        }
    } else {
        // The material as k_shade_gen evaluated it at this vertex (its
        // node-graph inputs in the path's own copy).
        const MaterialDev& m = s.materials[info.y].has_nodes ? P.node_mat[i] : s.materials[info.y];
        const M3 n2w = frame_from(n);
        const M3 w2n = inverse(n2w);
        const f3 wiL = mul(w2n, wi);
        const float u1 = get_1d(s.sampler, ss);
        float u2a, u2b;
        get_2d(s.sampler, ss, u2a, u2b);
        BSDFSample bs;
        if (material_sample(m, wiL, u1, u2a, u2b, bs)) {
            bs.wo = mul(n2w, bs.wo);
            if (!is_black(bs.f) && bs.pdf != 0.0f) {
                if constexpr (kSpec) {
                    const Spec4 fs = Lift4(s, P, i)(bs.f, abs_dot(bs.wo, n) / bs.pdf);
                    for (int j = 0; j < 4; j++)
                        beta4.v[j] *= fs.v[j];
                } else {
                    beta = beta * (abs_dot(bs.wo, n) * bs.f / bs.pdf);
                }
                P.bsdf_pdf[i] = bs.pdf;
                nextD = bs.wo;
                nextO = offset_ray_origin(p, face_same_hemisphere(n, nextD));
                next_med = medium_for(nextD, n, info.z, info.w);
                alive = true;
                if (depth > s.rr_depth) {
                    if constexpr (kSpec) {
                        // Russian roulette on the mean throughput (as k_shade_spec).
                        const float mean = 0.25f * (beta4.v[0] + beta4.v[1] + beta4.v[2] + beta4.v[3]);
                        const float q = pstd::clamp(1.0f - mean, 0.05f, 1.0f);
                        if (get_1d(s.sampler, ss) < q) {
                            alive = false;
                        } else {
                            for (int j = 0; j < 4; j++)
                                beta4.v[j] /= (1.0f - q);
                        }
                    } else {
                        const float q = pstd::clamp(1.0f - luminance(beta), 0.05f, 1.0f);
                        if (get_1d(s.sampler, ss) < q)
                            alive = false;
                        else
                            beta = beta / (1.0f - q);
                    }
                }
            }
        }
    }
    depth++;
    if constexpr (kSpec) {
        spec_store(P.beta4, i, beta4);
    } else {
        P.beta[3 * i + 0] = beta.x;
        P.beta[3 * i + 1] = beta.y;
        P.beta[3 * i + 2] = beta.z;
    }
    P.sdim[i] = ss.dim;
    P.depth[i] = depth;
    }
    const int32_t slot = block_slot(alive, &P.counters[0]);
    if (alive) {
        P.queue_out[slot] = i;
        P.medium[i] = next_med;
        store_ray(&P.rays_out[slot], nextO, nextD, 0.0f, kFloatMax);
    }
}


// --- learned path guiding -----------------------------------------------------

// Mixture parameters of every guided vertex: 16 vertices per wave through the
// register-chained MFMA network, all weights staged in LDS once per workgroup
// (59 KB: two 8-wave workgroups per CU), the head computed from the output
// registers.  kSrc 0: the split bounce's direction-sampling queue (queue_b,
// vtx, vinfo; records and gflag per path); kSrc 1: the fused bounce's queue
// (inputs and flags from k_guide_gather; records per queue slot).  Vertices
// the guide does not serve (misses, emitters, non-diffuse materials, the last
// depth, depth >= g_depth) have gflag 0.  Waves are independent after the
// staging barrier: a grid-stride loop over tiles of 16 per wave.
// Fused bounce, before inference: the network input of every hit that will
// sample a direction at a vertex the guide serves, in queue order (gx, gflag
// per queue slot).  Its dependent loads (queue, hit, shape, triangle,
// material) run here at full occupancy rather than in the network kernel.
__global__ __launch_bounds__(256) void k_guide_gather(SceneDev s, PathsDev P, int32_t n_in, const int32_t* d_n) {
    const int32_t n = d_n ? *d_n : n_in;
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int32_t i = P.queue_in[k];
        const pg_hit hit = P.hits[k];
        int use = 0;
        if (hit.prim >= 0 && P.depth[i] + 1 < s.max_depth && P.depth[i] < s.g_depth) {
            const float4* r4 = reinterpret_cast<const float4*>(&P.rays_in[k]);
            const float4 ra = r4[0], rb = r4[1];
            const f3 ro = mk(ra.x, ra.y, ra.z), rd = mk(rb.x, rb.y, rb.z);
            const pg_hit_uv uv = P.uvs[k];
            const Hit h = surface(s, hit.prim, hit.t, uv.u, uv.v, ro, rd);
            const int mat = s.shapes[h.shape].material;
            if (mat >= 0 && guidable(s, make_int4(0, mat, 0, 0))) {
                use = 1;
                guide_input(s, h.p, h.n, -rd, P.gx + guide::kIn * (int64_t)k);
            }
        }
        P.gflag[k] = use;
    }
}

constexpr int kInferThreads = 512;
template <int kSrc>
__global__ __launch_bounds__(kInferThreads) void k_guide_infer(SceneDev s, PathsDev P, int32_t n_in,
                                                               const int32_t* d_n) {
    using namespace guide;
    const int32_t n = d_n ? *d_n : n_in;
    if ((int64_t)blockIdx.x * (kInferThreads / 4) >= n)
        return;  // whole workgroup: no tile of it is in range, no barrier reached
    constexpr int kLd = kH + 8, kLd0 = kFeat + 8;
    __shared__ __attribute__((aligned(16))) __bf16 sw[(2 * kH + kOutPad) * kLd];
    __shared__ __attribute__((aligned(16))) __bf16 sw0[kH * kLd0];
    const int tid = threadIdx.x, l = tid & 63, r = l & 15, g = l >> 4;
    for (int k = tid; k < (2 * kH * kH + kOutPad * kH) / 8; k += kInferThreads) {
        const int row = k / (kH / 8), c8 = k % (kH / 8);
        *reinterpret_cast<bf16x8*>(sw + row * kLd + 8 * c8) =
            *reinterpret_cast<const bf16x8*>(s.g_wf + WF1 + row * kH + 8 * c8);
    }
    for (int k = tid; k < kH * kFeat / 8; k += kInferThreads) {
        const int row = k / (kFeat / 8), c8 = k % (kFeat / 8);
        *reinterpret_cast<bf16x8*>(sw0 + row * kLd0 + 8 * c8) =
            *reinterpret_cast<const bf16x8*>(s.g_wf + WF0 + row * kFeat + 8 * c8);
    }
    __syncthreads();
    const int32_t waves = gridDim.x * (kInferThreads / 64);
    for (int32_t tile0 = 16 * (blockIdx.x * (kInferThreads / 64) + (tid >> 6)); tile0 < n; tile0 += 16 * waves) {
        // The weight operands' LDS addresses do not depend on the tile: keep
        // the compiler from hoisting all 56 of them into registers (224 VGPRs,
        // then spills) by making each iteration's loads opaque.
        asm volatile("" ::: "memory");
        const int32_t k = tile0 + r;
        int32_t i = -1;
        bool use = false;
        float x[kIn] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (k < n) {
            if constexpr (kSrc == 1) {
                i = k;
                use = P.gflag[k] != 0;
                if (use)
                    for (int c = 0; c < kIn; c++)
                        x[c] = P.gx[kIn * (int64_t)k + c];
            } else {
                i = P.queue_b[k];
                const int4 info = *reinterpret_cast<const int4*>(P.vinfo + 4 * (int64_t)i);
                use = P.depth[i] < s.g_depth && guidable(s, info);
                if (use) {
                    const float* v = P.vtx + 9 * (int64_t)i;
                    guide_input(s, mk(v[0], v[1], v[2]), mk(v[3], v[4], v[5]), mk(v[6], v[7], v[8]), x);
                }
            }
        }
        f32x4 o[3];
        forward16(x, sw0, kLd0, sw, sw + kH * kLd, sw + 2 * kH * kLd, kLd, s.g_params, s.g_oct, o, nullptr,
                  nullptr, nullptr);
        if (kSrc == 0 && g == 0 && i >= 0)
            P.gflag[i] = use ? 1 : 0;
        head4(o, g, use, P.gparams + kParamStride * (int64_t)(i >= 0 ? i : 0));
    }
}

// Training weights of the recorded vertices once their paths are complete:
// incident radiance along the sampled direction, (L_final - L_before) /
// beta_after in luminance, times the BSDF pdf (for the diffuse vertices the
// guide serves, proportional to f cos: the guide learns the product), over the
// pdf the direction was sampled with.
__global__ void k_guide_finalize(PathsDev P, int64_t n_slots) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= n_slots)
        return;
    float w = 0.0f;
    const float pdf = P.rec_pdf[slot];
    if (pdf > 0.0f) {
        const int64_t i = slot / P.rec_depth;
        const f3 dl = mk(P.L[3 * i] - P.rec_lb[3 * slot], P.L[3 * i + 1] - P.rec_lb[3 * slot + 1],
                         P.L[3 * i + 2] - P.rec_lb[3 * slot + 2]);
        const float b = luminance(mk(P.rec_beta[3 * slot], P.rec_beta[3 * slot + 1],
                                     P.rec_beta[3 * slot + 2]));
        const float lin = b > 0.0f ? luminance(dl) / b : 0.0f;
        const float v = lin * P.rec_pb[slot] / pdf;
        w = lin > 0.0f && isfinite(v) ? v : 0.0f;
    }
    P.rec_w[slot] = w;
}

// Training-set compaction: the records with a nonzero weight, in slot order,
// with weights relative to their batch mean and clamped at kWeightClamp (a
// handful of high-radiance paths must not decide a whole step).  Every sum
// runs in a fixed order, so a batch trains to the same bits on every run.
constexpr float kWeightClamp = 20.0f;

// Per block of 256 slots: count and sum of the nonzero weights.
__global__ __launch_bounds__(256) void k_rec_count(const float* w, int64_t n, int32_t* bcnt, float* bsum) {
    __shared__ float ss[4];
    __shared__ int32_t sc[4];
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const float v = k < n ? w[k] : 0.0f;
    float s = v > 0.0f ? v : 0.0f;
    int32_t c = v > 0.0f ? 1 : 0;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off, 64);
        c += __shfl_down(c, off, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        ss[wave] = s;
        sc[wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = ((ss[0] + ss[1]) + ss[2]) + ss[3];
        bcnt[blockIdx.x] = sc[0] + sc[1] + sc[2] + sc[3];
    }
}

// One workgroup: exclusive scan of the block counts into boff, the total
// count into n_out and the mean weight into acc[0].
__global__ __launch_bounds__(1024) void k_rec_scan(const int32_t* bcnt, const float* bsum, int32_t nblk,
                                                   int32_t* boff, float* acc, int32_t* n_out) {
    __shared__ int32_t sc[1024];
    __shared__ float ss[1024];
    const int t = threadIdx.x;
    const int32_t chunk = (nblk + 1023) / 1024;
    const int32_t b0 = t * chunk, b1 = min(nblk, b0 + chunk);
    int32_t c = 0;
    float s = 0.0f;
    for (int32_t b = b0; b < b1; b++) {
        c += bcnt[b];
        s += bsum[b];
    }
    sc[t] = c;
    ss[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan of the counts
        const int32_t add = t >= off ? sc[t - off] : 0;
        __syncthreads();
        sc[t] += add;
        __syncthreads();
    }
    for (int off = 512; off > 0; off >>= 1) {  // tree sum of the weights
        if (t < off)
            ss[t] += ss[t + off];
        __syncthreads();
    }
    int32_t run = sc[t] - c;
    for (int32_t b = b0; b < b1; b++) {
        boff[b] = run;
        run += bcnt[b];
    }
    if (t == 0) {
        const int32_t total = sc[1023];
        n_out[0] = total;
        acc[0] = total > 0 ? ss[0] / (float)total : 0.0f;
    }
}

// Writes each nonzero record to its place in the compacted training set.
__global__ __launch_bounds__(256) void k_rec_scatter(PathsDev P, int64_t n) {
    __shared__ int32_t wc[4];
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const float v = k < n ? P.rec_w[k] : 0.0f;
    const bool keep = v > 0.0f;
    const unsigned long long m = __ballot(keep);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        wc[wave] = __popcll(m);
    __syncthreads();
    if (!keep)
        return;
    int32_t o = P.rec_boff[blockIdx.x] + __popcll(m & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; w++)
        o += wc[w];
    const float r = v / P.rec_acc[0];
    P.c_w[o] = r < kWeightClamp ? r : kWeightClamp;
    P.c_pb[o] = P.rec_pb[k];
    for (int c = 0; c < guide::kIn; c++)
        P.c_x[guide::kIn * (int64_t)o + c] = P.rec_x[guide::kIn * k + c];
    for (int c = 0; c < 3; c++)
        P.c_dir[3 * (int64_t)o + c] = P.rec_dir[3 * k + c];
}

__global__ void k_fill_f32(float* p, int64_t n, float v) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        p[k] = v;
}

// Film::AddSample (film.h:26-45) for every path of a pixel, in sample order,
// for filters of radius at most 1/2.  A workgroup takes G consecutive pixels
// (their samples are consecutive paths) and stages their samples' own-pixel
// contributions through LDS, kFilmTile at a time, so the path reads are
// coalesced; each pixel's thread then adds its samples to the film value in
// sample order, as pine's AddSample does.  A sample exactly on a pixel border
// also feeds a neighbour, whose own pass may be running: the pixel records the
// range of such samples and k_film_border adds them afterwards.
constexpr int kFilmTile = 2048;

// Pixels per k_film workgroup: at least 64 (a wave of pixel threads for the
// in-order sums), all of a tile's samples when a pixel has few.
int film_pixels_per_block(int spb) {
    return std::max(64, std::min(256, kFilmTile / std::max(spb, 1)));
}

// G pixels per workgroup; each tile holds C consecutive samples of each of
// them (C = kFilmTile / G, at most spb), so that G threads, one per pixel, add
// their pixel's samples in sample order, chunk after chunk.
__global__ __launch_bounds__(256) void k_film(SceneDev s, PathsDev P, int32_t n_pixels, int32_t spb, int32_t G,
                                              float* film) {
    __shared__ float4 contrib[kFilmTile];
    __shared__ uint8_t kind[kFilmTile];  // bit 0: the own pixel is in the footprint; bit 1: another pixel is
    const int32_t p0 = blockIdx.x * G;
    const int32_t np = min(G, n_pixels - p0);
    const int C = min(spb, kFilmTile / G);
    const int j = threadIdx.x;
    const bool owner = j < np;
    const float rx = s.filter_radius[0], ry = s.filter_radius[1];
    size_t own = 0;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int32_t first = -1, last = -1;
    if (owner) {
        const int32_t pixel = p0 + j;
        const int px = pixel % s.width, py = pixel / s.width;
        own = (size_t)(s.height - 1 - py) * s.width + px;
        for (int c = 0; c < 4; c++)
            acc[c] = film[4 * own + c];
    }
    for (int s0 = 0; s0 < spb; s0 += C) {
        const int cs = min(C, spb - s0);  // samples of each pixel in this tile
        __syncthreads();                  // the previous tile has been read
        for (int t = j; t < np * cs; t += 256) {
            const int pl = t / cs, sl = s0 + (t - pl * cs);
            const int32_t q = p0 + pl;
            const int64_t i = (int64_t)q * spb + sl;
            const int qx = q % s.width, qy = q / s.width;
            const float fx = P.pfilm[2 * i] * (float)s.width - 0.5f;
            const float fy = P.pfilm[2 * i + 1] * (float)s.height - 0.5f;
            int x0 = (int)pstd::ceil(fx - rx), y0 = (int)pstd::ceil(fy - ry);
            int x1 = (int)pstd::floor(fx + rx), y1 = (int)pstd::floor(fy + ry);
            x0 = x0 > 0 ? x0 : 0;
            y0 = y0 > 0 ? y0 : 0;
            x1 = x1 < s.width - 1 ? x1 : s.width - 1;
            y1 = y1 < s.height - 1 ? y1 : s.height - 1;
            uint8_t k = 0;
            float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (qx >= x0 && qx <= x1 && qy >= y0 && qy <= y1) {
                const float dx = fabsp((float)qx - fx) / rx, dy = fabsp((float)qy - fy) / ry;
                const int tx = (int)(16.0f * (dx < kOneMinusEps ? dx : kOneMinusEps));
                const int ty = (int)(16.0f * (dy < kOneMinusEps ? dy : kOneMinusEps));
                const float w = s.filter_table[ty * 16 + tx];
                c = make_float4(P.L[3 * i] * w, P.L[3 * i + 1] * w, P.L[3 * i + 2] * w, w);
                k = 1;
            }
            if (x0 <= x1 && y0 <= y1 && !(x0 == x1 && y0 == y1 && x0 == qx && y0 == qy))
                k |= 2;
            contrib[t] = c;
            kind[t] = k;
        }
        __syncthreads();
        if (owner) {
            // The pixel's cs samples of this tile, in sample order; the LDS
            // reads do not depend on the sums, so they are issued ahead.
            const int t0 = j * cs;
#pragma unroll 4
            for (int t = 0; t < cs; t++) {
                const uint8_t k = kind[t0 + t];
                const float4 c = contrib[t0 + t];
                if (k & 1) {
                    acc[0] += c.x;
                    acc[1] += c.y;
                    acc[2] += c.z;
                    acc[3] += c.w;
                }
                if (k & 2) {
                    last = s0 + t;
                    if (first < 0)
                        first = last;
                }
            }
        }
    }
    if (owner) {
        for (int c = 0; c < 4; c++)
            film[4 * own + c] = acc[c];
        P.border[p0 + j] = make_int2(first, last);
    }
}

// The border samples k_film left out, as a gather: each film pixel visits its
// (at most eight) neighbours' recorded ranges of border samples, finds those
// whose [Ceil(pf - r), Floor(pf + r)] footprint (computed as k_film computes
// it) covers it, and adds them in a fixed order (neighbour rows, columns,
// samples), so the film does not depend on scheduling.  Filters of radius at
// most 1/2: a sample lies within half a pixel (plus rounding) of its own
// pixel's centre, so its footprint reaches one pixel at most.
__global__ __launch_bounds__(256) void k_film_border(SceneDev s, PathsDev P, int32_t n_pixels, int32_t spb,
                                                     float* film) {
    const int32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= n_pixels)
        return;
    const int x = pixel % s.width, y = pixel / s.width;
    const int qy0 = y > 0 ? y - 1 : 0, qy1 = y < s.height - 1 ? y + 1 : y;
    const int qx0 = x > 0 ? x - 1 : 0, qx1 = x < s.width - 1 ? x + 1 : x;
    bool any = false;
    for (int qy = qy0; qy <= qy1; qy++)
        for (int qx = qx0; qx <= qx1; qx++)
            any |= (qx != x || qy != y) && P.border[qy * s.width + qx].x >= 0;
    if (!any)
        return;
    const float rx = s.filter_radius[0], ry = s.filter_radius[1];
    const size_t own = (size_t)(s.height - 1 - y) * s.width + x;
    float acc[4] = {film[4 * own + 0], film[4 * own + 1], film[4 * own + 2], film[4 * own + 3]};
    for (int qy = qy0; qy <= qy1; qy++)
        for (int qx = qx0; qx <= qx1; qx++) {
            const int32_t q = qy * s.width + qx;
            const int2 r = P.border[q];
            if ((qx == x && qy == y) || r.x < 0)
                continue;
            for (int sl = r.x; sl <= r.y; sl++) {
                const int64_t i = (int64_t)q * spb + sl;
                const float fx = P.pfilm[2 * i] * (float)s.width - 0.5f;
                const float fy = P.pfilm[2 * i + 1] * (float)s.height - 0.5f;
                const int x0 = (int)pstd::ceil(fx - rx), y0 = (int)pstd::ceil(fy - ry);
                const int x1 = (int)pstd::floor(fx + rx), y1 = (int)pstd::floor(fy + ry);
                if (x < x0 || x > x1 || y < y0 || y > y1)
                    continue;
                const float dx = fabsp((float)x - fx) / rx, dy = fabsp((float)y - fy) / ry;
                const int tx = (int)(16.0f * (dx < kOneMinusEps ? dx : kOneMinusEps));
                const int ty = (int)(16.0f * (dy < kOneMinusEps ? dy : kOneMinusEps));
                const float w = s.filter_table[ty * 16 + tx];
                acc[0] += P.L[3 * i] * w;
                acc[1] += P.L[3 * i + 1] * w;
                acc[2] += P.L[3 * i + 2] * w;
                acc[3] += w;
            }
        }
    film[4 * own + 0] = acc[0];
    film[4 * own + 1] = acc[1];
    film[4 * own + 2] = acc[2];
    film[4 * own + 3] = acc[3];
}

// Film::AddSample (film.h:26-45) for filters wider than a pixel, as a gather:
// each film pixel visits the samples of the pixels within reach (a sample
// lies inside its own pixel, so it can touch pixels up to ceil(radius + 1/2)
// away), tests whether it is inside that sample's [Ceil(pf - r), Floor(pf + r)]
// footprint exactly as AddSample computes it, and adds the filter-table
// weight.  Fixed order (rows, columns, samples), no atomics.
__global__ __launch_bounds__(256) void k_film_gather(SceneDev s, PathsDev P, int32_t n_pixels, int32_t spb,
                                                     float* film) {
    const int32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= n_pixels)
        return;
    const int x = pixel % s.width, y = pixel / s.width;
    const float rx = s.filter_radius[0], ry = s.filter_radius[1];
    const int Rx = (int)ceilf(rx + 0.5f), Ry = (int)ceilf(ry + 0.5f);
    const size_t own = (size_t)(s.height - 1 - y) * s.width + x;
    float acc[4] = {film[4 * own + 0], film[4 * own + 1], film[4 * own + 2], film[4 * own + 3]};
    const int qy0 = y - Ry > 0 ? y - Ry : 0, qy1 = y + Ry < s.height - 1 ? y + Ry : s.height - 1;
    const int qx0 = x - Rx > 0 ? x - Rx : 0, qx1 = x + Rx < s.width - 1 ? x + Rx : s.width - 1;
    for (int qy = qy0; qy <= qy1; qy++)
        for (int qx = qx0; qx <= qx1; qx++)
            for (int sl = 0; sl < spb; sl++) {
                const int32_t i = (qy * s.width + qx) * spb + sl;
                const float fx = P.pfilm[2 * i] * (float)s.width - 0.5f;
                const float fy = P.pfilm[2 * i + 1] * (float)s.height - 0.5f;
                const int x0 = (int)pstd::ceil(fx - rx), y0 = (int)pstd::ceil(fy - ry);
                const int x1 = (int)pstd::floor(fx + rx), y1 = (int)pstd::floor(fy + ry);
                if (x < x0 || x > x1 || y < y0 || y > y1)
                    continue;
                const float dx = fabsp((float)x - fx) / rx, dy = fabsp((float)y - fy) / ry;
                const int tx = (int)(16.0f * (dx < kOneMinusEps ? dx : kOneMinusEps));
                const int ty = (int)(16.0f * (dy < kOneMinusEps ? dy : kOneMinusEps));
                const float w = s.filter_table[ty * 16 + tx];
                acc[0] += P.L[3 * i] * w;
                acc[1] += P.L[3 * i + 1] * w;
                acc[2] += P.L[3 * i + 2] * w;
                acc[3] += w;
            }
    film[4 * own + 0] = acc[0];
    film[4 * own + 1] = acc[1];
    film[4 * own + 2] = acc[2];
    film[4 * own + 3] = acc[3];
}

}  // namespace

// ----------------------------------------------------------------------------
// Host: scene loading (fileio.cpp:251-281 and the Create* factories)

enum { kIntPath = 0, kIntAO = 1, kIntViz = 2, kIntRandomWalk = 3 };

// Per-path buffers of one batch.  Scenes do not own them: each device keeps a
// few sets (slot 0 for single-scene renders, one per stream of a batched
// render), reused by every scene rendered on that slot and grown on demand, so
// the memory held is bounded by the slots' largest batches, not by the number
// of live scenes.  Each kind of buffer has its own capacity, so one set serves
// fused, split-bounce, node-material, spectral and guided scenes alike.
struct PathState {
    PathsDev P{};
    int64_t core = 0, general = 0, nodes = 0, spectral = 0, guide = 0, records = 0;
    std::vector<void*> core_allocs, general_allocs, node_allocs, spectral_allocs, guide_allocs, record_allocs;
    int64_t gen = 0;  // changes with every reallocation, unique over all sets
    ~PathState() {
        for (const std::vector<void*>* l :
             {&core_allocs, &general_allocs, &node_allocs, &spectral_allocs, &guide_allocs, &record_allocs})
            for (void* p : *l)
                (void)hipFree(p);
    }
};

struct pg_scene {
    std::string dir;
    Params root;
    // host scene
    std::vector<ShapeDev> shapes;
    std::vector<MaterialDev> materials;
    std::vector<LightDev> lights;
    std::vector<float> tri_verts;
    std::vector<int32_t> prim_shape;
    std::vector<float> light_func, light_cdf;
    float light_func_int = 0.0f;
    bool light_uniform = false;
    SceneDev dev{};
    int spp = 1, max_depth = 4;
    int atmo_w = 0, atmo_h = 0;  // Atmosphere colour table size (0: no Atmosphere)
    bool atmo_accurate = false;  // table with correctly rounded exp/sqrt/pow (extension)
    std::vector<nodes::Inst> node_progs;  // material node-graph programs
    int sampler_kind = kSamplerHalton;
    bool halton_permute = true;            // HaltonSampler randomizeStrategy permutateDigits
    int sobol_randomize = kSobolFastOwen;  // SobolSampler randomizeStrategy
    std::string output = "result.png";
    bool tone_map = true;
    pg_accel* accel = nullptr;
    // device buffers owned by the scene
    std::vector<void*> allocations;
    float* d_film = nullptr;
    PathsDev paths{};   // the current render's view of its path-state slot
    int device = -1;    // device the scene was loaded on (-1 until registered)
    unsigned long long* d_lookups = nullptr;  // density lookups of the last render
    int32_t* d_viz_counts = nullptr;  // Viz Bvh: nodes and primitives per camera ray
    int64_t viz_counts_capacity = 0;
    double last_render_ms = 0.0;
    int64_t last_rays = 0;
    // media
    std::vector<MediumDev> media;
    std::vector<std::vector<float>> med_density, med_majorant;
    int lds_medium = -1;  // medium whose brick majorants the tracking kernels stage in LDS
    int64_t last_lookups = 0;
    bool general = false;
    int materialless_shapes = 0;  // shapes without a material (the split bounce's pass-throughs)
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};  // bounds of the finite shapes
    double last_guide_ms = 0.0;
    int64_t guide_steps = 0, guide_records = 0;
    // The fused pipeline's bounce loop recorded as a HIP graph (counts stay in
    // device memory), replayed for every batch while its key holds.
    struct GraphKey {
        int64_t capacity = -1, accel_gen = -1, paths_gen = -1;
        int guided = 0, oct = 0, max_depth = 0, g_depth = 0, rr_depth = 0;
        const void* g_params = nullptr;
        const void* g_wf = nullptr;
        bool operator==(const GraphKey& o) const {
            return capacity == o.capacity && accel_gen == o.accel_gen && paths_gen == o.paths_gen &&
                   guided == o.guided && oct == o.oct && max_depth == o.max_depth && g_depth == o.g_depth &&
                   rr_depth == o.rr_depth && g_params == o.g_params && g_wf == o.g_wf;
        }
    };
    bool use_graph = true;
    bool keep_film = false;  // the next render adds to the film instead of clearing it
    int integrator = 0;                  // kIntPath / kIntAO / kIntViz / kIntRandomWalk
    bool spectral = false;               // this render in spectral mode (PG_RENDER_SPECTRAL)
    GraphKey graph_key;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    int32_t* g_counts = nullptr;         // 4 per bounce after the batch's path count
    unsigned long long* g_rays = nullptr;  // rays traced: [0] extension + shadow, [1] extension
    int64_t last_ext_rays = 0;           // extension rays (path vertices) of the last render
    int64_t paths_gen = 0;               // generation of the path-state slot of the current render
    int64_t graph_captures = 0;
    // Paths per batch when not given (tunable), bounded by free device memory.
    // Large batches amortise each bounce's fixed cost (launches, and the split
    // bounce's count read-back) over the short tail bounces, and put a pixel's
    // samples in adjacent lanes, so a wave's rays start coherent.
    int64_t batch_paths = 256 << 20;
    pg_build_options build_opts{};       // the accel's build options (rebuilds keep them)
    // Guide inference kernel time and queries of the last launch-by-launch
    // render (events around each k_guide_infer launch; graph replays skip it).
    // Tracking-kernel variants of the current render: sampler (kSampler*) and
    // media (kMedia*), chosen per render from the scene and its sample range.
    int track_sampler = 0, track_media = 0;
    bool lane_tracking = false;  // brick-media tracking as persistent lanes (PG_RENDER_LANE_TRACKING)
    hipEvent_t infer_ev[2] = {nullptr, nullptr};
    double infer_ms = 0.0;
    int64_t infer_queries = 0;
    double guide_step_ms = 0.0;  // last pg_render_guided: GPU time of its Adam steps

    ~pg_scene() {
        for (hipEvent_t e : infer_ev)
            if (e)
                (void)hipEventDestroy(e);
        if (graph_exec)
            (void)hipGraphExecDestroy(graph_exec);
        if (capture_stream)
            (void)hipStreamDestroy(capture_stream);
        for (void* p : allocations)
            (void)hipFree(p);
        if (accel)
            pg_accel_release(accel);
        unregister_paths();
    }
    void unregister_paths();
};

namespace {

// The device's path-state slots and the lock that runs its renders one at a
// time (a render holds its slot from the first launch to the wait at its end).
struct DevicePaths {
    std::mutex render;
    std::vector<std::unique_ptr<PathState>> slots;
    int scenes = 0;
};
std::mutex g_paths_mu;
std::atomic<int64_t> g_paths_gen{0};

// Never destroyed: slots are freed when a device's last scene goes, not by
// static destructors that may run after the HIP runtime has shut down.
std::map<int, DevicePaths>& device_paths_map() {
    static auto* m = new std::map<int, DevicePaths>();
    return *m;
}

DevicePaths& device_paths(int dev) {
    std::lock_guard<std::mutex> lk(g_paths_mu);
    return device_paths_map()[dev];
}

PathState& path_slot(DevicePaths& D, int k) {
    std::lock_guard<std::mutex> lk(g_paths_mu);
    while ((int)D.slots.size() <= k)
        D.slots.push_back(std::make_unique<PathState>());
    return *D.slots[k];
}

int register_paths(pg_scene* S) {
    int dev = 0;
    PG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_paths_mu);
    device_paths_map()[dev].scenes++;
    S->device = dev;
    return PG_OK;
}

}  // namespace

void pg_scene::unregister_paths() {
    if (device < 0)
        return;
    std::lock_guard<std::mutex> lk(g_paths_mu);
    DevicePaths& D = device_paths_map()[device];
    if (--D.scenes == 0)
        D.slots.clear();
    device = -1;
}

namespace {

// Brick majorants staged in LDS by the tracking kernels: at most 32 KB.
constexpr size_t kMaxLdsMajorants = 8192;

std::runtime_error unsupported(const std::string& what) {
    return std::runtime_error("unsupported by the GPU path: " + what);
}

// A node input evaluated once: pine's CreateNode (node.cpp:33-98) for the
// constant forms a Path scene can use on the GPU.
struct ConstNode {
    float f;
    f3 v;
};
ConstNode const_node(const Params& p, const std::string& what) {
    std::string type = p.get_string("type");
    if (type == "Constant") {
        float vec[3];
        if (!p.has_value("float") || !p.has_value("vec3"))
            throw std::runtime_error(what + ": Constant needs float and vec3");
        p.get_vec3("vec3", vec);
        return {p.get_float("float"), mk(vec[0], vec[1], vec[2])};
    }
    std::string at = p.get_string("@");
    if (!at.empty() && type == at) {
        // A plain value: Constant(stof(v), stofs(v)).
        float vec[3];
        pine_stofs(at, vec, 3);
        float fv[3] = {vec[0], vec[1], vec[2]};
        if (p.has_value("vec3"))
            p.get_vec3("vec3", fv);
        return {p.has_value("float") ? p.get_float("float") : pine_stof(at), mk(fv[0], fv[1], fv[2])};
    }
    if (type.empty())
        throw std::runtime_error(what + ": missing value");
    throw unsupported("node type \"" + type + "\" for " + what);
}

// CreateNode (node.cpp:32-97) compiled for one evaluation mode (EvalFloat or
// EvalVec3, which node.h defines separately per node) into a postfix program
// of nodes::Inst.  Constant trees fold to one kConst.
struct NodeCompiler {
    std::vector<nodes::Inst>* out;
    int depth = 0, max_depth = 0;
    void emit(nodes::Op op, int arg = 0, float x = 0, float y = 0, float z = 0, int push = 0) {
        out->push_back({op, arg, x, y, z});
        depth += push;
        max_depth = std::max(max_depth, depth);
    }
    void constant(float f, f3 v, bool vec) {
        if (vec)
            emit(nodes::kConst, 0, v.x, v.y, v.z, 1);
        else
            emit(nodes::kConst, 0, f, f, f, 1);
    }
    void node(const Params& p, bool vec, const std::string& what) {
        using namespace nodes;
        const std::string type = p.get_string("type");
        auto in = [&](const char* k, bool v) { node(p.sub(k), v, what); };
        if (type == "Constant") {
            float v[3] = {0, 0, 0};
            p.get_vec3("vec3", v);
            constant(p.get_float("float"), mk(v[0], v[1], v[2]), vec);
        } else if (type == "Position" || type == "Normal" || type == "TexCoord") {
            if (!vec)  // Node::EvalFloat: 0
                constant(0.0f, mk(0, 0, 0), false);
            else
                emit(type == "Position" ? kPosition : type == "Normal" ? kNormal : kTexCoord, 0, 0, 0, 0, 1);
        } else if (type == "Decompose") {
            in("input", true);
            emit(kComponent, std::min(std::max(p.get_int("dimension"), 0), 2));
        } else if (type == "Composite") {
            if (!vec) {
                constant(0.0f, mk(0, 0, 0), false);
            } else {
                in("inputX", false), in("inputY", false), in("inputZ", false);
                emit(kCompose, 0, 0, 0, 0, -2);
            }
        } else if (type == "Add" || type == "Substract" || type == "Multiply" || type == "Divide") {
            in("input", vec), in("factor", vec);
            emit(type == "Add" ? kAdd : type == "Substract" ? kSub : type == "Multiply" ? kMul : kDiv, 0, 0, 0, 0, -1);
        } else if (type == "MultiplyAdd") {
            in("input", vec), in("mulFactor", vec), in("addFactor", vec);
            emit(kMulAdd, 0, 0, 0, 0, -2);
        } else if (type == "Length") {
            in("input", true);
            emit(kLength);
        } else if (type == "Sqr" || type == "Sqrt" || type == "Sin" || type == "Cos" || type == "Tan" ||
                   type == "Fract" || type == "Invert") {
            in("input", vec);
            emit(type == "Sqr" ? kSqr : type == "Sqrt" ? kSqrt : type == "Sin" ? kSin : type == "Cos" ? kCos
                 : type == "Tan" ? kTan : type == "Fract" ? kFract : kInvert);
        } else if (type == "Pow") {
            in("input", vec), in("exp", false);
            emit(kPow, 0, 0, 0, 0, -1);
        } else if (type == "Checkerboard") {
            in("input", true), in("frequency", false);
            emit(kChecker, 0, 0, 0, 0, -1);
        } else if (type == "Noise" || type == "Noise3D") {
            in("input", true), in("frequency", false), in("octaves", false);
            emit(type == "Noise" ? kNoise : kNoise3D, vec ? 1 : 0, 0, 0, 0, -2);
        } else if (type == "Texture") {
            throw unsupported("Texture node (image files) for " + what);
        } else if (type.empty() || !p.get_string("@").empty()) {
            // node.cpp:89-91: a plain value (or an unknown typed block) is a
            // Constant of its "float" / "vec3" keys, defaulting to the value.
            const std::string at = p.get_string("@");
            float v[3] = {0, 0, 0};
            pine_stofs(at, v, 3);
            if (p.has_value("vec3"))
                p.get_vec3("vec3", v);
            if (type.empty() && at.empty() && !p.has_value("float") && !p.has_value("vec3"))
                throw std::runtime_error(what + ": missing value");
            constant(p.has_value("float") ? p.get_float("float") : pine_stof(at), mk(v[0], v[1], v[2]), vec);
        } else {
            fprintf(stderr, "[NodeInput][Create]Unknown type \"%s\"\n", type.c_str());
            constant(p.get_float("float", 0.5f), mk(0.5f, 0.5f, 0.5f), vec);
        }
    }
};

// Compiles a material input; returns its program offset in `progs`, or -1 with
// the value in f / v when the input is a constant.
int compile_input(const Params& p, bool vec, const std::string& what, std::vector<nodes::Inst>& progs, float& f,
                  f3& v, bool always_program = false) {
    std::vector<nodes::Inst> body;
    NodeCompiler c{&body};
    c.node(p, vec, what);
    if (!always_program && body.size() == 1 && body[0].op == nodes::kConst) {
        f = body[0].x;
        v = mk(body[0].x, body[0].y, body[0].z);
        return -1;
    }
    if (c.max_depth > nodes::kStack)
        throw unsupported("node graph deeper than 16 values for " + what);
    const int at = (int)progs.size();
    progs.push_back({nodes::kProgram, (int)body.size(), 0, 0, 0});
    progs.insert(progs.end(), body.begin(), body.end());
    return at;
}

Layer make_layer(const Params& p, std::vector<nodes::Inst>& progs) {
    std::string type = p.get_string("type");
    Layer l{};
    l.prog_albedo = l.prog_roughness = l.prog_eta = -1;
    float f;
    f3 v;
    auto vec3_in = [&](const char* k, f3& dst, int& prog) {
        prog = compile_input(p.sub(k), true, k, progs, f, v);
        dst = v;
    };
    auto float_in = [&](const char* k, float& dst, int& prog) {
        prog = compile_input(p.sub(k), false, k, progs, f, v);
        dst = f;
    };
    if (type == "Dielectric") {
        l.type = kDielectric;
        if (p.has_subset("albedo"))
            vec3_in("albedo", l.albedo, l.prog_albedo);
        else
            l.albedo = mk(1, 1, 1);  // bxdf.cpp:176-180
        float_in("roughness", l.roughness, l.prog_roughness);
        float_in("eta", l.eta, l.prog_eta);
    } else if (type == "Conductor") {
        l.type = kConductor;
        vec3_in("albedo", l.albedo, l.prog_albedo);
        float_in("roughness", l.roughness, l.prog_roughness);
    } else {
        if (type != "Diffuse")
            fprintf(stderr, "[BSDF][Create]Unknown type \"%s\"\n", type.c_str());
        l.type = kDiffuse;
        vec3_in("albedo", l.albedo, l.prog_albedo);
    }
    return l;
}

MaterialDev make_material(const Params& p, std::vector<nodes::Inst>& progs) {
    MaterialDev m{};
    m.prog_bump = -1;
    std::string type = p.get_string("type");
    if (type == "Emissive") {
        m.kind = kMatEmissive;
        m.emission = const_node(p.sub("color"), "color").v;
    } else {
        if (type != "Layered")
            fprintf(stderr, "[Material][Create]Unknown type \"%s\"\n", type.c_str());
        m.kind = kMatLayered;
        // LayeredMaterial ctor (material.cpp:37-47): layers by descending index.
        std::vector<std::pair<int, const Params*>> layers;
        for (auto& kv : p.subsets)
            if (kv.first.compare(0, 5, "layer") == 0 && !kv.second.empty())
                layers.push_back({pine_stoi(kv.first.substr(5)), &kv.second.back()});
        std::stable_sort(layers.begin(), layers.end(),
                         [](const auto& a, const auto& b) { return a.first > b.first; });
        if ((int)layers.size() > kMaxLayers)
            throw unsupported("more than 4 material layers");
        m.nlayers = (int)layers.size();
        for (int i = 0; i < m.nlayers; i++)
            m.layers[i] = make_layer(*layers[i].second, progs);
    }
    // CreateMaterial (material.cpp:78-79): any material may carry a bump map.
    if (p.has_subset("bumpMap")) {
        float f;
        f3 v;
        // Always a program: even a constant map renormalises Cross(dpdu, dpdv).
        m.prog_bump = compile_input(p.sub("bumpMap"), false, "bumpMap", progs, f, v, true);
    }
    m.has_nodes = m.prog_bump >= 0;
    for (int i = 0; i < m.nlayers; i++)
        if (m.layers[i].prog_albedo >= 0 || m.layers[i].prog_roughness >= 0 || m.layers[i].prog_eta >= 0)
            m.has_nodes = 1;
    return m;
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot open \"" + path + "\"");
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// LoadObj (objloader.cpp:12-101) for the v / f forms it accepts.
void load_obj(const std::string& path, std::vector<float>& v, std::vector<uint32_t>& idx) {
    std::string text = read_file(path);
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        if (line.size() < 2)
            continue;
        if (line[0] == 'v' && line[1] != 't' && line[1] != 'n') {
            float p[3];
            pine_stofs(line.substr(2), p, 3);
            v.insert(v.end(), p, p + 3);
        } else if (line[0] == 'f') {
            std::istringstream fs(line.substr(2));
            std::vector<long> f;
            std::string tok;
            while (fs >> tok) {
                long k = pine_stoi(tok.substr(0, tok.find('/')));
                long nv = (long)(v.size() / 3);
                f.push_back(k < 0 ? nv + k : k - 1);
            }
            if (f.size() == 3) {
                idx.insert(idx.end(), {(uint32_t)f[0], (uint32_t)f[1], (uint32_t)f[2]});
            } else if (f.size() == 4) {
                idx.insert(idx.end(), {(uint32_t)f[0], (uint32_t)f[1], (uint32_t)f[2],
                                       (uint32_t)f[0], (uint32_t)f[2], (uint32_t)f[3]});
            }
        }
    }
    if (!idx.empty()) {
        uint32_t mn = *std::min_element(idx.begin(), idx.end());
        for (auto& k : idx)
            k -= mn;
    }
}

// LoadVolume (fileio.cpp:194-203): three int32 sizes, then the floats, x fastest.
void load_volume(const std::string& path, std::vector<float>& density, int32_t size[3]) {
    if (path.size() >= 11 && path.compare(path.size() - 11, 11, ".compressed") == 0)
        throw unsupported("compressed volume files (\"" + path + "\"); use a raw volume");
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot open volume \"" + path + "\"");
    f.read(reinterpret_cast<char*>(size), 3 * sizeof(int32_t));
    if (!f || size[0] <= 0 || size[1] <= 0 || size[2] <= 0 ||
        (int64_t)size[0] * size[1] * size[2] > ((int64_t)1 << 31))
        throw std::runtime_error("bad volume header in \"" + path + "\"");
    density.resize((size_t)size[0] * size[1] * size[2]);
    f.read(reinterpret_cast<char*>(density.data()), (std::streamsize)(density.size() * sizeof(float)));
    if (!f)
        throw std::runtime_error("truncated volume \"" + path + "\"");
}

// Restates Inverse(mat4) (vecmath.h:1147-1172) for the affine w2m of a grid;
// m[c][r] column-major as pine's Matrix4.
void inverse4(const float m[4][4], float r[4][4]) {
    float det = 0.0f;
    for (int i = 0; i < 4; i++)
        det += (m[(1 + i) % 4][0] * (m[(2 + i) % 4][1] * m[(3 + i) % 4][2] -
                                     m[(3 + i) % 4][1] * m[(2 + i) % 4][2]) +
                m[(2 + i) % 4][0] * (m[(3 + i) % 4][1] * m[(1 + i) % 4][2] -
                                     m[(1 + i) % 4][1] * m[(3 + i) % 4][2]) +
                m[(3 + i) % 4][0] * (m[(1 + i) % 4][1] * m[(2 + i) % 4][2] -
                                     m[(2 + i) % 4][1] * m[(1 + i) % 4][2])) *
               m[i % 4][3] * (i % 2 ? -1.0f : 1.0f);
    if (det == 0.0f)
        throw std::runtime_error("GridMedium: singular world-to-grid transform");
    const float inv = 1.0f / det;
    for (int v = 0; v < 4; v++)
        for (int i = 0; i < 4; i++)
            r[v][i] = (m[(1 + i) % 4][(1 + v) % 4] * (m[(2 + i) % 4][(2 + v) % 4] * m[(3 + i) % 4][(3 + v) % 4] -
                                                      m[(3 + i) % 4][(2 + v) % 4] * m[(2 + i) % 4][(3 + v) % 4]) +
                       m[(2 + i) % 4][(1 + v) % 4] * (m[(3 + i) % 4][(2 + v) % 4] * m[(1 + i) % 4][(3 + v) % 4] -
                                                      m[(1 + i) % 4][(2 + v) % 4] * m[(3 + i) % 4][(3 + v) % 4]) +
                       m[(3 + i) % 4][(1 + v) % 4] * (m[(1 + i) % 4][(2 + v) % 4] * m[(2 + i) % 4][(3 + v) % 4] -
                                                      m[(2 + i) % 4][(2 + v) % 4] * m[(1 + i) % 4][(3 + v) % 4])) *
                      ((v + i) % 2 ? 1.0f : -1.0f) * inv;
}

// CreateMedium (medium.cpp:185-196) with HomogeneousMedium / GridMedium::Create
// (:162-183).  Extension: `majorantBricks: B` on a Grid tracks against one
// majorant per B^3-voxel brick instead of the global one.
void make_medium(pg_scene& S, const Params& p) {
    std::string type = p.get_string("type");
    MediumDev m{};
    float a[3], sc[3];
    p.get_vec3("sigma_a", a);
    p.get_vec3("sigma_s", sc);
    m.sigma_a = mk(a[0], a[1], a[2]);
    m.sigma_s = mk(sc[0], sc[1], sc[2]);
    m.sigma_t3 = m.sigma_a + m.sigma_s;
    m.g = p.get_float("g", 0.0f);
    std::vector<float> density, majorant;
    if (type == "Grid") {
        m.type = kMediumGrid;
        std::string file = p.get_string("file");
        if (!file.empty() && file[0] != '/')
            file = S.dir + "/" + file;
        load_volume(file, density, m.size);
        std::string method = p.get_string("samplingMethod", "deltaTracking");
        m.method = kDeltaTracking;
        if (method == "rayMarching")
            m.method = kRayMarching;
        else if (method != "deltaTracking")
            fprintf(stderr, "[GridMedium]Unknown sampling method \"%s\"\n", method.c_str());
        float step = p.get_float("rayMarchingStepSize", 2.0f);
        if (step < 1.0f)
            step = 1.0f;
        const int32_t mx = std::max(m.size[0], std::max(m.size[1], m.size[2]));
        m.march_step = step / (float)mx;
        m.interpolate = p.get_bool("interpolate", true) ? 1 : 0;
        m.sigma_t = (m.sigma_a + m.sigma_s).x;
        float maxd = 0.0f;
        for (float d : density)
            maxd = std::max(d, maxd);
        m.inv_max_density = 1.0f / maxd;
        m.ln2_inv_max = pstd::Ln2 * m.inv_max_density;
        // w2m = Translate(0.5 - position) * Scale(1 / normalizedSize / scale)
        float pos[3];
        const float zero3[3] = {0.0f, 0.0f, 0.0f};
        p.get_vec3("position", pos, zero3);
        const float scale = p.get_float("scale", 1.0f);
        const float rm = 1.0f / (float)mx, rs = 1.0f / scale;
        float s3[3], t3[3];
        for (int k = 0; k < 3; k++) {
            s3[k] = (1.0f / ((float)m.size[k] * rm)) * rs;
            t3[k] = 0.5f - pos[k];
        }
        m.w2m_s = mk(s3[0], s3[1], s3[2]);
        m.w2m_t = mk(t3[0], t3[1], t3[2]);
        float w2m[4][4] = {{s3[0], 0, 0, 0}, {0, s3[1], 0, 0}, {0, 0, s3[2], 0}, {t3[0], t3[1], t3[2], 1}};
        float m2w[4][4];
        inverse4(w2m, m2w);
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 3; r++)
                m.m2w[3 * c + r] = m2w[c][r];
        m.brick = p.get_int("majorantBricks", 0);
        if (m.brick < 0)
            throw std::runtime_error("GridMedium: majorantBricks must be >= 0");
        if (m.brick > 0 && m.method == kDeltaTracking) {
            // Brick b bounds the trilinear density of every point inside it:
            // voxels b*B - 1 .. (b+1)*B (one voxel of halo on each side).
            const int B = m.brick;
            for (int k = 0; k < 3; k++)
                m.nbricks[k] = (m.size[k] + B - 1) / B;
            majorant.assign((size_t)m.nbricks[0] * m.nbricks[1] * m.nbricks[2], 0.0f);
            for (int z = 0; z < m.size[2]; z++)
                for (int y = 0; y < m.size[1]; y++)
                    for (int x = 0; x < m.size[0]; x++) {
                        const float d = density[(size_t)x + (size_t)m.size[0] * (y + (size_t)m.size[1] * z)];
                        if (!(d > 0.0f))
                            continue;
                        int lo[3], hi[3];
                        const int c[3] = {x, y, z};
                        for (int k = 0; k < 3; k++) {
                            lo[k] = std::max(0, (c[k] - 1) / B);
                            hi[k] = std::min(m.nbricks[k] - 1, (c[k] + 1) / B);
                        }
                        for (int bz = lo[2]; bz <= hi[2]; bz++)
                            for (int by = lo[1]; by <= hi[1]; by++)
                                for (int bx = lo[0]; bx <= hi[0]; bx++) {
                                    float& mj = majorant[(size_t)bx + (size_t)m.nbricks[0] * (by + (size_t)m.nbricks[1] * bz)];
                                    mj = std::max(mj, d);
                                }
                    }
        } else {
            m.brick = 0;
        }
    } else {
        if (type != "Homogeneous")
            fprintf(stderr, "[Medium][Create]Unknown type \"%s\"\n", type.c_str());
        m.type = kMediumHomogeneous;
    }
    S.media.push_back(m);
    S.med_density.push_back(std::move(density));
    S.med_majorant.push_back(std::move(majorant));
}

void build_scene(pg_scene& S) {
    const Params& root = S.root;
    // Integrator block (integrator.cpp:20-56): Path, and AO, Viz (Position,
    // Normal) and RandomWalk on the same kernels.
    const Params& ip = root.sub("Integrator");
    std::string itype = ip.get_string("type");
    if (itype == "Path") {
        S.integrator = kIntPath;
    } else if (itype == "AO") {
        S.integrator = kIntAO;
    } else if (itype == "RandomWalk") {
        S.integrator = kIntRandomWalk;
    } else if (itype == "Viz") {
        S.integrator = kIntViz;
        // VizIntegrator's constructor (viz.cpp:7-22): an unknown type is Normal.
        const std::string vt = ip.get_string("viztype", "");
        S.dev.viz_type = vt == "Position"   ? kVizPosition
                         : vt == "Texcoord" ? kVizTexcoord
                         : vt == "DpDu"     ? kVizDpDu
                         : vt == "DpDv"     ? kVizDpDv
                         : vt == "Bvh"      ? kVizBvh
                                            : kVizNormal;
        if (vt != "Normal" && S.dev.viz_type == kVizNormal)
            fprintf(stderr, "Unknown viztype %s\n", vt.c_str());
    } else {
        throw unsupported("integrator \"" + itype + "\" (Path, AO, Viz and RandomWalk are supported)");
    }
    S.max_depth = ip.get_int("maxDepth", 4);
    const Params& sp = ip.sub("sampler");
    std::string stype = sp.get_string("type");
    if (stype == "Halton") {
        // HaltonSampler::Create (sampler.cpp:253-260): an unknown strategy keeps the default.
        S.sampler_kind = kSamplerHalton;
        S.halton_permute = sp.get_string("randomizeStrategy", "permutateDigits") != "none";
    } else if (stype == "Sobol") {
        // SobolSampler::Create (sampler.cpp:268-278): an unknown strategy leaves
        // the initial BinaryPermutate in place.
        S.sampler_kind = kSamplerSobol;
        const std::string rs = sp.get_string("randomizeStrategy", "fastOwen");
        S.sobol_randomize = rs == "none"              ? kSobolNone
                            : rs == "binaryPermutate" ? kSobolBinaryPermutate
                            : rs == "Owen"            ? kSobolOwen
                            : rs == "fastOwen"        ? kSobolFastOwen
                                                      : kSobolBinaryPermutate;
    } else {
        throw unsupported("sampler \"" + stype + "\" (Halton and Sobol are supported)");
    }
    S.spp = sp.get_int("samplesPerPixel");
    std::string lstype = ip.sub("lightSampler").get_string("type", "Power");
    S.light_uniform = lstype == "Uniform";
    if (lstype != "Power" && lstype != "Uniform") {
        fprintf(stderr, "[LightSampler][Create]Unknown type \"%s\"\n", lstype.c_str());
        S.light_uniform = true;
    }

    // Materials by name.
    std::map<std::string, int> mat_index;
    for (const Params& p : root.all("Material")) {
        mat_index[p.get_string("name")] = (int)S.materials.size();
        S.materials.push_back(make_material(p, S.node_progs));
    }
    // Media by name (fileio.cpp:262-263).
    std::map<std::string, int> med_index;
    for (const Params& p : root.all("Medium")) {
        med_index[p.get_string("name")] = (int)S.media.size();
        make_medium(S, p);
    }
    auto find_medium = [&](const std::string& name) {
        auto it = med_index.find(name);
        if (it == med_index.end())
            throw std::runtime_error("medium \"" + name + "\" is not found");
        return it->second;
    };
    bool uses_media = false;
    S.dev.has_env = 0;
    // Lights from Light blocks (fileio.cpp:264-265).
    for (const Params& p : root.all("Light")) {
        std::string t = p.get_string("type");
        LightDev l{};
        float a[3], c[3];
        if (t == "Directional") {
            l.type = kLightDirectional;
            p.get_vec3("direction", a);
            l.v = normalize(mk(a[0], a[1], a[2]));
        } else if (t == "Point") {
            l.type = kLightPoint;
            p.get_vec3("position", a);
            l.v = mk(a[0], a[1], a[2]);
        } else if (t == "Environment") {
            // EnvironmentLight::Create (light.cpp:157-168): Sky, or Atmosphere,
            // which is also what an unknown type falls back to.
            const Params& ep = p.sub("environment");
            std::string et = ep.get_string("type");
            if (et == "Sky") {
                ep.get_vec3("sunDirection", a);
                ep.get_vec3("sunColor", c);
                l.type = kLightSky;
                S.dev.env_kind = kEnvSky;
                S.atmo_w = S.atmo_h = 0;
            } else if (et == "Constant") {
                // This library's extension (pine would fall back to its
                // Atmosphere): the same radiance from every direction.
                const float up[3] = {0.0f, 1.0f, 0.0f};
                ep.get_vec3("color", c);
                a[0] = up[0], a[1] = up[1], a[2] = up[2];
                l.type = kLightConstantEnv;
                S.dev.env_kind = kEnvConstant;
                S.atmo_w = S.atmo_h = 0;
            } else {
                if (et != "Atmosphere")
                    fprintf(stderr, "[EnvironmentLight][Create]Unknown type \"%s\"\n", et.c_str());
                // Atmosphere::Create (light.cpp:152-156) defaults.  `accurate`
                // is this library's extension (pine ignores unknown keys): the
                // colour table with correctly rounded exp/sqrt/pow instead of
                // pine's approximations, whose exp diverges here and leaves no
                // finite entry in pine's own table (NaN or +inf).
                const float dsun[3] = {2.0f, 6.0f, 3.0f}, dcol[3] = {1.0f, 1.0f, 1.0f};
                const int dsize[2] = {512, 512};
                int sz[2];
                ep.get_vec3("sunDirection", a, dsun);
                ep.get_vec3("sunColor", c, dcol);
                ep.get_vec2i("size", sz, dsize);
                if (sz[0] < 1 || sz[1] < 1 || (int64_t)sz[0] * sz[1] > (1 << 26))
                    throw std::runtime_error("Atmosphere: size out of range");
                l.type = kLightAtmosphere;
                S.dev.env_kind = kEnvAtmosphere;
                S.dev.atmo_interp = ep.get_bool("interpolate", true) ? 1 : 0;
                S.atmo_w = sz[0];
                S.atmo_h = sz[1];
                S.atmo_accurate = ep.get_bool("accurate", false);
            }
            l.v = normalize(mk(a[0], a[1], a[2]));
            l.color = mk(c[0], c[1], c[2]);
            // scene->envLight is the last environment light (fileio.cpp:272-274).
            S.dev.has_env = 1;
            S.dev.sky_dir = l.v;
            S.dev.sky_color = l.color;
            S.lights.push_back(l);
            continue;
        } else {
            throw unsupported("light type \"" + t + "\"");
        }
        p.get_vec3("color", c);
        l.color = mk(c[0], c[1], c[2]);
        S.lights.push_back(l);
    }

    // Shapes (geometry.cpp:366-411).
    pg_accel* a = nullptr;
    if (int rc = pg_accel_create(&a))
        throw std::runtime_error(pg_last_error());
    S.accel = a;
    std::vector<int> area_light_shapes;
    for (const Params& p : root.all("Shape")) {
        std::string t = p.get_string("type");
        ShapeDev sh{};
        sh.material = -1;
        std::string mname;
        if (p.try_string("material", mname)) {
            auto it = mat_index.find(mname);
            if (it == mat_index.end())
                throw std::runtime_error("material \"" + mname + "\" is not found");
            sh.material = it->second;
        }
        // geometry.cpp:392-409: "medium" and "mediumInside" set the inside,
        // "mediumOutside" the outside.
        sh.mi_inside = sh.mi_outside = -1;
        std::string mn;
        if (p.try_string("medium", mn))
            sh.mi_inside = find_medium(mn);
        if (p.try_string("mediumInside", mn))
            sh.mi_inside = find_medium(mn);
        if (p.try_string("mediumOutside", mn))
            sh.mi_outside = find_medium(mn);
        if (sh.mi_inside >= 0 || sh.mi_outside >= 0)
            uses_media = true;
        if (sh.material < 0) {
            S.general = true;
            S.materialless_shapes++;
        }
        float pos[3], e0[3], e1[3];
        int32_t first = 0;
        const int shape_index = (int)S.shapes.size();
        if (t == "Sphere") {
            p.get_vec3("position", pos);
            float r = p.get_float("radius");
            sh.type = kShapeSphere;
            sh.p0 = mk(pos[0], pos[1], pos[2]);
            sh.r = r;
            sh.area = 4.0f * kPi * r * r;
            if (pg_accel_add_sphere(a, pos, r, &first))
                throw std::runtime_error(pg_last_error());
        } else if (t == "Rect") {
            p.get_vec3("position", pos);
            p.get_vec3("ex", e0);
            p.get_vec3("ey", e1);
            f3 X = mk(e0[0], e0[1], e0[2]), Y = mk(e1[0], e1[1], e1[2]);
            sh.type = kShapeRect;
            sh.p0 = mk(pos[0], pos[1], pos[2]);
            sh.e0 = normalize(X);
            sh.e1 = normalize(Y);
            sh.n = normalize(cross(X, Y));
            sh.lx = length(X);
            sh.ly = length(Y);
            sh.area = sh.lx * sh.ly;
            if (pg_accel_add_rect(a, pos, e0, e1, &first))
                throw std::runtime_error(pg_last_error());
        } else if (t == "Plane") {
            p.get_vec3("position", pos);
            p.get_vec3("normal", e0);
            sh.type = kShapePlane;
            sh.p0 = mk(pos[0], pos[1], pos[2]);
            sh.n = normalize(mk(e0[0], e0[1], e0[2]));
            coordinate_system(sh.n, sh.e0, sh.e1);
            sh.area = kFloatMax;
            if (pg_accel_add_plane(a, pos, e0, &first))
                throw std::runtime_error(pg_last_error());
        } else if (t == "Triangle" || t == "TriangleMesh") {
            std::vector<float> v;
            std::vector<uint32_t> idx;
            if (t == "Triangle") {
                float q[3];
                for (const char* k : {"v0", "v1", "v2"}) {
                    p.get_vec3(k, q);
                    v.insert(v.end(), q, q + 3);
                }
                idx = {0, 1, 2};
            } else {
                std::string file = p.get_string("file");
                if (!file.empty() && file[0] != '/')
                    file = S.dir + "/" + file;
                load_obj(file, v, idx);
                float sc[3], of[3];
                if (p.try_vec3("scale", sc))
                    for (size_t k = 0; k < v.size(); k++)
                        v[k] *= sc[k % 3];
                if (p.try_vec3("position", of))
                    for (size_t k = 0; k < v.size(); k++)
                        v[k] += of[k % 3];
            }
            sh.type = kShapeMesh;
            sh.tri_offset = (int32_t)(S.tri_verts.size() / 9);
            const int64_t nt = (int64_t)idx.size() / 3;
            for (int64_t k = 0; k < 3 * nt; k++)
                for (int c = 0; c < 3; c++)
                    S.tri_verts.push_back(v[3 * (size_t)idx[k] + c]);
            if (pg_accel_add_mesh(a, v.data(), (int64_t)v.size() / 3, idx.data(), nt, &first))
                throw std::runtime_error(pg_last_error());
            if (t == "Triangle")
                sh.area = length(cross(mk(v[3] - v[0], v[4] - v[1], v[5] - v[2]),
                                       mk(v[6] - v[0], v[7] - v[1], v[8] - v[2]))) * 0.5f;
        } else {
            throw unsupported("shape type \"" + t + "\"");
        }
        sh.first_prim = first;
        int64_t nprims = (t == "Triangle" || t == "TriangleMesh") ? (int64_t)S.tri_verts.size() / 9 - sh.tri_offset : 1;
        for (int64_t k = 0; k < nprims; k++)
            S.prim_shape.push_back(shape_index);
        if (sh.material >= 0 && S.materials[sh.material].kind == kMatEmissive) {
            if (t == "TriangleMesh" || t == "Plane")
                throw unsupported("emissive " + t + " (pine cannot sample it)");
            area_light_shapes.push_back(shape_index);
        }
        S.shapes.push_back(sh);
    }
    // Scene bounds over the finite shapes (the guide's position input).
    {
        float lo[3] = {kFloatMax, kFloatMax, kFloatMax}, hi[3] = {-kFloatMax, -kFloatMax, -kFloatMax};
        auto grow = [&](f3 q) {
            const float c[3] = {q.x, q.y, q.z};
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], c[k]);
                hi[k] = std::max(hi[k], c[k]);
            }
        };
        for (size_t k = 0; k + 2 < S.tri_verts.size(); k += 3)
            grow(mk(S.tri_verts[k], S.tri_verts[k + 1], S.tri_verts[k + 2]));
        for (const ShapeDev& sh : S.shapes) {
            if (sh.type == kShapeSphere) {
                grow(sh.p0 - mk(sh.r, sh.r, sh.r));
                grow(sh.p0 + mk(sh.r, sh.r, sh.r));
            } else if (sh.type == kShapeRect) {
                for (int a = -1; a <= 1; a += 2)
                    for (int b = -1; b <= 1; b += 2)
                        grow(sh.p0 + sh.e0 * (0.5f * a * sh.lx) + sh.e1 * (0.5f * b * sh.ly));
            }
        }
        for (int k = 0; k < 3; k++) {
            if (!(lo[k] <= hi[k]))
                lo[k] = -1.0f, hi[k] = 1.0f;
            S.bmin[k] = lo[k];
            S.bmax[k] = hi[k];
        }
    }
    // Area lights follow the Light blocks (fileio.cpp:269-271).
    for (int si : area_light_shapes) {
        LightDev l{};
        l.type = kLightArea;
        l.shape = si;
        S.lights.push_back(l);
    }

    // PowerLightSampler (lightsampler.cpp:20-25): Power().y() per light.  An
    // area light's Power evaluates its emission with a zero normal, which
    // yields no emission in pine's build (observed: selection pdf 1 for one
    // area light), so area lights count as zero power.
    const int nl = (int)S.lights.size();
    S.light_func.resize(nl);
    for (int k = 0; k < nl; k++) {
        const LightDev& l = S.lights[k];
        S.light_func[k] = l.type == kLightPoint ? luminance(l.color * (kPi * 4.0f)) : 0.0f;
    }
    S.light_cdf.assign(nl + 1, 0.0f);
    if (nl > 0) {
        float dx = 1.0f / (float)nl;
        for (int k = 1; k <= nl; k++)
            S.light_cdf[k] = S.light_cdf[k - 1] + S.light_func[k - 1] * dx;
        S.light_func_int = S.light_cdf[nl];
        if (S.light_func_int != 0)
            for (int k = 0; k <= nl; k++)
                S.light_cdf[k] /= S.light_func_int;
    }

    // Camera (camera.cpp:8-12, camera.h:22-33) and film (film.cpp:8-28).
    const Params& cp = root.sub("Camera");
    std::string ctype = cp.get_string("type");
    if (ctype != "ThinLen")
        fprintf(stderr, "[Camera][Create]Unknown type \"%s\"\n", ctype.c_str());
    S.dev.camera_medium = -1;
    std::string cmed;
    if (cp.try_string("medium", cmed)) {
        S.dev.camera_medium = find_medium(cmed);
        uses_media = true;
    }
    if (uses_media)
        S.general = true;
    const Params& fp = cp.sub("film");
    int size[2];
    const int defsize[2] = {720, 480};
    fp.get_vec2i("size", size, defsize);
    S.output = fp.get_string("outputFileName", "result.png");
    S.tone_map = fp.get_bool("applyToneMapping", true);
    const Params& flp = fp.sub("filter");
    std::string ftype = flp.get_string("type", "LanczosSinc");
    float radius[2];
    const float defr[2] = {0.5f, 0.5f};
    flp.get_vec2("radius", radius, defr);
    // Film's filter table (film.cpp:22-28): the filter at the 16 x 16 cell
    // centres of [0, radius)^2; CreateFilter (filter.cpp:5-22) with filter.h's
    // Evaluate members.  An unknown type is a Box with a warning.
    if (ftype != "Box" && ftype != "Triangle" && ftype != "Gaussian" && ftype != "Mitchell" &&
        ftype != "LanczosSinc") {
        fprintf(stderr, "[Filter][Create]Unknown type \"%s\"\n", ftype.c_str());
        ftype = "Box";
    }
    const float tau = flp.get_float("tau", 3.0f);
    const float galpha = flp.get_float("alpha", 1.0f);
    const float mB = flp.get_float("B", 1.0f / 3.0f), mC = flp.get_float("C", 1.0f / 3.0f);
    // GaussianFilter's constructor; MitchellFilter's invRadius = 1.0 / radius (double).
    const float expX = pstd::exp(-galpha * (radius[0] * radius[0]));
    const float expY = pstd::exp(-galpha * (radius[1] * radius[1]));
    const float invRx = (float)(1.0 / (double)radius[0]), invRy = (float)(1.0 / (double)radius[1]);
    auto sinc = [](float x) {
        x = fabsp(x);
        if (x < 1e-5f)
            return 1.0f;
        return pstd::sin(kPi * x) / (kPi * x);
    };
    auto wsinc = [&](float x, float r) {  // LanczosSincFilter (filter.h:72-96)
        x = fabsp(x);
        if (x > r)
            return 0.0f;
        return sinc(x) * sinc(x / tau);
    };
    auto gauss = [&](float d, float expv) {  // GaussianFilter::Gaussian (filter.h:35-37)
        const float g = pstd::exp(-galpha * d * d) - expv;
        return g > 0.0f ? g : 0.0f;
    };
    auto mitchell = [&](float x) {  // MitchellFilter::Mitchell1D (filter.h:47-57)
        x = fabsp(2.0f * x);
        if (x > 1.0f)
            return ((-mB - 6.0f * mC) * x * x * x + (6.0f * mB + 30.0f * mC) * x * x + (-12.0f * mB - 48.0f * mC) * x +
                    (8.0f * mB + 24.0f * mC)) *
                   (1.0f / 6.0f);
        return ((12.0f - 9.0f * mB - 6.0f * mC) * x * x * x + (-18.0f + 12.0f * mB + 6.0f * mC) * x * x +
                (6.0f - 2.0f * mB)) *
               (1.0f / 6.0f);
    };
    float table[256];
    for (int y = 0, o = 0; y < 16; y++)
        for (int x = 0; x < 16; x++) {
            float px = ((float)x + 0.5f) / 16.0f * radius[0];
            float py = ((float)y + 0.5f) / 16.0f * radius[1];
            float w;
            if (ftype == "Box") {
                w = 1.0f;
            } else if (ftype == "Triangle") {
                const float tx = radius[0] - fabsp(px), ty = radius[1] - fabsp(py);
                w = (tx > 0.0f ? tx : 0.0f) * (ty > 0.0f ? ty : 0.0f);
            } else if (ftype == "Gaussian") {
                w = gauss(px, expX) * gauss(py, expY);
            } else if (ftype == "Mitchell") {
                // pine divides y by the inverse radius (filter.h:45)
                w = mitchell(px * invRx) * mitchell(py / invRy);
            } else {
                w = wsinc(px, radius[0]) * wsinc(py, radius[1]);
            }
            table[o++] = w;
        }
    float from[3], to[3];
    cp.get_vec3("from", from);
    cp.get_vec3("to", to);
    float fov = cp.get_float("fov");
    float lensRadius = cp.get_float("lensRadius", 0.0f);
    float focus = cp.get_float("focusDistance", 1.0f);
    // LookAt (vecmath.h:1212-1221)
    f3 F = mk(from[0], from[1], from[2]), T = mk(to[0], to[1], to[2]), up = mk(0, 1, 0);
    f3 z = normalize(T - F);
    if (fabsp(dot(z, up)) > 0.999f)
        z = normalize(z + mk(0.0f, 0.0f, 1e-5f));
    f3 x = normalize(cross(up, z));
    f3 y = cross(z, x);
    CameraDev cam;
    cam.cx = x;
    cam.cy = y;
    cam.cz = z;
    cam.cw = F;
    float tf = pstd::tan(fov / 2.0f);
    float aspect = (float)size[0] / (float)size[1];
    cam.fov2d[0] = tf * aspect;
    cam.fov2d[1] = tf;
    cam.lensRadius = lensRadius;
    cam.focalDistance = focus;

    SceneDev& d = S.dev;
    d.cam = cam;
    memcpy(d.filter_table, table, sizeof(table));
    d.filter_radius[0] = radius[0];
    d.filter_radius[1] = radius[1];
    d.width = size[0];
    d.height = size[1];
    d.max_depth = S.max_depth;
    d.rr_depth = 2;  // PathIntegrator::Li: Russian roulette after depth 2 (path.cpp:76)
    d.n_lights = nl;
    d.light_uniform = S.light_uniform ? 1 : 0;
    d.light_func_int = S.light_func_int;
}

// Halton tables: primes, prefix sums and RNG(0)-shuffled digit permutations
// (lowdiscrepancy.cpp:5-21, rng.h:39-105).
void halton_tables(std::vector<int>& primes, std::vector<int>& sums, std::vector<uint16_t>& perms) {
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
    perms.resize(acc);
    uint64_t seed = 0, st[2];
    auto splitmix = [&](uint64_t& s) {
        uint64_t r = s += 0x9E3779B97f4A7C15ULL;
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ULL;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBULL;
        return r ^ (r >> 31);
    };
    st[0] = splitmix(seed);
    st[1] = splitmix(seed);
    auto next32 = [&]() {
        const uint64_t s0 = st[0];
        uint64_t s1 = st[1];
        const uint64_t result = s0 + s1;
        s1 ^= s0;
        st[0] = ((s0 << 24) | (s0 >> 40)) ^ s1 ^ (s1 << 16);
        st[1] = (s1 << 37) | (s1 >> 27);
        return (uint32_t)(result ^ (result >> 32));
    };
    uint16_t* p = perms.data();
    for (int i = 0; i < 1000; i++) {
        for (int j = 0; j < primes[i]; j++)
            p[j] = (uint16_t)j;
        for (int k = 0; k < primes[i]; k++) {
            int other = k + (int)(next32() % (uint32_t)(primes[i] - k));
            std::swap(p[k], p[other]);
        }
        p += primes[i];
    }
}

void halton_setup(HaltonDev& h, int w, int hgt) {
    const int film[2] = {w, hgt};
    for (int i = 0; i < 2; i++) {
        int base = i == 0 ? 2 : 3, scale = 1, exp = 0;
        while (scale < std::min(film[i], 128)) {
            scale *= base;
            ++exp;
        }
        h.baseScales[i] = scale;
        h.baseExponents[i] = exp;
    }
    // multiplicativeInverse (sampler.cpp:9-30) by the extended Euclid algorithm.
    auto mulinv = [](int64_t a, int64_t n) {
        int64_t r0 = n, r1 = a % n, t0 = 0, t1 = 1;
        while (r1) {
            int64_t q = r0 / r1;
            int64_t r2 = r0 - q * r1, t2 = t0 - q * t1;
            r0 = r1, r1 = r2, t0 = t1, t1 = t2;
        }
        return (int)(((t0 % n) + n) % n);
    };
    h.multInverse[0] = mulinv(h.baseScales[0], h.baseScales[1]);
    h.multInverse[1] = mulinv(h.baseScales[1], h.baseScales[0]);
    h.stride = h.baseScales[0] * h.baseScales[1];
}

// Sobol generator matrices, dimension-major, 32 columns each: dimension 0 is
// van der Corput, dimensions 1..1023 expand Joe & Kuo's direction numbers by
// the Bratley-Fox recurrence in 32-bit fixed point (the numbers pine's table,
// src/util/sobolmetrices.cpp, holds expanded).
void sobol_matrices(std::vector<uint32_t>& C) {
    C.assign((size_t)kSobolDimensions * kSobolColumns, 0u);
    for (int k = 0; k < kSobolColumns; k++)
        C[k] = 1u << (31 - k);
    const uint16_t* q = kSobolDirnums;
    for (int d = 1; d < kSobolDimensions; d++) {
        const int deg = *q++, a = *q++;
        uint32_t* c = &C[(size_t)d * kSobolColumns];
        for (int k = 1; k <= kSobolColumns; k++) {
            if (k <= deg) {
                c[k - 1] = (uint32_t)q[k - 1] << (32 - k);
                continue;
            }
            uint32_t v = c[k - deg - 1] ^ (c[k - deg - 1] >> deg);
            for (int j = 1; j < deg; j++)
                if ((a >> (deg - 1 - j)) & 1)
                    v ^= c[k - j - 1];
            c[k - 1] = v;
        }
        q += deg;
    }
}

// The pixel -> Sobol index tables for a 2^m x 2^m pixel grid (the tables
// pine's SobolIntervalToIndex reads as VdCSobolMatrices[m-1] and
// VdCSobolMatricesInv[m-1]), derived here from the first two dimensions.
// With index = frame * 4^m + j, the pixel code b = (x << m) | y of an index
// (x, y: the top m bits of dimensions 0 and 1) is linear over GF(2) in the
// index bits.  Dimension 0 is the bit reversal (index bit k < m gives x bit
// m-1-k); dimension 1 is the Pascal matrix mod 2, whose column k has bit r
// (from the top) set iff (k & r) == r (Lucas).  vdc[c] is the code of index
// bit 2m + c (kept only where pine's 52-entry rows hold it, c < 52 - 2m), and
// vdc_inv[c] the index bits j whose code is bit c, from the inverse of the
// 2m x 2m map of j.
void sobol_pixel_tables(int m, std::vector<uint64_t>& vdc, std::vector<uint64_t>& vdc_inv) {
    vdc.assign(52, 0);
    vdc_inv.assign(52, 0);
    if (m == 0)
        return;
    auto code = [m](int k) -> uint64_t {
        uint64_t x = k < m ? (1ull << (m - 1 - k)) : 0, y = 0;
        for (int r = 0; r < m; r++)
            if ((k & r) == r)
                y |= 1ull << (m - 1 - r);
        return (x << m) | y;
    };
    for (int c = 0; c < 52 - 2 * m; c++)
        vdc[c] = code(2 * m + c);
    // Gauss-Jordan over GF(2) on the rows of A (row r: code bit r as a mask over j).
    const int n = 2 * m;
    std::vector<uint64_t> rowA(n, 0), rowI(n, 0);
    for (int r = 0; r < n; r++) {
        for (int j = 0; j < n; j++)
            if ((code(j) >> r) & 1)
                rowA[r] |= 1ull << j;
        rowI[r] = 1ull << r;
    }
    for (int col = 0; col < n; col++) {
        int piv = col;
        while (piv < n && !((rowA[piv] >> col) & 1))
            piv++;
        if (piv == n)
            throw std::runtime_error("Sobol pixel map is singular");
        std::swap(rowA[piv], rowA[col]);
        std::swap(rowI[piv], rowI[col]);
        for (int r = 0; r < n; r++)
            if (r != col && ((rowA[r] >> col) & 1)) {
                rowA[r] ^= rowA[col];
                rowI[r] ^= rowI[col];
            }
    }
    // rowI is now A^-1 by rows: bit j of the index for code bit c is bit c of rowI[j].
    for (int c = 0; c < n; c++)
        for (int j = 0; j < n; j++)
            if ((rowI[j] >> c) & 1)
                vdc_inv[c] |= 1ull << j;
}

template <typename T>
int upload(pg_scene& S, const std::vector<T>& v, const T** out) {
    if (v.empty()) {
        *out = nullptr;
        return PG_OK;
    }
    void* p = nullptr;
    PG_HIP(hipMalloc(&p, v.size() * sizeof(T)));
    S.allocations.push_back(p);
    PG_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = (const T*)p;
    return PG_OK;
}

template <typename T>
int dalloc(std::vector<void*>& list, int64_t n, T** out) {
    void* p = nullptr;
    PG_HIP(hipMalloc(&p, std::max<int64_t>(n, 1) * sizeof(T)));
    list.push_back(p);
    *out = (T*)p;
    return PG_OK;
}

template <typename T>
int dalloc(pg_scene& S, int64_t n, T** out) {
    return dalloc(S.allocations, n, out);
}

// Frees one list of per-path buffers before a larger set replaces it, after
// the device has finished every launch that reads them.  The recorded bounce
// graph that points at them is re-recorded first (paths_gen changes).
int release_state(std::vector<void*>& list) {
    if (list.empty())
        return PG_OK;
    PG_HIP(hipDeviceSynchronize());
    for (void* p : list)
        PG_HIP(hipFree(p));
    list.clear();
    return PG_OK;
}

// AtmosphereColor (color.cpp:37-89): single scattering along the view ray from
// the ground (8 samples), each lit through 4 samples towards the sun, Rayleigh
// and Mie coefficients and phase functions, x10 x sunColor.  M supplies the
// scalar type and exp/sqrt/pow/acos/sin/cos: pine's float approximations
// (pstd), or double precision with correctly rounded functions for the
// `accurate` table (where exp's overflow threshold, not float's, decides which
// texels below the horizon overflow).
struct PstdMath {
    using T = float;
    static __device__ float exp(float x) { return pstd::exp(x); }
    static __device__ float sqrt(float x) { return pstd::sqrt(x); }
    static __device__ float pow(float x, float e) { return pstd::pow(x, e); }
    static __device__ float acos(float x) { return pstd::acos(x); }
    static __device__ float sin(float x) { return pstd::sin(x); }
    static __device__ float cos(float x) { return pstd::cos(x); }
};
struct ExactMath {
    using T = double;
    static __device__ double exp(double x) { return ::exp(x); }
    static __device__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ double pow(double x, double e) { return ::pow(x, e); }
    static __device__ double acos(double x) { return ::acos(x); }
    static __device__ double sin(double x) { return ::sin(x); }
    static __device__ double cos(double x) { return ::cos(x); }
};

template <class T>
struct AV3 {
    T x, y, z;
};
template <class T>
__device__ AV3<T> av3(T x, T y, T z) { return AV3<T>{x, y, z}; }
template <class T>
__device__ AV3<T> operator+(AV3<T> a, AV3<T> b) { return av3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T>
__device__ AV3<T> operator*(AV3<T> a, AV3<T> b) { return av3(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T>
__device__ AV3<T> operator*(AV3<T> a, T s) { return av3(a.x * s, a.y * s, a.z * s); }
template <class T>
__device__ T adot(AV3<T> a, AV3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Sphere::ComputeT (geometry.cpp:114-126) about the origin, tmin 0.
template <class M, class T = typename M::T>
__device__ T atmo_sphere_t(AV3<T> ro, AV3<T> rd, T r) {
    const T a = adot(rd, rd), b = T(2) * adot(ro, rd);
    const T c = adot(ro, ro) - r * r;
    T d = b * b - T(4) * a * c;
    if (d <= T(0))
        return T(-1);
    d = M::sqrt(d);
    T t = (-b - d) / (T(2) * a);
    if (t < T(0))
        t = (-b + d) / (T(2) * a);
    return t;
}

template <class M, class T = typename M::T>
__device__ AV3<T> atmosphere_sky(AV3<T> dir, AV3<T> sun, AV3<T> sun_color) {
    const AV3<T> betaR = av3(T(3.8e-6f), T(13.5e-6f), T(33.1e-6f));
    const T betaM = T(21e-6f);
    const T atmosphereRadius = T(6420e3f), earthRadius = T(6360e3f);
    const bool along_sun = dir.x == sun.x && dir.y == sun.y && dir.z == sun.z;
    const T Hr = T(1) / T(7995), Hm = along_sun ? T(1) / T(10000) : T(1) / T(1200);
    const T mu = adot(dir, sun), g = T(0.76f), pi = T(3.14159265358979323846);
    const T phaseR = T(3) / (T(16) * pi) * (T(1) + mu * mu);
    const T phaseM = T(3) / (T(8) * pi) * (T(1) - g * g) * (T(1) + mu * mu) /
                     ((T(2) + g * g) * M::pow(T(1) + g * g - T(2) * g * mu, T(1.5)));
    const AV3<T> o = av3(T(0), earthRadius, T(0));
    const T seg = atmo_sphere_t<M>(o, dir, atmosphereRadius) / T(8);
    T tcur = T(0), odR = T(0), odM = T(0);
    AV3<T> sumR = av3(T(0), T(0), T(0)), sumM = sumR;
    for (int i = 0; i < 8; i++) {
        const AV3<T> sp = o + dir * (tcur + seg * T(0.5));
        const T height = M::sqrt(adot(sp, sp)) - earthRadius;
        const T hr = M::exp(-height * Hr) * seg, hm = M::exp(-height * Hm) * seg;
        odR += hr;
        odM += hm;
        const T segL = atmo_sphere_t<M>(sp, sun, atmosphereRadius) / T(4);
        T odRL = T(0), odML = T(0);
        const T lo2 = adot(sp, sp), ld2 = adot(sun, sun), lod = adot(sp, sun);
        T lt = segL * T(0.5);
        int j = 0;
        for (; j < 4; j++, lt += segL) {
            // color.cpp:66-72, as written (a squared distance minus a radius)
            const T hl = lo2 + lt * lt * ld2 + T(2) * lt * lod - earthRadius;
            if (hl < T(0))
                break;
            odRL += M::exp(-hl * Hr) * segL;
            odML += M::exp(-hl * Hm) * segL;
        }
        if (j == 4) {
            const AV3<T> tau = betaR * (odR + odRL) + av3(betaM, betaM, betaM) * (odM + odML);
            const AV3<T> tr = av3(M::exp(-tau.x), M::exp(-tau.y), M::exp(-tau.z));
            sumR = sumR + tr * hr;
            sumM = sumM + tr * hm;
        }
        tcur += seg;
    }
    return (sumR * betaR * phaseR + sumM * av3(betaM, betaM, betaM) * phaseM) * T(10) * sun_color;
}

// The Atmosphere constructor's table (light.cpp:81-92): texel (x, y) holds the
// colour towards UniformSphereSampling(x / size.x, y / size.y); one more
// thread computes the colour along the sun.  With pine's arithmetic, NaN
// entries are kept: pine's HasNaN() test folds to false under its -ffast-math
// build.  The accurate table applies the test as written (NaN -> 0).
template <class M>
__global__ void k_atmosphere_table(f3 sunf, f3 colf, int w, int h, float* out) {
    using T = typename M::T;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)w * h;
    if (t > n)
        return;
    const AV3<T> sun = av3(T(sunf.x), T(sunf.y), T(sunf.z)), col = av3(T(colf.x), T(colf.y), T(colf.z));
    AV3<T> c;
    if (t == n) {
        c = atmosphere_sky<M>(sun, sun, col);
    } else {
        const int y = (int)(t / w), x = (int)(t - (int64_t)y * w);
        const T u0 = T((float)x / (float)w), u1 = T((float)y / (float)h);
        const T phi = u0 * T(3.14159265358979323846) * T(2), theta = M::acos(T(1) - T(2) * u1);
        const T st = M::sin(theta);
        const AV3<T> d = av3(st * M::cos(phi), st * M::sin(phi), M::cos(theta));
        c = atmosphere_sky<M>(d, sun, col);
    }
    float r = (float)c.x, g = (float)c.y, b = (float)c.z;
    if (sizeof(T) == 8 && (isnan(r) || isnan(g) || isnan(b)))
        r = g = b = 0.0f;
    out[3 * t + 0] = r;
    out[3 * t + 1] = g;
    out[3 * t + 2] = b;
}

int atmosphere_setup(pg_scene& S) {
    const int64_t n = (int64_t)S.atmo_w * S.atmo_h;
    float* d = nullptr;
    PG_HIP(hipMalloc(&d, (n + 1) * 3 * sizeof(float)));
    S.allocations.push_back(d);
    const dim3 grid((unsigned)((n + 1 + 255) / 256));
    if (S.atmo_accurate)
        hipLaunchKernelGGL(k_atmosphere_table<ExactMath>, grid, dim3(256), 0, 0, S.dev.sky_dir, S.dev.sky_color,
                           S.atmo_w, S.atmo_h, d);
    else
        hipLaunchKernelGGL(k_atmosphere_table<PstdMath>, grid, dim3(256), 0, 0, S.dev.sky_dir, S.dev.sky_color,
                           S.atmo_w, S.atmo_h, d);
    PG_HIP(hipGetLastError());
    float sun[3];
    PG_HIP(hipMemcpy(sun, d + 3 * n, sizeof(sun), hipMemcpyDeviceToHost));
    S.dev.atmo_colors = d;
    S.dev.atmo_w = S.atmo_w;
    S.dev.atmo_h = S.atmo_h;
    S.dev.atmo_sun_le = mk(sun[0], sun[1], sun[2]);
    return PG_OK;
}

int finish_scene(pg_scene& S, const pg_build_options* opts) {
    if (opts)
        S.build_opts = *opts;
    spec::spectral_tables(S.dev.spec);
    if (int rc = pg_accel_build(S.accel, opts))
        return rc;
    std::vector<int> primes, sums;
    std::vector<uint16_t> perms;
    if (S.atmo_w > 0)
        if (int rc = atmosphere_setup(S))
            return rc;
    halton_tables(primes, sums, perms);
    if (!S.halton_permute) {
        // randomizeStrategy none: RadicalInverse (lowdiscrepancy.h:13-26) is the
        // scrambled form with the identity permutation.
        for (size_t d = 0, o = 0; d < primes.size(); o += primes[d], d++)
            for (int j = 0; j < primes[d]; j++)
                perms[o + j] = (uint16_t)j;
    }
    // Per-dimension records of the radical inverse (base, permutation offset,
    // 1/base, perm[0] / (base + 1); division magic; digit-chunk tables).
    std::vector<int4> dimtab(3 * primes.size());
    std::vector<uint16_t> chunk16;
    std::vector<uint32_t> chunk32;
    // udiv_magic's constants for a divisor b: l = ceil(log2 b), m, post-shift l - 1.
    auto magic = [](uint64_t b, uint32_t& m, int& post) {
        int l = 0;
        while ((uint64_t(1) << l) < b)
            l++;
        m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - b)) / b + 1);
        post = l - 1;
    };
    for (size_t d = 0; d < primes.size(); d++) {
        const float inv_base = 1.0f / (float)primes[d];
        const float series = (float)perms[sums[d]] / ((float)primes[d] + 1.0f);
        int ib, sb;
        std::memcpy(&ib, &inv_base, 4);
        std::memcpy(&sb, &series, 4);
        dimtab[3 * d] = make_int4(primes[d], sums[d], ib, sb);
        const uint64_t b = (uint64_t)primes[d];
        uint32_t m = 0;
        int post = 0;
        magic(b, m, post);
        dimtab[3 * d + 1] = make_int4((int32_t)m, post, 0, 0);
        dimtab[3 * d + 2] = make_int4(0, 0, 0, 0);
        // Chunk tables: k digits at a time, B = b^k <= kChunkEntries, for the
        // bases where k >= 2 and a 32-bit index has at most kMaxFullChunks
        // full chunks before its last part.
        int k = 0;
        uint64_t B = 1;
        while (B * b <= (uint64_t)kChunkEntries) {
            B *= b;
            k++;
        }
        int digits = 0;
        for (uint64_t v = 0xFFFFFFFFull; v; v /= b)
            digits++;
        if (k < 2 || (digits + k - 1) / k - 1 > kMaxFullChunks)
            continue;
        const uint16_t* perm = perms.data() + sums[d];
        const size_t off16 = chunk16.size(), off32 = chunk32.size();
        for (uint64_t x = 0; x < B; x++) {  // natural: the digits x has
            uint64_t v = 0, y = x;
            uint32_t nd = 0;
            for (; y; y /= b, nd++)
                v = v * b + perm[y % b];
            chunk16.push_back((uint16_t)(v | (nd << 12)));
        }
        for (uint64_t x = 0; x < B; x++) {  // full: all k digits
            uint64_t v = 0, y = x;
            for (int j = 0; j < k; j++, y /= b)
                v = v * b + perm[y % b];
            chunk16.push_back((uint16_t)v);
        }
        for (uint64_t n = 0, p = 1; n <= (uint64_t)k; n++, p *= b)
            chunk32.push_back((uint32_t)p);
        float ipow = 1.0f;
        for (int n = 0; n <= 32; n++) {
            uint32_t bits;
            std::memcpy(&bits, &ipow, 4);
            chunk32.push_back(bits);
            ipow *= inv_base;
        }
        uint32_t mB = 0;
        int postB = 0;
        magic(B, mB, postB);
        dimtab[3 * d + 1] = make_int4((int32_t)m, post, k, (int32_t)off32);
        dimtab[3 * d + 2] = make_int4((int32_t)B, (int32_t)mB, postB, (int32_t)off16);
    }
    SamplerDev& sd = S.dev.sampler;
    sd.kind = S.sampler_kind;
    halton_setup(sd.halton, S.dev.width, S.dev.height);
    int rc;
    if (S.sampler_kind == kSamplerSobol) {
        // SobolSampler (sampler.cpp:129-134): m = log2 of the larger film side
        // rounded up to a power of two.
        int m = 0;
        while ((1 << m) < std::max(S.dev.width, S.dev.height))
            m++;
        if (m > 25)
            return pg_fail(PG_E_PARSE, "Sobol sampler: film larger than 2^25 pixels a side");
        std::vector<uint32_t> C;
        std::vector<uint64_t> vdc, vdc_inv;
        sobol_matrices(C);
        sobol_pixel_tables(m, vdc, vdc_inv);
        sd.sobol.log2Scale = m;
        sd.sobol.randomize = S.sobol_randomize;
        if ((rc = upload(S, C, &sd.sobol.C)) || (rc = upload(S, vdc, &sd.sobol.vdc)) ||
            (rc = upload(S, vdc_inv, &sd.sobol.vdc_inv)))
            return rc;
    }
    if ((rc = upload(S, perms, &sd.halton.perms)) || (rc = upload(S, primes, &sd.halton.primes)) ||
        (rc = upload(S, dimtab, &sd.halton.dims)) || (rc = upload(S, chunk16, &sd.halton.chunk16)) ||
        (rc = upload(S, chunk32, &sd.halton.chunk32)) ||
        (rc = upload(S, sums, &sd.halton.primeSums)) || (rc = upload(S, S.prim_shape, &S.dev.prim_shape)) ||
        (rc = upload(S, S.tri_verts, &S.dev.tri_verts)) || (rc = upload(S, S.shapes, &S.dev.shapes)) ||
        (rc = upload(S, S.materials, &S.dev.materials)) || (rc = upload(S, S.lights, &S.dev.lights)) ||
        (rc = upload(S, S.light_func, &S.dev.light_func)) || (rc = upload(S, S.light_cdf, &S.dev.light_cdf)) ||
        (rc = upload(S, S.node_progs, &S.dev.node_prog)))
        return rc;
    // Media: density grids and brick majorants, then the MediumDev table.
    S.lds_medium = -1;
    int bricked = 0;
    for (size_t k = 0; k < S.media.size(); k++) {
        MediumDev& m = S.media[k];
        // Pine's x-fastest densities regrouped into 4 x 4 x 2-voxel blocks.
        std::vector<float> blocked;
        if (m.type == kMediumGrid) {
            int32_t nb[3];
            int64_t floats = 0;
            density_blocks(m.size, nb, floats);
            if (floats > (int64_t)INT32_MAX)
                return pg_fail(PG_E_ARG, "density grid too large (more than 2^31 voxels with padding)");
            m.dblocks[0] = nb[0];
            m.dblocks[1] = nb[1];
            blocked.assign((size_t)floats, 0.0f);
            const std::vector<float>& lin = S.med_density[k];
            for (int z = 0; z < m.size[2]; z++)
                for (int y = 0; y < m.size[1]; y++)
                    for (int x = 0; x < m.size[0]; x++)
                        blocked[(size_t)(dens_off_x(x) + dens_off_y(m, y) + dens_off_z(m, z))] =
                            lin[(size_t)x + (size_t)m.size[0] * (y + (size_t)m.size[1] * z)];
        }
        if ((rc = upload(S, blocked, &m.density)) || (rc = upload(S, S.med_majorant[k], &m.majorant)))
            return rc;
        if (m.brick > 0) {
            bricked++;
            if (S.med_majorant[k].size() <= kMaxLdsMajorants)
                S.lds_medium = (int)k;
        }
    }
    if (bricked != 1)
        S.lds_medium = -1;
    if ((rc = upload(S, S.media, &S.dev.media)))
        return rc;
    S.dev.n_media = (int32_t)S.media.size();
    if ((rc = dalloc(S, (int64_t)S.dev.width * S.dev.height * 4, &S.d_film)))
        return rc;
    PG_HIP(hipMemset(S.d_film, 0, (size_t)S.dev.width * S.dev.height * 4 * sizeof(float)));
    return PG_OK;
}

// Grows one kind of per-path buffer of a slot to n entries: frees the old
// set first (release_state waits for every launch that may read it) and moves
// the slot to a new generation, so recorded bounce graphs are re-recorded.
template <typename F>
int grow_state(PathState& W, int64_t& cap, std::vector<void*>& list, int64_t n, F&& alloc) {
    if (n <= cap)
        return PG_OK;
    W.gen = ++g_paths_gen;
    cap = 0;
    if (int rc = release_state(list))
        return rc;
    if (int rc = alloc(list))
        return rc;
    cap = n;
    return PG_OK;
}

int ensure_paths(PathState& W, const pg_scene& S, int64_t n) {
    PathsDev& P = W.P;
    int rc = grow_state(W, W.core, W.core_allocs, n, [&](std::vector<void*>& A) {
        int r;
        if ((r = dalloc(A, 3 * n, &P.L)) || (r = dalloc(A, 3 * n, &P.beta)) ||
            (r = dalloc(A, n, &P.bsdf_pdf)) || (r = dalloc(A, n, &P.sidx)) || (r = dalloc(A, n, &P.sdim)) ||
            (r = dalloc(A, n, &P.depth)) || (r = dalloc(A, 2 * n, &P.pfilm)) ||
            (r = dalloc(A, n, &P.queue_in)) || (r = dalloc(A, n, &P.queue_out)) ||
            (r = dalloc(A, n, &P.rays_in)) || (r = dalloc(A, n, &P.rays_out)) ||
            (r = dalloc(A, n, &P.hits)) || (r = dalloc(A, n, &P.uvs)) || (r = dalloc(A, 8, &P.counters)) ||
            (r = dalloc(A, n, &P.shadow_rays)) || (r = dalloc(A, n, &P.shadow_path)) ||
            (r = dalloc(A, 3 * n, &P.shadow_contrib)) || (r = dalloc(A, n, &P.occluded)) ||
            (r = dalloc(A, n, &P.border)))
            return r;
        return PG_OK;
    });
    if (rc)
        return rc;
    if (S.general) {
        rc = grow_state(W, W.general, W.general_allocs, n, [&](std::vector<void*>& A) {
            int r;
            if ((r = dalloc(A, n, &P.medium)) || (r = dalloc(A, 9 * n, &P.vtx)) ||
                (r = dalloc(A, 4 * n, &P.vinfo)) || (r = dalloc(A, n, &P.queue_b)) ||
                (r = dalloc(A, 16 * n, &P.sh_rec)) || (r = dalloc(A, 3 * n, &P.sh_tr)) ||
                (r = dalloc(A, n, &P.seg_rays[0])) || (r = dalloc(A, n, &P.seg_rays[1])) ||
                (r = dalloc(A, n, &P.seg_rec[0])) || (r = dalloc(A, n, &P.seg_rec[1])) ||
                (r = dalloc(A, n, &P.seg_med[0])) || (r = dalloc(A, n, &P.seg_med[1])) ||
                (r = dalloc(A, n, &P.seg_hits)) || (r = dalloc(A, n, &P.seg_uvs)) ||
                (r = dalloc(A, 2 * n, &P.mev)) || (r = dalloc(A, n, &P.seg_trf)))
                return r;
            return PG_OK;
        });
        if (rc)
            return rc;
    }
    if (!S.node_progs.empty()) {
        rc = grow_state(W, W.nodes, W.node_allocs, n, [&](std::vector<void*>& A) {
            int r;
            if ((r = dalloc(A, n, &P.node_mat)) || (r = dalloc(A, nodes::kStack * n, &P.node_stack)))
                return r;
            return PG_OK;
        });
    }
    return rc;
}

int ensure_spectral(PathState& W, int64_t n) {
    PathsDev& P = W.P;
    return grow_state(W, W.spectral, W.spectral_allocs, n, [&](std::vector<void*>& A) {
        int r;
        if ((r = dalloc(A, n, &P.lam)) || (r = dalloc(A, 4 * n, &P.L4)) || (r = dalloc(A, 4 * n, &P.beta4)) ||
            (r = dalloc(A, 4 * n, &P.sh4)))
            return r;
        return PG_OK;
    });
}

int ensure_guide(PathState& W, int64_t n, int rec_depth) {
    PathsDev& P = W.P;
    int rc = grow_state(W, W.guide, W.guide_allocs, n, [&](std::vector<void*>& A) {
        int r;
        if ((r = dalloc(A, (int64_t)guide::kParamStride * n, &P.gparams)) || (r = dalloc(A, n, &P.gflag)) ||
            (r = dalloc(A, n, &P.shadow_rec)) || (r = dalloc(A, (int64_t)guide::kIn * n, &P.gx)))
            return r;
        return PG_OK;
    });
    if (rc)
        return rc;
    const int64_t slots = n * rec_depth;
    return grow_state(W, W.records, W.record_allocs, slots, [&](std::vector<void*>& A) {
        int r;
        if ((r = dalloc(A, (int64_t)guide::kIn * slots, &P.rec_x)) || (r = dalloc(A, 3 * slots, &P.rec_dir)) ||
            (r = dalloc(A, slots, &P.rec_w)) || (r = dalloc(A, slots, &P.rec_pb)) ||
            (r = dalloc(A, slots, &P.rec_pdf)) || (r = dalloc(A, 3 * slots, &P.rec_lb)) ||
            (r = dalloc(A, 3 * slots, &P.rec_beta)) || (r = dalloc(A, 2, &P.rec_acc)) ||
            (r = dalloc(A, (int64_t)guide::kIn * slots, &P.c_x)) || (r = dalloc(A, 3 * slots, &P.c_dir)) ||
            (r = dalloc(A, slots, &P.c_w)) || (r = dalloc(A, slots, &P.c_pb)) ||
            (r = dalloc(A, (slots + 255) / 256, &P.rec_bcnt)) || (r = dalloc(A, (slots + 255) / 256, &P.rec_bsum)) ||
            (r = dalloc(A, (slots + 255) / 256, &P.rec_boff)) || (r = dalloc(A, 1, &P.rec_n)))
            return r;
        return PG_OK;
    });
}

// Device bytes per path of the buffers ensure_paths, ensure_spectral and
// ensure_guide allocate for a batch (rounded up; the small fixed buffers aside).
int64_t path_state_bytes(const pg_scene& S, bool guided, int rec_depth) {
    const int64_t ray = sizeof(pg_ray), hit = sizeof(pg_hit) + sizeof(pg_hit_uv);
    int64_t b = 4 * (3 + 3 + 1 + 2 + 3) + 8 + 4 * (1 + 1 + 2 + 1 + 1) + 3 * ray + hit + 8;
    if (S.general)
        b += 4 * (1 + 9 + 4 + 1 + 16 + 3 + 2 + 2 + 1) + 2 * ray + hit + 2 * sizeof(float4);
    if (!S.node_progs.empty())
        b += sizeof(MaterialDev) + nodes::kStack * sizeof(f3);
    if (S.spectral)
        b += 4 * (1 + 4 + 4 + 4);
    if (guided)
        b += 4 * (guide::kParamStride + 1 + 1 + guide::kIn) +
             (int64_t)rec_depth * 4 * (2 * guide::kIn + 3 + 3 + 1 + 1 + 1 + 3 + 3 + 1 + 1);
    return b;
}

// Samples per pixel of one batch when the caller leaves it open: the scene's
// batch_paths, at most half of the free device memory (counting what the
// slot's buffers would free on growing), at least one sample per pixel.
int auto_batch_spp(const PathState& W, const pg_scene& S, int spp, int64_t npix, bool guided, int rec_depth) {
    int64_t want = std::min<int64_t>((int64_t)spp * npix, S.batch_paths);
    const int64_t per_path = path_state_bytes(S, guided, rec_depth);
    if (want > W.core) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const int64_t held = W.core * per_path;
            want = std::min<int64_t>(want, ((int64_t)free_b + held) / 2 / per_path);
        }
    }
    want = std::min<int64_t>(want, INT32_MAX);
    return (int)std::max<int64_t>(1, std::min<int64_t>(spp, want / std::max<int64_t>(npix, 1)));
}

}  // namespace

// ----------------------------------------------------------------------------
// C ABI

static int load_common(pg_scene* S, const std::string& text, pg_scene** out,
                       const pg_build_options* opts) {
    try {
        S->root = parse_scene_text(text);
        build_scene(*S);
    } catch (const std::exception& e) {
        delete S;
        return pg_fail(PG_E_PARSE, e.what());
    }
    if (int rc = finish_scene(*S, opts)) {
        delete S;
        return rc;
    }
    if (int rc = register_paths(S)) {
        delete S;
        return rc;
    }
    *out = S;
    return PG_OK;
}

extern "C" int pg_scene_load_text(const char* text, const char* dir, const pg_build_options* opts,
                                  pg_scene** out) {
    if (!text || !out)
        return pg_fail(PG_E_ARG, "pg_scene_load_text: null argument");
    pg_scene* S = new pg_scene();
    S->dir = dir ? dir : ".";
    return load_common(S, text, out, opts);
}

extern "C" int pg_scene_load(const char* path, const pg_build_options* opts, pg_scene** out) {
    if (!path || !out)
        return pg_fail(PG_E_ARG, "pg_scene_load: null argument");
    std::string text;
    try {
        text = read_file(path);
    } catch (const std::exception& e) {
        return pg_fail(PG_E_IO, e.what());
    }
    pg_scene* S = new pg_scene();
    // GetFileDirectory (fileio.cpp:256): everything up to the last '/'.
    std::string p(path);
    size_t slash = p.find_last_of('/');
    S->dir = slash == std::string::npos ? "." : p.substr(0, slash);
    return load_common(S, text, out, opts);
}

extern "C" int pg_scene_get_info(const pg_scene* S, pg_scene_info* info) {
    if (!S || !info)
        return pg_fail(PG_E_ARG, "pg_scene_get_info: null argument");
    info->width = S->dev.width;
    info->height = S->dev.height;
    info->spp = S->spp;
    info->max_depth = S->max_depth;
    info->n_shapes = (int32_t)S->shapes.size();
    info->n_lights = (int32_t)S->lights.size();
    info->n_prims = (int64_t)S->prim_shape.size();
    info->last_render_ms = S->last_render_ms;
    info->last_rays = S->last_rays;
    info->n_media = (int32_t)S->media.size();
    info->general_mode = S->general ? 1 : 0;
    info->medium_steps = S->last_lookups;
    info->last_guided_ms = S->last_guide_ms;
    info->guide_steps = S->guide_steps;
    info->guide_records = S->guide_records;
    info->graph_captures = S->graph_captures;
    info->last_extension_rays = S->last_ext_rays;
    info->paths_in_flight = S->batch_paths;
    info->apply_tone_mapping = S->tone_map ? 1 : 0;
    info->guide_infer_ms = S->infer_ms;
    info->guide_infer_queries = S->infer_queries;
    info->guide_step_ms = S->guide_step_ms;
    pg_accel_stats as{};
    if (pg_accel_get_stats(S->accel, &as) == PG_OK) {
        info->bvh_width = as.width;
        info->trace_block = as.block;
        info->trace_flags = as.launch_flags;
    }
    return PG_OK;
}

constexpr int kMaxShadowSegments = 1024;
constexpr int kMaxBounces = 1 << 16;
// Guide inference: one workgroup per 128 vertices up to two per CU (each
// stages 59 KB of weights, then its waves stride over the rest).
int infer_blocks(int32_t n) {
    return (int)std::min<int64_t>((n + 127) / 128, 512);
}

// Grid of a grid-stride tracking kernel: enough workgroups to fill the chip,
// each staging the brick majorants once.
static int tracking_grid(int64_t n, int block) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, 4096));
}

// Grid of a persistent-lane tracking kernel: the chip's resident workgroups
// (four or five 256-lane workgroups per CU at these kernels' registers), fewer
// for short queues.
static int lane_grid(int64_t n, int block) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, 1280));
}

// Calls f(sampler variant, media variant) as integral constants, so the
// tracking kernels are launched in the instantiation that fits the render.
template <class F>
static void launch_tracking(int ks, int km, F&& f) {
    using S0 = std::integral_constant<int, kSamplerAny>;
    using S1 = std::integral_constant<int, kSamplerHalton32>;
    using S2 = std::integral_constant<int, kSamplerSobolOnly>;
    using M0 = std::integral_constant<int, kMediaAny>;
    using M1 = std::integral_constant<int, kMediaBricks>;
    using M2 = std::integral_constant<int, kMediaOneBrickGrid>;
    if (km == kMediaOneBrickGrid) {
        if (ks == kSamplerHalton32)
            f(S1{}, M2{});
        else if (ks == kSamplerSobolOnly)
            f(S2{}, M2{});
        else
            f(S0{}, M2{});
    } else if (km == kMediaBricks) {
        if (ks == kSamplerHalton32)
            f(S1{}, M1{});
        else if (ks == kSamplerSobolOnly)
            f(S2{}, M1{});
        else
            f(S0{}, M1{});
    } else {
        if (ks == kSamplerHalton32)
            f(S1{}, M0{});
        else if (ks == kSamplerSobolOnly)
            f(S2{}, M0{});
        else
            f(S0{}, M0{});
    }
}

// Shadow-segment passes of the split bounce run on device-held counts (no
// host round trip): pass 0 reads the record count (counters[1]), pass p
// counts its continuations at counters[4 + p].  A shadow ray crosses a convex
// material-less shape at most twice, so a scene with m such shapes gets
// 1 + 2m passes, at most kDeviceSegmentPasses; counters[4 + passes - 1] holds
// whatever is left after the last one (a non-convex shape, or more shapes),
// and the host then finishes those segments before the records are resolved.
constexpr int kDeviceSegmentPasses = 4;
static int device_segment_passes(const pg_scene* S) {
    return std::min(kDeviceSegmentPasses, 1 + 2 * S->materialless_shapes);
}

// One bounce in general mode; `active` becomes the next queue's length.  One
// host synchronisation per bounce: every launch after the shading kernel
// reads its count from device memory and is sized for `active` (records,
// segments and direction samples never outnumber the bounce's paths).
static int bounce_general(pg_scene* S, PathsDev& Q, int32_t& active, hipStream_t st, int64_t& rays) {
    const int B = 256;
    const size_t lds = S->lds_medium >= 0 ? S->med_majorant[S->lds_medium].size() * sizeof(float) : 0;
    // PG_RENDER_LANE_TRACKING: brick-tracked media run their tracking as
    // persistent lanes.
    const bool lanes = S->lane_tracking && S->track_media != kMediaAny;
    if (!S->media.empty())
        launch_tracking(S->track_sampler, S->track_media, [&](auto ks, auto km) {
            if constexpr (decltype(km)::value != kMediaAny) {
                if (lanes) {
                    hipLaunchKernelGGL((k_medium_track_lanes<decltype(ks)::value, decltype(km)::value>),
                                       dim3(lane_grid(active, B)), dim3(B), lds, st, S->dev, Q, active,
                                       S->lds_medium);
                    return;
                }
            }
            hipLaunchKernelGGL((k_medium_track<decltype(ks)::value, decltype(km)::value>),
                               dim3(tracking_grid(active, B)), dim3(B), lds, st, S->dev, Q, active, S->lds_medium);
        });
    {
        const bool nodes = !S->node_progs.empty();
        auto kern = S->spectral ? (nodes ? k_shade_gen<true, true> : k_shade_gen<true, false>)
                                : (nodes ? k_shade_gen<false, true> : k_shade_gen<false, false>);
        hipLaunchKernelGGL(kern, dim3((active + B - 1) / B), dim3(B), 0, st, S->dev, Q, active);
    }
    rays += active;
    const unsigned grid_all = (unsigned)((active + B - 1) / B);
    int32_t* C = Q.counters;
    const int passes = device_segment_passes(S);
    // IntersectTr: trace each segment, track its medium, continue through
    // material-less surfaces until every record is resolved.
    int src = 0;
    if (!lanes) {
        for (int pass = 0; pass < passes; pass++) {
            const int32_t* d_nseg = pass == 0 ? C + 1 : C + 4 + pass - 1;
            if (int rc = pg_trace_closest_dev(S->accel, Q.seg_rays[src], d_nseg, active, Q.seg_hits, Q.seg_uvs, st))
                return rc;
            launch_tracking(S->track_sampler, S->track_media, [&](auto ks, auto km) {
                hipLaunchKernelGGL((k_shadow_tr<decltype(ks)::value, decltype(km)::value, false>),
                                   dim3(tracking_grid(active, B)), dim3(B), lds, st, S->dev, Q, 0, src,
                                   S->lds_medium, d_nseg, C + 4 + pass);
            });
            src = 1 - src;
        }
        if (S->dev.guided)
            hipLaunchKernelGGL(k_guide_infer<0>, dim3(infer_blocks(active)), dim3(kInferThreads), 0, st, S->dev, Q,
                               active, C + 2);
        const int32_t* ovf = C + 4 + passes - 1;
        hipLaunchKernelGGL(S->spectral ? k_resolve_gen<true> : k_resolve_gen<false>, dim3(grid_all), dim3(B), 0, st,
                           S->dev, Q, 0, C + 1, ovf);
        hipLaunchKernelGGL(S->spectral ? k_shade_gen_b<true> : k_shade_gen_b<false>, dim3(grid_all), dim3(B), 0, st,
                           S->dev, Q, 0, C + 2, ovf);
    }
    int32_t counts[8];
    PG_HIP(hipMemcpyAsync(counts, C, sizeof(counts), hipMemcpyDeviceToHost, st));
    PG_HIP(hipStreamSynchronize(st));
    const int32_t n_rec = counts[1], n_b = counts[2];
    int32_t n_seg = n_rec;
    if (!lanes) {
        rays += n_rec;
        for (int pass = 0; pass + 1 < passes; pass++)
            rays += counts[4 + pass];
        n_seg = counts[4 + passes - 1];
        if (n_seg == 0) {
            active = counts[0];
            return PG_OK;
        }
    }
    // Host-driven segment passes: the lane-tracking variant throughout, or the
    // segments left after the device-driven passes (the resolve and direction
    // kernels above then did nothing).
    for (int pass = 0; n_seg > 0 && pass < kMaxShadowSegments; pass++) {
        rays += n_seg;
        PG_HIP(hipMemsetAsync(C + 3, 0, sizeof(int32_t), st));
        if (int rc = pg_trace_closest(S->accel, Q.seg_rays[src], n_seg, Q.seg_hits, Q.seg_uvs, st))
            return rc;
        launch_tracking(S->track_sampler, S->track_media, [&](auto ks, auto km) {
            if constexpr (decltype(km)::value != kMediaAny) {
                if (lanes) {
                    hipLaunchKernelGGL((k_shadow_track_lanes<decltype(ks)::value, decltype(km)::value>),
                                       dim3(lane_grid(n_seg, B)), dim3(B), lds, st, S->dev, Q, n_seg, src,
                                       S->lds_medium);
                    hipLaunchKernelGGL((k_shadow_tr<decltype(ks)::value, decltype(km)::value, true>),
                                       dim3((n_seg + B - 1) / B), dim3(B), 0, st, S->dev, Q, n_seg, src, -1,
                                       nullptr, nullptr);
                    return;
                }
            }
            hipLaunchKernelGGL((k_shadow_tr<decltype(ks)::value, decltype(km)::value, false>),
                               dim3(tracking_grid(n_seg, B)), dim3(B), lds, st, S->dev, Q, n_seg, src,
                               S->lds_medium, nullptr, nullptr);
        });
        int32_t next = 0;
        PG_HIP(hipMemcpyAsync(&next, C + 3, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        PG_HIP(hipStreamSynchronize(st));
        n_seg = next;
        src = 1 - src;
    }
    if (lanes && S->dev.guided && n_b > 0)
        hipLaunchKernelGGL(k_guide_infer<0>, dim3(infer_blocks(n_b)), dim3(kInferThreads), 0, st, S->dev, Q, n_b,
                           nullptr);
    if (n_rec > 0)
        hipLaunchKernelGGL(S->spectral ? k_resolve_gen<true> : k_resolve_gen<false>, dim3((n_rec + B - 1) / B),
                           dim3(B), 0, st, S->dev, Q, n_rec, nullptr, nullptr);
    if (n_b > 0)
        hipLaunchKernelGGL(S->spectral ? k_shade_gen_b<true> : k_shade_gen_b<false>, dim3((n_b + B - 1) / B),
                           dim3(B), 0, st, S->dev, Q, n_b, nullptr, nullptr);
    PG_HIP(hipMemcpyAsync(counts, C, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PG_HIP(hipStreamSynchronize(st));
    active = counts[0];
    return PG_OK;
}

// Per-batch training of the guide during a training pass.
__global__ void k_fill_i32(int32_t* p, int32_t n, int32_t v) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        p[k] = v;
}

// Rays of one bounce: total[0] += extension + shadow rays, total[1] += extension rays.
__global__ void k_count_rays(const int32_t* d_active, const int32_t* d_shadow, unsigned long long* total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        total[0] += (unsigned long long)*d_active + (unsigned long long)*d_shadow;
        total[1] += (unsigned long long)*d_active;
    }
}

// Workgroups of a launch over a queue whose length is read in device memory
// (at most cap): about two lanes per lane the GPU holds at once, striding over
// the queue, so the bounces that find few or no paths cost short launches.
constexpr int64_t kDevCountLanes = 1 << 20;

unsigned dev_count_grid(int64_t cap, int block) {
    return (unsigned)std::min<int64_t>((cap + block - 1) / block, kDevCountLanes / block);
}

// The fused pipeline's bounce loop for up to `cap` paths with every count in
// device memory: bounce b reads its path count from the previous bounce's
// counters (bounce 0 from g_counts[0], written by k_raygen) and writes its
// own (g_counts + 4(b + 1): next paths, shadow rays); every launch is sized
// for cap.  A fused vertex always adds one to the path depth, so max_depth
// bounces finish every path.  Counters are cleared by a fill kernel, not a
// memset node.
static int record_bounces(pg_scene* S, const PathsDev& P, int64_t cap, hipStream_t st) {
    const int B = 256;
    const int nb = S->max_depth;
    const unsigned grid = dev_count_grid(cap, B);
    // k_shade takes one tile per workgroup (a strided loop costs it registers:
    // 152 VGPRs, 3 waves per SIMD), so its grid covers the capacity.
    const unsigned shade_grid = (unsigned)((cap + kShadeBlock - 1) / kShadeBlock);
    hipLaunchKernelGGL(k_fill_i32, dim3((4 * nb + B - 1) / B), dim3(B), 0, st, S->g_counts + 4, 4 * nb, 0);
    PathsDev Q = P;
    for (int b = 0; b < nb; b++) {
        const int32_t* d_active = S->g_counts + 4 * b;
        Q.counters = S->g_counts + 4 * (b + 1);
        if (int rc = pg_trace_closest_dev(S->accel, Q.rays_in, d_active, cap, Q.hits, Q.uvs, st))
            return rc;
        if (S->dev.guided && b < S->dev.g_depth) {
            hipLaunchKernelGGL(k_guide_gather, dim3(grid), dim3(B), 0, st, S->dev, Q, 0, d_active);
            hipLaunchKernelGGL(k_guide_infer<1>, dim3(infer_blocks((int32_t)cap)), dim3(kInferThreads), 0, st,
                               S->dev, Q, 0, d_active);
            hipLaunchKernelGGL(k_shade<true>, dim3(shade_grid), dim3(kShadeBlock), 0, st, S->dev, Q, 0, d_active);
        } else {
            hipLaunchKernelGGL(k_shade<false>, dim3(shade_grid), dim3(kShadeBlock), 0, st, S->dev, Q, 0, d_active);
        }
        if (int rc = pg_trace_occluded_dev(S->accel, Q.shadow_rays, Q.counters + 1, cap, Q.occluded, st))
            return rc;
        hipLaunchKernelGGL(k_resolve_shadow, dim3(grid), dim3(B), 0, st, Q, 0, Q.counters + 1);
        hipLaunchKernelGGL(k_count_rays, dim3(1), dim3(64), 0, st, d_active, Q.counters + 1, S->g_rays);
        std::swap(Q.queue_in, Q.queue_out);
        std::swap(Q.rays_in, Q.rays_out);
    }
    PG_HIP(hipGetLastError());
    return PG_OK;
}

// Records the bounce loop for `cap` paths into the scene's graph unless the
// one it holds was recorded for the same state.
static int ensure_graph(pg_scene* S, const PathsDev& P, int64_t cap) {
    pg_scene::GraphKey key;
    key.capacity = cap;
    key.accel_gen = accel_generation(S->accel);
    key.paths_gen = S->paths_gen;
    key.guided = S->dev.guided;
    key.g_depth = S->dev.g_depth;
    key.rr_depth = S->dev.rr_depth;
    key.oct = S->dev.g_oct;
    key.max_depth = S->max_depth;
    key.g_params = S->dev.g_params;
    key.g_wf = S->dev.g_wf;
    if (S->graph_exec && key == S->graph_key)
        return PG_OK;
    if (S->graph_exec) {
        (void)hipGraphExecDestroy(S->graph_exec);
        S->graph_exec = nullptr;
    }
    if (!S->g_counts) {
        if (int rc = dalloc(*S, 4 * (S->max_depth + 1), &S->g_counts))
            return rc;
        if (int rc = dalloc(*S, 2, &S->g_rays))
            return rc;
    }
    if (!S->capture_stream)
        PG_HIP(hipStreamCreateWithFlags(&S->capture_stream, hipStreamNonBlocking));
    PG_HIP(hipStreamBeginCapture(S->capture_stream, hipStreamCaptureModeThreadLocal));
    const int rc = record_bounces(S, P, cap, S->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(S->capture_stream, &g);
    if (rc != PG_OK || e != hipSuccess) {
        if (g)
            (void)hipGraphDestroy(g);
        return rc != PG_OK ? rc : pg_fail(PG_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&S->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        S->graph_exec = nullptr;
        return pg_fail(PG_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    }
    S->graph_key = key;
    S->graph_captures++;
    return PG_OK;
}

constexpr int64_t kAutoSteps = 32;
struct GuideTraining {
    pg_guide* G;
    int rec_depth;
    int steps_per_batch;  // 0: one epoch over each batch's records
    int batch;
    int64_t steps;         // Adam steps taken
    int64_t last_records;  // nonzero-weight records trained on
    // Events around each batch's Adam steps, read once the render is done
    // (no extra synchronisation inside the passes).
    std::vector<std::pair<hipEvent_t, hipEvent_t>> step_events;
    double step_ms() const {
        double t = 0.0;
        for (const auto& e : step_events) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e.first, e.second) == hipSuccess)
                t += ms;
        }
        return t;
    }
    ~GuideTraining() {
        for (const auto& e : step_events) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
    }
};

static uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// Renders spp samples per pixel into the (cleared) film; with `tr`, records
// the guided vertices of every batch and trains the guide after it.
// A render enqueued on a stream: its bracketing events and what the host
// reads once it is done.
struct RenderTicket {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool graph = false;
    int64_t rays = 0, ext_rays = 0;
    void release() {
        if (e0)
            (void)hipEventDestroy(e0);
        if (e1)
            (void)hipEventDestroy(e1);
        e0 = e1 = nullptr;
    }
};

// Enqueues a whole render on st.  Without training and in the fused pipeline
// nothing waits on the host (graph replay per batch); training passes and the
// split bounce read counts back per bounce or per batch.
// W: the path-state slot the render uses; the caller holds the device's
// render lock until the render has finished on its stream.
static int render_enqueue(pg_scene* S, PathState& W, int spp, int spb_req, hipStream_t st, GuideTraining* tr,
                          int sample_offset, RenderTicket* t) {
    const int64_t npix = (int64_t)S->dev.width * S->dev.height;
    if (S->spectral && S->dev.guided)
        return pg_fail(PG_E_ARG, "spectral rendering is not combined with the learned guide");
    int spb = spb_req > 0 ? spb_req : auto_batch_spp(W, *S, spp, npix, S->dev.guided, tr ? tr->rec_depth : 0);
    spb = std::min(spb, spp);
    const int64_t n_paths = npix * spb;
    if (n_paths > (int64_t)INT32_MAX)
        return pg_fail(PG_E_ARG, "pg_render: batch too large");
    if (int rc = ensure_paths(W, *S, n_paths))
        return rc;
    if (S->dev.guided) {
        if (int rc = ensure_guide(W, n_paths, tr ? tr->rec_depth : 0))
            return rc;
    }
    if (S->spectral) {
        if (int rc = ensure_spectral(W, n_paths))
            return rc;
    }
    if (S->general && !S->d_lookups) {
        if (int rc = dalloc(*S, 1, &S->d_lookups))
            return rc;
    }
    // The scene's view of the slot: buffers of kinds this scene does not use
    // are left out (k_raygen writes a path's medium only where one is kept).
    S->paths = W.P;
    S->paths.lookups = S->general ? S->d_lookups : nullptr;
    if (!S->general)
        S->paths.medium = nullptr;
    S->paths_gen = W.gen;
    PG_HIP(hipEventCreate(&t->e0));
    PG_HIP(hipEventCreate(&t->e1));
    PG_HIP(hipEventRecord(t->e0, st));
    if (!S->keep_film)
        PG_HIP(hipMemsetAsync(S->d_film, 0, (size_t)npix * 4 * sizeof(float), st));
    PathsDev P = S->paths;
    P.rec_depth = tr ? tr->rec_depth : 0;
    if (!S->spectral)
        P.lam = nullptr;  // k_raygen initialises spectral state only in spectral mode
    if (P.lookups)
        PG_HIP(hipMemsetAsync(P.lookups, 0, sizeof(unsigned long long), st));
    // Fused pipeline without training: the bounce loop is a graph replay per
    // batch (no host round trip per bounce).
    const bool graph = S->use_graph && !S->general && !tr && !S->spectral && S->integrator == kIntPath;
    if (S->integrator != kIntPath && (S->spectral || S->dev.guided || tr))
        return pg_fail(PG_E_ARG, "spectral and guided rendering need the Path integrator");
    if (graph) {
        if (int rc = ensure_graph(S, P, n_paths))
            return rc;
        PG_HIP(hipMemsetAsync(S->g_rays, 0, 2 * sizeof(unsigned long long), st));
    }
    int64_t rays = 0, ext_rays = 0;
    S->infer_ms = 0.0;
    S->infer_queries = 0;
    // Tracking kernels compiled for this render: Halton indices stay below
    // 2^32 when stride * (last sample index + 1) does (a pixel's offset is
    // below the stride); brick-only kernels when every medium is a grid
    // tracked with brick majorants.
    if (S->dev.sampler.kind == kSamplerSobol)
        S->track_sampler = kSamplerSobolOnly;
    else
        S->track_sampler = (uint64_t)S->dev.sampler.halton.stride * ((uint64_t)sample_offset + (uint64_t)spp) <=
                                   (1ull << 32) ? kSamplerHalton32 : kSamplerAny;
    S->track_media = kMediaBricks;
    for (const MediumDev& m : S->media)
        if (!(m.type == kMediumGrid && m.method == kDeltaTracking && m.brick > 0))
            S->track_media = kMediaAny;
    if (S->track_media == kMediaBricks && S->media.size() == 1)
        S->track_media = kMediaOneBrickGrid;
    const int B = 256;
    for (int s0 = 0; s0 < spp; s0 += spb) {
        const int nb = std::min(spb, spp - s0);
        const int32_t np = (int32_t)(npix * nb);
        const int64_t n_slots = tr ? (int64_t)np * tr->rec_depth : 0;
        if (tr)
            hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n_slots + B - 1) / B)), dim3(B), 0, st,
                               P.rec_pdf, n_slots, 0.0f);
        hipLaunchKernelGGL(k_raygen, dim3((np + B - 1) / B), dim3(B), 0, st, S->dev, P, np, nb,
                           s0 + sample_offset, graph ? S->g_counts : nullptr);
        int32_t active = graph ? 0 : np;
        if (graph)
            PG_HIP(hipGraphLaunch(S->graph_exec, st));
        PathsDev Q = P;
        // Depth bounds every path except those crossing material-less surfaces,
        // which pine does not count (path.cpp:42-46); the bound guards those.
        for (int bounce = 0; active > 0 && bounce < kMaxBounces; bounce++) {
            ext_rays += active;
            PG_HIP(hipMemsetAsync(Q.counters, 0, 8 * sizeof(int32_t), st));
            const bool viz_bvh = S->integrator == kIntViz && S->dev.viz_type == kVizBvh;
            if (viz_bvh) {
                if (S->viz_counts_capacity < active) {
                    if (int rc = dalloc(*S, 2 * (int64_t)active, &S->d_viz_counts))
                        return rc;
                    S->viz_counts_capacity = active;
                }
                if (int rc = pg_trace_closest_counts(S->accel, Q.rays_in, active, Q.hits, S->d_viz_counts, st))
                    return rc;
            } else if (int rc = pg_trace_closest(S->accel, Q.rays_in, active, Q.hits, Q.uvs, st)) {
                return rc;
            }
            if (S->integrator == kIntAO) {
                hipLaunchKernelGGL(k_shade_ao, dim3((active + B - 1) / B), dim3(B), 0, st, S->dev, Q, active);
                int32_t counts[2];
                PG_HIP(hipMemcpyAsync(counts, Q.counters, sizeof(counts), hipMemcpyDeviceToHost, st));
                PG_HIP(hipStreamSynchronize(st));
                rays += active + counts[1];
                if (counts[1] > 0) {
                    if (int rc = pg_trace_any(S->accel, Q.shadow_rays, counts[1], Q.occluded, st))
                        return rc;
                    hipLaunchKernelGGL(k_resolve_ao, dim3((counts[1] + B - 1) / B), dim3(B), 0, st, Q, counts[1]);
                }
                active = 0;
            } else if (S->integrator == kIntViz) {
                hipLaunchKernelGGL(k_shade_viz, dim3((active + B - 1) / B), dim3(B), 0, st, S->dev, Q, active,
                                   viz_bvh ? S->d_viz_counts : nullptr);
                rays += active;
                active = 0;
            } else if (S->integrator == kIntRandomWalk) {
                if (S->general) {
                    const size_t lds = S->lds_medium >= 0 ? S->med_majorant[S->lds_medium].size() * sizeof(float) : 0;
                    hipLaunchKernelGGL(k_shade_rw_gen, dim3(tracking_grid(active, B)), dim3(B), lds, st, S->dev, Q,
                                       active, S->lds_medium);
                } else {
                    hipLaunchKernelGGL(k_shade_rw, dim3((active + B - 1) / B), dim3(B), 0, st, S->dev, Q, active);
                }
                int32_t next = 0;
                PG_HIP(hipMemcpyAsync(&next, Q.counters, sizeof(next), hipMemcpyDeviceToHost, st));
                PG_HIP(hipStreamSynchronize(st));
                rays += active;
                active = next;
            } else if (S->spectral && !S->general) {
                hipLaunchKernelGGL(k_shade_spec, dim3((active + B - 1) / B), dim3(B), 0, st, S->dev, Q, active);
                int32_t counts[2];
                PG_HIP(hipMemcpyAsync(counts, Q.counters, sizeof(counts), hipMemcpyDeviceToHost, st));
                PG_HIP(hipStreamSynchronize(st));
                rays += active + counts[1];
                if (counts[1] > 0) {
                    if (int rc = pg_trace_occluded(S->accel, Q.shadow_rays, counts[1], Q.occluded, st))
                        return rc;
                    hipLaunchKernelGGL(k_resolve_shadow_spec, dim3((counts[1] + B - 1) / B), dim3(B), 0, st, Q,
                                       counts[1]);
                }
                active = counts[0];
            } else if (!S->general) {
                const bool infer = S->dev.guided && bounce < S->dev.g_depth;
                if (infer) {
                    if (!S->infer_ev[0]) {
                        PG_HIP(hipEventCreate(&S->infer_ev[0]));
                        PG_HIP(hipEventCreate(&S->infer_ev[1]));
                    }
                    hipLaunchKernelGGL(k_guide_gather, dim3((active + B - 1) / B), dim3(B), 0, st, S->dev, Q,
                                       active, nullptr);
                    PG_HIP(hipEventRecord(S->infer_ev[0], st));
                    hipLaunchKernelGGL(k_guide_infer<1>, dim3(infer_blocks(active)), dim3(kInferThreads), 0, st,
                                       S->dev, Q, active, nullptr);
                    PG_HIP(hipEventRecord(S->infer_ev[1], st));
                    hipLaunchKernelGGL(k_shade<true>, dim3((active + kShadeBlock - 1) / kShadeBlock),
                                       dim3(kShadeBlock), 0, st, S->dev, Q, active, nullptr);
                } else {
                    hipLaunchKernelGGL(k_shade<false>, dim3((active + kShadeBlock - 1) / kShadeBlock),
                                       dim3(kShadeBlock), 0, st, S->dev, Q, active, nullptr);
                }
                int32_t counts[2];
                PG_HIP(hipMemcpyAsync(counts, Q.counters, sizeof(counts), hipMemcpyDeviceToHost, st));
                PG_HIP(hipStreamSynchronize(st));
                if (infer) {
                    float ims = 0.0f;
                    PG_HIP(hipEventElapsedTime(&ims, S->infer_ev[0], S->infer_ev[1]));
                    S->infer_ms += ims;
                    S->infer_queries += active;
                }
                rays += active + counts[1];
                if (counts[1] > 0) {
                    if (int rc = pg_trace_occluded(S->accel, Q.shadow_rays, counts[1], Q.occluded, st))
                        return rc;
                    hipLaunchKernelGGL(k_resolve_shadow, dim3((counts[1] + B - 1) / B), dim3(B), 0, st,
                                       Q, counts[1], nullptr);
                }
                active = counts[0];
            } else {
                if (int rc = bounce_general(S, Q, active, st, rays))
                    return rc;
            }
            std::swap(Q.queue_in, Q.queue_out);
            std::swap(Q.rays_in, Q.rays_out);
        }
        if (tr) {
            hipLaunchKernelGGL(k_guide_finalize, dim3((unsigned)((n_slots + B - 1) / B)), dim3(B), 0, st,
                               P, n_slots);
            const int32_t nblk = (int32_t)((n_slots + 255) / 256);
            hipLaunchKernelGGL(k_rec_count, dim3(nblk), dim3(256), 0, st, P.rec_w, n_slots, P.rec_bcnt,
                               P.rec_bsum);
            hipLaunchKernelGGL(k_rec_scan, dim3(1), dim3(1024), 0, st, P.rec_bcnt, P.rec_bsum, nblk,
                               P.rec_boff, P.rec_acc, P.rec_n);
            hipLaunchKernelGGL(k_rec_scatter, dim3(nblk), dim3(256), 0, st, P, n_slots);
            int32_t n_train = 0;
            PG_HIP(hipMemcpyAsync(&n_train, P.rec_n, sizeof(n_train), hipMemcpyDeviceToHost, st));
            PG_HIP(hipStreamSynchronize(st));
            tr->last_records += n_train;
            if (n_train > 0) {
                // Minibatches walk the records in a fixed pseudo-random order;
                // by default one epoch, so every step sees fresh paths.
                uint64_t a = 2654435761ull % (uint64_t)n_train;
                if (a < 2)
                    a = 1;
                while (a > 1 && gcd64(a, (uint64_t)n_train) != 1)
                    a++;
                // Default: one epoch in at most kAutoSteps steps of at least
                // `batch` records (larger films train on larger batches).
                int64_t count = std::min<int64_t>(tr->batch, n_train), steps = tr->steps_per_batch;
                if (steps <= 0) {
                    steps = std::min<int64_t>(kAutoSteps, (n_train + count - 1) / count);
                    count = (n_train + steps - 1) / steps;
                }
                int64_t first = 0;
                hipEvent_t e0 = nullptr, e1 = nullptr;
                PG_HIP(hipEventCreate(&e0));
                PG_HIP(hipEventCreate(&e1));
                tr->step_events.emplace_back(e0, e1);
                PG_HIP(hipEventRecord(e0, st));
                for (int64_t k = 0; k < steps; k++) {
                    const int64_t take = std::min<int64_t>(count, n_train - first);
                    if (int rc = guide_train_step(tr->G, P.c_x, P.c_dir, P.c_w, P.c_pb, n_train, first, take, a, st))
                        return rc;
                    first += take;
                    if (first >= n_train)
                        first = 0;
                }
                PG_HIP(hipEventRecord(e1, st));
                tr->steps += steps;
            }
        }
        if (S->spectral)
            hipLaunchKernelGGL(k_spec_to_rgb, dim3((np + B - 1) / B), dim3(B), 0, st, S->dev, P, np);
        if (S->dev.filter_radius[0] > 0.5f || S->dev.filter_radius[1] > 0.5f) {
            hipLaunchKernelGGL(k_film_gather, dim3((npix + B - 1) / B), dim3(B), 0, st, S->dev, P, (int32_t)npix,
                               nb, S->d_film);
        } else {
            const int G = film_pixels_per_block(nb);
            hipLaunchKernelGGL(k_film, dim3((unsigned)((npix + G - 1) / G)), dim3(B), 0, st, S->dev, P,
                               (int32_t)npix, nb, G, S->d_film);
            hipLaunchKernelGGL(k_film_border, dim3((npix + B - 1) / B), dim3(B), 0, st, S->dev, P, (int32_t)npix,
                               nb, S->d_film);
        }
    }
    PG_HIP(hipGetLastError());
    PG_HIP(hipEventRecord(t->e1, st));
    t->graph = graph;
    t->rays = rays;
    t->ext_rays = ext_rays;
    return PG_OK;
}

// Waits for an enqueued render; its time on its stream and rays traced.
static int render_finish(pg_scene* S, RenderTicket* t, double* ms_out, int64_t* rays_out) {
    PG_HIP(hipEventSynchronize(t->e1));
    float ms = 0.0f;
    PG_HIP(hipEventElapsedTime(&ms, t->e0, t->e1));
    t->release();
    int64_t rays = t->rays;
    S->last_ext_rays = t->ext_rays;
    if (t->graph) {
        unsigned long long n[2] = {0, 0};
        PG_HIP(hipMemcpy(n, S->g_rays, sizeof(n), hipMemcpyDeviceToHost));
        rays = (int64_t)n[0];
        S->last_ext_rays = (int64_t)n[1];
    }
    *ms_out = ms;
    *rays_out = rays;
    S->last_lookups = 0;
    if (S->d_lookups) {
        unsigned long long n = 0;
        PG_HIP(hipMemcpy(&n, S->d_lookups, sizeof(n), hipMemcpyDeviceToHost));
        S->last_lookups = (int64_t)n;
    }
    return PG_OK;
}

// Default cap on a scene's batch in pg_render_batch (its streams' path-state
// slots are live together): 64M paths, about 15 GB per slot for the fused
// bounce.
constexpr int64_t kBatchStreamPaths = 64 << 20;

// Holds the device's render lock for one entry point's renders.
struct RenderLock {
    DevicePaths& D;
    std::lock_guard<std::mutex> lk;
    explicit RenderLock(const pg_scene* S) : D(device_paths(S->device)), lk(D.render) {}
    PathState& slot(int k) { return path_slot(D, k); }
};

static int render_impl(pg_scene* S, PathState& W, int spp, int spb_req, hipStream_t st, GuideTraining* tr,
                       double* ms_out, int64_t* rays_out, int sample_offset = 0) {
    RenderTicket t;
    if (int rc = render_enqueue(S, W, spp, spb_req, st, tr, sample_offset, &t)) {
        (void)hipStreamSynchronize(st);
        t.release();
        return rc;
    }
    return render_finish(S, &t, ms_out, rays_out);
}

extern "C" int pg_render(pg_scene* S, const pg_render_options* opts, void* stream) {
    if (!S)
        return pg_fail(PG_E_ARG, "pg_render: null scene");
    if (opts && opts->rr_depth < 0)
        return pg_fail(PG_E_ARG, "pg_render: rr_depth must be >= 0");
    const int spp = opts && opts->spp > 0 ? opts->spp : S->spp;
    const int spb = opts && opts->samples_per_batch > 0 ? opts->samples_per_batch : 0;
    const int offset = opts && opts->sample_offset > 0 ? opts->sample_offset : 0;
    S->dev.guided = 0;
    S->dev.rr_depth = opts && opts->rr_depth > 0 ? opts->rr_depth : 2;
    S->use_graph = !(opts && (opts->flags & PG_RENDER_NO_GRAPH));
    S->spectral = opts && (opts->flags & PG_RENDER_SPECTRAL);
    S->lane_tracking = opts && (opts->flags & PG_RENDER_LANE_TRACKING);
    double ms = 0.0;
    int64_t rays = 0;
    RenderLock lock(S);
    const int rc = render_impl(S, lock.slot(0), spp, spb, (hipStream_t)stream, nullptr, &ms, &rays, offset);
    S->use_graph = true;
    S->spectral = false;
    S->lane_tracking = false;
    if (rc)
        return rc;
    S->last_render_ms = ms;
    S->last_rays = rays;
    S->last_guide_ms = 0.0;
    return PG_OK;
}

extern "C" int pg_spectral_tables(float* Q, float* M, float* w_scale) {
    if (!Q || !M || !w_scale)
        return pg_fail(PG_E_ARG, "pg_spectral_tables: null argument");
    spec::Tables t;
    spec::spectral_tables(t);
    for (int k = 0; k < 9; k++) {
        Q[k] = t.Q[k];
        M[k] = t.M[k];
    }
    *w_scale = t.w_scale;
    return PG_OK;
}

extern "C" int pg_render_batch(pg_scene* const* scenes, int32_t n, const int32_t* spp,
                               const pg_batch_options* o, void* stream, pg_batch_result* out) {
    if (n < 0 || (n > 0 && !scenes))
        return pg_fail(PG_E_ARG, "pg_render_batch: bad scene list");
    for (int32_t i = 0; i < n; i++)
        if (!scenes[i])
            return pg_fail(PG_E_ARG, "pg_render_batch: null scene");
        else if (scenes[i]->device != scenes[0]->device)
            return pg_fail(PG_E_ARG, "pg_render_batch: scenes loaded on different devices");
    if (n == 0) {
        if (out)
            *out = pg_batch_result{};
        return PG_OK;
    }
    // Stream k renders its scenes one after another on path-state slot k + 1.
    RenderLock lock(scenes[0]);
    const int K = o && o->streams > 0 ? std::min(o->streams, 16) : 4;
    const int flags = o ? o->flags : 0;
    if (o && o->paths_per_batch < 0)
        return pg_fail(PG_E_ARG, "pg_render_batch: paths_per_batch must be >= 0");
    // Each stream's path-state slot is live at once: by default a scene's
    // batch is at most kBatchStreamPaths paths.
    const int64_t cap = o && o->paths_per_batch > 0 ? o->paths_per_batch : kBatchStreamPaths;
    hipStream_t st = (hipStream_t)stream;
    std::vector<hipStream_t> ws(K, nullptr);
    std::vector<hipEvent_t> done(K, nullptr);
    hipEvent_t start = nullptr, stop = nullptr;
    std::vector<RenderTicket> tickets(n);
    auto cleanup = [&]() {
        for (auto& t : tickets)
            t.release();
        for (int k = 0; k < K; k++) {
            if (ws[k])
                (void)hipStreamDestroy(ws[k]);
            if (done[k])
                (void)hipEventDestroy(done[k]);
        }
        if (start)
            (void)hipEventDestroy(start);
        if (stop)
            (void)hipEventDestroy(stop);
    };
    int rc = PG_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == PG_OK)
            rc = pg_fail(PG_E_HIP, std::string("pg_render_batch: ") + what + ": " + hipGetErrorString(e));
        return rc == PG_OK;
    };
    hip(hipEventCreate(&start), "event") && hip(hipEventCreate(&stop), "event");
    for (int k = 0; k < K && rc == PG_OK; k++)
        hip(hipStreamCreateWithFlags(&ws[k], hipStreamNonBlocking), "stream") && hip(hipEventCreate(&done[k]), "event");
    if (rc == PG_OK)
        hip(hipEventRecord(start, st), "record");
    for (int k = 0; k < K && rc == PG_OK; k++)
        hip(hipStreamWaitEvent(ws[k], start, 0), "wait");
    // Longest first onto the least loaded stream (cost ~ camera samples).
    std::vector<int32_t> order(n);
    std::vector<int64_t> cost(n);
    for (int32_t i = 0; i < n; i++) {
        order[i] = i;
        const int sp = spp && spp[i] > 0 ? spp[i] : scenes[i]->spp;
        cost[i] = (int64_t)scenes[i]->dev.width * scenes[i]->dev.height * sp;
    }
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return cost[a] > cost[b]; });
    std::vector<int64_t> load(K, 0);
    for (int32_t j = 0; j < n && rc == PG_OK; j++) {
        const int32_t i = order[j];
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[k] += cost[i];
        pg_scene* S = scenes[i];
        const int sp = spp && spp[i] > 0 ? spp[i] : S->spp;
        S->dev.guided = 0;
        S->dev.rr_depth = 2;
        S->use_graph = !(flags & PG_RENDER_NO_GRAPH);
        S->lane_tracking = (flags & PG_RENDER_LANE_TRACKING) != 0;
        const int64_t own = S->batch_paths;
        S->batch_paths = std::min(own, cap);
        rc = render_enqueue(S, lock.slot(k + 1), sp, 0, ws[k], nullptr, 0, &tickets[i]);
        S->batch_paths = own;
        S->use_graph = true;
        S->lane_tracking = false;
    }
    for (int k = 0; k < K && rc == PG_OK; k++)
        hip(hipEventRecord(done[k], ws[k]), "record") && hip(hipStreamWaitEvent(st, done[k], 0), "wait");
    if (rc == PG_OK)
        hip(hipEventRecord(stop, st), "record");
    if (rc != PG_OK) {
        for (int k = 0; k < K; k++)
            if (ws[k])
                (void)hipStreamSynchronize(ws[k]);
        cleanup();
        return rc;
    }
    pg_batch_result r{};
    for (int32_t i = 0; i < n && rc == PG_OK; i++) {
        double ms = 0.0;
        int64_t rays = 0;
        rc = render_finish(scenes[i], &tickets[i], &ms, &rays);
        scenes[i]->last_render_ms = ms;
        scenes[i]->last_rays = rays;
        scenes[i]->last_guide_ms = 0.0;
        const int sp = spp && spp[i] > 0 ? spp[i] : scenes[i]->spp;
        r.samples += (int64_t)scenes[i]->dev.width * scenes[i]->dev.height * sp;
        r.rays += rays;
    }
    if (rc == PG_OK && hip(hipEventSynchronize(stop), "sync")) {
        float ms = 0.0f;
        if (hip(hipEventElapsedTime(&ms, start, stop), "elapsed"))
            r.ms = ms;
    }
    cleanup();
    if (rc == PG_OK && out)
        *out = r;
    return rc;
}

// One probe render of the scene as configured; its time on the stream.
static int probe(pg_scene* S, PathState& W, int spp, hipStream_t st, double* ms) {
    S->dev.guided = 0;
    S->dev.rr_depth = 2;
    int64_t rays = 0;
    return render_impl(S, W, spp, 0, st, nullptr, ms, &rays, 0);
}

extern "C" int pg_scene_autotune(pg_scene* S, const pg_tune_options* o, void* stream, pg_tune_result* out) {
    if (!S)
        return pg_fail(PG_E_ARG, "pg_scene_autotune: null scene");
    hipStream_t st = (hipStream_t)stream;
    const int pspp = std::max(1, std::min(o && o->probe_spp > 0 ? o->probe_spp : 1, S->spp));
    pg_tune_result r{};
    int rc;
    double ms = 0.0, best = 0.0;
    RenderLock lock(S);
    PathState& W = lock.slot(0);
    auto measure = [&](int spp, double* t) -> int {
        double a = 0.0, b = 0.0;
        if (int e = probe(S, W, spp, st, &a))  // first run records the graph
            return e;
        if (int e = probe(S, W, spp, st, &b))
            return e;
        r.probes += 2;
        *t = std::min(a, b);
        return PG_OK;
    };
    pg_accel_stats as{};
    if ((rc = pg_accel_get_stats(S->accel, &as)))
        return rc;
    if ((rc = measure(pspp, &best)))
        return rc;
    r.default_ms = best;
    // Node width: rebuild with the other width, keep it if faster.
    if (o && o->try_width) {
        pg_build_options bo = S->build_opts;
        const int other = as.width == 2 ? 4 : 2;
        bo.width = other;
        if ((rc = pg_accel_rebuild(S->accel, &bo)))
            return rc;
        pg_accel_set_launch(S->accel, as.block, as.launch_flags);
        if ((rc = measure(pspp, &ms)))
            return rc;
        if (ms < best) {
            best = ms;
            S->build_opts = bo;
        } else {
            bo.width = as.width;
            if ((rc = pg_accel_rebuild(S->accel, &bo)))
                return rc;
            pg_accel_set_launch(S->accel, as.block, as.launch_flags);
        }
    }
    // Traversal workgroup size x leaf postponing.
    int best_block = as.block, best_flags = as.launch_flags;
    for (int block : {64, 128, 256})
        for (int flags : {0, (int)PG_LAUNCH_POSTPONE_LEAVES}) {
            if (block == best_block && flags == best_flags)
                continue;
            if ((rc = pg_accel_set_launch(S->accel, block, flags)) || (rc = measure(pspp, &ms)))
                return rc;
            if (ms < best) {
                best = ms;
                best_block = block;
                best_flags = flags;
            }
        }
    pg_accel_set_launch(S->accel, best_block, best_flags);
    // Paths in flight: probes large enough to fill the largest candidate.
    const int64_t npix = (int64_t)S->dev.width * S->dev.height;
    const int64_t cands[4] = {8 << 20, 32 << 20, 64 << 20, 256 << 20};
    const int qspp = (int)std::min<int64_t>(S->spp, std::max<int64_t>(pspp, (cands[3] + npix - 1) / npix));
    if (qspp * npix > cands[0]) {
        const int64_t start = S->batch_paths;
        double qbest = 0.0;
        if ((rc = measure(qspp, &qbest)))
            return rc;
        int64_t qpaths = start;
        for (int64_t c : cands) {
            if (c == start)
                continue;
            S->batch_paths = c;
            if ((rc = measure(qspp, &ms)))
                return rc;
            if (ms < qbest) {
                qbest = ms;
                qpaths = c;
            }
        }
        S->batch_paths = qpaths;
    }
    if ((rc = pg_accel_get_stats(S->accel, &as)))
        return rc;
    r.trace_block = as.block;
    r.trace_flags = as.launch_flags;
    r.bvh_width = as.width;
    r.paths_in_flight = S->batch_paths;
    r.best_ms = best;
    if (out)
        *out = r;
    return PG_OK;
}

extern "C" int pg_render_guided(pg_scene* S, pg_guide* G, const pg_guided_options* o,
                                void* stream) {
    if (!S || !G)
        return pg_fail(PG_E_ARG, "pg_render_guided: null scene or guide");
    if (o && o->rr_depth < 0)
        return pg_fail(PG_E_ARG, "pg_render_guided: rr_depth must be >= 0");
    const int passes = o && o->train_passes > 0 ? o->train_passes : 16;
    const int train_spp = o && o->train_spp > 0 ? o->train_spp : 8;
    const int steps = o && o->steps_per_pass > 0 ? o->steps_per_pass : 0;
    const int batch = o && o->batch > 0 ? o->batch : 1024;
    const int depth = o && o->record_depth > 0 ? o->record_depth : 5;
    const int final_spp = o && o->final_spp > 0 ? o->final_spp : S->spp;
    const int offset = o && o->sample_offset > 0 ? o->sample_offset : 0;
    const int keep = o && o->keep_passes > 0 ? std::min(o->keep_passes, passes) : 0;
    RenderLock lock(S);
    PathState& W = lock.slot(0);
    SceneDev& d = S->dev;
    d.g_depth = o && o->guide_depth > 0 ? o->guide_depth : (1 << 30);
    d.rr_depth = o && o->rr_depth > 0 ? o->rr_depth : 2;
    d.g_params = guide_params_dev(G);
    d.g_wf = guide_wf_dev(G);
    d.g_oct = guide_freq_offset(G);
    d.g_bmin = mk(S->bmin[0], S->bmin[1], S->bmin[2]);
    d.g_binv = mk(1.0f / std::max(S->bmax[0] - S->bmin[0], 1e-6f),
                  1.0f / std::max(S->bmax[1] - S->bmin[1], 1e-6f),
                  1.0f / std::max(S->bmax[2] - S->bmin[2], 1e-6f));
    d.guided = 1;
    hipStream_t st = (hipStream_t)stream;
    double train_ms = 0.0, ms = 0.0;
    int64_t rays = 0, all_rays = 0;
    const int64_t npix = (int64_t)d.width * d.height;
    const int64_t per_batch = std::max<int64_t>(1, std::min<int64_t>(train_spp, (4 << 20) / std::max<int64_t>(npix, 1)));
    const int nbatches = (int)((train_spp + per_batch - 1) / per_batch);
    GuideTraining tr{G, depth, steps > 0 ? std::max(1, (steps + nbatches - 1) / nbatches) : 0, batch, 0, 0, {}};
    int rc = PG_OK;
    // The last `keep` training passes stay in the film: each is an unbiased
    // guided render of its own sample range, and with the final render they
    // cover one contiguous range of sample indices.
    for (int p = 0; p < passes && rc == PG_OK; p++) {
        S->keep_film = p > passes - keep;
        rc = render_impl(S, W, train_spp, (int)per_batch, st, &tr, &ms, &rays, offset + p * train_spp);
        train_ms += ms;
        all_rays += rays;
    }
    if (rc == PG_OK) {
        S->use_graph = !(o && (o->flags & PG_RENDER_NO_GRAPH));
        S->keep_film = keep > 0;
        rc = render_impl(S, W, final_spp, 0, st, nullptr, &ms, &rays, offset + passes * train_spp);
        S->use_graph = true;
    }
    S->keep_film = false;
    d.guided = 0;
    if (rc != PG_OK)
        return rc;
    S->last_render_ms = ms;
    S->last_rays = rays;
    S->last_guide_ms = train_ms + ms;
    S->guide_steps = tr.steps;
    S->guide_records = tr.last_records;
    S->guide_step_ms = tr.step_ms();  // every pass has finished (render_impl waited)
    return PG_OK;
}

extern "C" int pg_scene_film(const pg_scene* S, float* h_rgbw) {
    if (!S || !h_rgbw)
        return pg_fail(PG_E_ARG, "pg_scene_film: null argument");
    PG_HIP(hipMemcpy(h_rgbw, S->d_film, (size_t)S->dev.width * S->dev.height * 4 * sizeof(float),
                     hipMemcpyDeviceToHost));
    return PG_OK;
}

// One thread per (pixel, sample): StartPixel(p, sample) and ndims Get1D values
// in the harness's pixel spread (pixel k is (py0 * w + px0 + 977 k) mod film).
__global__ void k_sampler_values(SamplerDev sd, int w, int h, int px0, int py0, int npix, int spp,
                                 int ndims, float* out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)npix * spp)
        return;
    const int k = (int)(t / spp), si = (int)(t - (int64_t)k * spp);
    const int64_t idx = (int64_t)py0 * w + px0 + (int64_t)k * 977;
    const int px = (int)(idx % w), py = (int)((idx / w) % h);
    SamplerState ss;
    ss.index = sampler_start(sd, px, py, si);
    ss.dim = 2;
    float* o = out + t * ndims;
    for (int d = 0; d < ndims; d++)
        o[d] = get_1d(sd, ss);
}

extern "C" int pg_scene_sampler_values(const pg_scene* S, int32_t px0, int32_t py0, int32_t npix,
                                       int32_t spp, int32_t ndims, float* h_out) {
    if (!S || !h_out || px0 < 0 || py0 < 0 || npix <= 0 || spp <= 0 || ndims <= 0 ||
        (int64_t)npix * spp * ndims > (1ll << 28))
        return pg_fail(PG_E_ARG, "pg_scene_sampler_values: bad argument");
    const int64_t n = (int64_t)npix * spp, bytes = n * ndims * (int64_t)sizeof(float);
    float* d = nullptr;
    PG_HIP(hipMalloc(&d, bytes));
    hipLaunchKernelGGL(k_sampler_values, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, S->dev.sampler,
                       S->dev.width, S->dev.height, px0, py0, npix, spp, ndims, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpy(h_out, d, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess)
        return pg_fail(PG_E_HIP, std::string("pg_scene_sampler_values: ") + hipGetErrorString(e));
    return PG_OK;
}

extern "C" int pg_scene_output_name(const pg_scene* S, char* buf, int32_t len) {
    if (!S || !buf || len <= 0)
        return pg_fail(PG_E_ARG, "pg_scene_output_name: bad argument");
    snprintf(buf, (size_t)len, "%s", S->output.c_str());
    return PG_OK;
}

extern "C" int pg_scene_release(pg_scene* S) {
    delete S;
    return PG_OK;
}
