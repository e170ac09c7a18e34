/* pinehip.h - C ABI of the MI355X (gfx950) path-tracing hot path that sits
 * behind pine's Accel and Integrator plug-in points.
 *
 * Pine itself has no C ABI: its boundary is the C++ virtual interface
 *   class Accel { Initialize(const Scene*); bool Hit(Ray) const;
 *                 bool Intersect(Ray&, Interaction&) const; }   (src/core/accel.h:8-14)
 * called one ray at a time.  A per-ray virtual call is the wrong granularity for
 * a GPU, so this ABI takes whole ray batches.  Each entry point names the pine
 * function whose semantics it reproduces.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative PG_E* code; the message
 *     for the last failure on the calling thread is pg_last_error().
 *   - Pointers named d_* are device (HBM) pointers; h_* are host pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).  Work is
 *     stream-ordered: results are ready when the stream is synchronised.
 *   - An accel handle may be traced from several threads at once; building and
 *     releasing it must not overlap any other call on the same handle.
 */
#ifndef PINEHIP_H
#define PINEHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_OK 0
#define PG_E_ARG (-1)     /* invalid argument (null pointer, bad size, bad handle) */
#define PG_E_STATE (-2)   /* call out of order (trace before build, add after build) */
#define PG_E_HIP (-3)     /* a HIP runtime call failed */
#define PG_E_NOMEM (-4)   /* host or device allocation failed */
#define PG_E_IO (-5)      /* file could not be read or written */
#define PG_E_PARSE (-6)   /* scene text could not be parsed */

/* One ray, 32 bytes: pine's Ray{o, d, tmin, tmax} (src/core/ray.h:10-32) with
 * tmin moved next to o so the record is two 16-byte loads. */
typedef struct pg_ray {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} pg_ray;

/* Closest-hit result, 8 bytes.  prim is the global primitive id (see
 * pg_accel_add_*), -1 on a miss; t is the hit distance pine writes into
 * ray.tmax (bvh.cpp:534-610 + geometry.h:271-294). */
typedef struct pg_hit {
    int32_t prim;
    float t;
} pg_hit;

/* Barycentric / surface parameters of a closest hit (pine's Interaction::uv):
 * triangles (u,v) as in Triangle::Intersect, spheres (phi,theta), rects (u,v). */
typedef struct pg_hit_uv {
    float u;
    float v;
} pg_hit_uv;

typedef struct pg_accel pg_accel;

const char* pg_last_error(void);
const char* pg_version(void);

/* --- Accel (replaces pine::BVH, src/impl/accel/bvh.h:72-85) ---------------- */

/* Creates an empty accel (host only; the device is the one current at build). */
int pg_accel_create(pg_accel** out);

/* Adds a triangle mesh: pine's TriangleMesh{vertices, indices} (geometry.h:364-413).
 * Triangle k of the mesh gets global prim id *first_prim_out + k.  Copies the data. */
int pg_accel_add_mesh(pg_accel* a, const float* h_vertices, int64_t n_vertices,
                      const uint32_t* h_indices, int64_t n_triangles, int32_t* first_prim_out);

/* Analytic shapes: pine's Sphere{c, r} (geometry.h:143-168), Rect{position, ex, ey}
 * (geometry.h:334-362, ex/ey unnormalised as in the scene file), Plane{position,
 * normal} (geometry.h:118-141).  Each is one prim id. */
int pg_accel_add_sphere(pg_accel* a, const float c[3], float r, int32_t* prim_out);
int pg_accel_add_rect(pg_accel* a, const float position[3], const float ex[3],
                      const float ey[3], int32_t* prim_out);
int pg_accel_add_plane(pg_accel* a, const float position[3], const float normal[3],
                       int32_t* prim_out);

/* Build options.  width: 2 (binary) or 4 (4-wide nodes); leaf_max: 1..8 prims;
 * bins: SAH bins per axis; threads: host build threads (0 = all cores). */
typedef struct pg_build_options {
    int32_t width;
    int32_t leaf_max;
    int32_t bins;
    int32_t threads;
} pg_build_options;

/* Builds the hierarchy on the host and uploads it to HBM.  Replaces
 * BVH::Initialize (bvh.cpp:612-650).  opts may be NULL for defaults. */
int pg_accel_build(pg_accel* a, const pg_build_options* opts);

/* Builds the hierarchy again (other options, e.g. node width) from the same
 * primitives; prim ids do not change.  No pine counterpart. */
int pg_accel_rebuild(pg_accel* a, const pg_build_options* opts);

/* Closest hit for n rays: replaces BVH::Intersect (bvh.cpp:669-709).  Hit rule:
 * the smallest t in (tmin, tmax) over all prims, using pine's per-shape tests;
 * on an exact tie in t the smaller prim id wins (pine keeps the first found in
 * its own traversal order).  d_uv may be NULL. */
int pg_trace_closest(pg_accel* a, const pg_ray* d_rays, int64_t n, pg_hit* d_hits,
                     pg_hit_uv* d_uv, void* stream);

/* Any hit for n rays: replaces BVH::Hit (bvh.cpp:652-667) with pine's Hit-variant
 * tests (closed t interval for triangles, strict u+v<1).  d_occluded[i] = 0/1. */
int pg_trace_any(pg_accel* a, const pg_ray* d_rays, int64_t n, int32_t* d_occluded,
                 void* stream);

/* Shadow-ray visibility as pine's IntersectTr decides it for surfaces that all
 * carry a material (integrator.cpp:68-87): occluded when any prim's Intersect
 * test accepts a t in (tmin, tmax).  d_occluded[i] = 0/1. */
int pg_trace_occluded(pg_accel* a, const pg_ray* d_rays, int64_t n, int32_t* d_occluded,
                      void* stream);

/* pg_trace_closest plus, per ray, the work done: d_counts[2i] = nodes fetched,
 * d_counts[2i+1] = primitives tested.  For measurement (algorithmic bytes);
 * no pine counterpart beyond the per-node counter of bvh.cpp:553. */
int pg_trace_closest_counts(pg_accel* a, const pg_ray* d_rays, int64_t n, pg_hit* d_hits,
                            int32_t* d_counts, void* stream);

/* Statistics of the uploaded hierarchy (for reporting and roofline accounting). */
/* pg_trace_closest / pg_trace_occluded for a ray count held in device memory
 * (*d_n <= capacity, read when the kernel runs): the launch is sized for
 * capacity rays and does not depend on the count, so it can be recorded in a
 * HIP graph whose earlier nodes produce the rays.  No pine counterpart (pine
 * traces one ray per call, bvh.cpp:652-709). */
int pg_trace_closest_dev(pg_accel* a, const pg_ray* d_rays, const int32_t* d_n, int64_t capacity,
                         pg_hit* d_hits, pg_hit_uv* d_uv, void* stream);
int pg_trace_occluded_dev(pg_accel* a, const pg_ray* d_rays, const int32_t* d_n, int64_t capacity,
                          int32_t* d_occluded, void* stream);

typedef struct pg_accel_stats {
    int64_t n_prims;
    int64_t n_nodes;
    int64_t node_bytes;
    int64_t prim_bytes;
    int32_t width;
    int32_t max_depth;
    double build_ms;
    int32_t block;         /* traversal workgroup size (pg_accel_set_launch) */
    int32_t launch_flags;  /* PG_LAUNCH_* bits */
    double last_trace_ms;  /* the traversal kernel alone (not the reordering sort) of the last
                              launch made with PG_LAUNCH_TIMED, from HIP events recorded around
                              that kernel on its stream (get_stats waits for it); -1 if none */
} pg_accel_stats;
int pg_accel_get_stats(const pg_accel* a, pg_accel_stats* out);

/* Maps a global prim id back to (shape index in add order, prim within shape). */
int pg_accel_prim_shape(const pg_accel* a, int32_t prim, int32_t* shape, int32_t* local);

/* Kernel launch shape used by the trace calls (no counterpart in pine, whose
 * BVH::Intersect is one ray on one thread):
 *   block: threads per workgroup, 64, 128 or 256; 0 keeps the current value;
 *   flags: PG_LAUNCH_* bits, -1 keeps the current value.
 * Every setting returns identical hits; only the speed differs. */
#define PG_LAUNCH_POSTPONE_LEAVES 1 /* wave tests its leaves together (while-while loop) */
#define PG_LAUNCH_XCD_REMAP 2       /* each XCD traces one contiguous range of rays */
#define PG_LAUNCH_STACKLESS 4       /* binary trees of at most 31 levels: no stack; a per-ray trail of
                                       pending far children and a table of child codes by path key
                                       (constant-time backtracking); in the top levels a node shared
                                       by the whole wave is fetched once by scalar loads (ignores
                                       POSTPONE_LEAVES) */
#define PG_LAUNCH_REORDER 8         /* rays traced in the order of a key (direction octant, Morton
                                       code of the origin in the hierarchy's bounds) found by a
                                       counting sort on the device before each launch of 65,536
                                       rays or more; hits land in the callers' slots, identical to
                                       an unordered launch.  The sort's buffers belong to the accel
                                       (allocated at the first such launch, grown on demand; the
                                       _dev entry points reorder only when they already hold the
                                       capacity, and never while their stream is being captured
                                       into a graph: the sort clears memory with memset nodes), so
                                       concurrent launches on one accel from several streams must
                                       not use this flag */
#define PG_LAUNCH_TIMED 16          /* records HIP events around the traversal kernel of each
                                       launch (pg_accel_stats.last_trace_ms) */
#define PG_LAUNCH_DEFAULT 0
int pg_accel_set_launch(pg_accel* a, int32_t block, int32_t flags);

int pg_accel_release(pg_accel* a);

/* --- Scene + Path integrator (replaces LoadScene and PathIntegrator) --------- */

typedef struct pg_scene pg_scene;

/* Loads a pine scene file: the scene-file grammar of parser.cpp:45-110 with
 * pine's number parsing, the object order of LoadScene (fileio.cpp:251-281) and
 * the Create* factories, then builds the accel (opts may be NULL).  Supported
 * on the GPU: integrators Path, AO, Viz (Position, Normal) and RandomWalk
 * (without media) (CreateIntegrator, integrator.cpp:20-37) with the Halton
 * or Sobol sampler (CreateSampler, sampler.cpp:280-294), ThinLen camera, LanczosSinc/Box filters, Layered
 * (Diffuse/Conductor/Dielectric) and Emissive materials with constant inputs,
 * Sphere/Rect/Plane/Triangle/TriangleMesh shapes, Point, Directional, area and
 * Sky lights, Grid/Homogeneous media.  Anything else fails with PG_E_PARSE
 * naming it. */
int pg_scene_load(const char* path, const pg_build_options* opts, pg_scene** out);
/* Same from scene text; `dir` resolves relative file names (may be NULL). */
int pg_scene_load_text(const char* text, const char* dir, const pg_build_options* opts,
                       pg_scene** out);

typedef struct pg_scene_info {
    int32_t width, height;   /* film size */
    int32_t spp, max_depth;  /* sampler samplesPerPixel, integrator maxDepth */
    int32_t n_shapes, n_lights;
    int64_t n_prims;
    double last_render_ms;   /* wall time of the last pg_render on its stream */
    int64_t last_rays;       /* rays traced by the last pg_render (extension + shadow) */
    int32_t n_media;         /* Medium blocks */
    int32_t general_mode;    /* 1: media or material-less shapes (split bounce) */
    int64_t medium_steps;    /* density lookups (8 voxel taps each) of the last render */
    double last_guided_ms;   /* last pg_render_guided: training passes + final render */
    int64_t guide_steps;     /* last pg_render_guided: Adam steps taken */
    int64_t guide_records;   /* last pg_render_guided: nonzero-weight records trained on */
    int64_t graph_captures;  /* bounce loops recorded as HIP graphs so far */
    int64_t paths_in_flight; /* paths per batch when samples_per_batch is automatic */
    int32_t trace_block, trace_flags, bvh_width;  /* traversal launch and node width */
    int64_t last_extension_rays;  /* path vertices traced by the last render (closest-hit rays) */
    int32_t apply_tone_mapping;   /* the film's applyToneMapping (film.cpp:11), default 1 */
    double guide_infer_ms;        /* last guided render drawn launch by launch (PG_RENDER_NO_GRAPH):
                                     time of its guide-inference kernels on the GPU */
    int64_t guide_infer_queries;  /* ... and the network evaluations they made */
    double guide_step_ms;         /* last pg_render_guided: GPU time of its Adam steps (the
                                     training passes' network updates, not their rendering) */
} pg_scene_info;
int pg_scene_get_info(const pg_scene* s, pg_scene_info* info);

typedef struct pg_render_options {
    int32_t spp;                /* 0 = the scene's samplesPerPixel */
    int32_t samples_per_batch;  /* 0 = automatic (the scene's paths in flight, 256M unless tuned,
                                   at most half the free device memory) */
    int32_t sample_offset;      /* first sample index per pixel (progressive rendering) */
    int32_t flags;              /* PG_RENDER_* bits */
    int32_t rr_depth;           /* Russian roulette (path.cpp:76-82) only after this depth; 0 =
                                   pine's 2.  Other depths keep the estimator unbiased but are no
                                   longer pine's sample sequence */
} pg_render_options;
/* Run the fused pipeline's bounce loop launch by launch, counts read back per
 * bounce, instead of replaying it as a recorded HIP graph (the default; both
 * give the same image bit for bit). */
#define PG_RENDER_NO_GRAPH 1
/* Spectral transport (an extension; pine renders RGB): four wavelengths per
 * path by hero-wavelength sampling, the scene's RGB reflectances, emission and
 * sky upsampled to smooth spectra, radiance projected to RGB through the CIE
 * 1931 matching functions.  Gray scenes render as in RGB in expectation.  Both
 * pipelines (the split bounce of media and material-less shapes too); not
 * combined with the learned guide. */
#define PG_RENDER_SPECTRAL 2
/* Track brick-majorant media as persistent lanes (each lane takes the next ray
 * of its workgroup's share of the queue when its own is done) instead of one
 * ray per thread.  Same image bit for bit; measured 8% slower on the 256^3
 * cloud, so it is not the default. */
#define PG_RENDER_LANE_TRACKING 4

/* The spectral mode's tables: Q (RGB -> spectral basis coefficients), M (XYZ ->
 * RGB), both row-major 3x3, and the film weight scale 470 / (4 * integral of
 * ybar).  For tests; no pine counterpart. */
int pg_spectral_tables(float* Q, float* M, float* w_scale);

/* Renders the whole film: PixelIntegrator::Render + RadianceIntegrator::Compute
 * + PathIntegrator::Li (integrator.cpp:121-158, path.cpp:6-87) for every pixel
 * and sample, pine's Halton or Sobol sample sequence per pixel.  Synchronous. */
int pg_render(pg_scene* s, const pg_render_options* opts, void* stream);

/* Renders n scenes (each its own film; spp[i], or the scene's samplesPerPixel
 * when spp is NULL) as one batch: the scenes are assigned, longest first, to
 * `streams` concurrent streams (default 4) and enqueued without host round
 * trips (fused scenes replay their recorded bounce loops), so small and large
 * scenes share the GPU.  `stream` waits for the whole batch.  Every scene's
 * info then holds its own time on its stream and its rays; out gets the batch
 * wall time on `stream`, camera samples and rays.  A batch-level counterpart
 * of PixelIntegrator::Render (integrator.cpp:121-150) over many scenes; pine
 * renders one scene per process.  Synchronous. */
typedef struct pg_batch_options {
    int32_t streams;          /* concurrent streams, 1..16 (0 = 4) */
    int32_t flags;            /* PG_RENDER_* bits for every scene */
    int64_t paths_per_batch;  /* at most this many paths in one batch of any scene
                                 (0 = each scene's own paths_in_flight, at most 64M:
                                 every stream's path state is live at once) */
} pg_batch_options;
typedef struct pg_batch_result {
    double ms;        /* wall time of the batch on `stream` */
    int64_t samples;  /* camera samples: sum of width * height * spp */
    int64_t rays;     /* rays traced (extension + shadow) */
} pg_batch_result;
int pg_render_batch(pg_scene* const* scenes, int32_t n, const int32_t* spp, const pg_batch_options* o,
                    void* stream, pg_batch_result* out);

/* Chooses the scene's launch configuration by timing probe renders of the
 * scene itself: traversal workgroup size (64 / 128 / 256, i.e. waves per CU)
 * and leaf postponing, paths in flight per batch (queue size: 1M..8M), and,
 * with try_width, the BVH node width (2 / 4; the hierarchy is rebuilt).  The
 * best configuration stays set on the scene.  No pine counterpart. */
typedef struct pg_tune_options {
    int32_t probe_spp;  /* samples per pixel of a launch-configuration probe (0 = 1) */
    int32_t try_width;  /* 1: also try the other node width */
} pg_tune_options;
typedef struct pg_tune_result {
    int32_t trace_block, trace_flags, bvh_width;
    int32_t probes;           /* probe renders made */
    int64_t paths_in_flight;
    double default_ms;        /* probe time of the starting configuration */
    double best_ms;           /* probe time of the chosen configuration */
} pg_tune_result;
int pg_scene_autotune(pg_scene* s, const pg_tune_options* o, void* stream, pg_tune_result* out);

/* Film accumulators of the last render, h*w*4 floats (sum w*r, sum w*g, sum w*b,
 * sum w) in pine's pixel order (film.h:60-62: row 0 is the top of the image).
 * Linear radiance = rgb / w, as Film::CopyToRGBArray divides (film.cpp:72-89). */
int pg_scene_film(const pg_scene* s, float* h_rgbw);

/* The scene's sampler evaluated on the device, for parity tests: for npix
 * pixels spread as (py0 * width + px0 + 977 k) (wrapping over the film),
 * samples 0..spp-1 each start at the pixel and draw ndims values.  Replaces
 * Sampler::StartPixel / StartNextSample / Get1D (sampler.h:250-300) of pine's
 * HaltonSampler (sampler.cpp:32-96) and SobolSampler (sampler.cpp:129-220).
 * h_out: npix * spp * ndims floats, [pixel][sample][dimension].  Synchronous. */
int pg_scene_sampler_values(const pg_scene* s, int32_t px0, int32_t py0, int32_t npix, int32_t spp,
                            int32_t ndims, float* h_out);

/* The film's outputFileName (film.cpp:9), NUL-terminated, truncated to len. */
int pg_scene_output_name(const pg_scene* s, char* buf, int32_t len);

int pg_scene_release(pg_scene* s);

/* --- Deterministic benchmark workloads (no pine counterpart) ---------------- */

/* Geodesic icosphere of the given frequency (20*freq^2 triangles), displaced
 * radially by seeded 5-octave value-noise fBm of amplitude `amp`, truncated to
 * max_triangles (0 = all).  Call with h_vertices/h_indices NULL to get sizes. */
int pg_gen_icosphere(int32_t freq, float amp, uint64_t seed, int64_t max_triangles,
                     float* h_vertices, int64_t* n_vertices, uint32_t* h_indices,
                     int64_t* n_triangles);

/* Ray stream for the traversal benchmark: n_primary pinhole rays of a
 * res x res camera at `eye` looking at the origin with vertical fov `fov_deg`,
 * then n_bounce cosine-distributed rays leaving seeded points of the mesh
 * (origins offset with pine's OffsetRayOrigin, ray.h:34-45). */
int pg_gen_rays(const float* h_vertices, int64_t n_vertices, const uint32_t* h_indices,
                int64_t n_triangles, int32_t res, const float eye[3], float fov_deg,
                int64_t n_bounce, uint64_t seed, pg_ray* h_rays);

/* Heterogeneous-medium workload: an nx*ny*nz density grid, voxel value
 * max(0, fBm6(voxel centre in [0,1]^3) - threshold) * scale, fBm of value noise
 * with base frequency base_freq, x fastest (pine's raw volume order,
 * fileio.cpp:194-203).  Deterministic for a seed. */
int pg_gen_density_grid(int32_t nx, int32_t ny, int32_t nz, float base_freq, float threshold,
                        float scale, uint64_t seed, float* h_density);

/* Writes a triangle mesh as OBJ text (`v x y z` with six decimals, no
 * exponent, as pine's stof reads them; `f a b c` 1-based), formatted on all
 * host cores: workload files of tens of millions of triangles in seconds. */
int pg_write_obj(const char* path, const float* h_vertices, int64_t n_vertices,
                 const uint32_t* h_indices, int64_t n_triangles);

/* --- Learned path guiding (no counterpart in pine) ---------------------------
 * A 225-64-64-64-41 MLP (BF16 weights, FP32 accumulation on the matrix cores)
 * mapping a surface point to an 8-lobe von Mises-Fisher mixture over incident
 * directions plus the share alpha of directions drawn from it; the rest come
 * from the BSDF and the two are combined by one-sample MIS.  Inputs x[9]:
 * position normalised to the scene bounds, shading normal, outgoing direction.
 * Parameters: PG_GUIDE_N_PARAMS floats, the padded layout
 *   W0[64][256] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[48][64] b3[48]
 * (W0 columns >= 225 and the 7 W3/b3 rows no lobe or alpha uses are padding,
 * kept at zero). */
#define PG_GUIDE_N_PARAMS 27888
typedef struct pg_guide pg_guide;
typedef struct pg_guide_config {
    uint64_t seed;         /* initial weights */
    float learning_rate;   /* Adam step size; 0 = 1e-3 */
    int32_t freq_offset;   /* encoding frequencies pi 2^(k + freq_offset), k = 0..11 */
} pg_guide_config;
int pg_guide_create(const pg_guide_config* cfg, pg_guide** out);
int pg_guide_set_params(pg_guide* g, const float* h_params);
int pg_guide_get_params(const pg_guide* g, float* h_params);
/* Raw network outputs, n x 48 (41 live), for n inputs (n x 9), device buffers. */
int pg_guide_forward(pg_guide* g, const float* d_inputs, int64_t n, float* d_out, void* stream);
/* n x 3 out: guide density q(w), alpha, and q_mix = alpha q + (1 - alpha) p_bsdf. */
int pg_guide_pdf(pg_guide* g, const float* d_inputs, const float* d_dirs, const float* d_pbsdf,
                 int64_t n, float* d_out, void* stream);
/* n x 4 out: a direction drawn from the mixture with uniforms d_u (n x 3), and q there. */
int pg_guide_sample(pg_guide* g, const float* d_inputs, const float* d_u, int64_t n, float* d_out,
                    void* stream);
/* Gradient (host, PG_GUIDE_N_PARAMS floats) of the mean over the n samples of
 * -weight * log q_mix(dir), and that mean loss; weights unchanged.  Synchronous. */
int pg_guide_gradient(pg_guide* g, const float* d_inputs, const float* d_dirs,
                      const float* d_weight, const float* d_pbsdf, int64_t n, float* h_grad,
                      float* h_loss);
/* `steps` Adam steps over consecutive batches of `batch` samples (wrapping);
 * h_loss receives the last step's mean loss.  Synchronous. */
int pg_guide_train(pg_guide* g, const float* d_inputs, const float* d_dirs, const float* d_weight,
                   const float* d_pbsdf, int64_t n, int32_t steps, int32_t batch, float* h_loss);
/* pg_guide_train without allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in pine or elsewhere in this ABI.  Same
 * kernels and arithmetic as pg_guide_train, so the same parameters bit for
 * bit, but: stream-ordered (no synchronisation; nothing is read back), the
 * last step's mean loss goes to device memory d_loss (1 float, may be NULL),
 * and each step's partial-gradient rows live in a static device array of
 * PG_GUIDE_STATIC_MAX_BATCH / 64 rows (about 3.6 MB per GPU, held until the
 * process ends), so batch must not exceed PG_GUIDE_STATIC_MAX_BATCH
 * (PG_E_ARG otherwise).  Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams or threads) race. */
#define PG_GUIDE_STATIC_MAX_BATCH 2048
int pg_guide_train_static(pg_guide* g, const float* d_inputs, const float* d_dirs,
                          const float* d_weight, const float* d_pbsdf, int64_t n, int32_t steps,
                          int32_t batch, float* d_loss, void* stream);
int pg_guide_release(pg_guide* g);

/* Guided rendering (no counterpart in pine): `train_passes` passes of
 * `train_spp` samples per pixel sample directions from the current guide
 * (one-sample MIS with the BSDF at single-layer diffuse vertices), record the
 * radiance each sampled direction brought back, and train the guide on the
 * records that brought radiance (`steps_per_pass` Adam steps of `batch`
 * records; 0 = one epoch, every record once, in at most 32 steps of at least
 * `batch` records); then the film is cleared and `final_spp` samples per
 * pixel are rendered with the trained guide.  Every pass and the final render
 * take their own range of each pixel's sample sequence, from `sample_offset`
 * on, so the guide never estimates with the samples it learned from.  The
 * estimator stays unbiased for any guide.  Zero fields take the defaults
 * 16 / 8 / one epoch / 1024 / 5 / the scene's spp / 0.  Synchronous. */
typedef struct pg_guided_options {
    int32_t train_passes;
    int32_t train_spp;
    int32_t steps_per_pass;
    int32_t batch;
    int32_t record_depth;  /* path vertices recorded per path */
    int32_t final_spp;
    int32_t sample_offset;  /* first sample index of the first training pass */
    int32_t guide_depth;    /* vertices of depth < guide_depth use the guide (0 = all); deeper
                               ones sample their BSDF and cost no network evaluation */
    int32_t flags;          /* PG_RENDER_* bits for the final render */
    int32_t keep_passes;    /* the last keep_passes training passes stay in the final film (0 =
                               none): each is an unbiased guided render of its own sample range,
                               so the film holds train_spp * keep_passes + final_spp samples */
    int32_t rr_depth;       /* Russian roulette (pine's rule, path.cpp:76-82) only after this
                               depth in guided renders (0 = pine's 2); any depth keeps the
                               estimator unbiased */
} pg_guided_options;
int pg_render_guided(pg_scene* s, pg_guide* g, const pg_guided_options* o, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PINEHIP_H */
