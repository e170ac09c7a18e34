"""pine_amd: MI355X (gfx950) implementation of pine's path-tracing hot path.

The native library ``libpinehip.so`` (built in-tree by ``pine_amd.build``)
exports the C ABI declared in ``include/pinehip.h``; this module binds it with
ctypes for tests and the benchmark.  Device buffers are ``pine_amd.hip.DeviceArray``
allocations made through the same HIP runtime the library links.

There is no fallback: if the library is missing or fails to load, every entry
point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpinehip.so")

PG_OK = 0
# pg_accel_set_launch flags (include/pinehip.h)
LAUNCH_POSTPONE_LEAVES = 1
LAUNCH_XCD_REMAP = 2
LAUNCH_STACKLESS = 4
LAUNCH_REORDER = 8
LAUNCH_TIMED = 16
LAUNCH_DEFAULT = 0
EXPORTED = [
    "pg_last_error", "pg_version", "pg_accel_create", "pg_accel_add_mesh",
    "pg_accel_add_sphere", "pg_accel_add_rect", "pg_accel_add_plane", "pg_accel_build",
    "pg_trace_closest", "pg_trace_any", "pg_accel_get_stats", "pg_accel_prim_shape",
    "pg_accel_set_launch", "pg_accel_release", "pg_gen_icosphere", "pg_gen_rays",
    "pg_trace_closest_counts", "pg_trace_occluded", "pg_scene_load", "pg_scene_load_text",
    "pg_scene_get_info", "pg_render", "pg_scene_film", "pg_scene_output_name", "pg_scene_release",
    "pg_gen_density_grid", "pg_guide_create", "pg_guide_set_params", "pg_guide_get_params",
    "pg_guide_forward", "pg_guide_pdf", "pg_guide_sample", "pg_guide_gradient", "pg_guide_train",
    "pg_guide_release", "pg_render_guided", "pg_trace_closest_dev", "pg_trace_occluded_dev",
    "pg_accel_rebuild", "pg_render_batch", "pg_scene_autotune", "pg_write_obj",
    "pg_spectral_tables", "pg_scene_sampler_values", "pg_guide_train_static",
]
RENDER_NO_GRAPH = 1
RENDER_SPECTRAL = 2
RENDER_LANE_TRACKING = 4
GUIDE_N_PARAMS = 27888
GUIDE_STATIC_MAX_BATCH = 2048


class PineError(RuntimeError):
    pass


class _BuildOptions(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("leaf_max", ctypes.c_int32),
                ("bins", ctypes.c_int32), ("threads", ctypes.c_int32)]


class _Stats(ctypes.Structure):
    _fields_ = [("n_prims", ctypes.c_int64), ("n_nodes", ctypes.c_int64),
                ("node_bytes", ctypes.c_int64), ("prim_bytes", ctypes.c_int64),
                ("width", ctypes.c_int32), ("max_depth", ctypes.c_int32),
                ("build_ms", ctypes.c_double), ("block", ctypes.c_int32),
                ("launch_flags", ctypes.c_int32), ("last_trace_ms", ctypes.c_double)]


class _GuideConfig(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("learning_rate", ctypes.c_float),
                ("freq_offset", ctypes.c_int32)]


class _SceneInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("spp", ctypes.c_int32),
                ("max_depth", ctypes.c_int32), ("n_shapes", ctypes.c_int32),
                ("n_lights", ctypes.c_int32), ("n_prims", ctypes.c_int64),
                ("last_render_ms", ctypes.c_double), ("last_rays", ctypes.c_int64),
                ("n_media", ctypes.c_int32), ("general_mode", ctypes.c_int32),
                ("medium_steps", ctypes.c_int64), ("last_guided_ms", ctypes.c_double),
                ("guide_steps", ctypes.c_int64), ("guide_records", ctypes.c_int64),
                ("graph_captures", ctypes.c_int64), ("paths_in_flight", ctypes.c_int64),
                ("trace_block", ctypes.c_int32), ("trace_flags", ctypes.c_int32),
                ("bvh_width", ctypes.c_int32), ("last_extension_rays", ctypes.c_int64),
                ("apply_tone_mapping", ctypes.c_int32), ("guide_infer_ms", ctypes.c_double),
                ("guide_infer_queries", ctypes.c_int64), ("guide_step_ms", ctypes.c_double)]


class _GuidedOptions(ctypes.Structure):
    _fields_ = [("train_passes", ctypes.c_int32), ("train_spp", ctypes.c_int32),
                ("steps_per_pass", ctypes.c_int32), ("batch", ctypes.c_int32),
                ("record_depth", ctypes.c_int32), ("final_spp", ctypes.c_int32),
                ("sample_offset", ctypes.c_int32), ("guide_depth", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("keep_passes", ctypes.c_int32),
                ("rr_depth", ctypes.c_int32)]


class _RenderOptions(ctypes.Structure):
    _fields_ = [("spp", ctypes.c_int32), ("samples_per_batch", ctypes.c_int32),
                ("sample_offset", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("rr_depth", ctypes.c_int32)]


class _BatchOptions(ctypes.Structure):
    _fields_ = [("streams", ctypes.c_int32), ("flags", ctypes.c_int32), ("paths_per_batch", ctypes.c_int64)]


class _BatchResult(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double), ("samples", ctypes.c_int64), ("rays", ctypes.c_int64)]


class _TuneOptions(ctypes.Structure):
    _fields_ = [("probe_spp", ctypes.c_int32), ("try_width", ctypes.c_int32)]


class _TuneResult(ctypes.Structure):
    _fields_ = [("trace_block", ctypes.c_int32), ("trace_flags", ctypes.c_int32),
                ("bvh_width", ctypes.c_int32), ("probes", ctypes.c_int32),
                ("paths_in_flight", ctypes.c_int64), ("default_ms", ctypes.c_double),
                ("best_ms", ctypes.c_double)]


_lib = None


def lib():
    """Loads libpinehip.so once; raises PineError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PineError(f"{LIB_PATH} is missing: build it with pine_amd.build.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    fp = ctypes.POINTER(ctypes.c_float)
    L.pg_last_error.restype = ctypes.c_char_p
    L.pg_version.restype = ctypes.c_char_p
    L.pg_accel_create.argtypes = [ctypes.POINTER(vp)]
    L.pg_accel_add_mesh.argtypes = [vp, vp, i64, vp, i64, ctypes.POINTER(i32)]
    L.pg_accel_add_sphere.argtypes = [vp, fp, ctypes.c_float, ctypes.POINTER(i32)]
    L.pg_accel_add_rect.argtypes = [vp, fp, fp, fp, ctypes.POINTER(i32)]
    L.pg_accel_add_plane.argtypes = [vp, fp, fp, ctypes.POINTER(i32)]
    L.pg_accel_build.argtypes = [vp, ctypes.POINTER(_BuildOptions)]
    L.pg_trace_closest.argtypes = [vp, vp, i64, vp, vp, vp]
    L.pg_trace_any.argtypes = [vp, vp, i64, vp, vp]
    L.pg_trace_closest_counts.argtypes = [vp, vp, i64, vp, vp, vp]
    L.pg_trace_occluded.argtypes = [vp, vp, i64, vp, vp]
    L.pg_scene_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(_BuildOptions), ctypes.POINTER(vp)]
    L.pg_scene_load_text.argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.POINTER(_BuildOptions), ctypes.POINTER(vp)]
    L.pg_scene_get_info.argtypes = [vp, ctypes.POINTER(_SceneInfo)]
    L.pg_render.argtypes = [vp, ctypes.POINTER(_RenderOptions), vp]
    L.pg_scene_film.argtypes = [vp, vp]
    L.pg_scene_sampler_values.argtypes = [vp, i32, i32, i32, i32, i32, vp]
    L.pg_scene_output_name.argtypes = [vp, ctypes.c_char_p, i32]
    L.pg_scene_release.argtypes = [vp]
    L.pg_accel_get_stats.argtypes = [vp, ctypes.POINTER(_Stats)]
    L.pg_accel_prim_shape.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.pg_accel_set_launch.argtypes = [vp, i32, i32]
    L.pg_accel_release.argtypes = [vp]
    L.pg_gen_icosphere.argtypes = [i32, ctypes.c_float, ctypes.c_uint64, i64, vp,
                                   ctypes.POINTER(i64), vp, ctypes.POINTER(i64)]
    L.pg_gen_rays.argtypes = [vp, i64, vp, i64, i32, fp, ctypes.c_float, i64,
                              ctypes.c_uint64, vp]
    L.pg_gen_density_grid.argtypes = [i32, i32, i32, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_uint64, vp]
    L.pg_guide_create.argtypes = [ctypes.POINTER(_GuideConfig), ctypes.POINTER(vp)]
    L.pg_guide_set_params.argtypes = [vp, vp]
    L.pg_guide_get_params.argtypes = [vp, vp]
    L.pg_guide_forward.argtypes = [vp, vp, i64, vp, vp]
    L.pg_guide_pdf.argtypes = [vp, vp, vp, vp, i64, vp, vp]
    L.pg_guide_sample.argtypes = [vp, vp, vp, i64, vp, vp]
    L.pg_guide_gradient.argtypes = [vp, vp, vp, vp, vp, i64, vp, ctypes.POINTER(ctypes.c_float)]
    L.pg_guide_train.argtypes = [vp, vp, vp, vp, vp, i64, i32, i32, ctypes.POINTER(ctypes.c_float)]
    L.pg_guide_train_static.argtypes = [vp, vp, vp, vp, vp, i64, i32, i32, vp, vp]
    L.pg_guide_release.argtypes = [vp]
    L.pg_render_guided.argtypes = [vp, vp, ctypes.POINTER(_GuidedOptions), vp]
    L.pg_trace_closest_dev.argtypes = [vp, vp, vp, i64, vp, vp, vp]
    L.pg_trace_occluded_dev.argtypes = [vp, vp, vp, i64, vp, vp]
    L.pg_accel_rebuild.argtypes = [vp, ctypes.POINTER(_BuildOptions)]
    L.pg_render_batch.argtypes = [vp, i32, vp, ctypes.POINTER(_BatchOptions), vp,
                                  ctypes.POINTER(_BatchResult)]
    L.pg_scene_autotune.argtypes = [vp, ctypes.POINTER(_TuneOptions), vp, ctypes.POINTER(_TuneResult)]
    L.pg_write_obj.argtypes = [ctypes.c_char_p, vp, i64, vp, i64]
    L.pg_spectral_tables.argtypes = [vp, vp, vp]
    _lib = L
    return L


def check(rc):
    if rc != PG_OK:
        raise PineError(f"pinehip error {rc}: {lib().pg_last_error().decode()}")
    return rc


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _stream_handle():
    return None  # the default stream


def _dptr(a):
    return a.ptr if a is not None else None


# --- workloads -------------------------------------------------------------------

def gen_icosphere(freq, amp=0.1, seed=1, max_triangles=0):
    """Displaced geodesic icosphere; returns (vertices float32 [nv,3], indices uint32 [nt,3])."""
    L = lib()
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    check(L.pg_gen_icosphere(freq, amp, seed, max_triangles, None, ctypes.byref(nv), None,
                             ctypes.byref(nt)))
    v = np.zeros((nv.value, 3), np.float32)
    i = np.zeros((nt.value, 3), np.uint32)
    check(L.pg_gen_icosphere(freq, amp, seed, max_triangles, _ptr(v), ctypes.byref(nv), _ptr(i),
                             ctypes.byref(nt)))
    return v[: nv.value], i[: nt.value]


def gen_rays(vertices, indices, res, eye, fov_deg, n_bounce, seed=7):
    """Traversal-benchmark rays as float32 [n, 8] = (o.xyz, tmin, d.xyz, tmax)."""
    v = np.ascontiguousarray(vertices, np.float32)
    i = np.ascontiguousarray(indices, np.uint32)
    n = res * res + n_bounce
    rays = np.zeros((n, 8), np.float32)
    check(lib().pg_gen_rays(_ptr(v), len(v), _ptr(i), len(i), res, _f3(eye), fov_deg, n_bounce,
                            seed, _ptr(rays)))
    return rays


def gen_density_grid(n, base_freq=4.0, threshold=0.3, scale=40.0, seed=3):
    """fBm density grid, float32 [n, n, n] indexed [z, y, x] (x fastest in memory):
    max(0, fbm6 - threshold) * scale."""
    out = np.zeros((n, n, n), np.float32)
    check(lib().pg_gen_density_grid(n, n, n, base_freq, threshold, scale, seed, _ptr(out)))
    return out


def write_obj_fixed(path, vertices, indices):
    """Writes a mesh as OBJ text with six-decimal coordinates (pg_write_obj, all
    host cores).  For workload files; the scenes compared with pine use
    workloads.write_obj, whose decimals reproduce each float exactly."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
    check(lib().pg_write_obj(path.encode(), _ptr(v), v.shape[0], _ptr(i), i.shape[0]))


def spectral_tables():
    """(Q, M, w_scale) of the spectral mode (pg_spectral_tables)."""
    Q, M, w = np.zeros(9, np.float32), np.zeros(9, np.float32), np.zeros(1, np.float32)
    check(lib().pg_spectral_tables(_ptr(Q), _ptr(M), _ptr(w)))
    return Q.reshape(3, 3), M.reshape(3, 3), float(w[0])


def write_volume(path, density):
    """Writes pine's raw volume file (fileio.cpp:194-203): int32 size x, y, z, then
    the floats, x fastest.  `density` is indexed [z, y, x]."""
    d = np.ascontiguousarray(density, np.float32)
    nz, ny, nx = d.shape
    with open(path, "wb") as f:
        f.write(np.array([nx, ny, nz], np.int32).tobytes())
        f.write(d.tobytes())


# --- accel -----------------------------------------------------------------------

class Accel:
    """Owns a pg_accel handle (pine's Accel, src/core/accel.h:8-14)."""

    def __init__(self):
        self.h = ctypes.c_void_p()
        check(lib().pg_accel_create(ctypes.byref(self.h)))

    def add_mesh(self, vertices, indices):
        v = np.ascontiguousarray(vertices, np.float32)
        i = np.ascontiguousarray(indices, np.uint32)
        first = ctypes.c_int32()
        check(lib().pg_accel_add_mesh(self.h, _ptr(v), len(v), _ptr(i), len(i),
                                      ctypes.byref(first)))
        return first.value

    def add_sphere(self, c, r):
        p = ctypes.c_int32()
        check(lib().pg_accel_add_sphere(self.h, _f3(c), r, ctypes.byref(p)))
        return p.value

    def add_rect(self, position, ex, ey):
        p = ctypes.c_int32()
        check(lib().pg_accel_add_rect(self.h, _f3(position), _f3(ex), _f3(ey), ctypes.byref(p)))
        return p.value

    def add_plane(self, position, normal):
        p = ctypes.c_int32()
        check(lib().pg_accel_add_plane(self.h, _f3(position), _f3(normal), ctypes.byref(p)))
        return p.value

    def build(self, width=2, leaf_max=4, bins=32, threads=0):
        o = _BuildOptions(width, leaf_max, bins, threads)
        check(lib().pg_accel_build(self.h, ctypes.byref(o)))
        return self

    def set_launch(self, block=0, flags=-1):
        """Trace launch shape: block 64/128/256 (0 keeps), flags LAUNCH_* bits (-1 keeps)."""
        check(lib().pg_accel_set_launch(self.h, block, flags))

    def stats(self):
        s = _Stats()
        check(lib().pg_accel_get_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}

    def prim_shape(self, prim):
        s, l = ctypes.c_int32(), ctypes.c_int32()
        check(lib().pg_accel_prim_shape(self.h, prim, ctypes.byref(s), ctypes.byref(l)))
        return s.value, l.value

    def trace_closest(self, d_rays, d_hits=None, d_uv=None):
        """d_rays: DeviceArray float32 [n, 8].  Returns DeviceArray int32 [n, 2]
        (prim, bits of t)."""
        from .hip import DeviceArray
        n = d_rays.shape[0]
        if d_hits is None:
            d_hits = DeviceArray((n, 2), np.int32)
        check(lib().pg_trace_closest(self.h, d_rays.ptr, n, _dptr(d_hits), _dptr(d_uv),
                                     _stream_handle()))
        return d_hits

    def trace_counts(self, d_rays):
        """Closest hit plus per-ray (nodes fetched, prims tested), int32 [n, 2]."""
        from .hip import DeviceArray
        n = d_rays.shape[0]
        d_hits = DeviceArray((n, 2), np.int32)
        d_cnt = DeviceArray((n, 2), np.int32)
        check(lib().pg_trace_closest_counts(self.h, d_rays.ptr, n, d_hits.ptr, d_cnt.ptr,
                                            _stream_handle()))
        return d_hits, d_cnt

    def trace_any(self, d_rays, d_occ=None):
        from .hip import DeviceArray
        n = d_rays.shape[0]
        if d_occ is None:
            d_occ = DeviceArray((n,), np.int32)
        check(lib().pg_trace_any(self.h, d_rays.ptr, n, d_occ.ptr, _stream_handle()))
        return d_occ

    def trace_occluded(self, d_rays, d_occ=None):
        from .hip import DeviceArray
        n = d_rays.shape[0]
        if d_occ is None:
            d_occ = DeviceArray((n,), np.int32)
        check(lib().pg_trace_occluded(self.h, d_rays.ptr, n, d_occ.ptr, _stream_handle()))
        return d_occ

    def release(self):
        if self.h:
            lib().pg_accel_release(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# --- scene + Path integrator ---------------------------------------------------------

class Scene:
    """A loaded pine scene on the GPU (pine's LoadScene + Path integrator)."""

    def __init__(self, path=None, text=None, directory=None, width=2):
        self.h = ctypes.c_void_p()
        o = _BuildOptions(width, 4, 32, 0)
        if path is not None:
            check(lib().pg_scene_load(path.encode(), ctypes.byref(o), ctypes.byref(self.h)))
        else:
            check(lib().pg_scene_load_text(text.encode(), (directory or ".").encode(),
                                           ctypes.byref(o), ctypes.byref(self.h)))

    def info(self):
        s = _SceneInfo()
        check(lib().pg_scene_get_info(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _SceneInfo._fields_}

    def render(self, spp=0, samples_per_batch=0, sample_offset=0, graph=True, spectral=False,
               lane_tracking=False, rr_depth=0):
        flags = (0 if graph else RENDER_NO_GRAPH) | (RENDER_SPECTRAL if spectral else 0) | \
            (RENDER_LANE_TRACKING if lane_tracking else 0)
        o = _RenderOptions(spp, samples_per_batch, sample_offset, flags, rr_depth)
        check(lib().pg_render(self.h, ctypes.byref(o), _stream_handle()))
        return self.info()

    def autotune(self, probe_spp=1, try_width=False):
        """Picks the launch configuration by timing probe renders (pg_scene_autotune)."""
        o, r = _TuneOptions(probe_spp, 1 if try_width else 0), _TuneResult()
        check(lib().pg_scene_autotune(self.h, ctypes.byref(o), _stream_handle(), ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in _TuneResult._fields_}

    def render_guided(self, guide, train_passes=16, train_spp=8, steps_per_pass=0, batch=1024,
                      record_depth=5, final_spp=0, sample_offset=0, guide_depth=0, graph=True,
                      keep_passes=0, rr_depth=0):
        """Trains `guide` on this scene, then renders final_spp with it (pg_render_guided);
        the last keep_passes training passes stay in the film."""
        o = _GuidedOptions(train_passes, train_spp, steps_per_pass, batch, record_depth, final_spp,
                           sample_offset, guide_depth, 0 if graph else RENDER_NO_GRAPH, keep_passes,
                           rr_depth)
        check(lib().pg_render_guided(self.h, guide.h, ctypes.byref(o), _stream_handle()))
        return self.info()

    def film(self):
        """Film sums [h, w, 4] = (sum w*r, sum w*g, sum w*b, sum w), row 0 on top."""
        inf = self.info()
        out = np.zeros((inf["height"], inf["width"], 4), np.float32)
        check(lib().pg_scene_film(self.h, _ptr(out)))
        return out

    def sampler_values(self, px0, py0, npix, spp, ndims):
        """The scene's sampler on the device, [npix, spp, ndims] (pg_scene_sampler_values)."""
        out = np.zeros((npix, spp, ndims), np.float32)
        check(lib().pg_scene_sampler_values(self.h, px0, py0, npix, spp, ndims, _ptr(out)))
        return out

    def image(self):
        """Linear radiance [h, w, 3] as Film::CopyToRGBArray computes it."""
        f = self.film()
        w = f[..., 3:4]
        return np.where(w != 0, f[..., :3] / np.where(w != 0, w, 1), 0).astype(np.float32)

    def release(self):
        if self.h:
            lib().pg_scene_release(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def render_batch(scenes, spp=None, streams=4, graph=True, paths_per_batch=0):
    """Renders a list of Scene objects as one batch (pg_render_batch); returns
    the batch wall time, camera samples and rays.  Every scene's film and info
    hold its own result afterwards.  paths_per_batch > 0 caps every scene's
    batch (0: each scene's own paths in flight)."""
    n = len(scenes)
    arr = (ctypes.c_void_p * max(n, 1))(*[s.h for s in scenes])
    sp = None
    if spp is not None:
        sp = (ctypes.c_int32 * max(n, 1))(*[int(x) for x in spp])
    o, r = _BatchOptions(streams, 0 if graph else RENDER_NO_GRAPH, int(paths_per_batch)), _BatchResult()
    check(lib().pg_render_batch(arr, n, sp, ctypes.byref(o), _stream_handle(), ctypes.byref(r)))
    return {k: getattr(r, k) for k, _ in _BatchResult._fields_}


# --- learned path guiding ----------------------------------------------------------

class Guide:
    """Owns a pg_guide (the path-guiding MLP).  Batch methods take and return
    device arrays (pine_amd.hip.DeviceArray)."""

    def __init__(self, seed=1, learning_rate=1e-3, freq_offset=0):
        self.h = ctypes.c_void_p()
        cfg = _GuideConfig(seed, learning_rate, freq_offset)
        check(lib().pg_guide_create(ctypes.byref(cfg), ctypes.byref(self.h)))

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.shape == (GUIDE_N_PARAMS,)
        check(lib().pg_guide_set_params(self.h, _ptr(p)))

    def params(self):
        p = np.zeros(GUIDE_N_PARAMS, np.float32)
        check(lib().pg_guide_get_params(self.h, _ptr(p)))
        return p

    def forward(self, d_x):
        from .hip import DeviceArray
        out = DeviceArray((d_x.shape[0], 48), np.float32)
        check(lib().pg_guide_forward(self.h, d_x.ptr, d_x.shape[0], out.ptr, _stream_handle()))
        return out

    def pdf(self, d_x, d_dirs, d_pbsdf):
        from .hip import DeviceArray
        out = DeviceArray((d_x.shape[0], 3), np.float32)
        check(lib().pg_guide_pdf(self.h, d_x.ptr, d_dirs.ptr, d_pbsdf.ptr, d_x.shape[0], out.ptr,
                                 _stream_handle()))
        return out

    def sample(self, d_x, d_u):
        from .hip import DeviceArray
        out = DeviceArray((d_x.shape[0], 4), np.float32)
        check(lib().pg_guide_sample(self.h, d_x.ptr, d_u.ptr, d_x.shape[0], out.ptr,
                                    _stream_handle()))
        return out

    def gradient(self, d_x, d_dirs, d_weight, d_pbsdf):
        g = np.zeros(GUIDE_N_PARAMS, np.float32)
        loss = ctypes.c_float()
        check(lib().pg_guide_gradient(self.h, d_x.ptr, d_dirs.ptr, d_weight.ptr, d_pbsdf.ptr,
                                      d_x.shape[0], _ptr(g), ctypes.byref(loss)))
        return g, loss.value

    def train(self, d_x, d_dirs, d_weight, d_pbsdf, steps, batch):
        loss = ctypes.c_float()
        check(lib().pg_guide_train(self.h, d_x.ptr, d_dirs.ptr, d_weight.ptr, d_pbsdf.ptr,
                                   d_x.shape[0], steps, batch, ctypes.byref(loss)))
        return loss.value

    def train_static(self, d_x, d_dirs, d_weight, d_pbsdf, steps, batch, d_loss=None, stream=None):
        """pg_guide_train_static: the same steps with no device allocation,
        stream-ordered; the last step's mean loss goes to device memory
        d_loss (one float) when given.  batch <= GUIDE_STATIC_MAX_BATCH."""
        check(lib().pg_guide_train_static(self.h, d_x.ptr, d_dirs.ptr, d_weight.ptr, d_pbsdf.ptr,
                                          d_x.shape[0], steps, batch,
                                          d_loss.ptr if d_loss is not None else None, stream))

    def release(self):
        if self.h:
            lib().pg_guide_release(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
