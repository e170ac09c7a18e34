"""The oracle: a CPU restatement of pine's path-tracing hot path (pinecpu.cpp)
plus, where the reference sources are present, pine itself built from those
sources (oracle/_ref/).

TEST INFRASTRUCTURE ONLY.  It may be used by tests/, by __graft_entry__.smoke()
as the checker and by bench.py's cpu_baseline leg - never by the shipped
library or its Python bindings, and never as the thing measured on the GPU.

Parity status of each restated function (how it is pinned) is in the header
of pinecpu.cpp and in tests/test_oracle.py.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpinecpu.so")
HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_lib = None


def build():
    """Builds libpinecpu.so (and, when the reference sources are present, the
    reference library and harness under _ref/)."""
    targets = ["restatement"]
    if os.path.isdir("/root/reference/src"):
        targets = ["all"]
    subprocess.run(["make", "-s", "-C", HERE, "-j8", *targets], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        for nm in ["sqrt", "exp", "log", "cos", "sin", "acos"]:
            f = getattr(L, "po_pstd_" + nm)
            f.argtypes = [ctypes.c_float]
            f.restype = ctypes.c_float
        L.po_pstd_pow.argtypes = [ctypes.c_float, ctypes.c_float]
        L.po_pstd_pow.restype = ctypes.c_float
        L.po_pstd_atan2.argtypes = [ctypes.c_float, ctypes.c_float]
        L.po_pstd_atan2.restype = ctypes.c_float
        L.po_trace_brute.argtypes = [vp, i64, vp, i64, vp, i64, vp, vp, vp, vp, vp, ctypes.c_int]
        L.po_bvh_build.argtypes = [vp, i64]
        L.po_bvh_build.restype = vp
        L.po_bvh_free.argtypes = [vp]
        L.po_bvh_trace.argtypes = [vp, vp, i64, vp, vp, ctypes.c_int]
        L.po_bvh_trace.restype = ctypes.c_double
        L.po_halton_values.argtypes = [ctypes.c_int] * 7 + [vp]
        L.po_sobol_values.argtypes = [ctypes.c_int] * 8 + [vp]
        L.po_noise.argtypes = [vp, vp, vp, i64, vp]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def pstd(name, *args):
    return getattr(lib(), "po_pstd_" + name)(*[float(a) for a in args])


def tri9(vertices, indices):
    """Per-triangle vertex triples [n, 9] in mesh order."""
    return np.ascontiguousarray(vertices[indices].reshape(-1, 9), np.float32)


def trace_brute(tri, rays, analytic=None, threads=0):
    """Closest (smallest t, ties to smaller prim id) and any hit by testing every
    prim.  analytic: [m, 10] rows {type, a.xyz, b.xyz, c.xyz} (1 sphere {c, r},
    2 rect {pos, ex, ey}, 3 plane {pos, normal}); their ids follow the triangles."""
    rays = np.ascontiguousarray(rays, np.float32)
    tri = np.ascontiguousarray(tri if tri is not None else np.zeros((0, 9)), np.float32)
    ana = np.ascontiguousarray(analytic if analytic is not None else np.zeros((0, 10)),
                               np.float32)
    n = len(rays)
    prim = np.empty(n, np.int32)
    t, u, v = (np.empty(n, np.float32) for _ in range(3))
    occ = np.empty(n, np.int32)
    lib().po_trace_brute(_p(tri), len(tri), _p(ana), len(ana), _p(rays), n, _p(prim), _p(t),
                         _p(u), _p(v), _p(occ), threads)
    return prim, t, u, v, occ


def halton_values(w, h, spp, ndims, px0, py0, npix):
    """pine's HaltonSampler values, layout [npix, spp, ndims] (see pinecpu.cpp)."""
    out = np.zeros(npix * spp * ndims, np.float32)
    lib().po_halton_values(w, h, spp, ndims, px0, py0, npix, _p(out))
    return out.reshape(npix, spp, ndims)


def sobol_values(w, h, spp, ndims, px0, py0, npix, strategy):
    """pine's SobolSampler values (strategy 0 none, 1 binaryPermutate, 2 fastOwen,
    3 Owen), layout [npix, spp, ndims] (see pinecpu.cpp)."""
    out = np.zeros(npix * spp * ndims, np.float32)
    lib().po_sobol_values(w, h, spp, ndims, px0, py0, npix, strategy, _p(out))
    return out.reshape(npix, spp, ndims)


def noise(points, freq, octaves):
    """pine's PerlinNoise (seeds 0..2), Turbulence and Turbulence3D at points
    [n, 3]: returns [n, 7] (see pinecpu.cpp)."""
    pts = np.ascontiguousarray(points, np.float32)
    fr = np.ascontiguousarray(freq, np.float32)
    oc = np.ascontiguousarray(octaves, np.float32)
    out = np.zeros((len(pts), 8), np.float32)
    lib().po_noise(_p(pts), _p(fr), _p(oc), len(pts), _p(out))
    return out[:, :7]


class CpuBVH:
    """pine-order BVH2 traversal on the host (the CPU baseline)."""

    def __init__(self, tri):
        self.tri = np.ascontiguousarray(tri, np.float32)
        self.h = lib().po_bvh_build(_p(self.tri), len(self.tri))

    def trace(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        secs = lib().po_bvh_trace(self.h, _p(rays), n, _p(prim), _p(t), threads)
        return prim, t, secs

    def __del__(self):
        try:
            lib().po_bvh_free(self.h)
        except Exception:
            pass
