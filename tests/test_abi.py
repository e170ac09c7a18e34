"""The C ABI (include/pinehip.h) on the CPU: the library loads, exports every
declared symbol, validates arguments, and its host-only workload generators are
deterministic.  No GPU work is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(pg):
    L = pg.lib()
    names = declared_functions()
    assert len(names) >= 16
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/ but not exported"
    assert set(pg.EXPORTED) <= set(names)


def test_version_string(pg):
    assert b"gfx950" in pg.lib().pg_version()


def test_argument_errors_without_gpu(pg):
    L = pg.lib()
    assert L.pg_accel_create(None) == -1
    h = ctypes.c_void_p()
    assert L.pg_accel_create(ctypes.byref(h)) == 0
    v = np.zeros((3, 3), np.float32)
    bad = np.array([[0, 1, 7]], np.uint32)
    first = ctypes.c_int32()
    rc = L.pg_accel_add_mesh(h, ctypes.c_void_p(v.ctypes.data), 3, ctypes.c_void_p(bad.ctypes.data),
                             1, ctypes.byref(first))
    assert rc == -1 and b"out of range" in L.pg_last_error()
    assert L.pg_accel_add_mesh(h, None, 0, None, 0, None) == -1
    # Trace before build is a state error, caught before any device work.
    assert L.pg_trace_closest(h, ctypes.c_void_p(16), 1, ctypes.c_void_p(16), None, None) == -2
    assert L.pg_trace_any(h, ctypes.c_void_p(16), 1, ctypes.c_void_p(16), None) == -2
    assert L.pg_accel_set_launch(h, 96, -1) == -1
    assert L.pg_accel_set_launch(h, 64, 32) == -1 and b"flags" in L.pg_last_error()
    assert L.pg_accel_set_launch(h, 64, 8 | 16) == 0  # reorder + timed
    assert L.pg_accel_set_launch(h, 0, 3) == 0
    c = (ctypes.c_float * 3)(0, 0, 0)
    assert L.pg_accel_add_sphere(h, c, ctypes.c_float(-1.0), None) == -1
    assert L.pg_accel_release(h) == 0
    assert L.pg_accel_release(None) == 0


def test_icosphere_sizes_and_watertightness(pg):
    v, i = pg.gen_icosphere(8, 0.1, 3)
    assert i.shape == (20 * 64, 3) and v.shape == (10 * 64 + 2, 3)
    # Closed 2-manifold: every undirected edge is used by exactly two triangles,
    # once in each direction.
    e = np.concatenate([i[:, [0, 1]], i[:, [1, 2]], i[:, [2, 0]]])
    directed = {tuple(x) for x in e.tolist()}
    assert len(directed) == len(e)
    assert all((b, a) in directed for a, b in directed)
    r = np.linalg.norm(v, axis=1)
    assert r.min() > 0.85 and r.max() < 1.15


def test_icosphere_truncation(pg):
    v, i = pg.gen_icosphere(229, 0.1, 1, 1 << 20)
    assert len(i) == 1 << 20
    assert i.max() < len(v)


def test_generators_are_deterministic(pg):
    v1, i1 = pg.gen_icosphere(12, 0.1, 5)
    v2, i2 = pg.gen_icosphere(12, 0.1, 5)
    assert v1.tobytes() == v2.tobytes() and i1.tobytes() == i2.tobytes()
    r1 = pg.gen_rays(v1, i1, 16, (0, 0.6, 3), 45.0, 500, 9)
    r2 = pg.gen_rays(v1, i1, 16, (0, 0.6, 3), 45.0, 500, 9)
    assert r1.tobytes() == r2.tobytes()
    d = r1[:, 4:7]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    # Bounce rays leave the surface outward (cosine lobe around the normal).
    o = r1[256:, 0:3]
    assert (np.einsum("ij,ij->i", o, d[256:]) > -0.2).mean() > 0.95


def test_generator_rejects_bad_arguments(pg):
    L = pg.lib()
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    assert L.pg_gen_icosphere(0, ctypes.c_float(0.1), 1, 0, None, ctypes.byref(nv), None,
                              ctypes.byref(nt)) == -1


def test_density_grid_generator_is_deterministic_and_sparse(pg):
    """The volume workload (and the render_cloud fixture's grid) is a pure
    function of its parameters."""
    import hashlib
    a = pg.gen_density_grid(32, base_freq=2.0)
    b = pg.gen_density_grid(32, base_freq=2.0)
    assert np.array_equal(a, b)
    assert a.dtype == np.float32 and a.shape == (32, 32, 32)
    assert (a >= 0).all() and 0.02 < (a > 0).mean() < 0.3 and a.max() < 40 * 0.7
    assert not np.array_equal(a, pg.gen_density_grid(32, base_freq=2.0, seed=4))
    # Voxels at threshold 0 and scale 1 are the raw fBm clipped at zero.
    raw = pg.gen_density_grid(16, threshold=0.0, scale=1.0)
    assert raw.max() < 1.0
    h = hashlib.sha256(a.tobytes()).hexdigest()
    print("cloud32 sha256", h)


def test_volume_file_layout(pg, tmp_path):
    d = np.arange(2 * 3 * 4, dtype=np.float32).reshape(4, 3, 2)  # [z, y, x]
    pg.write_volume(str(tmp_path / "v.vol"), d)
    raw = (tmp_path / "v.vol").read_bytes()
    assert np.frombuffer(raw[:12], np.int32).tolist() == [2, 3, 4]
    vals = np.frombuffer(raw[12:], np.float32)
    assert vals[1] == d[0, 0, 1] and vals[2] == d[0, 1, 0] and vals[6] == d[1, 0, 0]


def test_write_obj_fixed_round_trips_to_six_decimals(pg, tmp_path):
    v = np.array([[0.0, -1.25, 3.0000004], [123456.5, -0.0000004, 1e-7], [-2.5, 0.1, 7.0]], np.float32)
    i = np.array([[0, 1, 2], [2, 1, 0]], np.uint32)
    p = str(tmp_path / "m.obj")
    pg.write_obj_fixed(p, v, i)
    lines = open(p).read().splitlines()
    assert len(lines) == 5 and all("e" not in l for l in lines)
    got = np.array([[float(x) for x in l.split()[1:]] for l in lines[:3]])
    assert np.abs(got - v).max() <= 5e-7 + 1e-7 * np.abs(v).max()
    assert lines[3] == "f 1 2 3" and lines[4] == "f 3 2 1"
    with pytest.raises(pg.PineError):
        pg.write_obj_fixed(p, v, np.array([[0, 1, 3]], np.uint32))


def test_batch_workload_is_seeded_and_spans_its_ranges():
    from pine_amd import workloads as W
    a, b = W.batch_workload(), W.batch_workload()
    assert [x["triangles"] for x in a] == [x["triangles"] for x in b] and len(a) == 64
    tr = np.array([x["triangles"] for x in a])
    assert tr.min() >= 10_000 and tr.max() <= 2_000_000
    assert {x["width"] for x in a} <= set(range(256, 1025, 128))
    assert {x["spp"] for x in a} <= set(range(16, 129, 8))
    v, i = W.make_mesh(a[0]["mesh"])
    assert len(i.reshape(-1, 3)) == a[0]["triangles"]
    assert "TriangleMesh" in W.batch_scene_text(a[0], "x.obj")
