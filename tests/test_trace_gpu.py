"""GPU parity of the trace kernels (pg_trace_closest / pg_trace_any) through the
C ABI, against pine's own BVH (golden fixtures) and the oracle.

Bar: t bit-exact on every hit; prim ids equal, except where pine's first-found
tie rule and the kernel's smaller-id rule pick different prims with bit-equal t
(certified ties); any-hit flags equal.
"""
import numpy as np
import pytest

from conftest import golden
from ties import certify_ties, mesh_prim_geom, single_prim

pytestmark = pytest.mark.gpu

WIDTHS = [2, 4]


def _accel(pg, v, i, width, block=64, flags=-1):
    a = pg.Accel()
    a.add_mesh(v, i)
    a.build(width=width)
    a.set_launch(block, flags)
    return a


def _trace(a, rays, cuda):
    d = cuda.DeviceArray.from_numpy(np.ascontiguousarray(rays, np.float32))
    h = a.trace_closest(d)
    occ = a.trace_any(d)
    cuda.synchronize()
    h = h.numpy()
    return h[:, 0].copy(), h[:, 1].view(np.float32).copy(), occ.numpy()


def _mesh_rays(fx):
    from pine_amd import workloads as W
    m = fx["mesh"]
    spec = W.MeshSpec(int(m[0]), float(fx["mesh_amp"]), int(m[1]), int(m[2]))
    v, i = W.make_mesh(spec)
    rs = fx["rayspec"]
    rspec = W.RaySpec(int(rs[0]), tuple(float(x) for x in fx["eye"]), float(fx["fov"]),
                      int(rs[1]), int(rs[2]))
    return v, i, W.make_rays(rspec, v, i)[fx["ray_index"]]


def check_vs_pine(prim, t, occ, gprim, gt, gocc, rays, prim_geom, oracle):
    """Bit-exact t and hit/miss against pine; every prim id that differs from
    pine's is a certified tie (tests/ties.py): the oracle finds pine's prim
    and this library's prim each hit at exactly that t, and the library's id is
    the smaller one."""
    hit = gprim >= 0
    assert np.array_equal(prim >= 0, hit), "hit/miss differs from pine"
    assert np.array_equal(t[hit].view(np.int32), gt[hit].view(np.int32)), "t not bit-exact"
    n = certify_ties(oracle, rays, prim, gprim, gt, prim_geom)
    assert np.array_equal(occ, gocc.astype(occ.dtype)), "any-hit differs from pine"
    return n


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("case", ["trace_5k", "trace_82k"])
def test_trace_matches_pine(pg, oracle, cuda, case, width):
    fx = golden(case)
    v, i, rays = _mesh_rays(fx)
    a = _accel(pg, v, i, width)
    prim, t, occ = _trace(a, rays, cuda)
    ties = check_vs_pine(prim, t, occ, fx["prim"], fx["t"], fx["occluded"], rays,
                         mesh_prim_geom(oracle, v, i), oracle)
    print(f"{case} W{width}: {ties} certified ties")
    # And exactly the oracle's lexicographic (t, id) minimum, ties included.
    bp, bt, _, _, bocc = oracle.trace_brute(oracle.tri9(v, i), rays)
    assert np.array_equal(prim, bp)
    assert np.array_equal(t.view(np.int32), bt.view(np.int32))


@pytest.mark.parametrize("width", WIDTHS)
def test_trace_1m_mesh_matches_pine(pg, oracle, cuda, width):
    fx = golden("trace_1m")
    v, i, rays = _mesh_rays(fx)
    a = _accel(pg, v, i, width)
    prim, t, occ = _trace(a, rays, cuda)
    ties = check_vs_pine(prim, t, occ, fx["prim"], fx["t"], fx["occluded"], rays,
                         mesh_prim_geom(oracle, v, i), oracle)
    print(f"trace_1m W{width}: {ties} certified ties")


def shapes_accel(pg, width):
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as mg
    shapes = mg.shapes_scene()
    a = pg.Accel()
    first = []
    for s in shapes:
        if s[0] == "sphere":
            first.append(a.add_sphere(s[1], s[2]))
        elif s[0] == "rect":
            first.append(a.add_rect(s[1], s[2], s[3]))
        elif s[0] == "plane":
            first.append(a.add_plane(s[1], s[2]))
        else:
            first.append(a.add_mesh(s[1], s[2]))
    a.build(width=width)
    return a, np.array(first), shapes


def shapes_prim_geom(oracle, a, shapes):
    """single_prim() of a global prim id of the shapes scene."""
    def geom(pid):
        s, l = a.prim_shape(pid)
        sh = shapes[s]
        if sh[0] == "mesh":
            return single_prim(tri=oracle.tri9(sh[1], sh[2][l: l + 1]))
        if sh[0] == "sphere":
            return single_prim(ana=[1, *sh[1], sh[2], 0, 0, 0, 0, 0])
        if sh[0] == "rect":
            return single_prim(ana=[2, *sh[1], *sh[2], *sh[3]])
        return single_prim(ana=[3, *sh[1], *sh[2], 0, 0, 0])
    return geom


@pytest.mark.parametrize("width", WIDTHS)
def test_analytic_shapes_match_pine(pg, oracle, cuda, width):
    fx = golden("shapes")
    a, first, shapes = shapes_accel(pg, width)
    prim, t, occ = _trace(a, fx["rays"], cuda)
    gprim = np.where(fx["shape"] >= 0, first[np.maximum(fx["shape"], 0)] + fx["prim"], -1)
    ties = check_vs_pine(prim, t, occ, gprim, fx["t"], fx["occluded"], fx["rays"],
                         shapes_prim_geom(oracle, a, shapes), oracle)
    print(f"shapes W{width}: {ties} certified ties")
    for p in np.unique(prim[prim >= 0])[:20]:
        s, l = a.prim_shape(int(p))
        assert first[s] + l == p


def test_uv_output_matches_pine(pg, cuda):
    fx = golden("trace_5k")
    v, i, rays = _mesh_rays(fx)
    a = _accel(pg, v, i, 2)
    d = cuda.DeviceArray.from_numpy(rays)
    uv = cuda.DeviceArray((len(rays), 2), np.float32)
    h = a.trace_closest(d, d_uv=uv).numpy()
    uv = uv.numpy()
    hit = (fx["prim"] >= 0) & (h[:, 0] == fx["prim"])
    assert np.array_equal(uv[hit, 0].view(np.int32), fx["u"][hit].view(np.int32))
    assert np.array_equal(uv[hit, 1].view(np.int32), fx["v"][hit].view(np.int32))


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("block", [64, 128, 256])
@pytest.mark.parametrize("flags", [0, 1, 2, 3, 4, 6])
def test_launch_shapes_agree(pg, oracle, cuda, width, block, flags):
    """Every block size, traversal loop (leaves postponed or not, stack or
    stackless with wave-wide node fetches) and XCD remap gives pine's hits;
    the any-hit and occlusion modes too."""
    fx = golden("trace_82k")
    v, i, rays = _mesh_rays(fx)
    a = _accel(pg, v, i, width, block, flags)
    prim, t, occ = _trace(a, rays, cuda)
    check_vs_pine(prim, t, occ, fx["prim"], fx["t"], fx["occluded"], rays,
                  mesh_prim_geom(oracle, v, i), oracle)
    d = cuda.DeviceArray.from_numpy(rays)
    occl = a.trace_occluded(d).numpy()
    cuda.synchronize()
    _, _, _, _, bocc = oracle.trace_brute(oracle.tri9(v, i), rays)
    # pg_trace_occluded uses pine's Intersect tests (open t interval).
    assert np.array_equal(occl, (prim >= 0).astype(occl.dtype))
    assert np.array_equal(occ, bocc)


def _tiled_82k(fx, reps=4, tail=777):
    """pine's 82k-triangle golden rays repeated, plus a ragged tail: above the
    65,536 rays from which PG_LAUNCH_REORDER sorts them."""
    v, i, rays = _mesh_rays(fx)
    def tile(x):
        return np.concatenate([x] * reps + [x[:tail]])
    big = tile(rays)
    assert len(big) >= 65536
    return v, i, big, tile(fx["prim"]), tile(fx["t"]), tile(fx["occluded"])


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("flags", [8, 9, 10, 12])
def test_reordered_launches_match_pine(pg, oracle, cuda, width, flags):
    """PG_LAUNCH_REORDER (alone, with postponed leaves, with the XCD remap,
    with the stackless traversal) sorts the rays on the device before tracing:
    every caller slot holds pine's hit, and uv, any-hit, occlusion and the
    per-ray work counts equal those of the unordered launch."""
    fx = golden("trace_82k")
    v, i, big, gprim, gt, gocc = _tiled_82k(fx)
    a = _accel(pg, v, i, width, 128, flags)
    b = _accel(pg, v, i, width, 128, flags & ~pg.LAUNCH_REORDER)
    prim, t, occ = _trace(a, big, cuda)
    check_vs_pine(prim, t, occ, gprim, gt, gocc, big, mesh_prim_geom(oracle, v, i), oracle)
    d = cuda.DeviceArray.from_numpy(big)
    uv_a = cuda.DeviceArray((len(big), 2), np.float32)
    uv_b = cuda.DeviceArray((len(big), 2), np.float32)
    ha = a.trace_closest(d, d_uv=uv_a).numpy()
    hb = b.trace_closest(d, d_uv=uv_b).numpy()
    assert np.array_equal(ha, hb)
    assert np.array_equal(uv_a.numpy().view(np.int32), uv_b.numpy().view(np.int32))
    assert np.array_equal(a.trace_occluded(d).numpy(), b.trace_occluded(d).numpy())
    assert np.array_equal(a.trace_any(d).numpy(), b.trace_any(d).numpy())
    ca, cb = a.trace_counts(d)[1].numpy(), b.trace_counts(d)[1].numpy()
    if not flags & pg.LAUNCH_POSTPONE_LEAVES:
        # Without speculative descent a ray's work does not depend on its wave.
        assert np.array_equal(ca, cb)
    else:
        assert ca.sum() > 0


def test_reordered_device_count_launch(pg, oracle, cuda):
    """pg_trace_closest_dev with reordering: only the first *d_n rays are
    sorted and traced (slots past the count keep their contents), and their
    hits equal the host-count launch's."""
    import ctypes
    fx = golden("trace_82k")
    v, i, big, _, _, _ = _tiled_82k(fx)
    a = _accel(pg, v, i, 2, 128, pg.LAUNCH_REORDER)
    d = cuda.DeviceArray.from_numpy(big)
    full = a.trace_closest(d).numpy()  # sizes the sort's buffers for len(big)
    for count in (len(big), 70001, 1000):
        hits = cuda.DeviceArray.from_numpy(np.full((len(big), 2), -7, np.int32))
        dn = cuda.DeviceArray.from_numpy(np.array([count], np.int32))
        pg.check(pg.lib().pg_trace_closest_dev(a.h, d.ptr, dn.ptr, ctypes.c_int64(len(big)),
                                               hits.ptr, None, None))
        h = hits.numpy()
        assert np.array_equal(h[:count], full[:count]), count
        assert (h[count:] == -7).all(), count


@pytest.mark.parametrize("width,flags", [(2, 0), (2, 1), (4, 0), (2, 4)])
def test_device_count_launch_strides_past_its_grid(pg, cuda, width, flags):
    """A device-count launch caps its grid at 2^22 lanes and strides over the
    count: with 5.2M rays (pine's golden rays repeated, ragged tail) some
    lanes trace two rays.  The first *d_n slots hold the host-count
    launch's hits and occlusion (one ray per lane; pine's t, bit for bit),
    for plain, postponed-leaf, 4-wide and stackless traversal, and the slots
    past the count are untouched."""
    import ctypes
    fx = golden("trace_82k")
    v, i, rays = _mesh_rays(fx)
    reps = -(-(5 << 20) // len(rays))
    def tile(x):
        return np.concatenate([x] * reps + [x[:333]])
    big, gprim, gt = tile(rays), tile(fx["prim"]), tile(fx["t"])
    assert len(big) > 5 << 20
    a = _accel(pg, v, i, width, 128, flags)
    d = cuda.DeviceArray.from_numpy(big)
    full, full_occ = a.trace_closest(d).numpy(), a.trace_occluded(d).numpy()
    hit = gprim >= 0
    assert np.array_equal(full[:, 0] >= 0, hit)
    assert np.array_equal(full[hit, 1], gt[hit].view(np.int32))
    for count in (len(big), (5 << 20) - 3, 1000, 0):
        dn = cuda.DeviceArray.from_numpy(np.array([count], np.int32))
        hits = cuda.DeviceArray.from_numpy(np.full((len(big), 2), -7, np.int32))
        pg.check(pg.lib().pg_trace_closest_dev(a.h, d.ptr, dn.ptr, ctypes.c_int64(len(big)),
                                               hits.ptr, None, None))
        occ = cuda.DeviceArray.from_numpy(np.full(len(big), -7, np.int32))
        pg.check(pg.lib().pg_trace_occluded_dev(a.h, d.ptr, dn.ptr, ctypes.c_int64(len(big)),
                                                occ.ptr, None))
        h, o = hits.numpy(), occ.numpy()
        assert np.array_equal(h[:count], full[:count]), count
        assert np.array_equal(o[:count], full_occ[:count]), count
        assert (h[count:] == -7).all() and (o[count:] == -7).all(), count


def test_timed_launch_reports_the_traversal_kernel(pg, cuda):
    """PG_LAUNCH_TIMED: last_trace_ms is -1 before any timed launch, then the
    traversal kernel's own duration (inside the whole reordered launch's),
    and timing changes no hit."""
    fx = golden("trace_82k")
    v, i, big, _, _, _ = _tiled_82k(fx)
    a = _accel(pg, v, i, 2, 128, pg.LAUNCH_REORDER)
    d = cuda.DeviceArray.from_numpy(big)
    plain = a.trace_closest(d).numpy()
    assert a.stats()["last_trace_ms"] == -1.0
    a.set_launch(128, pg.LAUNCH_REORDER | pg.LAUNCH_TIMED)
    s0, s1 = cuda.Event(), cuda.Event()
    s0.record()
    timed = a.trace_closest(d)
    s1.record()
    ms = a.stats()["last_trace_ms"]
    assert np.array_equal(timed.numpy(), plain)
    assert 0.0 < ms <= s0.elapsed_ms(s1) + 1e-3


def test_reordered_flag_in_a_recorded_graph(pg, cuda):
    """A device-count launch recorded into a HIP graph keeps the callers'
    order even with PG_LAUNCH_REORDER (the sort's library code would record
    memset nodes): every replay, with eager work and synchronisations between
    the replays, gives the eager launch's hits."""
    import ctypes
    fx = golden("trace_82k")
    v, i, big, _, _, _ = _tiled_82k(fx)
    a = _accel(pg, v, i, 2, 128, pg.LAUNCH_REORDER)
    d = cuda.DeviceArray.from_numpy(big)
    eager = a.trace_closest(d).numpy()
    R = cuda.rt()
    vp = ctypes.c_void_p
    R.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
    R.hipStreamBeginCapture.argtypes = [vp, ctypes.c_int]
    R.hipStreamEndCapture.argtypes = [vp, ctypes.POINTER(vp)]
    R.hipGraphInstantiate.argtypes = [ctypes.POINTER(vp), vp, vp, vp, ctypes.c_size_t]
    R.hipGraphLaunch.argtypes = [vp, vp]
    R.hipStreamSynchronize.argtypes = [vp]
    R.hipGraphExecDestroy.argtypes = [vp]
    R.hipGraphDestroy.argtypes = [vp]
    R.hipStreamDestroy.argtypes = [vp]
    st, g, ge = vp(), vp(), vp()
    cuda.check(R.hipStreamCreate(ctypes.byref(st)))
    dn = cuda.DeviceArray.from_numpy(np.array([len(big)], np.int32))
    hits = cuda.DeviceArray((len(big), 2), np.int32)
    cuda.check(R.hipStreamBeginCapture(st, 0))  # hipStreamCaptureModeGlobal
    rc = pg.lib().pg_trace_closest_dev(a.h, d.ptr, dn.ptr, ctypes.c_int64(len(big)), hits.ptr, None, st)
    cuda.check(R.hipStreamEndCapture(st, ctypes.byref(g)))
    pg.check(rc)
    cuda.check(R.hipGraphInstantiate(ctypes.byref(ge), g, None, None, 0))
    for _ in range(3):
        hits.zero()
        a.trace_closest(d)  # eager reordered work between replays
        cuda.synchronize()
        cuda.check(R.hipGraphLaunch(ge, st))
        cuda.check(R.hipStreamSynchronize(st))
        assert np.array_equal(hits.numpy(), eager)
    R.hipGraphExecDestroy(ge)
    R.hipGraphDestroy(g)
    R.hipStreamDestroy(st)


def test_edge_cases_vs_oracle(pg, oracle, cuda):
    rng = np.random.default_rng(5)
    # Two exact duplicates (prims 0 and 3) -> exact t ties resolved to id 0;
    # a degenerate (zero-area) triangle; an axis-aligned quad.
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0], [1, 0, 0], [0, 1, 0],
                  [2, 2, 2], [2, 2, 2], [3, 3, 3],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], np.float32)
    i = np.array([[0, 1, 2], [9, 10, 11], [9, 11, 12], [3, 4, 5], [6, 7, 8]], np.uint32)
    n = 4096
    o = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    o[:, 2] = rng.uniform(-3, -1, n)
    d = rng.normal(size=(n, 3))
    d[:, 2] = np.abs(d[:, 2]) + 0.5
    d[: n // 8, 0] = 0.0      # zero direction components (SafeRcp path)
    d[n // 8: n // 4, 1] = 0.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7] = o, d.astype(np.float32)
    rays[:, 7] = 3.4028235e38
    rays[n // 2: n // 2 + 256, 7] = 1.5           # short rays
    rays[n // 2 + 256: n // 2 + 512, 3] = 2.5      # tmin beyond the first plane
    # Rays straight down the z axis through the duplicated triangle's interior.
    rays[-64:, 0:3] = np.stack([rng.uniform(0.05, 0.4, 64), rng.uniform(0.05, 0.4, 64),
                                np.full(64, -1.0)], 1)
    rays[-64:, 4:7] = [0, 0, 1]
    for width in WIDTHS:
        a = _accel(pg, v, i, width)
        prim, t, occ = _trace(a, rays, cuda)
        bp, bt, _, _, bocc = oracle.trace_brute(oracle.tri9(v, i), rays)
        assert np.array_equal(prim, bp)
        assert np.array_equal(t.view(np.int32), bt.view(np.int32))
        assert np.array_equal(occ, bocc)
        assert (prim[-64:] == 0).all()   # tie between prims 0 and 3 -> 0


def test_single_triangle_and_empty_batch(pg, oracle, cuda):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    i = np.array([[0, 1, 2]], np.uint32)
    rays = np.array([[0.2, 0.2, -1, 0, 0, 0, 1, 3.4e38], [2, 2, -1, 0, 0, 0, 1, 3.4e38]],
                    np.float32)
    for width in WIDTHS:
        a = _accel(pg, v, i, width)
        prim, t, occ = _trace(a, rays, cuda)
        assert prim.tolist() == [0, -1] and t[0] == np.float32(1.0)
        assert occ.tolist() == [1, 0]
        empty = cuda.DeviceArray((0, 8), np.float32)
        assert a.trace_closest(empty).shape == (0, 2)


def test_full_benchmark_stream_is_tree_independent(pg, oracle, cuda):
    """All 16,777,216 benchmark rays: the closest hit does not depend on the
    hierarchy or the traversal, so the binary and 4-wide trees and the
    stackless traversal must agree bit for bit, and a sample must agree with
    the CPU port of pine's traversal."""
    from pine_amd import workloads as W
    v, i = W.make_mesh(W.ICOSPHERE_1M)
    rays = W.make_rays(W.RAYS_16M, v, i)
    d = cuda.DeviceArray.from_numpy(rays)
    outs = []
    configs = ((2, -1), (4, -1), (2, pg.LAUNCH_STACKLESS), (2, pg.LAUNCH_REORDER),
               (4, pg.LAUNCH_REORDER), (2, pg.LAUNCH_REORDER | pg.LAUNCH_STACKLESS))
    for width, flags in configs:
        a = _accel(pg, v, i, width, 128, flags)
        outs.append(a.trace_closest(d).numpy())
        a.release()
    for k in range(1, len(configs)):
        # other trees, the stackless traversal, rays sorted on the device:
        # the same hits bit for bit
        assert np.array_equal(outs[0], outs[k]), configs[k]
    hitfrac = (outs[0][:, 0] >= 0).mean()
    assert 0.05 < hitfrac < 0.9
    sel = np.arange(0, len(rays), 251)[:65536]
    cpu = oracle.CpuBVH(oracle.tri9(v, i))
    p, t, _ = cpu.trace(rays[sel])
    assert np.array_equal(t.view(np.int32), outs[0][sel, 1])
    # The port keeps pine's first-found prim on a tie: certified ties only.
    ties = certify_ties(oracle, rays[sel], outs[0][sel, 0], p, t, mesh_prim_geom(oracle, v, i))
    print(f"port vs GPU over {len(sel)} rays: {ties} certified ties")
