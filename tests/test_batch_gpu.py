"""The batched multi-scene path: the fused bounce loop replayed as a recorded
HIP graph, pg_render_batch over scenes of uneven size on concurrent streams,
and the launch autotuner.  None of them may change an image: the graph replay
runs the same kernels on the same data as the launch-by-launch loop, the batch
only interleaves independent scenes, and every tuned parameter (traversal
workgroup, leaf postponing, node width, paths per batch) leaves the hits and
the per-pixel sample order as they were."""
import os
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def linear(film):
    w = film[..., 3:4]
    return np.where(w != 0, film[..., :3] / np.where(w != 0, w, 1), 0)


def same_image(a, b):
    np.testing.assert_array_equal(a[..., 3], b[..., 3])
    la, lb = linear(a), linear(b)
    ok = np.isfinite(la).all(-1) & np.isfinite(lb).all(-1)
    assert np.array_equal(ok, np.isfinite(la).all(-1))
    e = float(np.mean((la[ok] - lb[ok]) ** 2 / (lb[ok] ** 2 + 1e-2)))
    assert e < 1e-10, e


@pytest.fixture(scope="module")
def scenes(tmp_path_factory, pg):
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as mg
    from pine_amd import workloads as W
    d = tmp_path_factory.mktemp("batch")
    v, i = mg.mixed_mesh()
    W.write_obj(str(d / "blob.obj"), v, i)
    mg.cloud_volume(str(d))
    texts = [golden("render_spheres")["scene"].item().decode(),
             golden("render_mixed")["scene"].item().decode(),
             golden("render_cloud")["scene"].item().decode(),  # media: the split bounce
             W.spheres_scene_text(seed=5, width=200, height=120, spp=24, max_depth=6)]
    return [t for t in texts], str(d)


def test_graph_replay_matches_the_launch_loop(pg, cuda, scenes):
    texts, d = scenes
    s = pg.Scene(text=texts[0], directory=d)
    s.render(spp=64, samples_per_batch=16, graph=False)
    ref, rays_ref = s.film(), s.info()["last_rays"]
    s.render(spp=64, samples_per_batch=16)
    inf = s.info()
    same_image(s.film(), ref)
    assert inf["last_rays"] == rays_ref
    caps = inf["graph_captures"]
    assert caps >= 1
    s.render(spp=64, samples_per_batch=16)  # same state: the recorded loop is reused
    assert s.info()["graph_captures"] == caps
    same_image(s.film(), ref)
    print(f"graph replay: {inf['last_render_ms']:.2f} ms, {rays_ref} rays, {caps} capture(s)")
    s.release()


def test_batch_renders_each_scene_as_alone(pg, cuda, scenes):
    texts, d = scenes
    alone = []
    for t in texts:
        s = pg.Scene(text=t, directory=d)
        s.render()
        alone.append((s.film(), s.info()["last_rays"]))
        s.release()
    ss = [pg.Scene(text=t, directory=d) for t in texts]
    res = pg.render_batch(ss, streams=3)
    for s, (film, rays) in zip(ss, alone):
        same_image(s.film(), film)
        assert s.info()["last_rays"] == rays
    inf = [s.info() for s in ss]
    assert res["samples"] == sum(i["width"] * i["height"] * i["spp"] for i in inf)
    assert res["rays"] == sum(r for _, r in alone)
    print(f"batch of {len(ss)}: {res['ms']:.2f} ms, {res['samples']} samples, {res['rays']} rays")
    # Per-scene spp override.
    res2 = pg.render_batch(ss, spp=[4, 8, 2, 6], streams=2)
    assert res2["samples"] == sum(i["width"] * i["height"] * k for i, k in zip(inf, [4, 8, 2, 6]))
    for s in ss:
        s.release()


def test_autotune_keeps_the_image(pg, cuda, scenes):
    texts, d = scenes
    s = pg.Scene(text=texts[1], directory=d)
    s.render(spp=32)
    ref = s.film()
    r = s.autotune(probe_spp=2, try_width=True)
    print("autotune:", r)
    inf = s.info()
    assert r["trace_block"] in (64, 128, 256) and r["bvh_width"] in (2, 4)
    assert (inf["trace_block"], inf["trace_flags"], inf["bvh_width"], inf["paths_in_flight"]) == \
        (r["trace_block"], r["trace_flags"], r["bvh_width"], r["paths_in_flight"])
    assert r["best_ms"] <= r["default_ms"] and r["probes"] >= 8
    s.render(spp=32)
    same_image(s.film(), ref)
    s.release()
