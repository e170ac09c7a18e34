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


def test_batch_paths_cap_matches_the_same_split_alone(pg, cuda, scenes):
    """paths_per_batch caps each scene's batch: a scene then renders in
    batches of floor(cap / pixels) samples per pixel, as it does alone with
    that samples_per_batch (same film bit for bit, same rays); one sample per
    pixel is the floor."""
    texts, d = scenes
    ss = [pg.Scene(text=t, directory=d) for t in texts]
    cap = 3 * max(s.info()["width"] * s.info()["height"] for s in ss) // 2
    pg.render_batch(ss, streams=2, paths_per_batch=cap)
    got = [(s.film(), s.info()["last_rays"]) for s in ss]
    for s, (film, rays) in zip(ss, got):
        inf = s.info()
        spb = max(1, cap // (inf["width"] * inf["height"]))
        s.render(samples_per_batch=spb)
        np.testing.assert_array_equal(s.film(), film)
        assert s.info()["last_rays"] == rays
    with pytest.raises(pg.PineError):
        pg.render_batch(ss, paths_per_batch=-1)
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


def test_batch_workload_scenes_match_pine(pg, cuda, tmp_path):
    """Five scenes of the 64-scene batched workload (icosphere, 16 spheres,
    instanced boxes), at reduced film, spp and mesh size, rendered together by
    pg_render_batch: each film matches pine's film of the same scene text and
    equals, bit for bit, the scene rendered alone."""
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as mg
    from test_render_gpu import check_film
    fx = golden("render_batch")
    items = mg.batch_golden_items()
    texts = []
    for it in items:
        mg.write_batch_files(str(tmp_path), it)
        texts.append(fx[f"b{it['index']}_scene"].item().decode())
    scenes = [pg.Scene(text=t, directory=str(tmp_path)) for t in texts]
    res = pg.render_batch(scenes, streams=3)
    films = [s.film() for s in scenes]
    rays = [s.info()["last_rays"] for s in scenes]
    for s in scenes:
        s.release()
    assert res["rays"] == sum(rays)
    for it, t, film, r in zip(items, texts, films, rays):
        # measured 1.6e-10 .. 9.1e-7, 0.54 .. 0.56 of the pixels bit-exact
        e = check_film(film, fx[f"b{it['index']}_film"], 1e-5, 0.45)
        print(f"batch scene {it['index']} ({it['triangles']} triangles, {len(it['boxes'])} boxes): "
              f"relMSE {e:.3e} vs pine")
        s = pg.Scene(text=t, directory=str(tmp_path))
        s.render()
        assert np.array_equal(s.film().view(np.int32), film.view(np.int32))
        assert s.info()["last_rays"] == r
        s.release()


def test_full_64_scene_batch_equals_each_scene_alone(pg, cuda, tmp_path):
    """The whole batched workload at full size (64 scenes, 22.6M icosphere
    triangles plus instanced boxes, 256^2..1024^2, 16..128 spp): one
    pg_render_batch gives every scene the film it gets rendered alone, bit
    for bit, with the same ray counts."""
    import hashlib
    from pine_amd import workloads as W
    items = W.batch_workload()
    W.write_box_obj(str(tmp_path / W.BOX_OBJ))
    scenes = []
    for it in items:
        v, i = W.make_mesh(it["mesh"])
        name = f"batch{it['index']}.obj"
        pg.write_obj_fixed(str(tmp_path / name), v, i)
        scenes.append(pg.Scene(text=W.batch_scene_text(it, name), directory=str(tmp_path)))
        os.remove(str(tmp_path / name))
    print(f"{len(scenes)} scenes loaded", flush=True)
    alone = []
    for s, it in zip(scenes, items):
        # The batch caps a scene's batch at 64M paths by default: alone, the
        # scene is rendered with the same split of its samples.
        s.render(samples_per_batch=max(1, min(it["spp"], (64 << 20) // (it["width"] * it["height"]))))
        alone.append((hashlib.sha256(s.film().tobytes()).hexdigest(), s.info()["last_rays"]))
    res = pg.render_batch(scenes, streams=4)
    for k, s in enumerate(scenes):
        assert hashlib.sha256(s.film().tobytes()).hexdigest() == alone[k][0], f"scene {k} differs"
        assert s.info()["last_rays"] == alone[k][1]
    assert res["rays"] == sum(r for _, r in alone)
    assert res["samples"] == sum(it["width"] * it["height"] * it["spp"] for it in items)
    print(f"64-scene batch: {res['ms']:.1f} ms, {res['samples']} samples, {res['rays']} rays")
    for s in scenes:
        s.release()


def test_film_border_samples_are_deterministic(pg, cuda, tmp_path):
    """Samples on a pixel border also feed the neighbouring pixel, after the
    batch's own-pixel pass; those additions run in a fixed order, so a batch of
    58.7M paths (1024^2 x 56 spp, one batch) gives the same film bit for bit on
    every render, on any path-state slot, alone or batched.  The same render in
    4-spp batches adds the border samples between batches instead, which
    changes the rounding only of the pixels that receive them."""
    import hashlib
    from pine_amd import workloads as W
    it = W.batch_workload()[63]
    W.write_box_obj(str(tmp_path / W.BOX_OBJ))
    v, i = W.make_mesh(it["mesh"])
    pg.write_obj_fixed(str(tmp_path / "b63.obj"), v, i)
    s = pg.Scene(text=W.batch_scene_text(it, "b63.obj"), directory=str(tmp_path))
    films = []
    for _ in range(3):
        s.render()
        films.append(s.film().copy())
    assert s.info()["paths_in_flight"] >= it["width"] * it["height"] * it["spp"]
    pg.render_batch([s], streams=1)
    films.append(s.film().copy())
    pg.render_batch([s], streams=2)
    films.append(s.film().copy())
    digests = {hashlib.sha256(f.tobytes()).hexdigest() for f in films}
    assert len(digests) == 1
    s.render(samples_per_batch=4)
    small = s.film()
    differ = np.any(small != films[0], axis=-1)
    assert 0 < differ.sum() < 0.01 * differ.size
    np.testing.assert_allclose(small, films[0], rtol=1e-5, atol=1e-6)
    s.release()
