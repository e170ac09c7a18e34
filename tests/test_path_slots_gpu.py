"""Per-path buffers shared by the scenes of a device (path-state slots): one
set serves renders one at a time, one set each stream of a batched render;
each grows on demand and is freed with the device's last scene.  Nothing a
render leaves in a slot may change a later render, whatever kind of scene
(fused bounce, split bounce with media, node-graph materials, spectral,
guided) used the slot before."""
import ctypes
import gc
import os
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _texts(tmp_path):
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as mg
    from pine_amd import workloads as W
    v, i = mg.mixed_mesh()
    W.write_obj(str(tmp_path / "blob.obj"), v, i)
    mg.cloud_volume(str(tmp_path))
    texts = {
        "spheres": golden("render_spheres")["scene"],
        "mixed": golden("render_mixed")["scene"],
        "cloud": golden("render_cloud")["scene"],
        "nodes": golden("render_nodes")["scene"],
        "fog": golden("render_nodes_media")["fog_scene"],
    }
    return {k: t.item().decode() for k, t in texts.items()}


def test_slot_reuse_across_scene_kinds(pg, cuda, tmp_path):
    """Each scene rendered alone (one scene alive at a time), then all of them
    alive together and rendered in turn on the same slot, with spectral and
    guided renders (their own buffer kinds) in between: every film is the
    first one bit for bit."""
    texts = _texts(tmp_path)
    d = str(tmp_path)
    first = {}
    for k, t in texts.items():
        s = pg.Scene(text=t, directory=d)
        s.render()
        first[k] = s.film().copy()
        s.release()
    scenes = {k: pg.Scene(text=t, directory=d) for k, t in texts.items()}
    for k in ["cloud", "spheres", "nodes", "mixed", "fog", "spheres", "cloud", "mixed"]:
        s = scenes[k]
        s.render()
        assert np.array_equal(s.film(), first[k]), k
        if k == "spheres":
            s.render(spectral=True)
        if k == "nodes":
            g = pg.Guide(seed=3, learning_rate=3e-3)
            s.render_guided(g, train_passes=2, train_spp=1, final_spp=2)
            g.release()
    res = pg.render_batch(list(scenes.values()), streams=2)
    assert res["samples"] > 0
    for k, s in scenes.items():
        assert np.array_equal(s.film(), first[k]), f"{k} in the batch"
    for s in scenes.values():
        s.release()


def _free_bytes(pg):
    hip = pg.hip.rt()
    hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    pg.hip.check(hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value


def test_slots_are_freed_with_the_last_scene(pg, cuda):
    """A 1024^2 x 64-spp render holds its 67M paths' buffers (about 14 GB)
    after it returns, for the next render; releasing the device's last scene
    gives them back."""
    from pine_amd import workloads as W
    gc.collect()
    before = _free_bytes(pg)
    s = pg.Scene(text=W.spheres_scene_text(seed=0, width=1024, height=1024, spp=64, max_depth=8))
    s.render()
    assert s.info()["last_rays"] > 0
    held = before - _free_bytes(pg)
    assert held > 8 << 30, held
    s.release()
    back = _free_bytes(pg) - (before - held)
    assert back > 0.9 * held, (back, held)
