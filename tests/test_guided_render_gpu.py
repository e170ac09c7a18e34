"""Learned path guiding inside the GPU path tracer (pg_render_guided) on the
occluded-light box: a five-wall room lit only through a 0.3 x 0.3 gap in a
partition, behind which the lamp faces away from the room.

What is checked:
  - unbiasedness: the guided estimator converges to the plain tracer's image
    (block means over 8 x 8 pixels), whatever the guide learned;
  - the gain: with the trained guide the image error at equal samples per
    pixel is lower than the plain tracer's (the plain tracer is the judge:
    the reference is its own robust high-spp image, pine_amd.metrics);
  - what the network learned: at points of the room, the trained mixture's
    density towards the gap exceeds the cosine density several-fold.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def box(tmp_path_factory, pg):
    from pine_amd import workloads as W
    d = tmp_path_factory.mktemp("box")
    v, i = W.occluded_box_clutter()
    W.write_obj(str(d / "clutter.obj"), v, i)
    return str(d), W


def scene(pg, box, res, spp=64):
    d, W = box
    return pg.Scene(text=W.occluded_box_scene_text("clutter.obj", width=res, height=res, spp=spp,
                                                   max_depth=8), directory=d)


def blocks(a):
    return a.reshape(8, 12, 8, 12, 3).mean(axis=(1, 3))


def check_converges(img, ref, what):
    """img and ref, two 2048-spp renders of the 96 x 96 spheres scene, agree
    in their 8 x 8 block means: within 1% over the image, 3% block by block."""
    ok = np.isfinite(img).all(-1) & np.isfinite(ref).all(-1)
    assert ok.all()
    bg, br = blocks(img), blocks(ref)
    rel = np.abs(bg - br).mean() / br.mean()
    print(f"{what}: block mean rel. difference {rel:.4f}, image means {bg.mean():.5f} vs {br.mean():.5f}")
    assert abs(bg.mean() - br.mean()) < 0.01 * br.mean()
    assert rel < 0.03


@pytest.mark.parametrize("rr_depth", [0, 4])
def test_guided_estimator_is_unbiased(pg, cuda, tmp_path, rr_depth):
    """The spheres scene (diffuse spheres are guided, dielectric and conductor
    ones are not): the guided image converges to the plain tracer's, block by
    block, whatever the guide learned, with pine's Russian roulette depth and
    with a later one."""
    from conftest import golden
    from pine_amd import metrics as M
    text = golden("render_spheres")["scene"].item().decode()
    s = pg.Scene(text=text, directory=str(tmp_path))
    s.render(spp=2048)
    ref = M.linear(s.film())
    g = pg.Guide(seed=3, learning_rate=3e-3)
    inf = s.render_guided(g, train_passes=4, train_spp=4, final_spp=2048, rr_depth=rr_depth)
    print(f"guided {inf['last_render_ms']:.1f} ms")
    check_converges(M.linear(s.film()), ref, f"guided, rr_depth {rr_depth}")
    g.release()
    s.release()


def test_russian_roulette_depth(pg, cuda, tmp_path):
    """rr_depth 0 is pine's depth 2, sample for sample; a later depth changes
    the samples but not what they converge to; a negative one is refused."""
    from conftest import golden
    from pine_amd import metrics as M
    text = golden("render_spheres")["scene"].item().decode()
    s = pg.Scene(text=text, directory=str(tmp_path))
    s.render(spp=64, sample_offset=77)
    f0 = s.film()
    s.render(spp=64, sample_offset=77, rr_depth=2)
    assert np.array_equal(s.film(), f0)
    s.render(spp=64, sample_offset=77, rr_depth=5)
    assert not np.array_equal(s.film(), f0)
    s.render(spp=2048)
    ref = M.linear(s.film())
    s.render(spp=2048, rr_depth=5)
    check_converges(M.linear(s.film()), ref, "plain, rr_depth 5")
    with pytest.raises(pg.PineError):
        s.render(spp=4, rr_depth=-1)
    s.release()


def test_trained_guide_lowers_error_at_equal_spp(pg, cuda, box):
    from pine_amd import metrics as M
    s = scene(pg, box, 128)
    ref = M.robust_reference(s, 1024, 8)
    s.render(spp=128, sample_offset=1 << 20)
    plain = M.relmse(M.linear(s.film()), ref)
    g = pg.Guide(seed=3, learning_rate=3e-3)
    s.render_guided(g, final_spp=128)
    guided = M.relmse(M.linear(s.film()), ref)
    print(f"relMSE at 128 spp: plain {plain:.3e}, guided {guided:.3e} ({plain / guided:.2f}x)")
    assert guided < 0.9 * plain
    g.release()
    s.release()


def test_guide_learns_the_gap(pg, cuda, box):
    """From floor points of the room, the trained guide's probability of a
    direction through the 0.3 x 0.3 gap is several times the cosine lobe's
    (both integrated over the gap's solid angle by quadrature on its area)."""
    s = scene(pg, box, 128)
    g = pg.Guide(seed=3, learning_rate=3e-3)
    s.render_guided(g, final_spp=4)
    bmin, bmax = np.array([-1.0, 0.0, -1.0]), np.array([1.0, 2.0, 1.0])
    cam = np.array([0.0, 1.0, 3.3])
    n = np.array([0.0, 1.0, 0.0])
    # Gap quadrature: 8 x 8 cell centres of the opening (x in [-0.15, 0.15],
    # y in [0.85, 1.15], z = -0.4, normal +z).
    u = (np.arange(8) + 0.5) / 8
    gx, gy = np.meshgrid(-0.15 + 0.3 * u, 0.85 + 0.3 * u)
    gap = np.stack([gx.ravel(), gy.ravel(), np.full(64, -0.4)], 1)
    ratios = []
    pts = [np.array([x, 0.001, z]) for x in (-0.6, -0.3, 0.0, 0.3, 0.6) for z in (0.0, 0.3, 0.6)]
    xs, ws, jac = [], [], []
    for p in pts:
        wo = (cam - p) / np.linalg.norm(cam - p)
        d = gap - p
        dist = np.linalg.norm(d, axis=1)
        w = d / dist[:, None]
        xs.append(np.repeat(np.concatenate([(p - bmin) / (bmax - bmin), n, wo])[None], 64, 0))
        ws.append(w)
        jac.append(np.abs(w[:, 2]) / dist ** 2)  # d(solid angle) / d(area)
    x = np.concatenate(xs).astype(np.float32)
    w = np.concatenate(ws).astype(np.float32)
    cosine = (np.maximum(w @ n, 0) / np.pi).astype(np.float32)
    out = g.pdf(cuda.DeviceArray.from_numpy(x), cuda.DeviceArray.from_numpy(w),
                cuda.DeviceArray.from_numpy(cosine)).numpy()
    q = out[:, 0].reshape(len(pts), 64)
    c = cosine.reshape(len(pts), 64)
    j = np.stack(jac)
    ratios = (q * j).sum(1) / (c * j).sum(1)
    print("guide / cosine probability through the gap:", np.round(ratios, 2),
          "alpha:", np.round(out[::64, 1], 3))
    assert np.median(ratios) > 3.0
    assert (out[:, 1] > 0.2).all()
    g.release()
    s.release()


def test_kept_training_passes_are_the_same_samples(pg, cuda, box):
    """keep_passes: the training passes' samples stay in the final film.  Film
    weights depend only on each sample's film position (pine's filter table at
    the jittered point, from sampler dimensions 2 and 3), so a guided render
    keeping every pass holds exactly the weights of a plain render over the same
    contiguous sample range; and its image is an unbiased guided estimate, close
    to the one without the kept passes."""
    from pine_amd import metrics as M
    s = scene(pg, box, 128)
    passes, tspp, fspp, off = 4, 2, 16, 1000
    s.render(spp=passes * tspp + fspp, sample_offset=off)
    plain = s.film()
    g = pg.Guide(seed=3, learning_rate=3e-3)
    s.render_guided(g, train_passes=passes, train_spp=tspp, final_spp=fspp, sample_offset=off,
                    keep_passes=passes)
    kept = s.film()
    g.release()
    g = pg.Guide(seed=3, learning_rate=3e-3)
    s.render_guided(g, train_passes=passes, train_spp=tspp, final_spp=fspp, sample_offset=off)
    final_only = s.film()
    g.release()
    s.release()
    assert np.array_equal(kept[..., 3], plain[..., 3])
    assert not np.array_equal(final_only[..., 3], plain[..., 3])
    a, b = M.linear(kept), M.linear(final_only)
    ok = np.isfinite(a).all(-1) & np.isfinite(b).all(-1)
    assert ok.mean() > 0.999
    ma, mb = a[ok].mean(), b[ok].mean()
    print(f"image means: kept passes {ma:.5f}, final only {mb:.5f}")
    assert abs(ma - mb) < 0.1 * mb
