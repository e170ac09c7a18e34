"""Spectral mode on the GPU (PG_RENDER_SPECTRAL): four wavelengths per path.

A scene whose reflectances, emission and sky are gray has flat spectra, and
the spectral estimator then has the RGB image as its expectation: block means
must agree within the two renders' noise.  A coloured scene differs by design
(products of spectra are not products of RGB triples) but must stay close.
Spectral renders are deterministic."""
import re

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def linear(film):
    w = film[..., 3:4]
    return np.where(w != 0, film[..., :3] / np.where(w != 0, w, 1), 0)


def grayed(text):
    def mean3(m):
        v = [float(x) for x in m.group(2).split()]
        g = sum(v) / len(v)
        return f"{m.group(1)}{g:.6f} {g:.6f} {g:.6f}" if len(v) == 3 else m.group(0)
    return re.sub(r"((?:albedo|color): )([-0-9. ]+)", mean3, text)


def blocks(a, b=12):
    h, w = a.shape[0] // b * b, a.shape[1] // b * b
    return a[:h, :w].reshape(h // b, b, w // b, b, 3).mean(axis=(1, 3))


def test_gray_scene_spectral_equals_rgb(pg, cuda, tmp_path):
    text = grayed(golden("render_spheres")["scene"].item().decode())
    assert "albedo" in text
    s = pg.Scene(text=text, directory=str(tmp_path))
    s.render(spp=2048)
    rgb = linear(s.film())
    s.render(spp=2048, spectral=True)
    spec = linear(s.film())
    s.render(spp=2048, spectral=True)
    assert np.array_equal(spec, linear(s.film()))
    br, bs = blocks(rgb), blocks(spec)
    rel = np.abs(bs - br).mean() / br.mean()
    print(f"gray scene: image means rgb {rgb.mean():.5f} spectral {spec.mean():.5f}, block rel. diff {rel:.4f}")
    assert abs(spec.mean() - rgb.mean()) < 0.01 * rgb.mean()
    assert rel < 0.03
    s.release()


def test_coloured_scene_stays_close(pg, cuda, tmp_path):
    text = golden("render_spheres")["scene"].item().decode()
    s = pg.Scene(text=text, directory=str(tmp_path))
    s.render(spp=1024)
    rgb = linear(s.film())
    s.render(spp=1024, spectral=True)
    spec = linear(s.film())
    means = rgb.reshape(-1, 3).mean(0), spec.reshape(-1, 3).mean(0)
    print("coloured scene channel means rgb", np.round(means[0], 5), "spectral", np.round(means[1], 5))
    assert np.isfinite(spec).all()
    np.testing.assert_allclose(means[1], means[0], rtol=0.15)
    s.release()


@pytest.mark.parametrize("view", ["through", "shadow"])
def test_spectral_media_closed_form(pg, cuda, tmp_path, view):
    """Spectral transport in the split bounce (participating media): the
    constant-density slab scenes are gray, so four-wavelength transport must
    give the closed-form transmittance exp(-sigma d L) in expectation, as the
    RGB tracer does (tests/test_medium_gpu.py), with wavelength noise on top."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as mg
    from conftest import golden
    from medium_closed_form import expected_slab
    mg.slab_volume(str(tmp_path))
    text = golden(f"slab_{view}")["scene"].item().decode()
    out = {}
    for spectral in (False, True):
        s = pg.Scene(text=text, directory=str(tmp_path))
        s.render(spectral=spectral)
        f = s.film()
        w = f[..., 3:4]
        out[spectral] = (np.where(w != 0, f[..., :3] / np.where(w != 0, w, 1), 0)).mean(-1)
        assert s.info()["general_mode"] == 1
        s.release()
    img = out[True]
    exp, mask, scale = expected_slab(view, img)
    d = img[mask] - exp[mask]
    drgb = out[True][mask] - out[False][mask]
    print(f"{view} spectral: mean diff {d.mean():.2e}, max |d| {np.abs(d).max():.3f}, "
          f"mean vs RGB {drgb.mean():.2e}")
    assert abs(d.mean()) < 3e-3
    assert abs(drgb.mean()) < 3e-3


def test_spectral_cloud_matches_rgb_in_expectation(pg, cuda, tmp_path):
    """The cloud scene (gray medium, brick tracking, camera in the medium,
    material-less rect) rendered with four wavelengths per path agrees with the
    RGB render block by block (8x8-pixel means); the scene's colours are
    near-gray, so upsampling changes the expectation only slightly."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as mg
    from pine_amd import workloads as W
    mg.cloud_volume(str(tmp_path))
    text = W.cloud_scene_text("cloud32.vol", width=64, height=64, spp=512, max_depth=16, bricks=8,
                              passthrough=True)
    imgs = []
    for spectral in (False, True):
        s = pg.Scene(text=text, directory=str(tmp_path))
        s.render(spectral=spectral)
        f = s.film()
        w = f[..., 3:4]
        imgs.append(np.where(w != 0, f[..., :3] / np.where(w != 0, w, 1), 0))
        s.release()
    blocks = [a.reshape(8, 8, 8, 8, 3).mean(axis=(1, 3)) for a in imgs]
    rel = np.abs(blocks[1] - blocks[0]) / (blocks[0] + 0.05)
    print(f"spectral vs RGB cloud: image means {imgs[0].mean():.5f} / {imgs[1].mean():.5f}, "
          f"block rel. diff max {rel.max():.3f}, mean {rel.mean():.4f}")
    assert np.isfinite(imgs[1]).all()
    np.testing.assert_allclose(imgs[1].mean(), imgs[0].mean(), rtol=0.03)
    assert rel.mean() < 0.03
