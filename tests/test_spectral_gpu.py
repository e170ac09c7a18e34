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


def test_spectral_needs_the_fused_pipeline(pg, cuda, tmp_path):
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as mg
    mg.cloud_volume(str(tmp_path))
    s = pg.Scene(text=golden("render_cloud")["scene"].item().decode(), directory=str(tmp_path))
    with pytest.raises(pg.PineError, match="spectral"):
        s.render(spp=1, spectral=True)
    s.release()
