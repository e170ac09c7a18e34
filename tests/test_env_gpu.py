"""The Constant environment light (this library's extension for the
heterogeneous-volume workload: pine has no constant environment, so there is
no pine film to compare with).  Pinned by closed forms: a camera ray that
escapes sees exactly the environment's radiance, and a diffuse convex object
under a constant environment reflects albedo x radiance (its whole upper
hemisphere sees the environment, nothing it reflects comes back to it)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCENE = """Integrator: Path{
    maxDepth: 4
    sampler: Halton{
        samplesPerPixel: 512
    }
}
Camera: ThinLen{
    film{
        size: 48 48
        filter: LanczosSinc{
            radius: 0.5
        }
    }
    from: 0 0 5
    to: 0 0 0
    fov: 0.5
}
Material m: Layered{
    layer0: Diffuse{
        albedo: 0.6
    }
}
Shape: Sphere{
    position: 0 0 0
    radius: 1
    material: m
}
Light: Environment{
    environment: Constant{
        color: 0.5 0.7 0.9
    }
}
"""


def linear(film):
    w = film[..., 3:4]
    return np.where(w != 0, film[..., :3] / np.where(w != 0, w, 1), 0)


@pytest.mark.parametrize("spectral", [False, True])
def test_constant_environment_closed_forms(pg, cuda, spectral):
    s = pg.Scene(text=SCENE)
    s.render(spectral=spectral)
    img = linear(s.film())
    s.release()
    c = np.array([0.5, 0.7, 0.9], np.float32)
    # Pixel centres' distance from the image centre, in units of the sphere's
    # image radius (fov 0.5 rad at distance 5: tan(0.25) * 5 = 1.277 half-height;
    # the sphere's silhouette subtends asin(1/5)).
    h = 48
    yy, xx = np.mgrid[0:h, 0:h]
    r = np.hypot(xx + 0.5 - h / 2, yy + 0.5 - h / 2) / (h / 2)
    r_sphere = np.tan(np.arcsin(0.2)) / np.tan(0.25)
    inner, outer = r < 0.8 * r_sphere, r > 1.2 * r_sphere
    bg = img[outer].mean(0)
    fg = img[inner].mean(0)
    print(f"spectral={spectral}: background {bg}, sphere {fg} (expected {c} and {0.6 * c})")
    if not spectral:
        # Camera rays that escape return exactly the environment radiance (up to
        # the rounding of the film's weighted sums over 512 samples).
        assert np.allclose(img[outer], c, rtol=1e-5, atol=0)
        np.testing.assert_allclose(fg, 0.6 * c, rtol=0.01)
    else:
        # Spectral transport reproduces RGB radiance in expectation only.
        np.testing.assert_allclose(bg, c, rtol=0.03)
        np.testing.assert_allclose(fg, 0.6 * c, rtol=0.03)
