"""The samplers on the GPU against pine's: the device Halton and Sobol samplers
(pg_scene_sampler_values) bit for bit against pine's HaltonSampler and
SobolSampler values (tests/golden/halton.npz, sobol.npz, written by the oracle
harness from pine's own sampler objects), and Path renders under the Sobol
sampler against pine's renders of the same scene text."""
import numpy as np
import pytest

from conftest import golden
from test_oracle import _sampler_key
from test_render_gpu import check_film, render_text

pytestmark = pytest.mark.gpu

SOBOL_STRATEGY = {0: "none", 1: "binaryPermutate", 2: "fastOwen", 3: "Owen"}


def sampler_scene(pg, w, h, spp, sampler, randomize=None):
    from pine_amd import workloads as W
    text = W.spheres_scene_text(seed=0, width=w, height=h, spp=spp, n_spheres=2, sampler=sampler,
                                randomize=randomize)
    return pg.Scene(text=text)


def test_device_halton_matches_pine(pg, cuda):
    g = golden("halton")
    for key in g.files:
        w, h, spp, nd, npix, _ = _sampler_key(key)
        s = sampler_scene(pg, w, h, spp, "Halton")
        got = s.sampler_values(3, 1, npix, spp, nd)
        s.release()
        assert np.array_equal(got.view(np.int32), g[key].view(np.int32)), key


def test_device_sobol_matches_pine(pg, cuda):
    g = golden("sobol")
    for key in g.files:
        w, h, spp, nd, npix, strategy = _sampler_key(key)
        s = sampler_scene(pg, w, h, spp, "Sobol", SOBOL_STRATEGY[strategy])
        got = s.sampler_values(3, 1, npix, spp, nd)
        s.release()
        bad = int((got.view(np.int32) != g[key].view(np.int32)).sum())
        assert bad == 0, f"{key}: {bad} of {got.size} values differ"


def test_device_sobol_pixels_tile_the_sequence(pg, cuda):
    """Sample k of the 4096 pixels of a 64x64 film (m = 6) uses Sobol indices
    4096k..4096k+4095: each dimension has one value per 1/4096 stratum.  (A
    scrambled 32-bit value just below a stratum's end can round up onto it when
    converted to float, so the j-th smallest value is checked against the closed
    stratum [j/n, (j+1)/n].)"""
    n = 64 * 64
    for strategy in (0, 2):
        s = sampler_scene(pg, 64, 64, 2, "Sobol", SOBOL_STRATEGY[strategy])
        vals = s.sampler_values(0, 0, n, 2, 16)
        s.release()
        assert len({(977 * k) % n for k in range(n)}) == n
        for sample in (0, 1):
            for d in range(16):
                v = np.sort(vals[:, sample, d].astype(np.float64)) * n
                j = np.arange(n)
                assert ((v >= j) & (v <= j + 1)).all(), (strategy, sample, d)


def test_halton_without_permutation_is_the_radical_inverse(pg, cuda):
    """randomizeStrategy none: dimension 2 (base 5) of Halton index i is the plain
    base-5 radical inverse of i, which a 1x1 film (stride 1) exposes directly."""
    s = sampler_scene(pg, 1, 1, 50, "Halton", "none")
    vals = s.sampler_values(0, 0, 1, 50, 1)[0, :, 0]
    s.release()

    def radical_inverse(i, b):
        r, f = 0.0, 1.0 / b
        while i:
            r += (i % b) * f
            i //= b
            f /= b
        return r

    ref = np.array([radical_inverse(i, 5) for i in range(50)])
    np.testing.assert_allclose(vals, ref, rtol=0, atol=1e-6)


@pytest.mark.parametrize("name", ["fastowen", "owen"])
def test_sobol_render_matches_pine(pg, cuda, tmp_path, name):
    fx = golden("render_sobol")
    film, info = render_text(pg, fx[name + "_scene"].item().decode(), str(tmp_path))
    # measured 8.4e-6 (fastOwen) and 2.4e-5 (Owen), 0.43 exact
    e = check_film(film, fx[name + "_film"], 2.5e-4, 0.35)
    print(f"Sobol {name} relMSE {e:.3e}")
