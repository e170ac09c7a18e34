"""The float64 Atmosphere table of oracle/atmosphere.py (the definition of the
library's `accurate: true` table) on the CPU: texel directions, the NaN rule,
linearity in the sun colour, and the zenith case in closed form."""
import numpy as np

import oracle.atmosphere as A
from conftest import golden


def test_texel_directions_are_pines_uniform_sphere_map():
    d = A.texel_directions(64, 32)
    np.testing.assert_allclose(np.linalg.norm(d, axis=-1), 1.0, atol=1e-12)
    y, x = np.mgrid[0:32, 0:64]
    np.testing.assert_allclose(d[..., 2], 1.0 - 2.0 * y / 32, atol=1e-6)   # z = cos(acos(1 - 2v))
    phi = np.mod(np.arctan2(d[..., 1], d[..., 0]), 2 * np.pi)
    ok = np.abs(d[..., 2]) < 0.999
    np.testing.assert_allclose(phi[ok & (x > 0)], (2 * np.pi * x / 64)[ok & (x > 0)], atol=1e-6)


def test_table_is_finite_and_linear_in_the_sun_colour():
    t1, s1 = A.atmosphere_table([2, 6, 3], [1, 1, 1], 48, 24)
    t2, s2 = A.atmosphere_table([2, 6, 3], [2, 0.5, 3], 48, 24)
    assert np.isfinite(t1).all() and np.isfinite(s1).all()
    np.testing.assert_allclose(t2, t1 * [2, 0.5, 3], rtol=1e-12, atol=0)
    np.testing.assert_allclose(s2, s1 * [2, 0.5, 3], rtol=1e-12)
    # Views that cross the planet's interior overflow exp and are zeroed
    # (HasNaN -> 0); every view into the upper hemisphere of the model's y axis
    # stays in the atmosphere and is positive.
    d = A.texel_directions(48, 24)
    assert (t1[d[..., 1] > 0.05] > 0).all()
    assert (t1[d[..., 1] < -0.5] == 0).all()


def test_zenith_view_with_the_sun_overhead_in_closed_form():
    """View and sun straight up (the model's +y): the view ray leaves the 60 km
    shell after 60 km, samples sit at heights (i + 1/2) 7.5 km, and the light
    path term pine writes (a squared distance minus a radius, about 4e13 m) makes
    exp() of the light optical depth vanish.  Colour = 10 sum_i T_i (betaR hr_i
    phaseR + betaM hm_i phaseM), T_i = exp(-betaR ODR_i - betaM ODM_i), with the
    Mie scale height 10 km along the sun and mu = 1."""
    up = np.array([0.0, 1.0, 0.0])
    got = A.atmosphere_color(up[None], up, [1, 1, 1])[0]
    seg = 60e3 / 8
    h = (np.arange(8) + 0.5) * seg
    hr, hm = np.exp(-h / 7995) * seg, np.exp(-h / 10000) * seg
    odr, odm = np.cumsum(hr), np.cumsum(hm)
    g = 0.76
    phase_r = 3 / (16 * np.pi) * 2
    phase_m = 3 / (8 * np.pi) * (1 - g * g) * 2 / ((2 + g * g) * (1 - g) ** 3)
    T = np.exp(-np.outer(odr, A.BETA_R) - (A.BETA_M * odm)[:, None])
    want = 10 * (T * (np.outer(hr, A.BETA_R) * phase_r + (A.BETA_M * hm * phase_m)[:, None])).sum(0)
    np.testing.assert_allclose(got, want, rtol=1e-9)


def test_pines_own_table_has_no_finite_entry():
    """The reason for the accurate table: pine's Atmosphere constructor, run by
    the harness (tests/golden/atmosphere.npz), leaves no finite entry (about
    60% NaN, the rest +inf) and a NaN sun colour, for two different suns; the
    float64 definition is finite."""
    g = golden("atmosphere")
    for key in [k for k in g.files if k.endswith("_table")]:
        stem = key[: -len("_table")]
        t = g[key]
        assert not np.isfinite(t).any() and np.isnan(t).any() and np.isposinf(t).any()
        assert np.isnan(g[stem + "_sun"]).all()
        args = g[stem + "_args"].astype(np.float64)
        h, w, _ = g[key].shape
        t, s = A.atmosphere_table(args[:3], args[3:], w, h)
        assert np.isfinite(t).all() and np.isfinite(s).all()
