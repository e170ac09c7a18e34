"""Pine's Atmosphere environment light on the GPU against pine's renders of the
same scene text (tests/golden/render_atmosphere.npz).

Pine's own colour table has no finite entry (NaN or +inf: its approximate exp diverges),
so pine's film is NaN wherever the environment contributes; the GPU reproduces
that with pine's arithmetic.  With `accurate: true` the GPU computes the table
in double precision with correctly rounded functions; the goldens for those
scenes are pine's PathIntegrator with pine's Atmosphere light fed the float64
table of oracle/atmosphere.py, so Sample (half uniform sphere with the nearest
texel, half the sun as a delta light), Color (bilinear) and Pdf (1/4pi in the
BSDF-side MIS weight) are pine's own."""
import numpy as np
import pytest

from conftest import golden
from test_render_gpu import check_film, linear, rel_mse, render_text

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["accurate_interp", "accurate_area"])
def test_atmosphere_matches_pine_with_the_accurate_table(pg, cuda, tmp_path, name):
    fx = golden("render_atmosphere")
    film, info = render_text(pg, fx[name + "_scene"].item().decode(), str(tmp_path))
    # measured 2.5e-4 (accurate_interp: pine's atan2 near the poles) and 6.5e-6
    e = check_film(film, fx[name + "_film"], {"accurate_interp": 1e-3}.get(name, 6e-5))
    print(f"Atmosphere {name}: relMSE {e:.3e}")


def test_atmosphere_with_pines_table_is_nan_where_pines_is(pg, cuda, tmp_path):
    fx = golden("render_atmosphere")
    film, _ = render_text(pg, fx["pine_table_scene"].item().decode(), str(tmp_path))
    ref = fx["pine_table_film"]
    assert np.array_equal(np.isnan(film[..., :3]), np.isnan(ref[..., :3]))
    np.testing.assert_allclose(film[..., 3], ref[..., 3], rtol=1e-5, atol=1e-6)


def test_nearest_texel_lookup_converges_to_the_bilinear_one(pg, cuda, tmp_path):
    """interpolate: false (pine's nearest-texel Color, which pine's own build
    cannot render on these scenes): finite, and its difference from the
    bilinear lookup is the floor texel's offset, so it shrinks as the table
    refines.  It concentrates at the table's horizon edge (a step), whose share
    of directions scales with the texel size, so 4x finer cuts the relMSE about
    4x (measured 3.4e-2 -> 8.0e-3); the bar is 2x.  Both modes trace the same
    paths; only Color differs."""
    fx = golden("render_atmosphere")
    base = fx["accurate_interp_scene"].item().decode()
    errs = []
    for size in ("256 128", "1024 512"):
        text = base.replace("size: 128 64", "size: " + size)
        a, _ = render_text(pg, text, str(tmp_path))
        b, _ = render_text(pg, text.replace("interpolate: true", "interpolate: false"), str(tmp_path))
        assert np.isfinite(a).all() and np.isfinite(b).all()
        errs.append(rel_mse(linear(b), linear(a)))
    print(f"nearest vs bilinear relMSE: {errs[0]:.3e} (256x128), {errs[1]:.3e} (1024x512)")
    assert 0 < errs[1] < 0.5 * errs[0]
