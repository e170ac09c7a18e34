"""Image-error helpers used to compare plain and guided rendering
(pine_amd.metrics): the reference estimate and the trimmed relMSE."""
import numpy as np

from pine_amd import metrics as M


def test_reference_is_the_chunk_mean_without_singular_chunks():
    rng = np.random.default_rng(0)
    imgs = rng.uniform(0.5, 1.5, (16, 4, 4, 3)).astype(np.float32)
    imgs[3, 1, 1] = 1e9       # a singular event
    imgs[5, 2, 2] = np.nan    # a non-finite sample range
    imgs[:, 3, 3] = 0.0
    imgs[7, 3, 3] = 5.0       # median zero: every range counts
    r = M.reference_from_chunks(imgs)
    np.testing.assert_allclose(r[0, 0], imgs[:, 0, 0].mean(0), rtol=1e-6)
    np.testing.assert_allclose(r[1, 1], np.delete(imgs[:, 1, 1], 3, 0).mean(0), rtol=1e-6)
    np.testing.assert_allclose(r[2, 2], np.delete(imgs[:, 2, 2], 5, 0).mean(0), rtol=1e-6)
    np.testing.assert_allclose(r[3, 3], 5.0 / 16, rtol=1e-6)
    # A bright but not singular range (rare light path) is kept: no low bias.
    imgs2 = np.full((16, 1, 1, 3), 0.01, np.float32)
    imgs2[0] = 2.0
    assert abs(M.reference_from_chunks(imgs2)[0, 0, 0] - (2.0 + 15 * 0.01) / 16) < 1e-7


def test_relmse_trims_the_worst_pixels_and_counts_non_finite_as_worst():
    ref = np.ones((10, 100, 3), np.float32)
    img = ref.copy()
    img[0, 0] = np.nan
    img[1, 1] = 3.0
    assert M.relmse(img, ref, trim=2e-3) == 0.0
    e = M.relmse(img, ref, trim=1e-3)  # drops only the non-finite pixel
    assert abs(e - (4.0 / 1.01) / 999) < 1e-9
    assert abs(M.time_to_error([10, 20, 40], [4.0, 2.0, 1.0], 2.0) - 20.0) < 1e-9
    assert abs(M.time_to_error([10, 20, 40], [4.0, 2.0, 1.0], np.sqrt(2.0)) - 20 * np.sqrt(2.0)) < 1e-9
    assert M.time_to_error([10, 20], [4.0, 2.0], 1.0) is None


def test_untrimmed_relmse_counts_nonfinite_pixels_apart():
    from pine_amd import metrics as M
    ref = np.ones((4, 4, 3), np.float32)
    img = ref.copy()
    img[0, 0] = 2.0                      # error 1 / 1.01 on one pixel of 15 finite ones
    img[3, 3, 1] = np.nan
    e, bad = M.relmse_untrimmed(img, ref)
    assert bad == 1
    assert abs(e - (1 / 1.01) / 15) < 1e-7
