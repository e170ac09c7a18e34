"""Spectral mode's colour math: the oracle (oracle/spectral.py) against closed
forms, and the library's tables (pg_spectral_tables, host code) against the
oracle.  No GPU."""
import numpy as np
import pytest

from oracle import spectral as S


def test_cmf_fit_reproduces_cie_1931_integrals_and_white_point():
    lam = S.grid()
    c = S.cmf(lam)
    # CIE 1931 2-degree tables sum to 106.857 (ybar) over 360..830 nm at 1 nm,
    # and the equal-energy spectrum has chromaticity (1/3, 1/3).
    integ = c.sum(0)
    np.testing.assert_allclose(integ, 106.857, rtol=3e-3)
    xy = integ[:2] / integ.sum()
    np.testing.assert_allclose(xy, [1 / 3, 1 / 3], atol=3e-3)
    # ybar peaks at 555 nm with value 1 (photopic maximum).
    assert 550 <= lam[np.argmax(c[:, 1])] <= 562 and abs(c[:, 1].max() - 1.0) < 0.03
    assert (c >= -0.02).all()


def test_basis_is_a_smooth_partition_of_unity():
    lam = np.linspace(360, 830, 1001)
    b = S.basis(lam)
    np.testing.assert_allclose(b.sum(1), 1.0, atol=1e-12)
    assert (b >= 0).all()


def test_round_trips():
    Q, M, w = S.tables()
    # The equal-energy spectrum is white.
    np.testing.assert_allclose(S.project(lambda l: np.ones_like(l), M), [1, 1, 1], atol=1e-12)
    # Primaries and gray project back to themselves; gray is a flat spectrum.
    for rgb in (np.eye(3)[0], np.eye(3)[1], np.eye(3)[2]):
        back = S.project(lambda l: S.basis(l) @ (Q @ rgb), M)
        np.testing.assert_allclose(back, rgb, atol=1e-9)
    lam = S.grid()
    np.testing.assert_allclose(S.upsample([0.4, 0.4, 0.4], lam, Q), 0.4, atol=1e-9)
    # The film estimate is unbiased: E over uniform wavelengths of
    # 4 * w_scale * M cmf(l) L(l) (four samples, each pdf 1/470) equals the
    # projection, here for the flat spectrum.
    est = 4 * w * (M @ S.cmf(lam).T).mean(1)
    np.testing.assert_allclose(est, [1, 1, 1], rtol=1e-9)


def test_library_tables_match_the_oracle(pg):
    Q, M, w = pg.spectral_tables()
    Qo, Mo, wo = S.tables()
    np.testing.assert_allclose(Q, Qo, rtol=2e-4, atol=2e-5)
    np.testing.assert_allclose(M, Mo, rtol=2e-5, atol=1e-6)
    assert abs(w - wo) < 1e-6 * wo
