"""Test infrastructure only: float64 restatement of the spectral mode's colour
math (pine_amd/csrc/pg_spectral.h; pine itself renders RGB, so there is no
pine counterpart).  Checked in tests/test_spectral.py against closed forms: the
CIE 1931 integrals, the equal-energy white point, exact round trips of the
primaries and of gray."""
import numpy as np

LAM_MIN, LAM_RANGE = 360.0, 470.0
XYZ_TO_SRGB = np.array([[3.2404542, -1.5371385, -0.4985314],
                        [-0.9692660, 1.8760108, 0.0415560],
                        [0.0556434, -0.2040259, 1.0572252]])


def lobe(x, mu, s1, s2):
    t = (x - mu) / np.where(x < mu, s1, s2)
    return np.exp(-0.5 * t * t)


def cmf(lam):
    """CIE 1931 2-degree matching functions, Wyman-Sloan-Shirley multi-lobe fit: [n, 3]."""
    lam = np.asarray(lam, np.float64)
    x = 1.056 * lobe(lam, 599.8, 37.9, 31.0) + 0.362 * lobe(lam, 442.0, 16.0, 26.7) - \
        0.065 * lobe(lam, 501.1, 20.4, 26.2)
    y = 0.821 * lobe(lam, 568.8, 46.9, 40.5) + 0.286 * lobe(lam, 530.9, 16.3, 31.1)
    z = 1.217 * lobe(lam, 437.0, 11.8, 36.0) + 0.681 * lobe(lam, 459.0, 26.0, 13.8)
    return np.stack([x, y, z], -1)


def basis(lam):
    """(red, green, blue) band weights: [n, 3], rows sum to 1."""
    lam = np.asarray(lam, np.float64)
    a = 1.0 / (1.0 + np.exp(-(lam - 490.0) / 12.0))
    c = 1.0 / (1.0 + np.exp(-(lam - 590.0) / 12.0))
    return np.stack([c, a - c, 1.0 - a], -1)


def grid():
    return LAM_MIN + np.arange(470) + 0.5


def tables():
    """Q (RGB -> basis coefficients), M (XYZ -> white-balanced RGB), w_scale."""
    lam = grid()
    c, b = cmf(lam), basis(lam)
    Y = c[:, 1].sum()
    white = XYZ_TO_SRGB @ (c.sum(0) / Y)
    M = XYZ_TO_SRGB / white[:, None]
    P = M @ (c.T @ b) / Y       # RGB of each band's spectrum (columns)
    return np.linalg.inv(P), M, LAM_RANGE / (4.0 * Y)


def upsample(rgb, lam, Q):
    """Spectrum of rgb [3] at wavelengths lam: max(0, b(lam)^T Q rgb)."""
    return np.maximum(0.0, basis(lam) @ (Q @ np.asarray(rgb, np.float64)))


def project(spectrum_fn, M):
    """RGB of a spectrum (callable on wavelengths) by the 1 nm rule."""
    lam = grid()
    c = cmf(lam)
    return M @ (c.T @ spectrum_fn(lam)) / c[:, 1].sum()
