"""Image-error measurement for comparing estimators (plain vs guided path
tracing) at equal time.

pine's Path estimator produces rare, extreme samples: non-finite pixels and
single paths carrying ~1e9 radiance after many bounces (pine's own CPU render
of the occluded-light box holds the same non-finite pixels at the same
sample indices).  So:
  - the reference is the per-pixel mean of the means of disjoint sample
    ranges of the plain tracer (pg_render's sample_offset), leaving out the
    ranges whose mean is non-finite or exceeds `outlier` times the pixel's
    median range mean: it converges to the expectation, minus only those
    singular events (a median would be biased low wherever rare bright paths
    carry the light, which are exactly the pixels guiding improves);
  - relMSE = mean over pixels of (img - ref)^2 / (ref^2 + 1e-2), computed after
    dropping the `trim` fraction of pixels with the largest error (default
    0.1%), the same way for every image compared.
"""
import numpy as np


def linear(film):
    w = film[..., 3:4]
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(w != 0, film[..., :3] / np.where(w != 0, w, 1), 0)


def relmse(img, ref, trim=1e-3):
    ok = np.isfinite(img).all(-1) & np.isfinite(ref).all(-1)
    e = ((img - ref) ** 2 / (ref ** 2 + 1e-2)).mean(-1)
    e = np.where(ok, e, np.inf)
    n = e.size
    keep = int(round(n * (1 - trim)))
    return float(np.sort(e.ravel())[:keep].mean())


def relmse_untrimmed(img, ref):
    """(relMSE over every pixel whose value and reference are finite, with no
    trimming; the number of the other pixels)."""
    ok = np.isfinite(img).all(-1) & np.isfinite(ref).all(-1)
    e = ((img - ref) ** 2 / (ref ** 2 + 1e-2)).mean(-1)
    return float(e[ok].mean()), int((~ok).sum())


def robust_reference(scene, spp_per_chunk, chunks, outlier=1000.0, progress=None):
    """Per-pixel mean over `chunks` renders of disjoint sample ranges, leaving
    out singular ranges (see the module docstring).  progress(c) is called
    after every 8th chunk."""
    imgs = []
    for c in range(chunks):
        scene.render(spp=spp_per_chunk, sample_offset=c * spp_per_chunk)
        imgs.append(linear(scene.film()))
        if progress and (c + 1) % 8 == 0:
            progress(c + 1)
    return reference_from_chunks(np.stack(imgs), outlier)


def reference_from_chunks(imgs, outlier=1000.0):
    """imgs [chunks, h, w, 3] -> [h, w, 3]: mean over the chunks whose pixel
    luminance is finite and at most `outlier` times the median (all chunks
    count where the median is zero)."""
    lum = imgs.mean(-1)
    with np.errstate(invalid="ignore"):
        med = np.nanmedian(np.where(np.isfinite(lum), lum, np.nan), axis=0)
        keep = np.isfinite(lum) & ((lum <= outlier * med[None]) | ~(med[None] > 0))
    w = keep[..., None].astype(np.float64)
    tot = np.where(keep[..., None], imgs, 0.0).sum(0)
    n = w.sum(0)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(n > 0, tot / np.maximum(n, 1), np.nan).astype(np.float32)


def time_to_error(times, errors, target):
    """Time at which an estimator reaches `target` relMSE, interpolating
    log(error) against log(time) between measured points (None if never)."""
    t, e = np.log(np.asarray(times, float)), np.log(np.asarray(errors, float))
    order = np.argsort(t)
    t, e = t[order], e[order]
    lt = np.log(target)
    for k in range(len(t)):
        if e[k] <= lt:
            if k == 0:
                return float(np.exp(t[0]))
            f = (lt - e[k - 1]) / (e[k] - e[k - 1])
            return float(np.exp(t[k - 1] + f * (t[k] - t[k - 1])))
    return None
