"""Image-error measurement for comparing estimators (plain vs guided path
tracing) at equal time.

pine's Path estimator produces rare, extreme samples (near-light geometry
terms; pine's own CPU render of the occluded-light box shows them at the same
pixels), so:
  - the reference is the per-pixel median of the means of disjoint sample
    ranges of the plain tracer (pg_render's sample_offset), which converges to
    the expectation without being dominated by one sample;
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


def robust_reference(scene, spp_per_chunk, chunks):
    """Per-pixel median over `chunks` renders of disjoint sample ranges."""
    imgs = []
    for c in range(chunks):
        scene.render(spp=spp_per_chunk, sample_offset=c * spp_per_chunk)
        imgs.append(linear(scene.film()))
    return np.median(np.stack(imgs), axis=0)


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
