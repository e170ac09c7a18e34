"""Closed-form expectations for the constant-density slab scenes of
pine_amd.workloads.slab_scene_text (test helper, no GPU).

pine's GridMedium (src/core/medium.cpp:139-160) places voxel i of an N-voxel
axis at model coordinate i/N and reads zero outside the grid, so a grid of
constant density 1 interpolates to rho(p) = f(x) f(y) f(z) with f(u) = 1 for
u*N <= N-1 and f(u) = N - u*N on the last voxel.  With scale 1 and position 0
the grid spans the world cube [-0.5, 0.5]^3 and model distance equals world
distance.  A ray's transmittance is exp(-sigma * integral of rho); delta
tracking (albedo 0) and ratio tracking are unbiased estimators of it.
"""
import numpy as np

N = 8  # voxels per axis of the constant grid


def falloff(u):
    v = u * N
    inside = (u >= 0.0) & (u <= 1.0)
    return np.where(inside, np.clip(np.where(v <= N - 1, 1.0, N - v), 0.0, 1.0), 0.0)


def _normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def camera_rays(frm, to, fov, w, h, sub=4):
    """Pinhole rays of ThinLenCamera::GenRay (camera.cpp:31-41, LookAt
    vecmath.h:1212-1221) at sub x sub stratified points of every pixel, film
    rows top first (row r samples pine's film y = h-1-r).  Returns origins
    [h, w, sub*sub, 3] and directions of the same shape."""
    F, T = np.array(frm, float), np.array(to, float)
    z = _normalize(T - F)
    x = _normalize(np.cross([0.0, 1.0, 0.0], z))
    y = np.cross(z, x)
    tf = np.tan(fov / 2.0)
    s = (np.arange(sub) + 0.5) / sub
    px = (np.arange(w)[:, None] + s[None, :]) / w                  # [w, sub]
    py = (np.arange(h)[::-1][:, None] + s[None, :]) / h            # [h, sub], row 0 = top
    fx = np.broadcast_to(px[None, :, None, :], (h, w, sub, sub))
    fy = np.broadcast_to(py[:, None, :, None], (h, w, sub, sub))
    sx = (fx - 0.5) * 2.0 * tf * (w / h)
    sy = (fy - 0.5) * 2.0 * tf
    d = _normalize(sx[..., None] * x + sy[..., None] * y + z).reshape(h, w, sub * sub, 3)
    o = np.broadcast_to(F, d.shape)
    return o, d


def optical_depth(o, d, n=512):
    """Integral of rho along rays (o, d) in world space, midpoint rule."""
    om = o + 0.5
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0, t1 = (0.0 - om) * inv, (1.0 - om) * inv
    tn = np.maximum(np.nanmax(np.minimum(t0, t1), axis=-1), 0.0)
    tf = np.nanmin(np.maximum(t0, t1), axis=-1)
    length = np.maximum(tf - tn, 0.0)
    k = (np.arange(n) + 0.5) / n
    tau = np.zeros(o.shape[:-1])
    for j in range(0, n, 64):
        t = tn[..., None] + length[..., None] * k[None, j:j + 64]
        p = om[..., None, :] + t[..., None] * d[..., None, :]
        rho = falloff(p[..., 0]) * falloff(p[..., 1]) * falloff(p[..., 2])
        tau += rho.sum(-1) * (length / n)
    return tau


def expected_slab(view, img, sigma=1.0):
    """Expected linear image of slab_scene_text(view) and the pixel mask where
    the closed form holds.  For "shadow" the lit ground radiance is taken from
    the image's fully lit pixels (returned as `scale`)."""
    h, w = img.shape[:2]
    if view == "through":
        o, d = camera_rays((0, 0, 3), (0, 0, 0), 0.25, w, h)
        T = np.exp(-sigma * optical_depth(o, d)).mean(-1)
        return T, np.ones((h, w), bool), 1.0
    o, d = camera_rays((0, -0.7, 3), (0, -1, 0), 0.9, w, h)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = (-1.0 - o[..., 1]) / d[..., 1]
    hitp = o + t[..., None] * d
    on_ground = (t > 0) & (np.abs(hitp[..., 0]) < 20) & (np.abs(hitp[..., 2]) < 20)
    # Straight up through the cube: falloff(x) falloff(z) times the integral of
    # falloff over y, (N-1)/N + 1/(2N).
    tau = falloff(hitp[..., 0] + 0.5) * falloff(hitp[..., 2] + 0.5) * ((N - 1) / N + 0.5 / N)
    T = np.where(on_ground, np.exp(-sigma * tau), 0.0).mean(-1)
    mask = on_ground.all(-1)
    lit = mask & (tau.max(-1) == 0.0)
    scale = float(img[lit].mean())
    return scale * T, mask, scale
