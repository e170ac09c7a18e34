"""The Atmosphere environment light's colour table in float64 with exact
exp/sqrt/pow/acos/cos/sin: the definition the library's `accurate: true`
table follows.  TEST INFRASTRUCTURE ONLY (the checker of tests/ and the golden
generator; never called by the library).

Pine computes the same table (Atmosphere::Atmosphere, light.cpp:81-92, and
AtmosphereColor, color.cpp:37-89) with its own approximate libm; there its
8-term Taylor exp of large negative arguments diverges and every entry comes
out NaN or +inf, which is why the accurate table exists.  Pins: the closed forms in
tests/test_atmosphere.py (the Rayleigh phase integrates to 1 over the sphere,
the per-channel ratio of an optically thin zenith path is betaR*phaseR*...,
the table's texel directions are UniformSphereSampling's) and the harness
render of pine's own Atmosphere light fed this table.
"""
import numpy as np

BETA_R = np.array([3.8e-6, 13.5e-6, 33.1e-6])
BETA_M = 21e-6
ATMOSPHERE_RADIUS, EARTH_RADIUS = 6420e3, 6360e3
N_VIEW, N_LIGHT = 8, 4


def sphere_t(o, d, r):
    """Sphere::ComputeT (geometry.cpp:114-126) about the origin, tmin 0; o, d [..., 3]."""
    a = np.sum(d * d, -1)
    b = 2.0 * np.sum(o * d, -1)
    c = np.sum(o * o, -1) - r * r
    disc = b * b - 4.0 * a * c
    sq = np.sqrt(np.maximum(disc, 0.0))
    t = (-b - sq) / (2.0 * a)
    t = np.where(t < 0.0, (-b + sq) / (2.0 * a), t)
    return np.where(disc <= 0.0, -1.0, t)


def atmosphere_color(d, sun, sun_color):
    """AtmosphereColor (color.cpp:37-89) for unit directions d [n, 3]."""
    d = np.asarray(d, np.float64)
    sun = np.asarray(sun, np.float64)
    along = np.all(d == sun, -1)
    hr_inv = 1.0 / 7995.0
    hm_inv = np.where(along, 1.0 / 10000.0, 1.0 / 1200.0)
    mu = d @ sun
    g = 0.76
    phase_r = 3.0 / (16.0 * np.pi) * (1.0 + mu * mu)
    phase_m = (3.0 / (8.0 * np.pi) * (1.0 - g * g) * (1.0 + mu * mu) /
               ((2.0 + g * g) * (1.0 + g * g - 2.0 * g * mu) ** 1.5))
    o = np.broadcast_to(np.array([0.0, EARTH_RADIUS, 0.0]), d.shape)
    seg = sphere_t(o, d, ATMOSPHERE_RADIUS) / N_VIEW
    od_r = np.zeros(len(d))
    od_m = np.zeros(len(d))
    sum_r = np.zeros((len(d), 3))
    sum_m = np.zeros((len(d), 3))
    for i in range(N_VIEW):
        sp = o + d * ((i + 0.5) * seg)[:, None]
        height = np.linalg.norm(sp, axis=-1) - EARTH_RADIUS
        hr = np.exp(-height * hr_inv) * seg
        hm = np.exp(-height * hm_inv) * seg
        od_r += hr
        od_m += hm
        seg_l = sphere_t(sp, np.broadcast_to(sun, sp.shape), ATMOSPHERE_RADIUS) / N_LIGHT
        lo2 = np.sum(sp * sp, -1)
        ld2 = sun @ sun
        lod = sp @ sun
        od_rl = np.zeros(len(d))
        od_ml = np.zeros(len(d))
        alive = np.ones(len(d), bool)
        for j in range(N_LIGHT):
            lt = (j + 0.5) * seg_l
            # color.cpp:66-72 as written: a squared distance minus a radius
            hl = lo2 + lt * lt * ld2 + 2.0 * lt * lod - EARTH_RADIUS
            alive &= hl >= 0.0
            od_rl += np.where(alive, np.exp(-np.maximum(hl, 0) * hr_inv) * seg_l, 0.0)
            od_ml += np.where(alive, np.exp(-np.maximum(hl, 0) * hm_inv) * seg_l, 0.0)
        tau = BETA_R * (od_r + od_rl)[:, None] + BETA_M * (od_m + od_ml)[:, None]
        tr = np.exp(-tau)
        sum_r += np.where(alive[:, None], tr * hr[:, None], 0.0)
        sum_m += np.where(alive[:, None], tr * hm[:, None], 0.0)
    return (sum_r * BETA_R * phase_r[:, None] + sum_m * BETA_M * phase_m[:, None]) * 10.0 * \
        np.asarray(sun_color, np.float64)


def texel_directions(w, h):
    """UniformSphereSampling((x / w, y / h)) (sampling.h:41-43) per texel, [h, w, 3];
    x / w and y / h rounded to float as pine computes them."""
    y, x = np.mgrid[0:h, 0:w]
    u0 = (x.astype(np.float32) / np.float32(w)).astype(np.float64)
    u1 = (y.astype(np.float32) / np.float32(h)).astype(np.float64)
    phi = u0 * 2.0 * np.pi
    theta = np.arccos(1.0 - 2.0 * u1)
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)


def atmosphere_table(sun_direction, sun_color, w, h):
    """(table [h, w, 3], colour along the sun [3]) as the Atmosphere constructor
    builds them (sun_direction normalised first), entries with a NaN channel set
    to 0 as its HasNaN() test says (light.cpp:88)."""
    sun = np.asarray(sun_direction, np.float64)
    sun = sun / np.linalg.norm(sun)
    dirs = texel_directions(w, h).reshape(-1, 3)
    with np.errstate(over="ignore", invalid="ignore"):
        table = atmosphere_color(dirs, sun, sun_color).reshape(h, w, 3)
        sun_le = atmosphere_color(sun[None, :], sun, sun_color)[0]
    table = np.where(np.isnan(table).any(-1, keepdims=True), 0.0, table)
    sun_le = np.where(np.isnan(sun_le).any(), 0.0, sun_le)
    return table, sun_le
