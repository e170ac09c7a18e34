"""The oracle (oracle/pinecpu.cpp) pinned against pine itself.

Golden vectors come from pine's own library, compiled from its sources by
oracle/Makefile and driven by oracle/_ref/ref_harness (tests/golden/make_golden.py).
Closed forms pin what the goldens leave open.
"""
import numpy as np
import pytest

from conftest import golden
from ties import certify_ties, mesh_prim_geom, single_prim


def ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


@pytest.mark.parametrize("name", ["sqrt", "exp", "log", "cos", "sin", "acos"])
def test_pstd_unary_bit_exact_vs_pine(oracle, name):
    # Pine's vectors are standalone calls (harness wrappers are out of line), so
    # they show how pine's -ffast-math build evaluates each function body.
    g = golden("pstd")[name]
    got = np.array([oracle.pstd(name, x) for x in g[:, 0]], np.float32)
    assert np.array_equal(got.view(np.int32), g[:, 1].view(np.int32)), \
        f"{name}: {(ulps(got, g[:, 1]) == 0).mean():.4f} bit-exact"


def test_pstd_pow_integer_exponent_matches_pine(oracle):
    # FrSchlick's pow(1 - cos, 5) (scattering.h:358-360): bit-exact.
    g = golden("pstd")["pow5"]
    got = np.array([oracle.pstd("pow", x, 5.0) for x in g[:, 0]], np.float32)
    assert np.array_equal(got.view(np.int32), g[:, 1].view(np.int32))


def test_pstd_pow_fractional_exponent_matches_pine(oracle):
    # pow(x, 1/2.2): the gamma of pine's 8-bit output (film.cpp:45-66).
    g = golden("pstd")["pow_gamma"]
    e = np.float32(1.0) / np.float32(2.2)
    got = np.array([oracle.pstd("pow", x, e) for x in g[:, 0]], np.float32)
    assert np.array_equal(got.view(np.int32), g[:, 1].view(np.int32))


def test_pstd_atan2_matches_pine(oracle):
    # Bit-exact for x > 0 and for y > 0.  In the y <= 0, x <= 0 branch the
    # golden wrapper's codegen differs from a standalone compile of the same
    # source by one ulp in a few percent of inputs, so 1 ulp is allowed there.
    # atan2 only feeds sphere (u, v) (geometry.cpp:140), unused by the shading
    # of constant materials.
    g = golden("pstd")["atan2"]
    got = np.array([oracle.pstd("atan2", y, x) for y, x, _ in g], np.float32)
    d = ulps(got, g[:, 2])
    upper = (g[:, 1] > 0) | (g[:, 0] > 0)
    assert d[upper].max() == 0
    assert d.max() <= 1 and (d == 0).mean() > 0.95


def test_pstd_closed_forms(oracle):
    # Sanity of the approximations against the true functions (catches a
    # dropped Newton step or a wrong constant even where pine has no vector).
    for x in [0.25, 1.0, 2.0, 9.0, 1e4]:
        assert abs(oracle.pstd("sqrt", x) - np.sqrt(x)) <= 1e-6 * np.sqrt(x)
    for x in [0.1, 0.5, 1.0]:
        # Horner loop from y=0: terms up to x^7/7!, so the error is ~x^8/8!.
        assert abs(oracle.pstd("exp", x) - np.exp(x)) < 3e-5
    for x in [0.25, 0.5, 0.9]:
        assert abs(oracle.pstd("log", x) - np.log(x)) < 5e-3   # 8 bisection steps
    for x in [0.0, 0.3, 1.2, 2.5, -4.0]:
        assert abs(oracle.pstd("cos", x) - np.cos(x)) < 1e-3
        assert abs(oracle.pstd("sin", x) - np.sin(x)) < 1e-3


def _mesh_rays(pg, fx):
    from pine_amd import workloads as W
    m = fx["mesh"]
    spec = W.MeshSpec(int(m[0]), float(fx["mesh_amp"]), int(m[1]), int(m[2]))
    v, i = W.make_mesh(spec)
    rs = fx["rayspec"]
    rspec = W.RaySpec(int(rs[0]), tuple(float(x) for x in fx["eye"]), float(fx["fov"]), int(rs[1]),
                      int(rs[2]))
    rays = W.make_rays(rspec, v, i)[fx["ray_index"]]
    return v, i, rays


def test_generator_reproduces_golden_rays(pg):
    import hashlib
    for case in ["trace_5k", "trace_82k"]:
        fx = golden(case)
        _, _, rays = _mesh_rays(pg, fx)
        assert hashlib.sha256(rays.tobytes()).hexdigest() == fx["ray_sha256"].item().decode()


@pytest.mark.parametrize("case", ["trace_5k", "trace_82k"])
def test_brute_force_matches_pine_bvh(pg, oracle, case):
    fx = golden(case)
    v, i, rays = _mesh_rays(pg, fx)
    prim, t, u, vv, occ = oracle.trace_brute(oracle.tri9(v, i), rays)
    hit = fx["prim"] >= 0
    assert np.array_equal(prim >= 0, hit)
    same = prim == fx["prim"]
    assert np.array_equal(t[hit].view(np.int32), fx["t"][hit].view(np.int32))
    # Certified ties only: both ids hit at pine's t, the smaller one kept.
    ties = certify_ties(oracle, rays, prim, fx["prim"], fx["t"], mesh_prim_geom(oracle, v, i))
    print(f"{case}: {ties} certified ties")
    assert np.array_equal(u[hit & same].view(np.int32), fx["u"][hit & same].view(np.int32))
    assert np.array_equal(vv[hit & same].view(np.int32), fx["v"][hit & same].view(np.int32))
    assert np.array_equal(occ, fx["occluded"].astype(np.int32))


def shapes_analytic():
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as mg
    shapes = mg.shapes_scene()
    ana, tri, order = [], None, []
    for k, s in enumerate(shapes):
        if s[0] == "sphere":
            ana.append([1, *s[1], s[2], 0, 0, 0, 0, 0])
        elif s[0] == "rect":
            ana.append([2, *s[1], *s[2], *s[3]])
        elif s[0] == "plane":
            ana.append([3, *s[1], *s[2], 0, 0, 0])
        else:
            tri = (s[1], s[2])
    return shapes, np.array(ana, np.float32), tri


def test_brute_force_shapes_match_pine(oracle):
    fx = golden("shapes")
    shapes, ana, (v, i) = shapes_analytic()
    tri = oracle.tri9(v, i)
    prim, t, u, vv, occ = oracle.trace_brute(tri, fx["rays"], ana)
    # Oracle ids: triangles first, then analytic in shape order (mesh is last shape).
    n_ana = len(ana)
    mesh_shape = len(shapes) - 1
    exp = np.where(fx["shape"] == mesh_shape, fx["prim"],
                   np.where(fx["shape"] >= 0, len(tri) + fx["shape"], -1))
    hit = exp >= 0
    assert np.array_equal(prim >= 0, hit)
    assert np.array_equal(t[hit].view(np.int32), fx["t"][hit].view(np.int32))

    def geom(pid):
        if pid < len(tri):
            return single_prim(tri=tri[pid])
        return single_prim(ana=ana[pid - len(tri)])
    ties = certify_ties(oracle, fx["rays"], prim, exp, fx["t"], geom)
    print(f"shapes: {ties} certified ties")
    assert np.array_equal(occ, fx["occluded"].astype(np.int32))


def test_closed_form_sphere_and_plane(oracle):
    # Axis ray through a sphere centre: t = |c - o| - r; plane at y = 0.
    rays = np.array([[0, 0, -10, 0, 0, 0, 1, 3.4e38],
                     [0, 5, 0, 0, 0, -1, 0, 3.4e38],
                     [0, 5, 0, 0, 0, 1, 0, 3.4e38]], np.float32)
    ana = np.array([[1, 0, 0, 0, 2.0, 0, 0, 0, 0, 0],
                    [3, 0, 0, 0, 0, 1, 0, 0, 0, 0]], np.float32)
    prim, t, _, _, occ = oracle.trace_brute(None, rays, ana)
    assert prim.tolist() == [0, 0, -1]       # sphere first (t=8 < plane's... ray 1 hits sphere at 3)
    assert abs(t[0] - 8.0) < 1e-5 and abs(t[1] - 3.0) < 1e-5
    assert occ.tolist() == [1, 1, 0]


def test_cpu_bvh_port_agrees_with_brute_force(pg, oracle):
    fx = golden("trace_5k")
    v, i, rays = _mesh_rays(pg, fx)
    tri = oracle.tri9(v, i)
    p1, t1, *_ = oracle.trace_brute(tri, rays)
    b = oracle.CpuBVH(tri)
    p2, t2, _ = b.trace(rays)
    assert np.array_equal(t1.view(np.int32), t2.view(np.int32))
    ties = certify_ties(oracle, rays, p1, p2, t1, mesh_prim_geom(oracle, v, i))
    print(f"port vs brute force: {ties} certified ties")


def test_halton_matches_pine(oracle):
    g = golden("halton")
    for key in g.files:
        w, h = (int(x) for x in key.split("_")[0].split("x"))
        spp = int(key.split("spp")[1].split("_")[0])
        nd = int(key.split("_d")[1].split("_")[0])
        npix = int(key.split("_n")[1])
        got = oracle.halton_values(w, h, spp, nd, 3, 1, npix)
        assert np.array_equal(got.view(np.int32), g[key].view(np.int32)), key


def test_halton_is_a_permuted_van_der_corput(oracle):
    # Dimension 2 uses base 5: over 5^k consecutive sample indices of one pixel
    # every stratum [j/5^k, (j+1)/5^k) holds exactly one value.
    # Pixel (1, 0): its first Halton index is not 0 (index 0 has no digits and
    # maps to the tail term alone).
    vals = oracle.halton_values(5, 5, 125, 3, 1, 0, 1)[0, :, 0]
    # Samples of one pixel step the Halton index by the stride (2*3 here), which
    # is coprime with 5, so the base-5 digits of 125 consecutive samples cover
    # all residues.
    # Coarse strata (2 digits) so that float rounding at a boundary cannot move
    # a value: each of the 25 strata holds exactly 5 of the 125 values.
    strata = np.floor(vals * 25).astype(int)
    assert np.bincount(strata, minlength=25).tolist() == [5] * 25


def _sampler_key(key):
    w, h = (int(x) for x in key.split("_")[0].split("x"))
    spp = int(key.split("spp")[1].split("_")[0])
    nd = int(key.split("_d")[1].split("_")[0])
    npix = int(key.split("_n")[1].split("_")[0])
    strategy = int(key.split("_s")[-1]) if "_s" in key.split("_n")[1] else None
    return w, h, spp, nd, npix, strategy


def test_sobol_matches_pine(oracle):
    """Pine's SobolSampler (all four scramblers, film sizes 1x1 to 1024x96, a run
    past dimension 1024) reproduced bit for bit."""
    g = golden("sobol")
    for key in g.files:
        w, h, spp, nd, npix, strategy = _sampler_key(key)
        got = oracle.sobol_values(w, h, spp, nd, 3, 1, npix, strategy)
        assert np.array_equal(got.view(np.int32), g[key].view(np.int32)), key


@pytest.mark.parametrize("strategy", [0, 1, 2, 3])
def test_sobol_pixels_tile_the_sequence(oracle, strategy):
    """On a 16x16 film (m = 4) sample 0 of the 256 pixels uses the Sobol indices
    0..255, each once, so every dimension's 256 values are the first 256 points
    of a Sobol coordinate: one value per stratum [i/256, (i+1)/256].  XOR and
    Owen scrambling permute dyadic intervals, so this holds for every scrambler.
    Sample 1 of the pixels uses indices 256..511: stratified again."""
    n = 16 * 16
    for sample in (0, 1):
        vals = oracle.sobol_values(16, 16, sample + 1, 10, 0, 0, n, strategy)[:, sample, :]
        # pixel k of the harness spread is (977 k) mod 256: every pixel once
        assert len({(977 * k) % n for k in range(n)}) == n
        for d in range(10):
            # j-th smallest value in the closed stratum [j/n, (j+1)/n]: float
            # rounding may carry a scrambled value onto its stratum's end
            v = np.sort(vals[:, d].astype(np.float64)) * n
            j = np.arange(n)
            assert ((v >= j) & (v <= j + 1)).all(), (sample, d)


def test_noise_matches_pine(oracle):
    """pine's PerlinNoise (seeds 0..2), Turbulence and Turbulence3D (noise.cpp)
    at 2,000 seeded points, frequencies 1..7, 1..5 octaves: bit for bit."""
    g = golden("noise")
    got = oracle.noise(g["points"], g["freq"], g["octaves"])
    ref = np.concatenate([g["perlin"], g["turbulence"][:, None], g["turbulence3d"]], 1)
    assert np.array_equal(got.view(np.int32), ref.view(np.int32))


def test_gradient_noise_is_one_half_on_the_lattice(oracle):
    """Closed form: at a lattice point the smoothstep weights select the corner
    itself, whose offset is zero, so every gradient term vanishes and
    PerlinNoise = 0.5 (1 + 0) exactly, for any seed; Turbulence of one octave
    is then (0.5 / (2 - 1))^2 = 0.25."""
    rng = np.random.default_rng(5)
    pts = rng.integers(-50, 50, size=(200, 3)).astype(np.float32)
    out = oracle.noise(pts, np.ones(200, np.float32), np.ones(200, np.float32))
    assert (out[:, :3] == 0.5).all()
    assert (out[:, 3] == 0.25).all() and (out[:, 4:] == 0.25).all()


def test_guide_oracle_encoding_is_the_frequency_basis():
    """Every (input, octave, sin/cos) pair and every raw input appears exactly
    once, in the first 225 columns; each column equals sin / cos(pi 2^(k+oct) x_c)
    evaluated without range reduction."""
    from oracle import guide as G
    lay = G.feature_layout()
    per = sorted((c, k, s) for kind, c, k, s in lay if kind == 2)
    assert per == sorted((c, k, s) for c in range(9) for k in range(12) for s in (0, 1))
    assert sorted(c for kind, c, _, _ in lay if kind == 1) == list(range(9))
    assert all(kind == 0 for kind, *_ in lay[225:]) and all(kind != 0 for kind, *_ in lay[:225])
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (50, 9))
    for oct in (0, -2):
        f = G.encode(x, oct)
        for col, (kind, c, k, s) in enumerate(lay):
            if kind == 2:
                a = np.pi * 2.0 ** (k + oct) * x[:, c]
                np.testing.assert_allclose(f[:, col], np.cos(a) if s else np.sin(a), atol=1e-9)
            elif kind == 1:
                assert np.array_equal(f[:, col], x[:, c])
            else:
                assert not f[:, col].any()


def test_guide_oracle_backward_pinned_by_finite_differences():
    """The guide network's float64 oracle: its analytic gradient (the GPU's
    checker) against central differences of its own loss."""
    from oracle import guide as G
    rng = np.random.default_rng(0)
    n = 12
    p = G.init_params(3) + G.live_mask() * rng.normal(scale=0.05, size=G.N_PARAMS)
    x = np.concatenate([rng.uniform(0, 1, (n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))], 1)
    w = rng.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    wt, pb = rng.uniform(0.1, 2, n), rng.uniform(0.05, 0.5, n)
    _, g = G.loss_and_grad(p, x, w, wt, pb)
    # The gradient holds q constant inside q_mix: difference that surrogate.
    q0, _, _ = G.pdf(p, x, w, pb)
    live = np.nonzero(G.live_mask())[0]
    sel = np.concatenate([rng.choice(live, 150, replace=False),
                          np.arange(G.N_PARAMS - 48 - 48 * 64, G.N_PARAMS - 48 * 64 + 64, 7),
                          np.arange(G.N_PARAMS - 48, G.N_PARAMS - 7)])
    eps = 1e-6
    for k in sel:
        d = np.zeros_like(p)
        d[k] = eps
        num = (G.loss_and_grad(p + d, x, w, wt, pb, q_fixed=q0)[0] -
               G.loss_and_grad(p - d, x, w, wt, pb, q_fixed=q0)[0]) / (2 * eps)
        assert abs(num - g[k]) <= 1e-4 * (abs(num) + abs(g[k])) + 1e-9, (k, num, g[k])
    assert np.all(g[G.live_mask() == 0] == 0)


def test_guide_oracle_vmf_closed_forms():
    """vMF density integrates to 1 and its samples have the closed-form mean
    cosine coth(kappa) - 1/kappa; PHI is a permutation."""
    from oracle import guide as G
    rng = np.random.default_rng(1)
    assert sorted(G.PHI.tolist()) == list(range(64))
    n = 400000
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    for kappa in (0.05, 1.0, 8.0, 50.0):
        mu = np.array([[[0.6, 0.0, 0.8]]])
        v = G.vmf(np.broadcast_to(mu, (n, 1, 3)), np.full((n, 1), kappa), u)[:, 0]
        if kappa <= 8:
            assert abs(4 * np.pi * v.mean() - 1) < 0.02
        s = G.sample_vmf(np.broadcast_to(mu[0], (n, 3)), np.full(n, kappa), rng.uniform(size=n),
                         rng.uniform(size=n))
        mean_cos = (s @ mu[0, 0]).mean()
        assert abs(mean_cos - (1 / np.tanh(kappa) - 1 / kappa)) < 4e-3


def test_tie_certifier_accepts_ties_and_rejects_other_ids(oracle):
    """certify_ties on two exact duplicate triangles (ids 0 and 3) and a third
    triangle behind them (id 1): id 3 in place of 0 is a certified tie, id 1
    (hit, but farther) is not, nor is the larger id kept on a tie."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float32)
    i = np.array([[0, 1, 2], [3, 4, 5], [2, 1, 0], [0, 1, 2]], np.uint32)
    rays = np.array([[0.2, 0.2, -1, 0, 0, 0, 1, 3.4e38]], np.float32)
    prim, t, *_ = oracle.trace_brute(oracle.tri9(v, i), rays)
    assert prim.tolist() == [0] and t[0] == np.float32(1.0)
    geom = mesh_prim_geom(oracle, v, i)
    assert certify_ties(oracle, rays, prim, np.array([3], np.int32), t, geom) == 1
    with pytest.raises(AssertionError, match="uncertified"):
        certify_ties(oracle, rays, prim, np.array([1], np.int32), t, geom)
    with pytest.raises(AssertionError, match="larger id"):
        certify_ties(oracle, rays, np.array([3], np.int32), prim, t, geom)
