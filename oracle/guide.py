"""Test infrastructure only: the float64 CPU oracle of the learned path-guiding
network (pine has none; its definition is this build's).  The GPU kernels in
pine_amd/csrc/guide.hip compute the same function in BF16 with FP32
accumulation; tests compare their outputs and gradients with this module, and
this module's backward pass is itself pinned by finite differences.

Network (padded parameter vector of 27,888 floats, 25,449 of them live):
  input  x[9] = (p - bmin) / (bmax - bmin), n, wo
  f[256] = x[0..8] and sin/cos(pi * 2^(k + oct) * x[c]) for frequencies
           k = 0..11, in the column order of feature_layout(); f[225..255] = 0
           (oct: the guide's freq_offset, 0 by default)
  h1 = relu(W0 f + b0)                W0 [64][256]
  h2 = relu(W1 h1[PHI] + b1)          W1 [64][64]
  h3 = relu(W2 h2[PHI] + b2)          W2 [64][64]
  o  = W3 h3[PHI] + b3                W3 [48][64] (41 live rows)
PHI is a fixed reordering of the 64 hidden units (the order in which the GPU's
matrix-core output registers feed the next layer).  Head, per lobe j < 8, with
m_j = o[OUT_UNIT[j]] (the output units of lobe j; OUT_UNIT places lobes 2g and
2g + 1 in the output registers of the GPU's lane group g):
  mu_j = normalize(m_j[0..2]), kappa_j = exp(clamp(m_j[3], -4, 7)),
  pi = softmax(m_j[4]); alpha = 0.5 * sigmoid(o[ALPHA_UNIT]) (guide share);
  vMF_j(w) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)),
  q(w) = sum_j pi_j vMF_j(w), q_mix = alpha q + (1 - alpha) p_bsdf.
Loss over a batch, mean over samples of
  -weight * (log q_u(w) + log q_mix(w)),  q_u = 0.9 q + 0.1 / (4 pi),
with q held constant inside q_mix: the guide density learns the radiance
samples' distribution by its own cross-entropy (with a defensive uniform part
that bounds its gradient), and alpha learns the best share of the mixture
(weight = radiance * BSDF pdf / sampling pdf).  Decoupled, alpha cannot starve
the guide of gradient when it drops.
"""
import numpy as np

N_IN, N_FEAT, N_FREQ, H, N_OUT, N_OUT_PAD, N_LOBES = 9, 256, 12, 64, 41, 48, 8
ALPHA_MAX = 0.5
Q_DEFENSIVE = 0.1  # uniform share inside the guide's own objective
KR_MIN, KR_MAX = -4.0, 7.0

# Padded parameter layout (floats): W0 | b0 | W1 | b1 | W2 | b2 | W3 | b3
SHAPES = [("W0", (H, N_FEAT)), ("b0", (H,)), ("W1", (H, H)), ("b1", (H,)), ("W2", (H, H)),
          ("b2", (H,)), ("W3", (N_OUT_PAD, H)), ("b3", (N_OUT_PAD,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def phi():
    """k -> hidden unit feeding input k of the next layer: for k = 32s + 8g + i,
    unit 16 (2s + i // 4) + 4g + i % 4."""
    k = np.arange(H)
    s, g, i = k // 32, (k // 8) % 4, k % 8
    return 16 * (2 * s + i // 4) + 4 * g + i % 4


PHI = phi()


def out_units():
    """Output unit of lobe j = 2g + h, component c: slot s = 5h + c of lane
    group g's twelve output registers, unit 16 (s // 4) + 4g + s % 4.  Alpha
    takes lane group 0's slot 10."""
    u = np.zeros((N_LOBES, 5), int)
    for j in range(N_LOBES):
        g, h = divmod(j, 2)
        for c in range(5):
            s = 5 * h + c
            u[j, c] = 16 * (s // 4) + 4 * g + s % 4
    return u, 16 * (10 // 4) + 10 % 4


OUT_UNIT, ALPHA_UNIT = out_units()
LIVE_OUT = np.sort(np.append(OUT_UNIT.ravel(), ALPHA_UNIT))


def unpack(params):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(params[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def live_mask():
    """1 for live parameters, 0 for the padding (W0 columns 225.., W3/b3 rows
    outside LIVE_OUT)."""
    m = {name: np.ones(shp) for name, shp in SHAPES}
    m["W0"][:, N_IN + 2 * N_FREQ * N_IN:] = 0
    dead = np.setdiff1d(np.arange(N_OUT_PAD), LIVE_OUT)
    m["W3"][dead, :] = 0
    m["b3"][dead] = 0
    return np.concatenate([m[name].ravel() for name, _ in SHAPES])


def init_params(seed=1):
    """He-uniform hidden layers, small output layer, zero biases (what
    pg_guide_create does, restated for tests that need a fixed start)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in SHAPES:
        if name.startswith("b"):
            p[name] = np.zeros(shp)
        else:
            fan_in = shp[1]
            lim = np.sqrt(6.0 / fan_in) * (0.1 if name == "W3" else 1.0)
            p[name] = rng.uniform(-lim, lim, shp)
    flat = np.concatenate([p[name].ravel() for name, _ in SHAPES])
    return flat * live_mask()


def feature_layout():
    """Column f = 8b + i of W0 as (kind, input c, octave k, cosine?) with kind
    0 zero, 1 raw input, 2 periodic.  Blocks b < 24 hold the sine (b even) or
    cosine (b odd) of frequency k = b // 2 for inputs 0..7; blocks 24..26 the
    24 (k, sin/cos) rows of input 8; block 27 inputs 0..7 raw; column 224 input
    8 raw; the rest is padding (columns 225.. are dead)."""
    out = []
    for f in range(N_FEAT):
        b, i = divmod(f, 8)
        if b < 24:
            out.append((2, i, b // 2, b % 2))
        elif b < 27:
            rho = 8 * (b - 24) + i
            out.append((2, 8, rho // 2, rho % 2))
        elif b == 27:
            out.append((1, i, 0, 0))
        elif b == 28 and i == 0:
            out.append((1, 8, 0, 0))
        else:
            out.append((0, 0, 0, 0))
    return out


def encode(x, oct=0):
    """x [n, 9] -> features [n, 256] in feature_layout() order:
    sin/cos(pi 2^(k+oct) x_c) evaluated as sin/cos(2 pi frac(2^(k+oct-1) x_c)),
    the range reduction the GPU uses, and the raw inputs."""
    x = np.asarray(x, np.float64)
    n = x.shape[0]
    f = np.zeros((n, N_FEAT))
    for col, (kind, c, k, cos) in enumerate(feature_layout()):
        if kind == 1:
            f[:, col] = x[:, c]
        elif kind == 2:
            r = x[:, c] * 2.0 ** (k + oct - 1)
            r = r - np.floor(r)
            f[:, col] = np.cos(2 * np.pi * r) if cos else np.sin(2 * np.pi * r)
    return f


def forward(params, x, oct=0):
    """Raw outputs o [n, 48] and the activations the backward pass needs."""
    P = unpack(params)
    f = encode(x, oct)
    z1 = f @ P["W0"].T + P["b0"]
    h1 = np.maximum(z1, 0)
    u1 = h1[:, PHI]
    z2 = u1 @ P["W1"].T + P["b1"]
    h2 = np.maximum(z2, 0)
    u2 = h2[:, PHI]
    z3 = u2 @ P["W2"].T + P["b2"]
    h3 = np.maximum(z3, 0)
    u3 = h3[:, PHI]
    o = u3 @ P["W3"].T + P["b3"]
    return o, (f, u1, u2, u3)


def head(o):
    """Mixture parameters from raw outputs: mu [n, 8, 3], kappa [n, 8], pi [n, 8], alpha [n]."""
    m = o[:, OUT_UNIT]
    raw = m[:, :, :3]
    norm = np.maximum(np.linalg.norm(raw, axis=-1), 1e-6)
    mu = raw / norm[..., None]
    kappa = np.exp(np.clip(m[:, :, 3], KR_MIN, KR_MAX))
    lg = m[:, :, 4]
    e = np.exp(lg - lg.max(axis=1, keepdims=True))
    pi = e / e.sum(axis=1, keepdims=True)
    alpha = ALPHA_MAX / (1.0 + np.exp(-o[:, ALPHA_UNIT]))
    return mu, kappa, pi, alpha


def vmf(mu, kappa, w):
    """vMF density of directions w [n, 3] for every lobe: [n, 8]."""
    c = np.einsum("njk,nk->nj", mu, w)
    return kappa / (2 * np.pi * (-np.expm1(-2 * kappa))) * np.exp(kappa * (c - 1.0))


def pdf(params, x, w, p_bsdf, oct=0):
    """Guide density q(w), mixing share alpha and mixture density q_mix."""
    o, _ = forward(params, x, oct)
    mu, kappa, pi, alpha = head(o)
    q = (pi * vmf(mu, kappa, w)).sum(axis=1)
    return q, alpha, alpha * q + (1 - alpha) * p_bsdf


def loss_and_grad(params, x, w, weight, p_bsdf, oct=0, q_fixed=None):
    """Mean loss over the batch and its gradient wrt the padded parameter vector.
    The gradient treats q as a constant inside q_mix; with q_fixed (the guide
    density at the point of differentiation) the returned loss is the
    surrogate whose exact gradient that is, for finite-difference checks."""
    P = unpack(params)
    n = x.shape[0]
    o, (f, u1, u2, u3) = forward(params, x, oct)
    m = o[:, OUT_UNIT]
    raw = m[:, :, :3]
    rn = np.linalg.norm(raw, axis=-1)
    norm = np.maximum(rn, 1e-6)
    mu = raw / norm[..., None]
    kr = m[:, :, 3]
    kappa = np.exp(np.clip(kr, KR_MIN, KR_MAX))
    lg = m[:, :, 4]
    e = np.exp(lg - lg.max(axis=1, keepdims=True))
    pi = e / e.sum(axis=1, keepdims=True)
    sig = 1.0 / (1.0 + np.exp(-o[:, ALPHA_UNIT]))
    alpha = ALPHA_MAX * sig
    c = np.einsum("njk,nk->nj", mu, w)
    v = vmf(mu, kappa, w)
    q = (pi * v).sum(axis=1)
    qm = alpha * q + (1 - alpha) * p_bsdf
    qu = (1 - Q_DEFENSIVE) * q + Q_DEFENSIVE / (4 * np.pi)
    qm_loss = qm if q_fixed is None else alpha * q_fixed + (1 - alpha) * p_bsdf
    loss = float(np.mean(-weight * (np.log(qu) + np.log(qm_loss))))
    g_q = -weight * (1 - Q_DEFENSIVE) / qu / n     # d loss / d q (guide objective)
    d_o = np.zeros_like(o)
    d_o[:, ALPHA_UNIT] = -weight / qm / n * (q - p_bsdf) * ALPHA_MAX * sig * (1 - sig)
    if q_fixed is not None:
        d_o[:, ALPHA_UNIT] = -weight / qm_loss / n * (q_fixed - p_bsdf) * ALPHA_MAX * sig * (1 - sig)
    d_lg = g_q[:, None] * pi * (v - q[:, None])
    em = np.exp(-2 * kappa)
    dlogv_dk = 1.0 / kappa - 2 * em / (-np.expm1(-2 * kappa)) + (c - 1.0)
    inside = (kr > KR_MIN) & (kr < KR_MAX)
    d_kr = g_q[:, None] * pi * v * dlogv_dk * kappa * inside
    d_mu = (g_q[:, None] * pi * v * kappa)[..., None] * w[:, None, :]   # dq/dmu
    # through mu = raw / |raw|: (I - mu mu^T) / |raw| (zero below the clamp)
    d_raw = (d_mu - np.einsum("njk,njk->nj", d_mu, mu)[..., None] * mu) / norm[..., None]
    d_raw = d_raw * (rn > 1e-6)[..., None]
    dm = np.zeros((n, N_LOBES, 5))
    dm[:, :, :3] = d_raw
    dm[:, :, 3] = d_kr
    dm[:, :, 4] = d_lg
    d_o[:, OUT_UNIT] = dm
    g = {}
    g["W3"] = d_o.T @ u3
    g["b3"] = d_o.sum(0)
    d_u3 = d_o @ P["W3"]
    d_h3 = np.zeros_like(d_u3)
    d_h3[:, PHI] = d_u3
    d_z3 = d_h3 * (u3[:, np.argsort(PHI)] > 0)
    g["W2"] = d_z3.T @ u2
    g["b2"] = d_z3.sum(0)
    d_u2 = d_z3 @ P["W2"]
    d_h2 = np.zeros_like(d_u2)
    d_h2[:, PHI] = d_u2
    d_z2 = d_h2 * (u2[:, np.argsort(PHI)] > 0)
    g["W1"] = d_z2.T @ u1
    g["b1"] = d_z2.sum(0)
    d_u1 = d_z2 @ P["W1"]
    d_h1 = np.zeros_like(d_u1)
    d_h1[:, PHI] = d_u1
    d_z1 = d_h1 * (u1[:, np.argsort(PHI)] > 0)
    g["W0"] = d_z1.T @ f
    g["b0"] = d_z1.sum(0)
    flat = np.concatenate([g[name].ravel() for name, _ in SHAPES])
    return loss, flat * live_mask()


def sample_vmf(mu, kappa, u0, u1):
    """Direction with density vMF(mu, kappa) from two uniforms (vectorised)."""
    e = np.exp(-2 * kappa)
    cw = 1.0 + np.log(u0 + (1 - u0) * e) / kappa
    sw = np.sqrt(np.maximum(0.0, 1 - cw * cw))
    ph = 2 * np.pi * u1
    # frame around mu (any orthonormal frame gives the same distribution)
    a = np.where(np.abs(mu[..., :1]) > 0.9, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    t = np.cross(mu, a)
    t /= np.linalg.norm(t, axis=-1, keepdims=True)
    b = np.cross(mu, t)
    return (sw * np.cos(ph))[..., None] * t + (sw * np.sin(ph))[..., None] * b + cw[..., None] * mu
