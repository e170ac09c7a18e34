"""The path-guiding network's GPU kernels (MFMA forward and backward, Adam)
against the float64 oracle (oracle/guide.py), through the C ABI.

Bars (BF16 operands, FP32 accumulation, against the oracle evaluated at the
same BF16-rounded weights):
  - raw outputs within 2% of their range; mixture pdf within 3%;
  - gradient per layer: cosine similarity > 0.995, norm within 3%;
  - sampling: the returned pdf equals the pdf kernel's; the mixture integrates
    to 1 and sample moments match the pdf's quadrature (within 4 sigma);
  - training reduces the cross-entropy towards a known target distribution.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bf16(a):
    """Round-to-nearest-even to BF16, returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def bf16_weights(G, params):
    """The parameter vector the GPU forward sees: weights in BF16, biases FP32."""
    p = np.array(params, np.float64)
    o = 0
    for name, shp in G.SHAPES:
        n = int(np.prod(shp))
        if name.startswith("W"):
            p[o:o + n] = bf16(p[o:o + n])
        o += n
    return p


def inputs(rng, n):
    x = np.zeros((n, 9))
    x[:, :3] = rng.uniform(0, 1, (n, 3))
    for a in (3, 6):
        v = rng.normal(size=(n, 3))
        x[:, a:a + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return x.astype(np.float32)


def unit(rng, n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


@pytest.fixture(scope="module")
def G():
    from oracle import guide
    return guide


def make_params(G, seed):
    rng = np.random.default_rng(seed)
    p = G.init_params(seed)
    # A livelier output layer than the initialisation, so lobes differ.
    u = G.unpack(p)
    o = G.N_PARAMS - 48 * 64 - 48
    p[o:o + 48 * 64] += rng.normal(scale=0.15, size=48 * 64)
    p[-48:] += rng.normal(scale=0.5, size=48)
    return (p * G.live_mask()).astype(np.float32)


def test_forward_and_pdf_match_oracle(pg, cuda, G):
    rng = np.random.default_rng(0)
    params = make_params(G, 1)
    g = pg.Guide()
    g.set_params(params)
    n = 1000
    x = inputs(rng, n)
    dx = cuda.DeviceArray.from_numpy(x)
    o = g.forward(dx).numpy()
    ref, _ = G.forward(bf16_weights(G, params), x.astype(np.float64))
    err = np.abs(o[:, :41] - ref[:, :41]).max() / np.abs(ref[:, :41]).max()
    print(f"forward max error / range {err:.2e}")
    assert err < 0.02
    assert np.corrcoef(o[:, :41].ravel(), ref[:, :41].ravel())[0, 1] > 0.9995
    w = unit(rng, n)
    pb = rng.uniform(0.05, 0.3, n).astype(np.float32)
    out = g.pdf(dx, cuda.DeviceArray.from_numpy(w), cuda.DeviceArray.from_numpy(pb)).numpy()
    q, alpha, qm = G.pdf(bf16_weights(G, params), x.astype(np.float64), w.astype(np.float64), pb)
    rel = np.abs(out[:, 2] - qm) / qm
    print(f"pdf max rel error {rel.max():.2e}, alpha max error {np.abs(out[:, 1] - alpha).max():.2e}")
    assert rel.max() < 0.03 and np.abs(out[:, 1] - alpha).max() < 0.01
    g.release()


def test_gradient_matches_oracle(pg, cuda, G):
    rng = np.random.default_rng(1)
    params = make_params(G, 2)
    g = pg.Guide()
    g.set_params(params)
    n = 1000  # not a multiple of the 64-sample workgroup tile
    x, w = inputs(rng, n), unit(rng, n)
    wt = rng.uniform(0.1, 2.0, n).astype(np.float32)
    wt[::17] = 0.0  # samples without radiance
    pb = rng.uniform(0.05, 0.3, n).astype(np.float32)
    d = [cuda.DeviceArray.from_numpy(a) for a in (x, w, wt, pb)]
    grad, loss = g.gradient(*d)
    ref_loss, ref = G.loss_and_grad(bf16_weights(G, params), x.astype(np.float64),
                                    w.astype(np.float64), wt.astype(np.float64), pb.astype(np.float64))
    print(f"loss gpu {loss:.6f} oracle {ref_loss:.6f}")
    assert abs(loss - ref_loss) < 0.01 * abs(ref_loss)
    assert np.all(grad[G.live_mask() == 0] == 0)
    o = 0
    for name, shp in G.SHAPES:
        k = int(np.prod(shp))
        a, b = grad[o:o + k].astype(np.float64), ref[o:o + k]
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30)
        nr = np.linalg.norm(a) / (np.linalg.norm(b) + 1e-30)
        print(f"{name}: cosine {cos:.5f}, norm ratio {nr:.4f}")
        assert cos > 0.995 and abs(nr - 1) < 0.03, name
        o += k
    # Deterministic: the same batch gives the same bits.
    grad2, _ = g.gradient(*d)
    assert np.array_equal(grad, grad2)
    g.release()


def test_sampling_matches_pdf(pg, cuda, G):
    rng = np.random.default_rng(2)
    params = make_params(G, 3)
    g = pg.Guide()
    g.set_params(params)
    n = 1 << 18
    x = np.repeat(inputs(rng, 1), n, axis=0)
    dx = cuda.DeviceArray.from_numpy(x)
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    s = g.sample(dx, cuda.DeviceArray.from_numpy(u)).numpy()
    dirs = s[:, :3]
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1, atol=1e-4)
    pb = np.full(n, 0.1, np.float32)
    pdfs = g.pdf(dx, cuda.DeviceArray.from_numpy(np.ascontiguousarray(dirs)),
                 cuda.DeviceArray.from_numpy(pb)).numpy()
    np.testing.assert_allclose(s[:, 3], pdfs[:, 0], rtol=1e-4, atol=1e-7)
    # Normalisation and moments by uniform-sphere quadrature.
    uni = unit(rng, n)
    qu = g.pdf(dx, cuda.DeviceArray.from_numpy(uni), cuda.DeviceArray.from_numpy(pb)).numpy()[:, 0]
    total = 4 * np.pi * qu.mean()
    print(f"integral of q over the sphere {total:.4f}")
    assert abs(total - 1) < 0.02
    for a in (np.array([1.0, 0.3, -0.2]), np.array([0.0, 0.0, 1.0])):
        a = a / np.linalg.norm(a)
        h_s = dirs @ a
        h_q = 4 * np.pi * np.mean((uni @ a) * qu)
        sig = h_s.std() / np.sqrt(n) + 4 * np.pi * np.std((uni @ a) * qu) / np.sqrt(n)
        print(f"E[w.a]: samples {h_s.mean():.4f}, quadrature {h_q:.4f} (sigma {sig:.4f})")
        assert abs(h_s.mean() - h_q) < 4 * sig + 1e-3
    g.release()


def test_training_learns_a_target_distribution(pg, cuda, G):
    """Samples drawn from a position-dependent vMF (kappa 20); weights 1, a
    uniform BSDF.  Adam on the GPU must cut the cross-entropy well below its
    starting value, towards the target's entropy."""
    rng = np.random.default_rng(3)
    n = 1 << 16
    x = inputs(rng, n)
    mu = np.stack([x[:, 0] - 0.5, np.ones(n), x[:, 2] - 0.5], 1)
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    kappa = np.full((n, 1), 20.0)
    w = G.sample_vmf(mu[:, None, :], kappa, rng.uniform(size=(n, 1)), rng.uniform(size=(n, 1)))[:, 0]
    w = w.astype(np.float32)
    wt = np.ones(n, np.float32)
    pb = np.full(n, 1.0 / (4 * np.pi), np.float32)
    d = [cuda.DeviceArray.from_numpy(a) for a in (x, w, wt, pb)]
    g = pg.Guide(seed=5, learning_rate=5e-3)
    _, l0 = g.gradient(*d)
    l1 = g.train(*d, steps=400, batch=4096)
    _, l2 = g.gradient(*d)
    # Two cross-entropy terms (guide with a defensive uniform share, mixture);
    # each starts near log(4 pi) = 2.53 for a broad guide.
    print(f"loss: start {l0:.3f}, last batch {l1:.3f}, end {l2:.3f}")
    assert l2 < l0 - 3.0
    assert l2 < 1.2
    g.release()
