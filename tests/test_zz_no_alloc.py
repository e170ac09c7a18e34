"""pg_guide_train_static, the guide-training entry point that allocates no
device memory (its partial-gradient rows live in a static device array), checked
through the C ABI against pg_guide_train, whose kernels it shares: the same
parameters bit for bit and the same loss, on two calls in a row, the second
with a smaller input (so stale rows left in the shared array by the first call
must not leak into the second)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def training_set(rng, n):
    x = np.zeros((n, 9), np.float32)
    x[:, :3] = rng.uniform(0, 1, (n, 3))
    for a in (3, 6):
        v = rng.normal(size=(n, 3))
        x[:, a:a + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    w = rng.uniform(0.1, 2.0, n).astype(np.float32)
    pb = rng.uniform(0.05, 1.0, n).astype(np.float32)
    return x, d, w, pb


def test_static_training_equals_allocating_training(pg, cuda):
    hip = cuda
    rng = np.random.default_rng(11)
    ga = pg.Guide(seed=5, learning_rate=3e-3)
    gb = pg.Guide(seed=5, learning_rate=3e-3)
    d_loss = hip.DeviceArray((1,), np.float32)
    # Second call: fewer samples, a smaller batch with a partial last workgroup.
    for n, steps, batch in ((5000, 3, 2048), (300, 2, 160)):
        print(f"call: n={n} steps={steps} batch={batch}", flush=True)
        arrs = [hip.DeviceArray.from_numpy(a) for a in training_set(rng, n)]
        loss_a = ga.train(*arrs, steps=steps, batch=batch)
        gb.train_static(*arrs, steps=steps, batch=batch, d_loss=d_loss)
        hip.synchronize()
        pa, pb = ga.params(), gb.params()
        assert np.array_equal(pa.view(np.int32), pb.view(np.int32)), \
            f"{int((pa != pb).sum())} parameters differ"
        loss_b = float(d_loss.numpy()[0])
        assert np.float32(loss_a).view(np.int32) == np.float32(loss_b).view(np.int32), (loss_a, loss_b)
        print(f"  same {pa.size} parameters, loss {loss_b:.6f}")
    ga.release()
    gb.release()


def test_static_training_rejects_batches_beyond_its_array(pg, cuda):
    hip = cuda
    rng = np.random.default_rng(3)
    g = pg.Guide(seed=1)
    arrs = [hip.DeviceArray.from_numpy(a) for a in training_set(rng, 64)]
    with pytest.raises(pg.PineError, match="PG_GUIDE_STATIC_MAX_BATCH"):
        g.train_static(*arrs, steps=1, batch=pg.GUIDE_STATIC_MAX_BATCH + 1)
    g.release()
