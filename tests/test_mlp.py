"""Surrogate MLP: the NumPy oracle is pinned by finite differences (CPU); the GPU BF16 MFMA
forward pass and gradients match the oracle run on the same BF16-rounded operands, and training
on the GPU reduces the loss on a fixed regression task."""
import numpy as np
import pytest

from tests import mlp_oracle as MO


def _small_layers(rng, dims):
    out = []
    for din, dout in dims:
        out.append((rng.normal(0, 1.0 / np.sqrt(din), (dout, din)), rng.normal(0, 0.1, dout)))
    return out


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    layers = _small_layers(rng, [(5, 7), (7, 6), (6, 3)])
    x = rng.normal(size=(4, 5))
    y = rng.uniform(size=(4, 3))
    loss, grads = MO.loss_and_grads(layers, x, y)
    eps = 1e-6
    for li, (W, b) in enumerate(layers):
        for arr, g in [(W, grads[li][0]), (b, grads[li][1])]:
            it = np.nditer(arr, flags=["multi_index"])
            for _ in it:
                idx = it.multi_index
                old = arr[idx]
                arr[idx] = old + eps
                lp, _ = MO.loss_and_grads(layers, x, y)
                arr[idx] = old - eps
                lm, _ = MO.loss_and_grads(layers, x, y)
                arr[idx] = old
                assert abs((lp - lm) / (2 * eps) - g[idx]) <= 1e-7 + 1e-5 * abs(g[idx])


def test_bf16_rounding_matches_bit_layout():
    v = np.array([1.0, 1.00390625, 1.005859375, 3.0e38, -2.5], np.float32)
    r = MO.to_bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # halfway, rounds to even
    assert r[2] == 1.0078125
    assert r[4] == -2.5


def _mlp(seed=0, n_in=100, hidden=128, layers=2, n_out=64, batch=128):
    from upmabossg_amd.mlp import Mlp

    return Mlp(n_in, hidden, layers, n_out, max_batch=batch, seed=seed)


@pytest.mark.gpu
def test_gpu_forward_matches_oracle():
    m = _mlp()
    rng = np.random.default_rng(1)
    x = rng.normal(size=(200, 100)).astype(np.float32)
    y, _ = m.predict(x)
    p = m.get_params()
    layers = [(W.astype(np.float64), b.astype(np.float64)) for W, b in m.unpack(p)]
    xp = np.zeros((200, m.padded_inputs), np.float32)
    xp[:, :100] = x
    want, _ = MO.forward(layers, xp, bf16=True)
    assert np.max(np.abs(y - want)) < 2e-3


@pytest.mark.gpu
def test_gpu_gradients_match_oracle():
    m = _mlp(seed=3)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(128, 100)).astype(np.float32)
    y = rng.uniform(size=(128, 64)).astype(np.float32)
    loss, g = m.gradients(x, y)
    layers = [(W.astype(np.float64), b.astype(np.float64)) for W, b in m.unpack(m.get_params())]
    xp = np.zeros((128, m.padded_inputs), np.float32)
    xp[:, :100] = x
    want_loss, want = MO.loss_and_grads(layers, xp, y, bf16=True)
    assert abs(loss - want_loss) <= 1e-4 * want_loss + 1e-7
    got = m.unpack(g)
    for (gw, gb), (ww, wb) in zip(got, want):
        for a, b in [(gw, ww), (gb, wb)]:
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = float(np.dot(a.ravel(), b.ravel()) / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
            assert rel < 0.02 and cos > 0.9995, (rel, cos)


@pytest.mark.gpu
def test_gpu_training_reduces_loss():
    m = _mlp(seed=5, n_in=32, hidden=128, layers=2, n_out=64, batch=128)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, size=(4096, 32)).astype(np.float32)
    A = rng.normal(size=(32, 64)) / 4
    y = (1 / (1 + np.exp(-(np.sin(x @ A) * 2)))).astype(np.float32)
    m.load_dataset(x, y)
    first = m.train(1, batch=128, lr=2e-3, seed=1)
    last = m.train(30, batch=128, lr=2e-3, seed=2)
    assert last < 0.25 * first
    pred, _ = m.predict(x[:512])
    assert np.mean((pred - y[:512]) ** 2) < 0.25 * first


@pytest.mark.gpu
def test_gpu_ensemble_predict_matches_members():
    from upmabossg_amd.mlp import ensemble_predict

    members = [_mlp(seed=s, n_in=100, hidden=128, layers=2, n_out=64, batch=128) for s in (7, 8, 9)]
    rng = np.random.default_rng(5)
    x = rng.normal(size=(300, 100)).astype(np.float32)  # 300 rows: three batches, the last one ragged
    mean, spread, ms = ensemble_predict(members, x)
    P = np.stack([m.predict(x)[0] for m in members])
    assert ms > 0
    assert np.max(np.abs(mean - P.mean(axis=0))) < 1e-6
    assert np.max(np.abs(spread - P.std(axis=0).max(axis=1))) < 1e-6
    one, s1, _ = ensemble_predict(members[:1], x)
    assert np.array_equal(one, P[0]) and not s1.any()


@pytest.mark.gpu
def test_gpu_large_tiles_match_oracle():
    """8,192 rows: the forward and input-gradient GEMMs take the 128 x 128 and 128 x 64 tiles."""
    m = _mlp(seed=9, n_in=100, hidden=512, layers=2, n_out=512, batch=8192)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(8192, 100)).astype(np.float32)
    y = rng.uniform(size=(8192, 512)).astype(np.float32)
    pred, _ = m.predict(x)
    layers = [(W.astype(np.float64), b.astype(np.float64)) for W, b in m.unpack(m.get_params())]
    xp = np.zeros((8192, m.padded_inputs), np.float32)
    xp[:, :100] = x
    want, _ = MO.forward(layers, xp, bf16=True)
    assert np.max(np.abs(pred - want)) < 2e-3
    loss, g = m.gradients(x, y)
    want_loss, want_g = MO.loss_and_grads(layers, xp, y, bf16=True)
    assert abs(loss - want_loss) <= 1e-4 * want_loss + 1e-7
    for (gw, gb), (ww, wb) in zip(m.unpack(g), want_g):
        for a, b in [(gw, ww), (gb, wb)]:
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = float(np.dot(a.ravel(), b.ravel()) / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
            assert rel < 0.02 and cos > 0.9995, (rel, cos)
