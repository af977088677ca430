"""The surrogate MLP's BF16 MFMA kernels against the CPU oracle (oracle/mlp_oracle.py):
forward pass, gradient of the loss, and a training run that must reduce the loss."""
import numpy as np
import pytest

from lgame.gpu import Mlp
from mlp_oracle import MlpOracle

pytestmark = pytest.mark.gpu


def _net(n_in=37, hidden=96, nh=3, n_out=11, seed=4):
    m = Mlp(n_in, hidden, nh, n_out, seed=seed)
    rng = np.random.default_rng(seed)
    xm, xs = rng.standard_normal(n_in).astype(np.float32), (0.5 + rng.random(n_in)).astype(np.float32)
    ym, ys = rng.standard_normal(n_out).astype(np.float32), (0.5 + rng.random(n_out)).astype(np.float32)
    m.set_norm(xm, xs, ym, ys)
    o = MlpOracle(n_in, hidden, nh, n_out, m.get_params(), xm, xs, ym, ys, bf16=True)
    return m, o, (xm, xs, ym, ys), rng


def test_mlp_forward_matches_oracle():
    m, o, (xm, xs, ym, ys), rng = _net()
    x = (xm + xs * rng.standard_normal((200, 37))).astype(np.float32)
    yg = m.forward(x)
    yo = o.forward(x)
    err = np.abs(yg - yo) / ys
    print(f"forward max normalised error {err.max():.3e}")
    assert err.max() < 2e-2
    # the device is not computing something unrelated: the outputs vary with the inputs
    assert np.std(yg[:, 0]) > 1e-3


def test_mlp_gradient_matches_oracle():
    m, o, (xm, xs, ym, ys), rng = _net()
    x = (xm + xs * rng.standard_normal((256, 37))).astype(np.float32)
    y = (ym + ys * rng.standard_normal((256, 11))).astype(np.float32)
    gg, lg = m.gradient(x, y)
    lo, go = o.loss_and_grad(x, y)
    rel = np.linalg.norm(gg - go) / np.linalg.norm(go)
    cos = float(np.dot(gg, go) / (np.linalg.norm(gg) * np.linalg.norm(go)))
    print(f"gradient relative error {rel:.3e}, cosine {cos:.6f}, loss {lg:.6f} vs {lo:.6f}")
    assert abs(lg - lo) <= 1e-2 * lo
    assert rel < 3e-2
    assert cos > 0.999
    # per parameter block (each weight matrix and bias)
    sizes = [37, 96, 96, 96, 11]
    off = 0
    for i, oo in zip(sizes[:-1], sizes[1:]):
        for n in (i * oo, oo):
            a, b = gg[off:off + n], go[off:off + n]
            assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b) + 1e-7
            off += n


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(1)
    n_in, n_out = 24, 8
    m = Mlp(n_in, 128, 2, n_out, seed=2)
    x = rng.standard_normal((16384, n_in)).astype(np.float32)
    w = rng.standard_normal((n_in, n_out)).astype(np.float32) / np.sqrt(n_in)
    y = np.tanh(x @ w) + 0.1 * np.sin(3.0 * x[:, :n_out])
    m.set_norm(np.zeros(n_in), np.ones(n_in), np.zeros(n_out), np.ones(n_out))
    _, loss0 = m.gradient(x[:4096], y[:4096].astype(np.float32))
    m.load_data(x, y.astype(np.float32))
    loss, sec = m.train(epochs=8, batch=512, lr=2e-3, seed=3)
    print(f"untrained loss {loss0:.5f}, loss per epoch {np.round(loss, 5)} in {sec:.3f} s")
    assert np.all(np.isfinite(loss))
    assert loss[-1] < 0.1 * loss0
    assert loss[-1] < 0.5 * loss[0]
    pred = m.forward(x[:1024])
    assert float(np.mean((pred - y[:1024]) ** 2)) < 0.2 * float(np.mean(y[:1024] ** 2))
