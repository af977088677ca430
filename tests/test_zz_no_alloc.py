"""lgame_mlp_forward_noalloc, the surrogate's inference without any device allocation, against
lgame_mlp_forward: the same outputs on two calls in a row, the second with a smaller input (the
static scratch keeps the first call's contents)."""
import numpy as np
import pytest

from lgame.gpu import LgameError, Mlp

pytestmark = pytest.mark.gpu


def _net(n_in=304, hidden=512, nh=3, n_out=62, seed=9):
    m = Mlp(n_in, hidden, nh, n_out, seed=seed)
    rng = np.random.default_rng(seed)
    xm, xs = rng.standard_normal(n_in).astype(np.float32), (0.5 + rng.random(n_in)).astype(np.float32)
    ym, ys = rng.standard_normal(n_out).astype(np.float32), (0.5 + rng.random(n_out)).astype(np.float32)
    m.set_norm(xm, xs, ym, ys)
    return m, rng, xm, xs


def test_forward_noalloc_equals_forward():
    m, rng, xm, xs = _net()
    for n in (5000, 777):  # several passes of the scratch, then one partial pass
        x = (xm + xs * rng.standard_normal((n, 304))).astype(np.float32)
        want = m.forward(x)
        got = m.forward_noalloc(x)
        print(f"n={n}: max |noalloc - forward| {np.abs(got - want).max():.3e}, bitwise={np.array_equal(got, want)}")
        assert np.all(np.isfinite(got))
        np.testing.assert_array_equal(got, want)
    m.close()


def test_forward_noalloc_small_network_and_rejection():
    m, rng, xm, xs = _net(n_in=37, hidden=96, nh=2, n_out=11, seed=3)
    x = (xm + xs * rng.standard_normal((300, 37))).astype(np.float32)
    np.testing.assert_array_equal(m.forward_noalloc(x), m.forward(x))
    m.close()
    big = Mlp(40000, 64, 1, 8, seed=1)  # 64 rows of 40000 inputs (FP32 and BF16) exceed the scratch
    big.set_norm(np.zeros(40000), np.ones(40000), np.zeros(8), np.ones(8))
    with pytest.raises(LgameError, match="scratch"):
        big.forward_noalloc(np.zeros((64, 40000), dtype=np.float32))
    big.close()
