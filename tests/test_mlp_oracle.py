"""The MLP oracle's backward pass against central finite differences (FP64, no rounding), and its
bfloat16 rounding against the IEEE definition."""
import numpy as np

from mlp_oracle import MlpOracle, bf16


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    sizes = (5, 7, 2, 3)
    n_par = 5 * 7 + 7 + 7 * 7 + 7 + 7 * 3 + 3
    p = rng.standard_normal(n_par) * 0.5
    x = rng.standard_normal((9, 5))
    y = rng.standard_normal((9, 3))
    o = MlpOracle(*sizes, p, bf16=False)
    loss, g = o.loss_and_grad(x, y)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd[i] = (MlpOracle(*sizes, pp, bf16=False).loss_and_grad(x, y)[0] -
                 MlpOracle(*sizes, pm, bf16=False).loss_and_grad(x, y)[0]) / (2 * h)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)
    assert loss > 0


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 65504.0, 1e-30], dtype=np.float32)
    r = bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # halfway, ties to even (mantissa of 1.0 is even)
    assert r[2] == 1.0 + 2 ** -7            # above halfway rounds up
    assert abs(r[3] + 3.140625) < 1e-12
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)
