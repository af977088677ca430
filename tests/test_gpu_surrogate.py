"""The learned radiation surrogate end to end: columns harvested from the model's own radiation
calls (gray two-stream solver as teacher), a 304 -> 3x512 GELU -> 62 network trained on the GPU
with the project's BF16 MFMA kernels, its accuracy against the method without it, and its use in
the step with the audit and the fallback to the solver."""
import dataclasses

import numpy as np
import pytest

from lgame import surrogate
from lgame.cases import make_case, moist3d_config, radiation_surface
from lgame.config import RAD_SURROGATE, RAD_TEACHER
from lgame.gpu import RAD_OUT_TENDENCY, Step

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    print("harvesting training columns", flush=True)
    xtr, ytr = surrogate.harvest(n=33, n_states=10, calls=2, times=6, seed=0)
    xte, yte = surrogate.harvest(n=33, n_states=2, calls=1, times=4, seed=1)
    print(f"training on {len(xtr)} columns", flush=True)
    norm = surrogate.fit_norm(xtr, ytr)
    m, loss, sec = surrogate.train(xtr, ytr, epochs=40, batch=512, lr=1e-3, norm=norm)
    print(f"loss {loss[0]:.4f} -> {loss[-1]:.5f} in {sec:.2f} s", flush=True)
    return m, norm, (xtr, ytr, xte, yte), loss


def test_surrogate_beats_the_method_without_it(trained):
    m, norm, (xtr, ytr, xte, yte), loss = trained
    pred = m.forward(xte)
    rmse = surrogate.heating_rmse(pred, yte)
    base = surrogate.baseline_rmse(ytr, yte)
    zero = surrogate.heating_rmse(np.zeros_like(yte), yte)
    print(f"held-out heating-rate RMSE: network {rmse:.3f} K/day, mean profile {base:.3f}, no radiation {zero:.3f}")
    assert np.all(np.isfinite(pred))
    assert loss[-1] < 0.2 * loss[0]
    assert rmse < 0.25 * base
    assert rmse < 0.25 * zero


def _pair(m, tol, frac):
    cfg = moist3d_config(n=33, n_layers=60, dtime=3.22, seed=11)
    cfg = dataclasses.replace(cfg, lrad=True, dtime_rad=4 * cfg.dt)
    g, st = make_case(cfg, seed=12, cloud_frac=0.004)
    sf = radiation_surface(cfg, g, st, seed=13)
    out = []
    for mode in (RAD_TEACHER, RAD_SURROGATE):
        s = Step(dataclasses.replace(cfg, rad_mode=mode), g)
        s.start(st)
        s.set_rad_surface(sf)
        if mode == RAD_SURROGATE:
            s.attach_mlp(m, tol_k_per_day=tol, audit_fraction=frac, seed=5)
        s.run(0, 1, t_init=surrogate.T_EPOCH + 3 * 3600.0)
        out.append((s, s.rad_get(RAD_OUT_TENDENCY)))
    rho_g = st["rho"][:, :, :, cfg.n_condensed_constituents]
    return out, rho_g


def test_surrogate_in_the_step(trained):
    m = trained[0]
    [(tch, rt_t), (sur, rt_s)], rho_g = _pair(m, tol=1.0, frac=0.05)
    k_per_day = 86400.0 / (rho_g * 1005.0)
    err = np.sqrt(np.mean(((rt_s - rt_t) * k_per_day) ** 2))
    st = sur.rad_stats()
    print(f"in-step heating RMSE vs solver {err:.3f} K/day, stats {st}")
    assert st["calls"] == 1
    assert 0 < st["audited"] < 33 * 33
    assert st["audit_rmse_k_per_day"] < 1.0
    assert err < 1.0
    assert sur.count_nonfinite(1) == 0


def test_surrogate_audit_falls_back_to_the_solver(trained):
    m = trained[0]
    [(tch, rt_t), (sur, rt_s)], _ = _pair(m, tol=1e-9, frac=1.0)
    st = sur.rad_stats()
    print(f"stats with a zero tolerance and every column audited: {st}")
    assert st["audited"] == 33 * 33 and st["fallbacks"] == 33 * 33
    np.testing.assert_array_equal(rt_s, rt_t)
