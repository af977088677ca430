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
from lgame.gpu import RAD_OUT_LW_OUT, RAD_OUT_SW_IN, RAD_OUT_TENDENCY, Step

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


def _surrogate_step(m, cfg, g, st, sf, tol=1.0, frac=0.01, seed=5):
    s = Step(dataclasses.replace(cfg, rad_mode=RAD_SURROGATE), g)
    s.start(st)
    s.set_rad_surface(sf)
    s.attach_mlp(m, tol_k_per_day=tol, audit_fraction=frac, seed=seed)
    return s


def _teacher(cfg, g, st, sf):
    t = Step(dataclasses.replace(cfg, rad_mode=RAD_TEACHER), g)
    t.start(st)
    t.set_rad_surface(sf)
    t.run(0, 1, t_init=surrogate.T_EPOCH + 3 * 3600.0)
    out = (t.rad_get(RAD_OUT_TENDENCY), t.rad_get(RAD_OUT_SW_IN), t.rad_get(RAD_OUT_LW_OUT))
    t.close()
    return out


def _case(nl=60, n=33):
    cfg = moist3d_config(n=n, n_layers=nl, dtime=3.22, seed=11)
    cfg = dataclasses.replace(cfg, lrad=True, dtime_rad=4 * cfg.dt)
    g, st = make_case(cfg, seed=12, cloud_frac=0.004)
    return cfg, g, st, radiation_surface(cfg, g, st, seed=13)


def test_surrogate_envelope_sends_every_column_to_the_solver(trained):
    """Predictions far outside the training envelope (here: the output bias of the last layer
    shifted by 50 standard deviations) go back to the solver even with no audit sample: every
    column falls back and the radiation equals the solver's."""
    m = trained[0]
    cfg, g, st, sf = _case()
    p = m.get_params()
    p_shift = p.copy()
    p_shift[-m.n_out:] += 50.0
    m.set_params(p_shift)
    try:
        s = _surrogate_step(m, cfg, g, st, sf, tol=1.0, frac=0.0)
        s.run(0, 1, t_init=surrogate.T_EPOCH + 3 * 3600.0)
        rt = s.rad_get(RAD_OUT_TENDENCY)
        stats = s.rad_stats()
        s.close()
    finally:
        m.set_params(p)
    rt_t = _teacher(cfg, g, st, sf)[0]
    print(f"stats with the output shifted out of the envelope: {stats}")
    assert stats["fallbacks"] == cfg.ny * cfg.nx
    assert stats["audited"] == 0
    np.testing.assert_array_equal(rt, rt_t)


def test_attach_needs_the_normalisation():
    from lgame.gpu import LgameError, Mlp

    cfg, g, st, sf = _case(n=17)
    m = Mlp(surrogate.N_IN, 64, 1, surrogate.N_OUT, seed=1)
    s = Step(dataclasses.replace(cfg, rad_mode=RAD_SURROGATE), g)
    s.set_rad_surface(sf)
    with pytest.raises(LgameError, match="normalisation"):
        s.attach_mlp(m)
    s.close()
    m.close()


@pytest.mark.parametrize("nl", [40, 70])
def test_surrogate_on_other_layer_counts(trained, nl):
    """A domain of another layer count: its columns are resampled to the network's 60 levels and
    the heating rates back to its layers.  With every column audited at a zero tolerance the
    radiation is exactly the solver's (the audit list and the resampled comparison cover every
    column); at the working tolerance the in-step heating stays close to the solver's."""
    m = trained[0]
    cfg, g, st, sf = _case(nl=nl)
    rt_t, sw_t, lw_t = _teacher(cfg, g, st, sf)
    s = _surrogate_step(m, cfg, g, st, sf, tol=1e-9, frac=1.0)
    s.run(0, 1, t_init=surrogate.T_EPOCH + 3 * 3600.0)
    np.testing.assert_array_equal(s.rad_get(RAD_OUT_TENDENCY), rt_t)
    np.testing.assert_array_equal(s.rad_get(RAD_OUT_SW_IN), sw_t)
    s.close()
    s = _surrogate_step(m, cfg, g, st, sf, tol=1.0, frac=0.05)
    s.run(0, 1, t_init=surrogate.T_EPOCH + 3 * 3600.0)
    rt = s.rad_get(RAD_OUT_TENDENCY)
    stats = s.rad_stats()
    assert s.count_nonfinite(1) == 0
    s.close()
    rho_g = st["rho"][:, :, :, cfg.n_condensed_constituents]
    err = np.sqrt(np.mean(((rt - rt_t) * 86400.0 / (rho_g * 1005.0)) ** 2))
    print(f"{nl} layers: in-step heating RMSE vs solver {err:.3f} K/day, stats {stats}")
    assert stats["calls"] == 1 and stats["audited"] > 0
    assert err < 2.5


def test_two_handles_share_one_network(trained):
    """Two domains attached to one network update their radiation at the same time on their own
    streams; each gets what it gets alone (the inference buffers are the handles' own)."""
    m = trained[0]
    cases = [_case(n=33), _case(n=29)]
    alone = []
    for cfg, g, st, sf in cases:
        s = _surrogate_step(m, cfg, g, st, sf)
        s.run(0, 2, t_init=surrogate.T_EPOCH + 3 * 3600.0)
        alone.append(s.get_state(0))
        s.close()
    steps = [_surrogate_step(m, cfg, g, st, sf) for cfg, g, st, sf in cases]
    for k in range(2):  # interleaved, no synchronisation between the handles
        for s in steps:
            t0 = surrogate.T_EPOCH + 3 * 3600.0 + k * s.cfg.dt
            s.pchevi(k % 2, 1 - k % 2, k, k == 0, t0)
    for s, want in zip(steps, alone):
        got = s.get_state(0)
        assert all(np.array_equal(got[f], want[f]) for f in want)
        s.close()
