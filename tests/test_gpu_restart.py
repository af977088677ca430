"""Fault detection and restarts through the C-ABI.

* The step's own non-finite check: after an unstable step (the Schaer case at dt 4 s, which the
  reference scheme also blows up) lgame_sync reports LGAME_ENONFINITE, as the reference's output
  check stops the run (mo_write_out.F90:63-67); a new state clears it.
* A restart from the two states plus the checkpoint of the carried fields continues the run bit
  for bit; without the checkpoint (the reference's own restart carries the state only) it does
  not.
"""
import dataclasses

import numpy as np
import pytest

from lgame import surrogate
from lgame.cases import make_case, moist3d_config, radiation_surface, schaer_config, tiny_config
from lgame.config import PREC_MIXED, RAD_SURROGATE, RAD_TEACHER
from lgame.gpu import LgameError, Step

pytestmark = pytest.mark.gpu
T0 = surrogate.T_EPOCH + 3 * 3600.0


def test_unstable_step_is_reported_at_sync():
    cfg = schaer_config(nx=201, n_layers=40, dtime=4.0)
    g, st = make_case(cfg)
    s = Step(cfg, g)
    s.start(st)
    s.run(0, 30)
    with pytest.raises(LgameError, match="non-finite"):
        s.sync()
    assert s.count_nonfinite(0) > 0
    # a new state clears the flag; a stable run keeps it clear
    s2 = Step(schaer_config(nx=201, n_layers=40), g)
    s2.start(st)
    s2.run(0, 30)
    s2.sync()
    s.start(st)
    s.sync()
    s.close()
    s2.close()


def _physics(cfg):
    return dataclasses.replace(cfg, lpbl=True, lprog_soil_temp=True, lsfc_sensible_heat_flux=True,
                               lsfc_phase_trans=True).set_diffusion()


def _restart_case(name):
    if name == "moist_physics_radiation":
        cfg = _physics(tiny_config(ny=9, nx=11, n_layers=8, moist=True, order=3, freeslip=True))
        cfg = dataclasses.replace(cfg, lrad=True, rad_mode=RAD_TEACHER, dtime_rad=3 * 2.0)
        g, st = make_case(cfg, cloud_frac=0.3)
        return cfg, g, st, radiation_surface(cfg, g, st), None
    if name == "mixed_surrogate":
        cfg = moist3d_config(n=33, n_layers=60, dtime=3.22, seed=21)
        cfg = dataclasses.replace(cfg, lrad=True, rad_mode=RAD_SURROGATE, dtime_rad=3 * cfg.dt, precision=PREC_MIXED)
        g, st = make_case(cfg, cloud_frac=0.01)
        return cfg, g, st, radiation_surface(cfg, g, st), surrogate.load(surrogate.DEFAULT_WEIGHTS)
    cfg = tiny_config(order=3, corio=True)
    g, st = make_case(cfg)
    return cfg, g, st, None, None


def _handle(cfg, g, sf, mlp):
    s = Step(cfg, g)
    if sf is not None:
        s.set_rad_surface(sf)
    if mlp is not None:
        s.attach_mlp(mlp, tol_k_per_day=1.0, audit_fraction=0.05, seed=3)
    return s


@pytest.mark.parametrize("name", ["dry", "moist_physics_radiation", "mixed_surrogate"])
def test_restart_from_checkpoint_is_bitwise(name):
    cfg, g, st, sf, mlp = _restart_case(name)
    k, m = 5, 6
    print(f"case {name}: {k} steps, checkpoint, {m} more", flush=True)
    a = _handle(cfg, g, sf, mlp)
    a.start(st)
    a.run(0, k, t_init=T0)
    slots = [a.get_state(0), a.get_state(1)]
    blob = a.checkpoint()
    a.run(k, m, t_init=T0)
    want = a.get_state((k + m) % 2)
    a.close()
    b = _handle(cfg, g, sf, mlp)
    b.set_state(0, slots[0])
    b.set_state(1, slots[1])
    b.restore(blob)
    b.run(k, m, t_init=T0)
    got = b.get_state((k + m) % 2)
    b.close()
    print(f"checkpoint of {len(blob) / 1e6:.2f} MB")
    assert all(np.array_equal(got[f], want[f]) for f in want)
    # the state alone (the reference's restart) does not continue the same run
    c = _handle(cfg, g, sf, mlp)
    c.set_state(0, slots[0])
    c.set_state(1, slots[1])
    c.run(k, m, t_init=T0)
    other = c.get_state((k + m) % 2)
    c.close()
    assert not all(np.array_equal(other[f], want[f]) for f in want)


def test_checkpoint_of_another_configuration_is_refused():
    cfg, g, st, _, _ = _restart_case("dry")
    a = Step(cfg, g)
    a.start(st)
    blob = a.checkpoint()
    cfg2 = dataclasses.replace(cfg, lmoist=True)
    g2, st2 = make_case(cfg2)
    b = Step(cfg2, g2)
    b.start(st2)
    with pytest.raises(LgameError, match="configuration"):
        b.restore(blob)
    with pytest.raises(LgameError, match="truncated|too short|configuration"):
        a.restore(blob[:-8])
    a.close()
    b.close()
