"""The radiation boundary on the GPU against the oracle: column inputs/outputs of the gray
two-stream solver, and whole steps with radiation updates (solver mode), FP64 form."""
import dataclasses

import numpy as np
import pytest

from golden_util import max_rel
from lgame.cases import make_case, radiation_surface, tiny_config
from lgame.config import RAD_TEACHER
from lgame.gpu import RAD_OUT_LW_OUT, RAD_OUT_SW_IN, RAD_OUT_TENDENCY, LgameError, Step
from pyoracle import CpuOracle

pytestmark = pytest.mark.gpu
T_EPOCH = 1576843200.0


def _case(moist, nl=12, cloud=0.02, **kw):
    cfg = dataclasses.replace(tiny_config(ny=7, nx=9, n_layers=nl, moist=moist), lrad=True, rad_mode=RAD_TEACHER,
                              dtime_rad=6.0, **kw)
    g, st = make_case(cfg, cloud_frac=cloud if moist else 0.0)
    return cfg, g, st, radiation_surface(cfg, g, st)


@pytest.mark.parametrize("moist", [False, True])
@pytest.mark.parametrize("hours", [0.0, 7.0])
def test_rad_columns_match_oracle(moist, hours):
    cfg, g, st, sf = _case(moist)
    print(f"rad columns moist={moist} hours={hours}", flush=True)
    t = T_EPOCH + 3600.0 * hours
    s = Step(cfg, g)
    s.start(st)
    s.set_rad_surface(sf)
    xg, yg = s.rad_columns(0, t)
    s.close()
    o = CpuOracle(cfg, g)
    o.start(st)
    o.set_rad_surface(sf)
    xo, yo = o.rad_columns(0, t)
    # the GPU hands the columns over in FP32
    np.testing.assert_allclose(xg, xo, rtol=2e-7, atol=1e-30)
    np.testing.assert_allclose(yg, yo, rtol=2e-6, atol=1e-6)


SOIL = {"lprog_soil_temp": True, "lsfc_sensible_heat_flux": True}


@pytest.mark.parametrize("moist,soil", [(False, False), (True, False), (True, True)])
def test_steps_with_radiation_match_oracle(moist, soil):
    """With the prognostic soil, the radiation sees the soil and sea surface temperature and the
    soil takes the radiation's surface fluxes."""
    cfg, g, st, sf = _case(moist, **(SOIL if soil else {}))
    n = 6
    print(f"steps with radiation moist={moist} n={n}", flush=True)
    s = Step(cfg, g)
    s.start(st)
    s.set_rad_surface(sf)
    s.run(0, n, t_init=T_EPOCH)
    got = s.get_state(n % 2)
    rt_g = s.rad_get(RAD_OUT_TENDENCY)
    sw_g, lw_g = s.rad_get(RAD_OUT_SW_IN), s.rad_get(RAD_OUT_LW_OUT)
    s.close()
    o = CpuOracle(cfg, g)
    o.start(st)
    o.set_rad_surface(sf)
    o.run_rad(0, n, T_EPOCH)
    want = o.get_state(n % 2)
    rt_o = o.get_diag("radiation_tendency", (cfg.ny, cfg.nx, cfg.n_layers))
    sw_o, lw_o = o.rad_sfc()
    rel = max_rel(got, want)
    print(f"  max relative deviation {rel:.3e}")
    assert rel <= 1e-12
    if soil:
        assert np.abs(got["temperature_soil"] - st["temperature_soil"]).max() > 0.0
    np.testing.assert_allclose(rt_g, rt_o, rtol=1e-12, atol=1e-12 * np.abs(rt_o).max())
    np.testing.assert_allclose(sw_g, sw_o, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(lw_g, lw_o, rtol=1e-12)
    assert np.abs(rt_o).max() > 0.0


def test_radiation_update_needs_surface():
    cfg, g, st, sf = _case(False)
    s = Step(cfg, g)
    s.start(st)
    with pytest.raises(LgameError):
        s.pchevi(0, 1, 0, True, T_EPOCH)
    s.close()
