"""The mixed form (FP32 stencils and explicit tendencies, FP64 column solves, density updates
and budget sums) against the FP64 oracle.

Bar: after N steps every prognostic field has an RMS error below RMS_TOL and a max error below
MAX_TOL relative to the oracle (metrics in lgame.metrics); 1e-4 is the RMS bar the mixed form is
held to over a long run, applied here after tens of steps.  The FP64 plain form on the same case
must stay bitwise, which checks that the two forms share the FP64 path correctly.
"""
import dataclasses

import numpy as np
import pytest

from lgame.cases import dry3d_config, make_case, schaer_config, tiny_config
from lgame.config import PREC_FP64, PREC_MIXED
from lgame.gpu import Step
from lgame.metrics import field_errors, worst
from pyoracle import CpuOracle

pytestmark = pytest.mark.gpu
RMS_TOL = 1e-4
MAX_TOL = 1e-4

CASES = {
    "tiny_moist_o3_freeslip": (tiny_config(moist=True, order=3, freeslip=True), 10, 0.3),
    "schaer_61x24": (schaer_config(nx=61, n_layers=24), 60, 0.0),
    "dry3d_33x33x20": (dry3d_config(n=33, n_layers=20), 20, 0.0),
    "moist3d_33x33x30": (None, 20, 0.2),
    # the reference's default diffusion (TKE, Smagorinsky viscosity, momentum/temperature/mass)
    "tiny_moist_o3_freeslip_diff": (dataclasses.replace(tiny_config(moist=True, order=3, freeslip=True)).set_diffusion(),
                                    10, 0.3),
    "dry3d_33x33x20_diff": (dataclasses.replace(dry3d_config(n=33, n_layers=20)).set_diffusion(), 20, 0.0),
    "dry3d_33x33x20_diff_pbl": (dataclasses.replace(dry3d_config(n=33, n_layers=20), lpbl=True).set_diffusion(), 20, 0.0),
    "dry3d_33x33x20_diff_pbl_soil": (dataclasses.replace(dry3d_config(n=33, n_layers=20), lpbl=True, lprog_soil_temp=True,
                                                         lsfc_sensible_heat_flux=True).set_diffusion(), 20, 0.0),
}


def _cfg(name):
    cfg, n, cf = CASES[name]
    if cfg is None:
        from lgame.cases import moist3d_config
        cfg = moist3d_config(n=33, n_layers=30, dtime=3.22)
    return cfg, n, cf


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_form_error(name):
    cfg, n, cf = _cfg(name)
    grid, st = make_case(cfg, cloud_frac=cf)
    print(f"mixed {name}: ny={cfg.ny} nx={cfg.nx} nl={cfg.n_layers} steps={n}", flush=True)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, n)
    want = o.get_state(n % 2)
    g = Step(dataclasses.replace(cfg, precision=PREC_MIXED), grid)
    g.start(st)
    g.run(0, n)
    got = g.get_state(n % 2)
    assert g.count_nonfinite(n % 2) == 0
    g.close()
    err = field_errors(got, want)
    for k, v in err.items():
        print(f"  {k:18s} rms {v['rms_rel']:.3e}  max {v['max_rel']:.3e}")
    assert worst(err, "rms_rel") < RMS_TOL
    assert worst(err, "max_rel") < MAX_TOL
    # the form differs from FP64: it is not silently running the plain path
    assert any(not np.array_equal(got[k], want[k]) for k in want)


def test_plain_form_still_bitwise_next_to_mixed():
    cfg, n, cf = _cfg("tiny_moist_o3_freeslip")
    grid, st = make_case(cfg, cloud_frac=cf)
    gm = Step(dataclasses.replace(cfg, precision=PREC_MIXED), grid)
    gp = Step(dataclasses.replace(cfg, precision=PREC_FP64), grid)
    for g in (gm, gp):
        g.start(st)
        g.run(0, n)
    got = gp.get_state(n % 2)
    gm.close()
    gp.close()
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, n)
    want = o.get_state(n % 2)
    assert all(np.array_equal(got[k], want[k]) for k in want)


def test_mixed_integrals_are_fp64_sums():
    """Budget sums in the mixed form use the FP64 state: equal to the oracle's at the start."""
    cfg, n, cf = _cfg("tiny_moist_o3_freeslip")
    grid, st = make_case(cfg, cloud_frac=cf)
    g = Step(dataclasses.replace(cfg, precision=PREC_MIXED), grid)
    g.start(st)
    i0 = g.integrals(0)
    g.close()
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.lib().orc_diagnose_temperature(o.h, 0)
    want = o.integrals(0)
    # mass and rhotheta integrals do not involve the FP32 temperature
    np.testing.assert_allclose(i0[[0, 1, 2]], want[[0, 1, 2]], rtol=1e-13)
    np.testing.assert_allclose(i0[4:], want[4:], rtol=1e-13)
    np.testing.assert_allclose(i0[3], want[3], rtol=1e-6)


def test_mixed_form_with_radiation_and_soil():
    """The mixed form with the two-stream radiation coupled to the prognostic soil (surface
    temperature from the soil and sea, surface fluxes into the soil) against the oracle."""
    from lgame.cases import radiation_surface
    from lgame.config import RAD_TEACHER

    cfg = dataclasses.replace(tiny_config(ny=7, nx=9, n_layers=12, moist=True), lrad=True, rad_mode=RAD_TEACHER,
                              dtime_rad=6.0, lprog_soil_temp=True, lsfc_sensible_heat_flux=True)
    grid, st = make_case(cfg, cloud_frac=0.02)
    sf = radiation_surface(cfg, grid, st)
    n, t0 = 8, 1576843200.0
    print(f"mixed with radiation and soil: {cfg.ny}x{cfg.nx}x{cfg.n_layers}, {n} steps", flush=True)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.set_rad_surface(sf)
    o.run_rad(0, n, t0)
    want = o.get_state(n % 2)
    g = Step(dataclasses.replace(cfg, precision=PREC_MIXED), grid)
    g.start(st)
    g.set_rad_surface(sf)
    g.run(0, n, t_init=t0)
    got = g.get_state(n % 2)
    assert g.count_nonfinite(n % 2) == 0
    g.close()
    err = field_errors(got, want)
    for k, v in err.items():
        print(f"  {k:18s} rms {v['rms_rel']:.3e}  max {v['max_rel']:.3e}")
    assert worst(err, "rms_rel") < RMS_TOL
    assert worst(err, "max_rel") < MAX_TOL
    assert np.abs(got["temperature_soil"] - st["temperature_soil"]).max() > 0.0
