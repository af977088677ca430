"""Parity of the HIP step (through the C-ABI) with the CPU oracle and the golden fixtures.

Bar: FP64 plain step within 1e-12 relative (max |gpu - oracle| / max |oracle| per prognostic
field) after N steps; the kernels restate the reference's expression trees without FMA, so the
fields are expected to agree bit for bit, which is reported separately.
"""
import dataclasses

import numpy as np
import pytest

from golden_util import NAMES, load, max_rel
from lgame.cases import dry3d_config, make_case, schaer_config, tiny_config
from lgame.gpu import LgameError, Step
from pyoracle import CpuOracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def run_both(cfg, nsteps, cloud_frac=0.0, seed=None):
    grid, st = make_case(cfg, seed=seed, cloud_frac=cloud_frac)
    print(f"case ny={cfg.ny} nx={cfg.nx} nl={cfg.n_layers} moist={cfg.lmoist} order={cfg.theta_adv_order} "
          f"freeslip={cfg.lfreeslip} steps={nsteps}", flush=True)
    g = Step(cfg, grid)
    g.start(st)
    g.run(0, nsteps)
    got = g.get_state(nsteps % 2)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, nsteps)
    want = o.get_state(nsteps % 2)
    g.close()
    return got, want, st


def diff(cfg, **kw):
    """cfg with the six diffusion switches on (the reference default), then the changes in kw."""
    cfg = dataclasses.replace(cfg).set_diffusion(True)
    return dataclasses.replace(cfg, **kw)


CASES = {
    "tiny_dry_o2": (tiny_config(), 4, 0.0),
    "tiny_dry_o3_freeslip": (tiny_config(order=3, freeslip=True), 5, 0.0),
    "tiny_plane_nocorio": (tiny_config(corio=False, plane=True), 4, 0.0),
    "tiny_linear": (tiny_config(ny=7, nx=5, n_layers=8, llinear=True), 3, 0.0),
    "tiny_moist_o3_freeslip": (tiny_config(moist=True, order=3, freeslip=True), 5, 0.3),
    "tiny_moist_o2": (tiny_config(ny=9, nx=11, n_layers=7, moist=True), 4, 0.2),
    "schaer_61x24": (schaer_config(nx=61, n_layers=24), 30, 0.0),
    "dry3d_33x33x20": (dry3d_config(n=33, n_layers=20), 8, 0.0),
    "dry3d_66x40x18_even": (dataclasses.replace(dry3d_config(n=40, n_layers=18), nx=66), 6, 0.0),
    # even nx: the potential vorticity with two cells per thread
    "tiny_even_moist_o3_freeslip": (tiny_config(nx=8, moist=True, order=3, freeslip=True), 5, 0.3),
    "tiny_even_linear": (tiny_config(ny=7, nx=6, n_layers=8, llinear=True), 3, 0.0),
    # no momentum advection at all: the second substep's wind update alone (k_vec_tend)
    "tiny_linear_nocorio": (tiny_config(ny=7, nx=6, n_layers=8, llinear=True, corio=False), 3, 0.0),
    "tiny_even_plane_nocorio": (tiny_config(nx=10, corio=False, plane=True), 4, 0.0),
    "schaer_60x24": (schaer_config(nx=60, n_layers=24), 30, 0.0),
    # diffusion and turbulence
    "tiny_dry_o2_diff": (diff(tiny_config()), 4, 0.0),
    "tiny_moist_o3_freeslip_diff": (diff(tiny_config(moist=True, order=3, freeslip=True)), 5, 0.3),
    "tiny_dry_o3_diff_tke": (diff(tiny_config(order=3), diff_coeff_scheme_h="tke"), 8, 0.0),
    # even nx: the diffusion coefficients with two cells per thread, both horizontal schemes
    "tiny_even_dry_o3_diff_tke": (diff(tiny_config(nx=8, order=3), diff_coeff_scheme_h="tke"), 8, 0.0),
    "tiny_even_moist_o2_diff": (diff(tiny_config(ny=6, nx=12, n_layers=7, moist=True)), 5, 0.2),
    "tiny_moist_diff_no_vertical": (diff(tiny_config(ny=9, nx=11, n_layers=7, moist=True), lmom_diff_v=False,
                                         ltemp_diff_v=False, lmass_diff_h=False), 5, 0.2),
    "tiny_plane_diff_mom_v_only": (diff(tiny_config(corio=False, plane=True), lmom_diff_h=False, ltemp_diff_h=False,
                                        lmass_diff_h=False), 4, 0.0),
    "schaer_61x24_diff": (diff(schaer_config(nx=61, n_layers=24)), 30, 0.0),
    "dry3d_33x33x20_diff": (diff(dry3d_config(n=33, n_layers=20)), 8, 0.0),
    "dry3d_66x40x18_even_diff": (diff(dataclasses.replace(dry3d_config(n=40, n_layers=18), nx=66)), 6, 0.0),
    # the PBL friction (with the device's log, pow and atan: agreement to rounding)
    "tiny_dry_pbl_mom_h": (dataclasses.replace(tiny_config(), lpbl=True, lmom_diff_h=True), 5, 0.0),
    "tiny_moist_o3_freeslip_diff_pbl": (diff(tiny_config(moist=True, order=3, freeslip=True), lpbl=True), 6, 0.3),
    "dry3d_33x33x20_diff_pbl": (diff(dry3d_config(n=33, n_layers=20), lpbl=True), 8, 0.0),
    # the prognostic soil: conduction alone, with the sensible heat flux (whose scalar flux
    # resistance uses the device's log and pow), with the PBL and all diffusion, 7 soil layers
    "tiny_dry_soil": (dataclasses.replace(tiny_config(), lprog_soil_temp=True), 5, 0.0),
    "tiny_moist_soil_sens": (dataclasses.replace(tiny_config(moist=True), lprog_soil_temp=True,
                                                 lsfc_sensible_heat_flux=True), 5, 0.3),
    "tiny_moist_o3_freeslip_diff_pbl_soil": (diff(tiny_config(moist=True, order=3, freeslip=True), lpbl=True,
                                                  lprog_soil_temp=True, lsfc_sensible_heat_flux=True,
                                                  lsfc_phase_trans=True), 6, 0.3),
    "dry3d_33x33x20_diff_pbl_soil7": (diff(dry3d_config(n=33, n_layers=20, nsoillays=7), lpbl=True,
                                           lprog_soil_temp=True, lsfc_sensible_heat_flux=True), 8, 0.0),
    "tiny_dry_phase_trans": (dataclasses.replace(tiny_config(), lsfc_phase_trans=True), 4, 0.0),
    # the smallest domains the step accepts
    "minimal_3x3x3": (tiny_config(ny=3, nx=3, n_layers=3), 4, 0.0),
    "minimal_3x4x4_moist_o3_diff": (diff(tiny_config(ny=3, nx=4, n_layers=4, moist=True, order=3)), 4, 0.2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_gpu_matches_oracle(name):
    cfg, n, cf = CASES[name]
    got, want, st = run_both(cfg, n, cloud_frac=cf)
    rel = max_rel(got, want)
    bitwise = all(np.array_equal(got[k], want[k]) for k in want)
    print(f"{name}: max relative deviation {rel:.3e}, bitwise={bitwise}")
    assert rel <= TOL
    assert np.abs(got["wind_w"] - st["wind_w"]).max() > 0.0
    if "temperature_soil" in want:
        assert np.abs(got["temperature_soil"] - st["temperature_soil"]).max() > 0.0


DIFF_DIAGS = ("tke", "viscosity", "mom_diff_tend_x", "mom_diff_tend_z", "temp_diff_heating", "heating_diss")


@pytest.mark.parametrize("name", ["tiny_moist_o3_freeslip_diff", "tiny_dry_o3_diff_tke", "tiny_even_dry_o3_diff_tke",
                                  "tiny_even_moist_o2_diff", "dry3d_33x33x20_diff", "dry3d_66x40x18_even_diff",
                                  "dry3d_33x33x20_diff_pbl", "dry3d_33x33x20_diff_pbl_soil7"])
def test_gpu_diffusion_diagnostics_match_oracle(name):
    """The carried t_diag fields of the diffusion and the PBL (TKE, viscosity, momentum diffusion
    tendencies, heating rates, roughness length, Monin-Obukhov length) after N steps, field by
    field."""
    cfg, n, cf = CASES[name]
    grid, st = make_case(cfg, cloud_frac=cf)
    print(f"case {name} steps={n}", flush=True)
    g = Step(cfg, grid)
    g.start(st)
    g.run(0, n)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, n)
    surface = ("roughness_length", "monin_obukhov_length") if cfg.lpbl else ()
    if cfg.lsfc_sensible_heat_flux:
        surface += ("scalar_flux_resistance", "power_flux_density_sens_soil", "power_flux_density_sens_sea")
    for k in DIFF_DIAGS + surface:
        a = g.diag(k)
        b = o.get_diag(k, a.shape)
        scale = max(float(np.abs(b).max()), 1e-300)
        rel = float(np.abs(a - b).max()) / scale
        print(f"{name} {k}: max relative deviation {rel:.3e}, bitwise={np.array_equal(a, b)}")
        assert rel <= TOL, k
        assert np.abs(b).max() > 0.0, k
    g.close()


@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_golden(name):
    cfg, grid, st_in, st_out, steps = load(name)
    g = Step(cfg, grid)
    g.start(st_in)
    g.run(0, steps)
    got = g.get_state(steps % 2)
    g.close()
    rel = max_rel(got, st_out)
    print(f"{name}: max relative deviation from the reference fixture {rel:.3e}")
    assert rel <= TOL


def test_host_drop_in_path_matches_resident_path():
    cfg = tiny_config(order=3, freeslip=True)
    grid, st = make_case(cfg)
    a = Step(cfg, grid)
    a.start(st)
    a.run(0, 3)
    res = a.get_state(1)
    b = Step(cfg, grid)
    new = {k: v.copy() for k, v in st.items()}
    old = {k: v.copy() for k, v in st.items()}
    from lgame.gpu import StateView
    import ctypes
    for s in range(3):
        ao = {k: np.asfortranarray(v) for k, v in old.items()}
        an = {k: np.asfortranarray(v) for k, v in new.items()}
        vo = StateView(*[ao[k].ctypes.data for k in ao], None)
        vn = StateView(*[an[k].ctypes.data for k in an], None)
        rc = b.lib.lgame_pchevi_host(b.h, ctypes.byref(vo), ctypes.byref(vn), s, 0, 0.0)
        assert rc == 0, b.lib.lgame_last_error(b.h)
        old, new = an, ao
    for k in res:
        assert np.array_equal(res[k], old[k]), k


def test_integrals_match_oracle():
    cfg = tiny_config(moist=True, order=3, freeslip=True)
    grid, st = make_case(cfg, cloud_frac=0.2)
    g = Step(cfg, grid)
    g.start(st)
    o = CpuOracle(cfg, grid)
    o.start(st)
    i0 = g.integrals(0)
    o.lib().orc_diagnose_temperature(o.h, 0)
    np.testing.assert_allclose(i0, o.integrals(0), rtol=1e-13)
    g.run(0, 3)
    o.run(0, 3)
    np.testing.assert_allclose(g.integrals(1), o.integrals(1), rtol=1e-13)


def test_check_finite_flags_injected_nan():
    cfg = tiny_config()
    grid, st = make_case(cfg)
    g = Step(cfg, grid)
    g.start(st)
    assert g.count_nonfinite(0) == 0
    bad = {k: v.copy() for k, v in st.items()}
    bad["exner_pert"][1, 2, 3] = np.nan
    bad["wind_u"][0, 0, 0] = np.inf
    g.set_state(1, bad)
    assert g.count_nonfinite(1) == 2


def test_step_argument_errors():
    cfg = tiny_config()
    grid, st = make_case(cfg)
    g = Step(cfg, grid)
    with pytest.raises(LgameError, match="holds no state"):
        g.pchevi(0, 1, 0)
    g.start(st)
    with pytest.raises(LgameError, match="distinct"):
        g.pchevi(0, 0, 0)


def test_linear_combine_matches_oracle():
    """linear_combine_two_states of the two time levels (the output-time interpolation of
    control.F90:539) on the device against the oracle, bit for bit; with coefficients (1, 0) it
    gives back the first state (theta_v_pert included: the solver forms it the same way)."""
    cfg = dataclasses.replace(tiny_config(moist=True), lprog_soil_temp=True)
    grid, st = make_case(cfg, cloud_frac=0.3)
    print("linear combination, tiny moist with soil, 3 steps", flush=True)
    g = Step(cfg, grid)
    g.start(st)
    g.run(0, 3)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, 3)
    for c1, c2 in ((0.3125, 0.6875), (-0.5, 1.5)):
        a, b = g.linear_combine(0, 1, c1, c2), o.linear_combine(0, 1, c1, c2)
        for k in b:
            assert np.array_equal(a[k], b[k]), (c1, c2, k)
    same = g.linear_combine(1, 0, 1.0, 0.0)
    state = g.get_state(1)
    for k in state:
        assert np.array_equal(same[k], state[k]), k
    g.close()
