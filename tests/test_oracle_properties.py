"""Properties that pin the oracle independently of the reference build.

* Thomas elimination (mo_column_solvers.F90:606-640) against a dense LU solve.
* Dry-mass conservation of the step to round-off on a periodic plane whose terrain vanishes at
  the domain edges (the flux-form divergence then telescopes; the reference Schaer run conserves
  mass to ~1e-13 as well).  The periodic lat-lon domain is not conservative in the reference:
  its extrapolated face areas at the duplicate edges differ (mo_grid_generator.F90:1113-1170).
* A resting, hydrostatically balanced atmosphere over flat ground stays at rest.
* Tracer densities stay non-negative (the solver clips round-off negatives,
  mo_column_solvers.F90:587-591).
"""
import dataclasses
import os

import numpy as np
import pytest

from lgame.cases import dry3d_config, make_case, tiny_config
from pyoracle import REF_LIB, CpuOracle, thomas_reference


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference build not present")
@pytest.mark.parametrize("n", [2, 3, 9, 59])
def test_reference_thomas_matches_dense_solve(n):
    rng = np.random.default_rng(n)
    c = rng.uniform(-1, 1, n - 1)
    e = rng.uniform(-1, 1, n - 1)
    d = 4.0 + rng.uniform(0, 1, n)
    r = rng.uniform(-1, 1, n)
    A = np.diag(d) + np.diag(c, -1) + np.diag(e, 1)
    x = thomas_reference(c, d, e, r)
    np.testing.assert_allclose(x, np.linalg.solve(A, r), rtol=1e-13, atol=1e-14)


def _mass(st, grid, jc):
    return float(np.sum(st["rho"][:, :, :, jc] * grid["volume"]))


def conservative_case(moist):
    cfg = tiny_config(ny=15, nx=15, n_layers=10, moist=moist, order=3, freeslip=True, plane=True)
    return dataclasses.replace(cfg, hills=((0.0, 0.0, 300.0, 3000.0),))


@pytest.mark.parametrize("moist", [False, True])
def test_oracle_conserves_mass(moist):
    cfg = conservative_case(moist)
    grid, st = make_case(cfg, cloud_frac=0.1 if moist else 0.0)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, 10)
    out = o.get_state(0)
    gc = cfg.n_condensed_constituents
    m0, m1 = _mass(st, grid, gc), _mass(out, grid, gc)
    assert abs(m1 - m0) / m0 < 1e-13
    assert np.abs(out["wind_w"]).max() > 1e-6  # the case is not trivially at rest
    if moist:
        assert out["rho"].min() >= 0.0


def test_oracle_rest_state_stays_at_rest():
    cfg = dataclasses.replace(tiny_config(), orography_id=0, hills=(), lcorio=True)
    grid, st = make_case(cfg, perturbation=False)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, 6)
    out = o.get_state(0)
    assert np.abs(out["wind_u"]).max() < 1e-10
    assert np.abs(out["wind_v"]).max() < 1e-10
    assert np.abs(out["wind_w"]).max() < 1e-10


@pytest.mark.parametrize("moist", [False, True])
def test_oracle_integrals_closed_form(moist):
    """write_out_integrals (mo_write_out.F90:226-304) on a uniform state over flat ground.

    Constituent masses are rho_c * V_total, the rhotheta integral rho*theta*V_total, kinetic
    energy 0.5*rho_gas*(u^2+v^2)*V_total (the inner-product weights of a flat Cartesian cell are
    1/2 on each face), potential energy rho_gas*sum(phi*V), and internal energy
    c_d_v*rho_gas*T*V_total with T = theta_v*Pi/(1 + rho_v/rho_gas*(m_d/m_v - 1)).
    """
    cfg = dataclasses.replace(tiny_config(moist=moist, plane=True), orography_id=0, hills=())
    grid, st = make_case(cfg, perturbation=False)
    nc, gc = cfg.n_constituents, cfg.n_condensed_constituents
    rho_c = np.linspace(1e-3, 7e-3, nc)
    rho_c[gc] = 1.1
    theta0, pi0, u0, v0 = 300.0, 0.9, 7.0, -3.0
    for jc in range(nc):
        st["rho"][:, :, :, jc] = rho_c[jc]
    st["rhotheta_v"][:] = rho_c[gc] * theta0
    st["theta_v_pert"][:] = theta0 - grid["theta_v_bg"]
    st["exner_pert"][:] = pi0 - grid["exner_bg"]
    st["wind_u"][:] = u0
    st["wind_v"][:] = v0
    st["wind_w"][:] = 0.0
    ipw = grid["inner_product_weights"]
    assert np.allclose(ipw[0] + ipw[2], 1.0, rtol=0, atol=1e-14)
    assert np.allclose(ipw[1] + ipw[3], 1.0, rtol=0, atol=1e-14)
    o = CpuOracle(cfg, grid)
    o.set_state(0, st)
    o.lib().orc_diagnose_temperature(o.h, 0)
    got = o.integrals(0)
    vol = float(np.sum(grid["volume"]))
    temp = theta0 * pi0
    if moist:
        m_d, m_v = 0.004810, 0.002991
        temp /= 1.0 + rho_c[nc - 1] / rho_c[gc] * (m_d / m_v - 1.0)
    want = np.zeros(4 + nc)
    want[0] = rho_c[gc] * theta0 * vol
    want[1] = 0.5 * rho_c[gc] * (u0 * u0 + v0 * v0) * vol
    want[2] = rho_c[gc] * float(np.sum(grid["gravity_potential"] * grid["volume"]))
    want[3] = 717.942189 * rho_c[gc] * temp * vol
    want[4:] = rho_c * vol
    np.testing.assert_allclose(got, want, rtol=1e-12)
