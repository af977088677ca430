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


def _shear_case(axis):
    """Flat periodic plane (dx != dy) at rest, with u = A sin(2 pi y/Ly) (axis 0) or
    v = A sin(2 pi x/Lx) (axis 1), every diffusion switch on, TKE horizontal scheme."""
    cfg = dataclasses.replace(tiny_config(ny=12, nx=10, n_layers=6, plane=True, corio=False), orography_id=0, hills=(),
                              dx=2000.0, dy=1500.0, diff_coeff_scheme_h="tke").set_diffusion()
    grid, st = make_case(cfg, perturbation=False)
    amp = 3.0
    if axis == 0:
        prof = amp * np.sin(2.0 * np.pi * np.arange(cfg.ny) / cfg.ny)
        st["wind_u"][:] = prof[:, None, None]
    else:
        prof = amp * np.sin(2.0 * np.pi * np.arange(cfg.nx) / cfg.nx)
        st["wind_v"][:, :, :] = prof[None, :, None]
    return cfg, grid, st


@pytest.mark.parametrize("axis", [0, 1])
def test_oracle_momentum_diffusion_is_the_discrete_laplacian(axis):
    """mom_diff_h (mo_momentum_diff_diss.F90:25-160) on a divergence-free shear flow over flat
    ground is the curl-of-vorticity term alone.  With a horizontally uniform viscosity (the TKE
    scheme with zero TKE leaves the molecular part) it is the second difference of the wind along
    the shear direction: the tendency over the wind is -visc*(4/d^2)*sin^2(pi/N) at every point,
    with visc the density-weighted viscosity of the cell (diag%viscosity; viscosity_dual times
    zeta divided by the averaged density reduces to it), d the grid distance across the shear
    (dy for u(y), dx for v(x)) and N the number of points along it."""
    cfg, grid, st = _shear_case(axis)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, 1)
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    visc = o.get_diag("viscosity", (ny, nx, nl))
    assert np.ptp(visc, axis=(0, 1)).max() <= 1e-12 * np.abs(visc).max()  # uniform per layer
    if axis == 0:
        tend = o.get_diag("mom_diff_tend_x", (ny, nx + 1, nl))[:, :nx]
        wind, d, n = st["wind_u"][:, :nx], cfg.dy, ny
    else:
        tend = o.get_diag("mom_diff_tend_y", (ny + 1, nx, nl))[:ny]
        wind, d, n = st["wind_v"][:ny], cfg.dx, nx
    mask = np.abs(wind) > 0.5
    ratio = tend[mask] / wind[mask]
    want = np.broadcast_to(-visc[0, 0, :] * 4.0 / d ** 2 * np.sin(np.pi / n) ** 2, wind.shape)[mask]
    np.testing.assert_allclose(ratio, want, rtol=1e-9)


def test_oracle_scalar_diffusion_conserves_heat_and_mass():
    """The diffusive temperature heating and mass tendencies are divergences of fluxes: their
    volume integrals vanish on the periodic plane whose terrain vanishes at the edges."""
    cfg = conservative_case(True).set_diffusion()
    grid, st = make_case(cfg, cloud_frac=0.1)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, 3)
    ny, nx, nl, nc = cfg.ny, cfg.nx, cfg.n_layers, cfg.n_constituents
    vol = grid["volume"]
    tdh = o.get_diag("temp_diff_heating", (ny, nx, nl))
    md = o.get_diag("mass_diff_tendency", (ny, nx, nl, nc))
    assert np.abs(tdh).max() > 0.0
    assert abs(np.sum(tdh * vol)) <= 1e-12 * np.sum(np.abs(tdh) * vol)
    for jc in range(nc):
        m = md[:, :, :, jc]
        if np.abs(m).max() > 0.0:
            assert abs(np.sum(m * vol)) <= 1e-12 * np.sum(np.abs(m) * vol), jc
    tke = o.get_diag("tke", (ny, nx, nl))
    assert tke.min() >= 0.0 and tke.max() > 0.0
