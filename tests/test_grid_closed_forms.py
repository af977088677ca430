"""The host grid generator (lgame/grid.py, the checker of the device generator) against what the
geometry and the background state's equations fix (mo_grid_generator.F90:29-1457).

* A flat periodic plane with equal layers: face areas dy dz, dx dz, dx dy, volumes dx dy dz,
  inner-product weights 1/2 on every face (half the cell's volume belongs to each face), TRSK
  weights 1/4 and 0 (the Coriolis term of a uniform flow is f times the four-point mean of the
  perpendicular wind, and the parallel wind does not enter).
* The sphere with flat ground: the columns' volumes add up to the spherical shell segment
  (R+toa)^3/3 - R^3/3 times the solid angle of the column, the surface areas to the area of the
  latitude band.
* The Schaer orography: h(x) = 250 m exp(-(x/5 km)^2) cos^2(pi x / 4 km) about the middle column
  (x = 0 at column nx/2+1), symmetric, 250 m at its centre, zero where cos vanishes.
* bg_setup: the background Exner pressure satisfies the discrete hydrostatic balance
  c_p (theta_l + theta_l+1)/2 (Pi_l - Pi_l+1) = -(phi_l - phi_l+1) between adjacent layers, with
  the ICAO temperature and, in the lowest layer, the ICAO pressure.
"""
import dataclasses

import numpy as np
import pytest

from lgame.cases import schaer_config, tiny_config
from lgame.grid import GRAVITY, R_E, bg_pres, bg_temp, grid_setup, P_0, R_D, C_D_P


def _flat_plane(ny=6, nx=7, nl=8, dx=2000.0, dy=1500.0, toa=12000.0):
    cfg = dataclasses.replace(tiny_config(ny=ny, nx=nx, n_layers=nl, plane=True), orography_id=0, hills=(),
                              stretching_parameter=1.0, dx=dx, dy=dy, toa=toa)
    return cfg, grid_setup(cfg)


def test_flat_plane_metrics():
    cfg, g = _flat_plane()
    dz = cfg.toa / cfg.n_layers
    np.testing.assert_allclose(np.diff(g["z_w"], axis=2), -dz, rtol=1e-12)
    np.testing.assert_allclose(g["area_x"], cfg.dy * dz, rtol=1e-12)
    np.testing.assert_allclose(g["area_y"], cfg.dx * dz, rtol=1e-12)
    np.testing.assert_allclose(g["area_z"], cfg.dx * cfg.dy, rtol=1e-12)
    np.testing.assert_allclose(g["volume"], cfg.dx * cfg.dy * dz, rtol=1e-12)
    np.testing.assert_allclose(g["dz"], dz, rtol=1e-12)
    np.testing.assert_allclose(g["inner_product_weights"], 0.5, rtol=1e-12)
    np.testing.assert_allclose(g["slope_x"], 0.0, atol=0)
    np.testing.assert_allclose(g["slope_y"], 0.0, atol=0)
    # the layer centres halfway between their interfaces
    np.testing.assert_allclose(g["z_scalar"], 0.5 * (g["z_w"][:, :, :-1] + g["z_w"][:, :, 1:]), rtol=1e-14)


def test_flat_plane_trsk_weights():
    """TRSK on a rectangle: the u-point term is 1/4 of the four surrounding v mass fluxes (weights
    1 and 3, each used twice) and none of the neighbouring u fluxes (weight 2); the v-point term is
    -1/4 of the four surrounding u fluxes.  For a uniform flow (u0, v0) with potential vorticity
    f/rho this gives the Coriolis acceleration (f v0, -f u0)."""
    cfg, g = _flat_plane(dx=2000.0, dy=1500.0)
    tu, tv = g["trsk_weights_u"], g["trsk_weights_v"]
    np.testing.assert_allclose(tu[0], 0.25, rtol=1e-12)
    np.testing.assert_allclose(tu[1], 0.0, atol=1e-15)
    np.testing.assert_allclose(tu[2], 0.25, rtol=1e-12)
    np.testing.assert_allclose(tv, -0.25, rtol=1e-12)
    f, rho, u0, v0 = 1e-4, 1.2, 5.0, -2.0
    # the six terms of mo_vorticity_flux.F90:31-44 and the four of :73-80 on a uniform field
    ax = (tu[0] + tu[1] * u0 / v0 + tu[2] + tu[2] + tu[1] * u0 / v0 + tu[0]) * rho * v0 * (f / rho)
    ay = (tv[0] + tv[0] + tv[1] + tv[1]) * rho * u0 * (f / rho)
    np.testing.assert_allclose(ax, f * v0, rtol=1e-12)
    np.testing.assert_allclose(ay, -f * u0, rtol=1e-12)


def test_sphere_flat_ground_volumes_and_areas():
    cfg = dataclasses.replace(tiny_config(ny=9, nx=11, n_layers=7, plane=False), orography_id=0, hills=(),
                              lat_center=0.0, lon_center=0.0)
    g = grid_setup(cfg)
    nl = cfg.n_layers
    dlat, dlon = cfg.dy / R_E, cfg.dx / R_E
    lat = (cfg.ny - 1.0) / 2.0 * dlat - dlat * np.arange(cfg.ny)
    band = R_E ** 2 * dlon * (np.sin(lat + 0.5 * dlat) - np.sin(lat - 0.5 * dlat))
    np.testing.assert_allclose(g["area_z"][:, :, nl], np.broadcast_to(band[:, None], (cfg.ny, cfg.nx)), rtol=1e-12)
    shell = ((R_E + cfg.toa) ** 3 - R_E ** 3) / 3.0 * dlon * (np.sin(lat + 0.5 * dlat) - np.sin(lat - 0.5 * dlat))
    np.testing.assert_allclose(g["volume"].sum(axis=2), np.broadcast_to(shell[:, None], (cfg.ny, cfg.nx)), rtol=1e-12)
    # faces grow with the square of the distance from the centre of the Earth
    zw = g["z_w"]
    np.testing.assert_allclose(g["area_z"], g["area_z"][:, :, nl:] * (R_E + zw) ** 2 / R_E ** 2, rtol=1e-12)


def test_schaer_orography():
    cfg = schaer_config(nx=201, n_layers=40)
    g = grid_setup(cfg)
    h = g["z_w"][0, :, cfg.n_layers]
    mid = cfg.nx // 2  # column nx/2+1 (1-based) holds x = 0
    x = cfg.dx * (np.arange(cfg.nx) - mid)
    assert h[mid] == pytest.approx(250.0, rel=1e-14)
    assert h.argmax() == mid
    np.testing.assert_allclose(h, h[::-1], rtol=0, atol=1e-9)
    np.testing.assert_allclose(h, 250.0 * np.exp(-(x / 5000.0) ** 2) * np.cos(np.pi * x / 4000.0) ** 2, rtol=1e-12,
                               atol=1e-9)
    # cos^2 vanishes at x = +-2 km, +-6 km, ...
    for xk in (2000.0, 6000.0, -2000.0):
        assert abs(h[mid + int(round(xk / cfg.dx))]) < 1e-9
    # the rows are identical (the orography depends on x only)
    assert np.array_equal(g["z_w"][0], g["z_w"][2])


@pytest.mark.parametrize("plane", [True, False])
def test_background_state_is_in_discrete_hydrostatic_balance(plane):
    cfg = dataclasses.replace(tiny_config(ny=5, nx=7, n_layers=12, plane=plane), toa=18000.0)
    g = grid_setup(cfg)
    z, pi, theta, phi = g["z_scalar"], g["exner_bg"], g["theta_v_bg"], g["gravity_potential"]
    lhs = C_D_P * 0.5 * (theta[:, :, :-1] + theta[:, :, 1:]) * (pi[:, :, :-1] - pi[:, :, 1:])
    rhs = -(phi[:, :, :-1] - phi[:, :, 1:])
    np.testing.assert_allclose(lhs, rhs, rtol=1e-10)
    np.testing.assert_allclose(theta * pi, bg_temp(z), rtol=1e-13)
    np.testing.assert_allclose(pi[:, :, -1], (bg_pres(z[:, :, -1]) / P_0) ** (R_D / C_D_P), rtol=1e-14)
    # the geopotential: g z on the plane, the inverse-square law on the sphere
    want = GRAVITY * z if plane else -GRAVITY * R_E ** 2 / (R_E + z) + GRAVITY * R_E
    np.testing.assert_allclose(phi, want, rtol=1e-13)


def _smooth_case():
    return dataclasses.replace(tiny_config(ny=9, nx=11, n_layers=6), orography_id=1, dx=3000.0, dy=2000.0)


def test_orography_smoothing_closed_forms():
    """smooth_hor_scalar2 (mo_grid_generator.F90:1512-1557): a constant stays constant (the
    weights are normalised), the smoothing is linear, and a single raised point p gives at every
    other point q the weight of the pair over the sum of q's weights, w = 1/(d + 1e-10)^1.5 with
    d the great-circle angle between the grid coordinates, here from the haversine formula (the
    reference uses the spherical law of cosines; the two agree to about 1e-8 at 2 km)."""
    from lgame.grid import grid_coordinates, smooth_hor_scalar2

    cfg = _smooth_case()
    ny, nx = cfg.ny, cfg.nx
    np.testing.assert_allclose(smooth_hor_scalar2(np.full((ny, nx), 731.5), cfg), 731.5, rtol=1e-14)
    rng = np.random.default_rng(2)
    a, b = rng.random((ny, nx)), rng.random((ny, nx))
    np.testing.assert_allclose(smooth_hor_scalar2(a + 2.0 * b, cfg),
                               smooth_hor_scalar2(a, cfg) + 2.0 * smooth_hor_scalar2(b, cfg), rtol=1e-13)
    lat, lon = grid_coordinates(cfg)
    pi_, pk = 3, 4
    delta = np.zeros((ny, nx))
    delta[pi_, pk] = 1.0
    got = smooth_hor_scalar2(delta, cfg)
    la, lo = np.meshgrid(lat, lon, indexing="ij")
    for qi, qk in ((0, 0), (3, 5), (8, 10), (pi_, pk)):
        hav = np.sin(0.5 * (la - lat[qi])) ** 2 + np.cos(la) * np.cos(lat[qi]) * np.sin(0.5 * (lo - lon[qk])) ** 2
        w = 1.0 / (2.0 * np.arcsin(np.sqrt(hav)) + 1e-10) ** 1.5
        w[qi, qk] = 0.0
        want = w[pi_, pk] / w.sum()
        np.testing.assert_allclose(got[qi, qk], want, rtol=1e-6)
    # the raised point itself gets only its neighbours' zeros
    assert got[pi_, pk] == 0.0


def test_smoothed_orography_in_the_grid():
    """orography_id 1: with SLEVE the surface is the smoothed field plus a tenth of the detail,
    without SLEVE the smoothed field (mo_grid_generator.F90:605-616)."""
    from lgame.grid import smooth_hor_scalar2

    cfg = _smooth_case()
    rng = np.random.default_rng(5)
    raw = 200.0 * rng.random((cfg.ny, cfg.nx))
    sm = smooth_hor_scalar2(raw, cfg)
    for sleve, want in ((True, sm + 0.1 * (raw - sm)), (False, sm)):
        g = grid_setup(dataclasses.replace(cfg, lsleve=sleve), oro_raw=raw)
        np.testing.assert_allclose(g["z_w"][:, :, -1], want, rtol=1e-14)
