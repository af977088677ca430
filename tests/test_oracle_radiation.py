"""Radiation boundary of the oracle: the coupler steps of mo_rterrtmgp_coupler.F90 (coszenith,
interfaces, water paths, calc_power_density) with the gray two-stream solver that replaces
RTE+RRTMGP (include/lgame_gpu.h).

RTE+RRTMGP is not in the reference tree, so nothing pins the reference's radiation outputs
(parity unpinned).  What is pinned here:
* coszenith against closed forms at the reference epoch (2019-12-20 12:00 UTC, a solstice);
* the solver's fluxes, heating rates and surface fluxes against an independent formulation of
  the same gray two-stream model (explicit sums and products over layers instead of the
  recursions of the oracle), built from the grid and state directly;
* flux conservation: the column integral of the heating equals the net flux through TOA minus
  the net flux through the surface.
"""
import dataclasses
import math

import numpy as np
import pytest

from lgame.cases import make_case, radiation_surface, tiny_config
from lgame.config import RAD_TEACHER
from pyoracle import CpuOracle

H = dict(S0=1361.0, HS=8.e3, TMIN=160.0, TMAX=355.0, EPS=0.98, D=1.66, P0=1e5, G=9.80616, SIG=5.670374419e-8,
         KLD=2.0e-4, KLV=0.08, KLL=0.13, KLI=0.07, KSD=1.0e-5, KSV=0.02, KSL=0.02, KSI=0.01)
R_D, C_D_P = 287.057811, 1005.0
T_EPOCH = 1576843200.0
OBLIQUITY = 0.409092592843564


def test_coszenith_closed_forms():
    cz = CpuOracle.coszenith
    # subsolar point at the solstice: latitude -obliquity, longitude 0 at 12 UTC
    assert cz(-OBLIQUITY, 0.0, T_EPOCH) == pytest.approx(1.0, abs=1e-12)
    assert cz(0.0, 0.0, T_EPOCH) == pytest.approx(math.cos(OBLIQUITY), abs=1e-12)
    assert cz(0.5, 0.0, T_EPOCH) == pytest.approx(math.cos(0.5 + OBLIQUITY), abs=1e-12)
    assert cz(math.pi / 2, 0.0, T_EPOCH) == 0.0          # polar night
    assert cz(0.0, math.pi, T_EPOCH) == 0.0              # midnight
    # a quarter turn of the earth later the subsolar meridian lies a quarter turn west
    q = (math.pi / 2) / 7.292115e-5
    assert cz(-OBLIQUITY, -math.pi / 2, T_EPOCH + q) == pytest.approx(1.0, abs=1e-5)


def _independent_columns(cfg, g, st, sf, t):
    """Heating rates [K/day] and surface fluxes of every column, formal solution of the model."""
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    gc = cfg.n_condensed_constituents
    T = (g["theta_v_bg"] + st["theta_v_pert"]) * (g["exner_bg"] + st["exner_pert"])
    rho = st["rho"]
    if cfg.lmoist:
        T = T / (1.0 + rho[:, :, :, gc + 1] / rho[:, :, :, gc] * (0.004810 / 0.002991 - 1.0))
    rg = rho[:, :, :, gc]
    p = R_D * rg * T
    zw, zs = g["z_w"], g["z_scalar"]
    thick = zw[:, :, :-1] - zw[:, :, 1:]
    if cfg.lmoist:
        lwp = thick * 1000.0 * (rho[:, :, :, 1] + rho[:, :, :, 3])
        iwp = thick * 1000.0 * (rho[:, :, :, 0] + rho[:, :, :, 2] + rho[:, :, :, 4])
        qv = rho[:, :, :, gc + 1] / rg
    else:
        lwp = iwp = qv = np.zeros_like(p)
    Tc = np.clip(T, H["TMIN"], H["TMAX"])
    pi = np.empty((ny, nx, nl + 1))
    pi[:, :, 0] = p[:, :, 0] * np.exp(-(zw[:, :, 0] - zs[:, :, 1]) / H["HS"])
    pi[:, :, nl] = p[:, :, nl - 1] * np.exp(-(zw[:, :, nl] - zs[:, :, nl - 1]) / H["HS"])
    pi[:, :, 1:nl] = 0.5 * (p[:, :, :-1] + p[:, :, 1:])
    dm = np.diff(pi, axis=2) / H["G"]
    tau = dm * (H["KLD"] * p / H["P0"] + H["KLV"] * qv) + H["KLL"] * lwp + H["KLI"] * iwp
    tsw = dm * (H["KSD"] + H["KSV"] * qv) + H["KSL"] * lwp + H["KSI"] * iwp
    B = H["SIG"] * Tc ** 4
    ts = np.clip(sf["temp_sfc"], H["TMIN"], H["TMAX"])
    mu = np.vectorize(CpuOracle.coszenith)(sf["lat_geo"], sf["lon_geo"], t)
    # optical depth from TOA down to each level, in units of the diffuse path
    ctau = np.concatenate([np.zeros((ny, nx, 1)), np.cumsum(tau, axis=2)], axis=2)
    dn = np.zeros((ny, nx, nl + 1))
    for k in range(1, nl + 1):   # level k: sources of layers j < k, transmitted through layers j+1..k-1
        src = B[:, :, :k] * (1.0 - np.exp(-H["D"] * tau[:, :, :k]))
        trans = np.exp(-H["D"] * (ctau[:, :, k:k + 1] - ctau[:, :, 1:k + 1]))
        dn[:, :, k] = np.sum(src * trans, axis=2)
    up_sfc = H["EPS"] * H["SIG"] * ts ** 4 + (1.0 - H["EPS"]) * dn[:, :, nl]
    up = np.zeros_like(dn)
    for k in range(nl + 1):
        up[:, :, k] = up_sfc * np.exp(-H["D"] * (ctau[:, :, nl] - ctau[:, :, k]))
        if k < nl:
            src = B[:, :, k:] * (1.0 - np.exp(-H["D"] * tau[:, :, k:]))
            trans = np.exp(-H["D"] * (ctau[:, :, k:nl] - ctau[:, :, k:k + 1]))
            up[:, :, k] += np.sum(src * trans, axis=2)
    q = (up[:, :, 1:] - up[:, :, :-1] + dn[:, :, :-1] - dn[:, :, 1:]) / thick
    lw_out = up[:, :, nl] - dn[:, :, nl]
    csw = np.concatenate([np.zeros((ny, nx, 1)), np.cumsum(tsw, axis=2)], axis=2)
    day = mu > 0.0
    mus = np.where(day, mu, 1.0)
    dsw = np.where(day[:, :, None], H["S0"] * mus[:, :, None] * np.exp(-csw / mus[:, :, None]), 0.0)
    usw = sf["sfc_albedo"][:, :, None] * dsw[:, :, nl:nl + 1] * np.exp(-H["D"] * (csw[:, :, nl:nl + 1] - csw))
    q = q + (usw[:, :, 1:] - usw[:, :, :-1] + dsw[:, :, :-1] - dsw[:, :, 1:]) / thick
    sw_in = dsw[:, :, nl] - usw[:, :, nl]
    toa_net = (dsw[:, :, 0] - usw[:, :, 0]) + (dn[:, :, 0] - up[:, :, 0])
    hr = 86400.0 * q / (rg * C_D_P)
    return hr, sw_in, lw_out, q, thick, toa_net


@pytest.mark.parametrize("moist", [False, True])
@pytest.mark.parametrize("hours", [0.0, 9.0])
def test_oracle_radiation_matches_independent_formulation(moist, hours):
    cfg = dataclasses.replace(tiny_config(moist=moist, n_layers=12), lrad=True, rad_mode=RAD_TEACHER)
    g, st = make_case(cfg, cloud_frac=0.02 if moist else 0.0)
    sf = radiation_surface(cfg, g, st)
    t = T_EPOCH + 3600.0 * hours
    o = CpuOracle(cfg, g)
    o.start(st)
    o.set_rad_surface(sf)
    x, y = o.rad_columns(0, t)
    hr, sw_in, lw_out, q, thick, toa_net = _independent_columns(cfg, g, st, sf, t)
    nl = cfg.n_layers
    # oracle columns are in Fortran order (ji fastest)
    hr_o = y[:, :nl].reshape(cfg.nx, cfg.ny, nl).transpose(1, 0, 2)
    np.testing.assert_allclose(hr_o, hr, rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(y[:, nl].reshape(cfg.nx, cfg.ny).T, sw_in, rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(y[:, nl + 1].reshape(cfg.nx, cfg.ny).T, lw_out, rtol=1e-10)
    # the column integral of the heating is the net flux into the column
    np.testing.assert_allclose(np.sum(q * thick, axis=2), toa_net - (sw_in - lw_out), rtol=1e-9, atol=1e-6)
    assert np.all(hr_o[:, :, 2:] < 0.0) or hours == 0.0  # longwave cooling dominates at night
    if hours == 0.0 and not moist:  # clear sky at noon: most of the beam reaches the ground
        assert float(sw_in.max()) > 100.0


def test_oracle_step_applies_radiation():
    """pchevi with lrad_update: the radiation tendency enters the rhotheta tendency."""
    cfg = dataclasses.replace(tiny_config(moist=True), lrad=True, rad_mode=RAD_TEACHER, dtime_rad=10.0)
    g, st = make_case(cfg, cloud_frac=0.3)
    sf = radiation_surface(cfg, g, st)
    runs = []
    for lrad in (False, True):
        o = CpuOracle(dataclasses.replace(cfg, lrad=lrad), g)
        o.start(st)
        o.set_rad_surface(sf)
        o.run_rad(0, 4, T_EPOCH)
        runs.append(o.get_state(0))
        if lrad:
            rt = o.get_diag("radiation_tendency", (cfg.ny, cfg.nx, cfg.n_layers))
    assert np.abs(rt).max() > 0.0
    d = np.abs(runs[1]["rhotheta_v"] - runs[0]["rhotheta_v"]).max()
    assert d > 0.0


def test_oracle_radiation_surface_temperature_with_the_soil():
    """With the prognostic soil the radiation sees the surface temperature of the sea and the soil
    (mo_manage_radiation_calls.F90:36-42): the sea surface temperature where there is neither land
    nor lake, the soil's top layer where land and lake cover the column, and in between the mean
    weighted by the sea fraction (the temperature clipping of the coupler, 160-350 K, does not
    act on these values)."""
    cfg = dataclasses.replace(tiny_config(ny=5, nx=6), lrad=True, rad_mode=RAD_TEACHER, dtime_rad=10.0,
                              lprog_soil_temp=True)
    g, st = make_case(cfg)
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    land = np.zeros((ny, nx))
    land[1] = 1.0
    land[2] = 0.25
    lake = np.zeros((ny, nx))
    lake[2] = 0.35
    lake[3] = 1.0
    g["land_fraction"], g["lake_fraction"] = np.asfortranarray(land), np.asfortranarray(lake)
    g["sst"] = np.asfortranarray(np.full((ny, nx), 291.5))
    st["temperature_soil"][:, :, 0] = 283.25
    o = CpuOracle(cfg, g)
    o.start(st)
    o.set_rad_surface(radiation_surface(cfg, g, st))
    x, _ = o.rad_columns(0, T_EPOCH)
    tsfc = x[:, 5 * nl].reshape(nx, ny).T  # columns in Fortran order (ji fastest)
    np.testing.assert_array_equal(tsfc[0], 291.5)  # open sea
    np.testing.assert_array_equal(tsfc[1], 283.25)  # land
    np.testing.assert_array_equal(tsfc[3], 283.25)  # lake
    np.testing.assert_allclose(tsfc[2], 0.4 * 291.5 + 0.6 * 283.25, rtol=1e-14)
