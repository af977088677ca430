"""Idealized initial states (host side, once per run).

Restates ``ideal_init`` / ``unessential_ideal_init`` (``io/mo_set_initial_state.F90:22-382``) for
the "standard", "schaer" and "mountain_rest" scenarios, plus seeded perturbations and a moist
humidity profile used by the larger workloads (set-up extensions, documented per function).
All arrays are float64, Fortran-ordered, with the reference extents.
"""
from __future__ import annotations

import numpy as np

from .config import RunConfig, state_shape
from .grid import C_D_P, P_0, R_D, bg_pres, bg_temp, geopot

C_D_V = 717.942189
R_V = 461.524879


def _zeros(cfg: RunConfig):
    ny, nx, nl, nc = cfg.ny, cfg.nx, cfg.n_layers, cfg.n_constituents
    st = {k: np.zeros(state_shape(k, ny, nx, nl, nc), order="F") for k in
          ("rho", "rhotheta_v", "theta_v_pert", "exner_pert", "wind_u", "wind_v", "wind_w")}
    if cfg.lprog_soil_temp:  # 280 K in every soil layer, mo_set_initial_state.F90:236-242
        st["temperature_soil"] = np.full((ny, nx, cfg.nsoillays), 280.0, order="F")
    return st


def w_free_slip(st: dict, g: dict) -> None:
    """w_free_slip, spatial_operators/mo_inner_product.F90:52-75."""
    ipw = g["inner_product_weights"]
    nl = st["wind_u"].shape[2]
    u, v, sx, sy = st["wind_u"], st["wind_v"], g["slope_x"], g["slope_y"]
    st["wind_w"][:, :, nl] = (ipw[0][:, :, nl - 1] * u[:, 1:, nl - 1] * sx[:, 1:, nl - 1]
                              + ipw[1][:, :, nl - 1] * v[:-1, :, nl - 1] * sy[:-1, :, nl - 1]
                              + ipw[2][:, :, nl - 1] * u[:, :-1, nl - 1] * sx[:, :-1, nl - 1]
                              + ipw[3][:, :, nl - 1] * v[1:, :, nl - 1] * sy[1:, :, nl - 1])


def _hydrostatic(st, g, cfg, temp, pres_lowest):
    """unessential_ideal_init, io/mo_set_initial_state.F90:322-382 (temp = virtual temperature)."""
    z = g["z_scalar"]
    nl = z.shape[2]
    phi = geopot(z, cfg.lplane)
    exner = np.zeros_like(z)
    exner[:, :, nl - 1] = (pres_lowest / P_0) ** (R_D / C_D_P)
    for l in range(nl - 2, -1, -1):
        b = -0.5 * exner[:, :, l + 1] / temp[:, :, l + 1] * (temp[:, :, l] - temp[:, :, l + 1] + 2.0 / C_D_P * (phi[:, :, l] - phi[:, :, l + 1]))
        c = exner[:, :, l + 1] ** 2 * temp[:, :, l] / temp[:, :, l + 1]
        exner[:, :, l] = b + np.sqrt(b ** 2 + c)
    theta = temp / exner
    gc = cfg.n_condensed_constituents
    st["rho"][:, :, :, gc] = P_0 * exner ** (C_D_P / R_D) / (R_D * temp)
    st["rhotheta_v"][:] = st["rho"][:, :, :, gc] * theta
    st["theta_v_pert"][:] = theta - g["theta_v_bg"]
    st["exner_pert"][:] = exner - g["exner_bg"]
    st["wind_w"][:] = 0.0


def ideal_init(cfg: RunConfig, g: dict) -> dict:
    """ideal_init, io/mo_set_initial_state.F90:22-257, scenarios standard / schaer / mountain_rest."""
    st = _zeros(cfg)
    z = g["z_scalar"]
    nl = z.shape[2]
    if cfg.scenario in ("standard", "seabreeze"):
        temp = bg_temp(z)
        pres_lowest = bg_pres(z[:, :, nl - 1])
    elif cfg.scenario in ("schaer", "mountain_rest"):
        if cfg.scenario == "schaer":
            st["wind_u"][:] = 10.0
        n2 = 0.01 ** 2
        t0 = 288.0
        theta = np.zeros_like(z)
        ex = np.zeros_like(z)
        dz = z[:, :, nl - 1]
        gl = geopot(z[:, :, nl - 1], cfg.lplane) / dz
        theta[:, :, nl - 1] = t0 * (1.0 + n2 * dz / (2.0 * gl)) / (1.0 - n2 * dz / (2.0 * gl))
        ex[:, :, nl - 1] = 1.0 - 2.0 * geopot(z[:, :, nl - 1], cfg.lplane) / (C_D_P * (theta[:, :, nl - 1] + t0))
        for l in range(nl - 2, -1, -1):
            dz = z[:, :, l] - z[:, :, l + 1]
            gl = (geopot(z[:, :, l], cfg.lplane) - geopot(z[:, :, l + 1], cfg.lplane)) / dz
            theta[:, :, l] = theta[:, :, l + 1] * (1.0 + n2 * dz / (2.0 * gl)) / (1.0 - n2 * dz / (2.0 * gl))
        for l in range(nl - 2, -1, -1):
            ex[:, :, l] = ex[:, :, l + 1] - 2.0 * (geopot(z[:, :, l], cfg.lplane) - geopot(z[:, :, l + 1], cfg.lplane)) / \
                (C_D_P * (theta[:, :, l] + theta[:, :, l + 1]))
        temp = ex * theta
        pres_lowest = P_0 * ex[:, :, nl - 1] ** (C_D_P / R_D)
    else:
        raise ValueError(f"scenario {cfg.scenario!r} is not set up here")
    _hydrostatic(st, g, cfg, temp, pres_lowest)
    if cfg.lfreeslip:
        w_free_slip(st, g)
    return st


def _smooth_periodic(field, sigma_cells):
    """Gaussian filter with periodic wrap along axes 0 and 1 (FFT), unit-variance rescaled."""
    ny, nx = field.shape[0], field.shape[1]
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    kern = np.exp(-2.0 * (np.pi * sigma_cells) ** 2 * (kx ** 2 + ky ** 2))
    if field.ndim == 3:
        kern = kern[:, :, None]
    out = np.fft.irfft2(np.fft.rfft2(field, axes=(0, 1)) * kern, s=(ny, nx), axes=(0, 1))
    s = out.std()
    return out / s if s > 0 else out


def perturb(st: dict, g: dict, cfg: RunConfig, seed: int, theta_sigma=0.5, wind_sigma=1.0,
            corr_length=20e3) -> None:
    """Seeded Gaussian-random perturbations of a resting state (set-up extension).

    theta' gets a horizontally correlated field (Gaussian kernel of half the correlation length,
    in cells) of standard deviation ``theta_sigma`` at constant density; rhotheta_v and the
    Exner pressure then follow from the equation of state as in ``restart``
    (io/mo_set_initial_state.F90:275-276).  u and v get correlated noise of ``wind_sigma``.
    """
    rng = np.random.default_rng(seed)
    gc = cfg.n_condensed_constituents
    sig = 0.5 * corr_length / cfg.dx
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    dth = theta_sigma * _smooth_periodic(rng.standard_normal((ny, nx, nl)), sig)
    rho = st["rho"][:, :, :, gc]
    theta = st["theta_v_pert"] + g["theta_v_bg"] + dth
    st["rhotheta_v"][:] = rho * theta
    st["theta_v_pert"][:] = theta - g["theta_v_bg"]
    exner = (R_D * st["rhotheta_v"] / P_0) ** (R_D / C_D_V)
    st["exner_pert"][:] = exner - g["exner_bg"]
    du = wind_sigma * _smooth_periodic(rng.standard_normal((ny, nx, nl)), sig)
    dv = wind_sigma * _smooth_periodic(rng.standard_normal((ny, nx, nl)), sig)
    st["wind_u"][:, :nx] += du
    st["wind_u"][:, nx] = st["wind_u"][:, 0]
    st["wind_v"][:ny] += dv
    st["wind_v"][ny] = st["wind_v"][0]
    if cfg.lfreeslip:
        w_free_slip(st, g)


def saturation_pressure_over_water(t):
    """Magnus-type saturation pressure used only to shape the seeded humidity (set-up extension)."""
    tc = t - 273.15
    return 610.94 * np.exp(17.625 * tc / (tc + 243.04))


def add_moisture(st: dict, g: dict, cfg: RunConfig, seed: int, rh_low=0.60, rh_high=0.95,
                 moist_depth=3000.0, decay=2000.0, cloud_frac=0.0) -> None:
    """Seeded water vapour (set-up extension): relative humidity U[rh_low, rh_high] in the lowest
    ``moist_depth`` metres, decaying exponentially above; the moist-air ("gas") density becomes
    dry density + vapour density; condensates start at zero unless ``cloud_frac`` > 0, which puts
    that fraction of the vapour into cloud liquid (jc=4) and cloud ice (jc=3) at saturation."""
    if not cfg.lmoist:
        return
    rng = np.random.default_rng(seed)
    nc = cfg.n_constituents
    gc = cfg.n_condensed_constituents
    z = g["z_scalar"]
    temp = (g["theta_v_bg"] + st["theta_v_pert"]) * (g["exner_bg"] + st["exner_pert"])
    rh = rh_low + (rh_high - rh_low) * 0.5 * (1.0 + np.tanh(_smooth_periodic(rng.standard_normal(z.shape[:2]), 4.0)))
    zsfc = g["z_w"][:, :, -1]
    hag = z - zsfc[:, :, None]
    prof = np.where(hag <= moist_depth, 1.0, np.exp(-(hag - moist_depth) / decay))
    rho_v = rh[:, :, None] * prof * saturation_pressure_over_water(temp) / (R_V * temp)
    theta = g["theta_v_bg"] + st["theta_v_pert"]
    st["rho"][:, :, :, nc - 1] = rho_v
    st["rho"][:, :, :, gc] = st["rho"][:, :, :, gc] + rho_v
    # keep theta_v; rhotheta_v and the Exner pressure follow the equation of state as in restart
    st["rhotheta_v"][:] = st["rho"][:, :, :, gc] * theta
    st["exner_pert"][:] = (R_D * st["rhotheta_v"] / P_0) ** (R_D / C_D_V) - g["exner_bg"]
    if cloud_frac > 0.0:
        st["rho"][:, :, :, 3] = np.where(temp > 273.15, cloud_frac * rho_v, 0.0)
        st["rho"][:, :, :, 2] = np.where(temp <= 273.15, cloud_frac * rho_v, 0.0)
