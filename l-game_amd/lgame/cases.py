"""Named workloads: run configuration + grid + initial state.

* ``schaer``: the Schaer mountain-wave test (run_scripts/schaer.sh:12-65) on the ny=3 plane with
  nx=201, 40 layers, dx=500 m, toa 21 km.  dtime defaults to the reference's own
  1.61*sqrt(dx*dy)/1000 = 0.805 s (mo_run_nml.F90:93-94); 4 s is unstable in the reference scheme.
* ``dry3d``: 3-D dry run on the rotated lat-lon grid of run_scripts/standard.sh:14-15, ICAO
  standard atmosphere, 5 seeded Gaussian hills (heights U[0,1500] m, half-widths U[5,30] km),
  seeded theta_v (0.5 K) and wind (1 m/s) perturbations with 20 km correlation length.
* ``moist3d``: as dry3d with the 7 constituents and seeded humidity (60-95 % RH below 3 km).
* ``tiny``: small random-ish 3-D case for fast parity tests.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .config import RunConfig
from .grid import grid_setup
from .init_state import add_moisture, ideal_init, perturb

LAT_CENTER_STD = 0.8929595951304794   # run_scripts/standard.sh:14
LON_CENTER_STD = 0.1199133716060684   # run_scripts/standard.sh:15


def schaer_config(nx=201, n_layers=40, dtime=None, **kw) -> RunConfig:
    return RunConfig(ny=3, nx=nx, n_layers=n_layers, n_oro_layers=n_layers - 1, toa=21000.0, dy=500.0,
                     dx=500.0, scenario="schaer", lcorio=False, stretching_parameter=1.0, lplane=True,
                     theta_adv_order=3, lmoist=False, klemp_begin_rel=0.5, orography_id=2, lsleve=True,
                     lperiodic=True, lfreeslip=True, dtime=dtime, **kw)


def random_hills(cfg: RunConfig, seed: int, n=5, hmax=1500.0, wmin=5e3, wmax=30e3):
    rng = np.random.default_rng(seed)
    Lx, Ly = cfg.nx * cfg.dx, cfg.ny * cfg.dy
    hills = []
    for _ in range(n):
        x0 = rng.uniform(-0.5 * Lx, 0.5 * Lx)
        y0 = rng.uniform(-0.5 * Ly, 0.5 * Ly)
        hills.append((x0, y0, rng.uniform(0.0, hmax), rng.uniform(wmin, wmax)))
    return tuple(hills)


def dry3d_config(n=129, n_layers=50, dtime=3.22, seed=2, **kw) -> RunConfig:
    cfg = RunConfig(ny=n, nx=n, n_layers=n_layers, n_oro_layers=int(round(0.8 * n_layers)), dy=2000.0,
                    dx=2000.0, toa=40000.0, stretching_parameter=1.3, scenario="standard",
                    lat_center=LAT_CENTER_STD, lon_center=LON_CENTER_STD, theta_adv_order=2,
                    lmoist=False, orography_id=5, lsleve=True, lperiodic=True, lfreeslip=False,
                    dtime=dtime, **kw)
    return dataclasses.replace(cfg, hills=random_hills(cfg, seed))


def moist3d_config(n=513, n_layers=60, dtime=3.22, seed=3, **kw) -> RunConfig:
    cfg = dry3d_config(n=n, n_layers=n_layers, dtime=dtime, seed=seed, **kw)
    return dataclasses.replace(cfg, lmoist=True)


def ensemble_members(n, seed0=5000):
    """The batched-ensemble workload: n moist members with seeded sizes nx, ny ~ U[64,384]
    (made odd, as the reference grid generator requires) and layers ~ U{40..70}, member i from
    seed seed0 + i (hills, perturbations and humidity), dx 2 km, dt 3.22 s."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed0 + i)
        nx = int(rng.integers(64, 385)) | 1
        ny = int(rng.integers(64, 385)) | 1
        nl = int(rng.integers(40, 71))
        cfg = moist3d_config(n=nx, n_layers=nl, dtime=3.22, seed=seed0 + i)
        out.append(dataclasses.replace(cfg, ny=ny, nx=nx))
    return out


def land_mask(cfg: RunConfig, seed: int = 7):
    """Seeded land/sea mask (set-up extension): land where a smoothed Gaussian field exceeds 0.3."""
    from .init_state import _smooth_periodic

    rng = np.random.default_rng(seed)
    return _smooth_periodic(rng.standard_normal((cfg.ny, cfg.nx)), max(2.0, 20e3 / cfg.dx)) > 0.3


def land_fraction(cfg: RunConfig, seed: int = 11):
    """Seeded fractional land cover (set-up extension): a smoothed Gaussian field rescaled to
    [0, 1], then 0 below 0.5 (sea), rising linearly to 1 at 0.9 (land), fractional between."""
    from .init_state import _smooth_periodic

    rng = np.random.default_rng(seed)
    f = _smooth_periodic(rng.standard_normal((cfg.ny, cfg.nx)), max(1.0, 10e3 / cfg.dx))
    f = (f - f.min()) / max(float(np.ptp(f)), 1e-300)
    return np.asfortranarray(np.clip(2.5 * (f - 0.5), 0.0, 1.0))


def soil_case(cfg: RunConfig, g: dict, st: dict, seed: int = 13) -> None:
    """Lakes and a varied soil temperature (set-up extension), so that every term of the soil
    budget is exercised: lakes cover 40 % of the non-land part of a seeded half of the columns
    (the heat capacity and conductivity follow from the fractions as in the reference), and the
    soil temperature is 280 K plus a seeded column offset of up to +-4 K at the surface, relaxing
    with depth; the sea surface temperature is the initial surface soil temperature, as in
    mo_set_initial_state.F90:246-247."""
    from .grid import soil_properties

    rng = np.random.default_rng(seed)
    lakes = rng.random((cfg.ny, cfg.nx)) < 0.5
    g["lake_fraction"] = np.asfortranarray(np.where(lakes, 0.4 * (1.0 - g["land_fraction"]), 0.0))
    soil_properties(g)
    ns = cfg.nsoillays
    offset = 4.0 * (2.0 * rng.random((cfg.ny, cfg.nx)) - 1.0)
    depth_weight = np.exp(-np.arange(ns) / 2.0)
    st["temperature_soil"] = np.asfortranarray(280.0 + offset[:, :, None] * depth_weight[None, None, :])
    g["sst"] = np.asfortranarray(st["temperature_soil"][:, :, 0].copy())


def make_case(cfg: RunConfig, seed: int | None = None, perturbation: bool = True, cloud_frac=0.0, grid=None):
    """Grid and initial state for ``cfg`` (perturbations only for the standard scenario).  With
    the surface physics on, the grid's land fraction is the seeded field of land_fraction; with
    the prognostic soil, soil_case sets lakes and the soil temperature.  ``grid`` is a grid of
    ``cfg`` made elsewhere (lgame.gpu.grid_setup_device), else the host generator makes it."""
    from .grid import soil_properties

    g = grid if grid is not None else grid_setup(cfg)
    if cfg.surface:
        g["land_fraction"] = land_fraction(cfg)
        soil_properties(g)
    st = ideal_init(cfg, g)
    if cfg.scenario == "standard" and perturbation:
        perturb(st, g, cfg, seed=(seed if seed is not None else 2))
    if cfg.lmoist:
        add_moisture(st, g, cfg, seed=(seed if seed is not None else 3) + 1000, cloud_frac=cloud_frac)
    if cfg.lprog_soil_temp:
        soil_case(cfg, g, st)
    return g, st


def cloud_patches(cfg: RunConfig, st: dict, seed: int = 17, offset: float = 0.0) -> None:
    """Patchy clouds (set-up extension): scales the condensates of every column by a smooth
    random cloud cover in [0, 1] (correlation length 2 grid points, the construction the
    radiation surrogate's training columns use)."""
    from .init_state import _smooth_periodic

    rng = np.random.default_rng(seed)
    cover = np.clip(_smooth_periodic(rng.standard_normal((cfg.ny, cfg.nx)), 2.0) * 0.6 + offset, 0.0, 1.0)
    for jc in range(cfg.n_condensed_constituents):
        st["rho"][:, :, :, jc] *= cover[:, :, None]


def tiny_config(ny=5, nx=7, n_layers=6, moist=False, order=2, freeslip=False, corio=True, plane=False, **kw) -> RunConfig:
    cfg = RunConfig(ny=ny, nx=nx, n_layers=n_layers, n_oro_layers=n_layers - 2, dy=2000.0, dx=2000.0,
                    toa=12000.0, stretching_parameter=1.3, scenario="standard", lplane=plane,
                    lat_center=LAT_CENTER_STD, lon_center=LON_CENTER_STD, theta_adv_order=order,
                    lmoist=moist, lcorio=corio, orography_id=5, lsleve=True, lperiodic=True,
                    lfreeslip=freeslip, dtime=2.0, klemp_begin_rel=0.5, **kw)
    return dataclasses.replace(cfg, hills=((0.0, 0.0, 400.0, 4000.0), (3000.0, -2000.0, 250.0, 3000.0)))


def radiation_surface(cfg: RunConfig, g: dict, st: dict, seed: int = 7) -> dict:
    """Per-column surface data of the radiation call (set-up extension): geographic coordinates
    of the grid, a land/sea mask from seeded random blobs (land where a smoothed Gaussian field
    exceeds 0.3), albedo 0.2 on land and 0.06 on sea, and a surface temperature 1 K (sea) or 3 K
    (land) above the temperature of the lowest layer of the initial state."""
    land = land_mask(cfg, seed)
    nl = cfg.n_layers
    t_low = (g["theta_v_bg"][:, :, nl - 1] + st["theta_v_pert"][:, :, nl - 1]) * (g["exner_bg"][:, :, nl - 1] +
                                                                              st["exner_pert"][:, :, nl - 1])
    return {
        "lat_geo": g["lat_geo_scalar"],
        "lon_geo": g["lon_geo_scalar"],
        "sfc_albedo": np.asfortranarray(np.where(land, 0.2, 0.06)),
        "temp_sfc": np.asfortranarray(t_low + np.where(land, 3.0, 1.0)),
        "land_fraction": np.asfortranarray(land.astype(np.float64)),
    }

