"""Run configuration: the reference namelist variables that the step and its set-up read.

Defaults follow the reference namelist set-up routines:
``mo_run_nml.F90:55-80`` (run), ``mo_diff_nml.F90:39-53`` (diff), ``mo_surface_nml.F90:31-37``
(surface), ``mo_bc_nml.F90:30-35`` (bc), ``mo_rad_nml.F90:31-36`` (rad),
``mo_constituents_nml.F90:26-43`` (constituents).  ``dtime`` is derived as in
``mo_run_nml.F90:93-94`` unless set explicitly.

``to_abi`` produces the ``lgame_dims`` / ``lgame_cfg`` C structs of ``include/lgame_gpu.h``.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

PREC_FP64 = 0
PREC_MIXED = 1
RAD_OFF = 0
RAD_TEACHER = 1
RAD_SURROGATE = 2


class LgameDims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "ny", "nx", "n_layers", "n_oro_layers", "n_constituents", "n_condensed_constituents",
        "nsoillays")]


class LgameCfg(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "dtime", "toa", "dx", "dy", "klemp_damp_max", "klemp_begin_rel", "dtime_rad", "c_s", "bg_shear",
        "h_prandtl", "karman")] + \
        [(n, ctypes.c_int) for n in (
            "theta_adv_order", "luse_bg_state", "llinear", "lcorio", "lmoist", "lklemp",
            "lmom_diff_h", "lmom_diff_v", "ltemp_diff_h", "ltemp_diff_v", "lmass_diff_h",
            "lmass_diff_v", "lprog_soil_temp", "lsfc_sensible_heat_flux", "lsfc_phase_trans",
            "lpbl", "lperiodic", "lfreeslip", "lrad", "precision", "rad_mode", "diff_coeff_scheme_h")]


@dataclasses.dataclass
class RunConfig:
    # &run
    ny: int = 35
    nx: int = 37
    n_layers: int = 50
    n_oro_layers: int = 40
    dy: float = 25e3
    dx: float = 25e3
    toa: float = 40000.0
    stretching_parameter: float = 1.3
    scenario: str = "standard"
    llinear: bool = False
    lcorio: bool = True
    lplane: bool = False
    lat_center: float = 0.0
    lon_center: float = 0.0
    theta_adv_order: int = 2
    luse_bg_state: bool = True
    dtime: float | None = None
    # &constituents
    lmoist: bool = False
    # &diff (the reference turns all six diffusion switches on; the cases here choose)
    lmom_diff_h: bool = False
    lmom_diff_v: bool = False
    ltemp_diff_h: bool = False
    ltemp_diff_v: bool = False
    lmass_diff_h: bool = False
    lmass_diff_v: bool = False
    diff_coeff_scheme_h: str = "smag"
    c_s: float = 0.2
    bg_shear: float = 1.5e-5
    h_prandtl: float = 100.0
    karman: float = 0.4
    lklemp: bool = True
    klemp_damp_max: float = 0.25
    klemp_begin_rel: float = 0.53
    # &surface
    orography_id: int = 0
    lsleve: bool = False
    nsoillays: int = 5
    lprog_soil_temp: bool = False
    lsfc_sensible_heat_flux: bool = False
    lsfc_phase_trans: bool = False
    lpbl: bool = False
    # &bc
    lperiodic: bool = True
    lfreeslip: bool = False
    # &rad
    lrad: bool = False
    dtime_rad: float | None = None
    # set-up extensions (not reference namelist variables)
    hills: tuple = ()            # ((x0, y0, height, half_width), ...) in metres, orography_id == 5
    precision: int = PREC_FP64
    rad_mode: int = RAD_OFF      # LGAME_RAD_* of the radiation call (with lrad)

    @property
    def n_levels(self) -> int:
        return self.n_layers + 1

    @property
    def n_flat_layers(self) -> int:
        return self.n_layers - self.n_oro_layers

    @property
    def n_condensed_constituents(self) -> int:
        return 5 if self.lmoist else 0

    @property
    def n_constituents(self) -> int:
        return 7 if self.lmoist else 1

    @property
    def dt(self) -> float:
        if self.dtime is not None:
            return float(self.dtime)
        # mo_run_nml.F90:93-94
        return 1.61 * math.sqrt(self.dx * self.dy) / 1000.0

    def set_diffusion(self, on: bool = True):
        """All six diffusion switches at once (the reference default is all on)."""
        for k in ("lmom_diff_h", "lmom_diff_v", "ltemp_diff_h", "ltemp_diff_v", "lmass_diff_h", "lmass_diff_v"):
            setattr(self, k, bool(on))
        return self

    @property
    def diffusion(self) -> bool:
        return any((self.lmom_diff_h, self.lmom_diff_v, self.ltemp_diff_h, self.lmass_diff_h))

    @property
    def surface(self) -> bool:
        """Whether the step reads the surface fields (land fraction, roughness length)."""
        return self.lpbl or self.lsfc_sensible_heat_flux or self.lsfc_phase_trans or self.lprog_soil_temp

    @property
    def state_fields(self) -> tuple:
        """The prognostic fields of a state: the soil temperature joins with the prognostic soil."""
        return STATE_FIELDS + (("temperature_soil",) if self.lprog_soil_temp else ())

    @property
    def gridpoints(self) -> int:
        return self.ny * self.nx * self.n_layers

    def to_abi(self):
        d = LgameDims(self.ny, self.nx, self.n_layers, self.n_oro_layers, self.n_constituents,
                      self.n_condensed_constituents, self.nsoillays)
        c = LgameCfg()
        c.dtime = self.dt
        c.toa = self.toa
        c.dx = self.dx
        c.dy = self.dy
        c.klemp_damp_max = self.klemp_damp_max
        c.klemp_begin_rel = self.klemp_begin_rel
        c.dtime_rad = self.dtime_rad if self.dtime_rad is not None else 60.0 * self.dy / 1000.0
        c.theta_adv_order = self.theta_adv_order
        c.luse_bg_state = int(self.luse_bg_state)
        c.llinear = int(self.llinear)
        c.lcorio = int(self.lcorio)
        c.lmoist = int(self.lmoist)
        c.lklemp = int(self.lklemp)
        for k in ("lmom_diff_h", "lmom_diff_v", "ltemp_diff_h", "ltemp_diff_v", "lmass_diff_h", "lmass_diff_v"):
            setattr(c, k, int(getattr(self, k)))
        c.c_s = self.c_s
        c.bg_shear = self.bg_shear
        c.h_prandtl = self.h_prandtl
        c.karman = self.karman
        c.lpbl = int(self.lpbl)
        c.lprog_soil_temp = int(self.lprog_soil_temp)
        c.lsfc_sensible_heat_flux = int(self.lsfc_sensible_heat_flux)
        c.lsfc_phase_trans = int(self.lsfc_phase_trans)
        if self.diff_coeff_scheme_h not in ("smag", "tke"):
            raise ValueError("diff_coeff_scheme_h must be either smag or tke")
        c.diff_coeff_scheme_h = 0 if self.diff_coeff_scheme_h == "smag" else 1
        c.lperiodic = int(self.lperiodic)
        c.lfreeslip = int(self.lfreeslip)
        c.lrad = int(self.lrad)
        c.precision = self.precision
        c.rad_mode = self.rad_mode
        return d, c


# Reference extents of every array that crosses the boundary, as functions of (ny, nx, nl).
GRID_FIELDS = {
    "z_scalar": lambda ny, nx, nl: (ny, nx, nl),
    "z_u": lambda ny, nx, nl: (ny, nx + 1, nl),
    "z_v": lambda ny, nx, nl: (ny + 1, nx, nl),
    "z_w": lambda ny, nx, nl: (ny, nx, nl + 1),
    "dx": lambda ny, nx, nl: (ny, nx + 1, nl),
    "dy": lambda ny, nx, nl: (ny + 1, nx, nl),
    "dz": lambda ny, nx, nl: (ny, nx, nl + 1),
    "area_x": lambda ny, nx, nl: (ny, nx + 1, nl),
    "area_y": lambda ny, nx, nl: (ny + 1, nx, nl),
    "area_z": lambda ny, nx, nl: (ny, nx, nl + 1),
    "volume": lambda ny, nx, nl: (ny, nx, nl),
    "slope_x": lambda ny, nx, nl: (ny, nx + 1, nl),
    "slope_y": lambda ny, nx, nl: (ny + 1, nx, nl),
    "inner_product_weights": lambda ny, nx, nl: (6, ny, nx, nl),
    "area_dual_x": lambda ny, nx, nl: (ny + 1, nx, nl + 1),
    "area_dual_y": lambda ny, nx, nl: (ny, nx + 1, nl + 1),
    "area_dual_z": lambda ny, nx, nl: (ny + 1, nx + 1, nl),
    "z_area_dual_z": lambda ny, nx, nl: (ny + 1, nx + 1, nl),
    "fvec_x": lambda ny, nx, nl: (ny + 1, nx),
    "fvec_y": lambda ny, nx, nl: (ny, nx + 1),
    "fvec_z": lambda ny, nx, nl: (ny + 1, nx + 1),
    "trsk_weights_u": lambda ny, nx, nl: (3, ny),
    "trsk_weights_v": lambda ny, nx, nl: (2, ny + 1),
    "exner_bg_grad_u": lambda ny, nx, nl: (ny, nx + 1, nl),
    "exner_bg_grad_v": lambda ny, nx, nl: (ny + 1, nx, nl),
    "exner_bg_grad_w": lambda ny, nx, nl: (ny, nx, nl + 1),
    "theta_v_bg": lambda ny, nx, nl: (ny, nx, nl),
    "exner_bg": lambda ny, nx, nl: (ny, nx, nl),
    "gravity_m_v": lambda ny, nx, nl: (ny, nx, nl + 1),
    "gravity_potential": lambda ny, nx, nl: (ny, nx, nl),
    "dx_dual": lambda ny, nx, nl: (ny + 1, nx, nl),
    "dy_dual": lambda ny, nx, nl: (ny, nx + 1, nl),
    "layer_thickness": lambda ny, nx, nl: (ny, nx, nl),
    "land_fraction": lambda ny, nx, nl: (ny, nx),
    "roughness_length": lambda ny, nx, nl: (ny, nx),
    "lake_fraction": lambda ny, nx, nl: (ny, nx),
    "sfc_rho_c": lambda ny, nx, nl: (ny, nx),
    "t_conduc_soil": lambda ny, nx, nl: (ny, nx),
    "t_const_soil": lambda ny, nx, nl: (ny, nx),
    "sst": lambda ny, nx, nl: (ny, nx),  # diag%sst, constant over a run (mo_set_initial_state.F90:246-247)
    "z_soil_interface": lambda ny, nx, nl, ns=5: (ns + 1,),
    "z_soil_center": lambda ny, nx, nl, ns=5: (ns,),
}
# scalar grid quantities that cross the boundary with the fields
GRID_SCALARS = ("mean_velocity_area", "z_t_const")

STATE_FIELDS = ("rho", "rhotheta_v", "theta_v_pert", "exner_pert", "wind_u", "wind_v", "wind_w")


def state_shape(name: str, ny: int, nx: int, nl: int, nc: int, ns: int = 5):
    return {
        "temperature_soil": (ny, nx, ns),
        "rho": (ny, nx, nl, nc),
        "rhotheta_v": (ny, nx, nl),
        "theta_v_pert": (ny, nx, nl),
        "exner_pert": (ny, nx, nl),
        "wind_u": (ny, nx + 1, nl),
        "wind_v": (ny + 1, nx, nl),
        "wind_w": (ny, nx, nl + 1),
    }[name]
