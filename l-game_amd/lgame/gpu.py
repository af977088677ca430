"""Python binding of the MI355X step library (``liblgame_gpu.so``, C-ABI in include/lgame_gpu.h).

Mirrors the reference's stepping interface: ``Step.pchevi(slot_old, slot_new,
total_step_counter)`` is the device-resident form of ``pchevi`` (mo_manage_pchevi.F90:27) and
``Step.run`` reproduces the ping-pong of the reference time loop (control.F90:520-524).
There is no CPU fallback: constructing a ``Step`` without the built library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .config import GRID_FIELDS, GRID_SCALARS, STATE_FIELDS, RunConfig, state_shape

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "liblgame_gpu.so")

EXPORTED = ("lgame_abi_version", "lgame_grid_setup", "lgame_init", "lgame_set_state", "lgame_get_state", "lgame_linear_combine", "lgame_pchevi",
            "lgame_pchevi_host", "lgame_sync", "lgame_integrals", "lgame_check_finite", "lgame_get_diag",
            "lgame_finalize", "lgame_last_error", "lgame_profile", "lgame_profile_read", "lgame_profile_reset",
            "lgame_rad_set_surface", "lgame_rad_columns", "lgame_rad_get", "lgame_rad_attach_mlp", "lgame_rad_stats",
            "lgame_mlp_create", "lgame_mlp_n_params", "lgame_mlp_set_params", "lgame_mlp_get_params",
            "lgame_mlp_set_norm", "lgame_mlp_forward", "lgame_mlp_forward_noalloc", "lgame_mlp_gradient", "lgame_mlp_load_data",
            "lgame_mlp_train", "lgame_mlp_last_error", "lgame_mlp_destroy", "lgame_autotune", "lgame_set_tiles",
            "lgame_set_segments", "lgame_get_segments", "lgame_set_column_passes", "lgame_get_column_passes",
            "lgame_ensemble_create", "lgame_ensemble_step", "lgame_ensemble_sync", "lgame_ensemble_last_error",
            "lgame_ensemble_destroy", "lgame_checkpoint_size", "lgame_checkpoint_save", "lgame_checkpoint_load")

N_TILES = 21
TILE_KERNELS = ("temperature", "theta_fill", "theta_grad", "theta_div", "theta_edge", "pgrad", "pot_vort",
                "mass_flux", "vort_flux", "ke", "ke_grad", "dissipation", "vec_tend", "scalar_flux", "scalar_div",
                "diff_zeta", "diff_coef", "diff_dual", "diff_mom", "diff_flux", "diff_div")

PROF_SOLVE_VER = 0
PROF_SOLVE_GEN = 1
PROF_STEP = 2
PROF_RADIATION = 3
PROF_KERNEL0 = 8  # + index in TILE_KERNELS, with profile(2)

RAD_OUT_TENDENCY = 0
RAD_OUT_SW_IN = 1
RAD_OUT_LW_OUT = 2
RAD_SURFACE_FIELDS = ("lat_geo", "lon_geo", "sfc_albedo", "temp_sfc")

LGAME_DIAG_TEMPERATURE = 0
# diagnostics readable with Step.diag: id and extent as a function of (ny, nx, nl)
DIAGS = {
    "temperature": (0, lambda ny, nx, nl: (ny, nx, nl)),
    "tke": (1, lambda ny, nx, nl: (ny, nx, nl)),
    "viscosity": (2, lambda ny, nx, nl: (ny, nx, nl)),
    "mom_diff_tend_x": (3, lambda ny, nx, nl: (ny, nx + 1, nl)),
    "mom_diff_tend_z": (4, lambda ny, nx, nl: (ny, nx, nl + 1)),
    "temp_diff_heating": (5, lambda ny, nx, nl: (ny, nx, nl)),
    "heating_diss": (6, lambda ny, nx, nl: (ny, nx, nl)),
    "roughness_length": (7, lambda ny, nx, nl: (ny, nx)),
    "monin_obukhov_length": (8, lambda ny, nx, nl: (ny, nx)),
    "scalar_flux_resistance": (9, lambda ny, nx, nl: (ny, nx)),
    "power_flux_density_sens_soil": (10, lambda ny, nx, nl: (ny, nx)),
    "power_flux_density_sens_sea": (11, lambda ny, nx, nl: (ny, nx)),
}


class LgameError(RuntimeError):
    pass


class GridView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in GRID_FIELDS] + [(n, ctypes.c_double) for n in GRID_SCALARS]


class GridParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("dx", "dy", "toa", "stretching_parameter", "lat_center", "lon_center")] + \
        [(n, ctypes.c_int) for n in ("orography_id", "lsleve", "lplane", "lfreeslip", "n_hills")] + \
        [("hills", ctypes.c_void_p), ("oro", ctypes.c_void_p)]


GRID_GEO = ("lat_geo_scalar", "lon_geo_scalar")


class GridOut(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in tuple(GRID_FIELDS) + GRID_GEO] + [(n, ctypes.c_double) for n in GRID_SCALARS]


def grid_setup_device(cfg: RunConfig, oro=None) -> dict:
    """The grid of cfg generated on the device (lgame_grid_setup, grid_setup and bg_setup of
    mo_grid_generator.F90:29-1457): the same dict as lgame.grid.grid_setup, including the
    geographic coordinates of the scalar points.  With orography_id 1, `oro` (ny, nx) is the
    orography before the smoothing."""
    if not cfg.lperiodic:
        raise ValueError("only periodic grids are generated")
    lib = load_library()
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    ns = max(cfg.nsoillays, 1)
    out = {k: np.zeros(GRID_FIELDS[k](ny, nx, nl, ns) if k.startswith("z_soil") else GRID_FIELDS[k](ny, nx, nl), order="F")
           for k in GRID_FIELDS}
    out.update({k: np.zeros((ny, nx), order="F") for k in GRID_GEO})
    hills = np.ascontiguousarray(np.array(cfg.hills, dtype=np.float64).reshape(-1))
    oro_f = _f64(oro) if oro is not None else None
    p = GridParams(cfg.dx, cfg.dy, cfg.toa, cfg.stretching_parameter, cfg.lat_center, cfg.lon_center,
                   cfg.orography_id, int(cfg.lsleve), int(cfg.lplane), int(cfg.lfreeslip), len(cfg.hills),
                   hills.ctypes.data if len(cfg.hills) else None, oro_f.ctypes.data if oro_f is not None else None)
    o = GridOut(*[out[k].ctypes.data for k in tuple(GRID_FIELDS) + GRID_GEO])
    d, _ = cfg.to_abi()
    rc = lib.lgame_grid_setup(ctypes.byref(d), ctypes.byref(p), ctypes.byref(o))
    if rc != 0:
        raise LgameError(f"lgame_grid_setup failed ({rc}): {lib.lgame_last_error(None).decode()}")
    out.update({k: float(getattr(o, k)) for k in GRID_SCALARS})
    return out


class RadView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in RAD_SURFACE_FIELDS]


class StateView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in STATE_FIELDS] + [("temperature_soil", ctypes.c_void_p)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load liblgame_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LgameError(f"{path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    vp, ci, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.lgame_abi_version.restype = ci
    lib.lgame_init.argtypes = [vp, vp, vp, ctypes.POINTER(vp)]
    lib.lgame_grid_setup.argtypes = [vp, vp, vp]
    lib.lgame_set_state.argtypes = [vp, ci, vp]
    lib.lgame_get_state.argtypes = [vp, ci, vp]
    lib.lgame_linear_combine.argtypes = [vp, ci, ci, cd, cd, vp]
    lib.lgame_pchevi.argtypes = [vp, ci, ci, ci, ci, cd]
    lib.lgame_pchevi_host.argtypes = [vp, vp, vp, ci, ci, cd]
    lib.lgame_sync.argtypes = [vp]
    lib.lgame_integrals.argtypes = [vp, ci, vp]
    lib.lgame_check_finite.argtypes = [vp, ci, ctypes.POINTER(ctypes.c_longlong)]
    lib.lgame_get_diag.argtypes = [vp, ci, vp]
    lib.lgame_finalize.argtypes = [vp]
    lib.lgame_finalize.restype = None
    lib.lgame_last_error.argtypes = [vp]
    lib.lgame_last_error.restype = ctypes.c_char_p
    lib.lgame_profile.argtypes = [vp, ci]
    lib.lgame_profile_read.argtypes = [vp, ci, ctypes.POINTER(cd), ctypes.POINTER(ctypes.c_longlong)]
    lib.lgame_profile_reset.argtypes = [vp]
    ll, ull, fp = ctypes.c_longlong, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_float)
    lib.lgame_rad_set_surface.argtypes = [vp, vp]
    lib.lgame_rad_columns.argtypes = [vp, ci, cd, vp, vp]
    lib.lgame_rad_get.argtypes = [vp, ci, vp]
    lib.lgame_rad_attach_mlp.argtypes = [vp, vp, cd, cd, ull]
    lib.lgame_rad_stats.argtypes = [vp, ctypes.POINTER(ll), ctypes.POINTER(ll), ctypes.POINTER(ll), ctypes.POINTER(cd)]
    lib.lgame_mlp_create.argtypes = [ci, ci, ci, ci, ull, ctypes.POINTER(vp)]
    lib.lgame_mlp_n_params.argtypes = [vp]
    lib.lgame_mlp_n_params.restype = ll
    lib.lgame_mlp_set_params.argtypes = [vp, vp, ll]
    lib.lgame_mlp_get_params.argtypes = [vp, vp, ll]
    lib.lgame_mlp_set_norm.argtypes = [vp, vp, vp, vp, vp]
    lib.lgame_mlp_forward.argtypes = [vp, vp, ll, vp]
    lib.lgame_mlp_forward_noalloc.argtypes = [vp, vp, ll, vp]
    lib.lgame_mlp_gradient.argtypes = [vp, vp, vp, ll, vp, fp]
    lib.lgame_mlp_load_data.argtypes = [vp, vp, vp, ll]
    lib.lgame_mlp_train.argtypes = [vp, ci, ci, ctypes.c_float, ull, vp, ctypes.POINTER(cd)]
    lib.lgame_mlp_last_error.argtypes = [vp]
    lib.lgame_mlp_last_error.restype = ctypes.c_char_p
    lib.lgame_mlp_destroy.argtypes = [vp]
    lib.lgame_mlp_destroy.restype = None
    lib.lgame_autotune.argtypes = [vp, ci, vp]
    lib.lgame_set_tiles.argtypes = [vp, vp]
    lib.lgame_set_segments.argtypes = [vp, ci]
    lib.lgame_get_segments.argtypes = [vp, ctypes.POINTER(ci)]
    lib.lgame_set_column_passes.argtypes = [vp, ci]
    lib.lgame_get_column_passes.argtypes = [vp, ctypes.POINTER(ci)]
    lib.lgame_ensemble_create.argtypes = [vp, ci, ctypes.POINTER(vp)]
    lib.lgame_ensemble_step.argtypes = [vp, ci, ci, cd, ci]
    lib.lgame_ensemble_sync.argtypes = [vp]
    lib.lgame_ensemble_last_error.argtypes = [vp]
    lib.lgame_ensemble_last_error.restype = ctypes.c_char_p
    lib.lgame_ensemble_destroy.argtypes = [vp]
    lib.lgame_ensemble_destroy.restype = None
    lib.lgame_checkpoint_size.argtypes = [vp, ctypes.POINTER(ll)]
    lib.lgame_checkpoint_save.argtypes = [vp, vp, ll]
    lib.lgame_checkpoint_load.argtypes = [vp, vp, ll]
    _lib = lib
    return lib


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class Step:
    """One handle of the device step for one configuration and grid."""

    def __init__(self, cfg: RunConfig, grid: dict):
        self.lib = load_library()
        self.cfg = cfg
        self.dims, self.ccfg = cfg.to_abi()
        keep = {n: _f64(grid[n]) for n in GRID_FIELDS}
        gv = GridView(*[keep[n].ctypes.data for n in GRID_FIELDS], *[float(grid[n]) for n in GRID_SCALARS])
        h = ctypes.c_void_p()
        rc = self.lib.lgame_init(ctypes.byref(self.dims), ctypes.byref(self.ccfg), ctypes.byref(gv), ctypes.byref(h))
        if rc != 0:
            raise LgameError(f"lgame_init failed ({rc}): {self.lib.lgame_last_error(None).decode()}")
        self.h = h

    def _check(self, rc, what):
        if rc != 0:
            raise LgameError(f"{what} failed ({rc}): {self.lib.lgame_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.lgame_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _view(self, st: dict):
        arrs = {k: _f64(st[k]) for k in self.cfg.state_fields}
        soil = arrs["temperature_soil"].ctypes.data if "temperature_soil" in arrs else None
        return arrs, StateView(*[arrs[k].ctypes.data for k in STATE_FIELDS], soil)

    def set_state(self, slot: int, st: dict):
        arrs, v = self._view(st)
        self._check(self.lib.lgame_set_state(self.h, slot, ctypes.byref(v)), "lgame_set_state")

    def get_state(self, slot: int) -> dict:
        c = self.cfg
        out = {k: np.zeros(state_shape(k, c.ny, c.nx, c.n_layers, c.n_constituents, c.nsoillays), order="F")
               for k in c.state_fields}
        soil = out["temperature_soil"].ctypes.data if "temperature_soil" in out else None
        v = StateView(*[out[k].ctypes.data for k in STATE_FIELDS], soil)
        self._check(self.lib.lgame_get_state(self.h, slot, ctypes.byref(v)), "lgame_get_state")
        return out

    def linear_combine(self, slot_1: int, slot_2: int, coeff_1: float, coeff_2: float) -> dict:
        """coeff_1*slot_1 + coeff_2*slot_2 as a host state (linear_combine_two_states,
        mo_linear_combination.F90:15-37; the output-time interpolation of control.F90:539)."""
        c = self.cfg
        out = {k: np.zeros(state_shape(k, c.ny, c.nx, c.n_layers, c.n_constituents, c.nsoillays), order="F")
               for k in c.state_fields}
        soil = out["temperature_soil"].ctypes.data if "temperature_soil" in out else None
        v = StateView(*[out[k].ctypes.data for k in STATE_FIELDS], soil)
        self._check(self.lib.lgame_linear_combine(self.h, slot_1, slot_2, coeff_1, coeff_2, ctypes.byref(v)),
                    "lgame_linear_combine")
        return out

    def start(self, st: dict):
        """Initial state into both slots (state_2 = state_1, control.F90:500)."""
        self.set_state(0, st)
        self.set_state(1, st)

    def pchevi(self, slot_old: int, slot_new: int, total_step_counter: int, lrad_update: bool = False, t_0: float = 0.0):
        self._check(self.lib.lgame_pchevi(self.h, slot_old, slot_new, total_step_counter, int(lrad_update), t_0), "lgame_pchevi")

    def run(self, first: int, nsteps: int, t_init: float | None = None):
        """Ping-pong steps first..first+nsteps-1.  With t_init, the radiation is updated as in
        control.F90:510-517 (when t_0 <= t_rad_update <= t_0 + dtime, t_rad_update starting at
        t_init and advancing by dtime_rad)."""
        dt = self.cfg.dt
        t_rad_update = None
        if t_init is not None:  # the first update time at or after the first step's t_0
            k = max(0, int(np.ceil((first * dt) / self.ccfg.dtime_rad - 1e-12)))
            t_rad_update = t_init + k * self.ccfg.dtime_rad
        for s in range(first, first + nsteps):
            lrad, t0 = False, 0.0
            if t_rad_update is not None:
                t0 = t_init + s * dt
                if self.cfg.lrad and t0 <= t_rad_update and t0 + dt >= t_rad_update:
                    lrad = True
                    t_rad_update += self.ccfg.dtime_rad
            if s % 2 == 0:
                self.pchevi(0, 1, s, lrad, t0)
            else:
                self.pchevi(1, 0, s, lrad, t0)

    def autotune(self, steps: int = 2) -> dict:
        """Pick the block shape of every stencil kernel by timing (before the first step)."""
        rows = np.zeros(N_TILES, dtype=np.int32)
        self._check(self.lib.lgame_autotune(self.h, steps, rows.ctypes.data), "lgame_autotune")
        return dict(zip(TILE_KERNELS, rows.tolist()))

    def set_tiles(self, rows):
        r = np.ascontiguousarray([rows[k] for k in TILE_KERNELS] if isinstance(rows, dict) else rows, dtype=np.int32)
        self._check(self.lib.lgame_set_tiles(self.h, r.ctypes.data), "lgame_set_tiles")

    @property
    def segments(self) -> int:
        """Vertical segments of the column-marching scalar kernel."""
        n = ctypes.c_int(0)
        self._check(self.lib.lgame_get_segments(self.h, ctypes.byref(n)), "lgame_get_segments")
        return n.value

    @segments.setter
    def segments(self, nseg: int):
        self._check(self.lib.lgame_set_segments(self.h, int(nseg)), "lgame_set_segments")

    @property
    def column_passes(self) -> int:
        """The vertical solver in one pass per column (1) or in three passes (3)."""
        n = ctypes.c_int(0)
        self._check(self.lib.lgame_get_column_passes(self.h, ctypes.byref(n)), "lgame_get_column_passes")
        return n.value

    @column_passes.setter
    def column_passes(self, passes: int):
        self._check(self.lib.lgame_set_column_passes(self.h, int(passes)), "lgame_set_column_passes")

    # ---- radiation ----
    def set_rad_surface(self, surface: dict):
        keep = {k: _f64(surface[k]) for k in RAD_SURFACE_FIELDS}
        v = RadView(*[keep[k].ctypes.data for k in RAD_SURFACE_FIELDS])
        self._check(self.lib.lgame_rad_set_surface(self.h, ctypes.byref(v)), "lgame_rad_set_surface")

    def rad_columns(self, slot: int, t: float):
        """(x, y): surrogate inputs and solver outputs of every column, Fortran column order."""
        c = self.cfg
        ncol = c.ny * c.nx
        x = np.zeros((ncol, 5 * c.n_layers + 4), dtype=np.float32)
        y = np.zeros((ncol, c.n_layers + 2), dtype=np.float32)
        self._check(self.lib.lgame_rad_columns(self.h, slot, t, x.ctypes.data, y.ctypes.data), "lgame_rad_columns")
        return x, y

    def rad_get(self, which: int) -> np.ndarray:
        c = self.cfg
        shape = (c.ny, c.nx, c.n_layers) if which == RAD_OUT_TENDENCY else (c.ny, c.nx)
        out = np.zeros(shape, order="F")
        self._check(self.lib.lgame_rad_get(self.h, which, out.ctypes.data), "lgame_rad_get")
        return out

    def attach_mlp(self, mlp: "Mlp", tol_k_per_day: float = 1.0, audit_fraction: float = 0.01, seed: int = 0):
        self._mlp = mlp  # the handle keeps a pointer to it
        self._check(self.lib.lgame_rad_attach_mlp(self.h, mlp.m, tol_k_per_day, audit_fraction, seed), "lgame_rad_attach_mlp")

    def rad_stats(self) -> dict:
        a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        r = ctypes.c_double()
        self._check(self.lib.lgame_rad_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(r)),
                    "lgame_rad_stats")
        return {"calls": a.value, "audited": b.value, "fallbacks": c.value, "audit_rmse_k_per_day": r.value}

    def sync(self):
        self._check(self.lib.lgame_sync(self.h), "lgame_sync")

    def checkpoint(self) -> bytes:
        """The carried fields besides the two states (lgame_checkpoint_save), as bytes."""
        n = ctypes.c_longlong(0)
        self._check(self.lib.lgame_checkpoint_size(self.h, ctypes.byref(n)), "lgame_checkpoint_size")
        buf = ctypes.create_string_buffer(n.value)
        self._check(self.lib.lgame_checkpoint_save(self.h, buf, n), "lgame_checkpoint_save")
        return buf.raw

    def restore(self, blob: bytes):
        """Load a checkpoint of a handle of the same configuration (lgame_checkpoint_load)."""
        self._check(self.lib.lgame_checkpoint_load(self.h, blob, len(blob)), "lgame_checkpoint_load")

    def integrals(self, slot: int) -> np.ndarray:
        out = np.zeros(4 + self.cfg.n_constituents)
        self._check(self.lib.lgame_integrals(self.h, slot, out.ctypes.data), "lgame_integrals")
        return out

    def count_nonfinite(self, slot: int) -> int:
        n = ctypes.c_longlong(0)
        rc = self.lib.lgame_check_finite(self.h, slot, ctypes.byref(n))
        if rc not in (0, 4):
            self._check(rc, "lgame_check_finite")
        return int(n.value)

    def profile(self, enable: int = 1):
        """1: time the solver, step and radiation groups; 2: also every stencil kernel."""
        self._check(self.lib.lgame_profile(self.h, int(enable)), "lgame_profile")

    def profile_read(self, group: int):
        """(total milliseconds, launch count) recorded for a kernel group since the last reset."""
        ms, n = ctypes.c_double(0.0), ctypes.c_longlong(0)
        self._check(self.lib.lgame_profile_read(self.h, group, ctypes.byref(ms), ctypes.byref(n)), "lgame_profile_read")
        return ms.value, n.value

    def profile_reset(self):
        self._check(self.lib.lgame_profile_reset(self.h), "lgame_profile_reset")

    def temperature(self) -> np.ndarray:
        return self.diag("temperature")

    def diag(self, name: str) -> np.ndarray:
        """A carried diagnostic field (DIAGS) in the reference layout."""
        which, shape = DIAGS[name]
        c = self.cfg
        out = np.zeros(shape(c.ny, c.nx, c.n_layers), order="F")
        self._check(self.lib.lgame_get_diag(self.h, which, out.ctypes.data), "lgame_get_diag")
        return out


class Mlp:
    """The per-column MLP of the radiation surrogate (BF16 MFMA kernels, FP32 Adam)."""

    def __init__(self, n_in: int, hidden: int, n_hidden_layers: int, n_out: int, seed: int = 0):
        self.lib = load_library()
        m = ctypes.c_void_p()
        rc = self.lib.lgame_mlp_create(n_in, hidden, n_hidden_layers, n_out, seed, ctypes.byref(m))
        if rc != 0:
            raise LgameError(f"lgame_mlp_create failed ({rc})")
        self.m = m
        self.n_in, self.n_out = n_in, n_out
        self.hidden, self.n_hidden_layers = hidden, n_hidden_layers
        self.n_params = int(self.lib.lgame_mlp_n_params(m))

    def _check(self, rc, what):
        if rc != 0:
            raise LgameError(f"{what} failed ({rc}): {self.lib.lgame_mlp_last_error(self.m).decode()}")

    def close(self):
        if getattr(self, "m", None):
            self.lib.lgame_mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self._check(self.lib.lgame_mlp_get_params(self.m, p.ctypes.data, self.n_params), "lgame_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.lib.lgame_mlp_set_params(self.m, p.ctypes.data, p.size), "lgame_mlp_set_params")

    def set_norm(self, x_mean, x_scale, y_mean, y_scale):
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (x_mean, x_scale, y_mean, y_scale)]
        self._norm = arrs
        self._check(self.lib.lgame_mlp_set_norm(self.m, *[a.ctypes.data for a in arrs]), "lgame_mlp_set_norm")

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        self._check(self.lib.lgame_mlp_forward(self.m, x.ctypes.data, x.shape[0], y.ctypes.data), "lgame_mlp_forward")
        return y

    def forward_noalloc(self, x) -> np.ndarray:
        """forward through lgame_mlp_forward_noalloc (no device allocation)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        self._check(self.lib.lgame_mlp_forward_noalloc(self.m, x.ctypes.data, x.shape[0], y.ctypes.data),
                    "lgame_mlp_forward_noalloc")
        return y

    def gradient(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self.lib.lgame_mlp_gradient(self.m, x.ctypes.data, y.ctypes.data, x.shape[0], g.ctypes.data,
                                                ctypes.byref(loss)), "lgame_mlp_gradient")
        return g, float(loss.value)

    def load_data(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        self._check(self.lib.lgame_mlp_load_data(self.m, x.ctypes.data, y.ctypes.data, x.shape[0]), "lgame_mlp_load_data")

    def train(self, epochs: int, batch: int, lr: float, seed: int = 0):
        loss = np.zeros(epochs, dtype=np.float32)
        sec = ctypes.c_double()
        self._check(self.lib.lgame_mlp_train(self.m, epochs, batch, lr, seed, loss.ctypes.data, ctypes.byref(sec)),
                    "lgame_mlp_train")
        return loss, sec.value


class Ensemble:
    """Independent domains advanced together: BATCHED launches every kernel phase once over all
    members (grouped by kernel sequence), GRAPH replays one graph with a branch per member,
    SERIAL runs the members one after another."""

    GRAPH, SERIAL, BATCHED = 0, 1, 2

    def __init__(self, members):
        self.lib = load_library()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h for m in self.members])
        e = ctypes.c_void_p()
        rc = self.lib.lgame_ensemble_create(arr, len(self.members), ctypes.byref(e))
        if rc != 0:
            raise LgameError(f"lgame_ensemble_create failed ({rc})")
        self.e = e

    def _check(self, rc, what):
        if rc != 0:
            raise LgameError(f"{what} failed ({rc}): {self.lib.lgame_ensemble_last_error(self.e).decode()}")

    def step(self, counter: int, lrad_update: bool = False, t_0: float = 0.0, mode: int = 0):
        self._check(self.lib.lgame_ensemble_step(self.e, counter, int(lrad_update), t_0, mode), "lgame_ensemble_step")

    def run(self, first: int, nsteps: int, mode: int = 0):
        for k in range(first, first + nsteps):
            self.step(k, mode=mode)

    def sync(self):
        self._check(self.lib.lgame_ensemble_sync(self.e), "lgame_ensemble_sync")

    def close(self):
        if getattr(self, "e", None):
            self.lib.lgame_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
