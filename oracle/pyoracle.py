"""ctypes bindings of the two CPU oracles (TEST INFRASTRUCTURE ONLY).

* ``CpuOracle``  -- oracle/liboracle.so, the C restatement (lgame_oracle.c).
* ``RefHarness`` -- oracle/_ref/libref.so, the reference's own modules driven by ref_harness.F90
  (exists only where the reference tree was available to build it).

Only the test suite, the smoke check and the CPU-baseline leg of the benchmark use this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

REF_GRID_IDS = {name: i + 1 for i, name in enumerate([
    "z_scalar", "z_u", "z_v", "z_w", "dx", "dy", "dz", "area_x", "area_y", "area_z", "volume",
    "slope_x", "slope_y", "inner_product_weights", "area_dual_x", "area_dual_y", "area_dual_z",
    "z_area_dual_z", "fvec_x", "fvec_y", "fvec_z", "trsk_weights_u", "trsk_weights_v",
    "exner_bg_grad_u", "exner_bg_grad_v", "exner_bg_grad_w", "theta_v_bg", "exner_bg",
    "gravity_m_v", "gravity_potential", "dx_dual", "dy_dual", "layer_thickness", "land_fraction",
    "roughness_length", "lake_fraction", "sfc_rho_c", "t_conduc_soil", "t_const_soil", "sst",
    "z_soil_interface", "z_soil_center", "z_t_const"])}
REF_STATE_IDS = {"rho": 1, "rhotheta_v": 2, "theta_v_pert": 3, "exner_pert": 4, "wind_u": 5,
                 "wind_v": 6, "wind_w": 7, "temperature_soil": 8}
REF_DIAG_IDS = {"temperature": 1, "heating_diss": 2, "pot_vort_tend_x": 3, "pot_vort_tend_z": 4,
                "condensates_sediment_heat": 5, "tke": 6, "mom_diff_tend_x": 7, "mom_diff_tend_y": 8,
                "mom_diff_tend_z": 9, "temp_diff_heating": 10, "mass_diff_tendency": 11, "viscosity": 12,
                "n_squared": 13, "roughness_length": 14, "roughness_velocity": 15, "monin_obukhov_length": 16,
                "scalar_flux_resistance": 17, "power_flux_density_sens_soil": 18,
                "power_flux_density_sens_sea": 19, "p_grad_acc_neg_nl_u": 20, "p_grad_acc_neg_nl_v": 21,
                "p_grad_acc_neg_nl_w": 22, "p_grad_acc_neg_l_u": 23, "p_grad_acc_neg_l_v": 24,
                "p_grad_acc_neg_l_w": 25, "p_grad_acc_old_u": 26, "p_grad_acc_old_v": 27,
                "p_grad_condensates_w": 28}


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _fortran(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class CpuOracle:
    """The C restatement of the reference step on one configuration."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_LIB):
                raise FileNotFoundError(f"{ORACLE_LIB} is not built (run `make -C oracle`)")
            lib = ctypes.CDLL(ORACLE_LIB)
            lib.orc_create.restype = ctypes.c_void_p
            lib.orc_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            for f in ("orc_set_grid", "orc_set_state", "orc_get_state", "orc_get_diag"):
                getattr(lib, f).restype = ctypes.c_int
            lib.orc_set_grid.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
            lib.orc_set_state.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p]
            lib.orc_get_state.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p]
            lib.orc_get_diag.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
            lib.orc_diag_size.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
            lib.orc_diag_size.restype = ctypes.c_longlong
            lib.orc_set_diag.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
            lib.orc_set_diag.restype = ctypes.c_int
            lib.orc_call.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            lib.orc_call.restype = ctypes.c_int
            lib.orc_finish_grid.argtypes = [ctypes.c_void_p]
            lib.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            lib.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            lib.orc_copy_slot.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            lib.orc_integrals.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
            lib.orc_diagnose_temperature.argtypes = [ctypes.c_void_p, ctypes.c_int]
            lib.orc_n_damping_levels.argtypes = [ctypes.c_void_p]
            lib.orc_destroy.argtypes = [ctypes.c_void_p]
            lib.orc_last_error.restype = ctypes.c_char_p
            lib.orc_step_rad.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
            lib.orc_set_rad_surface.argtypes = [ctypes.c_void_p] * 5
            lib.orc_set_rad_surface.restype = None
            lib.orc_rad_columns.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
            lib.orc_rad_columns.restype = None
            lib.orc_get_rad_sfc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
            lib.orc_coszenith.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
            lib.orc_coszenith.restype = ctypes.c_double
            cls._lib = lib
        return cls._lib

    def __init__(self, cfg, grid: dict):
        from lgame.config import GRID_FIELDS
        self.cfg = cfg
        self.dims, self.ccfg = cfg.to_abi()
        lib = self.lib()
        self.h = lib.orc_create(ctypes.byref(self.dims), ctypes.byref(self.ccfg))
        if not self.h:
            raise ValueError(lib.orc_last_error().decode())
        from lgame.config import GRID_SCALARS
        for name in list(GRID_FIELDS) + list(GRID_SCALARS):
            a = _fortran(np.atleast_1d(grid[name]))
            if lib.orc_set_grid(self.h, name.encode(), _ptr(a)) != 0:
                raise KeyError(name)
        lib.orc_finish_grid(self.h)

    def __del__(self):
        if getattr(self, "h", None) and self._lib is not None:
            self._lib.orc_destroy(self.h)
            self.h = None

    def set_state(self, slot: int, st: dict):
        for k, v in st.items():
            a = _fortran(v)
            if self.lib().orc_set_state(self.h, slot, k.encode(), _ptr(a)) != 0:
                raise KeyError(k)

    def get_state(self, slot: int) -> dict:
        from lgame.config import state_shape
        c = self.cfg
        out = {}
        for k in c.state_fields:
            a = np.zeros(state_shape(k, c.ny, c.nx, c.n_layers, c.n_constituents, c.nsoillays), order="F")
            self.lib().orc_get_state(self.h, slot, k.encode(), _ptr(a))
            out[k] = a
        return out

    def _check_size(self, name: str, count: int):
        n = self.lib().orc_diag_size(self.h, name.encode())
        if n == 0:
            raise KeyError(name)
        if n != count:
            raise ValueError(f"diagnostic {name} holds {n} values, not {count}")

    def get_diag(self, name: str, shape) -> np.ndarray:
        a = np.zeros(shape, order="F")
        self._check_size(name, a.size)
        if self.lib().orc_get_diag(self.h, name.encode(), _ptr(a)) != 0:
            raise KeyError(name)
        return a

    def set_diag(self, name: str, a):
        """A diagnostic field overwritten (test hook of the closed-form checks)."""
        a = _fortran(a)
        self._check_size(name, a.size)
        if self.lib().orc_set_diag(self.h, name.encode(), _ptr(a)) != 0:
            raise KeyError(name)

    def call(self, name: str, slot: int, rk_step: int = 1, counter: int = 1):
        """One subroutine of the step on a state slot (test hook of the closed-form checks):
        temperature_diagnostics, manage_p_grad (lfirst when counter is 0), vector_tend_expl,
        mom_diff_v, simple_dissipation_rate, calc_pot_vort."""
        lib = self.lib()
        if lib.orc_call(self.h, name.encode(), slot, rk_step, counter) != 0:
            raise KeyError(name)

    def linear_combine(self, slot_1: int, slot_2: int, coeff_1: float, coeff_2: float) -> dict:
        """coeff_1*slot_1 + coeff_2*slot_2 (linear_combine_two_states) as a state dict."""
        from lgame.config import state_shape
        c = self.cfg
        lib = self.lib()
        lib.orc_linear_combine.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_char_p, ctypes.c_void_p]
        lib.orc_linear_combine.restype = ctypes.c_int
        out = {}
        for k in c.state_fields:
            a = np.zeros(state_shape(k, c.ny, c.nx, c.n_layers, c.n_constituents, c.nsoillays), order="F")
            if lib.orc_linear_combine(self.h, slot_1, slot_2, coeff_1, coeff_2, k.encode(), _ptr(a)) != 0:
                raise KeyError(k)
            out[k] = a
        return out

    def start(self, st: dict):
        """Load the initial state into slot 0 and copy it to slot 1 (control.F90:500)."""
        self.set_state(0, st)
        self.lib().orc_copy_slot(self.h, 0, 1)

    def run(self, first: int, nsteps: int):
        self.lib().orc_run(self.h, first, nsteps)

    def step(self, old: int, new: int, counter: int):
        self.lib().orc_step(self.h, old, new, counter)

    # ---- radiation (coupler steps of mo_rterrtmgp_coupler.F90 with the gray two-stream solver) ----
    def set_rad_surface(self, surface: dict):
        self._rad = [_fortran(surface[k]) for k in ("lat_geo", "lon_geo", "sfc_albedo", "temp_sfc")]
        self.lib().orc_set_rad_surface(self.h, *[_ptr(a) for a in self._rad])

    def run_rad(self, first: int, nsteps: int, t_init: float):
        """Steps with the radiation schedule of control.F90:510-517 (same as lgame.gpu.Step.run)."""
        dt, drad = self.cfg.dt, self.ccfg.dtime_rad
        k = max(0, int(np.ceil((first * dt) / drad - 1e-12)))
        t_rad_update = t_init + k * drad
        for s in range(first, first + nsteps):
            t0 = t_init + s * dt
            lrad = 0
            if self.cfg.lrad and t0 <= t_rad_update and t0 + dt >= t_rad_update:
                lrad = 1
                t_rad_update += drad
            old, new = (0, 1) if s % 2 == 0 else (1, 0)
            self.lib().orc_step_rad(self.h, old, new, s, lrad, t0)

    def rad_columns(self, slot: int, t: float):
        c = self.cfg
        ncol = c.ny * c.nx
        x = np.zeros((ncol, 5 * c.n_layers + 4))
        y = np.zeros((ncol, c.n_layers + 2))
        self.lib().orc_rad_columns(self.h, slot, t, _ptr(x), _ptr(y))
        return x, y

    def rad_sfc(self):
        c = self.cfg
        sw, lw = np.zeros((c.ny, c.nx), order="F"), np.zeros((c.ny, c.nx), order="F")
        self.lib().orc_get_rad_sfc(self.h, _ptr(sw), _ptr(lw))
        return sw, lw

    @classmethod
    def coszenith(cls, lat: float, lon: float, t: float) -> float:
        return cls.lib().orc_coszenith(lat, lon, t)

    def integrals(self, slot: int) -> np.ndarray:
        out = np.zeros(4 + self.cfg.n_constituents)
        self.lib().orc_integrals(self.h, slot, _ptr(out))
        return out

    @property
    def n_damping_levels(self) -> int:
        return self.lib().orc_n_damping_levels(self.h)


class RefHarness:
    """The reference's compiled modules behind ref_harness.F90 (one instance per process)."""

    _lib = None
    _used = False

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_LIB)

    def __init__(self, cfg, grid: dict):
        from lgame.config import GRID_FIELDS
        if RefHarness._used:
            raise RuntimeError("the reference harness holds module state; use one per process")
        RefHarness._used = True
        lib = ctypes.CDLL(REF_LIB)
        lib.ref_set_grid.argtypes = [ctypes.c_int, ctypes.c_void_p]
        lib.ref_state_io.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        lib.ref_get_diag.argtypes = [ctypes.c_int, ctypes.c_void_p]
        lib.ref_run.argtypes = [ctypes.c_int, ctypes.c_int]
        self.lib = lib
        self.cfg = cfg
        idims = np.array([cfg.ny, cfg.nx, cfg.n_layers, cfg.n_oro_layers, cfg.n_constituents,
                          cfg.n_condensed_constituents, cfg.nsoillays], dtype=np.int32)
        icfg = np.array([cfg.theta_adv_order, int(cfg.luse_bg_state), int(cfg.llinear), int(cfg.lcorio),
                         int(cfg.lmoist), int(cfg.lklemp), int(cfg.lperiodic), int(cfg.lfreeslip),
                         int(cfg.lmom_diff_h), int(cfg.lmom_diff_v), int(cfg.ltemp_diff_h), int(cfg.ltemp_diff_v),
                         int(cfg.lmass_diff_h), int(cfg.lmass_diff_v), 0 if cfg.diff_coeff_scheme_h == "smag" else 1,
                         int(cfg.lpbl), int(cfg.lprog_soil_temp), int(cfg.lsfc_sensible_heat_flux),
                         int(cfg.lsfc_phase_trans)], dtype=np.int32)
        rcfg = np.array([cfg.dt, cfg.toa, cfg.dx, cfg.dy, cfg.klemp_damp_max, cfg.klemp_begin_rel, cfg.c_s,
                         cfg.bg_shear, grid["mean_velocity_area"], cfg.h_prandtl, cfg.karman])
        lib.ref_init(_ptr(idims), _ptr(icfg), _ptr(rcfg))
        for name in list(GRID_FIELDS) + ["z_t_const"]:
            a = _fortran(np.atleast_1d(np.asarray(grid[name], dtype=np.float64)))
            lib.ref_set_grid(REF_GRID_IDS[name], _ptr(a))
        lib.ref_finish_grid()

    def start(self, st: dict):
        for k, v in st.items():
            a = _fortran(v)
            self.lib.ref_state_io(0, REF_STATE_IDS[k], 0, _ptr(a))
        self.lib.ref_copy_slot0_to_1()

    def run(self, first: int, nsteps: int):
        self.lib.ref_run(first, nsteps)

    def set_state(self, slot: int, st: dict):
        for k, v in st.items():
            a = _fortran(v)
            self.lib.ref_state_io(slot, REF_STATE_IDS[k], 0, _ptr(a))

    def manage_p_grad(self, slot: int, lfirst: bool):
        """The reference's manage_p_grad and calc_p_grad_condensates_v on a state slot."""
        self.lib.ref_manage_p_grad.argtypes = [ctypes.c_int, ctypes.c_int]
        self.lib.ref_manage_p_grad(slot, int(lfirst))

    def get_state(self, slot: int) -> dict:
        from lgame.config import state_shape
        c = self.cfg
        out = {}
        for k in c.state_fields:
            a = np.zeros(state_shape(k, c.ny, c.nx, c.n_layers, c.n_constituents, c.nsoillays), order="F")
            self.lib.ref_state_io(slot, REF_STATE_IDS[k], 1, _ptr(a))
            out[k] = a
        return out

    def get_diag(self, name: str, shape) -> np.ndarray:
        a = np.zeros(shape, order="F")
        self.lib.ref_get_diag(REF_DIAG_IDS[name], _ptr(a))
        return a

    def linear_combine(self, slot_1: int, slot_2: int, coeff_1: float, coeff_2: float) -> dict:
        """The reference's linear_combine_two_states into the harness's third state, read back."""
        self.lib.ref_linear_combine.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        self.lib.ref_linear_combine(slot_1, slot_2, coeff_1, coeff_2)
        return self.get_state(2)


def thomas_reference(c, d, e, r):
    """The reference's thomas_algorithm (mo_column_solvers.F90:606-640) via the harness."""
    lib = ctypes.CDLL(REF_LIB)
    lib.ref_thomas.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int]
    n = len(d)
    x = np.zeros(n)
    c, d, e, r = (np.ascontiguousarray(v, dtype=np.float64) for v in (c, d, e, r))
    lib.ref_thomas(_ptr(c), _ptr(d), _ptr(e), _ptr(r), _ptr(x), n)
    return x
