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

from .config import GRID_FIELDS, STATE_FIELDS, RunConfig, state_shape

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "liblgame_gpu.so")

EXPORTED = ("lgame_abi_version", "lgame_init", "lgame_set_state", "lgame_get_state", "lgame_pchevi",
            "lgame_pchevi_host", "lgame_sync", "lgame_integrals", "lgame_check_finite", "lgame_get_diag",
            "lgame_finalize", "lgame_last_error", "lgame_profile", "lgame_profile_read", "lgame_profile_reset")

PROF_SOLVE_VER = 0
PROF_SOLVE_GEN = 1
PROF_STEP = 2

LGAME_DIAG_TEMPERATURE = 0


class LgameError(RuntimeError):
    pass


class GridView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in GRID_FIELDS]


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
    lib.lgame_set_state.argtypes = [vp, ci, vp]
    lib.lgame_get_state.argtypes = [vp, ci, vp]
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
        gv = GridView(*[keep[n].ctypes.data for n in GRID_FIELDS])
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
        arrs = {k: _f64(st[k]) for k in STATE_FIELDS}
        return arrs, StateView(*[arrs[k].ctypes.data for k in STATE_FIELDS], None)

    def set_state(self, slot: int, st: dict):
        arrs, v = self._view(st)
        self._check(self.lib.lgame_set_state(self.h, slot, ctypes.byref(v)), "lgame_set_state")

    def get_state(self, slot: int) -> dict:
        c = self.cfg
        out = {k: np.zeros(state_shape(k, c.ny, c.nx, c.n_layers, c.n_constituents), order="F") for k in STATE_FIELDS}
        v = StateView(*[out[k].ctypes.data for k in STATE_FIELDS], None)
        self._check(self.lib.lgame_get_state(self.h, slot, ctypes.byref(v)), "lgame_get_state")
        return out

    def start(self, st: dict):
        """Initial state into both slots (state_2 = state_1, control.F90:500)."""
        self.set_state(0, st)
        self.set_state(1, st)

    def pchevi(self, slot_old: int, slot_new: int, total_step_counter: int, lrad_update: bool = False, t_0: float = 0.0):
        self._check(self.lib.lgame_pchevi(self.h, slot_old, slot_new, total_step_counter, int(lrad_update), t_0), "lgame_pchevi")

    def run(self, first: int, nsteps: int):
        for s in range(first, first + nsteps):
            if s % 2 == 0:
                self.pchevi(0, 1, s)
            else:
                self.pchevi(1, 0, s)

    def sync(self):
        self._check(self.lib.lgame_sync(self.h), "lgame_sync")

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

    def profile(self, enable: bool = True):
        self._check(self.lib.lgame_profile(self.h, int(enable)), "lgame_profile")

    def profile_read(self, group: int):
        """(total milliseconds, launch count) recorded for a kernel group since the last reset."""
        ms, n = ctypes.c_double(0.0), ctypes.c_longlong(0)
        self._check(self.lib.lgame_profile_read(self.h, group, ctypes.byref(ms), ctypes.byref(n)), "lgame_profile_read")
        return ms.value, n.value

    def profile_reset(self):
        self._check(self.lib.lgame_profile_reset(self.h), "lgame_profile_reset")

    def temperature(self) -> np.ndarray:
        c = self.cfg
        out = np.zeros((c.ny, c.nx, c.n_layers), order="F")
        self._check(self.lib.lgame_get_diag(self.h, LGAME_DIAG_TEMPERATURE, out.ctypes.data), "lgame_get_diag")
        return out
