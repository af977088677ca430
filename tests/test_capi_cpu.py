"""C-ABI checks that need no GPU: the library loads, exports what include/lgame_gpu.h declares,
and rejects invalid or unsupported configurations before touching the device."""
import ctypes
import os
import re

import pytest

from lgame.cases import tiny_config
from lgame.config import GRID_FIELDS
from lgame.gpu import EXPORTED, LIB_PATH, GridView, load_library

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lgame_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|void|long long|const char \*)\s*(lgame_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_bound_functions():
    assert declared_functions() == sorted(EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def _init(mutate_dims=None, mutate_cfg=None):
    lib = load_library()
    cfg = tiny_config()
    d, c = cfg.to_abi()
    if mutate_dims:
        mutate_dims(d)
    if mutate_cfg:
        mutate_cfg(c)
    buf = ctypes.create_string_buffer(8 * 8)  # never read: validation fails first
    gv = GridView(*([ctypes.addressof(buf)] * len(GRID_FIELDS)))
    h = ctypes.c_void_p()
    rc = lib.lgame_init(ctypes.byref(d), ctypes.byref(c), ctypes.byref(gv), ctypes.byref(h))
    return rc, lib.lgame_last_error(None).decode(), h


@pytest.mark.parametrize("mutate_dims,mutate_cfg,msg", [
    (lambda d: setattr(d, "n_oro_layers", d.n_layers), None, "n_oro_layers"),
    (lambda d: setattr(d, "ny", 2), None, ">= 3"),
    (None, lambda c: setattr(c, "theta_adv_order", 4), "theta_adv_order"),
    (None, lambda c: setattr(c, "lperiodic", 0), "lperiodic"),
    (None, lambda c: setattr(c, "ltemp_diff_h", 1), "without horizontal momentum diffusion"),
    (None, lambda c: setattr(c, "diff_coeff_scheme_h", 2), "diff_coeff_scheme_h"),
    (None, lambda c: setattr(c, "lmom_diff_h", 1), "mean_velocity_area"),
    (None, lambda c: setattr(c, "lmoist", 1), "constituent"),
    (None, lambda c: setattr(c, "dtime", 0.0), "dtime"),
    (None, lambda c: setattr(c, "lpbl", 1), "lpbl needs lmom_diff_h"),
    (None, lambda c: setattr(c, "lsfc_sensible_heat_flux", 1), "without the prognostic soil"),
    (lambda d: setattr(d, "nsoillays", 1), lambda c: setattr(c, "lprog_soil_temp", 1), "nsoillays"),
    (lambda d: (setattr(d, "ny", 6000), setattr(d, "nx", 6000), setattr(d, "n_layers", 30), setattr(d, "n_oro_layers", 20)),
     None, "exceeds 4 GiB"),
    (lambda d: setattr(d, "nsoillays", 17), lambda c: setattr(c, "lprog_soil_temp", 1), "nsoillays"),
])
def test_init_rejects_invalid_configuration(mutate_dims, mutate_cfg, msg):
    rc, err, h = _init(mutate_dims, mutate_cfg)
    assert rc == 1  # LGAME_EINVAL
    assert msg in err
    assert not h.value


def test_abi_version():
    assert load_library().lgame_abi_version() == 4


def test_grid_setup_rejects_invalid_parameters():
    """lgame_grid_setup validates before touching the device."""
    from lgame.gpu import GridOut, GridParams
    lib = load_library()
    d, _ = tiny_config().to_abi()
    buf = ctypes.create_string_buffer(8 * 8)
    out = GridOut(*([ctypes.addressof(buf)] * 44))
    for p, msg in ((GridParams(2000.0, 2000.0, 12000.0, 1.3, 0.0, 0.0, 6, 1, 0, 0, 0, None), "orography_id"),
                   (GridParams(2000.0, 2000.0, 12000.0, 1.3, 0.0, 0.0, 1, 1, 0, 0, 0, None, None), "ETOPO"),
                   (GridParams(0.0, 2000.0, 12000.0, 1.3, 0.0, 0.0, 0, 1, 0, 0, 0, None), "positive"),
                   (GridParams(2000.0, 2000.0, 12000.0, 1.3, 0.0, 0.0, 5, 1, 0, 0, 2, None), "hills")):
        assert lib.lgame_grid_setup(ctypes.byref(d), ctypes.byref(p), ctypes.byref(out)) == 1
        assert msg in lib.lgame_last_error(None).decode()
    out = GridOut()
    p = GridParams(2000.0, 2000.0, 12000.0, 1.3, 0.0, 0.0, 0, 1, 0, 0, 0, None)
    assert lib.lgame_grid_setup(ctypes.byref(d), ctypes.byref(p), ctypes.byref(out)) == 1
    assert "null" in lib.lgame_last_error(None).decode()


def test_ensemble_rejects_a_member_listed_twice():
    """A handle listed twice would be stepped twice per ensemble step on one stream; the
    ensemble refuses it before touching any member (the pointers are compared, not used)."""
    lib = load_library()
    fake = ctypes.create_string_buffer(64)
    members = (ctypes.c_void_p * 3)(ctypes.addressof(fake), ctypes.addressof(fake) + 8, ctypes.addressof(fake))
    e = ctypes.c_void_p()
    assert lib.lgame_ensemble_create(members, 3, ctypes.byref(e)) == 1  # LGAME_EINVAL
    assert not e.value
