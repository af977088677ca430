"""Grid generation on the device (lgame_grid_setup) against the host restatement of grid_setup and
bg_setup (l-game_amd/lgame/grid.py, mo_grid_generator.F90:29-1457).

Bar: every field within 1e-11 of the host's, relative to the field's largest magnitude (the
device's sin, cos, exp, sinh, pow, asin and atan2 may differ from the host's in the last place,
and differences of nearly equal heights or pressures amplify that), and a step on the device
grid within 1e-10 of the same step on the host grid."""
import dataclasses

import numpy as np
import pytest

from golden_util import max_rel
from lgame.cases import dry3d_config, make_case, schaer_config, tiny_config
from lgame.grid import grid_setup
from lgame.gpu import LgameError, Step, grid_setup_device

pytestmark = pytest.mark.gpu
TOL = 1e-11

CASES = {
    "tiny_hills_sleve": tiny_config(),
    "tiny_plane_nocorio": tiny_config(corio=False, plane=True),
    "tiny_no_sleve_freeslip": dataclasses.replace(tiny_config(ny=6, nx=9, n_layers=8, freeslip=True), lsleve=False),
    "schaer_61x24": schaer_config(nx=61, n_layers=24),
    "dry3d_33x33x20": dry3d_config(n=33, n_layers=20),
    "dry3d_40x66x18_soil7": dataclasses.replace(dry3d_config(n=40, n_layers=18, nsoillays=7), nx=66),
    "oro3_plane": dataclasses.replace(schaer_config(nx=41, n_layers=20), orography_id=3, toa=30000.0),
    "oro4_rotated": dataclasses.replace(tiny_config(ny=7, nx=11, n_layers=10), orography_id=4, hills=()),
}


def _compare(a, b):
    worst = {}
    for k, want in b.items():
        got = np.asarray(a[k], dtype=np.float64)
        want = np.asarray(want, dtype=np.float64)
        assert got.shape == want.shape, k
        scale = float(np.abs(want).max())
        err = float(np.abs(got - want).max())
        worst[k] = err / scale if scale > 0.0 else err
    return worst


@pytest.mark.parametrize("name", list(CASES))
def test_device_grid_matches_host(name):
    cfg = CASES[name]
    print(f"grid {name}: ny={cfg.ny} nx={cfg.nx} nl={cfg.n_layers} orography {cfg.orography_id}", flush=True)
    a = grid_setup_device(cfg)
    b = grid_setup(cfg)
    worst = _compare(a, b)
    bad = {k: v for k, v in worst.items() if not v <= TOL}
    print(f"  largest relative deviation {max(worst.values()):.3e} ({max(worst, key=worst.get)})")
    assert not bad, bad


def test_step_on_device_grid_matches_host_grid():
    cfg = tiny_config(moist=True, order=3)
    print("step on the device grid, tiny moist, 5 steps", flush=True)
    grid_h, st = make_case(cfg, cloud_frac=0.2)
    grid_d = grid_setup_device(cfg)
    out = {}
    for tag, g in (("host", grid_h), ("device", grid_d)):
        s = Step(cfg, g)
        s.start(st)
        s.run(0, 5)
        out[tag] = s.get_state(1)
        s.close()
    rel = max_rel(out["device"], out["host"])
    print(f"  max relative deviation {rel:.3e}")
    assert rel <= 1e-10


def test_device_grid_rejects_orography_above_the_flat_levels():
    cfg = dataclasses.replace(tiny_config(), hills=((0.0, 0.0, 9000.0, 4000.0),))
    with pytest.raises(LgameError, match="vertical grid"):
        grid_setup_device(cfg)
