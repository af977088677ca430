"""Run the C oracle and the reference harness on one case and print the comparison as JSON.

TEST INFRASTRUCTURE ONLY.  The reference harness keeps Fortran module state, so each comparison
runs in its own process:  python oracle/compare_ref.py '<json case>'
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "l-game_amd"))
sys.path.insert(0, HERE)

from lgame.cases import make_case, schaer_config, tiny_config, dry3d_config  # noqa: E402
from pyoracle import CpuOracle, RefHarness  # noqa: E402


def build_case(spec):
    kind = spec.get("kind", "tiny")
    kw = spec.get("kw", {})
    if kind == "tiny":
        cfg = tiny_config(**kw)
    elif kind == "schaer":
        cfg = schaer_config(**kw)
    else:
        cfg = dry3d_config(**kw)
    if spec.get("diffusion"):
        cfg.set_diffusion(True)
    for k, v in spec.get("diff_kw", {}).items():
        setattr(cfg, k, v)
    return cfg, make_case(cfg, cloud_frac=spec.get("cloud_frac", 0.0))


def main():
    spec = json.loads(sys.argv[1])
    cfg, (g, st) = build_case(spec)
    n = spec.get("steps", 3)
    o = CpuOracle(cfg, g)
    o.start(st)
    o.run(0, n)
    r = RefHarness(cfg, g)
    r.start(st)
    r.run(0, n)
    slot = n % 2
    a, b = o.get_state(slot), r.get_state(slot)
    res = {}
    for k in a:
        d = np.abs(a[k] - b[k]).max()
        res[k] = {"bitwise": bool(np.array_equal(a[k], b[k])), "maxabs": float(d),
                  "moved": float(np.abs(a[k] - st[k]).max())}
    if spec.get("diffusion") or cfg.surface:
        ny, nx, nl, nc = cfg.ny, cfg.nx, cfg.n_layers, cfg.n_constituents
        for k, shp in (("tke", (ny, nx, nl)), ("viscosity", (ny, nx, nl)), ("n_squared", (ny, nx, nl)),
                       ("mom_diff_tend_x", (ny, nx + 1, nl)), ("mom_diff_tend_y", (ny + 1, nx, nl)),
                       ("mom_diff_tend_z", (ny, nx, nl + 1)), ("temp_diff_heating", (ny, nx, nl)),
                       ("heating_diss", (ny, nx, nl)), ("mass_diff_tendency", (ny, nx, nl, nc)),
                       ("roughness_length", (ny, nx)), ("roughness_velocity", (ny, nx)),
                       ("monin_obukhov_length", (ny, nx)), ("scalar_flux_resistance", (ny, nx)),
                       ("power_flux_density_sens_soil", (ny, nx)), ("power_flux_density_sens_sea", (ny, nx))):
            x, y = o.get_diag(k, shp), r.get_diag(k, shp)
            res["diag:" + k] = {"bitwise": bool(np.array_equal(x, y)), "maxabs": float(np.abs(x - y).max()),
                                "moved": float(np.abs(x).max())}
    if "combine" in spec:  # the output-time interpolation of control.F90:539
        c1, c2 = spec["combine"]
        x, y = o.linear_combine(0, 1, c1, c2), r.linear_combine(0, 1, c1, c2)
        for k in x:
            res["combine:" + k] = {"bitwise": bool(np.array_equal(x[k], y[k])), "maxabs": float(np.abs(x[k] - y[k]).max()),
                                   "moved": float(np.abs(x[k]).max())}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
