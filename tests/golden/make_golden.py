"""Writes the golden fixtures of this directory from the reference's own compiled step modules.

Each fixture holds one case: its run configuration, the grid metrics and initial state (inputs)
and the prognostic state after `steps` reference steps (outputs), produced by
oracle/_ref/libref.so (the reference modules built by oracle/Makefile, driven by
oracle/ref_harness.F90).  Run from the repository root where the reference tree is present:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import dataclasses
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "l-game_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

CASES = {
    "tiny_dry_o2": ({"kind": "tiny", "kw": {}}, 4),
    "tiny_dry_o3_freeslip": ({"kind": "tiny", "kw": {"order": 3, "freeslip": True}}, 5),
    "tiny_moist_o3_freeslip": ({"kind": "tiny", "kw": {"moist": True, "order": 3, "freeslip": True}, "cloud_frac": 0.3}, 5),
    "schaer_41x20": ({"kind": "schaer", "kw": {"nx": 41, "n_layers": 20}}, 12),
    # the six diffusion switches on (the reference default), Smagorinsky or TKE horizontal scheme
    "tiny_moist_o3_freeslip_diff": ({"kind": "tiny", "kw": {"moist": True, "order": 3, "freeslip": True},
                                     "cloud_frac": 0.3, "diffusion": True}, 5),
    "tiny_dry_o2_diff_tke": ({"kind": "tiny", "kw": {}, "diffusion": True,
                              "diff_kw": {"diff_coeff_scheme_h": "tke"}}, 6),
    "schaer_41x20_diff": ({"kind": "schaer", "kw": {"nx": 41, "n_layers": 20}, "diffusion": True}, 12),
    # with the PBL friction and the surface turbulence quantities as well
    "tiny_moist_o3_freeslip_diff_pbl": ({"kind": "tiny", "kw": {"moist": True, "order": 3, "freeslip": True},
                                         "cloud_frac": 0.3, "diffusion": True, "diff_kw": {"lpbl": True}}, 6),
    # and the prognostic soil with the surface sensible heat flux
    "tiny_moist_o3_freeslip_diff_pbl_soil": ({"kind": "tiny", "kw": {"moist": True, "order": 3, "freeslip": True},
                                              "cloud_frac": 0.3, "diffusion": True,
                                              "diff_kw": {"lpbl": True, "lprog_soil_temp": True,
                                                          "lsfc_sensible_heat_flux": True}}, 6),
}
# diagnostics stored with the diffusion cases (t_diag components carried between steps)
DIFF_DIAGS = ("tke", "viscosity", "mom_diff_tend_z", "heating_diss")


def one(name):
    from compare_ref import build_case
    from pyoracle import RefHarness
    from lgame.config import GRID_FIELDS, GRID_SCALARS

    spec, steps = CASES[name]
    cfg, (g, st) = build_case(spec)
    r = RefHarness(cfg, g)
    r.start(st)
    r.run(0, steps)
    out = r.get_state(steps % 2)
    arrays = {f"grid_{k}": g[k] for k in GRID_FIELDS}
    arrays.update({f"grid_{k}": np.array(g[k]) for k in GRID_SCALARS})
    arrays.update({f"in_{k}": v for k, v in st.items()})
    arrays.update({f"out_{k}": v for k, v in out.items()})
    if spec.get("diffusion"):
        ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
        shapes = {"tke": (ny, nx, nl), "viscosity": (ny, nx, nl), "mom_diff_tend_z": (ny, nx, nl + 1),
                  "heating_diss": (ny, nx, nl)}
        arrays.update({f"diag_{k}": r.get_diag(k, shapes[k]) for k in DIFF_DIAGS})
    meta = {"config": dataclasses.asdict(cfg), "steps": steps,
            "generator": "oracle/_ref/libref.so (reference modules, flang -O2, no OpenMP)"}
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), meta=np.array(json.dumps(meta)), **arrays)
    print(f"wrote {name}.npz")


if __name__ == "__main__":
    if len(sys.argv) > 1:
        one(sys.argv[1])
    else:
        for n in CASES:  # one process per case: the reference harness keeps module state
            subprocess.run([sys.executable, __file__, n], check=True)
