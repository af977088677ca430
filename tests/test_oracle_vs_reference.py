"""The C restatement (oracle/lgame_oracle.c) against the reference's own compiled modules.

Each case runs both on identical inputs for several steps in a fresh process (the reference
keeps Fortran module state) and requires bit-identical prognostic fields.  Skipped where the
reference build (oracle/_ref/libref.so) is absent.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libref.so")

CASES = [
    {"kind": "tiny", "kw": {}, "steps": 4},
    {"kind": "tiny", "kw": {"order": 3}, "steps": 4},
    {"kind": "tiny", "kw": {"order": 3, "freeslip": True}, "steps": 5},
    {"kind": "tiny", "kw": {"corio": False, "plane": True}, "steps": 4},
    {"kind": "tiny", "kw": {"moist": True}, "steps": 4},
    {"kind": "tiny", "kw": {"moist": True, "order": 3, "freeslip": True}, "steps": 5, "cloud_frac": 0.3},
    {"kind": "tiny", "kw": {"ny": 7, "nx": 5, "n_layers": 8, "llinear": True}, "steps": 3},
    {"kind": "schaer", "kw": {"nx": 41, "n_layers": 20}, "steps": 20},
    # diffusion and turbulence (the reference default switches), with the t_diag fields they carry
    {"kind": "tiny", "kw": {}, "steps": 4, "diffusion": True},
    {"kind": "tiny", "kw": {"moist": True, "order": 3, "freeslip": True}, "steps": 5, "cloud_frac": 0.3,
     "diffusion": True},
    {"kind": "tiny", "kw": {"order": 3}, "steps": 8, "diffusion": True, "diff_kw": {"diff_coeff_scheme_h": "tke"}},
    {"kind": "tiny", "kw": {"moist": True}, "steps": 5, "cloud_frac": 0.3, "diffusion": True,
     "diff_kw": {"lmom_diff_v": False, "ltemp_diff_v": False, "lmass_diff_h": False}},
    {"kind": "tiny", "kw": {"plane": True, "corio": False}, "steps": 4, "diffusion": True,
     "diff_kw": {"lmom_diff_h": False, "ltemp_diff_h": False, "lmass_diff_h": False}},
    {"kind": "schaer", "kw": {"nx": 41, "n_layers": 20}, "steps": 12, "diffusion": True},
    # the PBL friction with the surface turbulence quantities, with horizontal momentum diffusion
    # only (the reference requires it with lpbl) and with all diffusion switches
    {"kind": "tiny", "kw": {}, "steps": 5, "diff_kw": {"lpbl": True, "lmom_diff_h": True}},
    {"kind": "tiny", "kw": {"moist": True, "order": 3, "freeslip": True}, "steps": 6, "cloud_frac": 0.3,
     "diffusion": True, "diff_kw": {"lpbl": True}},
    {"kind": "dry3d", "kw": {"n": 17, "n_layers": 12}, "steps": 8, "diffusion": True, "diff_kw": {"lpbl": True}},
    # the prognostic soil: heat conduction alone, with the surface sensible heat flux, and with the
    # PBL and all diffusion; the surface phase-transition switch (surface turbulence quantities)
    {"kind": "tiny", "kw": {}, "steps": 5, "diff_kw": {"lprog_soil_temp": True}},
    {"kind": "tiny", "kw": {"moist": True}, "steps": 5, "cloud_frac": 0.3,
     "diff_kw": {"lprog_soil_temp": True, "lsfc_sensible_heat_flux": True}},
    {"kind": "tiny", "kw": {"moist": True, "order": 3, "freeslip": True}, "steps": 6, "cloud_frac": 0.3,
     "diffusion": True, "diff_kw": {"lpbl": True, "lprog_soil_temp": True, "lsfc_sensible_heat_flux": True,
                                    "lsfc_phase_trans": True}},
    {"kind": "dry3d", "kw": {"n": 17, "n_layers": 12, "nsoillays": 7}, "steps": 8, "diffusion": True,
     "diff_kw": {"lpbl": True, "lprog_soil_temp": True, "lsfc_sensible_heat_flux": True}},
    {"kind": "tiny", "kw": {}, "steps": 4, "diff_kw": {"lsfc_phase_trans": True}},
    # the smallest domains the step accepts
    {"kind": "tiny", "kw": {"ny": 3, "nx": 3, "n_layers": 3}, "steps": 4},
    {"kind": "tiny", "kw": {"ny": 3, "nx": 4, "n_layers": 4, "moist": True, "order": 3}, "steps": 4, "cloud_frac": 0.2,
     "diffusion": True},
    # the linear combination of the two time levels (output-time interpolation)
    {"kind": "tiny", "kw": {"moist": True}, "steps": 3, "cloud_frac": 0.3, "combine": [0.3125, 0.6875],
     "diff_kw": {"lprog_soil_temp": True}},
]


@pytest.mark.skipif(not os.path.exists(REF), reason="reference build oracle/_ref/libref.so not present")
@pytest.mark.parametrize("case", CASES, ids=[json.dumps(c, sort_keys=True) for c in CASES])
def test_oracle_matches_reference_bitwise(case):
    out = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "compare_ref.py"), json.dumps(case)],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    bad = {k: v for k, v in res.items() if not v["bitwise"]}
    assert not bad, bad
    # the case must actually evolve
    assert res["wind_w"]["moved"] > 0.0 and res["rhotheta_v"]["moved"] > 0.0
