"""Loading of the golden fixtures in tests/golden (written by tests/golden/make_golden.py)."""
import json
import os

import numpy as np

from lgame.config import GRID_FIELDS, GRID_SCALARS, RunConfig

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def load(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    c = dict(meta["config"])
    c["hills"] = tuple(tuple(h) for h in c["hills"])
    cfg = RunConfig(**c)
    grid = {k: np.asfortranarray(z[f"grid_{k}"]) for k in GRID_FIELDS}
    grid.update({k: float(z[f"grid_{k}"]) for k in GRID_SCALARS})
    st_in = {k: np.asfortranarray(z[f"in_{k}"]) for k in cfg.state_fields}
    st_out = {k: np.asfortranarray(z[f"out_{k}"]) for k in cfg.state_fields}
    return cfg, grid, st_in, st_out, meta["steps"]


def load_diags(name):
    """The t_diag fields stored with a fixture (diffusion cases), {} otherwise."""
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    return {k[5:]: np.asfortranarray(z[k]) for k in z.files if k.startswith("diag_")}


def max_rel(a, b):
    """max |a-b| / max |b| over the fields of two states; infinite if a field of either holds a
    NaN or Inf (a comparison with NaN would otherwise drop out of the maximum)."""
    worst = 0.0
    for k in b:
        if not (np.all(np.isfinite(a[k])) and np.all(np.isfinite(b[k]))):
            return float("inf")
        scale = max(float(np.abs(b[k]).max()), 1e-300)
        worst = max(worst, float(np.abs(a[k] - b[k]).max()) / scale)
    return worst
