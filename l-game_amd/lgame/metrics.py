"""Per-field error of a state against a reference state (the mixed form's report).

For every prognostic field: the RMS of the difference divided by the RMS of the reference field,
and the largest absolute difference divided by the largest absolute reference value.  Density
arrays are split per constituent.
"""
from __future__ import annotations

import numpy as np

CONSTITUENTS = ("snow", "rain", "cloud_ice", "cloud_water", "graupel", "gas", "vapour")


def _pieces(st: dict):
    rho = st["rho"]
    nc = rho.shape[3]
    names = CONSTITUENTS if nc == 7 else ("gas",)
    for jc in range(nc):
        yield f"rho_{names[jc]}", rho[:, :, :, jc]
    for k in ("rhotheta_v", "theta_v_pert", "exner_pert", "wind_u", "wind_v", "wind_w", "temperature_soil"):
        if k in st:
            yield k, st[k]


def field_errors(got: dict, want: dict) -> dict:
    """{field: {"rms_rel": ..., "max_rel": ...}} of `got` against `want`."""
    ref = dict(_pieces(want))
    out = {}
    for name, g in _pieces(got):
        w = ref[name]
        d = np.asarray(g, dtype=np.float64) - w
        rms_w = float(np.sqrt(np.mean(w * w)))
        max_w = float(np.abs(w).max())
        out[name] = {
            "rms_rel": float(np.sqrt(np.mean(d * d))) / rms_w if rms_w > 0 else float(np.sqrt(np.mean(d * d))),
            "max_rel": float(np.abs(d).max()) / max_w if max_w > 0 else float(np.abs(d).max()),
        }
    return out


def worst(errors: dict, key: str) -> float:
    """The largest error of kind `key`; infinite if any is NaN (max() would drop it)."""
    vals = [v[key] for v in errors.values()]
    return float("inf") if any(not np.isfinite(x) for x in vals) else max(vals)
