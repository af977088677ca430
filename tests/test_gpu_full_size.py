"""Properties of the step at the full size of the moist workload (512 x 512 x 60, 7 constituents),
where the oracle would take hours: budgets, reproducibility and the mixed form's error.

The domain is the periodic plane with hills that vanish towards its edges, where the flux-form
divergences telescope and the dry-air and vapour masses are conserved to round-off (on the
rotated lat-lon domain the reference's extrapolated edge face areas break that; the condensates
fall out through the surface).  The grid comes from the device generator."""
import dataclasses

import numpy as np
import pytest

from lgame.cases import make_case, moist3d_config
from lgame.config import PREC_MIXED
from lgame.gpu import Step, grid_setup_device
from lgame.metrics import field_errors, worst

pytestmark = pytest.mark.gpu
STEPS = 6
HILLS = ((0.0, 0.0, 1200.0, 25e3), (150e3, -100e3, 800.0, 20e3), (-200e3, 120e3, 600.0, 30e3))


@pytest.fixture(scope="module")
def case():
    cfg = dataclasses.replace(moist3d_config(n=512, n_layers=60, dtime=3.22), lplane=True, lcorio=False, hills=HILLS)
    print("full size: moist plane 512x512x60, device grid", flush=True)
    grid = grid_setup_device(cfg)
    grid, st = make_case(cfg, cloud_frac=0.1, grid=grid)
    return cfg, grid, st


def _run(cfg, grid, st, n, precision=None):
    s = Step(cfg if precision is None else dataclasses.replace(cfg, precision=precision), grid)
    s.start(st)
    m0 = s.integrals(0)
    s.run(0, n)
    out = s.get_state(n % 2)
    m1 = s.integrals(n % 2)
    bad = s.count_nonfinite(n % 2)
    s.close()
    return out, m0, m1, bad


def test_full_size_conserves_dry_air_and_vapour_and_is_reproducible(case):
    cfg, grid, st = case
    a, m0, m1, bad = _run(cfg, grid, st, STEPS)
    assert bad == 0
    g = 4 + cfg.n_condensed_constituents  # mass slots: LGAME_INT_MASS0 + constituent index
    for slot, name in ((g, "dry air"), (g + 1, "vapour")):
        rel = abs(m1[slot] - m0[slot]) / m0[slot]
        print(f"  {name} mass: relative change {rel:.3e} after {STEPS} steps")
        assert rel < 1e-12, name
    assert np.abs(a["wind_w"]).max() > 1e-4  # the flow over the hills is not trivial
    b, _, _, _ = _run(cfg, grid, st, STEPS)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_full_size_mixed_form_error(case):
    cfg, grid, st = case
    want, _, _, _ = _run(cfg, grid, st, STEPS)
    got, _, _, bad = _run(cfg, grid, st, STEPS, PREC_MIXED)
    assert bad == 0
    err = field_errors(got, want)
    print(f"  mixed form after {STEPS} steps: worst rms {worst(err, 'rms_rel'):.3e}, "
          f"worst max {worst(err, 'max_rel'):.3e}")
    assert worst(err, "rms_rel") < 1e-4
