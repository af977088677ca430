"""The five workloads at their own sizes on the GPU (through the C-ABI).

* Schaer mountain wave, 201 x 3 x 40 plane, dx 500 m: 1000 steps against the oracle within 1e-12
  (bitwise expected); the whole 5-hour run on the GPU alone, with the dry mass conserved and every
  field finite; dt 4 s (the value the workload names) turns the state non-finite, as it does in
  the reference scheme.
* Dry 3-D case, 129 x 129 x 50 rotated lat-lon: a few steps against the oracle within 1e-12; the
  128 x 128 x 50 mixed form against the FP64 form over 1000 steps, per-field RMS and max error
  reported; dt 10 s (the value the workload names) turns the state non-finite.
* Surrogate network 304 -> 3 x 512 -> 62 at batch 8192: the device gradient of the loss against
  the CPU oracle of the network.
* 48 moist members of uneven size in the mixed form with the surrogate: one launch per kernel phase
  over all members, and one graph launch per step, against serial launches, bitwise, over steps
  that include a radiation update.
"""
import dataclasses

import numpy as np
import pytest
import xxhash

from lgame.cases import dry3d_config, ensemble_members, make_case, radiation_surface, schaer_config
from lgame.config import PREC_FP64, PREC_MIXED, RAD_SURROGATE
from lgame.gpu import Ensemble, Mlp, Step, grid_setup_device
from lgame.metrics import field_errors, worst
from golden_util import max_rel
from pyoracle import CpuOracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _oracle(cfg, grid, st, n):
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.run(0, n)
    return o.get_state(n % 2)


def _gpu(cfg, grid, st, n):
    s = Step(cfg, grid)
    s.start(st)
    s.run(0, n)
    out = s.get_state(n % 2)
    s.close()
    return out


# ---- Schaer mountain wave (201 x 3 x 40, dx 500 m, toa 21 km, dt 0.805 s) ----------------------

@pytest.mark.timeout(300)
def test_schaer_201x3x40_1000_steps_match_oracle():
    cfg = schaer_config(nx=201, n_layers=40)
    grid, st = make_case(cfg)
    n = 1000
    print(f"schaer 201x3x40, dt {cfg.dt} s, {n} steps", flush=True)
    got = _gpu(cfg, grid, st, n)
    want = _oracle(cfg, grid, st, n)
    rel = max_rel(got, want)
    bitwise = all(np.array_equal(got[k], want[k]) for k in want)
    print(f"max relative deviation {rel:.3e}, bitwise={bitwise}")
    assert rel <= TOL
    assert np.abs(got["wind_w"]).max() > 1e-3  # the mountain wave has developed


@pytest.mark.timeout(300)
def test_schaer_201x3x40_five_hours_conserve_mass():
    """The whole run of the workload: 5 simulated hours plus the reference loop's 300 s
    (control.F90:507), 22,733 steps at dt 0.805 s, on the GPU alone.  Dry mass is conserved to
    round-off (the flux-form divergence telescopes on the periodic plane; the reference run drifts
    5.9e-14 over 1,119 steps), every field stays finite and the flow is not at rest."""
    cfg = schaer_config(nx=201, n_layers=40)
    grid, st = make_case(cfg)
    n = int(np.ceil((5 * 3600.0 + 300.0) / cfg.dt))
    print(f"schaer 201x3x40, dt {cfg.dt} s, {n} steps", flush=True)
    s = Step(cfg, grid)
    s.start(st)
    m0 = s.integrals(0)
    for k0 in range(0, n, 2000):
        s.run(k0, min(2000, n - k0))
        print(f"  {min(k0 + 2000, n)} steps", flush=True)
    m1 = s.integrals(n % 2)
    bad = s.count_nonfinite(n % 2)
    out = s.get_state(n % 2)
    s.close()
    drift = abs(m1[4] - m0[4]) / m0[4]
    print(f"dry mass drift {drift:.3e} after {n} steps, max |w| {np.abs(out['wind_w']).max():.3f} m/s")
    assert bad == 0
    assert drift < 1e-12
    assert np.abs(out["wind_w"]).max() > 1e-3


def test_schaer_dt_4s_is_unstable():
    """dt 4 s, the value the workload names, exceeds the horizontal sound-wave Courant limit of
    the scheme at dx 500 m (c dt/dx = 2.8): the reference scheme (here its oracle) turns the state
    non-finite within 10 steps; the device step does the same, and the step's own check reports
    it."""
    cfg = schaer_config(nx=201, n_layers=40, dtime=4.0)
    grid, st = make_case(cfg)
    s = Step(cfg, grid)
    s.start(st)
    s.run(0, 150)
    bad = s.count_nonfinite(0)
    s.close()
    print(f"non-finite values after 150 steps at dt 4 s: {bad}")
    assert bad > 0


# ---- dry 3-D case (129/128 x 129/128 x 50, dx 2 km) ---------------------------------------------

@pytest.mark.timeout(300)
def test_dry3d_129x129x50_matches_oracle():
    cfg = dry3d_config(n=129, n_layers=50)
    grid, st = make_case(cfg)
    n = 3
    print(f"dry3d 129x129x50, {n} steps", flush=True)
    got = _gpu(cfg, grid, st, n)
    want = _oracle(cfg, grid, st, n)
    rel = max_rel(got, want)
    bitwise = all(np.array_equal(got[k], want[k]) for k in want)
    print(f"max relative deviation {rel:.3e}, bitwise={bitwise}")
    assert rel <= TOL


@pytest.mark.timeout(300)
def test_dry3d_128x128x50_mixed_vs_fp64_1000_steps():
    """The mixed form (FP32 stencils and explicit tendencies, FP64 column solves, density updates
    and budget sums) against the plain FP64 form after 1000 steps of 3.22 s (54 simulated minutes):
    per-field RMS and max error relative to the FP64 fields."""
    base = dry3d_config(n=128, n_layers=50)
    grid, st = make_case(base)
    n = 1000
    runs = {}
    for form, prec in (("fp64", PREC_FP64), ("mixed", PREC_MIXED)):
        print(f"dry3d 128x128x50 {form}, {n} steps", flush=True)
        s = Step(dataclasses.replace(base, precision=prec), grid)
        s.start(st)
        s.run(0, n)
        assert s.count_nonfinite(n % 2) == 0
        runs[form] = s.get_state(n % 2)
        s.close()
    err = field_errors(runs["mixed"], runs["fp64"])
    for k, v in err.items():
        print(f"  {k:14s} rms {v['rms_rel']:.3e}  max {v['max_rel']:.3e}")
    print(f"worst rms {worst(err, 'rms_rel'):.3e}, worst max {worst(err, 'max_rel'):.3e}")
    assert worst(err, "rms_rel") < 1e-3
    assert worst(err, "max_rel") < 1e-2


def test_dry3d_dt_10s_is_unstable():
    """dt 10 s at dx 2 km, the value the workload names, is past the scheme's horizontal
    sound-wave limit (Courant number about 1.7): the oracle turns non-finite within 10 steps, and
    so does the device step."""
    cfg = dry3d_config(n=128, n_layers=50, dtime=10.0)
    grid, st = make_case(cfg)
    s = Step(cfg, grid)
    s.start(st)
    s.run(0, 40)
    bad = s.count_nonfinite(0)
    s.close()
    print(f"non-finite values after 40 steps at dt 10 s: {bad}")
    assert bad > 0


# ---- the surrogate network at its own size ----------------------------------------------------

@pytest.mark.timeout(300)
def test_mlp_304_3x512_62_gradient_at_batch_8192():
    from mlp_oracle import MlpOracle

    n_in, hidden, nh, n_out, batch = 304, 512, 3, 62, 8192
    m = Mlp(n_in, hidden, nh, n_out, seed=11)
    rng = np.random.default_rng(11)
    xm, xs = rng.standard_normal(n_in).astype(np.float32), (0.5 + rng.random(n_in)).astype(np.float32)
    ym, ys = rng.standard_normal(n_out).astype(np.float32), (0.5 + rng.random(n_out)).astype(np.float32)
    m.set_norm(xm, xs, ym, ys)
    o = MlpOracle(n_in, hidden, nh, n_out, m.get_params(), xm, xs, ym, ys, bf16=True)
    x = (xm + xs * rng.standard_normal((batch, n_in))).astype(np.float32)
    y = (ym + ys * rng.standard_normal((batch, n_out))).astype(np.float32)
    gg, lg = m.gradient(x, y)
    lo, go = o.loss_and_grad(x, y)
    rel = np.linalg.norm(gg - go) / np.linalg.norm(go)
    cos = float(np.dot(gg, go) / (np.linalg.norm(gg) * np.linalg.norm(go)))
    print(f"304->3x512->62, batch {batch}: gradient relative error {rel:.3e}, cosine {cos:.6f}, "
          f"loss {lg:.6f} vs {lo:.6f}")
    assert abs(lg - lo) <= 1e-2 * lo
    assert rel < 3e-2 and cos > 0.999
    sizes = [n_in] + [hidden] * nh + [n_out]
    off = 0
    for i, oo in zip(sizes[:-1], sizes[1:]):
        for cnt in (i * oo, oo):
            a, b = gg[off:off + cnt], go[off:off + cnt]
            assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b) + 1e-7
            off += cnt
    m.close()


# ---- 48 uneven members, mixed form with the surrogate ----------------------------------------

def _digest(st):
    h = xxhash.xxh3_128()
    for k in sorted(st):
        h.update(np.ascontiguousarray(st[k]).data)
    return h.hexdigest()


@pytest.mark.timeout(900)
def test_ensemble_48_uneven_members_graph_equals_serial():
    """48 moist members (nx, ny ~ U[64,384], 40-70 layers), mixed form with the radiation
    surrogate and its 1 % audit (members of other than 60 layers resampled to the network's
    levels): one launch per kernel phase over all members (the radiation update as one gather,
    one inference over every member's columns and one audit), and one graph launch per ensemble
    step, against serial launches, over 4 steps whose second updates the radiation; every
    member's state is bitwise the same."""
    from lgame import surrogate

    cfgs = [dataclasses.replace(c, precision=PREC_MIXED, lrad=True, dtime_rad=900.0, rad_mode=RAD_SURROGATE)
            for c in ensemble_members(48)]
    mlp = surrogate.load(surrogate.DEFAULT_WEIGHTS)
    n, t0 = 4, 1576843200.0 + 3 * 3600.0
    digests = {}
    for mode in (Ensemble.BATCHED, Ensemble.GRAPH, Ensemble.SERIAL):
        steps = []
        for i, c in enumerate(cfgs):
            g = grid_setup_device(c)
            g, st = make_case(c, grid=g)
            s = Step(c, g)
            s.set_rad_surface(radiation_surface(c, g, st))
            s.attach_mlp(mlp, tol_k_per_day=1.0, audit_fraction=0.01, seed=7)
            s.start(st)
            steps.append(s)
        print(f"mode {mode}: {len(steps)} members, {sum(c.gridpoints for c in cfgs) / 1e6:.1f} M points", flush=True)
        e = Ensemble(steps)
        for k in range(n):
            e.step(k, lrad_update=(k == 1), t_0=t0 + k * cfgs[0].dt, mode=mode)
        e.sync()
        digests[mode] = [_digest(s.get_state(n % 2)) for s in steps]
        for s in steps:
            assert s.count_nonfinite(n % 2) == 0
            assert s.rad_stats()["calls"] == 1
        e.close()
        for s in steps:
            s.close()
    for mode in (Ensemble.BATCHED, Ensemble.GRAPH):
        same = sum(a == b for a, b in zip(digests[mode], digests[Ensemble.SERIAL]))
        print(f"members bitwise equal between mode {mode} and serial launches: {same}/48")
        assert same == 48
