"""Launch tuning and ensembles: the autotuner's block shapes, the vertical segments of the
column-marching scalar kernel and the graph-captured ensemble step must not change any result
(every member stays bitwise equal to its own oracle run)."""
import dataclasses

import numpy as np
import pytest

from lgame.cases import dry3d_config, make_case, tiny_config
from lgame.gpu import TILE_KERNELS, Ensemble, LgameError, Step
from pyoracle import CpuOracle

pytestmark = pytest.mark.gpu


def _oracle(cfg, g, st, n):
    o = CpuOracle(cfg, g)
    o.start(st)
    o.run(0, n)
    return o.get_state(n % 2)


def test_autotune_keeps_results():
    cfg = tiny_config(ny=9, nx=11, n_layers=7, moist=True, order=3, freeslip=True)
    g, st = make_case(cfg, cloud_frac=0.2)
    s = Step(cfg, g)
    s.start(st)
    rows = s.autotune(2)
    print(f"tiles {rows}")
    assert set(rows) == set(TILE_KERNELS) and all(r in (1, 2, 4) for r in rows.values())
    s.run(0, 5)
    got = s.get_state(1)
    s.close()
    want = _oracle(cfg, g, st, 5)
    assert all(np.array_equal(got[k], want[k]) for k in want)


def test_scalar_kernel_segments_and_rows_keep_results():
    """The scalar divergence marches down 62-cell row tiles: a domain of 3 tiles (the last one
    partial) with an odd row count, cut into 1, 2 and 4 vertical segments on 1, 2 and 4 rows per
    block, every combination bitwise equal to the oracle over 4 steps (moist, 3rd-order theta,
    free slip, terrain-following layers from level 5, where the 4-segment cut starts its second
    segment)."""
    cfg = dataclasses.replace(tiny_config(ny=7, nx=150, n_layers=16, moist=True, order=3, freeslip=True), n_oro_layers=12)
    g, st = make_case(cfg, cloud_frac=0.3)
    n = 4
    want = _oracle(cfg, g, st, n)
    for rows in (1, 2, 4):
        for seg in (1, 2, 4):
            s = Step(cfg, g)
            s.set_tiles({k: rows for k in TILE_KERNELS})
            s.segments = seg
            assert s.segments == seg
            s.start(st)
            s.run(0, n)
            got = s.get_state(n % 2)
            s.close()
            print(f"rows {rows}, segments {seg}", flush=True)
            assert all(np.array_equal(got[k], want[k]) for k in want), (rows, seg)
    s = Step(cfg, g)
    for bad in (3, 8):  # not a power of two; fewer than 4 levels per segment
        with pytest.raises(LgameError, match="segment"):
            s.segments = bad
    s.close()


VER_CASES = {
    "dry_o2": (tiny_config(ny=6, nx=70, n_layers=9), 4, 0.0),
    "moist_o3_freeslip": (tiny_config(ny=9, nx=11, n_layers=7, moist=True, order=3, freeslip=True), 5, 0.3),
    "three_layers": (tiny_config(ny=3, nx=4, n_layers=3, moist=True), 4, 0.2),
    "diff_pbl_soil_sens": (dataclasses.replace(tiny_config(moist=True, order=3).set_diffusion(True), lpbl=True,
                                               lprog_soil_temp=True, lsfc_sensible_heat_flux=True), 6, 0.3),
}


@pytest.mark.parametrize("name", list(VER_CASES))
def test_vertical_solver_passes_keep_results(name):
    """The vertical solver in one pass per column and in three passes (rows, Thomas recursion,
    new state), each bitwise equal to the oracle; small domains default to three passes."""
    cfg, n, cf = VER_CASES[name]
    g, st = make_case(cfg, cloud_frac=cf)
    want = _oracle(cfg, g, st, n)
    for passes in (1, 3):
        s = Step(cfg, g)
        assert s.column_passes == 3  # fewer column wavefronts than SIMDs
        s.column_passes = passes
        assert s.column_passes == passes
        s.start(st)
        s.run(0, n)
        got = s.get_state(n % 2)
        s.close()
        print(f"{name}: passes {passes}", flush=True)
        assert all(np.array_equal(got[k], want[k]) for k in want), (name, passes)
    s = Step(cfg, g)
    with pytest.raises(LgameError, match="passes"):
        s.column_passes = 2
    s.close()


def test_vertical_solver_default_passes_by_size():
    """One pass per column once the domain has as many 64-column wavefronts as the GPU has SIMDs."""
    for n, want in ((128, 3), (256, 1)):
        cfg = dry3d_config(n=n, n_layers=10)
        g, st = make_case(cfg)
        s = Step(cfg, g)
        assert s.column_passes == want, n
        s.close()


@pytest.mark.parametrize("mode", [Ensemble.GRAPH, Ensemble.SERIAL, Ensemble.BATCHED])
def test_ensemble_members_match_oracle(mode):
    cfgs = [tiny_config(ny=5, nx=7, n_layers=6, moist=True, order=3),
            tiny_config(ny=9, nx=11, n_layers=8),
            # even nx: two cells per thread in the potential vorticity of a single-domain launch,
            # one cell per thread in the batched one; nx not a multiple of 64 in the column solvers
            tiny_config(ny=6, nx=70, n_layers=7, moist=True),
            dry3d_config(n=17, n_layers=10),
            # a member with the reference's default physics: diffusion, PBL, prognostic soil
            dataclasses.replace(tiny_config(ny=7, nx=9, n_layers=8, moist=True), lpbl=True, lprog_soil_temp=True,
                                lsfc_sensible_heat_flux=True, lsfc_phase_trans=True).set_diffusion()]
    cases = [make_case(c, cloud_frac=0.1 if c.lmoist else 0.0) for c in cfgs]
    steps = []
    for c, (g, st) in zip(cfgs, cases):
        s = Step(c, g)
        s.start(st)
        if mode == Ensemble.GRAPH:
            s.set_tiles({k: (1, 2, 4)[i % 3] for i, k in enumerate(TILE_KERNELS)})
        steps.append(s)
    e = Ensemble(steps)
    n = 7
    print(f"ensemble of {len(steps)} members, mode {mode}, {n} steps", flush=True)
    e.run(0, n, mode=mode)
    e.sync()
    for c, (g, st), s in zip(cfgs, cases, steps):
        got = s.get_state(n % 2)
        want = _oracle(c, g, st, n)
        assert all(np.array_equal(got[k], want[k]) for k in want), f"member {c.ny}x{c.nx}x{c.n_layers}"
    e.close()
    for s in steps:
        s.close()


def test_batched_radiation_set_up_after_the_first_step():
    """The batched groups are made at the first batched step; a radiation set up afterwards (surface
    data, surrogate) is taken up at the first radiation update: the batched steps stay bitwise equal
    to serial ones."""
    from lgame import surrogate
    from lgame.cases import moist3d_config, radiation_surface
    from lgame.config import PREC_MIXED, RAD_SURROGATE

    mlp = surrogate.load(surrogate.DEFAULT_WEIGHTS)
    cfgs = [dataclasses.replace(moist3d_config(n=n, n_layers=60, seed=30 + n), precision=PREC_MIXED, lrad=True,
                                rad_mode=RAD_SURROGATE, dtime_rad=900.0) for n in (33, 47)]
    cases = [make_case(c, cloud_frac=0.005) for c in cfgs]
    t0 = 1576843200.0 + 3 * 3600.0
    out = {}
    for mode in (Ensemble.BATCHED, Ensemble.SERIAL):
        steps = []
        for c, (g, st) in zip(cfgs, cases):
            s = Step(c, g)
            s.start(st)
            steps.append(s)
        e = Ensemble(steps)
        e.step(0, mode=mode)
        for c, (g, st), s in zip(cfgs, cases, steps):
            s.set_rad_surface(radiation_surface(c, g, st))
            s.attach_mlp(mlp, tol_k_per_day=1.0, audit_fraction=0.05, seed=3)
        for k in range(1, 4):
            e.step(k, lrad_update=(k == 1), t_0=t0 + k * cfgs[0].dt, mode=mode)
        e.sync()
        out[mode] = [s.get_state(0) for s in steps]
        assert all(s.rad_stats()["calls"] == 1 for s in steps)
        e.close()
        for s in steps:
            s.close()
    for a, b in zip(out[Ensemble.BATCHED], out[Ensemble.SERIAL]):
        assert all(np.array_equal(a[k], b[k]) for k in a)
