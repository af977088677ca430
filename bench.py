"""Benchmark of the MI355X HEVI predictor-corrector step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--form both|mixed|fp64]
                    [--diffusion] [--pbl] [--physics]

Runs the step on one GPU with the state resident in HBM, times exactly K steps between two
stream synchronizations, and prints one JSON line (last line of stdout).  Progress goes to stderr.

Workloads (synthetic initial states, see l-game_amd/lgame/cases.py):
  moist512        512x512x60 moist (7 constituents), dx 2 km, dt 3.22 s, radiation every 900 s
  dry128          128x128x50 dry, rotated lat-lon, dx 2 km, dt 3.22 s
  schaer          ny=3, nx=201, 40 layers, dx 500 m, dt 0.805 s
  surrogate-train harvest radiation columns from the model and train the 304 -> 3x512 -> 62 MLP
  ensemble        48 moist members of uneven size (nx, ny ~ U[64,384], 40-70 layers), mixed form with
                  the radiation surrogate, autotuned block shapes: one launch per kernel phase over
                  all members vs one graph launch per step vs serial launches
  grid-setup      grid generation (grid_setup, bg_setup) of the moist512 grid on the device
Block shapes of the stencil kernels are chosen by lgame_autotune before the timed steps
(--no-autotune keeps the default 64x4).  --diffusion turns on the reference's default diffusion
(momentum, temperature and mass, Smagorinsky horizontal and TKE vertical coefficients); --pbl the
PBL friction of the lowest-layer wind with the surface turbulence quantities; --physics all of
the reference's default namelist (diffusion, PBL, 5-layer prognostic soil, surface sensible heat
flux).  The default moist512 run also times both forms with that default physics
(`reference_physics` in the JSON line; --no-physics-extra skips it).

Two forms of the step are timed: the plain FP64 form with the gray two-stream radiation solver
and the mixed form (FP32 stencils and explicit tendencies, FP64 column solves, density updates
and budget sums) with the MLP radiation surrogate and its 1% audit.  Radiation updates follow
the model's cadence (dtime_rad / dtime steps), the first one inside the first timed step.  The
headline value is the mixed form; `forms` lists both and `speedup_mixed_vs_fp64` their ratio.
Every kernel is timed with HIP events on the library's own stream in a separate pass
(`forms.*.kernel_ms_per_step`).  `roofline` is the kernel that takes the most time per step in the
headline form, `roofline_fp64` the same in the plain FP64 form, each against its algorithmic
bytes per grid point (every array it must read or write counted once, at the width the form
stores it: KERNEL_BYTES); `roofline_solver_ver` keeps the implicit vertical solver
(three_band_solver_ver) of the FP64 form, 152 B (predictor) and 176 B (corrector) per grid point.
`step_roofline` puts the whole FP64 step against the 1296 B per grid-point step model, beside
the measured traffic.
The CPU baseline is the oracle (the C restatement) on all host cores, as independent instances of
a bounded sample of the same workload type, one per thread, with the same radiation schedule;
each GPU form is run on that sample and its per-field RMS and max error against the oracle are
reported in `field_error`.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "l-game_amd"))

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# algorithmic bytes of three_band_solver_ver per grid point (predictor, corrector): the FP64 form
# reads 14 (17) arrays and writes 5, all FP64; the mixed form reads its three tendency arrays
# (rho_gas, rhotheta_v, w) as FP32, 12 B fewer
SOLVER_BYTES = {"fp64": {1: 152.0, 2: 176.0}, "mixed": {1: 140.0, 2: 164.0}}
# Algorithmic bytes per grid point and launch of the kernels that can dominate a step (moist,
# theta order 2, no diffusion), per form; each is launched twice per step (predictor, corrector).
# scalar_div (k_scalar_col, scalar_tend_expl): reads the gas and six tracer densities,
#   theta_v_pert, exner_pert and the new u, v (11 FP64 arrays); theta_v_bg, area_x, area_y, area_z,
#   volume, slope_x, slope_y, the four horizontal inner-product weights and exner_bg (12 grid
#   arrays of the form's width); the heating terms (heating_diss, radiation, condensates),
#   temperature and the six tracer tendencies it updates (10 work arrays); writes the seven density
#   and the rhotheta_v tendencies (8 work arrays): FP64 88 + 96 + 80 + 64 = 328 B, mixed
#   88 + 48 + 40 + 32 = 208 B.
# solve_gen (k_solve_gen, three_band_solver_gen_densities): per tracer the old density, the new
#   density and the tendency read and the new density written (6 x 32 B in FP64), shared per
#   column old and new w, area_z, volume, the sink velocity of cloud water and the temperature
#   read and the latent heating written (56 B): 248 B; the mixed form reads the six tendencies,
#   the temperature and writes the heating at 4 B: 216 B.
# solve_ver: the mean of the predictor and corrector counts of SOLVER_BYTES.
KERNEL_BYTES = {"scalar_div": {"fp64": 328.0, "mixed": 208.0}, "solve_gen": {"fp64": 248.0, "mixed": 216.0},
                "solve_ver": {"fp64": 164.0, "mixed": 152.0}}
KERNEL_NAMES = {"scalar_div": "k_scalar_col (scalar_tend_expl, column-marching)",
                "solve_gen": "k_solve_gen (three_band_solver_gen_densities, FP64 solve)",
                "solve_ver": "k_solve_ver (three_band_solver_ver, FP64 solve)"}
STEP_BYTES = {"dry": 920.0, "moist": 1296.0}  # per grid-point step, whole-step traffic model (FP64)
# HBM bytes per k_solve_ver dispatch, averaged over predictor and corrector, from two rocprofv3
# --pmc passes (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0, the factors a 1 GiB device copy gives), per
# form of the step; profiles/ holds the commands and the per-kernel tables.
SOLVER_TRAFFIC = {("moist512", "fp64"): 4.369e+09, ("moist512", "mixed"): 4.128e+09}
# the same for the kernels of KERNEL_BYTES, (workload, kernel, form) -> bytes per launch
KERNEL_TRAFFIC = {("moist512", "solve_ver", "fp64"): 4.369e+09, ("moist512", "solve_ver", "mixed"): 4.128e+09,
                  ("moist512", "scalar_div", "fp64"): 5.932e+09, ("moist512", "scalar_div", "mixed"): 4.264e+09,
                  ("moist512", "solve_gen", "fp64"): 5.652e+09, ("moist512", "solve_gen", "mixed"): 5.263e+09}
# measured HBM bytes per grid-point step of the whole FP64 step (same counters, all kernels)
STEP_TRAFFIC = {"moist512": 3489.0}
# the reference's own CPU path, flang -O2 -fopenmp on 8 cores, measured on an 8-core Xeon host (the
# reference tree does not travel, so it cannot be timed beside the GPU)
REFERENCE_CPU = {"dry128": {"value": 3.19e6, "grid": "129x129x50 dry", "cores": 8},
                 "moist512": {"value": 2.05e6, "grid": "513x513x60 dry, no radiation", "cores": 8},
                 "schaer": {"value": 1.78e6, "grid": "201x3x40 Schaer", "cores": 8}}
METRIC = "grid-point-steps/s at 1 GPU: mixed+surrogate vs FP64 plain, at stated field error"
DEFAULT_SURROGATE = os.path.join(ROOT, "l-game_amd", "data", "rad_surrogate.npz")
SURROGATE_WEIGHTS = DEFAULT_SURROGATE
T_EPOCH = 1576843200.0 + 3 * 3600.0  # a time of day with sunlit columns
# the moist workloads start with cloud liquid (above 0 C) and cloud ice (below) (config 3 names both
# as separate mass densities): up to 0.5% of the vapour density at saturation, times a smooth
# random cloud cover per column (lgame.cases.cloud_patches), as in the surrogate's training columns
CLOUD_FRAC = 0.005
NO_AUTOTUNE = False
MIXED_RAD = "surrogate"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name):
    from lgame.cases import dry3d_config, moist3d_config, schaer_config
    from lgame.config import RAD_TEACHER

    if name == "moist512":
        rad = dict(lrad=True, rad_mode=RAD_TEACHER, dtime_rad=900.0)
        return (dataclasses.replace(moist3d_config(n=512, n_layers=60, dtime=3.22), **rad),
                dataclasses.replace(moist3d_config(n=65, n_layers=60, dtime=3.22), **rad))
    if name == "dry128":
        return dry3d_config(n=128, n_layers=50, dtime=3.22), dry3d_config(n=65, n_layers=50, dtime=3.22)
    if name == "schaer":
        return schaer_config(nx=201, n_layers=40), schaer_config(nx=201, n_layers=40)
    raise SystemExit(f"unknown workload {name}")


def describe(cfg):
    kind = f"moist (patchy cloud liquid and ice, up to {100 * CLOUD_FRAC:g}% of the vapour)" if cfg.lmoist else "dry"
    rad = f", radiation every {cfg.dtime_rad:g} s" if cfg.lrad else ", radiation off"
    diff = (f", diffusion on ({cfg.diff_coeff_scheme_h} horizontal, TKE vertical)" if cfg.diffusion else "")
    diff += ", PBL friction on" if cfg.lpbl else ""
    diff += f", prognostic soil ({cfg.nsoillays} layers) with the surface sensible heat flux" if cfg.lprog_soil_temp else ""
    return (f"{kind} {cfg.ny}x{cfg.nx}x{cfg.n_layers}, dx {cfg.dx:g} m, dt {cfg.dt:g} s, "
            f"theta order {cfg.theta_adv_order}{rad}{diff}")


def surrogate_mlp():
    """The trained radiation surrogate: the committed weights, or a quick training run."""
    from lgame import surrogate

    if os.path.exists(SURROGATE_WEIGHTS):
        log(f"radiation surrogate: {os.path.relpath(SURROGATE_WEIGHTS, ROOT)}")
        return surrogate.load(SURROGATE_WEIGHTS)
    if SURROGATE_WEIGHTS != DEFAULT_SURROGATE:
        raise SystemExit(f"no surrogate weights at {SURROGATE_WEIGHTS}")
    log("radiation surrogate: no trained weights, training a small one on the GPU")
    m, _, rep = surrogate.build(states=8, epochs=30)
    log(f"surrogate trained: {rep}")
    return m


# the vertical solver's passes per form (1: k_solve_ver, one column per lane; 3: k_ver_rows,
# k_ver_thomas, k_ver_update on domains with fewer column wavefronts than SIMDs)
SOLVER_PASSES = {}
VER_THREE_PASS = "k_ver_rows + k_ver_thomas + k_ver_update (three_band_solver_ver in three passes, FP64 solve)"


def make_step(cfg, grid, surface, form, mlp):
    from lgame.config import PREC_FP64, PREC_MIXED, RAD_SURROGATE, RAD_TEACHER
    from lgame.gpu import Step

    if form == "mixed":
        rad_mode = (RAD_TEACHER if MIXED_RAD == "solver" else RAD_SURROGATE) if cfg.lrad else cfg.rad_mode
        c = dataclasses.replace(cfg, precision=PREC_MIXED, rad_mode=rad_mode)
    else:
        c = dataclasses.replace(cfg, precision=PREC_FP64, rad_mode=RAD_TEACHER if cfg.lrad else cfg.rad_mode)
    s = Step(c, grid)
    SOLVER_PASSES[form] = s.column_passes
    if cfg.lrad:
        s.set_rad_surface(surface)
        if form == "mixed" and MIXED_RAD == "surrogate":
            s.attach_mlp(mlp, tol_k_per_day=1.0, audit_fraction=0.01, seed=7)
    return s


def run_steps(step, cfg, first, n):
    """n steps from counter `first`, radiation at the model's cadence starting at step `first`."""
    every = max(1, int(round(cfg.dtime_rad / cfg.dt))) if cfg.lrad else 0
    for k in range(first, first + n):
        lrad = bool(every) and (k - first) % every == 0
        step.pchevi(k % 2, 1 - k % 2, k, lrad, T_EPOCH + k * cfg.dt)


def time_form(cfg, grid, st, surface, form, mlp, steps, warmup):
    """Throughput of one form of the step plus event timing of the FP64 column solver."""
    from lgame.gpu import PROF_KERNEL0, PROF_RADIATION, PROF_SOLVE_GEN, PROF_SOLVE_VER, PROF_STEP, TILE_KERNELS

    step = make_step(cfg, grid, surface, form, mlp)
    step.start(st)
    tiles = None if NO_AUTOTUNE else step.autotune(2)
    run_steps(step, cfg, 0, warmup)
    step.sync()
    # timed region: exactly K steps between two synchronizations of the library's stream
    t0 = time.perf_counter()
    run_steps(step, cfg, warmup, steps)
    step.sync()
    elapsed = time.perf_counter() - t0
    counter = warmup + steps
    bad = step.count_nonfinite(counter % 2)
    rad = step.rad_stats() if cfg.lrad else {}
    # kernel timing with HIP events on the library's stream (separate pass, every kernel on its own)
    step.profile(2)
    nprof = max(2, min(steps, 6))
    run_steps(step, cfg, counter, nprof)
    prof = {g: step.profile_read(g) for g in (PROF_SOLVE_VER, PROF_SOLVE_GEN, PROF_STEP, PROF_RADIATION)}
    kernels = {"solve_ver": prof[PROF_SOLVE_VER], "solve_gen": prof[PROF_SOLVE_GEN]}
    for i, name in enumerate(TILE_KERNELS):
        ms, n = step.profile_read(PROF_KERNEL0 + i)
        if n:
            kernels[name] = (ms, n)
    prof["kernels"] = {k: (ms / nprof, n // nprof) for k, (ms, n) in kernels.items() if n}  # per step
    step.profile(False)
    step.close()
    return elapsed, bad, prof, dict(rad, tiles=tiles)


def host_threads():
    """Host cores this process may use, at most 16 (the share of the host a one-GPU job gets)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(cfg, nsteps, forms, mlp, threads):
    """The oracle on all host cores: `threads` independent instances of the same sample, one per
    thread (the ctypes calls release the interpreter lock, so they run in parallel), timed
    together; plus the GPU forms' per-field error against the oracle on that sample."""
    import threading

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from lgame.cases import cloud_patches, make_case, radiation_surface
    from lgame.metrics import field_errors, worst
    from pyoracle import CpuOracle

    grid, st = make_case(cfg, cloud_frac=CLOUD_FRAC)
    if cfg.lmoist:
        cloud_patches(cfg, st)
    surface = radiation_surface(cfg, grid, st)
    every = max(1, int(round(cfg.dtime_rad / cfg.dt))) if cfg.lrad else 0
    orcs = []
    for _ in range(threads):
        o = CpuOracle(cfg, grid)
        o.start(st)
        o.set_rad_surface(surface)
        orcs.append(o)
    lib = orcs[0].lib()

    def run(o):
        for k in range(nsteps):
            lrad = int(bool(every) and k % every == 0)
            lib.orc_step_rad(o.h, k % 2, 1 - k % 2, k, lrad, T_EPOCH + k * cfg.dt)

    pool = [threading.Thread(target=run, args=(o,)) for o in orcs]
    t0 = time.perf_counter()
    for t in pool:
        t.start()
    for t in pool:
        t.join()
    dt = time.perf_counter() - t0
    want = orcs[0].get_state(nsteps % 2)
    same = all(all(np.array_equal(o.get_state(nsteps % 2)[k], want[k]) for k in want) for o in orcs[1:])
    if not same:
        raise RuntimeError("the oracle instances disagree")
    del orcs
    errs = {}
    for form in forms:
        g = make_step(cfg, grid, surface, form, mlp)
        g.start(st)
        run_steps(g, cfg, 0, nsteps)
        got = g.get_state(nsteps % 2)
        g.close()
        e = field_errors(got, want)
        errs[form] = {"max_rel": worst(e, "max_rel"), "rms_rel": worst(e, "rms_rel"),
                      "per_field": {k: [v["rms_rel"], v["max_rel"]] for k, v in e.items()}}
    return threads * cfg.gridpoints * nsteps / dt, dt, errs


def copy_calibration(nbytes):
    """One device-to-device copy of a known size (the reference kernel for the HBM byte counters)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes} bytes device to device")


def bench_training(args):
    """The surrogate training run: harvest columns from the model, train with BF16 MFMA GEMMs."""
    from lgame import surrogate

    horizon = int(round(args.train_horizon_hours * 3600.0 / 3.22))
    m, norm, rep = surrogate.build(states=args.train_states, calls=args.train_calls, times=args.train_times,
                                   epochs=args.train_epochs, batch=args.train_batch, lr=args.train_lr, progress=log,
                                   horizon_steps=horizon, hidden=args.train_hidden)
    if args.save_surrogate:
        surrogate.save(args.save_surrogate, m, norm)
        log(f"saved {args.save_surrogate}")
    out = {
        "metric": "surrogate training samples/s at 1 GPU (BF16 MFMA, FP32 Adam)",
        "value": rep["samples_per_s"], "unit": "samples/s", "n_gpus": 1, "steps": rep["epochs"], "warmup": 0,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "surrogate-train: columns from 33x33x60 moist runs"
                               + (f" over {args.train_horizon_hours:g} simulated hours" if horizon else "")
                               + f", 304 -> 3x{args.train_hidden} GELU -> 62",
                   "batch": rep["batch"], "columns": rep["train_columns"]},
        "roofline": {"kernel": "k_gemm (all MLP GEMMs of training)", "bound": "mfma", "achieved": rep["train_tflops"],
                     "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": rep["train_tflops"] / BF16_PEAK_TFLOPS,
                     "traffic": None},
        "report": rep,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_accuracy(args, cfg, grid, st, surface, mlp):
    """The mixed form + surrogate against the plain FP64 form (bitwise equal to the oracle on every
    tested configuration) over args.accuracy_hours simulated hours from the same initial state,
    radiation at the model's cadence: per-field RMS and max error relative to the FP64 fields,
    every simulated hour.  Beside it, the same FP64 step from an initial state whose u differs by
    1e-7 relative: the growth of that error is the flow's own, which no form of the step can
    avoid."""
    from lgame.metrics import field_errors, worst

    n = int(round(args.accuracy_hours * 3600.0 / cfg.dt))
    per_hour = max(1, int(round(3600.0 / cfg.dt)))
    every = max(1, int(round(cfg.dtime_rad / cfg.dt))) if cfg.lrad else 0
    st_p = dict(st)
    st_p["wind_u"] = st["wind_u"] * (1.0 + 1e-7)
    runs = {"fp64": make_step(cfg, grid, surface, "fp64", None), "mixed": make_step(cfg, grid, surface, "mixed", mlp),
            "fp64_perturbed": make_step(cfg, grid, surface, "fp64", None)}
    for name, step in runs.items():
        step.start(st_p if name == "fp64_perturbed" else st)
    history = []
    t0 = time.perf_counter()
    for k0 in range(0, n, per_hour):
        m = min(per_hour, n - k0)
        for k in range(k0, k0 + m):
            lrad = bool(every) and k % every == 0
            for step in runs.values():
                step.pchevi(k % 2, 1 - k % 2, k, lrad, T_EPOCH + k * cfg.dt)
        k = k0 + m
        ref = runs["fp64"].get_state(k % 2)
        row = {"hours": k * cfg.dt / 3600.0}
        for name in ("mixed", "fp64_perturbed"):
            bad = runs[name].count_nonfinite(k % 2)
            err = field_errors(runs[name].get_state(k % 2), ref) if bad == 0 else None
            row[name] = {"nonfinite": bad, "worst_rms": worst(err, "rms_rel") if err else None,
                         "worst_max": worst(err, "max_rel") if err else None, "fields": err}
        history.append(row)
        log(f"{row['hours']:.2f} h: mixed worst rms {row['mixed']['worst_rms']}, perturbed FP64 worst rms "
            f"{row['fp64_perturbed']['worst_rms']}, {time.perf_counter() - t0:.0f} s")
    rad = runs["mixed"].rad_stats() if cfg.lrad else {}
    for step in runs.values():
        step.close()
    last = history[-1]
    out = {
        "metric": "mixed + surrogate field error against the plain FP64 step after simulated hours",
        "value": last["mixed"]["worst_rms"], "unit": "relative RMS error (worst field)", "n_gpus": 1, "steps": n,
        "warmup": 0, "higher_is_better": False, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {describe(cfg)}", "simulated_hours": args.accuracy_hours},
        "worst_max_rel": last["mixed"]["worst_max"],
        "perturbed_fp64_worst_rms": last["fp64_perturbed"]["worst_rms"],
        "surrogate": rad,
        "hourly": [{"hours": r["hours"], "mixed_worst_rms": r["mixed"]["worst_rms"],
                    "mixed_worst_max": r["mixed"]["worst_max"], "mixed_nonfinite": r["mixed"]["nonfinite"],
                    "perturbed_fp64_worst_rms": r["fp64_perturbed"]["worst_rms"]} for r in history],
        "fields": last["mixed"]["fields"],
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_grid(args):
    """Grid generation (grid_setup and bg_setup) on the device for the moist 512x512x60 grid:
    lgame_grid_setup including the copies of every field to host memory, against the host
    restatement (lgame.grid, numpy) on a 128x128x60 sample."""
    from lgame.cases import moist3d_config
    from lgame.grid import grid_setup
    from lgame.gpu import grid_setup_device

    cfg = moist3d_config(n=512, n_layers=60, dtime=3.22)
    points = cfg.gridpoints
    for _ in range(args.warmup):
        grid_setup_device(cfg)
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        grid_setup_device(cfg)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[len(times) // 2]
    log(f"device grid 512x512x60: median {1e3 * t:.1f} ms over {len(times)} calls")
    cpu = None
    if not args.no_cpu:
        sample = moist3d_config(n=128, n_layers=60, dtime=3.22)
        t0 = time.perf_counter()
        grid_setup(sample)
        tc = time.perf_counter() - t0
        cpu = {"value": sample.gridpoints / tc, "unit": "grid points/s", "cores": 1, "kind": "port",
               "sample": f"lgame.grid.grid_setup (numpy, one thread), moist 128x128x60, {tc:.1f} s"}
        log(f"host grid 128x128x60: {tc:.2f} s")
    out = {
        "metric": "grid points generated per second (grid_setup and bg_setup) at 1 GPU",
        "value": points / t, "unit": "grid points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * t, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "grid-setup: moist 512x512x60 grid, 5 Gaussian hills, SLEVE, rotated lat-lon, "
                               "including the copies of all fields to host memory", "grid_points": points},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_ensemble(args):
    """The 48-member ensemble in the mixed form with the radiation surrogate and its 1 % audit:
    aggregate throughput with one launch per kernel phase over all members (batched), with one
    graph launch per step, and with serial launches.  As in moist512, the timed steps start with a
    radiation update."""
    from lgame import surrogate
    from lgame.cases import ensemble_members, make_case, radiation_surface
    from lgame.config import PREC_MIXED, RAD_SURROGATE
    from lgame.gpu import Ensemble, Step, grid_setup_device

    cfgs = ensemble_members(args.members)
    mlp = surrogate.load(args.surrogate or surrogate.DEFAULT_WEIGHTS)
    steps, cells = [], []
    t0 = time.perf_counter()
    tiles = []
    t_grid = 0.0
    for i, c in enumerate(cfgs):
        c = dataclasses.replace(c, precision=PREC_MIXED, lrad=True, dtime_rad=900.0, rad_mode=RAD_SURROGATE)
        cfgs[i] = c
        tg = time.perf_counter()
        g = grid_setup_device(c)  # the member's grid, generated on the device
        t_grid += time.perf_counter() - tg
        g, st = make_case(c, grid=g, cloud_frac=0.01)
        s = Step(c, g)
        s.set_rad_surface(radiation_surface(c, g, st))
        s.attach_mlp(mlp, tol_k_per_day=1.0, audit_fraction=0.01, seed=7)
        s.start(st)
        del g
        tiles.append(s.autotune(2) if not args.no_autotune else None)
        s.start(st)
        del st
        steps.append(s)
        cells.append(c.gridpoints)
        log(f"member {i}: {c.ny}x{c.nx}x{c.n_layers} ({c.gridpoints / 1e6:.2f} M points), tiles {tiles[-1]}, "
            f"segments {s.segments}")
    t_setup = time.perf_counter() - t0
    log(f"set-up of {len(steps)} members in {t_setup:.1f} s ({t_grid:.1f} s of it device grid generation)")
    total = sum(cells)
    dt = cfgs[0].dt
    e = Ensemble(steps)
    res = {}
    k = 0
    for mode, name in ((Ensemble.BATCHED, "batched"), (Ensemble.GRAPH, "graph"), (Ensemble.SERIAL, "serial")):
        for _ in range(1 + args.warmup):  # counter 0 (eager in the graph mode) and the warm-up
            e.step(k, lrad_update=(k == 0), t_0=T_EPOCH + k * dt, mode=mode)
            k += 1
        e.sync()
        t1 = time.perf_counter()
        for q in range(args.steps):
            e.step(k, lrad_update=(q == 0), t_0=T_EPOCH + k * dt, mode=mode)
            k += 1
        e.sync()
        res[name] = time.perf_counter() - t1
        log(f"{name}: {args.steps} ensemble steps (one radiation update) in {res[name]:.3f} s")
    bad = sum(s.count_nonfinite(k % 2) for s in steps)
    rad = [s.rad_stats() for s in steps]
    # per-member wall time: each member alone, eager launches, no radiation update
    per = []
    for s, n in zip(steps, cells):
        s.sync()
        t1 = time.perf_counter()
        s.run(k, 4)
        s.sync()
        per.append(1e3 * (time.perf_counter() - t1) / 4)
    e.close()
    for s in steps:
        s.close()
    value = total * args.steps / res["batched"]
    out = {
        "metric": "aggregate grid-point-steps/s of a batched ensemble at 1 GPU",
        "value": value, "unit": "grid-point-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * res["batched"] / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"ensemble: {len(steps)} moist members, nx,ny ~ U[64,384], 40-70 layers, mixed form, "
                               f"radiation by the MLP surrogate with a 1% solver audit, one radiation update in the "
                               f"timed steps, {total / 1e6:.1f} M points in total, one launch per kernel phase over "
                               f"all members", "members": len(steps), "grid_points": total,
                   "autotuned": not args.no_autotune, "nonfinite_after_run": int(bad)},
        "per_member_value_mean": value / len(steps),
        "graph": {"value": total * args.steps / res["graph"], "ms_per_step": 1e3 * res["graph"] / args.steps},
        "serial": {"value": total * args.steps / res["serial"], "ms_per_step": 1e3 * res["serial"] / args.steps},
        "speedup_batched_vs_graph": res["graph"] / res["batched"],
        "speedup_batched_vs_serial": res["serial"] / res["batched"],
        "per_member_alone_ms_per_step": {"mean": sum(per) / len(per), "min": min(per), "max": max(per)},
        "radiation": {"calls": rad[0]["calls"], "audited": sum(r["audited"] for r in rad),
                      "fallbacks": sum(r["fallbacks"] for r in rad)},
        "setup_s": {"total": t_setup, "device_grid_generation": t_grid},
    }
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="moist512",
                    choices=["moist512", "dry128", "schaer", "surrogate-train", "ensemble", "grid-setup"])
    ap.add_argument("--members", type=int, default=48)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--diffusion", action="store_true",
                    help="turn on the six diffusion switches (the reference's default namelist)")
    ap.add_argument("--pbl", action="store_true",
                    help="turn on the PBL friction and surface turbulence quantities (lpbl)")
    ap.add_argument("--physics", action="store_true",
                    help="the reference's default namelist: the six diffusion switches, lpbl, the prognostic "
                         "soil (5 layers), the surface sensible heat flux and lsfc_phase_trans")
    ap.add_argument("--form", default="both", choices=["both", "mixed", "fp64"],
                    help="which forms of the step to time (the headline value is the mixed form)")
    ap.add_argument("--cpu-steps", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-physics-extra", action="store_true",
                    help="moist512: skip timing the same workload with the reference's default physics")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES (for HBM counter scaling)")
    ap.add_argument("--train-states", type=int, default=24)
    ap.add_argument("--train-epochs", type=int, default=60)
    ap.add_argument("--train-batch", type=int, default=8192)
    ap.add_argument("--train-lr", type=float, default=1e-3)
    ap.add_argument("--train-calls", type=int, default=2, help="radiation updates harvested per training run")
    ap.add_argument("--train-times", type=int, default=6, help="times of day per harvested update")
    ap.add_argument("--train-hidden", type=int, default=512, help="width of the surrogate's hidden layers")
    ap.add_argument("--train-horizon-hours", type=float, default=0.0,
                    help="spread the harvested updates of each run over this many simulated hours")
    ap.add_argument("--save-surrogate", default="")
    ap.add_argument("--surrogate", default="", help="surrogate weights (default: l-game_amd/data/rad_surrogate.npz)")
    ap.add_argument("--mixed-radiation", default="surrogate", choices=["surrogate", "solver"],
                    help="radiation of the mixed form: the MLP surrogate (default) or the gray two-stream solver")
    ap.add_argument("--accuracy-hours", type=float, default=0.0,
                    help="instead of timing: run both forms for this many simulated hours and report the mixed "
                         "form's field errors against the FP64 form")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the step runs on one GPU only (--gpus {args.gpus} requested); "
              "multi-GPU sharding is a design note, not implemented", file=sys.stderr)
        return 2
    global NO_AUTOTUNE, MIXED_RAD, SURROGATE_WEIGHTS
    NO_AUTOTUNE = args.no_autotune
    MIXED_RAD = args.mixed_radiation
    if args.surrogate:
        SURROGATE_WEIGHTS = os.path.abspath(args.surrogate)
    if args.workload == "surrogate-train":
        return bench_training(args)
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "grid-setup":
        return bench_grid(args)

    from lgame.cases import cloud_patches, make_case, radiation_surface
    from lgame.gpu import PROF_RADIATION, PROF_SOLVE_VER, PROF_STEP

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    cfg, sample_cfg = workload(args.workload)
    if args.physics:
        args.diffusion = args.pbl = True
        surface = {"lprog_soil_temp": True, "lsfc_sensible_heat_flux": True, "lsfc_phase_trans": True, "nsoillays": 5}
        cfg = dataclasses.replace(cfg, **surface)
        sample_cfg = dataclasses.replace(sample_cfg, **surface)
    if args.diffusion:
        cfg = dataclasses.replace(cfg).set_diffusion(True)
        sample_cfg = dataclasses.replace(sample_cfg).set_diffusion(True)
    if args.pbl:
        cfg = dataclasses.replace(cfg, lpbl=True)
        sample_cfg = dataclasses.replace(sample_cfg, lpbl=True)
    forms = {"both": ("fp64", "mixed"), "mixed": ("mixed",), "fp64": ("fp64",)}[args.form]
    mlp = surrogate_mlp() if (cfg.lrad and "mixed" in forms) else None
    t0 = time.perf_counter()
    log(f"setting up {describe(cfg)}")
    grid, st = make_case(cfg, cloud_frac=CLOUD_FRAC)
    if cfg.lmoist:
        cloud_patches(cfg, st)
    surface = radiation_surface(cfg, grid, st) if cfg.lrad else None
    log(f"set-up done in {time.perf_counter() - t0:.1f} s")
    if args.accuracy_hours > 0.0:
        return bench_accuracy(args, cfg, grid, st, surface, mlp if mlp is not None or not cfg.lrad else surrogate_mlp())
    cells = cfg.gridpoints
    res = {}
    for form in forms:
        elapsed, bad, prof, rad = time_form(cfg, grid, st, surface, form, mlp, args.steps, args.warmup)
        res[form] = (elapsed, bad, prof, rad)
        log(f"{form}: {args.steps} steps in {elapsed:.3f} s ({1e3 * elapsed / args.steps:.2f} ms/step), "
            f"non-finite values: {bad}, radiation {rad}")
    del grid, st
    head = "mixed" if "mixed" in res else "fp64"
    elapsed, bad, prof, rad = res[head]
    value = cells * args.steps / elapsed
    # the roofline objects: the kernel with the most time per step in each form
    def kernel_roofline(form):
        ks = res[form][2]["kernels"]
        name = max(ks, key=lambda k: ks[k][0])
        ms, n = ks[name]
        b = KERNEL_BYTES.get(name, {}).get(form)
        kname = VER_THREE_PASS if name == "solve_ver" and SOLVER_PASSES.get(form) == 3 else KERNEL_NAMES.get(name, name)
        out_r = {"kernel": f"{kname} of the {form} form", "bound": "hbm", "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "ms_per_step": ms, "launches_per_step": n, "avg_launch_us": 1e3 * ms / max(n, 1),
                 "share_of_step": ms / (1e3 * res[form][0] / args.steps)}
        if b is None:
            return dict(out_r, achieved=None, frac=None, traffic=None)
        gbs = b * cells * n / (ms * 1e-3) / 1e9
        return dict(out_r, achieved=gbs, frac=gbs / HBM_PEAK_GBS, algorithmic_bytes_per_point=b,
                    traffic=KERNEL_TRAFFIC.get((args.workload, name, form)) if not (cfg.diffusion or cfg.lpbl) else None,
                    traffic_unit="bytes per launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated)")

    rf = "fp64" if "fp64" in res else head
    ms_solver, n_solver = res[rf][2][PROF_SOLVE_VER]
    ms_step, n_step = prof[PROF_STEP]
    sb = SOLVER_BYTES[rf]
    solver_bytes = cells * (sb[1] + sb[2]) * (n_solver // 2)
    solver_gbs = solver_bytes / (ms_solver * 1e-3) / 1e9 if ms_solver > 0 else 0.0
    avg_launch_us = 1e3 * ms_solver / max(n_solver, 1)
    step_model = STEP_BYTES["moist" if cfg.lmoist else "dry"]
    form_text = ("mixed form (FP32 stencils and explicit tendencies, FP64 column solves, density updates and budget "
                 "sums), radiation by the MLP surrogate with a 1% solver audit" if head == "mixed" else
                 "plain FP64 form, radiation by the gray two-stream solver")
    if not cfg.lrad:
        form_text = form_text.split(", radiation")[0] + ", radiation off"
    every = max(1, int(round(cfg.dtime_rad / cfg.dt))) if cfg.lrad else 0
    rad_us = {}
    for f, r in res.items():
        ms_rad, n_rad = r[2][PROF_RADIATION]
        rad_us[f] = 1e3 * ms_rad / n_rad if n_rad else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {describe(cfg)}, {form_text}",
            "form": head,
            "diffusion": cfg.diffusion,
            "pbl": cfg.lpbl,
            "soil": cfg.lprog_soil_temp,
            "grid_points": cells,
            "nonfinite_after_run": bad,
            "radiation_updates_timed": (1 + (args.steps - 1) // every) if every else 0,
        },
        "forms": {f: {"value": cells * args.steps / r[0], "ms_per_step": 1e3 * r[0] / args.steps,
                      "solver_ver_us": 1e3 * r[2][PROF_SOLVE_VER][0] / max(r[2][PROF_SOLVE_VER][1], 1),
                      "radiation_update_us": rad_us[f], "radiation": r[3],
                      "kernel_ms_per_step": {k: v[0] for k, v in sorted(r[2]["kernels"].items(), key=lambda kv: -kv[1][0])}}
                  for f, r in res.items()},
        "roofline": kernel_roofline(head),
        "roofline_solver_ver": {
            "kernel": (f"{VER_THREE_PASS if SOLVER_PASSES.get(rf) == 3 else 'k_solve_ver (three_band_solver_ver, FP64 solve)'}"
                       f", both substeps, of the {rf} form"),
            "bound": "hbm",
            "achieved": solver_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": solver_gbs / HBM_PEAK_GBS,
            "traffic": SOLVER_TRAFFIC.get((args.workload, rf)) if not (cfg.diffusion or cfg.lpbl) else None,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated)",
            "algorithmic_bytes_per_point": {"predictor": sb[1], "corrector": sb[2]},
            "avg_launch_us": avg_launch_us,
            "launches_timed": n_solver,
        },
        "roofline_fp64": kernel_roofline("fp64") if "fp64" in res else None,
        "step_roofline": {
            "form": "fp64",
            "bytes_per_gp_step_model": step_model,
            "bytes_per_gp_step_measured": STEP_TRAFFIC.get(args.workload) if not (cfg.diffusion or cfg.lpbl) else None,
            "achieved_gbs_model": step_model * cells * args.steps / res["fp64"][0] / 1e9 if "fp64" in res else None,
            "frac_model": step_model * cells * args.steps / res["fp64"][0] / 1e9 / HBM_PEAK_GBS if "fp64" in res else None,
            "event_ms_per_step": ms_step / max(n_step, 1),
        },
    }
    if "fp64" in res and "mixed" in res:
        out["speedup_mixed_vs_fp64"] = res["fp64"][0] / res["mixed"][0]
    if args.workload == "moist512" and not (args.physics or args.diffusion or args.pbl or args.no_physics_extra):
        # the same workload with the reference's default namelist (diffusion, PBL, 5-layer soil,
        # surface sensible heat flux, lsfc_phase_trans), timed the same way beside the headline
        pcfg = dataclasses.replace(cfg, lprog_soil_temp=True, lsfc_sensible_heat_flux=True, lsfc_phase_trans=True,
                                   nsoillays=5, lpbl=True).set_diffusion(True)
        log(f"setting up {describe(pcfg)}")
        pgrid, pst = make_case(pcfg, cloud_frac=CLOUD_FRAC)
        cloud_patches(pcfg, pst)
        psurface = radiation_surface(pcfg, pgrid, pst)
        phys = {"workload": f"{args.workload}: {describe(pcfg)}"}
        for form in forms:
            el, bad_p, _, rad_p = time_form(pcfg, pgrid, pst, psurface, form, mlp, args.steps, args.warmup)
            phys[form] = {"value": cells * args.steps / el, "ms_per_step": 1e3 * el / args.steps, "nonfinite_after_run": bad_p,
                          "radiation_audit_rmse_k_per_day": rad_p.get("audit_rmse_k_per_day")}
            log(f"{form} with the reference's default physics: {1e3 * el / args.steps:.2f} ms/step")
        if "fp64" in phys and "mixed" in phys:
            phys["speedup_mixed_vs_fp64"] = phys["fp64"]["ms_per_step"] / phys["mixed"]["ms_per_step"]
        out["reference_physics"] = phys
        del pgrid, pst
    if not args.no_cpu:
        nt = host_threads()
        log(f"CPU baseline: oracle on {describe(sample_cfg)} for {args.cpu_steps} steps, {nt} instances on {nt} threads")
        v, secs, errs = cpu_baseline(sample_cfg, args.cpu_steps, forms, mlp, nt)
        out["cpu_baseline"] = {
            "value": v, "unit": "grid-point-steps/s", "cores": nt, "kind": "port",
            "sample": f"oracle/lgame_oracle.c (the C restatement), {nt} independent instances of {describe(sample_cfg)} "
                      f"on {nt} host threads, {args.cpu_steps} steps each, {secs:.1f} s",
            "reference_8_cores": REFERENCE_CPU.get(args.workload),
        }
        out["field_error"] = {"vs": "FP64 oracle (gray two-stream radiation) on the CPU-baseline sample",
                              "steps": args.cpu_steps, **errs}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
