"""Benchmark of the MI355X HEVI predictor-corrector step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Runs the step on one GPU with the state resident in HBM, times exactly K steps between two
stream synchronizations, and prints one JSON line (last line of stdout).  Progress goes to stderr.

Workloads (synthetic initial states, see l-game_amd/lgame/cases.py):
  moist512  512x512x60 moist (7 constituents), dx 2 km, dt 3.22 s   -- the operational-size config
  dry128    128x128x50 dry, rotated lat-lon, dx 2 km, dt 3.22 s
  schaer    ny=3, nx=201, 40 layers, dx 500 m, dt 0.805 s
Both forms of the step are timed: the plain FP64 form and the mixed form (FP32 stencils and
explicit tendencies, FP64 column solves, density updates and budget sums).  The headline value is
the mixed form; `forms` lists both and `speedup_mixed_vs_fp64` their ratio.
The roofline object is the implicit vertical solver (three_band_solver_ver, FP64 in both forms),
timed with HIP events on the library's own stream; its algorithmic bytes are 152 B (predictor) and 176 B
(corrector) per grid point (state, tendency and grid arrays streamed once, FP64).
The CPU baseline is the oracle (C restatement, one thread) on a bounded sample of the same
workload type; each GPU form is run on that sample and its per-field RMS and max error against
the oracle are reported in `field_error`.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "l-game_amd"))

HBM_PEAK_GBS = 8000.0
SOLVER_BYTES = {1: 152.0, 2: 176.0}          # per grid point, FP64, predictor / corrector
STEP_BYTES = {"dry": 920.0, "moist": 1296.0}  # per grid-point step, whole-step traffic model
# HBM bytes per k_solve_ver dispatch, averaged over predictor and corrector, from two rocprofv3
# --pmc passes (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0, the factors a 1 GiB device copy gives);
# profiles/moist512_fp64_hbm_traffic.md holds the commands and the per-kernel table.
SOLVER_TRAFFIC = {"moist512": 4.929e9}
METRIC = "grid-point-steps/s at 1 GPU: mixed+surrogate vs FP64 plain, at stated field error"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name):
    from lgame.cases import dry3d_config, moist3d_config, schaer_config

    if name == "moist512":
        return moist3d_config(n=512, n_layers=60, dtime=3.22), moist3d_config(n=65, n_layers=60, dtime=3.22)
    if name == "dry128":
        return dry3d_config(n=128, n_layers=50, dtime=3.22), dry3d_config(n=65, n_layers=50, dtime=3.22)
    if name == "schaer":
        from lgame.cases import schaer_config as sc
        return sc(nx=201, n_layers=40), sc(nx=201, n_layers=40)
    raise SystemExit(f"unknown workload {name}")


def describe(cfg):
    kind = "moist" if cfg.lmoist else "dry"
    return f"{kind} {cfg.ny}x{cfg.nx}x{cfg.n_layers}, dx {cfg.dx:g} m, dt {cfg.dt:g} s, theta order {cfg.theta_adv_order}"


def cpu_baseline(cfg, nsteps, forms):
    """Oracle timing on the host (one thread) plus the GPU forms' per-field error on the same sample."""
    import dataclasses

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from lgame.cases import make_case
    from lgame.config import PREC_FP64, PREC_MIXED
    from lgame.gpu import Step
    from lgame.metrics import field_errors, worst
    from pyoracle import CpuOracle

    grid, st = make_case(cfg)
    o = CpuOracle(cfg, grid)
    o.start(st)
    t0 = time.perf_counter()
    o.run(0, nsteps)
    dt = time.perf_counter() - t0
    want = o.get_state(nsteps % 2)
    errs = {}
    for form in forms:
        g = Step(dataclasses.replace(cfg, precision=PREC_MIXED if form == "mixed" else PREC_FP64), grid)
        g.start(st)
        g.run(0, nsteps)
        got = g.get_state(nsteps % 2)
        g.close()
        e = field_errors(got, want)
        errs[form] = {"max_rel": worst(e, "max_rel"), "rms_rel": worst(e, "rms_rel"),
                      "per_field": {k: [v["rms_rel"], v["max_rel"]] for k, v in e.items()}}
    return cfg.gridpoints * nsteps / dt, dt, errs


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


def time_form(cfg, grid, st, precision, steps, warmup):
    """Throughput of one form of the step plus event timing of the FP64 column solver."""
    import dataclasses

    from lgame.gpu import PROF_SOLVE_GEN, PROF_SOLVE_VER, PROF_STEP, Step

    step = Step(dataclasses.replace(cfg, precision=precision), grid)
    step.start(st)
    counter = 0
    for _ in range(warmup):
        step.pchevi(counter % 2, 1 - counter % 2, counter)
        counter += 1
    step.sync()
    # timed region: exactly K steps between two synchronizations of the library's stream
    t0 = time.perf_counter()
    for _ in range(steps):
        step.pchevi(counter % 2, 1 - counter % 2, counter)
        counter += 1
    step.sync()
    elapsed = time.perf_counter() - t0
    bad = step.count_nonfinite(counter % 2)
    # kernel timing with HIP events on the library's stream (separate pass)
    step.profile(True)
    nprof = max(2, min(steps, 6))
    for _ in range(nprof):
        step.pchevi(counter % 2, 1 - counter % 2, counter)
        counter += 1
    prof = {g: step.profile_read(g) for g in (PROF_SOLVE_VER, PROF_SOLVE_GEN, PROF_STEP)}
    step.profile(False)
    step.close()
    return elapsed, bad, prof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="moist512", choices=["moist512", "dry128", "schaer"])
    ap.add_argument("--form", default="both", choices=["both", "mixed", "fp64"],
                    help="which forms of the step to time (the headline value is the mixed form)")
    ap.add_argument("--cpu-steps", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES (for HBM counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "multi-GPU sharding is a design note, not part of the build", file=sys.stderr)
        return 2

    from lgame.cases import make_case
    from lgame.config import PREC_FP64, PREC_MIXED
    from lgame.gpu import PROF_SOLVE_VER, PROF_STEP

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    cfg, sample_cfg = workload(args.workload)
    t0 = time.perf_counter()
    log(f"setting up {describe(cfg)}")
    grid, st = make_case(cfg)
    log(f"set-up done in {time.perf_counter() - t0:.1f} s")
    cells = cfg.gridpoints
    forms = {"both": ("fp64", "mixed"), "mixed": ("mixed",), "fp64": ("fp64",)}[args.form]
    res = {}
    for form in forms:
        prec = PREC_MIXED if form == "mixed" else PREC_FP64
        elapsed, bad, prof = time_form(cfg, grid, st, prec, args.steps, args.warmup)
        res[form] = (elapsed, bad, prof)
        log(f"{form}: {args.steps} steps in {elapsed:.3f} s ({1e3 * elapsed / args.steps:.2f} ms/step), "
            f"non-finite values: {bad}")
    del grid, st
    head = "mixed" if "mixed" in res else "fp64"
    elapsed, bad, prof = res[head]
    value = cells * args.steps / elapsed
    ms_solver, n_solver = prof[PROF_SOLVE_VER]
    ms_step, n_step = prof[PROF_STEP]
    solver_bytes = cells * (SOLVER_BYTES[1] + SOLVER_BYTES[2]) * (n_solver // 2)
    solver_gbs = solver_bytes / (ms_solver * 1e-3) / 1e9 if ms_solver > 0 else 0.0
    avg_launch_us = 1e3 * ms_solver / max(n_solver, 1)
    step_model = STEP_BYTES["moist" if cfg.lmoist else "dry"]
    form_text = "mixed form (FP32 stencils and explicit tendencies, FP64 column solves, density updates " \
                "and budget sums)" if head == "mixed" else "plain FP64 form"
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
            "workload": f"{args.workload}: {describe(cfg)}, {form_text}, radiation off",
            "form": head,
            "grid_points": cells,
            "nonfinite_after_run": bad,
        },
        "forms": {f: {"value": cells * args.steps / r[0], "ms_per_step": 1e3 * r[0] / args.steps,
                      "solver_ver_us": 1e3 * r[2][PROF_SOLVE_VER][0] / max(r[2][PROF_SOLVE_VER][1], 1)}
                  for f, r in res.items()},
        "roofline": {
            "kernel": "k_solve_ver (three_band_solver_ver, both substeps, FP64)",
            "bound": "hbm",
            "achieved": solver_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": solver_gbs / HBM_PEAK_GBS,
            "traffic": SOLVER_TRAFFIC.get(args.workload),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-calibrated)",
            "algorithmic_bytes_per_point": {"predictor": SOLVER_BYTES[1], "corrector": SOLVER_BYTES[2]},
            "avg_launch_us": avg_launch_us,
            "launches_timed": n_solver,
        },
        "step_roofline": {
            "bytes_per_gp_step_model_fp64": step_model,
            "achieved_gbs_fp64_model": step_model * value / 1e9,
            "event_ms_per_step": ms_step / max(n_step, 1),
        },
    }
    if "fp64" in res and "mixed" in res:
        out["speedup_mixed_vs_fp64"] = res["fp64"][0] / res["mixed"][0]
    if not args.no_cpu:
        log(f"CPU baseline: oracle on {describe(sample_cfg)} for {args.cpu_steps} steps")
        v, secs, errs = cpu_baseline(sample_cfg, args.cpu_steps, forms)
        out["cpu_baseline"] = {
            "value": v, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle/lgame_oracle.c, {describe(sample_cfg)}, {args.cpu_steps} steps, {secs:.1f} s",
        }
        out["field_error"] = {"vs": "FP64 oracle on the CPU-baseline sample", "steps": args.cpu_steps, **errs}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
