"""Benchmark of the MI355X HEVI predictor-corrector step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Runs the step on one GPU with the state resident in HBM, times exactly K steps between two
stream synchronizations, and prints one JSON line (last line of stdout).  Progress goes to stderr.

Workloads (synthetic initial states, see l-game_amd/lgame/cases.py):
  moist512  512x512x60 moist (7 constituents), dx 2 km, dt 3.22 s   -- the operational-size config
  dry128    128x128x50 dry, rotated lat-lon, dx 2 km, dt 3.22 s
  schaer    ny=3, nx=201, 40 layers, dx 500 m, dt 0.805 s
The roofline object is the implicit vertical solver (three_band_solver_ver), timed with HIP
events on the library's own stream; its algorithmic bytes are 152 B (predictor) and 176 B
(corrector) per grid point (state, tendency and grid arrays streamed once, FP64).
The CPU baseline is the oracle (C restatement, one thread) on a bounded sample of the same
workload type; the GPU result on that sample is compared with it and the deviation reported.
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


def cpu_baseline(cfg, nsteps):
    """Oracle timing on the host (one thread) plus the GPU deviation on the same sample."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from lgame.cases import make_case
    from lgame.gpu import Step
    from pyoracle import CpuOracle

    grid, st = make_case(cfg)
    o = CpuOracle(cfg, grid)
    o.start(st)
    t0 = time.perf_counter()
    o.run(0, nsteps)
    dt = time.perf_counter() - t0
    want = o.get_state(nsteps % 2)
    g = Step(cfg, grid)
    g.start(st)
    g.run(0, nsteps)
    got = g.get_state(nsteps % 2)
    g.close()
    worst = 0.0
    for k in want:
        scale = max(float(np.abs(want[k]).max()), 1e-300)
        worst = max(worst, float(np.abs(got[k] - want[k]).max()) / scale)
    return cfg.gridpoints * nsteps / dt, dt, worst


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="moist512", choices=["moist512", "dry128", "schaer"])
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
    from lgame.gpu import PROF_SOLVE_VER, PROF_STEP, Step

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    cfg, sample_cfg = workload(args.workload)
    t0 = time.perf_counter()
    log(f"setting up {describe(cfg)}")
    grid, st = make_case(cfg)
    log(f"set-up done in {time.perf_counter() - t0:.1f} s; uploading")
    step = Step(cfg, grid)
    step.start(st)
    del grid, st
    counter = 0
    for _ in range(args.warmup):
        step.pchevi(counter % 2, 1 - counter % 2, counter)
        counter += 1
    step.sync()
    log(f"warm-up of {args.warmup} steps done")

    # timed region: exactly K steps between two synchronizations of the library's stream
    step.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step.pchevi(counter % 2, 1 - counter % 2, counter)
        counter += 1
    step.sync()
    elapsed = time.perf_counter() - t0
    bad = step.count_nonfinite(counter % 2)
    log(f"timed {args.steps} steps in {elapsed:.3f} s; non-finite values: {bad}")

    # kernel timing of the solver with HIP events on the library's stream (separate pass)
    step.profile(True)
    nprof = max(2, min(args.steps, 6))
    for _ in range(nprof):
        step.pchevi(counter % 2, 1 - counter % 2, counter)
        counter += 1
    ms_solver, n_solver = step.profile_read(PROF_SOLVE_VER)
    ms_step, n_step = step.profile_read(PROF_STEP)
    step.profile(False)
    step.close()
    cells = cfg.gridpoints
    solver_bytes = cells * (SOLVER_BYTES[1] + SOLVER_BYTES[2]) * (n_solver // 2)
    solver_gbs = solver_bytes / (ms_solver * 1e-3) / 1e9 if ms_solver > 0 else 0.0
    avg_launch_us = 1e3 * ms_solver / max(n_solver, 1)

    value = cells * args.steps / elapsed
    step_model = STEP_BYTES["moist" if cfg.lmoist else "dry"]
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
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {describe(cfg)}, FP64 plain form, radiation off",
            "form": "plain FP64",
            "grid_points": cells,
            "nonfinite_after_run": bad,
        },
        "roofline": {
            "kernel": "k_solve_ver (three_band_solver_ver, both substeps)",
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
            "bytes_per_gp_step_model": step_model,
            "achieved_gbs": step_model * value / 1e9,
            "frac": step_model * value / 1e9 / HBM_PEAK_GBS,
            "event_ms_per_step": ms_step / max(n_step, 1),
        },
    }
    if not args.no_cpu:
        log(f"CPU baseline: oracle on {describe(sample_cfg)} for {args.cpu_steps} steps")
        v, secs, dev = cpu_baseline(sample_cfg, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": v, "unit": "grid-point-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle/lgame_oracle.c, {describe(sample_cfg)}, {args.cpu_steps} steps, {secs:.1f} s",
        }
        out["field_error"] = {"vs": "FP64 oracle on the CPU-baseline sample", "max_rel": dev,
                              "steps": args.cpu_steps}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
