"""Radiation surrogate: harvesting training columns from the model's radiation calls, fitting
the normalisation, training the BF16 MLP on the GPU and evaluating it against the solver.

The solver (LGAME_RAD_TEACHER, the gray two-stream scheme in the place of RTE+RRTMGP) labels
the columns; the network maps the 5x60+4 column inputs to 60 heating rates [K/day] and the two
surface fluxes (include/lgame_gpu.h, LGAME_RAD_N_IN / LGAME_RAD_N_OUT).
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

from .cases import make_case, moist3d_config, radiation_surface
from .config import RAD_TEACHER
from .gpu import Mlp, Step

N_IN, N_OUT, LEVELS = 5 * 60 + 4, 60 + 2, 60
HIDDEN, N_HIDDEN = 512, 3
T_EPOCH = 1576843200.0
# the committed weights of the 304 -> 3x512 -> 62 network
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data", "rad_surrogate.npz")


def harvest(n=33, n_states=4, calls=2, times=6, steps_between=4, seed=0, cloud_max=0.01, progress=None,
            horizon_steps=0):
    """Columns (x, y) from runs of the moist n x n x 60 case with the solver as radiation.

    Each run starts from a seeded perturbed state whose cloud condensates, surface albedo and
    surface temperature vary from column to column (smooth random fields), and steps with
    radiation updates every `steps_between` steps.  At every update the inputs and solver outputs
    of all columns are taken for the update's own time and for `times - 1` further random times
    of day (the solver is a function of the state and the sun's position).  With horizon_steps,
    the `calls` updates are spread over that many steps of the run, so that the columns also
    come from the developed flow (stronger vertical motion, displaced temperature and moisture)
    that a long run reaches.
    """
    if horizon_steps > 0 and calls > 1:
        steps_between = max(steps_between, horizon_steps // (calls - 1))
    from .init_state import _smooth_periodic

    rng = np.random.default_rng(seed)
    xs, ys = [], []
    dropped = [0]
    for k in range(n_states):
        cfg = moist3d_config(n=n, n_layers=LEVELS, dtime=3.22, seed=int(rng.integers(1, 10 ** 6)))
        cfg = dataclasses.replace(cfg, lrad=True, rad_mode=RAD_TEACHER, dtime_rad=steps_between * cfg.dt)
        g, st = make_case(cfg, seed=int(rng.integers(1, 10 ** 6)), cloud_frac=float(rng.uniform(0.2, 1.0)) * cloud_max)
        cover = np.clip(_smooth_periodic(rng.standard_normal((n, n)), 2.0) * 0.6 + float(rng.uniform(-0.5, 0.5)), 0.0, 1.0)
        for jc in range(5):  # condensates
            st["rho"][:, :, :, jc] *= cover[:, :, None]
        sf = radiation_surface(cfg, g, st, seed=int(rng.integers(1, 10 ** 6)))
        alb = 0.05 + 0.45 * (0.5 + 0.5 * np.tanh(_smooth_periodic(rng.standard_normal((n, n)), 3.0)))
        sf["sfc_albedo"] = np.asfortranarray(alb)
        sf["temp_sfc"] = np.asfortranarray(sf["temp_sfc"] + 3.0 * _smooth_periodic(rng.standard_normal((n, n)), 3.0) +
                                           float(rng.uniform(-3.0, 6.0)))
        t0 = T_EPOCH + 3600.0 * float(rng.uniform(0.0, 24.0))
        s = Step(cfg, g)
        s.start(st)
        s.set_rad_surface(sf)
        step = 0
        for c in range(calls):
            slot = step % 2
            # the update at this step uses state_old = this slot at t_0 + dtime_rad/2
            t_call = t0 + step * cfg.dt + 0.5 * cfg.dtime_rad
            for q in range(times):
                t = t_call if q == 0 else T_EPOCH + 86400.0 * float(rng.uniform(0.0, 365.0))
                x, y = s.rad_columns(slot, t)
                xs.append(x)
                ys.append(y)
            s.run(step, steps_between, t_init=t0)
            step += steps_between
            if c + 1 < calls and s.count_nonfinite(step % 2):  # a run that went unstable gives no more columns
                dropped[0] += 1
                break
        s.close()
        if progress and (k + 1) % 16 == 0:
            progress(f"harvested {k + 1} of {n_states} states")
    x, y = np.concatenate(xs), np.concatenate(ys)
    ok = np.all(np.isfinite(x), axis=1) & np.all(np.isfinite(y), axis=1)
    if progress and (dropped[0] or not ok.all()):
        progress(f"harvest: {dropped[0]} runs stopped early (non-finite state), {int((~ok).sum())} columns dropped")
    return x[ok], y[ok]


def fit_norm(x, y):
    xm, xsd = x.mean(axis=0), x.std(axis=0)
    ym, ysd = y.mean(axis=0), y.std(axis=0)
    xsd = np.where(xsd > 1e-6 * np.maximum(np.abs(xm), 1e-30), xsd, np.maximum(np.abs(xm), 1.0))
    ysd = np.where(ysd > 1e-12, ysd, 1.0)
    return xm.astype(np.float32), xsd.astype(np.float32), ym.astype(np.float32), ysd.astype(np.float32)


def train(x, y, epochs=30, batch=512, lr=1e-3, seed=0, hidden=HIDDEN, n_hidden=N_HIDDEN, norm=None):
    m = Mlp(x.shape[1], hidden, n_hidden, y.shape[1], seed=seed)
    norm = norm if norm is not None else fit_norm(x, y)
    m.set_norm(*norm)
    m.load_data(x, y)
    loss, seconds = m.train(epochs, batch, lr, seed)
    return m, loss, seconds


def heating_rmse(pred, y, levels=LEVELS):
    """RMSE of the heating rates [K/day] over all columns and layers."""
    d = pred[:, :levels].astype(np.float64) - y[:, :levels]
    return float(np.sqrt(np.mean(d * d)))


def baseline_rmse(y_train, y_test, levels=LEVELS):
    """RMSE of the method without a network: the mean heating profile of the training columns."""
    return heating_rmse(np.broadcast_to(y_train.mean(axis=0), y_test.shape), y_test, levels)


def save(path, m: Mlp, norm):
    np.savez(path, params=m.get_params(), x_mean=norm[0], x_scale=norm[1], y_mean=norm[2], y_scale=norm[3],
             n_in=m.n_in, n_out=m.n_out, hidden=m.hidden, n_hidden=m.n_hidden_layers)


def load(path) -> Mlp:
    d = np.load(path, allow_pickle=False)
    m = Mlp(int(d["n_in"]), int(d["hidden"]), int(d["n_hidden"]), int(d["n_out"]))
    m.set_params(d["params"])
    m.set_norm(d["x_mean"], d["x_scale"], d["y_mean"], d["y_scale"])
    return m


def build(states=24, calls=2, times=6, epochs=60, batch=8192, lr=1e-3, seed=0, test_states=3, progress=None,
          horizon_steps=0, hidden=HIDDEN, n_hidden=N_HIDDEN):
    """Harvest, train and evaluate; returns (mlp, norm, report).  The test columns come from
    separate seeded runs over the same horizon."""
    import time

    t0 = time.perf_counter()
    xtr, ytr = harvest(n=33, n_states=states, calls=calls, times=times, seed=seed, progress=progress,
                       horizon_steps=horizon_steps)
    xte, yte = harvest(n=33, n_states=test_states, calls=calls if horizon_steps else 1, times=times, seed=seed + 1000,
                       horizon_steps=horizon_steps)
    t_harvest = time.perf_counter() - t0
    norm = fit_norm(xtr, ytr)
    m, loss, seconds = train(xtr, ytr, epochs=epochs, batch=batch, lr=lr, seed=seed, norm=norm, hidden=hidden,
                             n_hidden=n_hidden)
    pred = m.forward(xte)
    n_params = m.n_params
    flops = 6.0 * n_params * len(xtr) * epochs  # forward 2, backward 4 flops per parameter and sample
    report = {
        "train_columns": int(len(xtr)), "test_columns": int(len(xte)), "epochs": epochs, "batch": batch, "lr": lr,
        "harvest_calls_per_run": calls, "harvest_horizon_steps": horizon_steps,
        "harvest_s": t_harvest, "train_s": seconds, "samples_per_s": len(xtr) * epochs / seconds,
        "train_tflops": flops / seconds / 1e12, "params": n_params,
        "loss_first_epoch": float(loss[0]), "loss_last_epoch": float(loss[-1]),
        "heating_rmse_k_per_day": heating_rmse(pred, yte),
        "baseline_mean_profile_rmse_k_per_day": baseline_rmse(ytr, yte),
        "no_radiation_rmse_k_per_day": heating_rmse(np.zeros_like(yte), yte),
        "sw_in_rmse_w_m2": float(np.sqrt(np.mean((pred[:, LEVELS] - yte[:, LEVELS]) ** 2))),
        "lw_out_rmse_w_m2": float(np.sqrt(np.mean((pred[:, LEVELS + 1] - yte[:, LEVELS + 1]) ** 2))),
    }
    return m, norm, report


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="train the radiation surrogate on the GPU")
    ap.add_argument("--out", required=True)
    ap.add_argument("--states", type=int, default=24)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    m, norm, rep = build(states=a.states, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
    save(a.out, m, norm)
    print(json.dumps(rep), flush=True)
