"""The reference's own manage_p_grad on two consecutive states (TEST INFRASTRUCTURE ONLY).

    python oracle/ref_pgrad.py '<json case>' <in.npz> <out.npz>

in.npz holds the states s0_<field> and s1_<field> (step 0's old state and step 1's old state).
The compiled mo_p_grad (through ref_harness.F90) runs manage_p_grad(s0, lfirst) and then
manage_p_grad(s1, not lfirst), as pchevi calls it at the first substeps of steps 0 and 1
(mo_manage_pchevi.F90:71-73), each with calc_p_grad_condensates_v; out.npz holds the pressure-
gradient terms after each call (p0_<name>, p1_<name>).  The reference harness keeps Fortran
module state, so it runs in a process of its own.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "l-game_amd"))
sys.path.insert(0, HERE)

from compare_ref import build_case  # noqa: E402
from pyoracle import RefHarness  # noqa: E402

NAMES = ("p_grad_acc_neg_nl_u", "p_grad_acc_neg_nl_v", "p_grad_acc_neg_nl_w", "p_grad_acc_neg_l_u",
         "p_grad_acc_neg_l_v", "p_grad_acc_neg_l_w", "p_grad_acc_old_u", "p_grad_acc_old_v", "p_grad_condensates_w")


def shapes(cfg):
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    u, v, w = (ny, nx + 1, nl), (ny + 1, nx, nl), (ny, nx, nl + 1)
    return {n: {"u": u, "v": v, "w": w}[n[-1]] for n in NAMES}


def main():
    spec = json.loads(sys.argv[1])
    cfg, (g, _) = build_case(spec)
    z = np.load(sys.argv[2], allow_pickle=False)
    r = RefHarness(cfg, g)
    out = {}
    shp = shapes(cfg)
    for step in (0, 1):
        st = {k[3:]: z[k] for k in z.files if k.startswith(f"s{step}_")}
        r.set_state(0, st)
        r.manage_p_grad(0, lfirst=(step == 0))
        for n in NAMES:
            out[f"p{step}_{n}"] = r.get_diag(n, shp[n])
    np.savez(sys.argv[3], **out)


if __name__ == "__main__":
    main()
