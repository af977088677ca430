"""Checks that pin the oracle's restated glue by what the reference program or the mathematics
fixes, not by a second transcription of the same formulas.

The glue in question (oracle/lgame_oracle.c; mo_vector_tend_expl, mo_momentum_diff_diss and
mo_manage_pchevi do not build here, see oracle/ref_harness.F90):
* the step sequence of pchevi (mo_manage_pchevi.F90:27-118) and the tendency combination of
  vector_tend_expl (mo_vector_tend_expl.F90:58-180): the budgets of the reference program's own
  Schaer run (the whole unmodified control.F90, recorded from a flang build of the reference), the
  pressure-gradient extrapolation weights derived from c_v/c_p and applied to the compiled
  manage_p_grad, the inertial oscillation of a uniform flow on an f-plane (the predictor-corrector
  recurrence of the averaged, lagged Coriolis term), and the energy neutrality of the vorticity
  flux with the mass fluxes vector_tend_expl forms;
* mom_diff_v (mo_momentum_diff_diss.F90:162-331): the second derivative of a quadratic w(z);
* simple_dissipation_rate (mo_momentum_diff_diss.F90:333-388): the dissipation of a shear flow
  and of the Klemp layer, in closed form.
"""
import dataclasses
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from lgame.cases import make_case, schaer_config, tiny_config
from pyoracle import REF_LIB, CpuOracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# mo_constants.F90:24-28
C_D_V, C_D_P = 717.942189, 1005.0


# ---- the reference program's Schaer run --------------------------------------------------------

def test_oracle_schaer_run_matches_reference_program_budgets():
    """The Schaer case on the 201 x 3 x 40 plane for 1,119 steps of 0.805 s (a 10-minute span plus
    the time loop's 300 s, control.F90:507), as the whole reference program (control.F90 with
    every step module, built with flang) ran it: the dry mass went from 1492336736843.482 to
    1492336736843.394, the rhotheta_v integral drifted +2.7e-9 and the total energy -2.3e-8
    relative (write_out_integrals, mo_write_out.F90:226-304, 8 OpenMP threads).  The oracle must
    reproduce these to their printed precision; the last digit of the final mass may differ by the
    summation order of the threaded sums (a few units of 1e-3 in 1.5e12).  A mistake in the step
    sequence or the tendency combination moves them far outside (see the design notes of the
    oracle)."""
    cfg = schaer_config(nx=201, n_layers=40)
    grid, st = make_case(cfg)
    o = CpuOracle(cfg, grid)
    o.start(st)
    o.lib().orc_diagnose_temperature(o.h, 0)
    i0 = o.integrals(0)
    o.run(0, 1119)
    o.lib().orc_diagnose_temperature(o.h, 1)
    i1 = o.integrals(1)
    e0, e1 = i0[1] + i0[2] + i0[3], i1[1] + i1[2] + i1[3]
    print(f"mass {i0[4]!r} -> {i1[4]!r}, rhotheta drift {(i1[0] - i0[0]) / i0[0]:.4e}, "
          f"energy drift {(e1 - e0) / e0:.4e}")
    assert abs(i0[4] - 1492336736843.482) < 1e-3
    assert abs(i1[4] - 1492336736843.394) < 5e-3
    assert abs((i1[0] - i0[0]) / i0[0] - 2.7e-9) < 0.05e-9
    assert abs((e1 - e0) / e0 - (-2.3e-8)) < 0.05e-8


# ---- pressure-gradient extrapolation weights ---------------------------------------------------

PGRAD_CASES = [
    {"kind": "tiny", "kw": {"llinear": True, "corio": False}},
    {"kind": "tiny", "kw": {"llinear": True, "corio": False, "order": 3}},
    {"kind": "tiny", "kw": {"llinear": True, "corio": False, "moist": True}, "cloud_frac": 0.3},
]


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference build oracle/_ref/libref.so not present")
@pytest.mark.parametrize("case", PGRAD_CASES, ids=lambda c: json.dumps(c["kw"], sort_keys=True))
def test_oracle_wind_tendency_is_the_extrapolated_pressure_gradient(case, tmp_path):
    """Without momentum advection (llinear, no Coriolis) and without diffusion, the wind
    tendencies of a step are the pressure-gradient terms alone.  The weights follow from
    c_v/c_p (mo_constants.F90): the horizontal gradient of the current step enters with
    1/2 + c_v/c_p, the one of the previous step with 1 - (1/2 + c_v/c_p) (at step 0 the previous
    one is the current one), the vertical gradient with 1 - c_v/c_p, the condensate term with 1.
    The gradients come from the reference's compiled manage_p_grad and
    calc_p_grad_condensates_v on the states the oracle's steps 0 and 1 start from; the averaged
    tendency of the predictor and corrector equals that combination bit for bit (both substeps
    use the gradient of the first, and 0.5 x + 0.5 x = x)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from compare_ref import build_case

    cfg, (g, st) = build_case(case)
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    o = CpuOracle(cfg, g)
    o.start(st)
    states, tends = [], []
    for k, (a, b) in enumerate(((0, 1), (1, 0))):
        states.append(o.get_state(a))
        o.step(a, b, k)
        tends.append({c: o.get_diag(f"tend_wind_{c}", s) for c, s in
                      (("u", (ny, nx + 1, nl)), ("v", (ny + 1, nx, nl)), ("w", (ny, nx, nl + 1)))})
    inp = tmp_path / "states.npz"
    np.savez(inp, **{f"s{k}_{f}": v for k, s in enumerate(states) for f, v in s.items()})
    out = tmp_path / "pgrad.npz"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "ref_pgrad.py"), json.dumps(case), str(inp),
                        str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    p = np.load(out, allow_pickle=False)
    cur_h = 0.5 + C_D_V / C_D_P
    old_h = 1.0 - cur_h
    cur_v = 1.0 - C_D_V / C_D_P
    for k in (0, 1):
        q = {n[3:]: p[n] for n in p.files if n.startswith(f"p{k}_")}
        prev = q if k == 0 else {n[3:]: p[n] for n in p.files if n.startswith("p0_")}
        for c in ("u", "v"):
            # the previous gradient as the reference stores it: -nl - 1*l (use_bg_switch = 1)
            old = -prev[f"p_grad_acc_neg_nl_{c}"] - 1.0 * prev[f"p_grad_acc_neg_l_{c}"]
            assert np.array_equal(q[f"p_grad_acc_old_{c}"], old)
            want = old_h * old - cur_h * (q[f"p_grad_acc_neg_nl_{c}"] + q[f"p_grad_acc_neg_l_{c}"])
            assert np.array_equal(tends[k][c], want), (k, c, np.abs(tends[k][c] - want).max())
            assert np.abs(want).max() > 0.0
        want_w = -cur_v * (q["p_grad_acc_neg_nl_w"] + q["p_grad_acc_neg_l_w"]) + q["p_grad_condensates_w"]
        assert np.array_equal(tends[k]["w"], want_w), (k, np.abs(tends[k]["w"] - want_w).max())


# ---- inertial oscillation ----------------------------------------------------------------------

def test_oracle_inertial_oscillation_follows_the_predictor_corrector_recurrence():
    """A horizontally uniform wind (u0, v0) over flat ground on an f-plane (f constant, no
    horizontal Coriolis components), linear momentum equation: the pressure gradient, the kinetic
    energy gradient and the horizontal vorticities vanish, and the TRSK flux of the uniform mass
    flux with the potential vorticity f/rho is the Coriolis acceleration (f v, -f u).  With
    z = u + i v the wind then follows the step's time scheme applied to dz/dt = -i f z:
    at the first substep the tendency is the term of the previous corrector (lagged; at step 0
    formed from the old state), the predictor is z + dt T1; at the second substep the term is
    formed from the predictor and averaged with T1 (weights 1/2, 1/2), and z advances by dt times
    that average (mo_vector_tend_expl.F90:58 and 130-165, mo_manage_pchevi.F90:63-111).  A
    different weight, an unlagged or unaveraged term, or a term formed from the old state at the
    second substep changes z by O((f dt)^2) per step."""
    f0, u0, v0, n = 1.3e-3, 7.0, -3.0, 40
    cfg = dataclasses.replace(tiny_config(ny=5, nx=7, n_layers=6, corio=True, plane=True, llinear=True),
                              orography_id=0, hills=())
    g, st = make_case(cfg, perturbation=False)
    g["fvec_z"][:] = f0
    g["fvec_x"][:] = 0.0
    g["fvec_y"][:] = 0.0
    st["wind_u"][:] = u0
    st["wind_v"][:] = v0
    o = CpuOracle(cfg, g)
    o.start(st)
    o.run(0, n)
    out = o.get_state(n % 2)
    dt = cfg.dt
    z = complex(u0, v0)
    lag = None
    for k in range(n):
        t1 = -1j * f0 * z if k == 0 else lag
        pred = z + dt * t1
        t2 = -1j * f0 * pred
        z = z + dt * (0.5 * t1 + 0.5 * t2)
        lag = t2
    u, v = out["wind_u"], out["wind_v"]
    print(f"after {n} steps: u {u.mean():.15f} v {v.mean():.15f}, recurrence {z.real:.15f} {z.imag:.15f}")
    assert np.ptp(u) <= 1e-13 * abs(z) and np.ptp(v) <= 1e-13 * abs(z)
    np.testing.assert_allclose(u, z.real, rtol=0, atol=1e-12 * abs(z))
    np.testing.assert_allclose(v, z.imag, rtol=0, atol=1e-12 * abs(z))
    # the flow has turned through a measurable angle
    assert abs(np.angle(z / complex(u0, v0))) > 0.05


# ---- energy neutrality of the vorticity flux ---------------------------------------------------

@pytest.mark.parametrize("freeslip", [False, True])
def test_oracle_vorticity_flux_does_no_work(freeslip):
    """On a flat periodic plane the vorticity-flux term (calc_vorticity_flux_term with the
    potential vorticity and the mass fluxes that vector_tend_expl forms, mo_vector_tend_expl.F90:
    60-90) does no work: the volume integral of the inner product of the mass flux with the
    term, on the kinetic-energy inner-product weights, vanishes for any flow.  Flipping the sign
    or the stencil of any of its terms breaks this."""
    cfg = dataclasses.replace(tiny_config(ny=8, nx=9, n_layers=6, plane=True, corio=True, freeslip=freeslip),
                              orography_id=0, hills=())
    g, st = make_case(cfg)
    rng = np.random.default_rng(5)
    for k in ("wind_u", "wind_v", "wind_w"):
        st[k] = rng.standard_normal(st[k].shape)
    st["wind_u"][:, -1] = st["wind_u"][:, 0]
    st["wind_v"][-1] = st["wind_v"][0]
    st["wind_w"][:, :, 0] = 0.0
    st["wind_w"][:, :, -1] = 0.0  # flat ground: the kinematic surface condition (w_free_slip) gives 0
    o = CpuOracle(cfg, g)
    o.set_state(0, st)
    o.call("vector_tend_expl", 0, rk_step=2, counter=1)
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    su, sv, sw = (ny, nx + 1, nl), (ny + 1, nx, nl), (ny, nx, nl + 1)
    px, py, pz = o.get_diag("pot_vort_tend_x", su), o.get_diag("pot_vort_tend_y", sv), o.get_diag("pot_vort_tend_z", sw)
    fu, fv, fw = o.get_diag("u_placeholder", su), o.get_diag("v_placeholder", sv), o.get_diag("w_placeholder", sw)
    ipw, vol = g["inner_product_weights"], g["volume"]

    def inner(au, av, aw, bu, bv, bw):
        return (ipw[0] * au[:, 1:] * bu[:, 1:] + ipw[1] * av[:ny] * bv[:ny] + ipw[2] * au[:, :nx] * bu[:, :nx] +
                ipw[3] * av[1:] * bv[1:] + ipw[4] * aw[:, :, :nl] * bw[:, :, :nl] + ipw[5] * aw[:, :, 1:] * bw[:, :, 1:])

    work = float(np.sum(vol * inner(fu, fv, fw, px, py, pz)))
    scale = float(np.sum(vol * inner(np.abs(fu), np.abs(fv), np.abs(fw), np.abs(px), np.abs(py), np.abs(pz))))
    print(f"work {work:.3e}, scale {scale:.3e}")
    assert scale > 0.0 and abs(work) <= 1e-14 * scale
    # the horizontal part alone is not neutral: the vertical terms are needed
    hor = float(np.sum(vol * (ipw[0] * fu[:, 1:] * px[:, 1:] + ipw[2] * fu[:, :nx] * px[:, :nx] +
                              ipw[1] * fv[:ny] * py[:ny] + ipw[3] * fv[1:] * py[1:])))
    assert abs(hor) > 1e-6 * scale


# ---- mom_diff_v ----------------------------------------------------------------------------------

def _uniform_column_case(nl=8):
    """A flat periodic plane with equal layers (stretching 1): z_w(l) = toa (1 - (l-1)/nl)."""
    cfg = dataclasses.replace(tiny_config(ny=5, nx=6, n_layers=nl, plane=True, corio=False), orography_id=0,
                              hills=(), stretching_parameter=1.0, lmom_diff_v=True)
    g, st = make_case(cfg, perturbation=False)
    dz = cfg.toa / nl
    zw = g["z_w"][0, 0, :]
    assert np.allclose(np.diff(zw), -dz, rtol=1e-12)
    return cfg, g, st, dz


def test_oracle_vertical_momentum_diffusion_of_a_quadratic_profile():
    """mom_diff_v with w = a z^2, a uniform density rho0 and a uniform viscosity nu (molecular
    viscosity nu, no TKE): the diffusive flux rho nu dw/dz divided by the density differentiates
    to nu d2w/dz2 = 2 a nu at every level whose two adjacent layers are interior (the vertical
    divergence of a layer next to the top or the surface sees a zero boundary flux).  The
    horizontal wind is not diffused (vert_hor_viscosity_u/v stay zero in the reference:
    vert_hor_mom_viscosity is never called)."""
    cfg, g, st, dz = _uniform_column_case()
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    rho0, nu, a = 1.1, 0.37, 3.0e-7
    st["rho"][:] = rho0
    st["wind_w"][:] = a * g["z_w"] ** 2
    o = CpuOracle(cfg, g)
    o.set_state(0, st)
    o.set_diag("viscosity_molecular", np.full((ny, nx, nl), nu))
    o.set_diag("tke", np.zeros((ny, nx, nl)))
    o.set_diag("n_squared", np.full((ny, nx, nl), 1e-4))
    mark = np.random.default_rng(3).standard_normal((ny, nx + 1, nl))
    o.set_diag("mom_diff_tend_x", mark)
    o.call("mom_diff_v", 0)
    tz = o.get_diag("mom_diff_tend_z", (ny, nx, nl + 1))
    inner = tz[:, :, 2:nl - 1]  # Fortran levels 3..nl-1
    print(f"2 a nu = {2 * a * nu:.6e}, got {inner.min():.6e} .. {inner.max():.6e}")
    np.testing.assert_allclose(inner, 2.0 * a * nu, rtol=1e-9)
    assert np.array_equal(o.get_diag("mom_diff_tend_x", (ny, nx + 1, nl)), mark)


# ---- simple_dissipation_rate -------------------------------------------------------------------

def test_oracle_dissipation_of_a_shear_flow():
    """simple_dissipation_rate for u = A sin(2 pi y/Ly) over flat ground with the TKE viscosity
    scheme (uniform viscosity per layer): the momentum diffusion tendency is
    -visc (4/dy^2) sin^2(pi/ny) u, and the heating is minus the density times the inner product
    of the wind with it, rho visc (4/dy^2) sin^2(pi/ny) u^2 (the two u faces of a flat cell carry
    weight 1/2 each), non-negative everywhere."""
    cfg = dataclasses.replace(tiny_config(ny=12, nx=10, n_layers=6, plane=True, corio=False), orography_id=0, hills=(),
                              dx=2000.0, dy=1500.0, diff_coeff_scheme_h="tke").set_diffusion()
    g, st = make_case(cfg, perturbation=False)
    prof = 3.0 * np.sin(2.0 * np.pi * np.arange(cfg.ny) / cfg.ny)
    st["wind_u"][:] = prof[:, None, None]
    o = CpuOracle(cfg, g)
    o.start(st)
    o.run(0, 1)
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    visc = o.get_diag("viscosity", (ny, nx, nl))
    heat = o.get_diag("heating_diss", (ny, nx, nl))
    rho = st["rho"][:, :, :, 0]
    want = rho * visc * 4.0 / cfg.dy ** 2 * np.sin(np.pi / ny) ** 2 * st["wind_u"][:, :nx] ** 2
    print(f"heating {heat.min():.4e} .. {heat.max():.4e}")
    assert heat.max() > 0.0 and heat.min() >= -1e-12 * heat.max()
    np.testing.assert_allclose(heat, want, rtol=1e-9, atol=1e-12 * heat.max())


def test_oracle_dissipation_of_the_klemp_layer():
    """simple_dissipation_rate with only the Klemp layer (mo_momentum_diff_diss.F90:354-371): for
    w = w0 at every level, at rest otherwise, the damping acceleration at level l is
    (w0/p(l) - w0)/dt with p(z) = 1 + dt klemp_damp_max sin^2(pi/2 (z - z0)/(toa - z0)) above
    z0 = klemp_begin_rel toa, and the heating of a layer is minus the density times the inner
    product: -rho w0^2 (1/2 (1/p(top) - 1) + 1/2 (1/p(bottom) - 1))/dt, positive inside the layer
    and zero below it."""
    cfg, g, st, dz = _uniform_column_case(nl=10)
    cfg = dataclasses.replace(cfg, lmom_diff_v=False)
    ny, nx, nl = cfg.ny, cfg.nx, cfg.n_layers
    w0, dt = 0.8, cfg.dt
    st["wind_w"][:] = w0
    o = CpuOracle(cfg, g)
    o.set_state(0, st)
    o.set_diag("mom_diff_tend_z", np.zeros((ny, nx, nl + 1)))
    o.call("simple_dissipation_rate", 0)
    heat = o.get_diag("heating_diss", (ny, nx, nl))
    z0 = cfg.klemp_begin_rel * cfg.toa
    zw = g["z_w"]
    s = np.sin(0.5 * np.pi * np.clip(zw - z0, 0.0, None) / (cfg.toa - z0))
    p = 1.0 + dt * cfg.klemp_damp_max * s * s
    acc = (w0 / p - w0) / dt
    acc[:, :, 0] = 0.0  # the reference damps levels 2..n_damping_levels
    nd = o.n_damping_levels
    acc[:, :, nd:] = 0.0
    want = -st["rho"][:, :, :, 0] * (0.5 * w0 * acc[:, :, :nl] + 0.5 * w0 * acc[:, :, 1:])
    print(f"damping levels {nd}, heating {heat.max():.4e} at the top layers")
    assert heat.max() > 0.0 and heat.min() >= 0.0
    np.testing.assert_allclose(heat, want, rtol=1e-12, atol=1e-15 * heat.max())
    assert np.all(heat[:, :, -2:] == 0.0)
