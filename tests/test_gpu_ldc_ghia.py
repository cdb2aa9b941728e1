"""Lid-driven cavity at Re = 1000 on the planar time-step path, against the
Ghia, Ghia & Shin (1982) profiles that the reference's own LDC test
(tests/ldcTest: input/parameters.yaml, checkLDC.py) compares with.

The deck of that test: 128 x 128 cells on the unit square, double-sided tanh
stretching with beta = 1 on both axes, probType 1 (lid: Vx = 1 on the top
wall, no-slip elsewhere, hydro.cc:365-368), Re = 1000 (nu = 1/Re), central
differences (dScheme 1, non-linear term 1), Euler-CN with the CFL time step
(Courant 0.1, dt at most 1e-3) to t = 30, CN tolerance 1e-6, V(3,3) pressure
cycles (7 at most, relative L1 residual 1e-6, coarsest level smoothed).
Profiles as checkLDC.py takes them: u_x at x = 0.5 (mean of the two centre
columns) against z, u_z at z = 0.5 against x, interpolated linearly to
Ghia's points.

A GPU test (about 37 000 steps, under a minute on one MI355X; it prints a
progress line every 20 s); also runnable alone:
    python tests/test_gpu_ldc_ghia.py [--t-final 30]
which prints a final JSON line with the deviations.  Bar: the largest
deviation from Ghia's points at t = 30 is below 0.015 in u_x and 0.02 in
u_z (measured: 0.0070 and 0.0119), the rms deviations below 0.006 and 0.01.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "saras-cc_amd"))
import saras_amd as S  # noqa: E402


def ghia(name):
    return np.loadtxt(os.path.join(ROOT, "tests", "golden", name), comments="#")


def run_ldc(n=128, t_final=30.0, courant=0.1, dt_max=1e-3):
    pad = 1
    g = S.grid_desc((n, 1, n), dim=2, periodic=(0, 0, 0), gtype=(S.GRID_TANH_DOUBLE, 0, S.GRID_TANH_DOUBLE),
                    beta=(1.0, 1.0, 1.0), pad=pad)
    mg = S.mg_params(vc_count=7, vc_tol=1e-6, solve_flag=0, mg_tol=1e-6, pre=3, post=3, res_type=1,
                     bc=S.BC_NEUMANN, precision=S.PREC_FP64)
    D, N = S.FBC_DIRICHLET, S.FBC_NEUMANN
    # walls 0..5 = x-lo, x-hi, y-lo, y-hi, z-lo (bottom), z-hi (lid)
    vbc = [([D] * 6, [0.0, 0.0, 0.0, 0.0, 0.0, 1.0]), ([D] * 6, [0.0] * 6), ([D] * 6, [0.0] * 6)]
    prm = S.nse_params(dscheme=1, nlscheme=1, has_T=0, nu=1.0 / 1000, kappa=1.0 / 1000, dt=dt_max,
                       cn_tol=1e-6, scheme=1, vbc=vbc, pbc=([N] * 6, [0.0] * 6))
    s = S.Nse(g, mg, prm)
    zero = np.zeros(s.shape)
    s.set(vx=zero, vy=zero, vz=zero, p=zero, t=zero, pp=zero)
    t, steps, t0, last = 0.0, 0, time.time(), time.time()
    while t < t_final - 1e-12:
        dt = s.cfl(courant, dt_max)
        r, st = s.step()
        S.check(r)
        t += dt
        steps += 1
        if time.time() - last > 20:
            d = s.diag()
            print(f"t = {t:.3f}  steps {steps}  dt {dt:.3e}  KE {d['ke']:.6e}  max|div| {d['divmax']:.2e}  "
                  f"wall {time.time() - t0:.0f} s", flush=True)
            last = time.time()
    f = s.get()
    s.close()
    x = S.level_metrics(g, 0, 0)["x"][1:-1]
    z = S.level_metrics(g, 0, 2)["x"][1:-1]
    c = slice(pad, -pad)
    U, W = f["vx"][c, 0, c], f["vz"][c, 0, c]
    m = n // 2
    u_prof = (U[m, :] + U[m - 1, :]) / 2       # checkLDC.py: columns Nx/2 and Nx/2 - 1
    w_prof = (W[:, m] + W[:, m - 1]) / 2
    gu, gw = ghia("ghia_ldc_re1000_u.txt"), ghia("ghia_ldc_re1000_w.txt")
    # Ghia's wall points carry the wall values themselves
    zu = np.concatenate(([0.0], z, [1.0]))
    uu = np.concatenate(([0.0], u_prof, [1.0]))
    xw = np.concatenate(([0.0], x, [1.0]))
    ww = np.concatenate(([0.0], w_prof, [0.0]))
    du = np.interp(gu[:, 1], zu, uu) - gu[:, 2]
    dw = np.interp(gw[:, 1], xw, ww) - gw[:, 2]
    out = dict(t=round(t, 6), steps=steps, wall_s=round(time.time() - t0, 1),
               max_abs_dev_u=float(np.max(np.abs(du))), max_abs_dev_w=float(np.max(np.abs(dw))),
               rms_dev_u=float(np.sqrt(np.mean(du ** 2))), rms_dev_w=float(np.sqrt(np.mean(dw ** 2))),
               u_at_ghia=[round(float(v), 5) for v in np.interp(gu[:, 1], zu, uu)],
               w_at_ghia=[round(float(v), 5) for v in np.interp(gw[:, 1], xw, ww)])
    return out


@pytest.mark.gpu
def test_lid_driven_cavity_re1000_matches_ghia():
    out = run_ldc()
    print(json.dumps(out), flush=True)
    assert abs(out["t"] - 30.0) < 1e-3 * 30.0
    assert out["max_abs_dev_u"] < 0.015 and out["max_abs_dev_w"] < 0.02, out
    assert out["rms_dev_u"] < 0.006 and out["rms_dev_w"] < 0.01, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--t-final", type=float, default=30.0)
    ap.add_argument("--courant", type=float, default=0.1)
    ap.add_argument("--dt-max", type=float, default=1e-3)
    args = ap.parse_args()
    print(json.dumps(run_ldc(args.n, args.t_final, args.courant, args.dt_max)), flush=True)


if __name__ == "__main__":
    main()
