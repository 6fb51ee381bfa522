"""FP32-bulk against FP64 on a turbulence workload, every diagnostic output compared.

Default workload (decaying): periodic [0, 2 pi)^3, N = 256, nu = 2e-3, RK4 (the reference's
Crank-Nicolson viscous factor), divergence-free random-phase start with E(k) ~ k^4 exp(-2 (k/4)^2),
1/2 <|u|^2> = 0.5, Philox seed 1, 500 steps.  With --force-eps the forced workload instead (e.g.
the 512^3 configuration: --n 512 --nu 1e-3 --k0 2 --seed 2 --force-eps 0.1 --force-kf 2.5).  Both
runs take the same fixed dt (CFL 0.5 of the start field) so that they are compared at equal
times; both start from the identical FP64 state.  Every `every` steps
the totals (energy, enstrophy, palinstrophy, helicity, dissipation) and E(k) of both runs are
compared; E(k) errors are taken over the shells holding at least 1e-10 of the peak shell's energy
(below that FP32 round-off of the transforms is an absolute floor); helicity, which is near zero,
is compared relative to sqrt(energy x enstrophy).  The state's relative L2
difference is reported too: it grows with time because the flow is chaotic.

Usage (one GPU):  python 3d_navier_stokes_amd/fp32_error.py --out report.json
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from les import load_nsg, log  # noqa: E402

NAMES = ("energy", "enstrophy", "palinstrophy", "helicity", "dissipation")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nu", type=float, default=2e-3)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--k0", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--force-eps", type=float, default=0.0)
    ap.add_argument("--force-kf", type=float, default=2.5)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    nsg = load_nsg()
    N = a.n
    force = dict(force_eps=a.force_eps, force_kf=a.force_kf)
    s64 = nsg.Solver(N, nu=a.nu, dt=1e-3, precision="fp64", cfl=a.cfl, **force)
    s64.init("k4", k0=a.k0, energy=0.5, seed=a.seed)
    s64.step(1)  # one CFL step fixes dt from the start field's max|u|
    _, dt, _ = s64.time()
    s64.close()
    s64 = nsg.Solver(N, nu=a.nu, dt=dt, precision="fp64", **force)
    s32 = nsg.Solver(N, nu=a.nu, dt=dt, precision="fp32bulk", **force)
    s64.init("k4", k0=a.k0, energy=0.5, seed=a.seed)
    s32.set_state(s64.get_state())
    rows = []
    for n in range(0, a.steps + 1, a.every):
        if n:
            s64.step(a.every)
            s32.step(a.every)
        m64, ek64, ens64 = s64.measures()
        m32, ek32, ens32 = s32.measures()
        sel = ek64 >= 1e-10 * ek64.max()
        u64, u32 = s64.get_state(), s32.get_state()
        row = {"step": n, "time": s64.time()[0],
               "state_rel_l2": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
               "E_k_max_rel": float(np.max(np.abs(ek32[sel] - ek64[sel]) / ek64[sel])),
               "Ens_k_max_rel": float(np.max(np.abs(ens32[sel] - ens64[sel]) / ens64[sel])),
               "resolved_shells": int(np.count_nonzero(sel))}
        for name, v64, v32 in zip(NAMES, m64, m32):
            row[name + "_fp64"] = v64
            row[name + "_rel"] = abs(v32 - v64) / abs(v64) if name != "helicity" else abs(v32 - v64) / m64.enstrophy ** 0.5 / m64.energy ** 0.5
        rows.append(row)
        log(f"[fp32] step {n}: state {row['state_rel_l2']:.2e}, E(k) {row['E_k_max_rel']:.2e}, "
            f"energy {row['energy_rel']:.2e}, dissipation {row['dissipation_rel']:.2e}")
    report = {"config": {"N": N, "nu": a.nu, "dt": dt, "steps": a.steps,
                         "ic": f"k4 k0={a.k0:g} seed {a.seed} energy 0.5", "scheme": "rk4cn",
                         "force_eps": a.force_eps, "force_kf": a.force_kf}, "rows": rows}
    line = json.dumps(report)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line)


if __name__ == "__main__":
    main()
