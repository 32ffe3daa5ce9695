"""Multi-iteration FWI on a benchmark workload, through the C ABI.

    python -m fwi_cuda_amd.run_fwi --config C5 --iters 10 [--param lame] [--pcg] [--smooth HFS]

The observed data are modelled from the workload's true model; the inversion
starts from its smoothed start model (C5: the true model smoothed by a 40-cell
Gaussian) and runs simulate_fwi_PSV's iteration (src/cpu/n_simulate_PSV.cpp:249-582):
per-shot gradients preconditioned by the energy weights and summed, the taper,
the parabolic step-length search and the model update.  The unknowns are cp and cs
with rho held fixed (the lam/mu/rho gradient carried to them by the chain rule on the
device), or lam, mu, rho as the reference updates them (--param lame).  The gradient
sign is the GPU twin's (d_kernel_lib_PSV.cu:838-840), the one under which the loop
descends.  One line per iteration: misfit, step, wall time, and the model's
relative L2 distance to the true P and S velocities.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import configs, psv


def velocities(lam, mu, rho):
    """P and S velocities of a Lame model: cp = sqrt((lam + 2 mu) / rho), cs = sqrt(mu / rho)."""
    return np.sqrt((lam + 2.0 * mu) / rho), np.sqrt(mu / rho)


def rel_l2(a, b, window):
    z1, z2, x1, x2 = window
    a, b = a[z1:z2, x1:x2], b[z1:z2, x1:x2]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--config", default="C5", choices=("C1", "C2", "C5"))
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--nt", type=int, default=0, help="shorter time axis")
    ap.add_argument("--nshot", type=int, default=0, help="fewer shots")
    ap.add_argument("--param", choices=("velocity", "lame"), default="velocity")
    ap.add_argument("--pcg", action="store_true", help="conjugate-gradient directions (PCG_PSV)")
    ap.add_argument("--smooth", type=int, default=0, metavar="HFS", help="Gaussian smoothing half width")
    ap.add_argument("--batch", type=int, default=0,
                    help="shots per launch (0: as many as memory allows); fewer shots leave memory for "
                         "kept snapshot units, which cuts recomputation")
    ap.add_argument("--json", default="", help="also write the history as JSON to this file")
    args = ap.parse_args(argv)

    kw = {}
    if args.nt:
        kw["nt"] = args.nt
    if args.nshot:
        kw["nshot"] = args.nshot
    build = configs.BUILDERS[args.config]
    true = build(**kw)
    start = build(true_model=False, **kw) if args.config == "C5" else None
    if start is None:   # C1 / C2: the layered background without the inclusion, as their start model
        from scipy.ndimage import gaussian_filter
        start = {k: gaussian_filter(true[k], 10.0, mode="nearest") for k in ("lam", "mu", "rho")}
    cp_true, cs_true = velocities(true["lam"], true["mu"], true["rho"])
    z1, z2, x1, x2 = (int(v) for v in true["snap"][:4])
    window = (z1, z2 + 1, x1, x2 + 1)
    print(f"{args.config}: {true['nz']}x{true['nx']}, nt {true['nt']}, {true['nshot']} shots", flush=True)

    hist = []
    param = psv._lib.PARAM_VELOCITY if args.param == "velocity" else psv._lib.PARAM_LAME
    with psv.Simulation(true, grad_sign=1, param=param, pcg=int(args.pcg), smooth_hfs=args.smooth,
                        batch=args.batch) as sim:
        t0 = time.perf_counter()
        uz, ux = sim.forward()
        sim.set_observed(uz, ux)
        print(f"observed data: {time.perf_counter() - t0:.1f} s, plan {sim.plan()}", flush=True)
        sim.set_model(start["lam"], start["mu"], start["rho"])
        lam, mu, rho = sim.get_model()
        cp, cs = velocities(lam, mu, rho)
        e0 = (rel_l2(cp, cp_true, window), rel_l2(cs, cs_true, window))
        print(f"start model: cp rel-L2 {e0[0]:.5f}, cs rel-L2 {e0[1]:.5f}", flush=True)
        for k in range(args.iters):
            t0 = time.perf_counter()
            try:
                L2, step = sim.fwi_iteration()
            except psv._lib.FwiError as e:
                print(f"iteration {k}: {e}", flush=True)
                break
            dt = time.perf_counter() - t0
            lam, mu, rho = sim.get_model()
            cp, cs = velocities(lam, mu, rho)
            rec = {"iter": k, "misfit": L2, "step": step, "seconds": round(dt, 2),
                   "cp_rel_l2": rel_l2(cp, cp_true, window), "cs_rel_l2": rel_l2(cs, cs_true, window)}
            hist.append(rec)
            print(f"iteration {k}: misfit {L2:.6e}, step {step:.4e}, {dt:.1f} s, "
                  f"cp rel-L2 {rec['cp_rel_l2']:.5f}, cs rel-L2 {rec['cs_rel_l2']:.5f}", flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"config": args.config, "param": args.param, "pcg": args.pcg, "smooth": args.smooth,
                       "start": {"cp_rel_l2": e0[0], "cs_rel_l2": e0[1]},
                       "iterations": hist}, f, indent=1)
    return 0 if hist else 1


if __name__ == "__main__":
    sys.exit(main())
