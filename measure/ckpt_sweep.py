"""Checkpoint-budget sweep of one gradient, through the C ABI.

    python measure/ckpt_sweep.py --config C4 --budgets 240,120,60,30 [--slots 0,8,3]

For each device-memory budget (GB), and each cap on the checkpoint buffers
(--slots; 0 = one per segment boundary, else the binomial reversal over that
many held states), the planner picks the batch of shots and the checkpoint
interval that fit; the same gradient (observed data modelled
from the true model, start model = the true model smoothed) is timed and
checked against the first budget's result, which it must equal bit for bit
(checkpointed recomputation replays the forward exactly, and batches are
independent).  Every run after the first uses the stencil kernel, rows per
segment and work order the autotuner chose in the first (the tuner picks per
batch size, and different kernels round differently).  One line per
budget: plan, device bytes, wall time, useful Gcell-updates/s.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
from fwi_cuda_amd import _lib, configs, psv


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--config", default="C4", choices=tuple(configs.BUILDERS))
    ap.add_argument("--budgets", default="240,120,60,30", help="device-memory budgets in GB")
    ap.add_argument("--slots", default="0", help="checkpoint-buffer caps (0 = one per segment)")
    ap.add_argument("--batch", type=int, default=0, help="shots per launch (0 = the planner's)")
    ap.add_argument("--nt", type=int, default=0)
    ap.add_argument("--nshot", type=int, default=0)
    ap.add_argument("--json", default="")
    args = ap.parse_args(argv)
    kw = {}
    if args.nt:
        kw["nt"] = args.nt
    if args.nshot:
        kw["nshot"] = args.nshot
    from scipy.ndimage import gaussian_filter
    c = configs.BUILDERS[args.config](**kw)
    start = {k: gaussian_filter(c[k], 20.0, mode="nearest") for k in ("lam", "mu", "rho")}
    nz, nx, nt, nshot = int(c["nz"]), int(c["nx"]), int(c["nt"]), int(c["nshot"])
    useful = 2.0 * nz * nx * nt * nshot
    print(f"{args.config}: {nz}x{nx}, nt {nt}, {nshot} shots", flush=True)

    t0 = time.perf_counter()
    with psv.Simulation(c) as sim:
        obs = sim.forward()
    print(f"observed data: {time.perf_counter() - t0:.1f} s", flush=True)

    rows, ref = [], None
    runs = [(float(b), int(s)) for b in args.budgets.split(",") for s in args.slots.split(",")]
    pin = {}
    for gb, slots in runs:
        try:
            sim = psv.Simulation(c, storage=_lib.STORE_CHECKPOINT, mem_budget=int(gb * 1e9), ckpt_slots=slots,
                                 **({"batch": args.batch} if args.batch else {}), **pin)
        except _lib.FwiError as e:   # the budget holds less than one shot's arrays
            row = {"budget_GB": gb, "slots_cap": slots, "refused": str(e)}
            rows.append(row)
            print(json.dumps(row), flush=True)
            continue
        with sim:
            plan = sim.plan()
            sim.set_model(start["lam"], start["mu"], start["rho"])
            sim.set_observed(*obs)
            t0 = time.perf_counter()
            m = sim.gradient()
            sim.sync()
            wall = time.perf_counter() - t0
            G = sim.get_gradient()
            if not pin:
                import re
                chosen = re.search(r"chosen o(\d)/v(\d+)/tz(\d+)", sim.tuning()[1])
                if chosen:
                    pin = {"work_order": int(chosen.group(1)) + 1, "kernel_variant": int(chosen.group(2)),
                           "tile_tz": int(chosen.group(3))}
                    print(f"kernel of every later run: {pin}", flush=True)
        same, diff = None, None
        if ref is None:
            ref = (m, G)
        else:
            same = bool(np.array_equal(m, ref[0]) and all(np.array_equal(a, b) for a, b in zip(G, ref[1])))
            diff = {"misfit": float(np.max(np.abs(m - ref[0]) / np.abs(ref[0]))),
                    "gradient": [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(G, ref[1])]}
        row = {"budget_GB": gb, "slots_cap": slots, **plan, "seconds": round(wall, 2),
               "Gcell_per_s": round(useful / wall / 1e9, 3), "misfit": [float(v) for v in m],
               "identical_to_first": same, "rel_diff_to_first": diff}
        rows.append(row)
        print(json.dumps(row), flush=True)
        if same is False:
            print("gradient differs from the first budget's", flush=True)
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"config": args.config, "rows": rows}, f, indent=1)
    return 1 if any(r.get("identical_to_first") is False for r in rows) else 0


if __name__ == "__main__":
    sys.exit(main())
