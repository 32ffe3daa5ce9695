#!/usr/bin/env python3
"""Benchmark: forward + adjoint gradient throughput of the elastic P/SV path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C3] [--shots-per-step 8]

Workload (default C3): 1001 x 2001 grid, dx 0.5 m, 64 shots with 20-200
receivers each, nt = 4000, order-8 Holberg stencil, CPML 20, top free surface.
Synthetic data: a Gaussian-random-field true model generates the observed
seismograms (untimed); the gradient is taken at its smoothed version.

One step = the full gradient of one group of shots (--shots-per-step, 8 by
default): forward over all nt time steps with checkpoints, residual and misfit,
checkpointed adjoint with the fused correlation, energy-weighted
preconditioning added into the global gradient.  A group is one launch batch
(every stencil launch advances its shots together).  Consecutive steps walk
the workload's shot groups in order and wrap around, so 64 / 8 = 8 steps are
one complete C3 gradient.  W untimed steps, then K timed steps between two
device synchronisations; ms_per_step x K is the timed wall clock.

Printed JSON (one line): value = Gcell-updates/s = 2 nz nx nt x (shots of the
timed steps) / wall (forward and adjoint cells; recomputed steps are not
counted), from the timed run, whose launches also carry the HIP events the
roofline is computed from (on every --event-stride-th launch of each stencil
kernel, 8 by default: events on every launch cost ~4.6 % of the step).  With
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average duration (HIP events on the library's stream) against
                MI355X's 8 TB/s; beside it two measured ceilings, torch's
                elementwise multiply (copy_GBps) and measure/stream_probe, a
                kernel with the stencil's access shape and no arithmetic
                (shape_GBps); traffic = counter-measured HBM bytes per launch
                from profiles/roofline_traffic.txt when that file holds the
                same kernel instance at the same algorithmic bytes per launch;
  hbm_frac      the whole step's algorithmic bytes (fwd 68 B, adjoint 84 B,
                gradient 44 B and snapshot 20 B per snapshot step, recomputed
                step 60 B per cell-step, CPML strips 32 B per cell-step per
                direction) over wall time and 8 TB/s;
  graph         one step in a new context with time-step chunks captured in
                hipGraphs and replayed (wall clock only);
  plain         the reference-structured per-shot form on the same grid (one
                shot after a warm-up shot) and the speed-up in shots/s;
  cpu_baseline  the reference's CPU path (compiled from its sources into
                oracle/_ref) on a bounded sample of the same grid.
The optional legs (copy probe, graph, plain, cpu, shape probe) are dropped,
and the line still printed, when they would end past --deadline seconds.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
T_START = time.perf_counter()

HBM_PEAK = 8.0e12
METRIC = "fwd+adj gradient Gcell-updates/s & % HBM peak; batched-checkpointed vs plain"
# Per interior cell and launch, algorithmic bytes of each stencil kernel (FP32;
# every field and medium read once, every updated field written once).
# One-pass step (k_step):
B_STEP_FWD = 68        # RW vz vx szz szx sxx, R lam mu mu_zx rho_zp rho_xp, RW We
B_STEP_RECOMP = 60     # as forward without We
B_STEP_ADJ = 84        # forward + RW uz ux
B_SNAP = 20            # W 5 snapshot fields on a snapshot step (written by k_step or k_velocity)
B_GRAD = 44            # R 5 snapshot fields, RW 3 gradient sums on a snapshot step
# Two-kernel step (k_stress + k_velocity):
B_STRESS = 44          # R vz vx, RW szz szx sxx, R lam mu mu_zx
B_VEL_FWD = 44         # R szz szx sxx, RW vz vx, R rho_zp rho_xp, RW We
B_VEL_RECOMP = 36      # as forward without We
B_VEL_ADJ = 60         # + RW uz ux
B_VEL_GRAD = 52        # gradient: R 5 snapshot fields, RW 3 sums, R lam mu
# CPML memory: 4 psi RW per strip cell and direction per step (half in each kernel of the two-kernel step)
B_PML = 32


def log(*a):
    print(f"[{time.perf_counter() - T_START:6.1f} s]", *a, file=sys.stderr, flush=True)


def build_workload(name, nt, nshot):
    from fwi_cuda_amd import configs
    from scipy.ndimage import gaussian_filter
    kw = {}
    if nt:
        kw["nt"] = nt
    if nshot:
        kw["nshot"] = nshot
    c = configs.BUILDERS[name](**kw)
    # start model: the true model smoothed with a 20-cell Gaussian
    lam0, mu0, rho0 = (gaussian_filter(c[k], 20.0, mode="nearest") for k in ("lam", "mu", "rho"))
    return c, (lam0, mu0, rho0)


def shot_groups(nshot, per_step):
    """Consecutive groups of per_step shots (the last one may be smaller)."""
    return [(s, min(per_step, nshot - s)) for s in range(0, nshot, per_step)]


def kernel_instantiation(kind, tune_log, N):
    """The template instance the autotuner picked for a stencil kernel kind, as the
    profiler names it (without the argument list); a launch-mode template argument
    (forward, recomputation, adjoint) shows as *."""
    import re
    m = re.search(r"chosen o\d/v(\d+)/tz(\d+)", tune_log or "")
    if not m:
        return None
    var, tz = int(m.group(1)), int(m.group(2))
    if kind == "step":   # several launch modes share the prefix
        if var == 9:   # balanced launches: variant 8's layout, segments of 4 blocks
            return f"k_step7<{N}, 4, *, true>"
        if var in (7, 8):   # variant 8: the block's wavefronts side by side in x
            return f"k_step7<{N}, {max(1, tz // (2 * N + 2))}, *, {'true' if var == 8 else 'false'}>"
        return f"k_step<{N}, {tz}, *>"
    base = {"stress": "k_stress", "velocity": "k_velocity"}[kind]
    if var == 1:
        return f"{base}<{N}, {tz}>"
    if var in (5, 10, 11, 12, 13):   # 10-13: variant 5's kernels on full 64-column tiles
        return f"{base}{var}<{N}, {tz}>" if kind == "stress" else f"{base}{var}<{N}, {tz}, *>"
    if var == 6:   # segments of tz = M blocks of 2N + 2 rows, loads one row ahead
        m = max(1, tz // (2 * N + 2))
        return f"{base}6<{N}, 1, {m}>" if kind == "stress" else f"{base}6<{N}, 1, {m}, *>"
    return f"{base}2<{N}, {tz}, {1 if var == 2 else 6}>"


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """A measured ceiling for the roofline: GB/s of an elementwise multiply (one load and
    one store per element, torch's kernel) over two nbytes buffers on this GPU, or None."""
    try:
        import torch
        if not torch.cuda.is_available():
            log("copy probe skipped: torch sees no GPU")
            return None
        n = nbytes // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            torch.mul(src, 2.0, out=dst)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del src, dst
        return 8.0 * n / (ms / 1e3) / 1e9
    except Exception as e:   # the probe is a reported reference, not the measurement
        log(f"copy probe failed: {e!r}")
        return None


def shape_bandwidth(nz, P, B):
    """A second measured ceiling: GB/s of measure/stream_probe, a kernel with the
    stencil's access shape and no arithmetic (10 [B][nz][P] arrays read and 5 written,
    a wavefront marching 64 columns down 40 rows, the one-pass step's block layout), run
    as a child process; None when the probe is not built or fails."""
    exe = ROOT / "measure" / "stream_probe"
    if not exe.exists():
        log("shape probe skipped: measure/stream_probe is not built")
        return None
    P = (P + 255) // 256 * 256
    while B > 1 and nz * P * B * 4 >= 1 << 31:   # the probe keeps each array below 2 GiB
        B //= 2
    try:
        r = subprocess.run([str(exe), str(nz), str(P), str(B)], capture_output=True, text=True, timeout=120)
        if r.returncode != 0:
            log(f"shape probe failed: {r.stderr.strip()}")
            return None
        return json.loads(r.stdout.strip().splitlines()[-1])["shape_GBps"]
    except Exception as e:
        log(f"shape probe failed: {e!r}")
        return None


def measured_traffic(inst, algo_bytes, path=None, rtol=0.01):
    """HBM bytes per launch of a kernel instance from the committed counter summary
    (lines 'traffic <instance> <mix> ... bytes_per_launch=<B> algorithmic_bytes_per_launch=<A>').
    A line counts only when its instance matches and its algorithmic bytes per launch
    are within rtol of this run's (same grid, batch and launch mix); otherwise None."""
    path = Path(path) if path else ROOT / "profiles" / "roofline_traffic.txt"
    if not inst or not path.exists():
        return None
    for line in path.read_text().splitlines():
        if not line.startswith("traffic "):
            continue
        toks = line.split()
        kv = dict(t.split("=", 1) for t in toks if "=" in t)
        name = line[len("traffic "):].split(">", 1)[0] + ">"
        if name.strip() != inst.strip() or "bytes_per_launch" not in kv:
            continue
        a = float(kv.get("algorithmic_bytes_per_launch", "nan"))
        if not (abs(a - algo_bytes) <= rtol * algo_bytes):
            continue
        return float(kv["bytes_per_launch"])
    return None


class quiet_stdout:
    """Send file descriptor 1 to /dev/null (the reference prints progress from C++)."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        self.null = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self.null, 1)

    def __exit__(self, *exc):
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        os.close(self.null)


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or None


def cpu_baseline(c, budget_s=15.0, which=None):
    """The reference's kernel_PSV (compiled from its sources into oracle/_ref)
    on one shot of the same grid, forward then adjoint, over a bounded number
    of time steps chosen to take about budget_s seconds; which="oracle" times
    the FP64 restatement at the workload's own order instead."""
    from oracle import oracle as O
    if which is None:
        which = "ref" if O.ref_available() else "oracle"
    kind = "reference" if which == "ref" else "port"
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nz, nx = int(c["nz"]), int(c["nx"])
    cc = dict(c)
    cc.pop("rec_off", None)
    cc["z_rec"], cc["x_rec"] = c["z_rec"][:20].copy(), c["x_rec"][:20].copy()
    cc["z_src"], cc["x_src"] = c["z_src"][:1].copy(), c["x_src"][:1].copy()
    cc["shot_to_fire"] = np.zeros(1, np.int32)
    cc["nshot"] = 1
    if which == "ref":
        cc["fdorder"] = 2   # the reference's kernels compute order 2 whatever fdorder says
        cc["hc"] = O.holberg(2)
    med = O.mat_av2(cc, which="oracle")

    def run(nts):
        d = dict(cc)
        d["nt"] = nts
        d["stf_z"] = np.ascontiguousarray(cc["stf_z"][:1, :nts])
        d["stf_x"] = np.ascontiguousarray(cc["stf_x"][:1, :nts])
        t0 = time.perf_counter()
        f = O.kernel_psv(d, med, 0, grad=False, accu=True, which=which)
        adj = (np.ones_like(f["uz"]) * 1e-9, np.ones_like(f["ux"]) * 1e-9)
        O.kernel_psv(d, med, 0, grad=True, accu=False, snaps=f["snaps"], adj=adj, which=which)
        return time.perf_counter() - t0

    with quiet_stdout():
        t_probe = run(4)
        # bounded by time and by the snapshot memory of the reference (5 FP64 fields per step)
        nts = int(max(8, min(120, 4 * budget_s / max(t_probe, 1e-3))))
        t = run(nts)
    rate = 2.0 * nz * nx * nts / t / 1e9
    return {"value": rate, "unit": "Gcell-updates/s", "cores": threads, "kind": kind, "cpu_model": cpu_model(),
            "sample": f"one shot of the {nz}x{nx} grid, {nts} forward + {nts} adjoint steps "
                      f"({'order 2, the reference kernels, FP64, OpenMP' if which == 'ref' else 'order 8 FP64 oracle'}),"
                      f" {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed steps (8 = one full C3 gradient)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed steps before the timed ones")
    ap.add_argument("--config", default="C3")
    ap.add_argument("--shots-per-step", type=int, default=8,
                    help="shots whose full gradient one step computes (one launch batch)")
    ap.add_argument("--nt", type=int, default=0, help="shorter time axis (profiling runs)")
    ap.add_argument("--nshot", type=int, default=0)
    ap.add_argument("--storage", choices=("auto", "full", "checkpoint"), default="auto")
    ap.add_argument("--ckpt-interval", type=int, default=0)
    ap.add_argument("--event-stride", type=int, default=8,
                    help="HIP events on every k-th launch of the stencil kernels (1: every launch)")
    ap.add_argument("--ckpt-keep", type=int, default=0,
                    help="segments whose snapshots the forward pass keeps (0: as many as fit)")
    ap.add_argument("--tile-tz", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0, help="stencil kernel variant (0 = tuned)")
    ap.add_argument("--order", type=int, default=0, help="stencil work order 1..4 (0 = tuned)")
    ap.add_argument("--deadline", type=float, default=420.0,
                    help="seconds from start after which optional legs are dropped")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="skip the hipGraph-replay step")
    ap.add_argument("--no-probes", action="store_true", help="skip the copy and access-shape probes")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB five times (calibrates rocprof byte counters)")
    ap.add_argument("--copy-bandwidth", action="store_true",
                    help="only measure the elementwise-multiply bandwidth and print it (run as a child process)")
    args = ap.parse_args()
    if args.copy_bandwidth:
        print(json.dumps({"copy_GBps": copy_bandwidth()}))
        return
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)                  # the runtime's blit kernel
            torch.mul(src, 2.0, out=dst)    # an ordinary elementwise kernel: one load, one store per element
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_MiB": args.copy_probe, "bytes_read_per_copy": 4 * n,
                          "bytes_written_per_copy": 4 * n}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} (WORLD_SIZE {world}) is not run. Every GPU box "
              "here has one GPU; shots are independent, and the multi-GPU design shards shots and "
              "all-reduces the gradient once.", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0 or args.shots_per_step < 1:
        ap.error("--steps >= 1, --warmup >= 0, --shots-per-step >= 1")

    def have_time(est_s):
        return time.perf_counter() - T_START + est_s < args.deadline

    # The measured ceiling comes from a child process that uses only torch, started before
    # this process touches the GPU: torch's bundled HIP runtime does not initialise in a
    # process where the library's runtime already has.
    copy = None
    if not args.no_probes:
        try:
            r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--copy-bandwidth"],
                               capture_output=True, text=True, timeout=150)
            copy = json.loads(r.stdout.strip().splitlines()[-1])["copy_GBps"] if r.returncode == 0 else None
        except Exception as e:
            log(f"copy probe failed: {e!r}")
        log(f"copy probe: {copy} GB/s")

    from fwi_cuda_amd import _lib, psv

    c, start = build_workload(args.config, args.nt, args.nshot)
    nz, nx, nt, nshot = int(c["nz"]), int(c["nx"]), int(c["nt"]), int(c["nshot"])
    N = int(c["fdorder"]) // 2
    sps = min(args.shots_per_step, nshot)
    groups = shot_groups(nshot, sps)
    log(f"workload {args.config}: {nz}x{nx}, nt {nt}, {nshot} shots "
        f"({int(c['rec_off'][-1]) if 'rec_off' in c else ''} receivers), {len(groups)} groups of <= {sps} shots")
    storage = {"auto": _lib.STORE_AUTO, "full": _lib.STORE_FULL, "checkpoint": _lib.STORE_CHECKPOINT}[args.storage]
    kw = dict(batch=sps, tile_tz=args.tile_tz, kernel_variant=args.variant, work_order=args.order,
              storage=storage, ckpt_interval=args.ckpt_interval, ckpt_keep=args.ckpt_keep)
    sim = psv.Simulation(c, **kw)
    plan = sim.plan()
    log("plan", plan)
    # the observed-data forward (it also runs the autotuner) is timed per launch too, so
    # that launch counts and averages can be set beside a profiler's whole-run summary;
    # only the stencil kernels carry events (the roofline's candidates), the small
    # per-step kernels run untimed
    stencil = ("stress", "velocity", "step")
    sim.profile(True, kernels=stencil, stride=args.event_stride)
    t0 = time.perf_counter()
    obs = sim.forward()
    log(f"observed data generated in {time.perf_counter() - t0:.1f} s")
    kt_obs = sim.kernel_times()
    sim.profile(False)
    sim.set_observed(*obs)
    sim.set_model(*start)

    def step(k):
        """Step k: the full gradient of shot group k mod (number of groups)."""
        s0, ns = groups[k % len(groups)]
        if s0 == 0:
            sim.zero_gradient()   # a new pass over the workload's shots
        return s0, ns, sim.gradient(s0, ns)

    for w in range(args.warmup):
        t0 = time.perf_counter()
        step(w)
        sim.sync()
        log(f"warmup step {w}: {time.perf_counter() - t0:.2f} s")
    try:
        import torch
        sync_torch = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync_torch = lambda: None  # noqa: E731
    sim.profile(True, kernels=stencil, stride=args.event_stride)
    sim.sync()
    sync_torch()
    t0 = time.perf_counter()
    shots_timed = 0
    misfit = {}
    for k in range(args.steps):
        s0, ns, m = step(args.warmup + k)
        shots_timed += ns
        misfit[s0] = m
    sim.sync()
    sync_torch()
    wall = time.perf_counter() - t0
    kt = sim.kernel_times()
    sim.profile(False)
    tz, tune_log = sim.tuning()
    K = args.steps
    log(f"timed {K} step(s), {shots_timed} shot gradients: {wall:.2f} s ({wall / K * 1e3:.0f} ms per step); "
        f"tile rows {tz}")
    log("tuning:", tune_log)
    log("kernel times (launches, ms of the timed ones, timed)", {k: (n, round(ms, 1), t) for k, (n, ms, t) in kt.items()})

    cells = nz * nx
    interior = (nz - 2 * N) * (nx - 2 * N)

    # executed algorithmic bytes of the whole gradient over the full grid: forward 68,
    # adjoint 84, gradient 44 and snapshot store 20 per snapshot step, recomputed
    # step 60 when checkpointed, CPML 32 per strip cell and direction; checkpoint
    # copies (<0.3 % at C3) are left out
    ckpt = plan["storage"] == "checkpoint"
    # forward steps recomputed per time step: below 1 when the forward pass keeps the last
    # segments' snapshots, above 1 on a binomial schedule
    rec = float(plan["recompute"]) if ckpt else 0.0
    sdt = int(c["snap"][4])
    zs = int(np.count_nonzero((c["a_z"] != 0) | (c["a_half_z"] != 0))) if c["pml_z"] else 0
    xs = int(np.count_nonzero((c["a_x"] != 0) | (c["a_half_x"] != 0))) if c["pml_x"] else 0
    pml_cells = zs * nx + xs * nz
    model_bytes = (cells * (B_STEP_FWD + B_STEP_ADJ + (B_GRAD + B_SNAP) / sdt + B_STEP_RECOMP * rec)
                   + pml_cells * B_PML * (2 + rec))

    # roofline of the dominant kernel from its live per-launch times.  Every launch
    # advances one batch by one step; the bytes are summed over the launch modes the
    # gradient runs (forward, recomputation when checkpointed, adjoint) and divided
    # by the kernel's summed launch time.
    shot_steps = shots_timed * nt
    passes = 2 + rec
    zs_i = int(np.count_nonzero(((c["a_z"] != 0) | (c["a_half_z"] != 0))[N:nz - N])) if c["pml_z"] else 0
    xs_i = int(np.count_nonzero(((c["a_x"] != 0) | (c["a_half_x"] != 0))[N:nx - N])) if c["pml_x"] else 0
    strip = zs_i * (nx - 2 * N) + xs_i * (nz - 2 * N)   # interior cells inside a CPML strip (per direction)
    per_shot_step = {
        "step": interior * (B_STEP_FWD + B_STEP_ADJ + B_GRAD / sdt + B_SNAP / sdt
                            + B_STEP_RECOMP * rec) + strip * B_PML * passes,
        "stress": interior * B_STRESS * passes + strip * B_PML / 2 * passes,
        "velocity": interior * (B_VEL_FWD + B_VEL_ADJ + (B_VEL_GRAD + B_SNAP) / sdt
                                + B_VEL_RECOMP * rec) + strip * B_PML / 2 * passes,
    }
    # per kernel class: algorithmic bytes of all its launches, its launches, and the
    # average duration of the launches that carried events (every event_stride-th)
    kern = {k: (per_shot_step[k] * shot_steps, kt[k][1] / kt[k][2], kt[k][0], kt[k][2])
            for k in per_shot_step if kt[k][2] > 0}
    dom = max(kern, key=lambda k: kern[k][1] * kern[k][2])   # the most time: average x launches
    b, avg_ms, n, timed = kern[dom]
    achieved = b / n / (avg_ms / 1e3) / 1e9
    # every launch of the run that events timed (the observed-data forward and the timed
    # gradients): the launch set of a whole-run profiler summary of this command
    obs_shot_step = {"step": interior * B_STEP_FWD + strip * B_PML,
                     "stress": interior * B_STRESS + strip * B_PML / 2,
                     "velocity": interior * B_VEL_FWD + strip * B_PML / 2}
    b_all = b + obs_shot_step[dom] * nshot * nt
    n_all, timed_all = n + kt_obs[dom][0], timed + kt_obs[dom][2]
    avg_all = (avg_ms * timed + kt_obs[dom][1]) / max(timed_all, 1)
    inst = kernel_instantiation(dom, tune_log, N)
    algo_per_launch = b / max(n, 1)
    traffic = measured_traffic(inst, algo_per_launch)
    roofline = {"kernel": dom, "instantiation": inst, "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved * 1e9 / HBM_PEAK, 4),
                "traffic": traffic, "avg_launch_ms": round(avg_ms, 4),
                "launches": n, "timed_launches": timed, "event_stride": args.event_stride,
                "algorithmic_bytes_per_launch": int(algo_per_launch),
                "all_launches": {"launches": n_all, "timed_launches": timed_all, "avg_launch_ms": round(avg_all, 4),
                                 "algorithmic_bytes_per_launch": int(b_all / max(n_all, 1))}}
    if traffic:
        roofline["traffic_source"] = ("profiles/roofline_traffic.txt: rocprofv3 FETCH_SIZE and WRITE_SIZE per "
                                      "launch, calibrated against a copy kernel, averaged over the gradient's "
                                      "launch mix (forward, recomputation, adjoint)")
        roofline["traffic_over_algorithmic"] = round(traffic / max(algo_per_launch, 1.0), 3)
    per_kernel = {k: {"achieved_GBps": round(v[0] / v[2] / (v[1] / 1e3) / 1e9, 1) if v[1] else None,
                      "launches": v[2], "timed_launches": v[3], "avg_launch_ms": round(v[1], 4),
                      "est_total_ms": round(v[1] * v[2], 1)} for k, v in kern.items()}
    ms_step = wall / K * 1e3
    sim.close()
    if copy:
        roofline["copy_GBps"] = round(copy, 1)
        roofline["copy_probe"] = "torch elementwise multiply, 1 GiB in, 1 GiB out, this GPU"
        roofline["frac_of_copy"] = round(achieved / copy, 4)

    # One step in a new context with chunks of time steps captured in hipGraphs and
    # replayed (per-kernel events do not exist inside a graph: wall clock only).
    graph = None
    if not args.no_graph:
        if have_time(2.0 * ms_step / 1e3 + 20.0):
            with psv.Simulation(c, use_graph=1, **kw) as gs:
                gs.set_observed(*obs)
                gs.set_model(*start)
                s0, ns = groups[0]
                gs.sync()
                sync_torch()
                t0 = time.perf_counter()
                gm = gs.gradient(s0, ns)
                gs.sync()
                sync_torch()
                tg = time.perf_counter() - t0
            same = bool(s0 in misfit and np.allclose(gm, misfit[s0], rtol=1e-12, atol=0))
            graph = {"steps": 1, "ms_per_step": round(tg * 1e3, 2),
                     "Gcell_per_s": round(2.0 * cells * nt * ns / tg / 1e9, 3),
                     "speedup_vs_eager_with_events": round(ms_step / 1e3 * ns / sps / tg, 3),
                     "misfit_equal_to_eager": same}
            log("graph replay:", graph)
        else:
            graph = {"skipped": f"past the {args.deadline:.0f} s deadline"}

    value = 2.0 * cells * nt * shots_timed / wall / 1e9
    shots_per_s = shots_timed / wall
    hbm_frac = model_bytes * nt * shots_timed / wall / HBM_PEAK

    plain = None
    if not args.no_plain:
        est = 2.0 * 5.0 * ms_step / 1e3 / sps + 20.0   # the plain form runs ~4x slower per shot
        if have_time(est):
            try:
                with psv.Simulation(c, form=_lib.FORM_PLAIN, storage=_lib.STORE_FULL) as ps:
                    ps.set_observed(*obs)
                    ps.set_model(*start)
                    ps.zero_gradient()
                    ps.gradient(0, 1)   # warm-up shot
                    ps.zero_gradient()
                    ps.sync()
                    t0 = time.perf_counter()
                    ps.gradient(min(1, nshot - 1), 1)
                    ps.sync()
                    tp = time.perf_counter() - t0
                plain = {"shots_per_s": round(1.0 / tp, 4), "s_per_shot": round(tp, 3),
                         "Gcell_per_s": round(2.0 * cells * nt / tp / 1e9, 3),
                         "speedup_shots_per_s": round(shots_per_s * tp, 2)}
                log("plain per-shot form:", plain)
            except Exception as e:   # noqa: BLE001
                plain = {"error": str(e)}
                log("plain form failed:", e)
        else:
            plain = {"skipped": f"past the {args.deadline:.0f} s deadline"}

    cpu = None
    if not args.no_cpu:
        if have_time(30.0):
            try:
                cpu = cpu_baseline(c)
                log("cpu baseline:", cpu)
                if cpu.get("kind") == "reference" and have_time(25.0):
                    # the reference has no order-8 path: the FP64 restatement at the workload's order
                    o8 = cpu_baseline(c, budget_s=8.0, which="oracle")
                    cpu["oracle_same_order"] = {k: o8[k] for k in ("value", "unit", "cores", "sample")}
                    log("cpu oracle at the workload's order:", o8)
            except Exception as e:   # noqa: BLE001
                cpu = {"error": str(e)}
        else:
            cpu = {"skipped": f"past the {args.deadline:.0f} s deadline"}

    # the stencil's own access shape, on the batch's array size (after the contexts are closed)
    if not args.no_probes and have_time(20.0):
        shape = shape_bandwidth(nz, (nx + 1 + 63) // 64 * 64, int(plan["batch"]))
        if shape:
            roofline["shape_GBps"] = round(shape, 1)
            roofline["shape_probe"] = ("measure/stream_probe: 10 [B][nz][P] arrays read, 5 written, 64-column "
                                       "wavefronts marching 40 rows, no arithmetic, this GPU")
            roofline["frac_of_shape"] = round(achieved / shape, 4)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian-random-field model, observed data modelled on the GPU)",
        "config": {"workload": f"{args.config}: {nz}x{nx} grid, dx {c['dx']} m, nt {nt}, {nshot} shots, "
                               f"uneven receivers, order {int(c['fdorder'])}, gradient correlated every "
                               f"{sdt} steps (the reference's snap_dt)",
                   "step": f"full gradient of {sps} shots (one launch batch); {len(groups)} steps = one "
                           f"gradient over all {nshot} shots",
                   "shots_per_step": sps, "batch": plan["batch"], "storage": plan["storage"],
                   "ckpt_interval": plan["ckpt_interval"], "ckpt_units": plan["ckpt_units"],
                   "ckpt_kept": plan["ckpt_kept"], "recompute": plan["recompute"], "snap_dt": sdt, "tile_rows": tz,
                   "parallelism": "single GPU, shots batched per launch"},
        "shots_per_s": round(shots_per_s, 3),
        "hbm_frac": round(hbm_frac, 4),
        "roofline": roofline,
        "kernels": per_kernel,
        "timed_run": (f"eager launches; HIP events on every {args.event_stride}-th launch of each stencil kernel "
                      "(the roofline's average launch duration)"),
        "graph": graph,
        "plain": plain,
        "cpu_baseline": cpu,
        "bench_wall_s": round(time.perf_counter() - T_START, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
