"""Write the golden fixtures under tests/golden/ from the reference compiled here.

Run in the build container (needs oracle/_ref, built by ``make -C oracle ref``
from the reference's own CPU sources):

    python tests/golden/make_golden.py

Every fixture stores the inputs (keys ``in.*``) next to the reference's outputs,
so tests replay exactly the same case without rebuilding it.  Fixtures:

* ``pieces.npz``   -- holbergcoeff table, cpml_PSV, wavelet, mat_av2, taper2,
  update_mat2, energy_weights2, interpol_grad2, scale_grad_E2, adjsrc2;
* ``o2_pml.npz``   -- order-2 forward (2 shots), snapshots, one shot's adjoint
  gradient pipeline and step_length_PSV, CPML on all four sides;
* ``o2_surf.npz``  -- the same with a top free surface below the grid top
  (isurf = 3 > fpad, so the stress image rows are exercised);
* ``fwi_bin.npz``  -- the reference executable end to end on its binary files:
  forward data generation (seismograms and snapshot files), then FWI mode up to
  the executable's own exit in the first line search (the misfits it prints).
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent
SEED = 20201121


def base_model(nz, nx, seed, inclusion=True):
    """Homogeneous background (cp 3000, cs 1730, rho 2200) plus an optional seeded
    circular inclusion with cp, cs +10 % and rho +5 %."""
    rng = np.random.default_rng(seed)
    cp = np.full((nz, nx), 3000.0)
    cs = np.full((nz, nx), 1730.0)
    rho = np.full((nz, nx), 2200.0)
    if inclusion:
        zc, xc = rng.integers(nz // 3, 2 * nz // 3), rng.integers(nx // 3, 2 * nx // 3)
        zz, xx = np.mgrid[0:nz, 0:nx]
        m = (zz - zc) ** 2 + (xx - xc) ** 2 < 6 ** 2
        cp[m] *= 1.1
        cs[m] *= 1.1
        rho[m] *= 1.05
    mu = rho * cs * cs
    lam = rho * cp * cp - 2.0 * mu
    return lam, mu, rho


def flat(prefix, d):
    out = {}
    for k, v in d.items():
        if isinstance(v, dict):
            out.update(flat(f"{prefix}{k}.", v))
        elif isinstance(v, (tuple, list)) and len(v) and isinstance(v[0], np.ndarray):
            for i, a in enumerate(v):
                out[f"{prefix}{k}.{i}"] = a
        else:
            out[f"{prefix}{k}"] = np.asarray(v)
    return out


def small_case(surf: bool):
    nz, nx, nt = 48, 64, 300
    lam, mu, rho = base_model(nz, nx, SEED)
    hc = O.holberg(2)
    dt = 0.5 * O.dtstab(lam, mu, rho, hc, 1.0, 10)
    npml = (0, 10, 10, 10) if surf else (10, 10, 10, 10)
    isurf = (3, 0, 0, 0) if surf else (0, 0, 0, 0)
    snap = (3 if surf else 10, nz - 10, 10, nx - 10, 3, 1, 1)
    taper = (snap[0] + 2, snap[0] + 5, snap[1] - 2, snap[1] - 5, 12, 16, 52, 48)
    return O.make_case(nt=nt, nz=nz, nx=nx, dt=dt, dz=1.0, dx=1.0, fdorder=2, lam=lam, mu=mu, rho=rho,
                       z_src=[14, 16], x_src=[20, 43], shot_to_fire=[0, 1],
                       z_rec=[12] * 6, x_rec=list(range(12, 52, 7)), fc=150.0,
                       npml=npml, isurf=isurf, snap=snap, taper=taper)


def propagation_fixture(name, surf):
    c = small_case(surf)
    nz, nx = int(c["nz"]), int(c["nx"])
    med_true = O.mat_av2(c, which="ref")
    fwd = [O.kernel_psv(c, med_true, s, which="ref") for s in range(2)]
    obs = [(f["uz"], f["ux"]) for f in fwd]
    lam0, mu0, rho0 = base_model(nz, nx, SEED, inclusion=False)
    med0 = O.mat_av2(c, lam0, mu0, rho0, which="ref")
    sg = O.shot_gradient(c, med0, 1, *obs[1], which="ref")
    # global gradient over both shots, tapered, then the reference's step length
    G = [np.zeros((nz, nx)) for _ in range(3)]
    L2_0 = 0.0
    for s in range(2):
        g = sg if s == 1 else O.shot_gradient(c, med0, s, *obs[s], which="ref")
        L2_0 += g["L2"]
        for i, k in enumerate(("g_lam", "g_mu", "g_rho")):
            G[i] += g[k]
    for i in range(3):
        O.taper2(c, G[i], which="ref")
    # The CPU-sign gradient is an ascent direction here (the reference's line
    # search would exit(0)), so step_length_PSV is pinned along the model
    # difference m_true - m_start, a descent direction.
    D = [c["lam"] - lam0, c["mu"] - mu0, c["rho"] - rho0]
    step = O.ref_step_length(c, 0.01, L2_0, lam0, mu0, rho0, D, obs)
    snaps = fwd[1]["snaps"]
    pick = [0, 33, 66, 99]
    out = {"in": dict(c), "in0": dict(lam=lam0, mu=mu0, rho=rho0)}
    out["out"] = dict(
        seis0_uz=fwd[0]["uz"], seis0_ux=fwd[0]["ux"], seis1_uz=fwd[1]["uz"], seis1_ux=fwd[1]["ux"],
        state1={k: fwd[1]["state"][k] for k in ("vz", "vx", "uz", "ux", "szz", "szx", "sxx", "We")},
        snap_pick=np.asarray(pick), snaps1={k: v[pick] for k, v in snaps.items()},
        media_true=dict(mu_zx=med_true["mu_zx"], rho_zp=med_true["rho_zp"], rho_xp=med_true["rho_xp"],
                        C=med_true["C"]),
        grad1=dict(L2=sg["L2"], raw_lam=sg["raw"][0], raw_mu=sg["raw"][1], raw_rho=sg["raw"][2],
                   We_fwd=sg["We_fwd"], We_adj=sg["We_adj"], We=sg["We"],
                   g_lam=sg["g_lam"], g_mu=sg["g_mu"], g_rho=sg["g_rho"]),
        G_tapered=G, L2_0=L2_0, step_dir=D, step=step)
    np.savez_compressed(OUT / f"{name}.npz", **flat("", out))
    print(name, "L2_0", L2_0, "step", step)


def pieces_fixture():
    rng = np.random.default_rng(SEED + 1)
    out = {}
    table = np.zeros((5, 6, 7))
    for e in range(5):
        for n in range(6):
            table[e, n] = O.holberg(2 * (n + 1), e, which="ref")
    out["holberg"] = table
    cp = {}
    for i, (np1, np2, fpad, nh, dt, dh) in enumerate([(10, 10, 1, 64, 1e-4, 1.0), (20, 0, 4, 120, 8e-5, 1.0),
                                                       (0, 15, 2, 90, 3e-5, 0.5), (12, 9, 4, 80, 5e-5, 2.0)]):
        args = (2.0, 3300.0, 0.001, 1.0 + 0.5 * (i == 3), 100.0 + 50 * i, np1, np2, fpad, nh, dt, dh)
        cp[str(i)] = dict(args=np.asarray(args, np.float64), **O.cpml(*args, which="ref"))
    out["cpml"] = cp
    out["wavelet"] = dict(args=np.asarray([500, 7e-5, 1.0, 120.0, 0.0]),
                          s=O.wavelet(500, 7e-5, 1.0, 120.0, 0.0, which="ref"))
    nz, nx = 30, 41
    lam = rng.uniform(2e9, 9e9, (nz, nx))
    mu = rng.uniform(1e9, 5e9, (nz, nx))
    rho = rng.uniform(1500, 2800, (nz, nx))
    mu[5, 7] = 0.0
    rho[9, 10:13] = 0.0
    rho[10, 10] = 0.0
    c = dict(snap=np.asarray([3, 26, 2, 38, 1, 2, 3], np.int32),
             taper=np.asarray([5, 9, 24, 20, 4, 10, 36, 30], np.int32), nz=nz, nx=nx)
    med = O.mat_av2(c, lam, mu, rho, which="ref")
    out["mat"] = dict(lam=lam, mu=mu, rho=rho, mu_zx=med["mu_zx"], rho_zp=med["rho_zp"],
                      rho_xp=med["rho_xp"], C=med["C"])
    A = rng.standard_normal((nz, nx))
    out["taper"] = dict(snap=c["snap"], taper=c["taper"], A=A, out=O.taper2(c, A.copy(), which="ref"))
    grad = rng.standard_normal((nz, nx))
    out["update"] = dict(old=lam, grad=grad, out=O.update_mat2(lam, grad, 8e9, 2.5e9, 0.03, which="ref"))
    We, Wa = rng.uniform(0, 3, (nz, nx)), rng.uniform(0, 3, (nz, nx))
    out["energy"] = dict(We=We, Wa=Wa, out=O.energy_weights2(c, We, Wa, which="ref"))
    z1, z2, x1, x2, _, sdz, sdx = c["snap"]
    gs = rng.standard_normal((1 + (z2 - z1) // sdz, 1 + (x2 - x1) // sdx))
    out["interp"] = dict(gs=gs, out=O.interpol_grad2(c, gs, which="ref"))
    G = rng.standard_normal((nz, nx))
    gsf = rng.standard_normal((nz, nx))
    out["scale"] = dict(G=G, gs=gsf, We=We + 0.5, mat_av=np.asarray(3.3e9),
                        out=O.scale_grad_E2(c, G.copy(), gsf, 3.3e9, We + 0.5, which="ref"))
    muz, mux = rng.standard_normal((5, 70)), rng.standard_normal((5, 70))
    tuz, tux = rng.standard_normal((5, 70)), rng.standard_normal((5, 70))
    ruz, rux, L2 = O.adjsrc2(muz, mux, tuz, tux, 3e-5, which="ref")
    out["adjsrc"] = dict(muz=muz, mux=mux, tuz=tuz, tux=tux, ruz=ruz, rux=rux, L2=np.asarray(L2))
    np.savez_compressed(OUT / "pieces.npz", **flat("", out))
    print("pieces written")


# ---------------------------------------------------------------------------
# End to end through the reference executable and its binary files
# ---------------------------------------------------------------------------

def write_inputs(d: Path, p: dict, lam, mu, rho):
    """The four input files in the layout scripts/pre_proc_fwi.py:240-276 writes."""
    (d / "bin").mkdir(exist_ok=True)
    (d / "io").mkdir(exist_ok=True)
    metaint = np.array([0, p["surf"], 1, 1, p["accu_save"], 1, p["fwinv"], p["fwinv"], p["nt"], p["nz"], p["nx"],
                        0, p["nt"] - 1, *p["snap"][:4], *p["snap"][4:], *p["taper"],
                        p["nsrc"], p["nrec"], p["nshot"], p["stf_type"], 0, p["fdorder"], p["fdorder"] // 2,
                        p["mat_save_interval"], 1], dtype=np.int32)
    assert metaint.size == 37
    intarray = np.concatenate([np.asarray(p["npml"], np.int32), np.asarray(p["isurf"], np.int32),
                               p["z_src"], p["x_src"], p["fire"], p["z_rec"], p["x_rec"]]).astype(np.int32)
    metafloat = np.array([p["dt"], p["dz"], p["dx"], 2.0, p["damp_v"], 0.001, 1.0, p["fc"],
                          lam[0, 0], mu[0, 0], rho[0, 0]], dtype=np.float64)
    metaint.tofile(d / "bin" / "metaint.bin")
    intarray.tofile(d / "bin" / "intarray.bin")
    metafloat.tofile(d / "bin" / "metafloat.bin")
    np.concatenate([lam.ravel(), mu.ravel(), rho.ravel()]).astype(np.float64).tofile(d / "bin" / "mat.bin")
    return dict(metaint=metaint, intarray=intarray, metafloat=metafloat)


def run_until(d: Path, stop_pattern: str | None, timeout=600):
    """Run the reference executable in d; stop it once stop_pattern has printed
    (its FWI loop has no working stop rule, src/cpu/n_simulate_PSV.cpp:576-581)."""
    proc = subprocess.Popen([str(O.REF_BIN)], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                            text=True, env=dict(os.environ, OMP_NUM_THREADS="4"))
    lines = []
    try:
        for line in proc.stdout:
            if "Time step" in line:
                continue
            lines.append(line.rstrip("\n"))
            if stop_pattern and re.search(stop_pattern, line):
                break
    finally:
        if proc.poll() is None:
            proc.kill()
        proc.wait(timeout=timeout)
    return lines


def fwi_bin_fixture():
    nz, nx, nt = 40, 56, 250
    lam, mu, rho = base_model(nz, nx, SEED + 7)
    lam0, mu0, rho0 = base_model(nz, nx, SEED + 7, inclusion=False)
    hc = O.holberg(2)
    dt = float(np.float64(0.5 * O.dtstab(lam, mu, rho, hc, 1.0, 8)))
    npml = (8, 8, 8, 8)
    snap = [8, nz - 8, 8, nx - 8, 2, 1, 1]
    p = dict(surf=0, fwinv=0, accu_save=1, nt=nt, nz=nz, nx=nx, snap=snap, taper=[10, 13, 30, 27, 10, 14, 46, 42],
             nsrc=2, nrec=5, nshot=2, stf_type=1, fdorder=2, mat_save_interval=1, npml=npml, isurf=[0, 0, 0, 0],
             z_src=np.array([12, 12], np.int32), x_src=np.array([18, 38], np.int32),
             fire=np.array([0, 1], np.int32), z_rec=np.full(5, 10, np.int32),
             x_rec=np.array([12, 20, 28, 36, 44], np.int32), dt=dt, dz=1.0, dx=1.0,
             damp_v=3300.0, fc=180.0)
    with tempfile.TemporaryDirectory() as tmp:
        d = Path(tmp)
        files_fwd = write_inputs(d, p, lam, mu, rho)
        log_fwd = run_until(d, None)
        obs = {}
        for s in range(2):
            for comp in ("uz", "ux"):
                f = d / "bin" / f"shot{s}_rtf_{comp}.bin"
                obs[f"shot{s}_{comp}"] = np.fromfile(f, np.float64).reshape(5, nt)
                shutil.copy(f, d / "io" / f.name)
        snt = 1 + (nt - 1) // snap[4]
        snz, snx = 1 + (snap[1] - snap[0]) // snap[5], 1 + (snap[3] - snap[2]) // snap[6]
        pick = np.array([0, 40, 80, 124])
        accu = {f: np.fromfile(d / "bin" / f"shot1_{f}.bin", np.float64).reshape(snt, snz, snx)[pick]
                for f in ("vz", "vx", "szz", "szx", "sxx")}
        # FWI mode: with the CPU gradient sign the first line search fails and the
        # executable exits (src/cpu/n_step_length_PSV.cpp:173-178); the misfits it
        # prints on the way (6 significant digits) are kept.
        p["fwinv"] = 1
        p["accu_save"] = 0
        files_fwi = write_inputs(d, p, lam0, mu0, rho0)
        log = run_until(d, r"ITERATION 1 <DONE>")
    l2 = [float(m.group(1)) for line in log for m in [re.search(r"^L2 NORM: \S+, (\S+)", line)] if m]
    trials = [(float(m.group(1)), float(m.group(2))) for line in log
              for m in [re.search(r"^Step Length: (\S+), L2 = (\S+), counter", line)] if m][::2]
    out = dict(files_fwd=files_fwd, files_fwi=files_fwi, true=dict(lam=lam, mu=mu, rho=rho),
               start=dict(lam=lam0, mu=mu0, rho=rho0), obs=obs, accu_pick=pick, accu1=accu,
               l2_printed=np.asarray(l2), trials_printed=np.asarray(trials))
    np.savez_compressed(OUT / "fwi_bin.npz", **flat("", out))
    (OUT / "fwi_bin_stdout.txt").write_text("\n".join(log_fwd[-5:] + ["----"] + log) + "\n")
    print("fwi_bin: L2", l2, "trials", len(trials))


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("build the reference first: make -C oracle ref")
    pieces_fixture()
    propagation_fixture("o2_pml", surf=False)
    propagation_fixture("o2_surf", surf=True)
    fwi_bin_fixture()
