"""Seeded synthetic inputs for the five benchmark configurations.

C1  201 x 401, dx 1 m, circular inclusion, 1 shot, 40 receivers, 100 Hz, nt 2000
C2  301 x 601, dx 1 m, 10 layers, 8 shots, 120 receivers, 250 Hz, nt 3000
C3  1001 x 2001, dx 0.5 m, Gaussian random field, 64 shots, 20-200 receivers each, 100 Hz, nt 4000
C4  2001 x 4001, dx 0.5 m, 15 layers + dipping dyke, 16 shots, 300 receivers, nt 8000
C5  1001 x 2001, 8 layers + 3 ellipses, 32 shots, 150 receivers, 150 Hz, nt 4000

Common set-up (order-8 Holberg stencil, 20-cell CPML left / right / bottom, a
free surface at the top row of the interior (isurf = fpad), vertical-force
sources two cells below the surface, receivers on the surface row, dt = 0.5
of the stencil's stability limit, snapshots and the gradient's correlation on
every 3rd time step as the reference's FWI set-up has them,
scripts/pre_proc_fwi.py:68).  ``nt`` / ``nshot`` / ``snap_dt`` can be changed for
tests; every model is reproducible from its seed.
"""
from __future__ import annotations

import numpy as np

from . import psv

NPML = 20
FDORDER = 8
SNAP_DT = 3   # scripts/pre_proc_fwi.py:68


def gardner(cp):
    return 310.0 * cp ** 0.25


def lame(cp, cs, rho):
    mu = rho * cs * cs
    return rho * cp * cp - 2.0 * mu, mu, rho


def grf(nz, nx, corr, mean, std, seed):
    """Gaussian random field: white noise smoothed by a Gaussian of the given
    correlation length (cells), standardised to mean / std."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal((nz, nx)), corr / np.sqrt(2.0), mode="reflect")
    f = (f - f.mean()) / f.std()
    return mean + std * f


def layers(nz, nx, n, lo, hi, seed, monotone=True):
    rng = np.random.default_rng(seed)
    b = np.sort(rng.choice(np.arange(nz // 20, nz - nz // 20), n - 1, replace=False))
    v = np.sort(rng.uniform(lo, hi, n)) if monotone else rng.uniform(lo, hi, n)
    cp = np.empty((nz, nx))
    edges = [0, *b, nz]
    for i in range(n):
        cp[edges[i]:edges[i + 1]] = v[i]
    return cp


def _case(nz, nx, dx, nt, cp, cs, rho, src_x, rec, fc, seed, snap_dt=SNAP_DT):
    lam, mu, rho = lame(cp, cs, rho)
    N = FDORDER // 2
    hc = psv.holbergcoeff(FDORDER)
    dt = 0.5 * psv.checkfd_dtstab(lam, mu, rho, hc, dx, NPML)
    nshot = len(src_x)
    if isinstance(rec, tuple):
        z_rec, x_rec, rec_off = rec
    else:
        z_rec, x_rec, rec_off = np.full(len(rec), N, np.int32), np.asarray(rec, np.int32), None
    c = psv.make_case(nt=nt, nz=nz, nx=nx, dt=dt, dz=dx, dx=dx, fdorder=FDORDER, lam=lam, mu=mu, rho=rho,
                      z_src=np.full(nshot, N + 2, np.int32), x_src=np.asarray(src_x, np.int32),
                      shot_to_fire=np.arange(nshot), z_rec=z_rec, x_rec=x_rec, rec_off=rec_off,
                      stf_type=0, fc=fc, amp_z=1.0, amp_x=0.0, npml=(0, NPML, NPML, NPML), isurf=(N, 0, 0, 0),
                      snap=(0, nz - NPML, NPML, nx - NPML, snap_dt, 1, 1),
                      taper=(N + 2, N + 12, nz - NPML - 2, nz - NPML - 12, NPML + 2, NPML + 12,
                             nx - NPML - 2, nx - NPML - 12))
    c["seed"] = seed
    return c


def c1(nt=2000, seed=1, true_model=True, snap_dt=SNAP_DT):
    nz, nx = 201, 401
    rng = np.random.default_rng(seed)
    cp = np.full((nz, nx), 3000.0)
    cs = np.full((nz, nx), 1730.0)
    if true_model:
        r = 20
        zc = rng.integers(r + 5 + FDORDER // 2, nz - NPML - r)
        xc = rng.integers(NPML + r, nx - NPML - r)
        zz, xx = np.mgrid[0:nz, 0:nx]
        m = (zz - zc) ** 2 + (xx - xc) ** 2 < r * r
        cp[m] *= 1.1
        cs[m] *= 1.1
    rec = np.linspace(NPML, nx - NPML - 1, 40).astype(np.int32)
    return _case(nz, nx, 1.0, nt, cp, cs, np.full((nz, nx), 2200.0), [nx // 2], rec, 100.0, seed, snap_dt)


def c2(nt=3000, seed=2, nshot=8, snap_dt=SNAP_DT):
    nz, nx = 301, 601
    cp = layers(nz, nx, 10, 2000.0, 4500.0, seed, monotone=False)
    src = np.linspace(NPML + 10, nx - NPML - 11, nshot).astype(np.int32)
    rec = np.linspace(NPML, nx - NPML - 1, 120).astype(np.int32)
    return _case(nz, nx, 1.0, nt, cp, cp / np.sqrt(3.0), gardner(cp), src, rec, 250.0, seed, snap_dt)


def uneven_receivers(nx, nshot, src_x, seed, lo=20, hi=200):
    rng = np.random.default_rng(seed + 1000)
    N = FDORDER // 2
    zr, xr, off = [], [], [0]
    for s in range(nshot):
        n = int(rng.integers(lo, hi + 1))
        o1 = int(rng.integers(-nx // 2, nx // 4))
        o2 = o1 + int(rng.integers(max(n, nx // 8), nx // 2))
        a = int(np.clip(src_x[s] + o1, NPML, nx - NPML - 1))
        b = int(np.clip(src_x[s] + o2, a + 1, nx - NPML - 1))
        x = np.unique(np.linspace(a, b, n).astype(np.int32))
        xr += list(x)
        zr += [N] * len(x)
        off.append(off[-1] + len(x))
    return np.asarray(zr, np.int32), np.asarray(xr, np.int32), np.asarray(off, np.int32)


def c3(nt=4000, seed=3, nshot=64, nz=1001, nx=2001, snap_dt=SNAP_DT):
    cp = grf(nz, nx, 30.0, 3500.0, 300.0, seed)
    rng = np.random.default_rng(seed)
    src = np.sort(rng.integers(NPML + 10, nx - NPML - 10, nshot)).astype(np.int32)
    rec = uneven_receivers(nx, nshot, src, seed)
    return _case(nz, nx, 0.5, nt, cp, cp / 1.8, gardner(cp), src, rec, 100.0, seed, snap_dt)


def c4(nt=8000, seed=4, nshot=16, nz=2001, nx=4001, snap_dt=SNAP_DT):
    cp = layers(nz, nx, 15, 1800.0, 5000.0, seed)
    cs = cp / np.sqrt(3.0)
    rho = gardner(cp)
    # dipping dyke: width 6 cells, 60 degrees from horizontal
    zz, xx = np.mgrid[0:nz, 0:nx]
    x0 = nx // 3
    d = np.abs((xx - x0) * np.sin(np.deg2rad(60)) - zz * np.cos(np.deg2rad(60)))
    m = d < 3.0
    cp[m], cs[m], rho[m] = 5500.0, 3100.0, 2900.0
    src = np.linspace(NPML + 10, nx - NPML - 11, nshot).astype(np.int32)
    rec = np.linspace(NPML, nx - NPML - 1, 300).astype(np.int32)
    return _case(nz, nx, 0.5, nt, cp, cs, rho, src, rec, 100.0, seed, snap_dt)


def c5(nt=4000, seed=5, nshot=32, true_model=True, nz=1001, nx=2001, nrec=150, snap_dt=SNAP_DT):
    """C5; a smaller nz x nx keeps the layout with the anomalies and the start model's
    smoothing scaled by nz / 1001 (the seeded draws are the same at full size)."""
    from scipy.ndimage import gaussian_filter
    f_sz = nz / 1001.0
    cp = layers(nz, nx, 8, 2000.0, 4500.0, seed)
    cs = cp / np.sqrt(3.0)
    rng = np.random.default_rng(seed)
    zz, xx = np.mgrid[0:nz, 0:nx]
    for k in range(3):   # three elliptical anomalies, +8 % / -8 % / +8 % in cp and cs
        zc, xc = rng.integers(nz // 5, nz - nz // 4), rng.integers(nx // 5, nx - nx // 5)
        a, b = rng.integers(30, 80) * f_sz, rng.integers(40, 120) * f_sz
        m = ((zz - zc) / a) ** 2 + ((xx - xc) / b) ** 2 < 1
        f = 1.08 if k % 2 == 0 else 0.92
        cp[m] *= f
        cs[m] *= f
    if not true_model:   # the start model: the true model smoothed with a 40-cell Gaussian
        cp = gaussian_filter(cp, 40.0 * f_sz, mode="nearest")
        cs = gaussian_filter(cs, 40.0 * f_sz, mode="nearest")
    src = np.linspace(NPML + 10, nx - NPML - 11, nshot).astype(np.int32)
    rec = np.linspace(NPML, nx - NPML - 1, nrec).astype(np.int32)
    return _case(nz, nx, 0.5, nt, cp, cs, gardner(cp), src, rec, 150.0, seed, snap_dt)


BUILDERS = {"C1": c1, "C2": c2, "C3": c3, "C4": c4, "C5": c5}
