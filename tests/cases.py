"""Seeded test cases shared by the CPU and GPU tests (built with the oracle's
preprocessing so the HIP path and the oracle see identical inputs)."""
import numpy as np


def layered_model(nz, nx, seed, nlayer=4, cp_range=(2200.0, 3600.0), inclusion=True, vs_ratio=1.73):
    rng = np.random.default_rng(seed)
    cp = np.empty((nz, nx))
    bounds = np.sort(rng.integers(nz // 8, nz, nlayer - 1))
    vals = np.sort(rng.uniform(*cp_range, nlayer))
    z = np.arange(nz)[:, None]
    cp[:] = vals[0]
    for b, v in zip(bounds, vals[1:]):
        cp[np.broadcast_to(z >= b, (nz, nx))] = v
    if inclusion:
        # shallow enough that its reflection reaches surface receivers within the test records
        zc, xc = rng.integers(nz // 4, nz // 2), rng.integers(nx // 3, 2 * nx // 3)
        zz, xx = np.mgrid[0:nz, 0:nx]
        cp[(zz - zc) ** 2 + (xx - xc) ** 2 < (min(nz, nx) // 7) ** 2] *= 1.15
    cs = cp / vs_ratio
    rho = 310.0 * cp ** 0.25
    mu = rho * cs * cs
    return rho * cp * cp - 2 * mu, mu, rho


def small_case(O, *, nz=72, nx=96, nt=400, fdorder=8, nshot=3, surface=True, seed=7, snap_dt=2,
               uneven=False, npml=12, inclusion=True, fc=None):
    """A C1-shaped miniature: CPML on three sides and a free surface at the top
    (isurf = fpad), or CPML on all four sides.  ``uneven`` gives every shot its
    own receiver count and offset window (CSR rec_off)."""
    rng = np.random.default_rng(seed)
    lam, mu, rho = layered_model(nz, nx, seed, inclusion=inclusion)
    N = fdorder // 2
    hc = O.holberg(fdorder)
    dt = 0.5 * O.dtstab(lam, mu, rho, hc, 1.0, npml)
    if fc is None:
        fc = 300.0  # S wavelength ~ 4 cells at the slowest layer, above Holberg's 3.69 at order 8
    top = 0 if surface else npml
    xs = np.linspace(npml + 8, nx - npml - 8, nshot).astype(np.int32)
    zs = np.full(nshot, N + 2 if surface else npml + 3, np.int32)
    kw = {}
    if uneven:
        zr, xr, off = [], [], [0]
        for s in range(nshot):
            n = int(rng.integers(3, 11))
            lo = int(rng.integers(npml + 2, nx // 2))
            hi = int(min(nx - npml - 2, lo + rng.integers(n, nx // 2)))
            xr += list(np.linspace(lo, hi, n).astype(int))
            zr += [N + 1 if surface else npml + 1] * n
            off.append(off[-1] + n)
        kw = dict(z_rec=zr, x_rec=xr)
        rec_off = np.asarray(off, np.int32)
    else:
        kw = dict(z_rec=[N + 1 if surface else npml + 1] * 8, x_rec=list(np.linspace(npml + 2, nx - npml - 3, 8).astype(int)))
        rec_off = None
    c = O.make_case(nt=nt, nz=nz, nx=nx, dt=dt, dz=1.0, dx=1.0, fdorder=fdorder, lam=lam, mu=mu, rho=rho,
                    z_src=zs, x_src=xs, shot_to_fire=np.arange(nshot), fc=fc, amp_z=1.0, amp_x=0.3,
                    npml=(top, npml, npml, npml), isurf=(N if surface else 0, 0, 0, 0),
                    snap=(top, nz - npml, npml, nx - npml, snap_dt, 1, 1),
                    taper=(top + 2, top + 6, nz - npml - 2, nz - npml - 6, npml + 2, npml + 6,
                           nx - npml - 2, nx - npml - 6), **kw)
    if rec_off is not None:
        c["rec_off"] = rec_off
    return c


def adjoint_case(O, fdorder, surf, stf_type, nz=40, nx=52, nt=160, nshot=1, uneven=False):
    """A small case for the exact adjoint: a layered model with an inclusion, CPML or a free
    surface ("none", "top", or "four" sides with image rows and columns), shots 9 cells
    apart, five receivers (shot 1 of an uneven pair keeps three of them)."""
    N = fdorder // 2
    lam, mu, rho = layered_model(nz, nx, 11)
    hc = O.holberg(fdorder)
    npml = 8
    dt = 0.5 * O.dtstab(lam, mu, rho, hc, 1.0, npml)
    if surf == "none":
        isurf, pml = (0, 0, 0, 0), (npml, npml, npml, npml)
    elif surf == "top":
        isurf, pml = (N, 0, 0, 0), (0, npml, npml, npml)
    else:   # all four sides, with image rows / columns
        isurf, pml = (N + 2, nz - N - 3, N + 1, nx - N - 2), (0, 0, 0, 0)
    top = max(isurf[0], pml[0], N)
    zr = [top + 1, top + 2, nz // 2, top + 1, nz - 12]
    xr = [12, 20, 30, 40, 26]
    extra = {}
    if nshot > 1 and uneven:
        zr, xr = zr + zr[:3], xr + xr[:3]
        extra["rec_off"] = np.array([0, 5, 8], np.int32)
    c = O.make_case(nt=nt, nz=nz, nx=nx, dt=dt, dz=1.0, dx=1.0, fdorder=fdorder, lam=lam, mu=mu, rho=rho,
                    z_src=[top + 5 + s for s in range(nshot)], x_src=[nx // 2 - 3 + 9 * s for s in range(nshot)],
                    shot_to_fire=list(range(nshot)), fc=350.0, amp_z=1.0, amp_x=0.6,
                    stf_type=stf_type, npml=pml, isurf=isurf, z_rec=zr, x_rec=xr,
                    snap=(top, nz - 10, 10, nx - 10, 1, 1, 1))
    c.update(extra)
    return c


def start_model(c, seed=7):
    """The background without the inclusion (for gradient tests)."""
    return layered_model(int(c["nz"]), int(c["nx"]), seed, inclusion=False)


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def per_shot_obs(c, O, media):
    """Oracle forward seismograms per shot, receivers from rec_off when present."""
    out = []
    for s in range(int(c["nshot"])):
        cs = shot_case(c, s)
        f = O.kernel_psv(cs, media, s, accu=False)
        out.append((f["uz"], f["ux"]))
    return out


def matched_obs(c, O, media0, syn, obs):
    """Observed data for the oracle that gives it the GPU's residual.

    The adjoint source is d_syn - d_obs.  Where d_obs is FP64 and the residual is a
    small part of the data, the FP32 error of d_syn (rel ~1e-6) is a large part of
    the residual, and no FP32 forward can do better.  With
    d_obs' = d_syn_oracle + d_obs - d_syn_gpu the oracle's residual equals the
    GPU's, so the gradients compare the adjoint pass and the correlation alone.
    syn: per-shot (uz, ux) of the GPU at the start model media0.
    """
    out = []
    for s in range(int(c["nshot"])):
        f = O.kernel_psv(shot_case(c, s), media0, s, accu=False)
        out.append((f["uz"] + obs[s][0] - syn[s][0], f["ux"] + obs[s][1] - syn[s][1]))
    return out


def residual_fraction(syn, obs):
    """|d_syn - d_obs| / |d_obs| over every shot."""
    r = sum(np.sum((a - b) ** 2) for p, q in zip(syn, obs) for a, b in zip(p, q))
    d = sum(np.sum(b ** 2) for q in obs for b in q)
    return float(np.sqrt(r / d))


def shot_case(c, s):
    """The case restricted to shot s's receivers (the oracle's receivers are shared)."""
    if "rec_off" not in c:
        return c
    cs = dict(c)
    a, b = int(c["rec_off"][s]), int(c["rec_off"][s + 1])
    cs["z_rec"] = np.ascontiguousarray(c["z_rec"][a:b])
    cs["x_rec"] = np.ascontiguousarray(c["x_rec"][a:b])
    cs.pop("rec_off")
    return cs
