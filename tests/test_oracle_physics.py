"""Pins for the oracle's order-2N generalisation, which the reference (order 2
only, src/cpu/n_kernel_lib_PSV.cpp:125-128) cannot pin.

* staggered derivative exact on polynomials of degree <= 2N-1 (Taylor row);
* Taylor convergence rate h^(2N) on a sine;
* zero source -> identically zero field; linearity in the source amplitude;
* CPML absorbs: energy left in the grid after the wave has passed is small.
A wrong tap offset, a sign flip or a dropped coefficient fails one of these.
"""
import ctypes as C

import numpy as np
import pytest


def _grid_with_fields(O, nz, nx, dz, dx, fdorder, coeffs):
    c = O.make_case(nt=1, nz=nz, nx=nx, dt=1e-4, dz=dz, dx=dx, fdorder=fdorder,
                    lam=np.full((nz, nx), 1e9), mu=np.full((nz, nx), 5e8), rho=np.full((nz, nx), 2000.0),
                    z_src=[nz // 2], x_src=[nx // 2], shot_to_fire=[0], z_rec=[nz // 2], x_rec=[nx // 2],
                    npml=(0, 0, 0, 0))
    c["hc"] = coeffs
    return c


def _derivs(O, c, fields, fn="or_vdiff"):
    lib = O.load("oracle")
    keep = O._Keep()
    g = O._grid(c, keep)
    st = O.new_state(c)
    st.update({k: np.ascontiguousarray(v, np.float64) for k, v in fields.items()})
    s = O.OrState(*[keep.d(st[k]) for k in O.STATE_FIELDS])
    getattr(lib, fn)(C.byref(g), C.byref(s))
    return st


def _taylor(O, n):
    c = np.zeros(7)
    O.load("oracle").or_taylorcoeff(C.c_int32(n), O._d(c))
    return c


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 6])
def test_velocity_derivative_polynomial_exact(oracle_mod, N):
    O = oracle_mod
    nz = nx = 4 * N + 8
    h = 0.5
    c = _grid_with_fields(O, nz, nx, h, h, 2 * N, _taylor(O, N))
    z = np.arange(nz)[:, None] * h
    x = np.arange(nx)[None, :] * h
    p = 2 * N - 1
    # vz lives at (z + h/2, x), vx at (z, x + h/2)
    vz = ((z + h / 2) / 3.0) ** p + 0 * x
    vx = ((x + h / 2) / 3.0) ** p + 0 * z
    st = _derivs(O, c, dict(vz=vz, vx=vx))
    I = (slice(N, nz - N), slice(N, nx - N))
    dvz_dz = p * (z / 3.0) ** (p - 1) / 3.0 + 0 * x  # at integer z
    dvx_dx = p * (x / 3.0) ** (p - 1) / 3.0 + 0 * z  # at integer x
    np.testing.assert_allclose(st["dz_z"][I], dvz_dz[I], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(st["dx_x"][I], dvx_dx[I], rtol=1e-9, atol=1e-9)
    # cross terms of 1-D fields vanish
    assert np.abs(st["dz_x"][I]).max() < 1e-9 and np.abs(st["dx_z"][I]).max() < 1e-9


@pytest.mark.parametrize("N", [1, 2, 4])
def test_stress_derivative_polynomial_exact(oracle_mod, N):
    O = oracle_mod
    nz = nx = 4 * N + 8
    h = 1.0
    c = _grid_with_fields(O, nz, nx, h, h, 2 * N, _taylor(O, N))
    z = np.arange(nz)[:, None] * h
    x = np.arange(nx)[None, :] * h
    p = 2 * N - 1
    szz = (z / 4.0) ** p + 0 * x          # at integer points
    sxx = (x / 4.0) ** p + 0 * z
    szx = ((z + 0.5) / 4.0) ** p + ((x + 0.5) / 4.0) ** p   # at (z+1/2, x+1/2)
    st = _derivs(O, c, dict(szz=szz, sxx=sxx, szx=szx), fn="or_sdiff")
    I = (slice(N, nz - N), slice(N, nx - N))
    dz_half = p * ((z + 0.5) / 4.0) ** (p - 1) / 4.0 + 0 * x
    dx_half = p * ((x + 0.5) / 4.0) ** (p - 1) / 4.0 + 0 * z
    dz_int = p * (z / 4.0) ** (p - 1) / 4.0 + 0 * x
    dx_int = p * (x / 4.0) ** (p - 1) / 4.0 + 0 * z
    np.testing.assert_allclose(st["dz_z"][I], dz_half[I], rtol=1e-9, atol=1e-9)   # szz_z at z+1/2
    np.testing.assert_allclose(st["dx_z"][I], dz_int[I], rtol=1e-9, atol=1e-9)    # szx_z at z
    np.testing.assert_allclose(st["dz_x"][I], dx_int[I], rtol=1e-9, atol=1e-9)    # szx_x at x
    np.testing.assert_allclose(st["dx_x"][I], dx_half[I], rtol=1e-9, atol=1e-9)   # sxx_x at x+1/2


def test_taylor_convergence_order8(oracle_mod):
    O = oracle_mod
    errs = []
    for h in (0.2, 0.1):
        n = int(round(8.0 / h)) + 10
        c = _grid_with_fields(O, 12, n, 1.0, h, 8, _taylor(O, 4))
        x = np.arange(n)[None, :] * h + 0 * np.arange(12)[:, None]
        st = _derivs(O, c, dict(vx=np.sin(x + h / 2)))
        I = (slice(4, 8), slice(4, n - 4))
        errs.append(np.abs(st["dx_x"][I] - np.cos(x[I])).max())
    rate = np.log2(errs[0] / errs[1])
    assert 7.5 < rate < 8.6, (errs, rate)


def _small(O, fdorder, amp=1.0, npml=(8, 8, 8, 8), nt=220):
    nz, nx = 44, 52
    lam = np.full((nz, nx), 2200 * 3000.0 ** 2 - 2 * 2200 * 1730.0 ** 2)
    mu = np.full((nz, nx), 2200 * 1730.0 ** 2)
    rho = np.full((nz, nx), 2200.0)
    hc = O.holberg(fdorder)
    dt = 0.5 * O.dtstab(lam, mu, rho, hc, 1.0, 8)
    return O.make_case(nt=nt, nz=nz, nx=nx, dt=dt, dz=1.0, dx=1.0, fdorder=fdorder, lam=lam, mu=mu, rho=rho,
                       z_src=[22], x_src=[26], shot_to_fire=[0], z_rec=[12, 30], x_rec=[20, 34], fc=250.0,
                       amp_z=0.6 * amp, amp_x=amp, npml=npml)


@pytest.mark.parametrize("fdorder", [2, 8])
def test_zero_source_and_linearity(oracle_mod, fdorder):
    O = oracle_mod
    c0 = _small(O, fdorder, amp=0.0)
    f0 = O.kernel_psv(c0, O.mat_av2(c0), 0)
    assert not np.any(f0["uz"]) and not np.any(f0["state"]["vx"])
    c1, c2 = _small(O, fdorder, 1.0), _small(O, fdorder, 2.0)
    f1, f2 = O.kernel_psv(c1, O.mat_av2(c1), 0), O.kernel_psv(c2, O.mat_av2(c2), 0)
    np.testing.assert_allclose(f2["uz"], 2 * f1["uz"], rtol=1e-12, atol=1e-25)


def test_cpml_absorbs(oracle_mod):
    """After the pulse has crossed the grid, the energy left with CPML is a small
    fraction of the energy left with rigid edges."""
    O = oracle_mod
    e = {}
    for key, npml in (("pml", (10, 10, 10, 10)), ("rigid", (0, 0, 0, 0))):
        c = _small(O, 8, npml=npml, nt=700)
        f = O.kernel_psv(c, O.mat_av2(c), 0, accu=False)
        st = f["state"]
        e[key] = float((st["vx"] ** 2 + st["vz"] ** 2).sum())
    assert e["pml"] < 0.05 * e["rigid"], e


def test_reference_correlation_against_misfit_derivative(oracle_mod):
    """What the reference's gradient (fwi_grad2 with the CPU sign, restated by the
    oracle) is: for lam it is +dJ/dlam of J = sum 0.5 dt r^2 (directional
    derivatives along smooth bumps agree with central differences to a few
    per cent, so the CPU path's update m + eps f G climbs the misfit); its density
    term is not dJ/drho (1e-2 .. 1e-1 of it, either sign)."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).parent))
    import cases
    O = oracle_mod
    c = cases.small_case(O, nshot=1, nt=450, snap_dt=1)
    obs = cases.per_shot_obs(c, O, O.mat_av2(c))
    lam0, mu0, rho0 = cases.start_model(c)
    sg = O.shot_gradient(c, O.mat_av2(c, lam0, mu0, rho0), 0, *obs[0])
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    nz, nx = int(c["nz"]), int(c["nx"])
    zz, xx = np.mgrid[0:nz, 0:nx]
    ratios = {}
    for p, name in ((0, "lam"), (2, "rho")):
        G = np.zeros((nz, nx))
        G[z1:z2 + 1, x1:x2 + 1] = sg["raw"][p]
        r = []
        for zc, xc in ((42, 26), (44, 52), (11, 26)):
            base = [lam0, mu0, rho0]
            dm = 1e-3 * base[p] * np.exp(-((zz - zc) ** 2 + (xx - xc) ** 2) / 8.0)
            mp = [m.copy() for m in base]
            mm = [m.copy() for m in base]
            mp[p] += dm
            mm[p] -= dm
            fd = (O.misfit_all(c, *mp, obs) - O.misfit_all(c, *mm, obs)) / 2
            r.append(fd / float(np.sum(G * dm)))
        ratios[name] = np.array(r)
    assert np.all(np.abs(ratios["lam"] - 1) < 0.06), ratios
    assert np.all(np.abs(ratios["rho"]) < 0.2), ratios
