"""The oracle's exact discrete adjoint, pinned by what the mathematics fixes.

  - dot-product test: for the forward map F (media -> receiver displacement
    seismograms), <(F(m + h dm) - F(m - h dm)) / 2h, y> against <dm, F'(m)^T y> from
    the adjoint, FP64, to 1e-7 relative, at orders 2 and 8, with and without a free
    surface (top only, or all four sides with image rows), CPML, and type-0 / type-1
    sources (a type-0 source's amplitude carries the inverse density of its cell);
  - the chain rule through mat_av2 against central differences of mat_av2 itself.
"""
import numpy as np
import pytest

import cases


@pytest.mark.parametrize("fdorder,surf,stf_type", [(2, "top", 1), (8, "top", 0), (8, "none", 1), (4, "four", 0)])
def test_exact_adjoint_dot_product(oracle_mod, fdorder, surf, stf_type):
    O = oracle_mod
    c = cases.adjoint_case(O, fdorder, surf, stf_type)
    med = O.mat_av2(c)
    rng = np.random.default_rng(5)
    nz, nx, nt = int(c["nz"]), int(c["nx"]), int(c["nt"])
    keys = ("lam", "mu", "mu_zx", "rho_zp", "rho_xp")
    dm = {}
    from scipy.ndimage import gaussian_filter
    for k in keys:
        d = gaussian_filter(rng.standard_normal((nz, nx)), 2.0)
        dm[k] = d / np.abs(d).max() * np.abs(med[k]).max()
    nrec = len(c["z_rec"])
    y = (rng.standard_normal((nrec, nt)), rng.standard_normal((nrec, nt)))

    def F(sign, h):
        m = dict(med)
        for k in keys:
            m[k] = np.ascontiguousarray(med[k] + sign * h * dm[k])
        f = O.kernel_psv(c, m, 0, accu=False)
        return f["uz"], f["ux"]

    h = 1e-6
    p, q = F(1, h), F(-1, h)
    lhs = float(np.sum((p[0] - q[0]) / (2 * h) * y[0]) + np.sum((p[1] - q[1]) / (2 * h) * y[1]))
    g = O.adjoint_exact(c, med, 0, *y)
    rhs = float(sum(np.sum(dm[k] * g[k]) for k in keys))
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"order {fdorder} surface {surf} stf {stf_type}: <J dm, y> {lhs:.12e}, <dm, J^T y> {rhs:.12e}, rel {rel:.2e}")
    assert rel < 1e-7
    # each media part on its own (a wrong term in one would hide in the sum)
    for k in keys:
        m1, m2 = dict(med), dict(med)
        m1[k] = np.ascontiguousarray(med[k] + h * dm[k])
        m2[k] = np.ascontiguousarray(med[k] - h * dm[k])
        a, b = O.kernel_psv(c, m1, 0, accu=False), O.kernel_psv(c, m2, 0, accu=False)
        l_k = float(np.sum((a["uz"] - b["uz"]) / (2 * h) * y[0]) + np.sum((a["ux"] - b["ux"]) / (2 * h) * y[1]))
        r_k = float(np.sum(dm[k] * g[k]))
        assert abs(l_k - r_k) <= 1e-6 * abs(lhs), (k, l_k, r_k)


def test_media_chain_matches_mat_av2_differences(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(9)
    nz, nx = 9, 12
    lam, mu, rho = cases.layered_model(nz, nx, 3)
    mu = mu * rng.uniform(0.8, 1.2, (nz, nx))
    rho = rho * rng.uniform(0.8, 1.2, (nz, nx))
    G = {k: rng.standard_normal((nz, nx)) for k in ("lam", "mu", "mu_zx", "rho_zp", "rho_xp")}
    c = dict(nz=nz, nx=nx)
    med = O.mat_av2(c, lam, mu, rho)

    def F(p, l, m, r):   # the terms that depend on parameter p only (their scales differ by 1e14)
        md = O.mat_av2(c, l, m, r)
        if p == 0:
            return float(np.sum(G["lam"] * l))
        if p == 1:
            return float(np.sum(G["mu"] * m) + np.sum((G["mu_zx"] * md["mu_zx"])[:-1, :-1]))
        return float(np.sum((G["rho_zp"] * md["rho_zp"])[:-1, :-1]) + np.sum((G["rho_xp"] * md["rho_xp"])[:-1, :-1]))

    G2 = dict(G)
    for k in ("mu_zx", "rho_zp", "rho_xp"):   # the chain rule never reads the undefined edge
        G2[k] = G[k].copy()
        G2[k][-1, :] = 0.0
        G2[k][:, -1] = 0.0
    d = O.media_chain(G2, med)
    base = [lam, mu, rho]
    for p in range(3):
        for iz, ix in [(0, 0), (4, 5), (8, 11), (3, 10), (8, 2)]:
            hh = 1e-6 * base[p][iz, ix]
            up = [a.copy() for a in base]
            dn = [a.copy() for a in base]
            up[p][iz, ix] += hh
            dn[p][iz, ix] -= hh
            fd = (F(p, *up) - F(p, *dn)) / (2 * hh)
            assert d[p][iz, ix] == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(d[p]).max()), (p, iz, ix)
