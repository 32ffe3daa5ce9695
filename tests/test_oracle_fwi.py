"""The oracle's FWI iteration and its cp / cs parameterisation, pinned without a GPU.

  - the chain rule to cp / cs against central differences of the parameter map;
  - update_model("lame") against the compiled reference's update_mat2, bit for bit;
  - update_model("velocity"): rho kept, bounds honoured, update_mat2's scaling of the
    largest change;
  - fwi_iteration's accumulation over shots and taper against the reference's own
    tapered global gradient (tests/golden/o2_pml.npz, made by the compiled reference);
    with the CPU sign the search fails as the reference's exit(0) does; the global
    gradient is never reset between iterations.
"""
import numpy as np
import pytest

from conftest import case_from_fixture, load_fixture


def _model(rng, shape=(9, 11)):
    cp = rng.uniform(2000.0, 4500.0, shape)
    cs = cp / rng.uniform(1.6, 2.0, shape)
    rho = 310.0 * cp ** 0.25
    mu = rho * cs * cs
    return rho * cp * cp - 2.0 * mu, mu, rho


def test_velocity_chain_rule_matches_finite_differences(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(21)
    lam, mu, rho = _model(rng)
    cp, cs = O.velocity_of(lam, mu, rho)
    np.testing.assert_allclose(rho * cs * cs, mu, rtol=1e-14)
    np.testing.assert_allclose(rho * cp * cp - 2.0 * rho * cs * cs, lam, rtol=1e-12)
    Gl, Gm = rng.standard_normal(lam.shape), rng.standard_normal(lam.shape)
    g_cp, g_cs = O.velocity_gradient(Gl, Gm, rho, cp, cs)

    def F(cp_, cs_):   # a functional linear in lam, mu, with dF/dlam = Gl, dF/dmu = Gm
        m = rho * cs_ * cs_
        return float(np.sum(Gl * (rho * cp_ * cp_ - 2.0 * m) + Gm * m))

    for iz, ix in [(0, 0), (3, 7), (8, 10), (5, 2)]:
        for which, g in (("cp", g_cp), ("cs", g_cs)):
            a, b = cp.copy(), cs.copy()
            base = (a if which == "cp" else b)
            h = 1e-5 * base[iz, ix]
            base[iz, ix] += h
            fp = F(a, b)
            base[iz, ix] -= 2 * h
            fm = F(a, b)
            assert (fp - fm) / (2 * h) == pytest.approx(g[iz, ix], rel=1e-6, abs=1e-9 * np.abs(g).max())


def test_update_model_lame_is_the_reference_update(oracle_mod):
    O = oracle_mod
    if not O.ref_available():
        pytest.skip("oracle/_ref is not built (make -C oracle ref)")
    rng = np.random.default_rng(22)
    old = _model(rng)
    G = [rng.standard_normal(old[0].shape) * s for s in (1e3, 1e3, 1e-3)]
    new = O.update_model(old, G, 0.02, param="lame", rho_factor=0.5, rho_min=1.25)
    want = (O.update_mat2(old[0], G[0], O.LAM_MAX, 0.0, 0.02, which="ref"),
            O.update_mat2(old[1], G[1], O.MU_MAX, 0.0, 0.02, which="ref"),
            O.update_mat2(old[2], G[2], O.RHO_MAX, 1.25, 0.01, which="ref"))
    for a, b in zip(new, want):
        np.testing.assert_array_equal(a, b)


def test_update_model_velocity_keeps_rho_and_scales_like_update_mat2(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(23)
    old = _model(rng)
    G = [rng.standard_normal(old[0].shape) for _ in range(3)]
    step = 0.01
    lam, mu, rho = O.update_model(old, G, step, param="velocity")
    np.testing.assert_array_equal(rho, old[2])
    cp0, cs0 = O.velocity_of(*old)
    cp1, cs1 = O.velocity_of(lam, mu, rho)
    g_cp, g_cs = O.velocity_gradient(G[0], G[1], old[2], cp0, cs0)
    for v0, v1, g, vmax in ((cp0, cp1, g_cp, O.VEL_BOUNDS[1]), (cs0, cs1, g_cs, O.VEL_BOUNDS[3])):
        # update_mat2: the largest |gradient| cell moves by step * max(max|v|, v_max), signed with g
        k = np.unravel_index(np.argmax(np.abs(g)), g.shape)
        assert v1[k] - v0[k] == pytest.approx(np.sign(g[k]) * step * max(np.abs(v0).max(), vmax), rel=1e-9)
        assert np.all(np.abs(v1 - v0) <= step * max(np.abs(v0).max(), vmax) * (1 + 1e-9))
    # bounds clip
    lam, mu, rho = O.update_model(old, G, 10.0, param="velocity")
    cp2, cs2 = O.velocity_of(lam, mu, rho)
    assert cp2.min() >= O.VEL_BOUNDS[0] * (1 - 1e-12) and cp2.max() <= O.VEL_BOUNDS[1] * (1 + 1e-12)
    assert cs2.min() >= O.VEL_BOUNDS[2] * (1 - 1e-12) and cs2.max() <= O.VEL_BOUNDS[3] * (1 + 1e-12)


def test_fwi_iteration_on_the_reference_fixture(oracle_mod):
    """fwi_iteration's global gradient over the two shots, tapered, is the compiled
    reference's (o2_pml G_tapered); with the CPU sign (an ascent direction here) the
    step search fails where the reference exits; with the GPU twin's sign it finds a
    step that lowers the misfit, and a second iteration adds to the tapered gradient
    of the first instead of starting from zero."""
    O = oracle_mod
    fx = load_fixture("o2_pml")
    c = case_from_fixture(fx)
    out = fx["out"]
    obs = [(out[f"seis{s}_uz"], out[f"seis{s}_ux"]) for s in range(2)]
    m0 = (fx["in0"]["lam"], fx["in0"]["mu"], fx["in0"]["rho"])
    nz, nx = m0[0].shape
    G = [np.zeros((nz, nx)) for _ in range(3)]
    with pytest.raises(RuntimeError):
        O.fwi_iteration(c, m0, G, obs, 0.01, grad_sign=-1)
    for i in range(3):
        np.testing.assert_allclose(G[i], out["G_tapered"][str(i)], rtol=1e-13, atol=0)

    G = [np.zeros((nz, nx)) for _ in range(3)]
    L2, step, m1, trials = O.fwi_iteration(c, m0, G, obs, 0.01, grad_sign=+1)
    assert L2 == pytest.approx(float(out["L2_0"]), rel=1e-13)
    for i in range(3):
        np.testing.assert_allclose(G[i], -out["G_tapered"][str(i)], rtol=1e-13, atol=0)
    assert step > 0 and len(trials) >= 2
    assert O.misfit_all(c, *m1, obs) < L2
    G1 = [g.copy() for g in G]
    L2b, _, _, _ = O.fwi_iteration(c, m1, G, obs, step, grad_sign=+1)
    assert L2b < L2
    med1 = O.mat_av2(c, *m1)
    own = [np.zeros((nz, nx)) for _ in range(3)]
    for s in range(2):
        sg = O.shot_gradient(c, med1, s, *obs[s])
        for i, k in enumerate(("g_lam", "g_mu", "g_rho")):
            own[i] -= sg[k]
    for i in range(3):
        want = O.taper2(c, G1[i] + own[i])
        np.testing.assert_allclose(G[i], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        alone = O.taper2(c, own[i].copy())
        assert np.linalg.norm(G[i] - alone) > 1e-3 * np.linalg.norm(alone)


def test_pcg_matches_the_reference(oracle_mod):
    """PCG_PSV (src/cpu/n_PCG_PSV.cpp:13-44): with a previous direction in place, the
    restatement and the compiled reference agree bit for bit over two calls; on the first
    call (old = 0) the reference divides by zero and returns NaN, the restatement takes
    beta = 0 (the steepest-descent direction -g)."""
    O = oracle_mod
    if not O.ref_available():
        pytest.skip("oracle/_ref is not built (make -C oracle ref)")
    rng = np.random.default_rng(31)
    shape = (7, 9)
    d0, o0 = rng.standard_normal(shape), rng.standard_normal(shape)
    gs = [rng.standard_normal(shape) for _ in range(2)]
    outs = {}
    for which in ("oracle", "ref"):
        d, o = d0.copy(), o0.copy()
        got = []
        for g0 in gs:
            g = g0.copy()
            O.pcg_psv(d, o, g, which=which)
            got.append((d.copy(), o.copy(), g.copy()))
        outs[which] = got
    for a, b in zip(outs["oracle"], outs["ref"]):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    g = rng.standard_normal(shape)
    d, o = np.zeros(shape), np.zeros(shape)
    O.pcg_psv(d, o, g.copy(), which="ref")
    assert np.isnan(d).all()
    d, o, gg = np.zeros(shape), np.zeros(shape), g.copy()
    O.pcg_psv(d, o, gg)
    np.testing.assert_array_equal(gg, -g)
    np.testing.assert_array_equal(o, -g)


@pytest.mark.parametrize("hfs", [2, 3, 5])
def test_gauss_filter_kernel_and_smoothing(oracle_mod, hfs):
    """gauss_filter_kernel against the compiled reference (bit for bit; sigma = hfs / 2 in
    integer division, normalised); apply_gauss_filter against a direct numpy correlation
    over the reference's index range.  The reference's own apply_gauss_filter accumulates in
    an int, so a gradient-sized field (|values| < 1) comes out 0 wherever it filters."""
    O = oracle_mod
    if not O.ref_available():
        pytest.skip("oracle/_ref is not built (make -C oracle ref)")
    k = O.gauss_filter_kernel(hfs)
    np.testing.assert_array_equal(k, O.gauss_filter_kernel(hfs, which="ref"))
    assert k.sum() == pytest.approx(1.0, rel=1e-14)
    np.testing.assert_array_equal(k, k.T)
    rng = np.random.default_rng(hfs)
    A = rng.standard_normal((24, 31)) * 1e-3
    got = O.apply_gauss_filter(A.copy(), hfs)
    want = A.copy()
    for iz in range(hfs, 24 - hfs - 1):
        for ix in range(hfs, 31 - hfs - 1):
            want[iz, ix] = np.sum(A[iz - hfs:iz + hfs + 1, ix - hfs:ix + hfs + 1] * k)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-18)
    ref = O.apply_gauss_filter(A.copy(), hfs, which="ref")
    assert np.all(ref[hfs:24 - hfs - 1, hfs:31 - hfs - 1] == 0.0)
    np.testing.assert_array_equal(ref[:hfs], A[:hfs])
