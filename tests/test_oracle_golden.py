"""The FP64 restatement (oracle/) against the reference compiled in this
container (tests/golden/*.npz) and against closed forms.

Bit-exact (==) where the reference's operation order is deterministic; the
only tolerance is on mat_av2's scalar means, which the reference sums with an
OpenMP reduction in scheduler order (src/cpu/n_kernel_lib_PSV.cpp:866-909).
"""
import numpy as np
import pytest

from conftest import case_from_fixture, load_fixture


def test_holberg_table(oracle_mod):
    fx = load_fixture("pieces")
    for e in range(5):
        for n in range(6):
            np.testing.assert_array_equal(oracle_mod.holberg(2 * (n + 1), e), fx["holberg"][e, n])


def test_taylor_row_is_closed_form(oracle_mod):
    """Row 0 of the table equals the Taylor closed form (an independent derivation)."""
    import ctypes as C
    lib = oracle_mod.load("oracle")
    fx = load_fixture("pieces")
    for n in range(1, 7):
        c = np.zeros(7)
        lib.or_taylorcoeff(C.c_int32(n), oracle_mod._d(c))
        np.testing.assert_allclose(c[1:], fx["holberg"][0, n - 1, 1:], rtol=1e-15, atol=1e-17)
        # consistency: sum_k c_k (2k-1) = 1 (exact first derivative of a linear function)
        assert abs(sum(c[k] * (2 * k - 1) for k in range(1, n + 1)) - 1.0) < 1e-14


def test_holberg_rows_are_consistent_first_derivatives():
    """Each Holberg row approximates d/dx: sum_k c_k (2k-1) within the design error
    (the first-order consistency condition), so a mistyped entry fails."""
    fx = load_fixture("pieces")
    tol = {1: 2e-3, 2: 1e-2, 3: 2e-2, 4: 6e-2}
    for e in (1, 2, 3, 4):
        for n in range(1, 7):
            c = fx["holberg"][e, n - 1]
            s = sum(c[k] * (2 * k - 1) for k in range(1, n + 1))
            assert abs(s - 1.0) < tol[e], (e, n, s)


def test_cpml_profiles(oracle_mod):
    fx = load_fixture("pieces")
    for key, ref in fx["cpml"].items():
        a = ref["args"]
        args = (a[0], a[1], a[2], a[3], a[4], int(a[5]), int(a[6]), int(a[7]), int(a[8]), a[9], a[10])
        got = oracle_mod.cpml(*args)
        for k in ("a", "b", "K", "a_half", "b_half", "K_half"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"cpml {key} {k}")


def test_cpml_invariants(oracle_mod):
    """Outside the strips a = 0, b = 1, K = 1; inside 0 < b <= 1 and a <= 0;
    the profile grows towards the outer edge (d ~ x^N)."""
    p = oracle_mod.cpml(2.0, 3300.0, 1e-3, 1.0, 100.0, 20, 20, 4, 200, 8e-5, 1.0)
    mid = slice(4 + 21, 200 - 4 - 22)
    assert np.all(p["a"][mid] == 0) and np.all(p["b"][mid] == 1) and np.all(p["K"][mid] == 1)
    assert np.all((p["b"] > 0) & (p["b"] <= 1)) and np.all(p["a"] <= 0)
    left = p["a"][4:24]
    assert np.all(np.diff(np.abs(left)) <= 0)  # |a| largest at the outer edge


def test_wavelet(oracle_mod):
    fx = load_fixture("pieces")
    n, dt, amp, fc, ts = fx["wavelet"]["args"]
    s = oracle_mod.wavelet(int(n), dt, amp, fc, ts)
    np.testing.assert_array_equal(s, fx["wavelet"]["s"])
    # closed form with the reference's truncated pi
    t = np.arange(int(n)) * dt
    tau = 3.14159265 * (t - 1.5 / fc - ts) / (1.5 / fc)
    np.testing.assert_allclose(s, amp * (1 - 2 * tau ** 2) * np.exp(-2 * tau ** 2), rtol=1e-14, atol=1e-16)


def test_mat_av2(oracle_mod):
    fx = load_fixture("pieces")["mat"]
    c = dict(nz=fx["lam"].shape[0], nx=fx["lam"].shape[1])
    m = oracle_mod.mat_av2(c, fx["lam"], fx["mu"], fx["rho"])
    nz, nx = fx["lam"].shape
    core = (slice(0, nz - 1), slice(0, nx - 1))  # the reference leaves the last row/col unwritten
    for k in ("mu_zx", "rho_zp", "rho_xp"):
        np.testing.assert_array_equal(m[k][core], fx[k][core], err_msg=k)
    np.testing.assert_allclose(m["C"], fx["C"], rtol=1e-14)
    # closed form on the zero-mu cell and the zero-density guard quirk
    assert m["mu_zx"][5, 7] == 0 and m["mu_zx"][4, 6] == 0
    assert m["rho_zp"][9, 10] == 0  # written by the rho_xp guard (src/cpu/n_kernel_lib_PSV.cpp:894-897)


def test_taper_update_energy_interp_scale_adjsrc(oracle_mod):
    fx = load_fixture("pieces")
    c = dict(snap=fx["taper"]["snap"], taper=fx["taper"]["taper"], nz=30, nx=41)
    np.testing.assert_array_equal(oracle_mod.taper2(c, fx["taper"]["A"].copy()), fx["taper"]["out"])
    u = fx["update"]
    np.testing.assert_array_equal(oracle_mod.update_mat2(u["old"], u["grad"], 8e9, 2.5e9, 0.03), u["out"])
    e = fx["energy"]
    np.testing.assert_array_equal(oracle_mod.energy_weights2(c, e["We"], e["Wa"]), e["out"])
    np.testing.assert_array_equal(oracle_mod.interpol_grad2(c, fx["interp"]["gs"]), fx["interp"]["out"])
    s = fx["scale"]
    got = oracle_mod.scale_grad_E2(c, s["G"].copy(), s["gs"], float(s["mat_av"]), s["We"])
    np.testing.assert_array_equal(got, s["out"])
    a = fx["adjsrc"]
    ruz, rux, L2 = oracle_mod.adjsrc2(a["muz"], a["mux"], a["tuz"], a["tux"], 3e-5)
    np.testing.assert_array_equal(ruz, a["ruz"])
    np.testing.assert_array_equal(rux, a["rux"])
    assert L2 == float(a["L2"])
    # closed form of the misfit
    assert abs(L2 - 0.5 * 3e-5 * ((a["muz"] - a["tuz"]) ** 2 + (a["mux"] - a["tux"]) ** 2).sum()) < 1e-12 * L2


@pytest.mark.parametrize("name", ["o2_pml", "o2_surf"])
def test_forward_bitexact(oracle_mod, name):
    fx = load_fixture(name)
    c = case_from_fixture(fx)
    out = fx["out"]
    med = oracle_mod.mat_av2(c)
    for k in ("mu_zx", "rho_zp", "rho_xp"):
        np.testing.assert_array_equal(med[k][:-1, :-1], out["media_true"][k][:-1, :-1])
    for s in range(2):
        f = oracle_mod.kernel_psv(c, med, s)
        np.testing.assert_array_equal(f["uz"], out[f"seis{s}_uz"])
        np.testing.assert_array_equal(f["ux"], out[f"seis{s}_ux"])
    for k, v in out["state1"].items():
        np.testing.assert_array_equal(f["state"][k], v, err_msg=k)
    pick = out["snap_pick"]
    for k, v in out["snaps1"].items():
        np.testing.assert_array_equal(f["snaps"][k][pick], v, err_msg=k)


@pytest.mark.parametrize("name", ["o2_pml", "o2_surf"])
def test_gradient_pipeline_bitexact(oracle_mod, name):
    fx = load_fixture(name)
    c = case_from_fixture(fx)
    out = fx["out"]
    obs = [(out[f"seis{s}_uz"], out[f"seis{s}_ux"]) for s in range(2)]
    med0 = oracle_mod.mat_av2(c, fx["in0"]["lam"], fx["in0"]["mu"], fx["in0"]["rho"])
    sg = oracle_mod.shot_gradient(c, med0, 1, *obs[1])
    g = out["grad1"]
    assert sg["L2"] == float(g["L2"])
    for i, k in enumerate(("raw_lam", "raw_mu", "raw_rho")):
        np.testing.assert_array_equal(sg["raw"][i], g[k], err_msg=k)
    for k in ("We_fwd", "We_adj", "We"):
        np.testing.assert_array_equal(sg[k], g[k], err_msg=k)
    # the scalar means enter here; the reference sums them in OpenMP order
    for k in ("g_lam", "g_mu", "g_rho"):
        np.testing.assert_allclose(sg[k], g[k], rtol=1e-13, atol=0, err_msg=k)


@pytest.mark.parametrize("name", ["o2_pml", "o2_surf"])
def test_step_length_matches_reference(oracle_mod, name):
    fx = load_fixture(name)
    c = case_from_fixture(fx)
    out = fx["out"]
    obs = [(out[f"seis{s}_uz"], out[f"seis{s}_ux"]) for s in range(2)]
    D = [out["step_dir"][str(i)] for i in range(3)]
    step, trials, _ = oracle_mod.step_length(c, 0.01, float(out["L2_0"]), fx["in0"]["lam"], fx["in0"]["mu"],
                                             fx["in0"]["rho"], D, obs)
    assert step == pytest.approx(float(out["step"]), rel=1e-12)
    assert len(trials) >= 2


def test_parabola_known_answers(oracle_mod):
    """Vertices the reference printed in its own CPU log (Output2.txt:209-211,
    340-342, 447-449, 557-559) and the monotone clamp (Output.txt:41-43)."""
    cases = [([9.6096e-07, 8.89441e-07, 8.94124e-07], [0, 0.01, 0.050625], 0.029911)]
    for L2, sl, want in cases:
        assert oracle_mod.calc_opt_step(L2, sl) == pytest.approx(want, rel=2e-5)
    L2 = [2.08539e-11, 2.07842e-11, 2.07123e-11]
    sl = [0, 7.8125e-05, 0.000593262]
    assert oracle_mod.calc_opt_step(L2, sl) == sl[2]


def test_parabola_closed_form(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(20):
        a, e0 = rng.uniform(0.5, 2), rng.uniform(0.01, 0.05)
        sl = [0.0, rng.uniform(0.001, 0.02), rng.uniform(0.03, 0.08)]
        L2 = [a * (s - e0) ** 2 + 1.0 for s in sl]
        got = oracle_mod.calc_opt_step(L2, sl)
        want = min(max(e0, 0), sl[2])
        if L2[1] < L2[0] and L2[2] < L2[1]:
            want = sl[2]
        assert got == pytest.approx(want, rel=1e-9)
