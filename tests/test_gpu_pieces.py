"""HIP path for the pieces of the gradient that the main parity tests leave at their
identity settings: taper2 on the device, the strided snapshot lattice with its
interpol_grad2, and free surfaces on the bottom, left and right sides.

Tolerances:
  taper2 (FP64 on the device)                 rel 1e-12 against the reference's own output
  seismograms, order 8                        rel-L2 <= 1e-4 against the FP64 oracle
  final stresses and velocities               rel-L2 <= 5e-4 (the field has partly left through
                                              the CPML: its FP32 error is relative to less energy)
  raw and preconditioned gradients            rel-L2 <= 2e-3 against the FP64 oracle (the
                                              residual is about 1e-3 of the data here)
"""
import numpy as np
import pytest

import cases
from conftest import gpu_available, load_fixture

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


@pytest.fixture(scope="module")
def psv():
    from fwi_cuda_amd import psv as P
    return P


@pytest.fixture(scope="module")
def O():
    from oracle import oracle as O_
    O_.load("oracle")
    return O_


def _print(*a):
    print(*a, flush=True)


def test_taper_matches_reference_output(psv, O):
    """taper2 of the reference (src/cpu/n_kernel_lib_PSV.cpp:942-1010), run by the
    compiled reference on a 30 x 41 array with an uneven snap box and ramp widths
    (tests/golden/pieces.npz), against fwi_taper_gradient on the three gradient
    arrays; and against the oracle on a second, seeded array and box."""
    p = load_fixture("pieces")["taper"]
    A, snap, taper, want = p["A"], p["snap"], p["taper"], p["out"]
    nz, nx = A.shape
    c = cases.small_case(O, nz=nz, nx=nx, nt=4, fdorder=2, nshot=1, npml=4, surface=False)
    c["snap"] = np.array([snap[0], snap[1], snap[2], snap[3], 1, 1, 1], np.int32)   # taper2 ignores strides
    c["taper"] = np.asarray(taper, np.int32)
    with psv.Simulation(c) as sim:
        sim.set_gradient(A, 2.0 * A, -A)
        sim.taper2()
        G = sim.get_gradient()
    np.testing.assert_allclose(G[0], want, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(G[1], 2.0 * want, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(G[2], -want, rtol=1e-12, atol=1e-300)
    assert np.count_nonzero(want == 0.0) > 0 and np.count_nonzero((want != A) & (want != 0.0)) > 0

    # a second box with ramps on every side, on the oracle
    c = cases.small_case(O, nz=64, nx=90, nt=4, fdorder=8, nshot=1, npml=10, surface=True)
    c["taper"] = np.array([6, 14, 50, 41, 13, 21, 77, 70], np.int32)
    B = np.random.default_rng(11).standard_normal((64, 90))
    want = O.taper2(c, B.copy())
    with psv.Simulation(c) as sim:
        sim.set_gradient(B, B, B)
        sim.taper2()
        G = sim.get_gradient()
    for g in G:
        np.testing.assert_allclose(g, want, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("variant", [1, 5, 7, 8, 9, 10, 11, 12, 13])
def test_strided_snapshot_gradient_vs_oracle(psv, O, variant):
    """Snapshots on a lattice of every 2nd row and 3rd column (snap_dz = 2,
    snap_dx = 3, snap_dt = 2, the strides of the golden interpol_grad2 piece): the
    correlation is taken on the lattice only and interpol_grad2
    (src/cpu/n_kernel_lib_PSV.cpp:628-693) fills the box between lattice points on
    the device.  Raw sums on the lattice and the preconditioned gradient against
    the FP64 oracle (its own or_interpol_grad2), two shots in one batch."""
    c = cases.small_case(O, nshot=2, nt=450)
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    assert (z2 - z1) % 2 == 0 and (x2 - x1) % 3 == 0
    c["snap"] = np.array([z1, z2, x1, x2, 2, 2, 3], np.int32)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    med0 = O.mat_av2(c, lam0, mu0, rho0)
    G = [np.zeros((int(c["nz"]), int(c["nx"]))) for _ in range(3)]
    L2 = 0.0
    for s in range(2):
        sg = O.shot_gradient(c, med0, s, *obs[s])
        L2 += sg["L2"]
        for i, k in enumerate(("g_lam", "g_mu", "g_rho")):
            G[i] += sg[k]
    # raw lattice sums of shot 1 for given adjoint sources (the residual of the start model)
    fwd = O.kernel_psv(c, med0, 1, grad=False, accu=True)
    rz, rx = fwd["uz"] - obs[1][0], fwd["ux"] - obs[1][1]
    adj = O.kernel_psv(c, med0, 1, grad=True, accu=False, snaps=fwd["snaps"], adj=(rz, rx))
    raw = [adj["g_lam"], adj["g_mu"], adj["g_rho"]]
    tz = 40 if variant in (7, 8) else 0
    with psv.Simulation(c, kernel_variant=variant, tile_tz=tz) as sim:
        sim.set_model(lam0, mu0, rho0)
        sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
        m = sim.gradient()
        Gg = sim.get_gradient()
        rg = sim.gradient_raw(rz, rx, 1, 1)
    assert m.sum() == pytest.approx(L2, rel=1e-4)
    assert rg[0].shape == raw[0].shape == (1 + (z2 - z1) // 2, 1 + (x2 - x1) // 3)
    for i, k in enumerate(("lam", "mu", "rho")):
        e = cases.rel_l2(Gg[i], G[i])
        er = cases.rel_l2(rg[i], raw[i])
        _print(f"variant {variant} strided lattice {k}: preconditioned rel-L2 {e:.2e}, raw lattice sums {er:.2e}")
        assert e < 2e-3 and er < 2e-3
        # between lattice points the gradient is the linear interpolation of its neighbours
        zz = z1 + 1
        np.testing.assert_allclose(Gg[i][zz, x1 + 3:x2 - 3], G[i][zz, x1 + 3:x2 - 3],
                                   atol=2e-3 * np.abs(G[i]).max())


def _surface_case(O, isurf, npml, nt=360, nshot=2):
    """A 72 x 96 order-8 case with free surfaces at the given rows / columns and
    CPML on the sides given by npml (top, bottom, left, right); sources and
    receivers inside the region the surfaces enclose."""
    nz, nx, N = 72, 96, 4
    lam, mu, rho = cases.layered_model(nz, nx, 5)
    hc = O.holberg(8)
    dt = 0.5 * O.dtstab(lam, mu, rho, hc, 1.0, max(npml))
    top = isurf[0] if isurf[0] else max(npml[0], N)
    bot = isurf[1] if isurf[1] else nz - max(npml[1], N) - 1
    left = isurf[2] if isurf[2] else max(npml[2], N)
    right = isurf[3] if isurf[3] else nx - max(npml[3], N) - 1
    xs = np.linspace(left + 12, right - 12, nshot).astype(np.int32)
    zs = np.full(nshot, top + 6, np.int32)
    zr = [top + 2] * 6 + [bot - 2] * 4
    xr = list(np.linspace(left + 3, right - 3, 6).astype(int)) + list(np.linspace(left + 5, right - 5, 4).astype(int))
    zr += list(np.linspace(top + 4, bot - 4, 4).astype(int)) * 2
    xr += [left + 2] * 4 + [right - 2] * 4
    snap = (top, bot, left, right, 2, 1, 1)
    return O.make_case(nt=nt, nz=nz, nx=nx, dt=dt, dz=1.0, dx=1.0, fdorder=8, lam=lam, mu=mu, rho=rho,
                       z_src=zs, x_src=xs, shot_to_fire=np.arange(nshot), fc=300.0, amp_z=1.0, amp_x=0.5,
                       npml=npml, isurf=isurf, z_rec=zr, x_rec=xr, snap=snap,
                       taper=(top + 2, top + 6, bot - 2, bot - 6, left + 2, left + 6, right - 2, right - 6))


SURFACES = {
    # bottom, left and right surfaces with their image rows / columns, top CPML
    "bottom-left-right": ((0, 64, 7, 88), (10, 0, 0, 0)),
    # all four sides, no CPML (rigid edges beyond the image rows)
    "all-four": ((6, 65, 5, 90), (0, 0, 0, 0)),
    # a bottom surface at the last interior row (no image rows inside the interior)
    "bottom-edge": ((4, 67, 0, 0), (0, 0, 10, 10)),
}


@pytest.mark.parametrize("variant", [1, 2, 5, 6, 10, 11, 12, 13])
@pytest.mark.parametrize("name", sorted(SURFACES))
def test_free_surfaces_on_every_side_vs_oracle(psv, O, name, variant):
    """surf_mirror on the bottom, left and right sides (src/cpu/n_kernel_lib_PSV.cpp:317-394):
    the tangential stress set (not incremented) on the surface line, the normal and
    shear stresses zeroed, and the negated images between the interior edge and the
    surface; seismograms of receivers near every surface and the final stresses
    against the FP64 oracle.  The one-pass variants refuse these grids."""
    isurf, npml = SURFACES[name]
    c = _surface_case(O, isurf, npml)
    med = O.mat_av2(c)
    ref = [O.kernel_psv(c, med, s, accu=False) for s in range(2)]
    with psv.Simulation(c, kernel_variant=variant) as sim:
        uz, ux = sim.forward()
        for s, (a, b) in enumerate(zip(sim.split(uz), sim.split(ux))):
            ez, ex = cases.rel_l2(a, ref[s]["uz"]), cases.rel_l2(b, ref[s]["ux"])
            _print(f"surfaces {name} variant {variant} shot {s}: rel-L2 uz {ez:.2e} ux {ex:.2e}")
            assert ez < 1e-4 and ex < 1e-4
        for k in ("szz", "szx", "sxx", "vz", "vx"):
            e = cases.rel_l2(sim.wavefield(1, k), ref[1]["state"][k])
            assert e < 5e-4, (k, e)
        # the surface line carries the assigned stress and zeros
        st = {k: sim.wavefield(1, k) for k in ("szz", "szx", "sxx")}
    if isurf[1]:   # (the left and right sides, processed after it, set szz where they cross it)
        assert np.all(st["szx"][isurf[1], 4:-4] == 0)
    if isurf[2]:
        assert np.all(st["sxx"][4:-4, isurf[2]] == 0) and np.all(st["szx"][4:-4, isurf[2]] == 0)


@pytest.mark.parametrize("variant", [1, 6, 10, 11, 12, 13])
def test_free_surfaces_gradient_vs_oracle(psv, O, variant):
    """A gradient with surfaces on the bottom, left and right (the adjoint pass
    applies them too): misfit and preconditioned gradient against the FP64 oracle, to
    1e-4 given the GPU's own residual, and end to end within what the residual's FP32
    error alone does to the oracle's gradient (the start model differs from the true one
    by the inclusion only, so the residual is a small part of the data)."""
    isurf, npml = SURFACES["bottom-left-right"]
    c = _surface_case(O, isurf, npml, nt=420)
    obs = cases.per_shot_obs(c, O, O.mat_av2(c))
    lam0, mu0, rho0 = cases.layered_model(int(c["nz"]), int(c["nx"]), 5, inclusion=False)
    med0 = O.mat_av2(c, lam0, mu0, rho0)
    G = [np.zeros((int(c["nz"]), int(c["nx"]))) for _ in range(3)]
    L2 = 0.0
    for s in range(2):
        sg = O.shot_gradient(c, med0, s, *obs[s])
        L2 += sg["L2"]
        for i, k in enumerate(("g_lam", "g_mu", "g_rho")):
            G[i] += sg[k]
    with psv.Simulation(c, kernel_variant=variant) as sim:
        sim.set_model(lam0, mu0, rho0)
        uz0, ux0 = sim.forward()
        syn = list(zip(sim.split(uz0), sim.split(ux0)))
        sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
        m = sim.gradient()
        Gg = sim.get_gradient()
    assert m.sum() == pytest.approx(L2, rel=1e-4)
    # the oracle given the GPU's own residual: the adjoint pass and correlation alone
    obs_m = cases.matched_obs(c, O, med0, syn, obs)
    Gm = [np.zeros_like(g) for g in G]
    for s in range(2):
        sg = O.shot_gradient(c, med0, s, *obs_m[s])
        for i, k in enumerate(("g_lam", "g_mu", "g_rho")):
            Gm[i] += sg[k]
    r = cases.residual_fraction(syn, obs)
    for i, k in enumerate(("lam", "mu", "rho")):
        em, e = cases.rel_l2(Gg[i], Gm[i]), cases.rel_l2(Gg[i], G[i])
        d = cases.rel_l2(Gm[i], G[i])   # the oracle's own gradient moved by the GPU's residual
        _print(f"surfaces bottom-left-right variant {variant}: preconditioned {k} rel-L2 same residual {em:.2e}, "
               f"end to end {e:.2e} (residual {r:.1e} of the data; its FP32 error alone: {d:.2e})")
        assert em < 1e-4
        assert e <= d * (1 + 1e-3) + em * np.linalg.norm(Gm[i]) / np.linalg.norm(G[i]) + 1e-12


def test_pcg_direction_vs_oracle(psv, O):
    """PCG_PSV (src/cpu/n_PCG_PSV.cpp:13-44) on the library's global gradient, three calls in a
    row (the first with no previous gradient: beta 0), against the oracle per parameter."""
    c = cases.small_case(O, nshot=1, nt=50)
    nz, nx = int(c["nz"]), int(c["nx"])
    rng = np.random.default_rng(17)
    state = [(np.zeros((nz, nx)), np.zeros((nz, nx))) for _ in range(3)]
    with psv.Simulation(c) as sim:
        for call in range(3):
            G = [rng.standard_normal((nz, nx)) * 10.0 ** (3 * p - call) for p in range(3)]
            sim.set_gradient(*G)
            sim.PCG_PSV()
            Gg = sim.get_gradient()
            for p in range(3):
                d, o = state[p]
                g = G[p].copy()
                O.pcg_psv(d, o, g)
                np.testing.assert_allclose(Gg[p], g, rtol=1e-12, atol=1e-12 * np.abs(g).max(),
                                           err_msg=f"call {call} parameter {p}")
    _print("PCG direction: three calls equal the oracle's to 1e-12")


@pytest.mark.parametrize("hfs", [2, 5, 8])
def test_gradient_smoothing_vs_oracle(psv, O, hfs):
    """apply_gauss_filter (src/cpu/n_kernel_lib_PSV.cpp:1051-1084) on the global gradient,
    FP64 accumulation, against the oracle: interior smoothed, border rows kept."""
    c = cases.small_case(O, nshot=1, nt=50)
    nz, nx = int(c["nz"]), int(c["nx"])
    rng = np.random.default_rng(hfs)
    G = [rng.standard_normal((nz, nx)) * 10.0 ** p for p in range(3)]
    with psv.Simulation(c) as sim:
        sim.set_gradient(*G)
        sim.apply_gauss_filter(hfs)
        Gg = sim.get_gradient()
    filt = O.gauss_filter_kernel(hfs)
    for p in range(3):
        ref = G[p].copy()
        O.apply_gauss_filter(ref, hfs)
        np.testing.assert_allclose(Gg[p], ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())
        assert not np.allclose(ref, G[p])
    assert filt.sum() == pytest.approx(1.0)
    _print(f"Gaussian smoothing hfs {hfs}: equal to the oracle's to 1e-12")


@pytest.mark.parametrize("stride", [1, 3])
def test_sampled_kernel_timing(psv, O, stride):
    """fwi_profile_stride: every launch of a timed kernel class is counted and every
    stride-th one carries events; classes left untimed count nothing."""
    nt = 50
    c = cases.small_case(O, nshot=2, nt=nt)
    with psv.Simulation(c, kernel_variant=13) as sim:
        assert sim.plan()["batch"] == 2
        sim.forward()   # the tuner's launches stay out of the counts below
        sim.profile(True, kernels=("stress", "velocity"), stride=stride)
        sim.forward()
        kt = sim.kernel_times()
        sim.profile(False)
    for k in ("stress", "velocity"):
        n, ms, timed = kt[k]
        assert n == nt and timed == (nt + stride - 1) // stride and ms > 0, (k, kt[k])
    assert kt["source"] == (0, 0.0, 0) and kt["receiver"] == (0, 0.0, 0)
    _print(f"stride {stride}: {kt['stress']}, {kt['velocity']}")
