"""The torch.autograd.Function over the C ABI (fwi_cuda_amd.autograd).

Both backward passes against central finite differences of the L2 misfit at 20
seeded interior cells, relative step 1e-3, with the misfits taken from the FP64
oracle's forwards, and along smooth directions on the C2 grid; the correlation
form's lam part against the reference's correlation sum (fwi_gradient_raw),
which it shares.

The exact adjoint (the default) is the derivative of the FP32-modelled misfit:
tolerance 1e-2 for every parameter (what is left is the synthetics' FP32 error
in the residual and the finite difference's own error).  Tolerances of the
correlation form (rel-L2 over the 20 cells) state how close it comes to the
derivative of the discrete misfit: lam 0.15, mu 0.1, rho 0.05.  The
lam and mu terms pair the forward and adjoint stresses at the same time level,
an O(dt) departure from the discrete adjoint; mu also goes through the harmonic
mean of mat_av2 (the chain rule takes its derivative, the correlation itself
is evaluated with the cell's own mu).  The reference's own density term
(adjoint displacement times forward velocity / dt) is 1e-2 .. 1e-1 of the
finite difference and of either sign; the backward pass replaces it.
"""
import numpy as np
import pytest

import cases
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def setup():
    from oracle import oracle as O
    O.load("oracle")
    c = cases.small_case(O, nshot=2, nt=450, snap_dt=1)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    return O, c, obs, cases.start_model(c)


TOL_CELLS = {"correlation": (0.15, 0.1, 0.05), "exact": (1e-2, 1e-2, 1e-2)}


@pytest.mark.parametrize("adjoint", ["correlation", "exact"])
def test_backward_matches_misfit_finite_differences(setup, adjoint):
    import torch
    from fwi_cuda_amd import autograd as A, psv
    O, c, obs, (lam0, mu0, rho0) = setup
    dt = float(c["dt"])
    obs_z = torch.from_numpy(np.concatenate([o[0] for o in obs]))
    obs_x = torch.from_numpy(np.concatenate([o[1] for o in obs]))
    model = [torch.tensor(m, dtype=torch.float64, requires_grad=True) for m in (lam0, mu0, rho0)]
    with psv.Simulation(c) as sim:
        uz, ux = A.seismograms(sim, *model, adjoint=adjoint)
        loss = A.l2_misfit(uz, ux, obs_z, obs_x, dt)
        loss.backward()
    J0 = O.misfit_all(c, lam0, mu0, rho0, obs)
    assert float(loss) == pytest.approx(J0, rel=1e-4)
    grads = [m.grad.numpy() for m in model]

    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    rng = np.random.default_rng(2026)
    cells = [(int(rng.integers(z1 + 8, z2 - 8)), int(rng.integers(x1 + 8, x2 - 8))) for _ in range(20)]
    base = [lam0, mu0, rho0]
    # correlation form measured: lam 0.087, mu 0.019, rho 0.026 (the FP64 oracle's correlation
    # gives the same lam figure)
    for p, (name, tol) in enumerate(zip(("lam", "mu", "rho"), TOL_CELLS[adjoint])):
        fd, ad = [], []
        for iz, ix in cells:
            h = 1e-3 * base[p][iz, ix]
            mp = [m.copy() for m in base]
            mm = [m.copy() for m in base]
            mp[p][iz, ix] += h
            mm[p][iz, ix] -= h
            fd.append((O.misfit_all(c, *mp, obs) - O.misfit_all(c, *mm, obs)) / (2 * h))
            ad.append(grads[p][iz, ix])
        fd, ad = np.array(fd), np.array(ad)
        e = cases.rel_l2(ad, fd)
        _print(f"{adjoint} backward, d misfit / d {name} at 20 cells: rel-L2 vs finite differences {e:.3e}, "
               f"median ratio {np.median(ad / fd):.4f}")
        assert e < tol, (name, e)


@pytest.mark.parametrize("adjoint,seis64", [("correlation", 0), ("exact", 0), ("exact", 1)])
def test_c2_directional_derivatives(adjoint, seis64):
    """C2 (301 x 601, 10 layers, 8 shots, 120 receivers, 250 Hz, nt 3000): the
    backward pass along smooth random directions against central differences of
    the misfit along the same directions (relative step 1e-3), all on the GPU; with
    seis64 the misfits (and the adjoint sources) come from the FP64 propagation."""
    import torch
    from scipy.ndimage import gaussian_filter
    from fwi_cuda_amd import autograd as A, configs, psv
    c = configs.c2(snap_dt=1)   # the correlation form on every step (its tolerances were measured so)
    dt = float(c["dt"])
    true = [c["lam"], c["mu"], c["rho"]]
    start = [gaussian_filter(m, 10.0, mode="nearest") for m in true]
    nz, nx = int(c["nz"]), int(c["nx"])
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    rng = np.random.default_rng(22)
    errs = {}
    with psv.Simulation(c, seis_fp64=seis64) as sim:
        obs = sim.forward()
        sim.set_observed(*obs)
        model = [torch.tensor(m, requires_grad=True) for m in start]
        uz, ux = A.seismograms(sim, *model, adjoint=adjoint)
        loss = A.l2_misfit(uz, ux, torch.from_numpy(obs[0]), torch.from_numpy(obs[1]), dt)
        loss.backward()
        for p, name in enumerate(("lam", "mu", "rho")):
            d = np.zeros((nz, nx))
            d[z1 + 10:z2 - 10, x1 + 10:x2 - 10] = 1.0
            d *= gaussian_filter(rng.standard_normal((nz, nx)), 8.0)
            d *= start[p] / np.abs(d).max()
            h = 1e-3
            Js = []
            for sgn in (1, -1):
                m = [a.copy() for a in start]
                m[p] = m[p] + sgn * h * d
                sim.set_model(*m)
                Js.append(sim.misfit())
            fd = (Js[0] - Js[1]) / (2 * h)
            ad = float(np.sum(model[p].grad.numpy() * d))
            e = abs(ad - fd) / abs(fd)
            _print(f"C2 directional derivative ({adjoint}, FP64 seismograms {seis64}) along a smooth {name} "
                   f"direction: backward {ad:.6e}, finite difference {fd:.6e}, relative difference {e:.3e}")
            errs[name] = e
    # lam's derivative along this direction (3.42e-27) is 1e3 times smaller than mu's and
    # rho's and below what FP32 misfits resolve at this step (their rounding noise is ~1e-29
    # in J, 1e-7 of it, against a change of 7e-30): with FP32 seismograms the lam finite
    # difference is noise (7.8e-27), so only its sign is asserted there.  With the misfits
    # from the FP64 propagation (seis64) the finite difference resolves it, and the exact
    # adjoint matches all three: measured lam 7.3e-5, mu 1.8e-5, rho 2.7e-6.  The correlation
    # form measured against the FP32 differences: lam 0.22, mu 5.4e-3, rho 1.1e-4; against the
    # FP64 ones its lam derivative along this direction is 1.76 times the true one (the O(dt)
    # pairing of forward and adjoint stresses), mu 5.4e-3, rho 1.1e-4.
    tols = {("correlation", 0): (0.3, 0.05, 0.01), ("exact", 0): (1.0, 1e-2, 1e-2),
            ("exact", 1): (1e-3, 1e-3, 1e-3)}[(adjoint, seis64)]
    for name, tol in zip(("lam", "mu", "rho"), tols):
        assert errs[name] < tol, (name, errs[name])


def test_lam_part_is_the_reference_correlation(setup):
    from fwi_cuda_amd import psv
    O, c, obs, (lam0, mu0, rho0) = setup
    with psv.Simulation(c) as sim:
        sim.set_model(lam0, mu0, rho0)
        uz, ux = sim.forward()
        rz = uz - np.concatenate([o[0] for o in obs])
        rx = ux - np.concatenate([o[1] for o in obs])
        d_lam, _, _ = sim.misfit_gradient(rz, rx)
        raw_lam, _, _ = sim.gradient_raw(rz, rx)
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    np.testing.assert_array_equal(d_lam[z1:z2 + 1, x1:x2 + 1], raw_lam)
    assert not d_lam[:z1].any() and not d_lam[z2 + 1:].any()


def test_device_pointers_equal_the_host_path(setup):
    """CUDA tensors cross the ABI as device pointers (fwi_set_model_device,
    fwi_forward_device, fwi_misfit_gradient_device; torch was loaded first, so the
    library runs on torch's HIP runtime): seismograms and the misfit derivative equal
    the host-array calls bit for bit, and the autograd function keeps them on the GPU."""
    import torch
    from fwi_cuda_amd import autograd as A, psv
    O, c, obs, (lam0, mu0, rho0) = setup
    dev = torch.device("cuda")
    dt = float(c["dt"])
    obs_z = np.concatenate([o[0] for o in obs])
    obs_x = np.concatenate([o[1] for o in obs])
    with psv.Simulation(c) as sim:
        sim.set_model(lam0, mu0, rho0)
        uz_h, ux_h = sim.forward()
        g_h = sim.misfit_gradient(uz_h - obs_z, ux_h - obs_x)
        m = [torch.tensor(a, device=dev) for a in (lam0, mu0, rho0)]
        uz = torch.empty(uz_h.shape, dtype=torch.float64, device=dev)
        ux = torch.empty_like(uz)
        torch.cuda.synchronize()
        sim.set_model_device(*[t.data_ptr() for t in m])
        sim.forward_device(uz.data_ptr(), ux.data_ptr())
        np.testing.assert_array_equal(uz.cpu().numpy(), uz_h)
        np.testing.assert_array_equal(ux.cpu().numpy(), ux_h)
        rz = (uz - torch.tensor(obs_z, device=dev)).contiguous()
        rx = (ux - torch.tensor(obs_x, device=dev)).contiguous()
        d = [torch.empty_like(t) for t in m]
        torch.cuda.synchronize()
        sim.misfit_gradient_device(rz.data_ptr(), rx.data_ptr(), *[t.data_ptr() for t in d])
        for a, b in zip(d, g_h):
            np.testing.assert_array_equal(a.cpu().numpy(), b)
        # the autograd function with CUDA tensors against CPU tensors
        out = {}
        for where in ("cpu", "cuda"):
            model = [torch.tensor(a, device=where, requires_grad=True) for a in (lam0, mu0, rho0)]
            uz_t, ux_t = A.seismograms(sim, *model)
            loss = A.l2_misfit(uz_t, ux_t, torch.tensor(obs_z, device=where), torch.tensor(obs_x, device=where), dt)
            loss.backward()
            assert uz_t.device.type == where and model[0].grad.device.type == where
            out[where] = (float(loss.detach()), [t.grad.cpu().numpy() for t in model])
    assert out["cuda"][0] == pytest.approx(out["cpu"][0], rel=1e-12)
    for a, b in zip(out["cuda"][1], out["cpu"][1]):
        np.testing.assert_array_equal(a, b)


def test_c2_twenty_cells_vs_fp64_finite_differences():
    """C2 itself (301 x 601, 10 layers, 8 shots, 120 receivers, 250 Hz, nt 3000): the exact
    adjoint (the torch backward) at 20 seeded interior cells against central differences
    (relative step 1e-3) of the misfit, taken from the FP64 propagation (seis_fp64) so that a
    single cell's change is resolved; all on the GPU."""
    import time
    import torch
    from scipy.ndimage import gaussian_filter
    from fwi_cuda_amd import autograd as A, configs, psv
    c = configs.c2()
    dt = float(c["dt"])
    true = [c["lam"], c["mu"], c["rho"]]
    start = [gaussian_filter(m, 10.0, mode="nearest") for m in true]
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    rng = np.random.default_rng(2)
    cells = [(int(rng.integers(z1 + 10, z2 - 10)), int(rng.integers(x1 + 10, x2 - 10))) for _ in range(20)]
    t0 = time.perf_counter()
    with psv.Simulation(c, seis_fp64=1) as sim:
        obs = sim.forward()
        sim.set_observed(*obs)
        model = [torch.tensor(m, requires_grad=True) for m in start]
        uz, ux = A.seismograms(sim, *model, adjoint="exact")
        loss = A.l2_misfit(uz, ux, torch.from_numpy(obs[0]), torch.from_numpy(obs[1]), dt)
        loss.backward()
        for p, name in enumerate(("lam", "mu", "rho")):
            fd, ad = [], []
            for iz, ix in cells:
                h = 1e-3 * start[p][iz, ix]
                Js = []
                for sgn in (1, -1):
                    m = [a.copy() for a in start]
                    m[p][iz, ix] += sgn * h
                    sim.set_model(*m)
                    Js.append(sim.misfit())
                fd.append((Js[0] - Js[1]) / (2 * h))
                ad.append(float(model[p].grad[iz, ix]))
            fd, ad = np.array(fd), np.array(ad)
            e = cases.rel_l2(ad, fd)
            _print(f"C2, 20 cells, d misfit / d {name}: exact backward against FP64 finite differences rel-L2 "
                   f"{e:.3e}, median ratio {np.median(ad / fd):.6f} ({time.perf_counter() - t0:.0f} s)")
            assert np.all(np.abs(fd) > 0)
            assert e < 1e-4, (name, e)   # measured lam 1.1e-5, mu 9.9e-7, rho 6.6e-7
