"""The exact discrete adjoint on the GPU (fwi_misfit_gradient_exact) against the oracle's.

The oracle's adjoint_exact is the reverse mode of its FP64 time step, pinned by the
dot-product test (tests/test_oracle_adjoint.py).  Given the same cotangents the GPU's
gradient (FP32 forward and adjoint fields, FP64 products and sums) is compared after the
chain rule through mat_av2, inside the inversion box: rel-L2 <= 1e-4 per parameter, at
orders 2, 4 and 8, CPML or a free surface (top, or four sides with image rows), type-0
and type-1 sources, one shot or two shots with uneven receivers, host and device
pointers, and with the forward recomputed from checkpoints; the dot-product test <J dm, y> = <dm, J^T y> with J dm from central
differences of the FP64 oracle's forward, relative 1e-5.  Then the autograd backward in exact mode against central differences of the
misfit along smooth directions (relative step 1e-3), tolerance 1e-2.
"""
import numpy as np
import pytest

import cases
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _print(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def O():
    from oracle import oracle as O_
    O_.load("oracle")
    return O_


def _oracle_exact(O, c, med, rz, rx):
    """sum over shots of the oracle's adjoint for cotangents dt * r, chained to (lam, mu, rho)
    and cut to the inversion box (inclusive bounds)."""
    dt = float(c["dt"])
    keys = ("lam", "mu", "mu_zx", "rho_zp", "rho_xp")
    tot = {k: 0.0 for k in keys}
    nrec = len(c["z_rec"])   # shared receivers unless rec_off gives each shot its own rows
    off = c["rec_off"] if "rec_off" in c else np.arange(int(c["nshot"]) + 1) * nrec
    for s in range(int(c["nshot"])):
        a, b = int(off[s]), int(off[s + 1])
        g = O.adjoint_exact(cases.shot_case(c, s), med, s, dt * rz[a:b], dt * rx[a:b])
        for k in keys:
            tot[k] = tot[k] + g[k]
    d = O.media_chain(tot, med)
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    out = []
    for a in d:
        m = np.zeros_like(a)
        m[z1:z2 + 1, x1:x2 + 1] = a[z1:z2 + 1, x1:x2 + 1]
        out.append(m)
    return out


CASES = [(2, "top", 1, 1, False), (8, "top", 0, 1, False), (8, "none", 1, 2, True), (4, "four", 0, 2, True),
         (8, "four", 1, 1, False)]


@pytest.mark.parametrize("fdorder,surf,stf_type,nshot,uneven", CASES)
def test_exact_adjoint_vs_oracle(O, fdorder, surf, stf_type, nshot, uneven):
    from fwi_cuda_amd import psv
    c = cases.adjoint_case(O, fdorder, surf, stf_type, nshot=nshot, uneven=uneven)
    med = O.mat_av2(c)
    nt = int(c["nt"])
    with psv.Simulation(c) as sim:
        rows = sim._rows(0, nshot)
        rng = np.random.default_rng(fdorder * 10 + nshot)
        rz, rx = rng.standard_normal((rows, nt)), rng.standard_normal((rows, nt))
        g = sim.misfit_gradient_exact(rz, rx)
    go = _oracle_exact(O, c, med, rz, rx)
    for k, a, b in zip(("lam", "mu", "rho"), g, go):
        e = cases.rel_l2(a, b)
        _print(f"exact adjoint order {fdorder} surface {surf} stf {stf_type} shots {nshot}: d{k} rel-L2 {e:.2e}")
        assert np.abs(b).max() > 0
        assert e < 1e-4, (k, e)


@pytest.mark.parametrize("fdorder,surf,stf_type", [(8, "top", 1), (4, "four", 0)])
def test_exact_adjoint_dot_product(O, fdorder, surf, stf_type):
    """<J dm, y> with J dm from central differences of the FP64 oracle's seismograms
    (model -> mat_av2 -> kernel_psv) against <dm, J^T y> from the GPU's exact adjoint,
    for a smooth dm in (lam, mu, rho) inside the inversion box: relative 1e-5."""
    from scipy.ndimage import gaussian_filter
    from fwi_cuda_amd import psv
    c = cases.adjoint_case(O, fdorder, surf, stf_type, nshot=2, uneven=True)
    nz, nx, nt, dt = int(c["nz"]), int(c["nx"]), int(c["nt"]), float(c["dt"])
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    rng = np.random.default_rng(41)
    base = [c["lam"], c["mu"], c["rho"]]
    dm = []
    for m in base:
        d = np.zeros((nz, nx))
        d[z1:z2 + 1, x1:x2 + 1] = gaussian_filter(rng.standard_normal((z2 - z1 + 1, x2 - x1 + 1)), 2.0)
        dm.append(d / np.abs(d).max() * np.abs(m).max())
    with psv.Simulation(c) as sim:
        rows = sim._rows(0, 2)
        y = rng.standard_normal((rows, nt)), rng.standard_normal((rows, nt))
        g = sim.misfit_gradient_exact(*y)
    off = c["rec_off"]
    h = 1e-6

    def seis(sign):
        m = [a + sign * h * d for a, d in zip(base, dm)]
        med = O.mat_av2(c, *m)
        out = []
        for s in range(2):
            f = O.kernel_psv(cases.shot_case(c, s), med, s, accu=False)
            out.append((f["uz"], f["ux"]))
        return out

    p, q = seis(1), seis(-1)
    lhs = 0.0
    for s in range(2):
        a, b = int(off[s]), int(off[s + 1])
        for k in range(2):
            lhs += float(np.sum((p[s][k] - q[s][k]) / (2 * h) * dt * y[k][a:b]))
    rhs = float(sum(np.sum(d * gk) for d, gk in zip(dm, g)))
    rel = abs(lhs - rhs) / abs(lhs)
    _print(f"dot product order {fdorder} surface {surf} stf {stf_type}: <J dm, y> {lhs:.10e} (oracle differences), "
           f"<dm, J^T y> {rhs:.10e} (GPU), relative {rel:.2e}")
    assert rel < 1e-5


@pytest.mark.parametrize("S", [17, 40])
def test_exact_adjoint_checkpointed_equals_full_storage(O, S):
    """The exact adjoint with the forward's intermediates of only S steps held (a checkpoint
    of the forward state at each segment start, each segment recomputed before its reverse
    sweep; the last one short) equals the run that holds every step's, bit for bit."""
    from fwi_cuda_amd import psv
    c = cases.adjoint_case(O, 8, "top", 0, nshot=2, uneven=True)
    nt = int(c["nt"])
    rng = np.random.default_rng(S)
    out = {}
    for st, kw in (("full", dict(storage=psv._lib.STORE_FULL)),
                   ("ckpt", dict(storage=psv._lib.STORE_CHECKPOINT, ckpt_interval=S))):
        with psv.Simulation(c, **kw) as sim:
            rows = sim._rows(0, 2)
            if not out:
                y = rng.standard_normal((rows, nt)), rng.standard_normal((rows, nt))
            out[st] = sim.misfit_gradient_exact(*y)
    for a, b in zip(out["ckpt"], out["full"]):
        np.testing.assert_array_equal(a, b)
    assert np.abs(out["full"][0]).max() > 0
    _print(f"exact adjoint, segments of {S} of {nt} steps recomputed: equal to full storage bit for bit")


def test_exact_adjoint_device_pointers_equal_host(O):
    import torch
    from fwi_cuda_amd import psv
    c = cases.adjoint_case(O, 8, "top", 1, nshot=2, uneven=True)
    nt = int(c["nt"])
    with psv.Simulation(c) as sim:
        rows = sim._rows(0, 2)
        rng = np.random.default_rng(3)
        rz, rx = rng.standard_normal((rows, nt)), rng.standard_normal((rows, nt))
        g = sim.misfit_gradient_exact(rz, rx)
        dz, dx = torch.from_numpy(rz).cuda(), torch.from_numpy(rx).cuda()
        d = [torch.empty((sim.nz, sim.nx), dtype=torch.float64, device="cuda") for _ in range(3)]
        torch.cuda.synchronize()
        sim.misfit_gradient_exact_device(dz.data_ptr(), dx.data_ptr(), *[t.data_ptr() for t in d])
        torch.cuda.synchronize()
    for a, b in zip(g, d):
        np.testing.assert_array_equal(a, b.cpu().numpy())


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_exact_backward_matches_misfit_differences(O, device):
    """The autograd backward in exact mode along smooth random directions, against central
    differences (relative step 1e-3) of the FP64 oracle's misfit for FP64 observed data:
    tolerance 1e-2 for every parameter (the correlation form is 0.05 .. 0.3 off here).
    What is left is the synthetics' FP32 error in the residual: the finite differences of
    the GPU's own misfit are no reference for lam, whose derivative along such a direction
    is ~1e3 times smaller than mu's and comparable to the FP32 forward's rounding noise."""
    import torch
    from scipy.ndimage import gaussian_filter
    from fwi_cuda_amd import autograd as A, psv
    c = cases.small_case(O, nshot=2, nt=450, snap_dt=1)
    dt = float(c["dt"])
    obs = cases.per_shot_obs(c, O, O.mat_av2(c))
    start = list(cases.start_model(c))
    nz, nx = int(c["nz"]), int(c["nx"])
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    rng = np.random.default_rng(31)
    obs_z = torch.from_numpy(np.concatenate([o[0] for o in obs])).to(device)
    obs_x = torch.from_numpy(np.concatenate([o[1] for o in obs])).to(device)
    with psv.Simulation(c) as sim:
        model = [torch.tensor(m, requires_grad=True, device=device) for m in start]
        uz, ux = A.seismograms(sim, *model, adjoint="exact")
        loss = A.l2_misfit(uz, ux, obs_z, obs_x, dt)
        loss.backward()
    assert float(loss) == pytest.approx(O.misfit_all(c, *start, obs), rel=1e-4)
    for p, name in enumerate(("lam", "mu", "rho")):
        d = np.zeros((nz, nx))
        d[z1 + 4:z2 - 4, x1 + 4:x2 - 4] = 1.0
        d *= gaussian_filter(rng.standard_normal((nz, nx)), 4.0)
        d *= start[p] / np.abs(d).max()
        h = 1e-3
        Js = []
        for sgn in (1, -1):
            m = [a.copy() for a in start]
            m[p] = m[p] + sgn * h * d
            Js.append(O.misfit_all(c, *m, obs))
        fd = (Js[0] - Js[1]) / (2 * h)
        ad = float(np.sum(model[p].grad.cpu().numpy() * d))
        e = abs(ad - fd) / abs(fd)
        _print(f"exact backward ({device}) along a smooth {name} direction: {ad:.8e} vs FP64 finite difference "
               f"{fd:.8e}, relative difference {e:.2e}")
        assert e < 1e-2, (name, e)
