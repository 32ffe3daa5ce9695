"""seis_fp64: seismograms from the FP64 propagation (psv_fp64.hip) against the FP64 oracle.

With FP32 wavefields the seismograms carry ~1e-6 of relative error, which becomes ~1e-3 of
the gradient where the residual is ~1e-2 of the data (C1). With seis_fp64 the seismograms
(fwi_forward, misfits, adjoint sources) come from an FP64 propagation and the FP32 stencil
still runs the wavefields the gradient correlates:

  - forward seismograms against the oracle's to 1e-10 relative (orders 2 and 8, CPML, a top
    surface or surfaces on all four sides with image rows, type-0 and type-1 sources, uneven
    shots); the only differences are FMA contraction and the order of FP64 roundings;
  - C1 at full size: misfit to 1e-9 and the preconditioned gradient end to end within 1e-4
    of the FP64 oracle's (1.3e-3 .. 2.3e-3 with FP32 seismograms), and one FWI iteration
    (gradient, taper, step-length search, update) with the GPU twin's sign.
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


@pytest.mark.parametrize("fdorder,surf,stf_type,nshot,uneven", [(2, "top", 1, 1, False), (8, "top", 0, 2, True),
                                                                (8, "none", 1, 2, False), (4, "four", 0, 2, True)])
def test_fp64_seismograms_vs_oracle(O, fdorder, surf, stf_type, nshot, uneven):
    from fwi_cuda_amd import psv
    c = cases.adjoint_case(O, fdorder, surf, stf_type, nshot=nshot, uneven=uneven, nt=300)
    med = O.mat_av2(c)
    ref = [O.kernel_psv(cases.shot_case(c, s), med, s, accu=False) for s in range(nshot)]
    out = {}
    for seis64 in (0, 1):
        with psv.Simulation(c, seis_fp64=seis64) as sim:
            uz, ux = sim.forward()
            out[seis64] = list(zip(sim.split(uz), sim.split(ux)))
    for s in range(nshot):
        e64 = max(cases.rel_l2(out[1][s][0], ref[s]["uz"]), cases.rel_l2(out[1][s][1], ref[s]["ux"]))
        e32 = max(cases.rel_l2(out[0][s][0], ref[s]["uz"]), cases.rel_l2(out[0][s][1], ref[s]["ux"]))
        _print(f"order {fdorder} surface {surf} stf {stf_type} shot {s}: seismograms rel-L2 FP64 {e64:.2e}, "
               f"FP32 {e32:.2e}")
        assert e64 < 1e-10
        assert np.abs(ref[s]["uz"]).max() > 0


def _oracle_gradient(O, c, model, obs):
    med = O.mat_av2(c, *model)
    nz, nx = int(c["nz"]), int(c["nx"])
    G = [np.zeros((nz, nx)) for _ in range(3)]
    L2 = 0.0
    for s in range(int(c["nshot"])):
        sg = O.shot_gradient(cases.shot_case(c, s), med, s, *obs[s])
        L2 += sg["L2"]
        for i, k in enumerate(("g_lam", "g_mu", "g_rho")):
            G[i] += sg[k]
    return G, L2


def test_c1_gradient_with_fp64_seismograms(O):
    from fwi_cuda_amd import configs, psv
    c = dict(configs.c1())
    c.pop("seed", None)
    start = dict(configs.c1(true_model=False))
    m0 = (start["lam"], start["mu"], start["rho"])
    obs = cases.per_shot_obs(c, O, O.mat_av2(c))
    G, L2 = _oracle_gradient(O, c, m0, obs)
    with psv.Simulation(c, seis_fp64=1) as sim:
        sim.set_model(*m0)
        sim.set_observed(*obs[0])
        m = sim.gradient()
        Gg = sim.get_gradient()
    _print(f"C1 with FP64 seismograms: misfit {m[0]:.12e} vs {L2:.12e}")
    assert m[0] == pytest.approx(L2, rel=1e-9)
    for i, k in enumerate(("lam", "mu", "rho")):
        e = cases.rel_l2(Gg[i], G[i])
        _print(f"C1 with FP64 seismograms: preconditioned gradient {k} rel-L2 vs FP64 oracle {e:.2e}")
        assert e < 1e-4
    # one FWI iteration with the GPU twin's sign against the oracle's
    Gz = [np.zeros_like(g) for g in G]
    L2_it, step_o, m1_o, trials = O.fwi_iteration(c, m0, Gz, obs, 0.01, grad_sign=+1)
    with psv.Simulation(c, grad_sign=+1, seis_fp64=1) as sim:
        sim.set_model(*m0)
        sim.set_observed(*obs[0])
        L2_d, step_d = sim.fwi_iteration()
        m1 = sim.get_model()
    _print(f"C1 iteration with FP64 seismograms: misfit {L2_d:.10e} vs {L2_it:.10e}; step {step_d:.8g} vs "
           f"{step_o:.8g} ({len(trials)} oracle trials)")
    assert L2_d == pytest.approx(L2_it, rel=1e-9)
    assert step_d == pytest.approx(step_o, rel=1e-3)
    for i, k in enumerate(("lam", "mu", "rho")):
        du = cases.rel_l2(m1[i] - m0[i], m1_o[i] - m0[i])
        _print(f"C1 iteration with FP64 seismograms: {k} update rel-L2 {du:.2e}")
        assert du < 1e-3
