"""HIP path against the reference's golden vectors and the FP64 oracle.

Tolerances (FP32 fields, FP64 receiver sums and gradient preconditioning):
  seismograms            rel-L2 <= 2e-5 (order 2, 300 steps) / 1e-4 (order 8)
  raw shot gradients     rel-L2 <= 1e-3
  preconditioned G       rel-L2 <= 1e-3
  checkpointed vs full, batched vs per-shot: bit-identical (same kernels, same inputs)
"""
import numpy as np
import pytest

import cases
from conftest import case_from_fixture, gpu_available, load_fixture

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


@pytest.mark.parametrize("name", ["o2_pml", "o2_surf"])
def test_forward_matches_reference_golden(psv, name):
    fx = load_fixture(name)
    c = case_from_fixture(fx)
    out = fx["out"]
    with psv.Simulation(c) as sim:
        med = sim.mat_av2()
        for k in ("mu_zx", "rho_zp", "rho_xp"):
            assert cases.rel_l2(med[k][:-1, :-1], out["media_true"][k][:-1, :-1]) < 1e-7
        np.testing.assert_allclose(med["C"], out["media_true"]["C"], rtol=1e-13)
        uz, ux = sim.forward()
        segs_z, segs_x = sim.split(uz), sim.split(ux)
        for s in range(2):
            ez = cases.rel_l2(segs_z[s], out[f"seis{s}_uz"])
            ex = cases.rel_l2(segs_x[s], out[f"seis{s}_ux"])
            _print(f"{name} shot {s}: seismogram rel-L2 uz {ez:.2e} ux {ex:.2e}")
            assert ez < 2e-5 and ex < 2e-5
        # shot 1 was the last in the batch; its final state
        for k in ("vz", "vx", "szz", "szx", "sxx", "We"):
            e = cases.rel_l2(sim.wavefield(1, k), out["state1"][k])
            assert e < 2e-5, (k, e)


@pytest.mark.parametrize("name", ["o2_pml", "o2_surf"])
def test_shot_gradient_matches_reference_golden(psv, name):
    fx = load_fixture(name)
    c = case_from_fixture(fx)
    out = fx["out"]
    g = out["grad1"]
    with psv.Simulation(c) as sim:
        sim.set_model(fx["in0"]["lam"], fx["in0"]["mu"], fx["in0"]["rho"])
        obs_z = np.concatenate([out["seis0_uz"], out["seis1_uz"]])
        obs_x = np.concatenate([out["seis0_ux"], out["seis1_ux"]])
        sim.set_observed(obs_z, obs_x)
        sim.zero_gradient()
        m = sim.gradient(1, 1)
        assert m[0] == pytest.approx(float(g["L2"]), rel=1e-4)
        G = sim.get_gradient()
        for i, k in enumerate(("g_lam", "g_mu", "g_rho")):
            e = cases.rel_l2(G[i], g[k])
            _print(f"{name}: preconditioned {k} rel-L2 {e:.2e}")
            assert e < 1e-3
        # raw correlation sums with the reference's residuals as adjoint sources
        fz, fx_ = sim.forward(1, 1)
        rz, rx = fz - out["seis1_uz"], fx_ - out["seis1_ux"]
        raw = sim.gradient_raw(rz, rx, 1, 1)
        for i, k in enumerate(("raw_lam", "raw_mu", "raw_rho")):
            e = cases.rel_l2(raw[i], g[k])
            _print(f"{name}: raw {k} rel-L2 {e:.2e}")
            assert e < 1e-3


@pytest.mark.parametrize("surface", [True, False])
@pytest.mark.parametrize("variant", [1, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13])
def test_order8_forward_vs_oracle(psv, O, surface, variant):
    c = cases.small_case(O, surface=surface, nshot=2)
    med = O.mat_av2(c)
    ref = [O.kernel_psv(c, med, s, accu=False) for s in range(2)]
    with psv.Simulation(c, kernel_variant=variant) as sim:
        uz, ux = sim.forward()
        for s, (a, b) in enumerate(zip(sim.split(uz), sim.split(ux))):
            ez, ex = cases.rel_l2(a, ref[s]["uz"]), cases.rel_l2(b, ref[s]["ux"])
            _print(f"order 8 surface={surface} shot {s}: rel-L2 uz {ez:.2e} ux {ex:.2e}")
            assert ez < 1e-4 and ex < 1e-4
        e = cases.rel_l2(sim.wavefield(1, "sxx"), ref[1]["state"]["sxx"])
        assert e < 1e-4


@pytest.mark.parametrize("variant", [1, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13])
def test_order8_gradient_vs_oracle(psv, O, variant):
    c = cases.small_case(O, nshot=2, nt=450)
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
    with psv.Simulation(c, kernel_variant=variant) as sim:
        sim.set_model(lam0, mu0, rho0)
        sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
        m = sim.gradient()
        assert m.sum() == pytest.approx(L2, rel=1e-4)
        Gg = sim.get_gradient()
        for i, k in enumerate(("lam", "mu", "rho")):
            e = cases.rel_l2(Gg[i], G[i])
            _print(f"order 8 preconditioned gradient {k}: rel-L2 vs FP64 oracle {e:.2e}")
            assert e < 2e-3


@pytest.mark.parametrize("variant", [1, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13])
def test_checkpointed_equals_full_storage(psv, O, variant):
    c = cases.small_case(O, nshot=2, nt=300, snap_dt=2)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    res = {}
    # checkpoints every 14 steps with the last 1, 5 or (budget) all 22 segments' snapshots kept
    # by the forward pass, and one segment of 300 steps
    for storage, interval, keep in (("full", 0, 0), ("ckpt", 14, 1), ("ckpt", 14, 5), ("ckpt", 14, 0),
                                    ("ckpt", 300, 0)):
        st = psv._lib.STORE_FULL if storage == "full" else psv._lib.STORE_CHECKPOINT
        with psv.Simulation(c, storage=st, ckpt_interval=interval, ckpt_keep=keep, kernel_variant=variant) as sim:
            plan = sim.plan()
            assert plan["storage"] == ("full" if storage == "full" else "checkpoint")
            if storage == "ckpt":
                assert plan["ckpt_kept"] == (keep or plan["ckpt_units"])
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
            m = sim.gradient()
            res[(storage, interval, keep)] = (m, sim.get_gradient())
    ref = res[("full", 0, 0)]
    for key, (m, G) in res.items():
        np.testing.assert_array_equal(m, ref[0])
        for i in range(3):
            np.testing.assert_array_equal(G[i], ref[1][i], err_msg=str(key))


@pytest.mark.parametrize("variant,graph,sdt,S", [(1, 0, 2, 14), (7, 0, 2, 14), (8, 1, 2, 14), (5, 1, 2, 14),
                                                (9, 1, 2, 14), (10, 0, 3, 15), (10, 1, 3, 18), (5, 0, 3, 15),
                                                (8, 0, 3, 15)])
def test_binomial_checkpoints_equal_full_storage(psv, O, variant, graph, sdt, S):
    """Fewer checkpoint buffers than segment boundaries (ckpt_slots 1, 2, 5 for the
    segments of S steps, the last one short, snapshots every sdt-th step): the reversal
    advances from held checkpoints on the binomial schedule and recomputes more, and the
    misfits and gradient still equal full snapshot storage bit for bit (the same kernels
    replay the same steps), with and without graph replay.  With a checkpoint at every
    boundary the forward pass keeps the snapshots of the last 1, 4 or (budget) all
    segments, which are then not recomputed: the same results again."""
    nt = 302
    c = cases.small_case(O, nshot=3, nt=nt, snap_dt=sdt, uneven=True)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    res = {}
    units = (nt + S - 1) // S
    for storage, slots, keep in (("full", 0, 0), ("ckpt", 0, 1), ("ckpt", 0, 4), ("ckpt", 0, 0), ("ckpt", 5, 0),
                                 ("ckpt", 2, 0), ("ckpt", 1, 0)):
        st = psv._lib.STORE_FULL if storage == "full" else psv._lib.STORE_CHECKPOINT
        with psv.Simulation(c, storage=st, ckpt_interval=S, ckpt_slots=slots, ckpt_keep=keep, kernel_variant=variant,
                            use_graph=graph, batch=2) as sim:
            plan = sim.plan()
            if storage == "ckpt":
                kept = (keep or units) if slots == 0 else 1   # fewer slots: only the last unit kept
                assert plan["ckpt_units"] == units
                assert plan["ckpt_kept"] == kept
                assert plan["ckpt_slots"] == (max(units - kept - 1, 0) if slots == 0 else slots)
                recomputed = (units - kept) * S / nt
                if slots == 0:
                    assert plan["recompute"] == pytest.approx(recomputed, abs=1e-4)
                else:
                    assert plan["recompute"] > recomputed
            _print(f"variant {variant} graph {graph} {storage} slots {slots} keep {keep}: {plan}")
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
            m = sim.gradient()
            res[(storage, slots, keep)] = (m, sim.get_gradient(), plan["recompute"])
    ref = res[("full", 0, 0)]
    for key, (m, G, _) in res.items():
        np.testing.assert_array_equal(m, ref[0])
        for i in range(3):
            np.testing.assert_array_equal(G[i], ref[1][i], err_msg=str(key))
    r = {k[1:]: v[2] for k, v in res.items() if k[0] == "ckpt"}
    assert r[(1, 0)] > r[(2, 0)] > r[(5, 0)] > r[(0, 1)] > r[(0, 4)] > r[(0, 0)] == 0.0


@pytest.mark.parametrize("variant", [5, 10])
def test_binomial_checkpoints_many_units(psv, O, variant):
    """The C4 sweep's shape of reversal on a small grid: 51 units of 15 steps (the last
    one 5 steps), snapshots every 3rd step, one shot per launch, 50 (one per boundary),
    32, 8 and 3 checkpoint buffers: every schedule gives full storage's misfit and
    gradient bit for bit, and recomputes more steps the fewer buffers it holds."""
    S, nunits, sdt = 15, 51, 3
    nt = (nunits - 1) * S + 5
    c = cases.small_case(O, nshot=1, nt=nt, snap_dt=sdt, nz=56, nx=80, npml=10)
    obs = cases.per_shot_obs(c, O, O.mat_av2(c))
    lam0, mu0, rho0 = cases.start_model(c)
    res = {}
    for storage, slots in (("full", 0), ("ckpt", 0), ("ckpt", 32), ("ckpt", 8), ("ckpt", 3)):
        st = psv._lib.STORE_FULL if storage == "full" else psv._lib.STORE_CHECKPOINT
        with psv.Simulation(c, storage=st, ckpt_interval=S, ckpt_slots=slots, ckpt_keep=1, kernel_variant=variant,
                            tile_tz=16, work_order=1) as sim:
            plan = sim.plan()
            if storage == "ckpt":   # the last unit kept: 50 units reversed, 49 boundaries
                assert plan["ckpt_units"] == nunits and plan["ckpt_kept"] == 1
                assert plan["ckpt_slots"] == (nunits - 2 if slots == 0 else slots)
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed(*obs[0])
            m = sim.gradient()
            res[(storage, slots)] = (m, sim.get_gradient(), plan["recompute"])
            _print(f"variant {variant} {storage} slots {slots}: recompute {plan['recompute']}, misfit {m[0]:.17e}")
    ref = res[("full", 0)]
    for key, (m, G, _) in res.items():
        np.testing.assert_array_equal(m, ref[0], err_msg=str(key))
        for i in range(3):
            np.testing.assert_array_equal(G[i], ref[1][i], err_msg=str(key))
    assert res[("ckpt", 3)][2] > res[("ckpt", 8)][2] > res[("ckpt", 32)][2] > res[("ckpt", 0)][2]
    assert res[("ckpt", 0)][2] == pytest.approx((nunits - 1) * S / nt, abs=1e-4)


def test_c4_window_checkpointed_equals_full_storage(psv, O):
    """The C4 model (15 layers and a dipping dyke, order 8) cropped to a 1001 x 2001
    window, two shots, 1600 steps (the 100 Hz wavelet peaks at step ~620): checkpointed
    recomputation at two intervals (every segment but the last recomputed, and the last
    segments' snapshots kept as the budget allows) against full snapshot storage of the
    same window (64 GB of snapshots per shot), bit for bit (misfits and gradient)."""
    from scipy.ndimage import gaussian_filter
    from fwi_cuda_amd import configs
    c = configs.c4(nt=1600, nshot=2, nz=1001, nx=2001)
    start = [gaussian_filter(c[k], 20.0, mode="nearest") for k in ("lam", "mu", "rho")]
    with psv.Simulation(c) as sim:
        obs = sim.forward()
    res = {}
    for storage, interval, keep in (("full", 0, 0), ("ckpt", 26, 1), ("ckpt", 100, 0)):
        st = psv._lib.STORE_FULL if storage == "full" else psv._lib.STORE_CHECKPOINT
        with psv.Simulation(c, storage=st, ckpt_interval=interval, ckpt_keep=keep) as sim:
            plan = sim.plan()
            assert plan["storage"] == ("full" if storage == "full" else "checkpoint")
            sim.set_model(*start)
            sim.set_observed(*obs)
            m = sim.gradient()
            res[(storage, interval, keep)] = (m, sim.get_gradient())
            _print(f"C4 window {storage} {interval}: plan {plan}, misfits {m}")
    ref = res[("full", 0, 0)]
    assert np.all(ref[0] > 0)
    for key, (m, G) in res.items():
        np.testing.assert_array_equal(m, ref[0])
        for i in range(3):
            np.testing.assert_array_equal(G[i], ref[1][i], err_msg=str(key))


@pytest.mark.parametrize("variant", [1, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13])
def test_batched_uneven_equals_per_shot(psv, O, variant):
    c = cases.small_case(O, nshot=4, nt=260, uneven=True)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    out = {}
    for batch in (4, 1, 3):
        with psv.Simulation(c, batch=batch, kernel_variant=variant) as sim:
            assert sim.plan()["batch"] <= batch
            seis = sim.forward()
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
            m = sim.gradient()
            out[batch] = (seis, m, sim.get_gradient())
    for batch in (1, 3):
        for i in range(2):
            np.testing.assert_array_equal(out[batch][0][i], out[4][0][i])
        np.testing.assert_array_equal(out[batch][1], out[4][1])
        for i in range(3):
            np.testing.assert_array_equal(out[batch][2][i], out[4][2][i])
    # and the uneven receivers agree with the oracle shot by shot
    segs = psv.Simulation  # noqa: F841 (readability)
    with psv.Simulation(c) as sim:
        uz, _ = sim.forward()
        for s, a in enumerate(sim.split(uz)):
            assert a.shape == obs[s][0].shape
            assert cases.rel_l2(a, obs[s][0]) < 1e-4


@pytest.mark.parametrize("variant", [5, 8, 10, 11, 12, 13])
def test_checkpointed_gradient_independent_of_batch_and_tuning(psv, O, variant):
    """Checkpointed storage (segments of 30 steps, the last three kept by the forward pass,
    snapshots every 3rd step): the misfits and gradient of three uneven shots are the same
    bits whether the shots run 3, 2 + 1
    or 1 at a time per launch, and whether the autotuner timed its candidates in the
    context first (on the context's own buffers) or the kernel and work order were
    given, so nothing the tuner leaves behind reaches a result."""
    c = cases.small_case(O, nshot=3, nt=270, snap_dt=3, uneven=True)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    out, pin = {}, None
    for batch, pinned in ((3, False), (1, True), (2, True), (1, False), (3, True)):
        kw = dict(kernel_variant=variant)
        if pinned:
            kw.update(pin)
        with psv.Simulation(c, batch=batch, storage=psv._lib.STORE_CHECKPOINT, ckpt_interval=30, ckpt_keep=3,
                            **kw) as sim:
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
            m = sim.gradient()
            out[(batch, pinned)] = (m, sim.get_gradient())
            if pin is None:
                import re
                g = re.search(r"chosen o(\d)/v(\d+)/tz(\d+)", sim.tuning()[1])
                pin = dict(work_order=int(g.group(1)) + 1, tile_tz=int(g.group(3)))
    ref = out[(3, False)]
    for key, (m, G) in out.items():
        np.testing.assert_array_equal(m, ref[0], err_msg=str(key))
        for i in range(3):
            np.testing.assert_array_equal(G[i], ref[1][i], err_msg=str(key))
    _print(f"variant {variant}: batches 3, 2 + 1, 1 x 3, tuned or pinned {pin}: identical")


def test_plain_form_matches_fused(psv, O):
    c = cases.small_case(O, nshot=2, nt=240)
    with psv.Simulation(c) as fused, psv.Simulation(c, form=psv._lib.FORM_PLAIN) as plain:
        assert plain.plan()["batch"] == 1
        a, b = fused.forward(), plain.forward()
        for i in range(2):
            assert cases.rel_l2(b[i], a[i]) < 1e-5


def test_step_length_vs_oracle(psv, O):
    """The bracketing search and parabola on the device reproduce the restated
    reference search along a descent direction (the model difference)."""
    c = cases.small_case(O, nshot=2, nt=450)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    D = [c["lam"] - lam0, c["mu"] - mu0, c["rho"] - rho0]
    L2_0 = O.misfit_all(c, lam0, mu0, rho0, obs)
    step_o, trials_o, _ = O.step_length(c, 0.01, L2_0, lam0, mu0, rho0, D, obs)
    with psv.Simulation(c) as sim:
        sim.set_model(lam0, mu0, rho0)
        sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
        assert sim.misfit() == pytest.approx(L2_0, rel=1e-4)
        sim.copy_mat()
        sim.set_gradient(*D)
        step, n = sim.step_length_PSV(0.01, L2_0)
    _print(f"step length: device {step:.6g} ({n} trials), oracle {step_o:.6g} ({len(trials_o)} trials)")
    assert n == len(trials_o)
    assert step == pytest.approx(step_o, rel=1e-3)


def test_fwi_iteration_runs_and_reduces_misfit_with_gpu_sign(psv, O):
    """Two simulate_fwi_PSV iterations with the GPU twin's gradient sign
    (descent): the misfit decreases from iteration to iteration."""
    c = cases.small_case(O, nshot=2, nt=450)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    with psv.Simulation(c, grad_sign=+1) as sim:
        sim.set_model(lam0, mu0, rho0)
        sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
        hist = sim.simulate_fwi_PSV(2)
        L2_after = sim.misfit()
        lam1, _, _ = sim.get_model()
    _print("fwi history", hist, "final misfit", L2_after)
    assert hist[1][0] < hist[0][0] and L2_after < hist[1][0]
    assert np.abs(lam1 - lam0).max() > 0


def test_nonfinite_model_is_reported(psv, O):
    c = cases.small_case(O, nshot=1, nt=50)
    lam = c["lam"].copy()
    lam[30, 40] = np.nan
    with psv.Simulation(c) as sim:
        sim.set_model(lam, c["mu"], c["rho"])
        with pytest.raises(psv._lib.FwiError) as e:
            sim.forward()
        assert e.value.status == -4


def test_shared_cells_and_empty_shot(psv, O):
    """Two sources on one cell in one shot (added in source order, as vsrc2),
    two receivers on one cell, and a shot without receivers."""
    c = cases.small_case(O, nshot=2, nt=200)
    c["z_src"] = np.array([c["z_src"][0], c["z_src"][0], c["z_src"][1]], np.int32)
    c["x_src"] = np.array([c["x_src"][0], c["x_src"][0], c["x_src"][1]], np.int32)
    c["stf_z"] = np.stack([c["stf_z"][0], 0.5 * c["stf_z"][0], c["stf_z"][1]])
    c["stf_x"] = np.stack([c["stf_x"][0], 0.5 * c["stf_x"][0], c["stf_x"][1]])
    c["shot_to_fire"] = np.array([0, 0, 1], np.int32)
    c["z_rec"] = np.array([5, 5, 6], np.int32)
    c["x_rec"] = np.array([40, 40, 50], np.int32)
    c["rec_off"] = np.array([0, 3, 3], np.int32)
    med = O.mat_av2(c)
    ref = O.kernel_psv(cases.shot_case(c, 0), med, 0, accu=False)
    with psv.Simulation(c) as sim:
        uz, ux = sim.forward()
        assert uz.shape == (3, int(c["nt"]))
        assert cases.rel_l2(uz, ref["uz"]) < 1e-4
        np.testing.assert_array_equal(uz[0], uz[1])


@pytest.mark.parametrize("fdorder", [2, 8])
@pytest.mark.parametrize("variant,blocks", [(5, 16), (5, 32), (6, 1), (6, 2)])
def test_uniform_addressing_variant_matches_variant2(psv, O, fdorder, variant, blocks):
    """Variants 5 (wave-uniform addressing, ordered loads, strip profiles from the
    kernel arguments, halo stores into the pad column) and 6 (the same, every
    load one row ahead in register rings; segments of one or two blocks of
    fdorder + 2 rows) against variant 2, the same stencil: seismograms, stored
    snapshots, the wavefield, the shot gradient and the misfit-derivative sums,
    over uneven shots in one batch.  The 72-row test grid leaves a short last
    segment, so the guarded remainder rows run too."""
    c = cases.small_case(O, nshot=3, nt=240, fdorder=fdorder, uneven=True)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    tz = blocks if variant == 5 else blocks * (fdorder + 2)
    out = {}
    for v in (2, variant):
        with psv.Simulation(c, kernel_variant=v, tile_tz=tz if v == variant else 16,
                            storage=psv._lib.STORE_FULL) as sim:
            seis = sim.forward()
            wf = sim.wavefield(1, "szx")
            _, snaps = sim.forward_snapshots(1)
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
            m = sim.gradient()
            G = sim.get_gradient()
            adj = [np.asarray(o, np.float64) * 1e-3 for o in obs[0]]
            mg = sim.misfit_gradient(*adj, shot0=0, nshots=1)
            out[v] = (seis, wf, snaps, m, G, mg)
    got, ref = out[variant], out[2]
    tag = f"order {fdorder} variant {variant} tz {tz}"
    for i in range(2):
        e = cases.rel_l2(got[0][i], ref[0][i])
        _print(f"{tag}: seismogram {i} rel-L2 vs variant 2 {e:.2e}")
        assert e < 1e-6
    assert cases.rel_l2(got[1], ref[1]) < 1e-6
    for k in ref[2]:
        assert cases.rel_l2(got[2][k], ref[2][k]) < 1e-6, k
    np.testing.assert_allclose(got[3], ref[3], rtol=1e-5)
    for i in range(3):
        e = cases.rel_l2(got[4][i], ref[4][i])
        e2 = cases.rel_l2(got[5][i], ref[5][i])
        _print(f"{tag}: gradient {i} rel-L2 vs variant 2 {e:.2e}, misfit derivative {e2:.2e}")
        assert e < 1e-4 and e2 < 1e-4


@pytest.mark.parametrize("fdorder,nx,nshot", [(2, 96, 3), (4, 133, 5), (8, 96, 3), (8, 196, 4), (8, 300, 4)])
@pytest.mark.parametrize("tz", [16, 32])
def test_full_tiles_equal_variant5(psv, O, fdorder, nx, nshot, tz):
    """Variant 10 runs variant 5's arithmetic on 64-column tiles from column 0 with the x
    halo from a separate load (lane_off_h) instead of lanes given to it, variant 11 runs
    variant 10 with a block's four wavefronts on four shots of one tile, variant 12 with
    them on four adjacent tiles of one segment, and variant 13 is variant 12 with the
    non-temporal hint on single-use loads and stores: every tap reads the same value in
    the same order, so seismograms, stored snapshots, the wavefield, misfits, the gradient
    and the misfit-derivative sums equal variant 5's bit for bit.  Widths where nx - N
    is, and is not, a multiple of 64 (a last tile with lanes past the grid) and 2 to 5
    tiles (blocks of variant 12 with idle wavefronts past the last tile); CPML on the left
    and right (x strips inside the first and last tiles); uneven shots in one batch, 3, 4
    or 5 of them (a last group of four with idle wavefronts)."""
    c = cases.small_case(O, nshot=nshot, nt=200, fdorder=fdorder, uneven=True, nx=nx)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    out = {}
    for v in (5, 10, 11, 12, 13):
        with psv.Simulation(c, kernel_variant=v, tile_tz=tz, storage=psv._lib.STORE_FULL) as sim:
            seis = sim.forward()
            wf = sim.wavefield(2, "sxx")
            _, snaps = sim.forward_snapshots(1)
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
            m = sim.gradient()
            G = sim.get_gradient()
            adj = [np.asarray(o, np.float64) * 1e-3 for o in obs[0]]
            mg = sim.misfit_gradient(*adj, shot0=0, nshots=1)
            out[v] = (seis, wf, snaps, m, G, mg)
    b = out[5]
    for v in (10, 11, 12, 13):
        a = out[v]
        for i in range(2):
            np.testing.assert_array_equal(a[0][i], b[0][i], err_msg=f"variant {v}")
        np.testing.assert_array_equal(a[1], b[1], err_msg=f"variant {v}")
        for k in b[2]:
            np.testing.assert_array_equal(a[2][k], b[2][k], err_msg=f"variant {v} {k}")
        np.testing.assert_array_equal(a[3], b[3], err_msg=f"variant {v}")
        for i in range(3):
            np.testing.assert_array_equal(a[4][i], b[4][i], err_msg=f"variant {v}")
            np.testing.assert_array_equal(a[5][i], b[5][i], err_msg=f"variant {v}")
    assert np.abs(b[0][0]).max() > 0 and np.abs(b[4][0]).max() > 0
    _print(f"order {fdorder} nx {nx} shots {nshot} tz {tz}: variants 10 and 11 equal variant 5 bit for bit")


@pytest.mark.parametrize("variant,tz,fdorder", [(4, 0, 8), (7, 4, 8), (7, 8, 8), (7, 4, 2), (7, 8, 4),
                                                 (8, 4, 8), (8, 8, 8), (8, 4, 2)])
def test_one_pass_step_matches_two_kernel_step(psv, O, variant, tz, fdorder):
    """The one-pass variants (stress, surface and velocity in one pass, two time
    levels; 7 and 8 with segments of tz blocks of fdorder + 2 rows, 8 with a
    block's wavefronts side by side in x) against variant 1:
    seismograms, stored snapshots (the source cells are rewritten by the
    injection), the shot gradient and the misfit-derivative sums, over uneven
    shots in one batch.  Same arithmetic, so the differences are FP32 contraction
    choices only.  The 72-row grid leaves a short last segment."""
    c = cases.small_case(O, nshot=3, nt=300, fdorder=fdorder, uneven=True)
    med_true = O.mat_av2(c)
    obs = cases.per_shot_obs(c, O, med_true)
    lam0, mu0, rho0 = cases.start_model(c)
    out = {}
    for v in (1, variant):
        t = tz * (fdorder + 2) if v in (7, 8) else 0
        with psv.Simulation(c, kernel_variant=v, tile_tz=t, storage=psv._lib.STORE_FULL) as sim:
            seis = sim.forward()
            _, snaps = sim.forward_snapshots(1)
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
            m = sim.gradient()
            G = sim.get_gradient()
            adj = [np.asarray(o, np.float64) * 1e-3 for o in obs[0]]
            mg = sim.misfit_gradient(*adj, shot0=0, nshots=1)
            out[v] = (seis, snaps, m, G, mg)
    got, ref = out[variant], out[1]
    for i in range(2):
        e = cases.rel_l2(got[0][i], ref[0][i])
        _print(f"variant {variant} tz {tz} order {fdorder} vs variant 1: seismogram {i} rel-L2 {e:.2e}")
        assert e < 1e-6
    for k in ref[1]:
        e = cases.rel_l2(got[1][k], ref[1][k])
        assert e < 1e-6, (k, e)
    np.testing.assert_allclose(got[2], ref[2], rtol=1e-5)
    for i in range(3):
        e = cases.rel_l2(got[3][i], ref[3][i])
        e2 = cases.rel_l2(got[4][i], ref[4][i])
        _print(f"variant {variant} tz {tz} order {fdorder} vs variant 1: gradient {i} rel-L2 {e:.2e}, "
               f"misfit derivative {e2:.2e}")
        assert e < 1e-4 and e2 < 1e-4


def test_work_orders_agree_bitwise(psv, O):
    """The work order only changes which workgroup computes which tile (variant 8:
    with a block's wavefronts side by side in x, the 96-column grid leaves a block
    with wavefronts past the last tile)."""
    c = cases.small_case(O, nshot=3, nt=120)
    ref = None
    for v, tz in ((4, 32), (1, 8), (7, 40), (8, 40)):
        for order in (1, 2, 3, 4):
            with psv.Simulation(c, kernel_variant=v, tile_tz=tz, work_order=order) as sim:
                uz, ux = sim.forward()
            if ref is None or ref[0] != v:
                ref = (v, uz, ux)
            np.testing.assert_array_equal(uz, ref[1])
            np.testing.assert_array_equal(ux, ref[2])


@pytest.mark.parametrize("variant", [1, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13])
@pytest.mark.parametrize("storage", ["full", "ckpt"])
def test_graph_replay_equals_eager(psv, O, variant, storage):
    """use_graph: chunks of time steps captured once per propagation and replayed
    (64 steps, or one checkpoint segment), with eager steps for the remainder.  The
    same kernels see the same arguments, so seismograms, misfits and gradients are
    bit-identical to the eager launches; two gradients in a row re-capture.  With
    checkpoints the last two of the eight segments are kept by the forward pass (the
    last one short: eager), the other six recomputed."""
    c = cases.small_case(O, nshot=3, nt=300, snap_dt=2, uneven=True)
    obs = cases.per_shot_obs(c, O, O.mat_av2(c))
    lam0, mu0, rho0 = cases.start_model(c)
    st = psv._lib.STORE_FULL if storage == "full" else psv._lib.STORE_CHECKPOINT
    out = {}
    for g in (0, 1):
        with psv.Simulation(c, kernel_variant=variant, storage=st, ckpt_interval=40, use_graph=g,
                            batch=2, ckpt_keep=2) as sim:
            seis = sim.forward()
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
            m1 = sim.gradient()
            m2 = sim.gradient()
            out[g] = (seis, m1, m2, sim.get_gradient())
    for i in range(2):
        np.testing.assert_array_equal(out[1][0][i], out[0][0][i])
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][2], out[0][2])
    for i in range(3):
        np.testing.assert_array_equal(out[1][3][i], out[0][3][i])


@pytest.mark.parametrize("variant", [4, 7, 8, 9])
def test_one_pass_step_refused_where_it_does_not_apply(psv, O, variant):
    c = cases.small_case(O, nshot=1, nt=20)
    c["isurf"] = np.array([int(c["fdorder"]) // 2, int(c["nz"]) - 20, 0, 0], np.int32)   # a bottom surface too
    with psv.Simulation(c, kernel_variant=variant) as sim:
        with pytest.raises(psv._lib.FwiError) as e:
            sim.forward()
        assert e.value.status == -1


def test_kept_units_follow_the_memory_budget(psv, O):
    """The planner keeps as many final units' snapshots as the budget holds beside the batch
    and one checkpoint per remaining boundary: over rising budgets (at the same batch) the kept
    count rises to all units, the allocation stays within the budget, and at a budget that keeps some
    but not all the gradient is still bit-identical to full snapshot storage."""
    c = cases.small_case(O, nshot=2, nt=300, snap_dt=3, uneven=True)
    obs = cases.per_shot_obs(c, O, O.mat_av2(c))
    lam0, mu0, rho0 = cases.start_model(c)
    S, units = 15, 20
    sizes = {}
    for keep in (1, 2):
        with psv.Simulation(c, storage=psv._lib.STORE_CHECKPOINT, ckpt_interval=S, ckpt_keep=keep) as sim:
            p = sim.plan()
            assert p["ckpt_units"] == units and p["ckpt_kept"] == keep
            sizes[keep] = p["device_bytes"]
    per_unit = sizes[2] - sizes[1]   # one unit of snapshots less one checkpoint
    assert per_unit > 0
    kept, middle = [], None
    for j in range(0, 40, 2):        # the planner's estimate includes a fixed reserve above the arrays
        budget = int(sizes[1] + (64 << 20) + j * per_unit)
        try:
            sim = psv.Simulation(c, storage=psv._lib.STORE_CHECKPOINT, ckpt_interval=S, mem_budget=budget)
        except psv._lib.FwiError:    # below what one unit kept with a checkpoint per boundary needs
            continue
        with sim:
            p = sim.plan()
            assert p["device_bytes"] <= budget and p["ckpt_slots"] == max(units - p["ckpt_kept"] - 1, 0)
            if p["batch"] < 2:       # the budget holds one shot only (which then keeps more units)
                continue
            kept.append(p["ckpt_kept"])
            if middle is None and 1 < p["ckpt_kept"] < units:
                middle = budget
    _print(f"kept units over rising budgets (batch 2): {kept}")
    assert kept == sorted(kept) and kept[-1] == units and middle is not None
    res = {}
    for name, kw in (("full", dict(storage=psv._lib.STORE_FULL)),
                     ("budget", dict(storage=psv._lib.STORE_CHECKPOINT, ckpt_interval=S, mem_budget=middle))):
        with psv.Simulation(c, **kw) as sim:
            sim.set_model(lam0, mu0, rho0)
            sim.set_observed([o[0] for o in obs], [o[1] for o in obs])
            res[name] = (sim.gradient(), sim.get_gradient())
    np.testing.assert_array_equal(res["budget"][0], res["full"][0])
    for i in range(3):
        np.testing.assert_array_equal(res["budget"][1][i], res["full"][1][i])
