"""The C-ABI library without a GPU: it loads, exports every entry point the
header declares with the layouts the Python binding assumes, its host-side
preprocessing reproduces the compiled reference bit for bit (golden
``pieces.npz``), and the device entry points refuse to run without a device
instead of falling back to the CPU."""
import ctypes as C
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, gpu_available

HEADER = ROOT / "include" / "fwi_psv.h"


@pytest.fixture(scope="module")
def lib():
    from fwi_cuda_amd import _lib
    return _lib.load()


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fwi_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(lib):
    from fwi_cuda_amd import _lib
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fwi_[a-z0-9_]+)", out))
    assert set(names) <= exported, set(names) - exported


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof / offsetof of the ABI structs from the C compiler against ctypes."""
    from fwi_cuda_amd import _lib
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "fwi_psv.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu %d\\n", sizeof(fwi_grid), sizeof(fwi_acq), sizeof(fwi_opts),\n'
        "         sizeof(fwi_kernel_time), (int)FWI_K_COUNT);\n"
        '  printf("%zu %zu %zu %zu %zu %zu\\n", offsetof(fwi_grid, taper), offsetof(fwi_opts, work_order),\n'
        "         offsetof(fwi_acq, rec_x), offsetof(fwi_opts, param), offsetof(fwi_opts, vel_bounds),\n"
        "         offsetof(fwi_opts, ckpt_keep));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(v) for v in lines[0].split()]
    offs = [int(v) for v in lines[1].split()]
    assert sizes[:4] == [C.sizeof(_lib.FwiGrid), C.sizeof(_lib.FwiAcq), C.sizeof(_lib.FwiOpts),
                         C.sizeof(_lib.FwiKernelTime)]
    assert sizes[4] == len(_lib.KERNELS)
    assert offs == [_lib.FwiGrid.taper.offset, _lib.FwiOpts.work_order.offset, _lib.FwiAcq.rec_x.offset,
                    _lib.FwiOpts.param.offset, _lib.FwiOpts.vel_bounds.offset, _lib.FwiOpts.ckpt_keep.offset]


def test_default_options(lib):
    from fwi_cuda_amd import _lib
    o = _lib.FwiOpts()
    lib.fwi_default_opts(C.byref(o))
    assert (o.storage, o.form, o.grad_sign, o.check_finite, o.kernel_variant, o.work_order) == \
        (_lib.STORE_AUTO, _lib.FORM_FUSED, -1, 1, 0, 0)
    assert lib.fwi_abi_version() == _lib.ABI_VERSION == 6
    assert (o.pcg, o.smooth_hfs) == (0, 0)
    assert o.param == _lib.PARAM_LAME and list(o.vel_bounds) == [0.0] * 4 and o.ckpt_slots == 0
    assert o.seis_fp64 == 0 and o.ckpt_keep == 0
    assert lib.fwi_strerror(-8)


def test_host_preprocessing_matches_reference():
    """holbergcoeff, cpml_PSV and wavelet of the library against the values the
    reference's own CPU sources computed (bit for bit)."""
    from fwi_cuda_amd import psv
    z = np.load(GOLDEN / "pieces.npz", allow_pickle=False)
    table = z["holberg"]
    for e in range(table.shape[0]):
        for n in range(table.shape[1]):
            np.testing.assert_array_equal(psv.holbergcoeff(2 * (n + 1), e), table[e, n])
    for i in range(4):
        args = z[f"cpml.{i}.args"]
        out = psv.cpml_PSV(*[float(a) if k not in (5, 6, 7, 8) else int(a) for k, a in enumerate(args)])
        for key in ("a", "b", "K", "a_half", "b_half", "K_half"):
            np.testing.assert_array_equal(out[key], z[f"cpml.{i}.{key}"], err_msg=f"cpml {i} {key}")
    nt, dt, amp, fc, ts = z["wavelet.args"]
    np.testing.assert_array_equal(psv.wavelet(int(nt), dt, amp, fc, ts), z["wavelet.s"])


def test_stability_limit_closed_form():
    from fwi_cuda_amd import psv
    nz, nx = 20, 30
    rho = np.full((nz, nx), 2000.0)
    mu = rho * 1500.0 ** 2
    lam = rho * 3000.0 ** 2 - 2 * mu
    lam[5, 7] = rho[5, 7] * 4000.0 ** 2 - 2 * mu[5, 7]    # the fastest cell sets the limit
    hc = psv.holbergcoeff(8)
    dts = psv.checkfd_dtstab(lam, mu, rho, hc, 1.0, 0)
    want = 1.0 / (np.sqrt(2.0) * np.abs(hc[1:5]).sum() * 4000.0)
    assert dts == pytest.approx(want, rel=1e-12)
    # the oracle's restatement: the same closed form; cells in the skipped first row and
    # column, and inside the fw-cell margin, do not set the limit
    import sys
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as O
    O.load("oracle")
    assert O.dtstab(lam, mu, rho, hc, 1.0, 0) == pytest.approx(want, rel=1e-12)
    lam2 = lam.copy()
    lam2[0, 3] = rho[0, 3] * 5000.0 ** 2 - 2 * mu[0, 3]     # row 0 is outside the scan
    lam2[8, 28] = rho[8, 28] * 6000.0 ** 2 - 2 * mu[8, 28]  # column nx - 2, past nx - fw for fw = 3
    assert O.dtstab(lam2, mu, rho, hc, 1.0, 3) == pytest.approx(want, rel=1e-12)
    assert O.dtstab(lam2, mu, rho, hc, 1.0, 0) == pytest.approx(want * 4000.0 / 6000.0, rel=1e-12)
    assert O.dtstab(lam, mu, rho, hc, 0.5, 0) == pytest.approx(want / 2, rel=1e-12)


@pytest.mark.skipif(gpu_available(), reason="checks the behaviour without a device")
def test_no_device_no_fallback():
    """Without a HIP device the context is refused (FWI_ENODEV): the HIP path has no CPU fallback."""
    import sys
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle import oracle as O
    import cases
    from fwi_cuda_amd import _lib, psv
    O.load("oracle")
    c = cases.small_case(O, nshot=1, nt=10)
    with pytest.raises(_lib.FwiError) as e:
        psv.Simulation(c)
    assert e.value.status == -8


def test_invalid_options_are_rejected_before_any_device_work():
    """Bad tile rows / variants fail as FWI_EINVAL (validation precedes the device check)."""
    import sys
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle import oracle as O
    import cases
    from fwi_cuda_amd import _lib, psv
    O.load("oracle")
    c = cases.small_case(O, nshot=1, nt=10)
    for kw in ({"tile_tz": 5}, {"kernel_variant": 14}, {"work_order": 5}, {"kernel_variant": 9, "tile_tz": 3},
               {"ckpt_slots": -1}, {"ckpt_keep": -1}, {"smooth_hfs": 1}):
        with pytest.raises(_lib.FwiError) as e:
            psv.Simulation(c, **kw)
        assert e.value.status == -1, kw


@pytest.mark.parametrize("name,kw,shape,nshot", [
    ("C1", dict(nt=20), (201, 401), 1),
    ("C2", dict(nt=20), (301, 601), 8),
    ("C3", dict(nt=20, nshot=6, nz=201, nx=401), (201, 401), 6),
    ("C4", dict(nt=20, nshot=3, nz=301, nx=601), (301, 601), 3),
    ("C5", dict(nt=20, nshot=4, nz=201, nx=401), (201, 401), 4),
])
def test_workload_builders(name, kw, shape, nshot):
    """The seeded workloads: shapes, shot counts, receivers inside the grid, dt at
    half the stencil's limit, a free surface at the top row of the interior."""
    from fwi_cuda_amd import configs, psv
    c = configs.BUILDERS[name](**kw)
    assert (int(c["nz"]), int(c["nx"])) == shape and int(c["nshot"]) == nshot
    hc = psv.holbergcoeff(8)
    dts = psv.checkfd_dtstab(c["lam"], c["mu"], c["rho"], hc, float(c["dx"]), configs.NPML)
    assert float(c["dt"]) == pytest.approx(0.5 * dts, rel=1e-12)
    assert list(c["isurf"]) == [4, 0, 0, 0]
    if "rec_off" in c:
        counts = np.diff(c["rec_off"])
        assert len(counts) == nshot and counts.min() >= 1
    zr, xr = np.asarray(c["z_rec"]), np.asarray(c["x_rec"])
    assert (zr >= 4).all() and (xr >= configs.NPML).all() and (xr < shape[1] - configs.NPML).all()
    # identical inputs from the same seed
    c2 = configs.BUILDERS[name](**kw)
    np.testing.assert_array_equal(c["lam"], c2["lam"])


def test_velocity_parameter_chain_rule():
    """(lam, mu, rho) gradient -> (cp, cs, rho) gradient against central differences
    of a nonlinear test functional J(lam, mu, rho) evaluated through cp, cs."""
    from fwi_cuda_amd import psv
    rng = np.random.default_rng(3)
    cp = rng.uniform(2000, 4500, 50)
    cs = cp / rng.uniform(1.6, 2.0, 50)
    rho = rng.uniform(1800, 2600, 50)
    w = rng.standard_normal((3, 50))

    def lame(cp_, cs_, rho_):
        mu_ = rho_ * cs_ ** 2
        return rho_ * cp_ ** 2 - 2 * mu_, mu_, rho_

    def J(l, m, r):
        return np.sum(w[0] * np.sin(l / 1e10) + w[1] * (m / 1e10) ** 2 + w[2] * np.log(r))

    l, m, r = lame(cp, cs, rho)
    g_l = w[0] * np.cos(l / 1e10) / 1e10
    g_m = w[1] * 2 * m / 1e20
    g_r = w[2] / r
    g_cp, g_cs, g_rho = psv.lame_to_velocity_gradient(g_l, g_m, g_r, l, m, r)
    for k, (g, v) in enumerate(((g_cp, cp), (g_cs, cs), (g_rho, rho))):
        h = 1e-4 * v
        up, dn = [cp.copy(), cs.copy(), rho.copy()], [cp.copy(), cs.copy(), rho.copy()]
        fd = np.empty(50)
        for i in range(50):
            up[k][i] += h[i]
            dn[k][i] -= h[i]
            fd[i] = (J(*lame(*up)) - J(*lame(*dn))) / (2 * h[i])
            up[k][i] -= h[i]
            dn[k][i] += h[i]
        # central differences of the summed functional: truncation and cancellation ~1e-5
        np.testing.assert_allclose(g, fd, rtol=1e-4, atol=1e-12 * np.abs(fd).max())


def test_bench_names_the_tuned_instance_and_finds_its_traffic(tmp_path):
    """bench.py's roofline names the kernel instance the tuner chose (as the profiler
    names it) and takes that instance's counter traffic out of a summary only when the
    summary's line was measured at the same algorithmic bytes per launch (same grid,
    batch and launch mix): a line of the same instance at another batch is refused."""
    import sys
    sys.path.insert(0, str(ROOT))
    import bench
    log = "chosen o2/v8/tz40/d1; o2/v7/tz40/d1 4.880; "
    assert bench.kernel_instantiation("step", log, 4) == "k_step7<4, 4, *, true>"
    assert bench.kernel_instantiation("step", "chosen o2/v7/tz80/d1; ", 4) == "k_step7<4, 8, *, false>"
    assert bench.kernel_instantiation("step", "chosen o1/v9/tz2/d2; ", 4) == "k_step7<4, 4, *, true>"
    assert bench.kernel_instantiation("step", "chosen o0/v4/tz32/d1; ", 4) == "k_step<4, 32, *>"
    assert bench.kernel_instantiation("velocity", "chosen o1/v6/tz20/d1; ", 4) == "k_velocity6<4, 1, 2, *>"
    assert bench.kernel_instantiation("velocity", "chosen o1/v5/tz16/d1; ", 4) == "k_velocity5<4, 16, *>"
    assert bench.kernel_instantiation("velocity", "chosen o1/v10/tz32/d1; ", 4) == "k_velocity10<4, 32, *>"
    assert bench.kernel_instantiation("stress", "chosen o1/v10/tz16/d1; ", 4) == "k_stress10<4, 16>"
    assert bench.kernel_instantiation("velocity", "chosen o0/v11/tz16/d1; ", 4) == "k_velocity11<4, 16, *>"
    assert bench.kernel_instantiation("stress", "chosen o2/v2/tz16/d1; ", 4) == "k_stress2<4, 16, 1>"
    assert bench.kernel_instantiation("velocity", "chosen o3/v13/tz16/d1; ", 4) == "k_velocity13<4, 16, *>"
    assert bench.kernel_instantiation("stress", "chosen o3/v12/tz32/d1; ", 4) == "k_stress12<4, 32>"
    f = tmp_path / "traffic.txt"
    f.write_text("header\n"
                 "traffic k_step7<4, 4, *, false> gradient-mix bytes_per_launch=9.0e9 "
                 "algorithmic_bytes_per_launch=5.9e9 launches=10\n"
                 "traffic k_step7<4, 4, *, true> gradient-mix bytes_per_launch=8.0e9 "
                 "algorithmic_bytes_per_launch=5.9e9 launches=10\n"
                 "traffic k_step7<4, 4, *, true> gradient-mix bytes_per_launch=2.1e9 "
                 "algorithmic_bytes_per_launch=1.48e9 launches=10\n")
    assert bench.measured_traffic("k_step7<4, 4, *, false>", 5.9e9, f) == 9.0e9
    assert bench.measured_traffic("k_step7<4, 4, *, true>", 5.91e9, f) == 8.0e9
    # the same instance at a quarter of the batch picks its own line, and none at another size
    assert bench.measured_traffic("k_step7<4, 4, *, true>", 1.476e9, f) == 2.1e9
    assert bench.measured_traffic("k_step7<4, 4, *, true>", 2.95e9, f) is None
    assert bench.measured_traffic("k_step7<4, 8, *, true>", 5.9e9, f) is None
    # the committed summary parses and every line carries both byte counts
    lines = [l for l in (ROOT / "profiles" / "roofline_traffic.txt").read_text().splitlines()
             if l.startswith("traffic ")]
    assert lines and all("bytes_per_launch=" in l and "algorithmic_bytes_per_launch=" in l for l in lines)
    # the bench's default run (batch 8, 22 of 35 units kept) finds the tuned instance's line
    assert bench.measured_traffic("k_velocity13<4, 16, *>", 952501703.0) == pytest.approx(8.888624e8)


def test_bench_shot_groups_cover_the_workload():
    """A bench step is the gradient of one shot group; the groups tile the shots."""
    import sys
    sys.path.insert(0, str(ROOT))
    import bench
    assert bench.shot_groups(64, 8) == [(8 * k, 8) for k in range(8)]
    g = bench.shot_groups(10, 4)
    assert g == [(0, 4), (4, 4), (8, 2)] and sum(n for _, n in g) == 10


def test_reference_input_files_round_trip_the_fixture(tmp_path):
    """fwi_cuda_amd.refio reads the input files the reference executable consumed
    when the fwi_bin fixture was made (metaint, intarray, metafloat, mat in the
    layout of scripts/pre_proc_fwi.py:240-276) and writes them back byte for byte."""
    from conftest import load_fixture
    from fwi_cuda_amd import refio
    fx = load_fixture("fwi_bin")
    src, dst = tmp_path / "src", tmp_path / "dst"
    (src / "bin").mkdir(parents=True)
    files = fx["files_fwd"]
    files["metaint"].astype(np.int32).tofile(src / "bin" / "metaint.bin")
    files["intarray"].astype(np.int32).tofile(src / "bin" / "intarray.bin")
    files["metafloat"].astype(np.float64).tofile(src / "bin" / "metafloat.bin")
    lam, mu, rho = (fx["true"][k].astype(np.float64) for k in ("lam", "mu", "rho"))
    np.concatenate([lam.ravel(), mu.ravel(), rho.ravel()]).tofile(src / "bin" / "mat.bin")
    meta, arrays, l2, m2, r2 = refio.read_inputs(src)
    assert (meta["nt"], meta["nz"], meta["nx"], meta["nsrc"], meta["nrec"], meta["nshot"], meta["fdorder"]) == \
        (250, 40, 56, 2, 5, 2, 2)
    assert list(arrays["x_rec"]) == [12, 20, 28, 36, 44] and meta["fpad"] == 1
    np.testing.assert_array_equal(l2, lam)
    refio.write_inputs(dst, meta, arrays, l2, m2, r2)
    for f in ("metaint", "intarray", "metafloat", "mat"):
        assert (src / "bin" / f"{f}.bin").read_bytes() == (dst / "bin" / f"{f}.bin").read_bytes(), f


def test_reference_input_files_from_a_workload(tmp_path):
    """A seeded workload (C1) as the reference's input files: the metadata carries
    its grid, acquisition, snapshot box, taper and CPML parameters; a workload
    whose shots have different receivers (C3) is refused."""
    from fwi_cuda_amd import configs, refio
    c = configs.c1(nt=30)
    meta, arrays = refio.inputs_from_case(c, fwinv=True)
    refio.write_inputs(tmp_path, meta, arrays, c["lam"], c["mu"], c["rho"])
    m2, a2, lam, mu, rho = refio.read_inputs(tmp_path)
    assert m2 == meta
    for k in refio.ARRAYS:
        np.testing.assert_array_equal(a2[k], arrays[k])
    np.testing.assert_array_equal(lam, c["lam"])
    assert (m2["nz"], m2["nx"], m2["nrec"], m2["nshot"], m2["fdorder"], m2["surf"]) == (201, 401, 40, 1, 8, 1)
    assert list(a2["isurf"]) == [4, 0, 0, 0] and list(a2["npml"]) == [0, 20, 20, 20]
    assert m2["freq_pml"] == c["fc"] and m2["dt"] == c["dt"] and m2["fwinv"] == 1
    with pytest.raises(ValueError):
        refio.inputs_from_case(configs.c3(nt=10, nshot=3, nz=101, nx=201))
