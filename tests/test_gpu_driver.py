"""The fwi_psv host driver end to end on the reference's binary files.

The fixture ``fwi_bin.npz`` holds the reference executable's inputs (metaint,
intarray, metafloat, mat) and outputs for a 40 x 56 grid, 2 shots, order 2:
the seismogram files of its forward mode, snapshot cubes of shot 1 (write_accu),
and, in FWI mode from the homogeneous start model, the per-shot cumulative L2
it prints before its line search fails (the CPU gradient sign climbs the misfit,
src/cpu/n_step_length_PSV.cpp:173-178).  The driver reads the same files and
writes the same layouts (src/h_io_PSV.cpp).
"""
import subprocess

import numpy as np
import pytest

import cases
from conftest import ROOT, gpu_available, load_fixture

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

DRIVER = ROOT / "fwi_cuda_amd" / "fwi_psv"


def _write(d, files, lam, mu, rho):
    (d / "bin").mkdir(exist_ok=True)
    (d / "io").mkdir(exist_ok=True)
    files["metaint"].astype(np.int32).tofile(d / "bin" / "metaint.bin")
    files["intarray"].astype(np.int32).tofile(d / "bin" / "intarray.bin")
    files["metafloat"].astype(np.float64).tofile(d / "bin" / "metafloat.bin")
    np.concatenate([lam.ravel(), mu.ravel(), rho.ravel()]).tofile(d / "bin" / "mat.bin")


def _run(d, *args):
    assert DRIVER.exists(), "build the driver: python -m fwi_cuda_amd.build --driver"
    return subprocess.run([str(DRIVER), str(d), *args], capture_output=True, text=True, timeout=300)


def test_forward_files_match_reference_executable(tmp_path):
    fx = load_fixture("fwi_bin")
    nt = int(fx["files_fwd"]["metaint"][8])
    _write(tmp_path, fx["files_fwd"], fx["true"]["lam"], fx["true"]["mu"], fx["true"]["rho"])
    r = _run(tmp_path)
    assert r.returncode == 0, r.stderr
    for s in range(2):
        for comp in ("uz", "ux"):
            got = np.fromfile(tmp_path / "bin" / f"shot{s}_rtf_{comp}.bin", np.float64).reshape(-1, nt)
            e = cases.rel_l2(got, fx["obs"][f"shot{s}_{comp}"])
            print(f"driver shot {s} {comp}: rel-L2 vs reference executable {e:.2e}", flush=True)
            assert e < 2e-5
    pick = fx["accu_pick"]
    for f in ("vz", "vx", "szz", "szx", "sxx"):
        ref = fx["accu1"][f]
        got = np.fromfile(tmp_path / "bin" / f"shot1_{f}.bin", np.float64).reshape((-1,) + ref.shape[1:])[pick]
        assert cases.rel_l2(got, ref) < 1e-4, f


def test_fwi_mode_reproduces_reference_misfits_and_failed_line_search(tmp_path):
    fx = load_fixture("fwi_bin")
    _write(tmp_path, fx["files_fwi"], fx["start"]["lam"], fx["start"]["mu"], fx["start"]["rho"])
    for s in range(2):
        for comp in ("uz", "ux"):
            fx["obs"][f"shot{s}_{comp}"].tofile(tmp_path / "io" / f"shot{s}_rtf_{comp}.bin")
    r = _run(tmp_path, "--niter", "1")
    assert r.returncode == 2, (r.returncode, r.stdout[-2000:], r.stderr)
    l2 = [float(line.split(",")[1]) for line in r.stdout.splitlines() if line.startswith("L2 NORM:")]
    np.testing.assert_allclose(l2, fx["l2_printed"], rtol=2e-4)
    assert "Steplength estimation failed" in r.stdout


def test_fwi_mode_with_fp64_seismograms_matches_the_reference_to_rounding(tmp_path):
    """--seis-fp64: the misfits the driver prints come from the FP64 propagation, so they equal
    the reference executable's in all six digits it prints (with FP32 seismograms the sixth
    digit differs: 4.88013e-13 against 4.88014e-13), and the line search fails where the
    reference's does."""
    fx = load_fixture("fwi_bin")
    _write(tmp_path, fx["files_fwi"], fx["start"]["lam"], fx["start"]["mu"], fx["start"]["rho"])
    for s in range(2):
        for comp in ("uz", "ux"):
            fx["obs"][f"shot{s}_{comp}"].tofile(tmp_path / "io" / f"shot{s}_rtf_{comp}.bin")
    r = _run(tmp_path, "--niter", "1", "--seis-fp64")
    assert r.returncode == 2, (r.returncode, r.stdout[-2000:], r.stderr)
    l2 = [float(line.split(",")[1]) for line in r.stdout.splitlines() if line.startswith("L2 NORM:")]
    print("driver L2 norms with FP64 seismograms:", l2, "reference:", list(fx["l2_printed"]), flush=True)
    assert [f"{v:.5e}" for v in l2] == [f"{v:.5e}" for v in fx["l2_printed"]]
    assert "Steplength estimation failed" in r.stdout


def test_fwi_mode_with_descent_sign_writes_models(tmp_path):
    """With the GPU twin's sign the same inversion descends: two iterations, the
    misfit drops and iter*_mat.bin (lam, mu, rho FP64) are written each iteration."""
    fx = load_fixture("fwi_bin")
    _write(tmp_path, fx["files_fwi"], fx["start"]["lam"], fx["start"]["mu"], fx["start"]["rho"])
    for s in range(2):
        for comp in ("uz", "ux"):
            fx["obs"][f"shot{s}_{comp}"].tofile(tmp_path / "io" / f"shot{s}_rtf_{comp}.bin")
    r = _run(tmp_path, "--niter", "2", "--grad-sign", "1")
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr)
    line = next(ln for ln in r.stdout.splitlines() if ln.startswith("L2 norms:"))
    norms = [float(v) for v in line[len("L2 norms:"):].split(",") if v.strip()]
    assert len(norms) == 2 and norms[1] < norms[0]
    nz, nx = fx["start"]["lam"].shape
    for it in range(2):
        m = np.fromfile(tmp_path / "bin" / f"iter{it}_mat.bin", np.float64)
        assert m.size == 3 * nz * nx and np.isfinite(m).all()


def test_driver_runs_input_files_written_from_a_workload(tmp_path):
    """fwi_cuda_amd.refio writes a seeded workload (C1, shortened) as the reference's
    input files; the driver's forward mode reads them and writes seismograms equal
    to the library's forward of the same case with the reference's horizontal force
    (src/h_simulate_PSV.cpp:183-184)."""
    from fwi_cuda_amd import configs, psv, refio
    nt = 300
    c = configs.c1(nt=nt)
    meta, arrays = refio.inputs_from_case(c)
    refio.write_inputs(tmp_path, meta, arrays, c["lam"], c["mu"], c["rho"])
    r = _run(tmp_path)
    assert r.returncode == 0, r.stderr
    h = dict(c)
    w = psv.wavelet(nt, c["dt"], 1.0, c["fc"])
    h["stf_z"] = np.zeros_like(c["stf_z"])
    h["stf_x"] = np.tile(w, (len(c["z_src"]), 1))
    with psv.Simulation(h) as sim:
        uz, ux = sim.forward()
    for comp, ref in (("uz", uz), ("ux", ux)):
        got = np.fromfile(tmp_path / "bin" / f"shot0_rtf_{comp}.bin", np.float64).reshape(-1, nt)
        e = cases.rel_l2(got, ref)
        print(f"driver on refio inputs, {comp}: rel-L2 vs the library's forward {e:.2e}", flush=True)
        assert e < 1e-6


def test_stream_probe_reports_its_bandwidth():
    """bench.py's access-shape probe runs, checks its arguments and prints one JSON
    line whose byte count is the 15 arrays it streams."""
    import json
    probe = ROOT / "measure" / "stream_probe"
    assert probe.exists(), "build it: make -C measure"
    r = subprocess.run([str(probe), "201", "512", "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["bytes_per_launch"] == 15 * 201 * 512 * 4 * 4
    assert d["shape_GBps"] > 100
    bad = subprocess.run([str(probe), "201", "500", "4"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 2
