"""The reference's binary input files, written and read from Python.

The reference reads four files from ``./bin/`` (written by
scripts/pre_proc_fwi.py:240-276, read by src/h_io_PSV.cpp):

  metaint.bin    int32 x 37: 8 switches (metabool) then 29 integers (META_INT)
  intarray.bin   int32: npml (top, bottom, left, right), isurf (same order),
                 z_src, x_src, src_shot_to_fire [nsrc], z_rec, x_rec [nrec]
  metafloat.bin  float64 x 11 (META_FLOAT)
  mat.bin        float64: lam, mu, rho, each [nz][nx]

and, in FWI mode, the observed seismograms ``./io/shot<s>_rtf_u<z|x>.bin``
(float64 [nrec][nt] per shot).  The fwi_psv driver (fwi_cuda_amd/fwi_psv)
reads the same files.

    meta, arrays = inputs_from_case(configs.c1(), fwinv=True)
    write_inputs(directory, meta, arrays, lam, mu, rho)

A case's receiver set must be the same for every shot: the files hold one.  The
reference executable, and the fwi_psv driver with it, fires a horizontal
force whatever the case's source components (src/h_simulate_PSV.cpp:183-184).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

# scripts/pre_proc_fwi.py:240-248
META_BOOL = ("cuda_computation", "surf", "pml_z", "pml_x", "accu_save", "seismo_save", "fwinv", "rtf_meas_true")
META_INT = ("nt", "nz", "nx", "snap_t1", "snap_t2", "snap_z1", "snap_z2", "snap_x1", "snap_x2", "snap_dt",
            "snap_dz", "snap_dx", "taper_t1", "taper_t2", "taper_b1", "taper_b2", "taper_l1", "taper_l2",
            "taper_r1", "taper_r2", "nsrc", "nrec", "nshot", "stf_type", "rtf_type", "fdorder", "fpad",
            "mat_save_interval", "mat_grid")
# scripts/pre_proc_fwi.py:260-261
META_FLOAT = ("dt", "dz", "dx", "pml_npower", "damp_v_pml", "rcoef", "k_max_pml", "freq_pml", "scalar_lam",
              "scalar_mu", "scalar_rho")
ARRAYS = ("npml", "isurf", "z_src", "x_src", "src_shot_to_fire", "z_rec", "x_rec")


def inputs_from_case(case: dict, *, fwinv: bool = False, accu_save: bool = False, seismo_save: bool = True,
                     mat_save_interval: int = 1) -> tuple[dict, dict]:
    """The reference's metadata and integer arrays for a case built by
    psv.make_case (or configs.C1..C5).  Raises ValueError for a case whose
    shots have different receiver sets."""
    c = case
    nshot = int(c["nshot"])
    z_rec, x_rec = np.asarray(c["z_rec"], np.int32), np.asarray(c["x_rec"], np.int32)
    if "rec_off" in c:
        off = np.asarray(c["rec_off"])
        n = int(off[1] - off[0])
        same = all(int(off[s + 1] - off[s]) == n and
                   np.array_equal(z_rec[off[s]:off[s + 1]], z_rec[:n]) and
                   np.array_equal(x_rec[off[s]:off[s + 1]], x_rec[:n]) for s in range(nshot))
        if not same:
            raise ValueError("the reference's files hold one receiver set for all shots")
        z_rec, x_rec = z_rec[:n], x_rec[:n]
    snap, taper = [int(v) for v in c["snap"]], [int(v) for v in c["taper"]]
    nt, fdorder = int(c["nt"]), int(c["fdorder"])
    meta = dict(cuda_computation=0, surf=int(c["surf"]), pml_z=int(c["pml_z"]), pml_x=int(c["pml_x"]),
                accu_save=int(accu_save), seismo_save=int(seismo_save), fwinv=int(fwinv), rtf_meas_true=int(fwinv),
                nt=nt, nz=int(c["nz"]), nx=int(c["nx"]), snap_t1=0, snap_t2=nt - 1,
                snap_z1=snap[0], snap_z2=snap[1], snap_x1=snap[2], snap_x2=snap[3],
                snap_dt=snap[4], snap_dz=snap[5], snap_dx=snap[6],
                taper_t1=taper[0], taper_t2=taper[1], taper_b1=taper[2], taper_b2=taper[3],
                taper_l1=taper[4], taper_l2=taper[5], taper_r1=taper[6], taper_r2=taper[7],
                nsrc=len(c["z_src"]), nrec=len(z_rec), nshot=nshot, stf_type=int(c["stf_type"]), rtf_type=0,
                fdorder=fdorder, fpad=fdorder // 2, mat_save_interval=int(mat_save_interval), mat_grid=1)
    npower, damp_v, rcoef, k_max = c["cpml_par"]
    meta.update(dt=float(c["dt"]), dz=float(c["dz"]), dx=float(c["dx"]), pml_npower=npower, damp_v_pml=damp_v,
                rcoef=rcoef, k_max_pml=k_max, freq_pml=float(c["fc"]), scalar_lam=float(c["lam"][0, 0]),
                scalar_mu=float(c["mu"][0, 0]), scalar_rho=float(c["rho"][0, 0]))
    arrays = dict(npml=np.asarray(c["npml"], np.int32), isurf=np.asarray(c["isurf"], np.int32),
                  z_src=np.asarray(c["z_src"], np.int32), x_src=np.asarray(c["x_src"], np.int32),
                  src_shot_to_fire=np.asarray(c["shot_to_fire"], np.int32), z_rec=z_rec, x_rec=x_rec)
    return meta, arrays


def write_inputs(directory, meta: dict, arrays: dict, lam, mu, rho) -> None:
    """metaint.bin, intarray.bin, metafloat.bin and mat.bin under directory/bin
    (directory/io is created for the observed data)."""
    d = Path(directory)
    (d / "bin").mkdir(parents=True, exist_ok=True)
    (d / "io").mkdir(exist_ok=True)
    np.array([meta[k] for k in META_BOOL + META_INT], np.int32).tofile(d / "bin" / "metaint.bin")
    np.concatenate([np.asarray(arrays[k], np.int32).ravel() for k in ARRAYS]).tofile(d / "bin" / "intarray.bin")
    np.array([meta[k] for k in META_FLOAT], np.float64).tofile(d / "bin" / "metafloat.bin")
    np.concatenate([np.asarray(a, np.float64).ravel() for a in (lam, mu, rho)]).tofile(d / "bin" / "mat.bin")


def read_inputs(directory) -> tuple[dict, dict, np.ndarray, np.ndarray, np.ndarray]:
    """The four files back: (meta, arrays, lam, mu, rho)."""
    d = Path(directory) / "bin"
    mi = np.fromfile(d / "metaint.bin", np.int32)
    mf = np.fromfile(d / "metafloat.bin", np.float64)
    if mi.size != len(META_BOOL) + len(META_INT) or mf.size != len(META_FLOAT):
        raise ValueError(f"metaint holds {mi.size} values, metafloat {mf.size}")
    meta = {k: int(v) for k, v in zip(META_BOOL + META_INT, mi)}
    meta.update({k: float(v) for k, v in zip(META_FLOAT, mf)})
    ia = np.fromfile(d / "intarray.bin", np.int32)
    sizes = dict(npml=4, isurf=4, z_src=meta["nsrc"], x_src=meta["nsrc"], src_shot_to_fire=meta["nsrc"],
                 z_rec=meta["nrec"], x_rec=meta["nrec"])
    if ia.size != sum(sizes.values()):
        raise ValueError(f"intarray holds {ia.size} values, the metadata {sum(sizes.values())}")
    arrays, p = {}, 0
    for k in ARRAYS:
        arrays[k] = ia[p:p + sizes[k]].copy()
        p += sizes[k]
    nz, nx = meta["nz"], meta["nx"]
    mat = np.fromfile(d / "mat.bin", np.float64)
    if mat.size != 3 * nz * nx:
        raise ValueError(f"mat.bin holds {mat.size} values, 3 x {nz} x {nx} expected")
    lam, mu, rho = (m.reshape(nz, nx) for m in np.split(mat, 3))
    return meta, arrays, lam, mu, rho


def write_observed(directory, seis_z, seis_x) -> None:
    """Observed seismograms [nshot][nrec][nt] (z and x components) as
    io/shot<s>_rtf_uz.bin and io/shot<s>_rtf_ux.bin."""
    d = Path(directory) / "io"
    d.mkdir(parents=True, exist_ok=True)
    for s, (z, x) in enumerate(zip(seis_z, seis_x)):
        np.asarray(z, np.float64).tofile(d / f"shot{s}_rtf_uz.bin")
        np.asarray(x, np.float64).tofile(d / f"shot{s}_rtf_ux.bin")
