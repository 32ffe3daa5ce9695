"""ctypes front end of the test-only checkers.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg import this module.  It drives two
libraries with one interface:

* ``liboracle_fwi.so`` -- the FP64 C restatement in ``oracle/fwi_oracle.c``;
* ``_ref/libref_psv.so`` -- the reference's own CPU sources compiled in place
  by ``oracle/Makefile`` (present only where ``/root/reference`` was there to
  build it, or where the built file travelled with the tree).

A *case* is a plain dict of numpy arrays and scalars (see :func:`make_case`),
so golden fixtures round-trip through ``numpy.savez``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle_fwi.so"
REF_SO = HERE / "_ref" / "libref_psv.so"
REF_BIN = HERE / "_ref" / "seis_fwi"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class OrGrid(C.Structure):
    _fields_ = [("nt", C.c_int32), ("nz", C.c_int32), ("nx", C.c_int32),
                ("dt", C.c_double), ("dz", C.c_double), ("dx", C.c_double),
                ("fdorder", C.c_int32), ("hc", C.c_double * 7),
                ("surf", C.c_int32), ("isurf", C.c_int32 * 4),
                ("pml_z", C.c_int32), ("pml_x", C.c_int32)] + \
        [(n, _dp) for n in ("a_z", "b_z", "K_z", "a_half_z", "b_half_z", "K_half_z",
                            "a_x", "b_x", "K_x", "a_half_x", "b_half_x", "K_half_x")] + \
        [(n, C.c_int32) for n in ("snap_z1", "snap_z2", "snap_x1", "snap_x2",
                                  "snap_dt", "snap_dz", "snap_dx")]


class OrMedia(C.Structure):
    _fields_ = [(n, _dp) for n in ("lam", "mu", "mu_zx", "rho_zp", "rho_xp")]


STATE_FIELDS = ("vz", "vx", "uz", "ux", "szz", "szx", "sxx", "We",
                "dz_z", "dx_z", "dz_x", "dx_x",
                "mem_vz_z", "mem_vx_z", "mem_szz_z", "mem_szx_z",
                "mem_vz_x", "mem_vx_x", "mem_szx_x", "mem_sxx_x")


class OrState(C.Structure):
    _fields_ = [(n, _dp) for n in STATE_FIELDS]


class OrSources(C.Structure):
    _fields_ = [("nsrc", C.c_int32), ("stf_type", C.c_int32), ("stf_z", _dp), ("stf_x", _dp),
                ("z_src", _ip), ("x_src", _ip), ("shot_to_fire", _ip)]


class OrReceivers(C.Structure):
    _fields_ = [("nrec", C.c_int32), ("rtf_type", C.c_int32), ("rtf_uz", _dp), ("rtf_ux", _dp),
                ("z_rec", _ip), ("x_rec", _ip)]


class OrSnaps(C.Structure):
    _fields_ = [(n, _dp) for n in ("vz", "vx", "szz", "szx", "sxx")]


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _i(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


_libs: dict = {}


def load(which: str = "oracle"):
    """Load ``"oracle"`` (the restatement) or ``"ref"`` (the compiled reference)."""
    if which in _libs:
        return _libs[which]
    path = ORACLE_SO if which == "oracle" else REF_SO
    if not path.exists():
        raise FileNotFoundError(f"{path} is not built (make -C oracle{' ref' if which == 'ref' else ''})")
    lib = C.CDLL(str(path))
    p = "or_" if which == "oracle" else "ref_"
    if which == "oracle":
        lib.or_checkfd_dtstab.restype = C.c_double
        lib.or_calc_opt_step.restype = C.c_double
        lib.or_energy_weights2.restype = C.c_double
    lib.__dict__["prefix"] = p
    getattr(lib, p + "adjsrc2").restype = C.c_double
    if which == "ref":
        lib.ref_step_length_psv.restype = C.c_double
    _libs[which] = lib
    return lib


def ref_available() -> bool:
    return REF_SO.exists()


# --------------------------------------------------------------------------- #
# Case construction                                                            #
# --------------------------------------------------------------------------- #

def holberg(fdorder: int, maxrelerror: int = 1, which: str = "oracle") -> np.ndarray:
    lib = load(which)
    hc = np.zeros(7)
    getattr(lib, lib.prefix + "holbergcoeff")(C.c_int32(fdorder), C.c_int32(maxrelerror), _d(hc))
    return hc


def cpml(npower, damp_v, rcoef, k_max, freq, npml1, npml2, fpad, nh, dt, dh, which="oracle"):
    lib = load(which)
    arrs = [np.zeros(nh) for _ in range(6)]
    getattr(lib, lib.prefix + "cpml")(*[_d(a) for a in arrs], C.c_double(npower), C.c_double(damp_v),
                                      C.c_double(rcoef), C.c_double(k_max), C.c_double(freq),
                                      C.c_int32(npml1), C.c_int32(npml2), C.c_int32(fpad), C.c_int32(nh),
                                      C.c_double(dt), C.c_double(dh))
    return dict(zip(("a", "b", "K", "a_half", "b_half", "K_half"), arrs))


def wavelet(nt, dt, amp, fc, ts=0.0, which="oracle"):
    lib = load(which)
    s = np.zeros(nt)
    getattr(lib, lib.prefix + "wavelet")(_d(s), C.c_int32(nt), C.c_double(dt), C.c_double(amp),
                                         C.c_double(fc), C.c_double(ts), C.c_int32(1))
    return s


def dtstab(lam, mu, rho, hc, dh, fw):
    lib = load("oracle")
    nz, nx = lam.shape
    return lib.or_checkfd_dtstab(C.c_int32(nz), C.c_int32(nx), C.c_double(dh), C.c_int32(fw),
                                 _d(rho), _d(lam), _d(mu), _d(hc), None, None)


def make_case(*, nt, nz, nx, dt, dz, dx, fdorder, lam, mu, rho, z_src, x_src, shot_to_fire,
              z_rec, x_rec, stf_z=None, stf_x=None, stf_type=1, fc=100.0, amp_z=0.0, amp_x=1.0,
              npml=(10, 10, 10, 10), isurf=(0, 0, 0, 0), npower=2.0, damp_v=None, rcoef=0.001,
              k_max=1.0, snap=None, taper=None, maxrelerror=1, nshot=None):
    """Assemble a case the way simulate_PSV does (src/h_simulate_PSV.cpp:101-185).

    ``snap`` = (z1, z2, x1, x2, dt, dz, dx) and ``taper`` = (t1, t2, b1, b2, l1, l2, r1, r2)
    default to the pre_proc_fwi.py conventions (scripts/pre_proc_fwi.py:61-80).
    """
    fpad = fdorder // 2
    hc = holberg(fdorder, maxrelerror)
    if damp_v is None:
        damp_v = float(np.sqrt((lam + 2 * mu) / rho).max())
    pz = cpml(npower, damp_v, rcoef, k_max, fc, npml[0], npml[1], fpad, nz, dt, dz)
    px = cpml(npower, damp_v, rcoef, k_max, fc, npml[2], npml[3], fpad, nx, dt, dx)
    nsrc = len(z_src)
    if stf_z is None:
        stf_z = np.stack([wavelet(nt, dt, amp_z, fc) for _ in range(nsrc)])
        stf_x = np.stack([wavelet(nt, dt, amp_x, fc) for _ in range(nsrc)])
    if snap is None:
        snap = (npml[0], nz - npml[1], npml[2], nx - npml[3], 1, 1, 1)
    if taper is None:
        taper = (snap[0], snap[0], snap[1], snap[1], snap[2], snap[2], snap[3], snap[3])
    c = dict(nt=nt, nz=nz, nx=nx, dt=float(dt), dz=float(dz), dx=float(dx), fdorder=fdorder, hc=hc,
             surf=int(any(s > 0 for s in isurf)), isurf=np.asarray(isurf, np.int32),
             pml_z=int(npml[0] > 0 or npml[1] > 0), pml_x=int(npml[2] > 0 or npml[3] > 0),
             npml=np.asarray(npml, np.int32), fc=float(fc),
             snap=np.asarray(snap, np.int32), taper=np.asarray(taper, np.int32),
             lam=np.ascontiguousarray(lam, np.float64), mu=np.ascontiguousarray(mu, np.float64),
             rho=np.ascontiguousarray(rho, np.float64),
             stf_type=int(stf_type), stf_z=np.ascontiguousarray(stf_z, np.float64),
             stf_x=np.ascontiguousarray(stf_x, np.float64),
             z_src=np.asarray(z_src, np.int32), x_src=np.asarray(x_src, np.int32),
             shot_to_fire=np.asarray(shot_to_fire, np.int32),
             nshot=int(nshot if nshot is not None else max(shot_to_fire) + 1),
             z_rec=np.asarray(z_rec, np.int32), x_rec=np.asarray(x_rec, np.int32))
    for k, v in pz.items():
        c[k + "_z"] = v
    for k, v in px.items():
        c[k + "_x"] = v
    return c


def snap_dims(c):
    z1, z2, x1, x2, sdt, sdz, sdx = (int(v) for v in c["snap"])
    return 1 + (int(c["nt"]) - 1) // sdt, 1 + (z2 - z1) // sdz, 1 + (x2 - x1) // sdx


# --------------------------------------------------------------------------- #
# Runners (identical for the restatement and the reference)                    #
# --------------------------------------------------------------------------- #

class _Keep:
    """Holds numpy arrays alive while a ctypes struct points into them."""

    def __init__(self):
        self.refs = []

    def d(self, a):
        self.refs.append(a)
        return _d(a)

    def i(self, a):
        self.refs.append(a)
        return _i(a)


def _grid(c, keep):
    g = OrGrid()
    g.nt, g.nz, g.nx = int(c["nt"]), int(c["nz"]), int(c["nx"])
    g.dt, g.dz, g.dx = float(c["dt"]), float(c["dz"]), float(c["dx"])
    g.fdorder = int(c["fdorder"])
    g.hc[:] = [float(v) for v in c["hc"]]
    g.surf = int(c["surf"])
    g.isurf[:] = [int(v) for v in c["isurf"]]
    g.pml_z, g.pml_x = int(c["pml_z"]), int(c["pml_x"])
    for ax in ("z", "x"):
        for n in ("a", "b", "K", "a_half", "b_half", "K_half"):
            setattr(g, f"{n}_{ax}", keep.d(np.ascontiguousarray(c[f"{n}_{ax}"], np.float64)))
    (g.snap_z1, g.snap_z2, g.snap_x1, g.snap_x2, g.snap_dt, g.snap_dz, g.snap_dx) = (int(v) for v in c["snap"])
    return g


def mat_av2(c, lam=None, mu=None, rho=None, which="oracle"):
    lib = load(which)
    lam = c["lam"] if lam is None else lam
    mu = c["mu"] if mu is None else mu
    rho = c["rho"] if rho is None else rho
    nz, nx = lam.shape
    out = {k: np.zeros((nz, nx)) for k in ("mu_zx", "rho_zp", "rho_xp")}
    Cm = np.zeros(3)
    args = [lam, mu, rho] if which == "oracle" else [lam.copy(), mu.copy(), rho.copy()]
    getattr(lib, lib.prefix + "mat_av2")(*[_d(a) for a in args], _d(out["mu_zx"]), _d(out["rho_zp"]),
                                         _d(out["rho_xp"]), _d(Cm), C.c_int32(nz), C.c_int32(nx))
    out.update(lam=lam, mu=mu, rho=rho, C=Cm)
    return out


def new_state(c):
    nz, nx = int(c["nz"]), int(c["nx"])
    return {k: np.zeros((nz, nx)) for k in STATE_FIELDS}


def kernel_psv(c, media, ishot, *, grad=False, accu=True, state=None, snaps=None, adj=None,
               which="oracle", record=True):
    """kernel_PSV for one shot (src/cpu/n_kernel_PSV.cpp:20-213).

    Forward (``grad=False``): returns seismograms ``uz, ux`` [nrec][nt], the
    final ``state`` and (``accu``) the snapshot cubes.  Adjoint (``grad=True``):
    ``adj`` = (r_z, r_x) residuals injected at the receivers as type-0 sources,
    as simulate_fwi_PSV does (src/cpu/n_simulate_PSV.cpp:336-356); returns the
    shot gradients [snap_nz][snap_nx] and the state (We holds We_adj).
    """
    lib = load(which)
    keep = _Keep()
    g = _grid(c, keep)
    nt = int(c["nt"])
    m = OrMedia(keep.d(media["lam"]), keep.d(media["mu"]), keep.d(media["mu_zx"]),
                keep.d(media["rho_zp"]), keep.d(media["rho_xp"]))
    st = new_state(c) if state is None else state
    s = OrState(*[keep.d(st[k]) for k in STATE_FIELDS])
    nrec = len(c["z_rec"])
    rec_uz = np.zeros((nrec, nt))
    rec_ux = np.zeros((nrec, nt))
    rec = OrReceivers(nrec, 0, keep.d(rec_uz), keep.d(rec_ux), keep.i(c["z_rec"]), keep.i(c["x_rec"]))
    if grad:
        rz, rx = (np.ascontiguousarray(a, np.float64) for a in adj)
        fire = np.full(nrec, ishot, np.int32)
        src = OrSources(nrec, 0, keep.d(rz), keep.d(rx), keep.i(c["z_rec"]), keep.i(c["x_rec"]), keep.i(fire))
    else:
        src = OrSources(len(c["z_src"]), int(c["stf_type"]), keep.d(c["stf_z"]), keep.d(c["stf_x"]),
                        keep.i(c["z_src"]), keep.i(c["x_src"]), keep.i(c["shot_to_fire"]))
    snt, snz, snx = snap_dims(c)
    if snaps is None and (accu or grad):
        snaps = {k: np.zeros((snt, snz, snx)) for k in ("vz", "vx", "szz", "szx", "sxx")}
    sn = OrSnaps(*[keep.d(snaps[k]) for k in ("vz", "vx", "szz", "szx", "sxx")]) if snaps is not None else None
    gl, gm, gr = (np.zeros((snz, snx)) for _ in range(3))
    fn = getattr(lib, lib.prefix + "kernel_psv")
    fn(C.byref(g), C.byref(m), C.byref(s), C.c_int32(ishot), C.c_int32(int(grad)), _d(gl), _d(gm), _d(gr),
       C.byref(src), C.byref(rec), C.c_int32(int(accu)), C.byref(sn) if sn is not None else None)
    out = dict(state=st, snaps=snaps)
    if grad:
        out.update(g_lam=gl, g_mu=gm, g_rho=gr)
    else:
        out.update(uz=rec_uz, ux=rec_ux)
    return out


def adjoint_exact(c, media, ishot, cot_uz, cot_ux):
    """The exact discrete adjoint of kernel_psv's forward for one shot: the gradient of a
    loss L with respect to the media (lam, mu, mu_zx, rho_zp, rho_xp) given the cotangents
    dL/d(seismogram) of the receivers' displacement seismograms [nrec][nt] (reverse mode
    through every operation of the time step, FP64, the forward states of every step kept)."""
    lib = load("oracle")
    keep = _Keep()
    g = _grid(c, keep)
    m = OrMedia(keep.d(media["lam"]), keep.d(media["mu"]), keep.d(media["mu_zx"]),
                keep.d(media["rho_zp"]), keep.d(media["rho_xp"]))
    src = OrSources(len(c["z_src"]), int(c["stf_type"]), keep.d(c["stf_z"]), keep.d(c["stf_x"]),
                    keep.i(c["z_src"]), keep.i(c["x_src"]), keep.i(c["shot_to_fire"]))
    nz, nx = int(c["nz"]), int(c["nx"])
    out = [np.zeros((nz, nx)) for _ in range(5)]
    zr, xr = np.ascontiguousarray(c["z_rec"], np.int32), np.ascontiguousarray(c["x_rec"], np.int32)
    cz = np.ascontiguousarray(cot_uz, np.float64)
    cx = np.ascontiguousarray(cot_ux, np.float64)
    lib.or_adjoint_exact(C.byref(g), C.byref(m), C.c_int32(ishot), C.byref(src), C.c_int32(len(zr)), _i(zr), _i(xr),
                         _d(cz), _d(cx), *[_d(a) for a in out])
    return dict(zip(("lam", "mu", "mu_zx", "rho_zp", "rho_xp"), out))


def media_chain(grads, media):
    """The media gradient carried to (lam, mu, rho) through mat_av2's averages."""
    lib = load("oracle")
    nz, nx = media["mu"].shape
    out = [np.zeros((nz, nx)) for _ in range(3)]
    g = [np.ascontiguousarray(grads[k], np.float64) for k in ("lam", "mu", "mu_zx", "rho_zp", "rho_xp")]
    lib.or_media_chain(*[_d(a) for a in g], _d(media["mu"]), _d(media["mu_zx"]), _d(media["rho_zp"]),
                       _d(media["rho_xp"]), C.c_int32(nz), C.c_int32(nx), *[_d(a) for a in out])
    return tuple(out)


def pcg_psv(dir_, old, grad, which="oracle"):
    """PCG_PSV (src/cpu/n_PCG_PSV.cpp:13-44) on one parameter, in place on all three arrays."""
    lib = load(which)
    nz, nx = grad.shape
    fn = lib.or_pcg_psv if which == "oracle" else lib.ref_pcg_psv
    fn(_d(dir_), _d(old), _d(grad), C.c_int32(nz), C.c_int32(nx))
    return dir_, old, grad


def gauss_filter_kernel(hfs, which="oracle"):
    """gauss_filter_kernel (src/cpu/n_kernel_lib_PSV.cpp:1012-1049): [2 hfs + 1]^2."""
    lib = load(which)
    k = np.zeros((2 * hfs + 1, 2 * hfs + 1))
    (lib.or_gauss_filter_kernel if which == "oracle" else lib.ref_gauss_filter_kernel)(_d(k), C.c_int32(hfs))
    return k


def apply_gauss_filter(mat, hfs, which="oracle"):
    """apply_gauss_filter (:1051-1084) in place; which="ref" runs the reference as written."""
    lib = load(which)
    nz, nx = mat.shape
    filt = gauss_filter_kernel(hfs, which=which)
    if which == "oracle":
        lib.or_apply_gauss_filter(_d(mat), _d(filt), C.c_int32(nz), C.c_int32(nx), C.c_int32(hfs))
    else:
        tmp = np.zeros_like(mat)
        lib.ref_apply_gauss_filter(_d(mat), _d(tmp), _d(filt), C.c_int32(nz), C.c_int32(nx), C.c_int32(hfs))
    return mat


def adjsrc2(mod_uz, mod_ux, true_uz, true_ux, dt, which="oracle"):
    """adjsrc2: residuals (returned) and the L2 misfit."""
    lib = load(which)
    nrec, nt = mod_uz.shape
    ruz, rux = mod_uz.copy(), mod_ux.copy()
    tuz, tux = np.ascontiguousarray(true_uz, np.float64), np.ascontiguousarray(true_ux, np.float64)
    if which == "oracle":
        L2 = lib.or_adjsrc2(_d(ruz), _d(rux), _d(ruz), _d(rux), _d(tuz), _d(tux), C.c_double(dt),
                            C.c_int32(nrec), C.c_int32(nt))
    else:
        L2 = lib.ref_adjsrc2(_d(ruz), _d(rux), _d(tuz.copy()), _d(tux.copy()), C.c_double(dt),
                             C.c_int32(nrec), C.c_int32(nt))
    return ruz, rux, L2


def energy_weights2(c, We, We_adj, which="oracle"):
    lib = load(which)
    We = We.copy()
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    nz, nx = We.shape
    if which == "oracle":
        lib.or_energy_weights2(_d(We), _d(We_adj), C.c_int32(nx), z1, z2, x1, x2)
    else:
        lib.ref_energy_weights2(_d(We), _d(We_adj.copy()), nz, nx, z1, z2, x1, x2)
    return We


def interpol_grad2(c, grad_shot, which="oracle"):
    lib = load(which)
    nz, nx = int(c["nz"]), int(c["nx"])
    z1, z2, x1, x2, _, sdz, sdx = (int(v) for v in c["snap"])
    out = np.zeros((nz, nx))
    gs = np.ascontiguousarray(grad_shot, np.float64)
    if which == "oracle":
        lib.or_interpol_grad2(_d(out), _d(gs), nx, z1, z2, x1, x2, sdz, sdx)
    else:
        lib.ref_interpol_grad2(_d(out), _d(gs.copy()), nz, nx, z1, z2, x1, x2, sdz, sdx)
    return out


def scale_grad_E2(c, grad, grad_full_shot, mat_av, We, which="oracle"):
    lib = load(which)
    nz, nx = grad.shape
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    if which == "oracle":
        lib.or_scale_grad_E2(_d(grad), _d(grad_full_shot), C.c_double(mat_av), _d(We), nx, z1, z2, x1, x2)
    else:
        lib.ref_scale_grad_E2(_d(grad), _d(grad_full_shot.copy()), C.c_double(mat_av), _d(We.copy()),
                              nz, nx, z1, z2, x1, x2)
    return grad


def taper2(c, A, which="oracle"):
    lib = load(which)
    nz, nx = A.shape
    z1, z2, x1, x2 = (int(v) for v in c["snap"][:4])
    t = np.ascontiguousarray(c["taper"], np.int32)
    if which == "oracle":
        lib.or_taper2(_d(A), nz, nx, z1, z2, x1, x2, _i(t))
    else:
        lib.ref_taper2(_d(A), nz, nx, z1, z2, x1, x2, _i(t))
    return A


def update_mat2(mat_old, grad, mat_max, mat_min, step, which="oracle"):
    lib = load(which)
    nz, nx = mat_old.shape
    out = np.zeros_like(mat_old)
    if which == "oracle":
        f = C.c_double(0)
        lib.or_update_mat2(_d(out), _d(mat_old), _d(grad), C.c_double(mat_max), C.c_double(mat_min),
                           C.c_double(step), nz, nx, C.byref(f))
    else:
        lib.ref_update_mat2(_d(out), _d(mat_old.copy()), _d(grad.copy()), C.c_double(mat_max),
                            C.c_double(mat_min), C.c_double(step), nz, nx)
    return out


def calc_opt_step(L2, sl):
    lib = load("oracle")
    L2 = np.ascontiguousarray(L2, np.float64)
    sl = np.ascontiguousarray(sl, np.float64)
    return lib.or_calc_opt_step(_d(L2), _d(sl))


# --------------------------------------------------------------------------- #
# Orchestration restated (simulate_fwi_PSV / step_length_PSV)                  #
# --------------------------------------------------------------------------- #

LAM_MAX, MU_MAX, RHO_MAX = 4.8e10, 2.7e10, 3000.0


def shot_gradient(c, media, ishot, obs_uz, obs_ux, which="oracle"):
    """One shot of the simulate_fwi_PSV loop body (src/cpu/n_simulate_PSV.cpp:270-418)
    up to the shot's preconditioned gradient, returned as full-grid contributions."""
    fwd = kernel_psv(c, media, ishot, grad=False, accu=True, which=which)
    We_fwd = fwd["state"]["We"].copy()
    ruz, rux, L2 = adjsrc2(fwd["uz"], fwd["ux"], obs_uz, obs_ux, float(c["dt"]), which=which)
    adj = kernel_psv(c, media, ishot, grad=True, accu=False, snaps=fwd["snaps"], adj=(ruz, rux), which=which)
    We_adj = adj["state"]["We"].copy()
    We = energy_weights2(c, We_fwd, We_adj, which=which)
    nz, nx = int(c["nz"]), int(c["nx"])
    contrib = {}
    for key, ci in (("g_lam", 0), ("g_mu", 1), ("g_rho", 2)):
        full = interpol_grad2(c, adj[key], which=which)
        acc = np.zeros((nz, nx))
        scale_grad_E2(c, acc, full, float(media["C"][ci]), We, which=which)
        contrib[key] = acc
    return dict(L2=L2, seis=(fwd["uz"], fwd["ux"]), raw=(adj["g_lam"], adj["g_mu"], adj["g_rho"]),
                We_fwd=We_fwd, We_adj=We_adj, We=We, **contrib)


def misfit_all(c, lam, mu, rho, obs, which="oracle"):
    media = mat_av2(c, lam, mu, rho, which=which)
    L2 = 0.0
    for ishot in range(int(c["nshot"])):
        f = kernel_psv(c, media, ishot, grad=False, accu=False, which=which)
        _, _, l = adjsrc2(f["uz"], f["ux"], obs[ishot][0], obs[ishot][1], float(c["dt"]), which=which)
        L2 += l
    return L2


VEL_BOUNDS = (1000.0, 6000.0, 500.0, 3500.0)   # cp_min, cp_max, cs_min, cs_max (m/s)


def velocity_of(lam, mu, rho):
    """cp = sqrt((lam + 2 mu) / rho), cs = sqrt(mu / rho)."""
    return np.sqrt((lam + 2.0 * mu) / rho), np.sqrt(mu / rho)


def velocity_gradient(G_lam, G_mu, rho, cp, cs):
    """The lam / mu gradient carried to cp / cs at fixed rho by the chain rule through
    lam = rho (cp^2 - 2 cs^2), mu = rho cs^2: dJ/dcp = 2 rho cp dJ/dlam,
    dJ/dcs = 2 rho cs (dJ/dmu - 2 dJ/dlam)."""
    return 2.0 * rho * cp * G_lam, 2.0 * rho * cs * (G_mu - 2.0 * G_lam)


def update_model(old, G, step, *, param="lame", rho_factor=0.5, rho_min=1.25, bounds=VEL_BOUNDS):
    """The model update of simulate_fwi_PSV (src/cpu/n_simulate_PSV.cpp:542-546):
    update_mat2 on lam, mu (step) and rho (rho_factor * step, floor rho_min); or, for
    param "velocity", update_mat2 on cp and cs with the chain-ruled gradient and the
    bounds (cp_min, cp_max, cs_min, cs_max), rho kept, lam and mu from cp, cs, rho."""
    lam_old, mu_old, rho_old = old
    if param == "lame":
        return (update_mat2(lam_old, G[0], LAM_MAX, 0.0, step), update_mat2(mu_old, G[1], MU_MAX, 0.0, step),
                update_mat2(rho_old, G[2], RHO_MAX, rho_min, rho_factor * step))
    cp_old, cs_old = velocity_of(lam_old, mu_old, rho_old)
    g_cp, g_cs = velocity_gradient(G[0], G[1], rho_old, cp_old, cs_old)
    cp = update_mat2(cp_old, g_cp, bounds[1], bounds[0], step)
    cs = update_mat2(cs_old, g_cs, bounds[3], bounds[2], step)
    mu = rho_old * cs * cs
    return rho_old * cp * cp - 2.0 * mu, mu, rho_old.copy()


def step_length(c, est, L2_0, lam_old, mu_old, rho_old, G, obs, which="oracle", log=None, param="lame",
                bounds=VEL_BOUNDS):
    """step_length_PSV with update_param 0 (src/cpu/n_step_length_PSV.cpp:13-209):
    bracketing with scalefac 2, at most 30 halvings and 4 enlargements, then the
    parabola vertex.  Trial models as update_model (rho at 0.5 step, floor 1.5).
    Returns (step, trials) or raises RuntimeError where the reference calls exit(0)."""
    scalefac, stepmax_1, stepmax_2 = 2.0, 30, 4
    L2t = [L2_0, 0.0, 0.0]
    sl = [0.0, 0.0, 0.0]
    step1 = step2 = 0
    itests = iteste = 1
    countstep = 0
    trials = []
    while step2 != 1 or step1 != 1:
        for itest in range(itests, iteste + 1):
            lam, mu, rho = update_model((lam_old, mu_old, rho_old), G, est, param=param, rho_factor=0.5,
                                        rho_min=1.5, bounds=bounds)
            L2tmp = misfit_all(c, lam, mu, rho, obs, which=which)
            L2t[itest] = L2tmp
            sl[itest] = est
            sl[0] = 0.0
            trials.append((est, L2tmp))
        if step1 == 0 and L2t[0] <= L2t[1]:
            est = est / scalefac
            countstep += 1
        if step1 == 1 and L2t[1] < L2t[2]:
            sl[2] = est
            step2 = 1
        if step1 == 1 and L2t[1] >= L2t[2]:
            sl[2] = est
            est = est + est / scalefac
            countstep += 1
        if step1 == 0 and L2t[0] > L2t[1]:
            sl[1] = est
            step1 = 1
            iteste = itests = 2
            countstep = 0
            est = est + est / scalefac
        if step1 == 0 and countstep > stepmax_1:
            raise RuntimeError("step length estimation failed")
        if step1 == 1 and countstep > stepmax_2:
            step1 = step2 = 1
    if step1 == 1:
        est = calc_opt_step(L2t, sl)
    return est, trials, (list(L2t), list(sl))


def fwi_iteration(c, model, G, obs, est, *, param="lame", grad_sign=-1, bounds=VEL_BOUNDS, which="oracle"):
    """One iteration of simulate_fwi_PSV (src/cpu/n_simulate_PSV.cpp:249-582) composed from
    the pieces above: every shot's preconditioned gradient added into the global gradient
    G (modified in place and never reset, as in the reference), taper2 on the accumulated
    G, the step-length search from est, and the update at the chosen step (rho at half the
    step, floor 1.25).  grad_sign -1 is the CPU path's "grad -= ..." (its shot gradients
    as shot_gradient returns them), +1 the GPU twin's "+=" (the same sums negated).
    Returns (L2, step, new model, trials)."""
    lam_old, mu_old, rho_old = (np.asarray(m, np.float64).copy() for m in model)
    media = mat_av2(c, lam_old, mu_old, rho_old, which=which)
    L2 = 0.0
    for s in range(int(c["nshot"])):
        sg = shot_gradient(c, media, s, *obs[s], which=which)
        L2 += sg["L2"]
        for i, k in enumerate(("g_lam", "g_mu", "g_rho")):
            G[i] += -grad_sign * sg[k]
    for i in range(3):
        G[i][...] = taper2(c, G[i], which=which)
    step, trials, _ = step_length(c, est, L2, lam_old, mu_old, rho_old, G, obs, which=which, param=param,
                                  bounds=bounds)
    new = update_model((lam_old, mu_old, rho_old), G, step, param=param, rho_factor=0.5, rho_min=1.25,
                       bounds=bounds)
    return L2, step, new, trials


def ref_step_length(c, est, L2_0, lam_old, mu_old, rho_old, G, obs):
    """The compiled reference's step_length_PSV (update_param 0) on the same inputs."""
    lib = load("ref")
    keep = _Keep()
    g = _grid(c, keep)
    nz, nx, nt = int(c["nz"]), int(c["nx"]), int(c["nt"])
    nshot = int(c["nshot"])
    lam, mu, rho = lam_old.copy(), mu_old.copy(), rho_old.copy()
    lc, mc, rc = lam_old.copy(), mu_old.copy(), rho_old.copy()
    mzx, rz, rx = (np.zeros((nz, nx)) for _ in range(3))
    gl, gm, gr = (np.ascontiguousarray(a, np.float64).copy() for a in G)
    st = new_state(c)
    s = OrState(*[keep.d(st[k]) for k in STATE_FIELDS])
    nrec = len(c["z_rec"])
    rec_uz, rec_ux = np.zeros((nrec, nt)), np.zeros((nrec, nt))
    rec = OrReceivers(nrec, 0, keep.d(rec_uz), keep.d(rec_ux), keep.i(c["z_rec"]), keep.i(c["x_rec"]))
    src = OrSources(len(c["z_src"]), int(c["stf_type"]), keep.d(c["stf_z"]), keep.d(c["stf_x"]),
                    keep.i(c["z_src"]), keep.i(c["x_src"]), keep.i(c["shot_to_fire"]))
    tuz = np.ascontiguousarray(np.stack([o[0] for o in obs]), np.float64)
    tux = np.ascontiguousarray(np.stack([o[1] for o in obs]), np.float64)
    snt, snz, snx = snap_dims(c)
    snaps = {k: np.zeros((snt, snz, snx)) for k in ("vz", "vx", "szz", "szx", "sxx")}
    sn = OrSnaps(*[keep.d(snaps[k]) for k in ("vz", "vx", "szz", "szx", "sxx")])
    return lib.ref_step_length_psv(C.c_double(est), C.c_double(L2_0), nshot, C.byref(g), _d(lam), _d(mu), _d(rho),
                                   _d(lc), _d(mc), _d(rc), _d(mzx), _d(rz), _d(rx), _d(gl), _d(gm), _d(gr),
                                   C.byref(s), C.byref(src), C.byref(rec), _d(tuz), _d(tux), C.byref(sn))
