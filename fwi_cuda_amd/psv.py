"""Host-side mirror of the reference's P/SV interface over the C ABI.

Names follow the reference (src/h_simulate_PSV.cpp, src/cpu/n_simulate_PSV.cpp,
src/cpu/n_step_length_PSV.cpp, src/h_preproc_PSV.cpp).  A problem is a plain
dict ("case") with the keys the reference's input files carry:

    nt nz nx dt dz dx fdorder hc isurf pml_z pml_x
    a_z b_z K_z a_half_z b_half_z K_half_z  a_x ... K_half_x
    snap = (z1, z2, x1, x2, dt, dz, dx)   taper = (t1, t2, b1, b2, l1, l2, r1, r2)
    stf_type stf_z stf_x z_src x_src shot_to_fire nshot z_rec x_rec
    lam mu rho

Receivers may also be given per shot (``rec_off`` CSR plus concatenated
``z_rec``/``x_rec``), which the reference cannot express.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FwiAcq, FwiGrid, FwiKernelTime, FwiOpts, check, dptr, iptr

# ---------------------------------------------------------------------------
# preprocessing (src/h_preproc_PSV.cpp, src/h_holbergcoeff.cpp, src/h_checkfd_ssg_elastic.cpp)
# ---------------------------------------------------------------------------


def holbergcoeff(fdorder: int, maxrelerror: int = 1) -> np.ndarray:
    hc = np.zeros(7)
    check(_lib.load().fwi_holbergcoeff(fdorder, maxrelerror, dptr(hc)))
    return hc


def cpml_PSV(npower, damp_v_pml, rcoef, k_max_pml, freq, npml_h1, npml_h2, fpad, nh, dt, dh) -> dict:
    arrs = [np.zeros(nh) for _ in range(6)]
    check(_lib.load().fwi_cpml(*[dptr(a) for a in arrs], float(npower), float(damp_v_pml), float(rcoef),
                               float(k_max_pml), float(freq), int(npml_h1), int(npml_h2), int(fpad), int(nh),
                               float(dt), float(dh)))
    return dict(zip(("a", "b", "K", "a_half", "b_half", "K_half"), arrs))


def wavelet(nt, dt, amp, fc, ts=0.0, shape=1) -> np.ndarray:
    s = np.zeros(nt)
    check(_lib.load().fwi_wavelet(dptr(s), int(nt), float(dt), float(amp), float(fc), float(ts), int(shape)))
    return s


def checkfd_dtstab(lam, mu, rho, hc, dh, fw=0) -> float:
    lam, mu, rho = (np.ascontiguousarray(a, np.float64) for a in (lam, mu, rho))
    nz, nx = lam.shape
    return _lib.load().fwi_dtstab(nz, nx, float(dh), int(fw), dptr(rho), dptr(lam), dptr(mu),
                                  dptr(np.ascontiguousarray(hc, np.float64)), None, None)


def make_case(*, nt, nz, nx, dt, dz, dx, fdorder, lam, mu, rho, z_src, x_src, shot_to_fire=None,
              z_rec, x_rec, rec_off=None, stf_z=None, stf_x=None, stf_type=1, fc=100.0, amp_z=0.0, amp_x=1.0,
              npml=(10, 10, 10, 10), isurf=(0, 0, 0, 0), npower=2.0, damp_v=None, rcoef=0.001, k_max=1.0,
              snap=None, taper=None, maxrelerror=1, nshot=None) -> dict:
    """Assemble a case the way simulate_PSV does (src/h_simulate_PSV.cpp:101-185),
    with this library's own preprocessing.  Defaults: the reference's Holberg
    E = 0.1 % coefficients (MAXRELERROR = 1), a wavelet per source with the given
    component amplitudes (the reference hard-wires amp_z = 0, amp_x = 1)."""
    fpad = fdorder // 2
    hc = holbergcoeff(fdorder, maxrelerror)
    lam, mu, rho = (np.ascontiguousarray(a, np.float64) for a in (lam, mu, rho))
    if damp_v is None:
        damp_v = float(np.sqrt((lam + 2 * mu) / rho).max())
    pz = cpml_PSV(npower, damp_v, rcoef, k_max, fc, npml[0], npml[1], fpad, nz, dt, dz)
    px = cpml_PSV(npower, damp_v, rcoef, k_max, fc, npml[2], npml[3], fpad, nx, dt, dx)
    nsrc = len(z_src)
    if shot_to_fire is None:
        shot_to_fire = np.arange(nsrc)
    if stf_z is None:
        w = wavelet(nt, dt, 1.0, fc)
        stf_z = np.tile(amp_z * w, (nsrc, 1))
        stf_x = np.tile(amp_x * w, (nsrc, 1))
    if snap is None:
        snap = (npml[0], nz - npml[1], npml[2], nx - npml[3], 1, 1, 1)
    if taper is None:
        taper = (snap[0], snap[0], snap[1], snap[1], snap[2], snap[2], snap[3], snap[3])
    c = dict(nt=int(nt), nz=int(nz), nx=int(nx), dt=float(dt), dz=float(dz), dx=float(dx), fdorder=int(fdorder),
             hc=hc, surf=int(any(s > 0 for s in isurf)), isurf=np.asarray(isurf, np.int32),
             pml_z=int(npml[0] > 0 or npml[1] > 0), pml_x=int(npml[2] > 0 or npml[3] > 0),
             npml=np.asarray(npml, np.int32), fc=float(fc),
             cpml_par=(float(npower), float(damp_v), float(rcoef), float(k_max)),
             snap=np.asarray(snap, np.int32), taper=np.asarray(taper, np.int32),
             lam=lam, mu=mu, rho=rho, stf_type=int(stf_type),
             stf_z=np.ascontiguousarray(stf_z, np.float64), stf_x=np.ascontiguousarray(stf_x, np.float64),
             z_src=np.asarray(z_src, np.int32), x_src=np.asarray(x_src, np.int32),
             shot_to_fire=np.asarray(shot_to_fire, np.int32),
             nshot=int(nshot if nshot is not None else int(np.max(shot_to_fire)) + 1),
             z_rec=np.asarray(z_rec, np.int32), x_rec=np.asarray(x_rec, np.int32))
    if rec_off is not None:
        c["rec_off"] = np.asarray(rec_off, np.int32)
    for k, v in pz.items():
        c[k + "_z"] = v
    for k, v in px.items():
        c[k + "_x"] = v
    return c


def lame_to_velocity_gradient(g_lam, g_mu, g_rho, lam, mu, rho):
    """Gradient with respect to (cp, cs, rho) from the gradient with respect to
    (lam, mu, rho), by the chain rule through lam = rho (cp^2 - 2 cs^2), mu = rho cs^2:
    dJ/dcp = 2 rho cp dJ/dlam, dJ/dcs = 2 rho cs (dJ/dmu - 2 dJ/dlam),
    dJ/drho|cp,cs = (cp^2 - 2 cs^2) dJ/dlam + cs^2 dJ/dmu + dJ/drho."""
    lam, mu, rho = (np.asarray(v, np.float64) for v in (lam, mu, rho))
    cp = np.sqrt((lam + 2.0 * mu) / rho)
    cs = np.sqrt(mu / rho)
    g_cp = 2.0 * rho * cp * g_lam
    g_cs = 2.0 * rho * cs * (g_mu - 2.0 * g_lam)
    g_r = (cp * cp - 2.0 * cs * cs) * g_lam + cs * cs * g_mu + g_rho
    return g_cp, g_cs, g_r


def acquisition_csr(case: dict):
    """Sources grouped by shot (stable in source order, src_shot_to_fire of
    src/h_io_PSV.cpp:143-145) and receivers per shot.  Without ``rec_off`` every
    shot records the same receivers, as the reference does."""
    nshot = int(case["nshot"])
    fire = np.asarray(case["shot_to_fire"], np.int64)
    order = np.argsort(fire, kind="stable")
    src_off = np.zeros(nshot + 1, np.int32)
    np.add.at(src_off, fire + 1, 1)
    src_off = np.cumsum(src_off).astype(np.int32)
    src = dict(src_off=src_off, src_z=np.ascontiguousarray(case["z_src"][order], np.int32),
               src_x=np.ascontiguousarray(case["x_src"][order], np.int32),
               stf_z=np.ascontiguousarray(case["stf_z"][order], np.float64),
               stf_x=np.ascontiguousarray(case["stf_x"][order], np.float64))
    if "rec_off" in case:
        rec = dict(rec_off=np.ascontiguousarray(case["rec_off"], np.int32),
                   rec_z=np.ascontiguousarray(case["z_rec"], np.int32),
                   rec_x=np.ascontiguousarray(case["x_rec"], np.int32))
    else:
        nrec = len(case["z_rec"])
        rec = dict(rec_off=(np.arange(nshot + 1) * nrec).astype(np.int32),
                   rec_z=np.ascontiguousarray(np.tile(case["z_rec"], nshot), np.int32),
                   rec_x=np.ascontiguousarray(np.tile(case["x_rec"], nshot), np.int32))
    return src, rec


class Simulation:
    """One device context: grid, acquisition, model, observed data, gradient.

    Mirrors simulate_PSV's state (src/h_simulate_PSV.cpp:17-90) and the
    per-shot entry points of the CPU path.  All heavy work runs in the HIP
    library; this class only moves numpy arrays across the C ABI.
    """

    def __init__(self, case: dict, *, batch=0, storage=_lib.STORE_AUTO, ckpt_interval=0, mem_budget=0,
                 form=_lib.FORM_FUSED, grad_sign=-1, tile_tz=0, device=0, check_finite=1, kernel_variant=0,
                 work_order=0, use_graph=0, param=_lib.PARAM_LAME, vel_bounds=None, ckpt_slots=0, pcg=0,
                 smooth_hfs=0, seis_fp64=0, ckpt_keep=0):
        lib = _lib.load()
        self.lib = lib
        self.case = case
        self._keep = []
        g = FwiGrid()
        g.nt, g.nz, g.nx = int(case["nt"]), int(case["nz"]), int(case["nx"])
        g.dt, g.dz, g.dx = float(case["dt"]), float(case["dz"]), float(case["dx"])
        g.fdorder = int(case["fdorder"])
        g.hc[:] = [float(v) for v in case["hc"]]
        g.isurf[:] = [int(v) for v in case["isurf"]] if int(case.get("surf", 1)) else [0, 0, 0, 0]
        g.pml_z, g.pml_x = int(case["pml_z"]), int(case["pml_x"])
        for ax in ("z", "x"):
            for n in ("a", "b", "K", "a_half", "b_half", "K_half"):
                arr = np.ascontiguousarray(case[f"{n}_{ax}"], np.float64)
                self._keep.append(arr)
                setattr(g, f"{n}_{ax}", dptr(arr))
        (g.snap_z1, g.snap_z2, g.snap_x1, g.snap_x2, g.snap_dt, g.snap_dz, g.snap_dx) = \
            (int(v) for v in case["snap"])
        g.taper[:] = [int(v) for v in case["taper"]]
        src, rec = acquisition_csr(case)
        self.src, self.rec = src, rec
        q = FwiAcq()
        q.nshot = int(case["nshot"])
        q.stf_type = int(case["stf_type"])
        q.src_off, q.src_z, q.src_x = iptr(src["src_off"]), iptr(src["src_z"]), iptr(src["src_x"])
        q.stf_z, q.stf_x = dptr(src["stf_z"]), dptr(src["stf_x"])
        q.rec_off, q.rec_z, q.rec_x = iptr(rec["rec_off"]), iptr(rec["rec_z"]), iptr(rec["rec_x"])
        o = FwiOpts()
        lib.fwi_default_opts(C.byref(o))
        o.batch, o.storage, o.ckpt_interval, o.mem_budget = int(batch), int(storage), int(ckpt_interval), int(mem_budget)
        o.form, o.grad_sign, o.tile_tz, o.device = int(form), int(grad_sign), int(tile_tz), int(device)
        o.check_finite = int(check_finite)
        o.kernel_variant, o.work_order = int(kernel_variant), int(work_order)
        o.use_graph = int(use_graph)
        o.param = int(param)
        o.ckpt_slots = int(ckpt_slots)
        o.pcg, o.smooth_hfs = int(pcg), int(smooth_hfs)
        o.seis_fp64 = int(seis_fp64)
        o.ckpt_keep = int(ckpt_keep)
        if vel_bounds is not None:
            o.vel_bounds[:] = [float(v) for v in vel_bounds]
        h = C.c_void_p()
        check(lib.fwi_create(C.byref(g), C.byref(q), C.byref(o), C.byref(h)), None)
        self.ctx = h
        self.nt, self.nz, self.nx = g.nt, g.nz, g.nx
        self.nshot = q.nshot
        self.rec_off = rec["rec_off"]
        z1, z2, x1, x2, _, sdz, sdx = (int(v) for v in case["snap"])
        self.snap_shape = (1 + (z2 - z1) // sdz, 1 + (x2 - x1) // sdx)
        if "lam" in case:
            self.set_model(case["lam"], case["mu"], case["rho"])

    # -- lifetime --------------------------------------------------------------
    def close(self):
        if getattr(self, "ctx", None):
            self.lib.fwi_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, status):
        check(status, self.ctx)

    def plan(self) -> dict:
        b, s, k = C.c_int32(), C.c_int32(), C.c_int32()
        n = C.c_int64()
        self._chk(self.lib.fwi_plan(self.ctx, C.byref(b), C.byref(s), C.byref(k), C.byref(n)))
        u, sl, rf = C.c_int32(), C.c_int32(), C.c_double()
        self._chk(self.lib.fwi_checkpoint_plan(self.ctx, C.byref(u), C.byref(sl), C.byref(rf)))
        kept = C.c_int32()
        self._chk(self.lib.fwi_checkpoint_kept(self.ctx, C.byref(kept)))
        return dict(batch=b.value, storage={1: "full", 2: "checkpoint"}.get(s.value, s.value),
                    ckpt_interval=k.value, device_bytes=n.value, ckpt_units=u.value, ckpt_slots=sl.value,
                    ckpt_kept=kept.value, recompute=round(rf.value, 4))

    # -- model (mat_av2) --------------------------------------------------------
    def set_model(self, lam, mu, rho):
        lam, mu, rho = (np.ascontiguousarray(a, np.float64) for a in (lam, mu, rho))
        self._chk(self.lib.fwi_set_model(self.ctx, dptr(lam), dptr(mu), dptr(rho)))

    def set_model_device(self, lam_ptr: int, mu_ptr: int, rho_ptr: int):
        """Model from device arrays [nz][nx] FP64 (raw HIP pointers, e.g. tensor.data_ptr())."""
        self._chk(self.lib.fwi_set_model_device(self.ctx, C.c_void_p(lam_ptr), C.c_void_p(mu_ptr),
                                                C.c_void_p(rho_ptr)))

    def get_model(self):
        out = [np.zeros((self.nz, self.nx)) for _ in range(3)]
        self._chk(self.lib.fwi_get_model(self.ctx, *[dptr(a) for a in out]))
        return tuple(out)

    def mat_av2(self):
        out = [np.zeros((self.nz, self.nx)) for _ in range(3)]
        Cm = np.zeros(3)
        self._chk(self.lib.fwi_get_media(self.ctx, *[dptr(a) for a in out], dptr(Cm)))
        return dict(mu_zx=out[0], rho_zp=out[1], rho_xp=out[2], C=Cm)

    def set_observed(self, obs_z, obs_x):
        """Observed seismograms: per-shot [nrec][nt] arrays or the CSR stack."""
        if isinstance(obs_z, (list, tuple)):
            obs_z, obs_x = np.concatenate(obs_z), np.concatenate(obs_x)
        self._obs = (np.ascontiguousarray(obs_z, np.float64), np.ascontiguousarray(obs_x, np.float64))
        self._chk(self.lib.fwi_set_observed(self.ctx, dptr(self._obs[0]), dptr(self._obs[1])))

    # -- propagation --------------------------------------------------------------
    def _rows(self, shot0, nshots):
        return int(self.rec_off[shot0 + nshots] - self.rec_off[shot0])

    def forward(self, shot0=0, nshots=None):
        """kernel_PSV forward for the shots; returns (uz, ux) stacked [rows][nt]."""
        nshots = self.nshot - shot0 if nshots is None else nshots
        n = self._rows(shot0, nshots)
        uz, ux = np.zeros((n, self.nt)), np.zeros((n, self.nt))
        self._chk(self.lib.fwi_forward(self.ctx, shot0, nshots, dptr(uz), dptr(ux)))
        return uz, ux

    def forward_device(self, seis_z_ptr: int, seis_x_ptr: int, shot0=0, nshots=None):
        """forward() into device arrays [rows][nt] FP64 (raw HIP pointers)."""
        nshots = self.nshot - shot0 if nshots is None else nshots
        self._chk(self.lib.fwi_forward_device(self.ctx, shot0, nshots, C.c_void_p(seis_z_ptr),
                                              C.c_void_p(seis_x_ptr)))

    def split(self, stacked, shot0=0):
        """Stacked CSR rows -> list of per-shot [nrec][nt] arrays."""
        out, base = [], int(self.rec_off[shot0])
        nshots = None
        for s in range(shot0, self.nshot):
            a, b = int(self.rec_off[s]) - base, int(self.rec_off[s + 1]) - base
            if b > stacked.shape[0]:
                break
            out.append(stacked[a:b])
        return out

    def wavefield(self, shot, field):
        names = ("vz", "vx", "uz", "ux", "szz", "szx", "sxx", "We", "We_adj")
        idx = names.index(field) if isinstance(field, str) else int(field)
        out = np.zeros((self.nz, self.nx))
        self._chk(self.lib.fwi_get_wavefield(self.ctx, shot, idx, dptr(out)))
        return out

    def gradient(self, shot0=0, nshots=None):
        """Energy-weighted gradient of the shots added to the global gradient
        (simulate_fwi_PSV's shot loop); returns the per-shot misfits."""
        nshots = self.nshot - shot0 if nshots is None else nshots
        m = np.zeros(nshots)
        self._chk(self.lib.fwi_gradient(self.ctx, shot0, nshots, dptr(m)))
        return m

    def gradient_raw(self, adj_z, adj_x, shot0=0, nshots=None):
        """fwi_grad2 sums for caller-given adjoint sources (snap grid)."""
        nshots = self.nshot - shot0 if nshots is None else nshots
        adj_z, adj_x = (np.ascontiguousarray(a, np.float64) for a in (adj_z, adj_x))
        out = [np.zeros(self.snap_shape) for _ in range(3)]
        self._chk(self.lib.fwi_gradient_raw(self.ctx, shot0, nshots, dptr(adj_z), dptr(adj_x),
                                            *[dptr(a) for a in out]))
        return tuple(out)

    def misfit_gradient(self, adj_z, adj_x, shot0=0, nshots=None):
        """dJ/d(lam, mu, rho) [nz][nx] of J = sum 0.5 dt r^2 for adjoint sources r
        (CSR rows like the seismograms): the backward pass of forward()."""
        nshots = self.nshot - shot0 if nshots is None else nshots
        adj_z, adj_x = (np.ascontiguousarray(a, np.float64) for a in (adj_z, adj_x))
        out = [np.zeros((self.nz, self.nx)) for _ in range(3)]
        self._chk(self.lib.fwi_misfit_gradient(self.ctx, shot0, nshots, dptr(adj_z), dptr(adj_x),
                                               *[dptr(a) for a in out]))
        return tuple(out)

    def misfit_gradient_exact(self, adj_z, adj_x, shot0=0, nshots=None):
        """The exact derivative of J = sum 0.5 dt r^2 for adjoint sources r (the discrete
        adjoint of the time step), dJ/d(lam, mu, rho) [nz][nx]."""
        nshots = self.nshot - shot0 if nshots is None else nshots
        adj_z, adj_x = (np.ascontiguousarray(a, np.float64) for a in (adj_z, adj_x))
        out = [np.zeros((self.nz, self.nx)) for _ in range(3)]
        self._chk(self.lib.fwi_misfit_gradient_exact(self.ctx, shot0, nshots, adj_z.ctypes.data, adj_x.ctypes.data,
                                                     *[a.ctypes.data for a in out], 0))
        return tuple(out)

    def misfit_gradient_exact_device(self, adj_z_ptr: int, adj_x_ptr: int, d_lam_ptr: int, d_mu_ptr: int,
                                     d_rho_ptr: int, shot0=0, nshots=None):
        """misfit_gradient_exact() on device arrays (raw HIP pointers)."""
        nshots = self.nshot - shot0 if nshots is None else nshots
        self._chk(self.lib.fwi_misfit_gradient_exact(self.ctx, shot0, nshots, *[C.c_void_p(p) for p in (
            adj_z_ptr, adj_x_ptr, d_lam_ptr, d_mu_ptr, d_rho_ptr)], 1))

    def PCG_PSV(self):
        """The conjugate-gradient direction in place of the global gradient (PCG_PSV)."""
        self._chk(self.lib.fwi_pcg_direction(self.ctx))

    def apply_gauss_filter(self, hfs):
        """Gaussian smoothing of the global gradient (apply_gauss_filter, FP64 accumulator)."""
        self._chk(self.lib.fwi_smooth_gradient(self.ctx, int(hfs)))

    def misfit_gradient_device(self, adj_z_ptr: int, adj_x_ptr: int, d_lam_ptr: int, d_mu_ptr: int, d_rho_ptr: int,
                               shot0=0, nshots=None):
        """misfit_gradient() on device arrays (raw HIP pointers): adjoint sources [rows][nt]
        in, dJ/d(lam, mu, rho) [nz][nx] out, all FP64."""
        nshots = self.nshot - shot0 if nshots is None else nshots
        self._chk(self.lib.fwi_misfit_gradient_device(self.ctx, shot0, nshots, *[C.c_void_p(p) for p in (
            adj_z_ptr, adj_x_ptr, d_lam_ptr, d_mu_ptr, d_rho_ptr)]))

    def get_gradient(self):
        out = [np.zeros((self.nz, self.nx)) for _ in range(3)]
        self._chk(self.lib.fwi_get_gradient(self.ctx, *[dptr(a) for a in out]))
        return tuple(out)

    def set_gradient(self, g_lam, g_mu, g_rho):
        g = [np.ascontiguousarray(a, np.float64) for a in (g_lam, g_mu, g_rho)]
        self._chk(self.lib.fwi_set_gradient(self.ctx, *[dptr(a) for a in g]))

    def zero_gradient(self):
        self._chk(self.lib.fwi_zero_gradient(self.ctx))

    def taper2(self):
        self._chk(self.lib.fwi_taper_gradient(self.ctx))

    def misfit(self, shot0=0, nshots=None) -> float:
        nshots = self.nshot - shot0 if nshots is None else nshots
        L2 = C.c_double()
        self._chk(self.lib.fwi_misfit(self.ctx, shot0, nshots, C.byref(L2)))
        return L2.value

    def copy_mat(self):
        self._chk(self.lib.fwi_save_model(self.ctx))

    def update_mat(self, step, rho_factor=0.5, rho_min=1.25):
        self._chk(self.lib.fwi_update_model(self.ctx, float(step), float(rho_factor), float(rho_min)))

    def step_length_PSV(self, est_step_length, L2_norm_0):
        step, n = C.c_double(), C.c_int32()
        self._chk(self.lib.fwi_step_length(self.ctx, float(est_step_length), float(L2_norm_0), C.byref(step),
                                           C.byref(n)))
        return step.value, n.value

    def fwi_iteration(self):
        L2, step = C.c_double(), C.c_double()
        self._chk(self.lib.fwi_iteration(self.ctx, C.byref(L2), C.byref(step)))
        return L2.value, step.value

    # -- reference-named drivers -------------------------------------------------
    def simulate_fwd_PSV(self):
        """Forward modelling of every shot (src/cpu/n_simulate_PSV.cpp:18-149):
        per-shot displacement seismograms."""
        uz, ux = self.forward(0, self.nshot)
        return list(zip(self.split(uz), self.split(ux)))

    def simulate_fwi_PSV(self, niter, callback=None):
        """niter iterations of simulate_fwi_PSV (src/cpu/n_simulate_PSV.cpp:249-582);
        the global gradient accumulates across iterations as in the reference."""
        hist = []
        for k in range(niter):
            L2, step = self.fwi_iteration()
            hist.append((L2, step))
            if callback:
                callback(k, L2, step)
        return hist

    # -- profiling --------------------------------------------------------------
    def tuning(self):
        tz = C.c_int32()
        buf = C.create_string_buffer(8192)
        self._chk(self.lib.fwi_tuning(self.ctx, C.byref(tz), buf, 8192))
        return tz.value, buf.value.decode()

    def forward_snapshots(self, shot):
        """One shot's forward with its snapshot cubes (write_accu's arrays)."""
        z1, z2, x1, x2, sdt, sdz, sdx = (int(v) for v in self.case["snap"])
        snt = 1 + (self.nt - 1) // sdt
        cubes = [np.zeros((snt,) + self.snap_shape) for _ in range(5)]
        n = self._rows(shot, 1)
        uz, ux = np.zeros((n, self.nt)), np.zeros((n, self.nt))
        self._chk(self.lib.fwi_forward_snapshots(self.ctx, shot, dptr(uz), dptr(ux), *[dptr(a) for a in cubes]))
        return (uz, ux), dict(zip(("vz", "vx", "szz", "szx", "sxx"), cubes))

    def profile(self, on=True, kernels=None, stride=1):
        """Per-launch HIP-event timing of every kernel class, or of the named ones only
        (names of _lib.KERNELS, e.g. ("stress", "velocity", "step")); with stride k only
        every k-th launch of a class carries events."""
        self._chk(self.lib.fwi_profile_stride(self.ctx, int(stride)))
        if kernels is None or not on:
            self._chk(self.lib.fwi_profile(self.ctx, int(on)))
            return
        mask = 0
        for k in kernels:
            mask |= 1 << _lib.KERNELS.index(k)
        self._chk(self.lib.fwi_profile_kernels(self.ctx, mask))

    def kernel_times(self) -> dict:
        """{kernel class: (launches, summed ms of the timed launches, timed launches)}."""
        t = (FwiKernelTime * len(_lib.KERNELS))()   # FWI_K_COUNT entries
        self._chk(self.lib.fwi_kernel_times(self.ctx, t))
        return {n: (t[i].launches, t[i].total_ms, t[i].timed) for i, n in enumerate(_lib.KERNELS)}

    def sync(self):
        self._chk(self.lib.fwi_sync(self.ctx))
