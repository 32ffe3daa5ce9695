"""ctypes binding of include/fwi_psv.h (libfwi_psv.so, built in place).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfwi_psv.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

FWI_OK = 0
STATUS = {0: "FWI_OK", -1: "FWI_EINVAL", -2: "FWI_ENOMEM", -3: "FWI_EHIP", -4: "FWI_ENONFINITE",
          -5: "FWI_ECFL", -6: "FWI_ESTEP", -7: "FWI_ESTATE", -8: "FWI_ENODEV"}
STORE_AUTO, STORE_FULL, STORE_CHECKPOINT = 0, 1, 2
FORM_FUSED, FORM_PLAIN = 0, 1
PARAM_LAME, PARAM_VELOCITY = 0, 1
ABI_VERSION = 6
KERNELS = ("stress", "velocity", "surface", "source", "receiver", "store", "other", "step", "graph")


class FwiGrid(C.Structure):
    _fields_ = [("nt", C.c_int32), ("nz", C.c_int32), ("nx", C.c_int32),
                ("dt", C.c_double), ("dz", C.c_double), ("dx", C.c_double),
                ("fdorder", C.c_int32), ("hc", C.c_double * 7), ("isurf", C.c_int32 * 4),
                ("pml_z", C.c_int32), ("pml_x", C.c_int32)] + \
        [(n, _dp) for n in ("a_z", "b_z", "K_z", "a_half_z", "b_half_z", "K_half_z",
                            "a_x", "b_x", "K_x", "a_half_x", "b_half_x", "K_half_x")] + \
        [(n, C.c_int32) for n in ("snap_z1", "snap_z2", "snap_x1", "snap_x2", "snap_dt", "snap_dz", "snap_dx")] + \
        [("taper", C.c_int32 * 8)]


class FwiAcq(C.Structure):
    _fields_ = [("nshot", C.c_int32), ("stf_type", C.c_int32), ("src_off", _ip), ("src_z", _ip),
                ("src_x", _ip), ("stf_z", _dp), ("stf_x", _dp), ("rec_off", _ip), ("rec_z", _ip), ("rec_x", _ip)]


class FwiOpts(C.Structure):
    _fields_ = [("batch", C.c_int32), ("storage", C.c_int32), ("ckpt_interval", C.c_int32),
                ("mem_budget", C.c_int64), ("form", C.c_int32), ("grad_sign", C.c_int32),
                ("use_graph", C.c_int32), ("tile_tz", C.c_int32), ("device", C.c_int32),
                ("check_finite", C.c_int32), ("kernel_variant", C.c_int32), ("work_order", C.c_int32),
                ("param", C.c_int32), ("vel_bounds", C.c_double * 4), ("pcg", C.c_int32), ("smooth_hfs", C.c_int32),
                ("ckpt_slots", C.c_int32),
                ("seis_fp64", C.c_int32),
                ("ckpt_keep", C.c_int32)]


class FwiKernelTime(C.Structure):
    _fields_ = [("launches", C.c_int64), ("total_ms", C.c_double), ("timed", C.c_int64)]


class FwiError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise FileNotFoundError(f"{LIB_PATH} is not built; run `python -m fwi_cuda_amd.build` "
                                "(the HIP path has no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    sig = {
        "fwi_abi_version": (C.c_int32, []),
        "fwi_strerror": (C.c_char_p, [C.c_int32]),
        "fwi_last_error": (C.c_char_p, [vp]),
        "fwi_default_opts": (None, [C.POINTER(FwiOpts)]),
        "fwi_holbergcoeff": (C.c_int32, [C.c_int32, C.c_int32, _dp]),
        "fwi_cpml": (C.c_int32, [_dp] * 6 + [C.c_double] * 5 + [C.c_int32] * 4 + [C.c_double] * 2),
        "fwi_wavelet": (C.c_int32, [_dp, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32]),
        "fwi_dtstab": (C.c_double, [C.c_int32, C.c_int32, C.c_double, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp]),
        "fwi_create": (C.c_int32, [C.POINTER(FwiGrid), C.POINTER(FwiAcq), C.POINTER(FwiOpts), C.POINTER(vp)]),
        "fwi_destroy": (None, [vp]),
        "fwi_plan": (C.c_int32, [vp, _ip, _ip, _ip, C.POINTER(C.c_int64)]),
        "fwi_checkpoint_plan": (C.c_int32, [vp, _ip, _ip, _dp]),
        "fwi_checkpoint_kept": (C.c_int32, [vp, _ip]),
        "fwi_set_model_device": (C.c_int32, [vp, vp, vp, vp]),
        "fwi_forward_device": (C.c_int32, [vp, C.c_int32, C.c_int32, vp, vp]),
        "fwi_misfit_gradient_device": (C.c_int32, [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp]),
        "fwi_misfit_gradient_exact": (C.c_int32, [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, C.c_int32]),
        "fwi_pcg_direction": (C.c_int32, [vp]),
        "fwi_smooth_gradient": (C.c_int32, [vp, C.c_int32]),
        "fwi_set_model": (C.c_int32, [vp, _dp, _dp, _dp]),
        "fwi_get_model": (C.c_int32, [vp, _dp, _dp, _dp]),
        "fwi_get_media": (C.c_int32, [vp, _dp, _dp, _dp, _dp]),
        "fwi_set_observed": (C.c_int32, [vp, _dp, _dp]),
        "fwi_forward": (C.c_int32, [vp, C.c_int32, C.c_int32, _dp, _dp]),
        "fwi_get_wavefield": (C.c_int32, [vp, C.c_int32, C.c_int32, _dp]),
        "fwi_forward_snapshots": (C.c_int32, [vp, C.c_int32] + [_dp] * 7),
        "fwi_gradient": (C.c_int32, [vp, C.c_int32, C.c_int32, _dp]),
        "fwi_gradient_raw": (C.c_int32, [vp, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, _dp]),
        "fwi_misfit_gradient": (C.c_int32, [vp, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, _dp]),
        "fwi_zero_gradient": (C.c_int32, [vp]),
        "fwi_get_gradient": (C.c_int32, [vp, _dp, _dp, _dp]),
        "fwi_set_gradient": (C.c_int32, [vp, _dp, _dp, _dp]),
        "fwi_taper_gradient": (C.c_int32, [vp]),
        "fwi_misfit": (C.c_int32, [vp, C.c_int32, C.c_int32, _dp]),
        "fwi_save_model": (C.c_int32, [vp]),
        "fwi_update_model": (C.c_int32, [vp, C.c_double, C.c_double, C.c_double]),
        "fwi_step_length": (C.c_int32, [vp, C.c_double, C.c_double, _dp, _ip]),
        "fwi_iteration": (C.c_int32, [vp, _dp, _dp]),
        "fwi_tuning": (C.c_int32, [vp, _ip, C.c_char_p, C.c_int32]),
        "fwi_profile": (C.c_int32, [vp, C.c_int32]),
        "fwi_profile_kernels": (C.c_int32, [vp, C.c_uint32]),
        "fwi_kernel_times": (C.c_int32, [vp, C.POINTER(FwiKernelTime)]),
        "fwi_profile_stride": (C.c_int32, [vp, C.c_int32]),
        "fwi_sync": (C.c_int32, [vp]),
        "fwi_mat_av2_noalloc": (C.c_int32, [vp, vp, vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.fwi_abi_version() != ABI_VERSION:
        raise RuntimeError("libfwi_psv.so ABI version mismatch")
    _lib = lib
    return lib


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags.c_contiguous)
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def check(status: int, ctx=None):
    if status != FWI_OK:
        lib = load()
        msg = lib.fwi_last_error(ctx)
        raise FwiError(status, msg.decode() if msg else lib.fwi_strerror(status).decode())


EXPORTED = ("fwi_abi_version", "fwi_strerror", "fwi_last_error", "fwi_default_opts", "fwi_holbergcoeff",
            "fwi_cpml", "fwi_wavelet", "fwi_dtstab", "fwi_create", "fwi_destroy", "fwi_plan", "fwi_set_model",
            "fwi_get_model", "fwi_get_media", "fwi_set_observed", "fwi_forward", "fwi_get_wavefield",
            "fwi_gradient", "fwi_gradient_raw", "fwi_misfit_gradient", "fwi_zero_gradient", "fwi_get_gradient",
            "fwi_set_gradient",
            "fwi_taper_gradient", "fwi_misfit", "fwi_save_model", "fwi_update_model", "fwi_step_length",
            "fwi_iteration", "fwi_tuning", "fwi_profile", "fwi_profile_kernels", "fwi_profile_stride", "fwi_kernel_times", "fwi_sync", "fwi_forward_snapshots",
            "fwi_mat_av2_noalloc", "fwi_checkpoint_plan", "fwi_checkpoint_kept", "fwi_set_model_device", "fwi_forward_device",
            "fwi_misfit_gradient_device", "fwi_misfit_gradient_exact", "fwi_pcg_direction", "fwi_smooth_gradient")
