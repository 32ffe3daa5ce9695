"""fwi_mat_av2_noalloc (the allocation-free entry point) against fwi_set_model's mat_av2.

Both run the same two kernels; the no-alloc call keeps its block partial sums in the
library's fixed device array instead of the context's buffer, so the media and the
three means must be bit-identical.  Device buffers come from the HIP runtime the
library itself links (loaded through ctypes), on two calls in a row, the second on a
smaller grid, then a size the entry point must refuse.
"""
import ctypes as C

import numpy as np
import pytest

import cases
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    from fwi_cuda_amd import _lib
    _lib.load()   # brings in the library's HIP runtime (libamdhip64.so.7)
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
        f.restype = C.c_int
    return h


class Dev:
    """A device buffer from the library's HIP runtime."""

    def __init__(self, hip, nbytes):
        self.hip, self.n = hip, nbytes
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0

    @classmethod
    def of(cls, hip, a):
        a = np.ascontiguousarray(a)
        d = cls(hip, a.nbytes)
        assert hip.hipMemcpy(d.p, a.ctypes.data, a.nbytes, H2D) == 0
        return d

    def get(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def test_noalloc_mat_av2_equals_set_model(hip):
    from fwi_cuda_amd import _lib, psv
    from oracle import oracle as O
    O.load("oracle")
    lib = _lib.load()
    for nz, nx, seed in ((120, 200, 3), (50, 70, 4)):
        print(f"no-alloc mat_av2: {nz} x {nx}", flush=True)
        lam, mu, rho = cases.layered_model(nz, nx, seed)
        c = cases.small_case(O, nz=nz, nx=nx, nt=4, nshot=1, npml=8)
        with psv.Simulation(c) as sim:
            sim.set_model(lam, mu, rho)
            want = sim.mat_av2()
        ins = [Dev.of(hip, a) for a in (lam, mu, rho)]
        outs = [Dev(hip, 4 * nz * nx) for _ in range(5)]
        means = Dev(hip, 3 * 8)
        rc = lib.fwi_mat_av2_noalloc(*[d.p for d in ins], nz, nx, *[d.p for d in outs], means.p, None)
        assert rc == 0, lib.fwi_last_error(None)
        assert hip.hipDeviceSynchronize() == 0
        got = [d.get(np.float32, (nz, nx)) for d in outs]
        m = means.get(np.float64, (3,))
        for d in ins + outs + [means]:
            d.free()
        np.testing.assert_array_equal(got[0], lam.astype(np.float32))
        np.testing.assert_array_equal(got[1], mu.astype(np.float32))
        for k, g in zip(("mu_zx", "rho_zp", "rho_xp"), got[2:]):
            np.testing.assert_array_equal(g.astype(np.float64), want[k], err_msg=k)
        np.testing.assert_array_equal(m, want["C"])
    # sizes the entry point refuses, before touching the device
    rc = lib.fwi_mat_av2_noalloc(*([C.c_void_p(16)] * 3), 1, 70, *([C.c_void_p(16)] * 6), None)
    assert rc == -1 and b"below 2 x 2" in lib.fwi_last_error(None)
    rc = lib.fwi_mat_av2_noalloc(*([C.c_void_p(16)] * 3), 1 << 16, 1 << 15, *([C.c_void_p(16)] * 6), None)
    assert rc == -1 and b"above the maximum" in lib.fwi_last_error(None)
