// psv_noalloc.hip -- mat_av2 on caller-owned device arrays without any allocation.
//
// fwi_set_model derives the stencil's media from a model with one pass of k_mat_av
// (src/cpu/n_kernel_lib_PSV.cpp:853-922) whose block partial sums of lam, mu and rho
// go to the context's reduction buffer before a fixed-order sum gives the three means.
// fwi_mat_av2_noalloc runs the same two kernels on arrays the caller owns and puts the
// partial sums in `scratch` below, a fixed device array that lives as long as the
// process: it calls no hipMalloc, hipFree or allocating library routine, so graph
// capture and loops that must not allocate can call it.  Each GPU holds its own copy
// of the array; its address on a device comes from hipGetSymbolAddress, taken once per
// device on the first call there.  Calls that overlap on one GPU share the array and
// race; each call writes every partial it reads, so nothing needs clearing.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "psv_internal.h"

namespace {

// kMatAvMaxParts partial triples (lam, mu, rho sums of one block each): 12 KiB
__device__ __align__(256) double scratch[3 * psv::kMatAvMaxParts];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};   // the array's address on each device, once known

int fail(int code, const char *fmt, long long a = 0, long long b = 0) {
    char buf[256];
    std::snprintf(buf, sizeof buf, fmt, a, b);
    psv::set_global_error(buf);
    return code;
}

}  // namespace

extern "C" int32_t fwi_mat_av2_noalloc(const double *lam, const double *mu, const double *rho, int32_t nz,
                                       int32_t nx, float *lam_f, float *mu_f, float *mu_zx, float *rho_zp,
                                       float *rho_xp, double *means, void *stream) {
    if (!lam || !mu || !rho || !lam_f || !mu_f || !mu_zx || !rho_zp || !rho_xp || !means)
        return fail(FWI_EINVAL, "fwi_mat_av2_noalloc: null array");
    if (nz < 2 || nx < 2) return fail(FWI_EINVAL, "fwi_mat_av2_noalloc: grid %lld x %lld below 2 x 2", nz, nx);
    if ((long long)nz * nx > FWI_NOALLOC_MAX_CELLS)
        return fail(FWI_EINVAL, "fwi_mat_av2_noalloc: %lld cells above the maximum of %lld",
                    (long long)nz * nx, (long long)FWI_NOALLOC_MAX_CELLS);
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return fail(FWI_ENODEV, "fwi_mat_av2_noalloc: no HIP device");
    if (dev >= kMaxDevices) return fail(FWI_ENODEV, "fwi_mat_av2_noalloc: device %lld beyond %lld", dev, kMaxDevices);
    double *part = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_scratch[dev]) {   // once per device, with that device current
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p) {
                (void)hipGetLastError();
                return fail(FWI_EHIP, "fwi_mat_av2_noalloc: the device scratch array has no address (error %lld)",
                            (long long)e);
            }
            g_scratch[dev] = static_cast<double *>(p);
        }
        part = g_scratch[dev];
    }
    // the partial count never exceeds the array: mat_av_parts caps it at kMatAvMaxParts
    const hipStream_t st = static_cast<hipStream_t>(stream);
    psv::dev_mat_av_into(lam, mu, rho, lam_f, mu_f, mu_zx, rho_zp, rho_xp, part, means, nz, nx, nx, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FWI_EHIP, "fwi_mat_av2_noalloc: launch failed (error %lld)", (long long)e);
    return FWI_OK;
}
