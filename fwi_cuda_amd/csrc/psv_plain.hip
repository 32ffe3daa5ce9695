// psv_plain.hip -- the reference GPU path's step structure, rebuilt for gfx950.
//
// The reference's dormant CUDA path runs each sub-step of kernel_PSV as its own
// kernel over the whole grid with 32x32 thread blocks, keeps the CPML memory
// and the four derivative arrays at full grid size, and synchronises the
// device after every launch (src/cuda/d_kernel_PSV.cu:16-220,
// src/cuda/d_kernel_lib_PSV.cu:289-576).  This file reproduces that structure
// (FWI_FORM_PLAIN) so the fused form can be measured against it on the same
// hardware, with the same arithmetic.  One shot at a time.
#include "psv_internal.h"

namespace psv {

namespace {

constexpr int kT = 32;

struct PArgs {
    KArgs a;
    PlainBufs p;
    int N;
};

// vdiff2 / sdiff2: derivatives into the scratch slots dz_z, dx_z, dz_x, dx_x
__global__ __launch_bounds__(kT *kT) void k_pdiff(const PArgs q, int stress) {
    const KArgs &a = q.a;
    const int N = q.N;
    const int ix = N + blockIdx.x * kT + threadIdx.x, iz = N + blockIdx.y * kT + threadIdx.y;
    if (ix >= a.nx - N || iz >= a.nz - N) return;
    const int P = a.P;
    const size_t r = (size_t)iz * P + ix;
    float d0 = 0, d1 = 0, d2 = 0, d3 = 0;
    if (!stress) {
        for (int k = 1; k <= N; ++k) {
            d0 += a.cz[k - 1] * (a.vz[r + (size_t)(k - 1) * P] - a.vz[r - (size_t)k * P]);
            d1 += a.cz[k - 1] * (a.vx[r + (size_t)k * P] - a.vx[r - (size_t)(k - 1) * P]);
            d2 += a.cx[k - 1] * (a.vz[r + k] - a.vz[r - k + 1]);
            d3 += a.cx[k - 1] * (a.vx[r + k - 1] - a.vx[r - k]);
        }
    } else {
        for (int k = 1; k <= N; ++k) {
            d0 += a.cz[k - 1] * (a.szz[r + (size_t)k * P] - a.szz[r - (size_t)(k - 1) * P]);
            d1 += a.cz[k - 1] * (a.szx[r + (size_t)(k - 1) * P] - a.szx[r - (size_t)k * P]);
            d2 += a.cx[k - 1] * (a.szx[r + k - 1] - a.szx[r - k]);
            d3 += a.cx[k - 1] * (a.sxx[r + k] - a.sxx[r - k + 1]);
        }
    }
    q.p.dz_z[r] = d0;
    q.p.dx_z[r] = d1;
    q.p.dz_x[r] = d2;
    q.p.dx_x[r] = d3;
}

// pml_diff2 over the whole interior with full-size memory arrays
__global__ __launch_bounds__(kT *kT) void k_ppml(const PArgs q, float *mzz, float *mxz, float *mzx, float *mxx,
                                                  int pml_z, int pml_x) {
    const KArgs &a = q.a;
    const PlainBufs &p = q.p;
    const int N = q.N;
    const int ix = N + blockIdx.x * kT + threadIdx.x, iz = N + blockIdx.y * kT + threadIdx.y;
    if (ix >= a.nx - N || iz >= a.nz - N) return;
    const size_t r = (size_t)iz * a.P + ix;
    if (pml_z) {
        mzz[r] = p.fbz[iz] * mzz[r] + p.faz[iz] * p.dz_z[r];
        mxz[r] = p.fbhz[iz] * mxz[r] + p.fahz[iz] * p.dx_z[r];
        p.dz_z[r] = p.dz_z[r] * p.fikz[iz] + mzz[r];
        p.dx_z[r] = p.dx_z[r] * p.fikhz[iz] + mxz[r];
    }
    if (pml_x) {
        mxx[r] = p.fbx[ix] * mxx[r] + p.fax[ix] * p.dx_x[r];
        mzx[r] = p.fbhx[ix] * mzx[r] + p.fahx[ix] * p.dz_x[r];
        p.dx_x[r] = p.dx_x[r] * p.fikx[ix] + mxx[r];
        p.dz_x[r] = p.dz_x[r] * p.fikhx[ix] + mzx[r];
    }
}

__global__ __launch_bounds__(kT *kT) void k_pupdate_s(const PArgs q) {
    const KArgs &a = q.a;
    const PlainBufs &p = q.p;
    const int N = q.N;
    const int ix = N + blockIdx.x * kT + threadIdx.x, iz = N + blockIdx.y * kT + threadIdx.y;
    if (ix >= a.nx - N || iz >= a.nz - N) return;
    const size_t r = (size_t)iz * a.P + ix;
    const float vz_z = p.dz_z[r], vx_z = p.dx_z[r], vz_x = p.dz_x[r], vx_x = p.dx_x[r];
    const float lam = a.lam[r], mu = a.mu[r];
    a.szz[r] += a.dt * (lam * (vx_x + vz_z) + 2.f * mu * vz_z);
    a.szx[r] += a.dt * a.muzx[r] * (vz_x + vx_z);
    a.sxx[r] += a.dt * (lam * (vx_x + vz_z) + 2.f * mu * vx_x);
}

// surf_mirror reading the stored derivative slots, one side per launch
__global__ void k_psurf(const PArgs q, int side, int is) {
    const KArgs &a = q.a;
    const PlainBufs &p = q.p;
    const int N = q.N, P = a.P;
    const int t = N + blockIdx.x * blockDim.x + threadIdx.x;
    if (side < 2) {
        if (t >= a.nx - N) return;
        const size_t r = (size_t)is * P + t;
        const float l = a.lam[r], u = a.mu[r];
        a.szz[r] = 0.f;
        a.szx[r] = 0.f;
        a.sxx[r] = 4.f * a.dt * p.dx_x[r] * (l * u + u * u) / (l + 2.f * u);
        if (side == 0)
            for (int k = 1; k < is - N + 1; ++k) {
                a.szx[r - (size_t)k * P] = -a.szx[r + (size_t)k * P];
                a.szz[r - (size_t)k * P] = -a.szz[r + (size_t)k * P];
            }
        else
            for (int k = 1; k <= a.nz - N - is; ++k) {
                a.szx[r + (size_t)k * P] = -a.szx[r - (size_t)k * P];
                a.szz[r + (size_t)k * P] = -a.szz[r - (size_t)k * P];
            }
    } else {
        if (t >= a.nz - N) return;
        const size_t r = (size_t)t * P + is;
        const float l = a.lam[r], u = a.mu[r];
        a.sxx[r] = 0.f;
        a.szx[r] = 0.f;
        a.szz[r] = 4.f * a.dt * p.dz_z[r] * (l * u + u * u) / (l + 2.f * u);
        if (side == 2)
            for (int k = 1; k < is - N + 1; ++k) {
                a.szx[r - k] = -a.szx[r + k];
                a.sxx[r - k] = -a.sxx[r + k];
            }
        else
            for (int k = 1; k <= a.nx - N - is; ++k) {
                a.szx[r + k] = -a.szx[r - k];
                a.sxx[r + k] = -a.sxx[r - k];
            }
    }
}

__global__ __launch_bounds__(kT *kT) void k_pupdate_v(const PArgs q) {
    const KArgs &a = q.a;
    const PlainBufs &p = q.p;
    const int N = q.N;
    const int ix = N + blockIdx.x * kT + threadIdx.x, iz = N + blockIdx.y * kT + threadIdx.y;
    if (ix >= a.nx - N || iz >= a.nz - N) return;
    const size_t r = (size_t)iz * a.P + ix;
    const float vz0 = a.vz[r], vx0 = a.vx[r];
    // the reference path updates u in every pass (src/cuda/d_kernel_lib_PSV.cu:526-554)
    a.uz[r] += a.dt * vz0;
    a.ux[r] += a.dt * vx0;
    const float vz1 = vz0 + a.dt * a.rhoz[r] * (p.dz_x[r] + p.dz_z[r]);
    const float vx1 = vx0 + a.dt * a.rhox[r] * (p.dx_x[r] + p.dx_z[r]);
    a.vz[r] = vz1;
    a.vx[r] = vx1;
    a.We[r] += vx1 * vx1 + vz1 * vz1;
}

// fwi_grad2 as its own pass over the snap grid
__global__ void k_pgrad(const KArgs a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.snz * a.snx) return;
    const int jz = q / a.snx, jx = q - jz * a.snx;
    const size_t r = (size_t)(a.snap_z1 + jz * a.snap_dz) * a.P + a.snap_x1 + jx * a.snap_dx;
    const float Szz = a.sn_szz[q], Sxx = a.sn_sxx[q], Szx = a.sn_szx[q], Vz = a.sn_vz[q], Vx = a.sn_vx[q];
    const float lam = a.lam[r], mu = a.mu[r], lm = lam + mu;
    const float szz = a.szz[r], sxx = a.sxx[r], szx = a.szx[r];
    const float il = a.gain_lm / (lm * lm), im = a.gain_lm / (mu * mu);
    const float s1 = 0.25f * il * (Szz + Sxx) * (szz + sxx);
    const float s2 = 0.25f * im * (Szz - Sxx) * (szz - sxx);
    const float s3 = im * (Szx * szx);
    const float s4 = a.ux[r] * Vx + a.uz[r] * Vz;
    a.g_lam[q] += a.gw * s1;
    a.g_mu[q] += a.gw * (s3 + s1 + s2);
    a.g_rho[q] += a.gw * s4;
}

}  // namespace

void plain_step(const KArgs &a, const PlainBufs &p, int N, const int isurf[4], int pml_z, int pml_x,
                hipStream_t st) {
    PArgs q{a, p, N};
    const dim3 block(kT, kT), grid((a.nx - 2 * N + kT - 1) / kT, (a.nz - 2 * N + kT - 1) / kT);
    hipLaunchKernelGGL(k_pdiff, grid, block, 0, st, q, 0);
    (void)hipStreamSynchronize(st);
    if (pml_z || pml_x) {
        hipLaunchKernelGGL(k_ppml, grid, block, 0, st, q, p.mvz_z, p.mvx_z, p.mvz_x, p.mvx_x, pml_z, pml_x);
        (void)hipStreamSynchronize(st);
    }
    hipLaunchKernelGGL(k_pupdate_s, grid, block, 0, st, q);
    (void)hipStreamSynchronize(st);
    for (int side = 0; side < 4; ++side) {
        if (isurf[side] <= 0) continue;
        const int n = side < 2 ? a.nx - 2 * N : a.nz - 2 * N;
        hipLaunchKernelGGL(k_psurf, dim3((n + 255) / 256), dim3(256), 0, st, q, side, isurf[side]);
        (void)hipStreamSynchronize(st);
    }
    hipLaunchKernelGGL(k_pdiff, grid, block, 0, st, q, 1);
    (void)hipStreamSynchronize(st);
    if (pml_z || pml_x) {
        hipLaunchKernelGGL(k_ppml, grid, block, 0, st, q, p.mszz_z, p.mszx_z, p.mszx_x, p.msxx_x, pml_z, pml_x);
        (void)hipStreamSynchronize(st);
    }
    hipLaunchKernelGGL(k_pupdate_v, grid, block, 0, st, q);
    (void)hipStreamSynchronize(st);
}

void plain_grad(const KArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_pgrad, dim3((a.snz * a.snx + 255) / 256), dim3(256), 0, st, a);
    (void)hipStreamSynchronize(st);
}

}  // namespace psv
