// psv_fp64.hip -- an FP64 forward propagation for the seismograms.
//
// With FP32 wavefields the synthetic seismograms carry a relative error of ~1e-6, and
// where the residual d_syn - d_obs is a small part of the data (late FWI iterations, a
// weak scatterer) that error is a visible part of the adjoint source: a gradient that is
// exact for the FP32-modelled residual differs from the FP64 reference's by ~1e-3.  When
// fwi_opts.seis_fp64 is set, the seismograms the misfit and the adjoint sources are made
// of come from this propagation instead: the same step as the exact path's forward
// kernels (vdiff2, pml_diff2, update_s2, surf_mirror, sdiff2, pml_diff2, update_v2,
// src/cpu/n_kernel_lib_PSV.cpp:100-395), vsrc2 (:487-536) and urec2 (:538-576), all in FP64
// with FP64 media and CPML profiles, one thread per cell.  The FP32 stencil still runs the
// wavefields the correlation uses (checkpoints, snapshots, energy).
#include <algorithm>

#include "psv_internal.h"

namespace psv {
namespace {

constexpr int kDB = 256;

__device__ __forceinline__ bool d_cell(const DArgs &d, size_t t, int &s, int &iz, int &ix) {
    const size_t per = (size_t)d.nz * d.nx;
    if (t >= per * d.B) return false;
    s = (int)(t / per);
    const size_t r = t - (size_t)s * per;
    iz = (int)(r / d.nx);
    ix = (int)(r - (size_t)iz * d.nx);
    return iz >= d.N && iz < d.nz - d.N && ix >= d.N && ix < d.nx - d.N;
}

// mat_av2 (src/cpu/n_kernel_lib_PSV.cpp:853-922) into FP64 [nz][P] arrays: harmonic mu_zx
// (0 where any of the four mu is 0), inverse mean densities, the rho_xp guard writing
// rho_zp as the reference's does, the last row and column 0
__global__ void dk_media(const double *lam, const double *mu, const double *rho, double *dlam, double *dmu,
                         double *dmuzx, double *drhoz, double *drhox, int nz, int nx, int P) {
    const size_t n = (size_t)nz * nx;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(i / nx), ix = (int)(i - (size_t)iz * nx);
        const size_t o = (size_t)iz * P + ix;
        dlam[o] = lam[i];
        dmu[o] = mu[i];
        if (iz < nz - 1 && ix < nx - 1) {
            const double m00 = mu[i], m01 = mu[i + 1], m10 = mu[i + nx], m11 = mu[i + nx + 1];
            double v = 4.0 / ((1.0 / m00) + (1.0 / m01) + (1.0 / m10) + (1.0 / m11));
            if (m00 == 0.0 || m01 == 0.0 || m10 == 0.0 || m11 == 0.0) v = 0.0;
            const double r00 = rho[i], r10 = rho[i + nx], r01 = rho[i + 1];
            double rz = 1.0 / (0.5 * (r00 + r10));
            const double rx = 1.0 / (0.5 * (r00 + r01));
            if (r00 < 1e-4 && r10 < 1e-4) rz = 0.0;
            if (r00 < 1e-4 && r01 < 1e-4) rz = 0.0;
            dmuzx[o] = v;
            drhoz[o] = rz;
            drhox[o] = rx;
        } else {
            dmuzx[o] = drhoz[o] = drhox[o] = 0.0;
        }
    }
}

// velocity derivatives, CPML (integer profile for d_z vz and d_x vx, half for the other
// two), stress update
__global__ void dk_stress(const DArgs d) {
    int s, iz, ix;
    if (!d_cell(d, (size_t)blockIdx.x * blockDim.x + threadIdx.x, s, iz, ix)) return;
    const size_t P = d.P, c = (size_t)iz * P + ix, o = (size_t)s * d.fs + c;
    const double *vz = d.vz + (size_t)s * d.fs, *vx = d.vx + (size_t)s * d.fs;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;   // vz_z, vx_z, vz_x, vx_x
    for (int k = 1; k <= d.N; ++k) {
        d0 += d.cz[k - 1] * (vz[c + (size_t)(k - 1) * P] - vz[c - (size_t)k * P]);
        d1 += d.cz[k - 1] * (vx[c + (size_t)k * P] - vx[c - (size_t)(k - 1) * P]);
        d2 += d.cx[k - 1] * (vz[c + k] - vz[c - (k - 1)]);
        d3 += d.cx[k - 1] * (vx[c + (k - 1)] - vx[c - k]);
    }
    if (d.pml_z) {
        const double *p = d.zp + (size_t)iz * 6;
        const double a = p[0] * d.pv[0][o] + p[1] * d0, b = p[3] * d.pv[1][o] + p[4] * d1;
        d.pv[0][o] = a;
        d.pv[1][o] = b;
        d0 = d0 * p[2] + a;
        d1 = d1 * p[5] + b;
    }
    if (d.pml_x) {
        const double *p = d.xp + (size_t)ix * 6;
        const double a = p[0] * d.pv[3][o] + p[1] * d3, b = p[3] * d.pv[2][o] + p[4] * d2;
        d.pv[3][o] = a;
        d.pv[2][o] = b;
        d3 = d3 * p[2] + a;
        d2 = d2 * p[5] + b;
    }
    d.E0[o] = d0;
    d.E3[o] = d3;
    const double l = d.lam[c], u = d.mu[c], dt = d.dt;
    d.szz[o] += dt * (l * (d3 + d0) + 2.0 * u * d0);
    d.szx[o] += dt * d.muzx[c] * (d2 + d1);
    d.sxx[o] += dt * (l * (d3 + d0) + 2.0 * u * d3);
}

// surf_mirror for one side with the strain rate the stress kernel kept
__global__ void dk_surface(const DArgs d, int side, int is) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
    const int N = d.N;
    const size_t P = d.P, f0 = (size_t)s * d.fs;
    if (side < 2) {
        const int ix = N + t;
        if (ix >= d.nx - N) return;
        const size_t r = (size_t)is * P + ix;
        const double l = d.lam[r], u = d.mu[r];
        d.szz[f0 + r] = 0.0;
        d.szx[f0 + r] = 0.0;
        d.sxx[f0 + r] = 4.0 * d.dt * d.E3[f0 + r] * (l * u + u * u) / (l + 2.0 * u);
        const int nk = side == 0 ? is - N : d.nz - N - is;
        for (int k = 1; k <= nk; ++k) {
            const size_t dst = side == 0 ? r - (size_t)k * P : r + (size_t)k * P;
            const size_t src = side == 0 ? r + (size_t)k * P : r - (size_t)k * P;
            d.szx[f0 + dst] = -d.szx[f0 + src];
            d.szz[f0 + dst] = -d.szz[f0 + src];
        }
    } else {
        const int iz = N + t;
        if (iz >= d.nz - N) return;
        const size_t r = (size_t)iz * P + is;
        const double l = d.lam[r], u = d.mu[r];
        d.sxx[f0 + r] = 0.0;
        d.szx[f0 + r] = 0.0;
        d.szz[f0 + r] = 4.0 * d.dt * d.E0[f0 + r] * (l * u + u * u) / (l + 2.0 * u);
        const int nk = side == 2 ? is - N : d.nx - N - is;
        for (int k = 1; k <= nk; ++k) {
            const size_t dst = side == 2 ? r - k : r + k, src = side == 2 ? r + k : r - k;
            d.szx[f0 + dst] = -d.szx[f0 + src];
            d.sxx[f0 + dst] = -d.sxx[f0 + src];
        }
    }
}

// stress derivatives, CPML (the reference's slot mapping), velocity update
__global__ void dk_velocity(const DArgs d) {
    int s, iz, ix;
    if (!d_cell(d, (size_t)blockIdx.x * blockDim.x + threadIdx.x, s, iz, ix)) return;
    const size_t P = d.P, c = (size_t)iz * P + ix, o = (size_t)s * d.fs + c;
    const double *zz = d.szz + (size_t)s * d.fs, *zx = d.szx + (size_t)s * d.fs, *xx = d.sxx + (size_t)s * d.fs;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;   // szz_z, szx_z, szx_x, sxx_x
    for (int k = 1; k <= d.N; ++k) {
        d0 += d.cz[k - 1] * (zz[c + (size_t)k * P] - zz[c - (size_t)(k - 1) * P]);
        d1 += d.cz[k - 1] * (zx[c + (size_t)(k - 1) * P] - zx[c - (size_t)k * P]);
        d2 += d.cx[k - 1] * (zx[c + (k - 1)] - zx[c - k]);
        d3 += d.cx[k - 1] * (xx[c + k] - xx[c - (k - 1)]);
    }
    if (d.pml_z) {
        const double *p = d.zp + (size_t)iz * 6;
        const double a = p[0] * d.ps[0][o] + p[1] * d0, b = p[3] * d.ps[1][o] + p[4] * d1;
        d.ps[0][o] = a;
        d.ps[1][o] = b;
        d0 = d0 * p[2] + a;
        d1 = d1 * p[5] + b;
    }
    if (d.pml_x) {
        const double *p = d.xp + (size_t)ix * 6;
        const double a = p[0] * d.ps[3][o] + p[1] * d3, b = p[3] * d.ps[2][o] + p[4] * d2;
        d.ps[3][o] = a;
        d.ps[2][o] = b;
        d3 = d3 * p[2] + a;
        d2 = d2 * p[5] + b;
    }
    d.vz[o] += d.dt * d.rhoz[c] * (d2 + d0);
    d.vx[o] += d.dt * d.rhox[c] * (d3 + d1);
}

// vsrc2: a group's sources added in source order
__global__ void dk_inject(const DArgs d, const SrcArgs sa, int it) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= sa.ngroups) return;
    const int s = sa.grp_shot[g] - sa.shot0, cell = sa.grp_cell[g];
    const size_t o = (size_t)s * d.fs + cell;
    double vz = d.vz[o], vx = d.vx[o];
    const double sc = sa.scale_shot ? sa.scale_shot[sa.grp_shot[g]] : 1.0;
    for (int j = sa.grp_off[g]; j < sa.grp_off[g + 1]; ++j) {
        const size_t row = (size_t)sa.grp_idx[j] * sa.nt + it;
        const double fz = sa.stf_z[row] * sc, fx = sa.stf_x[row] * sc;
        if (sa.type == 0) {
            vz += sa.dt * d.rhoz[cell] * fz / sa.dzdx;
            vx += sa.dt * d.rhox[cell] * fx / sa.dzdx;
        } else {
            vz += fz;
            vx += fx;
        }
    }
    d.vz[o] = vz;
    d.vx[o] = vx;
}

// urec2: displacement as the running sum of dt v / (dz dx)
__global__ void dk_record(const DArgs d, const RecArgs ra, int it) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ra.nrec) return;
    const size_t o = (size_t)(ra.rec_shot[r] - ra.shot0) * d.fs + ra.rec_cell[r];
    const double uz = ra.acc_z[r] + ra.dt * d.vz[o] / ra.dzdx;
    const double ux = ra.acc_x[r] + ra.dt * d.vx[o] / ra.dzdx;
    ra.acc_z[r] = uz;
    ra.acc_x[r] = ux;
    ra.seis_z[(size_t)r * ra.nt + it] = uz;
    ra.seis_x[(size_t)r * ra.nt + it] = ux;
}

unsigned cells_grid(const DArgs &d) { return (unsigned)(((size_t)d.B * d.nz * d.nx + kDB - 1) / kDB); }

}  // namespace

void d_media(const double *lam, const double *mu, const double *rho, double *dlam, double *dmu, double *dmuzx,
             double *drhoz, double *drhox, int nz, int nx, int P, hipStream_t st) {
    const size_t n = (size_t)nz * nx;
    const unsigned g = (unsigned)std::min<size_t>((n + kDB - 1) / kDB, 4096);
    hipLaunchKernelGGL(dk_media, dim3(g), dim3(kDB), 0, st, lam, mu, rho, dlam, dmu, dmuzx, drhoz, drhox, nz, nx, P);
}

void d_forward_step(const DArgs &d, const int isurf[4], hipStream_t st) {
    hipLaunchKernelGGL(dk_stress, dim3(cells_grid(d)), dim3(kDB), 0, st, d);
    for (int side = 0; side < 4; ++side) {
        if (isurf[side] <= 0) continue;
        const int len = side < 2 ? d.nx - 2 * d.N : d.nz - 2 * d.N;
        hipLaunchKernelGGL(dk_surface, dim3((len + kDB - 1) / kDB, d.B), dim3(kDB), 0, st, d, side, isurf[side]);
    }
    hipLaunchKernelGGL(dk_velocity, dim3(cells_grid(d)), dim3(kDB), 0, st, d);
}

void d_inject(const DArgs &d, const SrcArgs &sa, int it, hipStream_t st) {
    if (sa.ngroups <= 0) return;
    hipLaunchKernelGGL(dk_inject, dim3((sa.ngroups + kDB - 1) / kDB), dim3(kDB), 0, st, d, sa, it);
}

void d_record(const DArgs &d, const RecArgs &ra, int it, hipStream_t st) {
    if (ra.nrec <= 0) return;
    hipLaunchKernelGGL(dk_record, dim3((ra.nrec + kDB - 1) / kDB), dim3(kDB), 0, st, d, ra, it);
}

}  // namespace psv
