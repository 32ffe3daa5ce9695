// psv_exact.hip -- the exact discrete adjoint of the time step (reverse mode).
//
// The library's gradient (fwi_gradient, fwi_misfit_gradient) follows the reference's
// correlation (src/cpu/n_kernel_lib_PSV.cpp:439-484), which pairs forward and adjoint
// fields at one time level and is not the derivative of the discrete misfit.  These
// kernels give that derivative: the forward step is run once more storing, per step and
// cell, the five quantities the media enter through -- the CPML-corrected strain rates
// e_zz = d_z vz, e_xx = d_x vx, e_zx = d_x vz + d_z vx of the stress update and the
// stress-divergence sums S_z, S_x of the velocity update -- and the reverse sweep applies
// the transpose of every operation of the step in reverse order: the receivers' running
// sum (its adjoint is the reverse cumulative sum of the cotangents, injected as a
// velocity source), the velocity update, the CPML recursion (psi' = b psi + a d,
// d~ = d / K + psi' transposes to P = carry + d~', d' = d~' / K + a P, carry = b P), the
// staggered derivatives (as gathers: no two threads write one cell), the free surface
// (sides in reverse order; an assignment passes no adjoint to what it overwrote) and the
// stress update.  The media gradient is accumulated in FP64 per shot:
//   lam  += dt (s'_zz + s'_xx)(e_zz + e_xx)       mu    += 2 dt (s'_zz e_zz + s'_xx e_xx)
//   mu_zx += dt s'_zx e_zx                          rho_z^-1 += dt v'_z S_z,  rho_x^-1 likewise
// plus the free-surface terms of the set stress 4 dt e (lam mu + mu^2) / (lam + 2 mu) and the
// type-0 sources' inverse-density factor.  One thread per cell, every interior cell
// stored: this path serves the autograd backward and its checks, not the FWI loop.
#include <algorithm>

#include "psv_internal.h"

namespace psv {
namespace {

constexpr int kXB = 256;

__device__ __forceinline__ bool interior(const XArgs &x, int iz, int ix) {
    return iz >= x.N && iz < x.nz - x.N && ix >= x.N && ix < x.nx - x.N;
}

// one cell per thread over [B][nz][nx]
__device__ __forceinline__ bool cell_of(const XArgs &x, size_t t, int &s, int &iz, int &ix) {
    const size_t per = (size_t)x.nz * x.nx;
    if (t >= per * x.B) return false;
    s = (int)(t / per);
    const size_t r = t - (size_t)s * per;
    iz = (int)(r / x.nx);
    ix = (int)(r - (size_t)iz * x.nx);
    return true;
}

// ---- forward: strain rates, stress update, intermediates e_zz e_xx e_zx --------------
__global__ void xk_stress(const XArgs x, int n) {
    int s, iz, ix;
    if (!cell_of(x, (size_t)blockIdx.x * blockDim.x + threadIdx.x, s, iz, ix) || !interior(x, iz, ix)) return;
    const size_t P = x.P, o = (size_t)s * x.fs + (size_t)iz * P + ix, m = (size_t)iz * P + ix;
    const float *vz = x.vz + (size_t)s * x.fs, *vx = x.vx + (size_t)s * x.fs;
    const size_t c = (size_t)iz * P + ix;
    float d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f;   // vz_z, vx_z, vz_x, vx_x
    for (int k = 1; k <= x.N; ++k) {
        d0 += x.cz[k - 1] * (vz[c + (size_t)(k - 1) * P] - vz[c - (size_t)k * P]);
        d1 += x.cz[k - 1] * (vx[c + (size_t)k * P] - vx[c - (size_t)(k - 1) * P]);
        d2 += x.cx[k - 1] * (vz[c + k] - vz[c - (k - 1)]);
        d3 += x.cx[k - 1] * (vx[c + (k - 1)] - vx[c - k]);
    }
    if (x.pml_z) {
        const CpmlPoint p = x.zp[iz];
        const float a = p.b * x.pv[0][o] + p.a * d0, b = p.bh * x.pv[1][o] + p.ah * d1;
        x.pv[0][o] = a;
        x.pv[1][o] = b;
        d0 = d0 * p.ik + a;
        d1 = d1 * p.ikh + b;
    }
    if (x.pml_x) {
        const CpmlPoint p = x.xp[ix];
        const float a = p.b * x.pv[3][o] + p.a * d3, b = p.bh * x.pv[2][o] + p.ah * d2;
        x.pv[3][o] = a;
        x.pv[2][o] = b;
        d3 = d3 * p.ik + a;
        d2 = d2 * p.ikh + b;
    }
    x.E[0][o] = d0;
    x.E[1][o] = d1;
    x.E[2][o] = d2;
    x.E[3][o] = d3;
    const float l = x.lam[m], u = x.mu[m], dt = x.dt;
    x.szz[o] += dt * (l * (d3 + d0) + 2.f * u * d0);
    x.szx[o] += dt * x.muzx[m] * (d2 + d1);
    x.sxx[o] += dt * (l * (d3 + d0) + 2.f * u * d3);
    float *h = x.hist + ((size_t)(n - x.hist_base) * x.B + s) * 5 * x.fs;
    h[0 * x.fs + c] = d0;
    h[1 * x.fs + c] = d3;
    h[2 * x.fs + c] = d2 + d1;
}

// surf_mirror for one side (src/cpu/n_kernel_lib_PSV.cpp:272-395) with the strain rate the
// stress kernel kept; one thread per line of the side
__global__ void xk_surface(const XArgs x, int side, int is) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
    const int N = x.N;
    const size_t P = x.P, f0 = (size_t)s * x.fs;
    if (side < 2) {
        const int ix = N + t;
        if (ix >= x.nx - N) return;
        const size_t r = (size_t)is * P + ix;
        const float l = x.lam[r], u = x.mu[r];
        x.szz[f0 + r] = 0.f;
        x.szx[f0 + r] = 0.f;
        x.sxx[f0 + r] = 4.f * x.dt * x.E[3][f0 + r] * (l * u + u * u) / (l + 2.f * u);
        const int nk = side == 0 ? is - N : x.nz - N - is;
        for (int k = 1; k <= nk; ++k) {
            const size_t dst = side == 0 ? r - (size_t)k * P : r + (size_t)k * P;
            const size_t src = side == 0 ? r + (size_t)k * P : r - (size_t)k * P;
            x.szx[f0 + dst] = -x.szx[f0 + src];
            x.szz[f0 + dst] = -x.szz[f0 + src];
        }
    } else {
        const int iz = N + t;
        if (iz >= x.nz - N) return;
        const size_t r = (size_t)iz * P + is;
        const float l = x.lam[r], u = x.mu[r];
        x.sxx[f0 + r] = 0.f;
        x.szx[f0 + r] = 0.f;
        x.szz[f0 + r] = 4.f * x.dt * x.E[0][f0 + r] * (l * u + u * u) / (l + 2.f * u);
        const int nk = side == 2 ? is - N : x.nx - N - is;
        for (int k = 1; k <= nk; ++k) {
            const size_t dst = side == 2 ? r - k : r + k, src = side == 2 ? r + k : r - k;
            x.szx[f0 + dst] = -x.szx[f0 + src];
            x.sxx[f0 + dst] = -x.sxx[f0 + src];
        }
    }
}

// stress derivatives, CPML, velocity update, intermediates S_z S_x
__global__ void xk_velocity(const XArgs x, int n) {
    int s, iz, ix;
    if (!cell_of(x, (size_t)blockIdx.x * blockDim.x + threadIdx.x, s, iz, ix) || !interior(x, iz, ix)) return;
    const size_t P = x.P, c = (size_t)iz * P + ix, o = (size_t)s * x.fs + c;
    const float *zz = x.szz + (size_t)s * x.fs, *zx = x.szx + (size_t)s * x.fs, *xx = x.sxx + (size_t)s * x.fs;
    float d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f;   // szz_z, szx_z, szx_x, sxx_x
    for (int k = 1; k <= x.N; ++k) {
        d0 += x.cz[k - 1] * (zz[c + (size_t)k * P] - zz[c - (size_t)(k - 1) * P]);
        d1 += x.cz[k - 1] * (zx[c + (size_t)(k - 1) * P] - zx[c - (size_t)k * P]);
        d2 += x.cx[k - 1] * (zx[c + (k - 1)] - zx[c - k]);
        d3 += x.cx[k - 1] * (xx[c + k] - xx[c - (k - 1)]);
    }
    if (x.pml_z) {   // the reference's slot mapping: integer profile for szz_z, half for szx_z
        const CpmlPoint p = x.zp[iz];
        const float a = p.b * x.ps[0][o] + p.a * d0, b = p.bh * x.ps[1][o] + p.ah * d1;
        x.ps[0][o] = a;
        x.ps[1][o] = b;
        d0 = d0 * p.ik + a;
        d1 = d1 * p.ikh + b;
    }
    if (x.pml_x) {
        const CpmlPoint p = x.xp[ix];
        const float a = p.b * x.ps[3][o] + p.a * d3, b = p.bh * x.ps[2][o] + p.ah * d2;
        x.ps[3][o] = a;
        x.ps[2][o] = b;
        d3 = d3 * p.ik + a;
        d2 = d2 * p.ikh + b;
    }
    const float Sz = d2 + d0, Sx = d3 + d1;
    x.vz[o] += x.dt * x.rhoz[c] * Sz;
    x.vx[o] += x.dt * x.rhox[c] * Sx;
    float *h = x.hist + ((size_t)(n - x.hist_base) * x.B + s) * 5 * x.fs;
    h[3 * x.fs + c] = Sz;
    h[4 * x.fs + c] = Sx;
}

// ---- reverse -----------------------------------------------------------------------
// type-0 sources: the injected amount dt rho^-1 stf / (dz dx) carries the cell's inverse density
__global__ void xa_source_rho(const XArgs x, const SrcArgs sa, int n) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= sa.ngroups) return;
    const int s = sa.grp_shot[g] - sa.shot0, cell = sa.grp_cell[g];
    const size_t o = (size_t)s * x.fs + cell;
    double fz = 0.0, fx = 0.0;
    for (int j = sa.grp_off[g]; j < sa.grp_off[g + 1]; ++j) {
        fz += sa.stf_z[(size_t)sa.grp_idx[j] * sa.nt + n];
        fx += sa.stf_x[(size_t)sa.grp_idx[j] * sa.nt + n];
    }
    double *G = x.G + (size_t)s * 5 * x.fs;
    G[3 * x.fs + cell] += (double)x.bvz[o] * sa.dt * fz / sa.dzdx;
    G[4 * x.fs + cell] += (double)x.bvx[o] * sa.dt * fx / sa.dzdx;
}

// velocity update reversed: the density terms, the stress-derivative adjoints and the
// velocity pass's CPML reversed (0 outside the interior)
__global__ void xa_velocity(const XArgs x, int n) {
    int s, iz, ix;
    if (!cell_of(x, (size_t)blockIdx.x * blockDim.x + threadIdx.x, s, iz, ix)) return;
    const size_t c = (size_t)iz * x.P + ix, o = (size_t)s * x.fs + c;
    if (!interior(x, iz, ix)) {
        x.D[0][o] = x.D[1][o] = x.D[2][o] = x.D[3][o] = 0.f;
        return;
    }
    const float *h = x.hist + ((size_t)(n - x.hist_base) * x.B + s) * 5 * x.fs;
    const float az = x.bvz[o], ax = x.bvx[o];
    double *G = x.G + (size_t)s * 5 * x.fs;
    G[3 * x.fs + c] += (double)az * x.dt * h[3 * x.fs + c];
    G[4 * x.fs + c] += (double)ax * x.dt * h[4 * x.fs + c];
    const float tz = x.dt * x.rhoz[c] * az, tx = x.dt * x.rhox[c] * ax;
    float d0 = tz, d2 = tz, d1 = tx, d3 = tx;   // szz_z, szx_x; szx_z, sxx_x
    if (x.pml_z) {
        const CpmlPoint p = x.zp[iz];
        float q = x.cs[0][o] + d0;
        d0 = d0 * p.ik + p.a * q;
        x.cs[0][o] = p.b * q;
        q = x.cs[1][o] + d1;
        d1 = d1 * p.ikh + p.ah * q;
        x.cs[1][o] = p.bh * q;
    }
    if (x.pml_x) {
        const CpmlPoint p = x.xp[ix];
        float q = x.cs[3][o] + d3;
        d3 = d3 * p.ik + p.a * q;
        x.cs[3][o] = p.b * q;
        q = x.cs[2][o] + d2;
        d2 = d2 * p.ikh + p.ah * q;
        x.cs[2][o] = p.bh * q;
    }
    x.D[0][o] = d0;
    x.D[1][o] = d1;
    x.D[2][o] = d2;
    x.D[3][o] = d3;
}

__device__ __forceinline__ float at(const float *f, const XArgs &x, int iz, int ix) {
    return (iz < 0 || iz >= x.nz || ix < 0 || ix >= x.nx) ? 0.f : f[(size_t)iz * x.P + ix];
}

// stress-derivative transposes, gathered onto every cell (the derivative adjoints are 0
// outside the interior)
__global__ void xa_stress_gather(const XArgs x) {
    int s, iz, ix;
    if (!cell_of(x, (size_t)blockIdx.x * blockDim.x + threadIdx.x, s, iz, ix)) return;
    const size_t so = (size_t)s * x.fs, o = so + (size_t)iz * x.P + ix;
    const float *D0 = x.D[0] + so, *D1 = x.D[1] + so, *D2 = x.D[2] + so, *D3 = x.D[3] + so;
    float zz = 0.f, zx = 0.f, xx = 0.f;
    for (int k = 1; k <= x.N; ++k) {
        zz += x.cz[k - 1] * (at(D0, x, iz - k, ix) - at(D0, x, iz + k - 1, ix));
        zx += x.cz[k - 1] * (at(D1, x, iz - k + 1, ix) - at(D1, x, iz + k, ix));
        zx += x.cx[k - 1] * (at(D2, x, iz, ix - k + 1) - at(D2, x, iz, ix + k));
        xx += x.cx[k - 1] * (at(D3, x, iz, ix - k) - at(D3, x, iz, ix + k - 1));
    }
    x.bzz[o] += zz;
    x.bzx[o] += zx;
    x.bxx[o] += xx;
}

// surf_mirror reversed for one side: images undone, then the assignment (its adjoint goes
// to the strain rate it was made from, slot 3 for top / bottom, 0 for left / right, and to
// lam, mu); the surface line's stress adjoints end at 0
__global__ void xa_surface(const XArgs x, int side, int is, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
    const int N = x.N;
    const size_t P = x.P, f0 = (size_t)s * x.fs;
    size_t r;
    if (side < 2) {
        const int ix = N + t;
        if (ix >= x.nx - N) return;
        r = (size_t)is * P + ix;
        const int nk = side == 0 ? is - N : x.nz - N - is;
        for (int k = 1; k <= nk; ++k) {
            const size_t img = side == 0 ? r - (size_t)k * P : r + (size_t)k * P;
            const size_t src = side == 0 ? r + (size_t)k * P : r - (size_t)k * P;
            x.bzx[f0 + src] -= x.bzx[f0 + img];
            x.bzx[f0 + img] = 0.f;
            x.bzz[f0 + src] -= x.bzz[f0 + img];
            x.bzz[f0 + img] = 0.f;
        }
    } else {
        const int iz = N + t;
        if (iz >= x.nz - N) return;
        r = (size_t)iz * P + is;
        const int nk = side == 2 ? is - N : x.nx - N - is;
        for (int k = 1; k <= nk; ++k) {
            const size_t img = side == 2 ? r - k : r + k, src = side == 2 ? r + k : r - k;
            x.bzx[f0 + src] -= x.bzx[f0 + img];
            x.bzx[f0 + img] = 0.f;
            x.bxx[f0 + src] -= x.bxx[f0 + img];
            x.bxx[f0 + img] = 0.f;
        }
    }
    const double l = x.lam[r], u = x.mu[r], den = l + 2.0 * u;
    const double f = (l * u + u * u) / den, f_l = u * u / (den * den),
                 f_u = (l * l + 2.0 * l * u + 2.0 * u * u) / (den * den);
    const float *h = x.hist + ((size_t)(n - x.hist_base) * x.B + s) * 5 * x.fs;
    const double a = side < 2 ? x.bxx[f0 + r] : x.bzz[f0 + r];
    const double e = side < 2 ? h[1 * x.fs + r] : h[0 * x.fs + r];
    float *Ds = side < 2 ? x.Dv[3] : x.Dv[0];
    Ds[f0 + r] += (float)(a * 4.0 * x.dt * f);
    double *G = x.G + (size_t)s * 5 * x.fs;
    G[0 * x.fs + r] += a * 4.0 * x.dt * e * f_l;
    G[1 * x.fs + r] += a * 4.0 * x.dt * e * f_u;
    x.bzz[f0 + r] = 0.f;
    x.bzx[f0 + r] = 0.f;
    x.bxx[f0 + r] = 0.f;
}

// stress adjoints outside the interior: the forward never reads them again (0)
__global__ void xa_zero_outside(const XArgs x) {
    int s, iz, ix;
    if (!cell_of(x, (size_t)blockIdx.x * blockDim.x + threadIdx.x, s, iz, ix) || interior(x, iz, ix)) return;
    const size_t o = (size_t)s * x.fs + (size_t)iz * x.P + ix;
    x.bzz[o] = x.bzx[o] = x.bxx[o] = 0.f;
}

// stress update reversed: media gradient and the strain-rate adjoints, then the stress
// pass's CPML reversed
__global__ void xa_stress(const XArgs x, int n) {
    int s, iz, ix;
    if (!cell_of(x, (size_t)blockIdx.x * blockDim.x + threadIdx.x, s, iz, ix) || !interior(x, iz, ix)) return;
    const size_t c = (size_t)iz * x.P + ix, o = (size_t)s * x.fs + c;
    const float *h = x.hist + ((size_t)(n - x.hist_base) * x.B + s) * 5 * x.fs;
    const float zz = x.bzz[o], zx = x.bzx[o], xx = x.bxx[o];
    const float e_zz = h[c], e_xx = h[x.fs + c], e_zx = h[2 * x.fs + c];
    const float l = x.lam[c], u = x.mu[c], mzx = x.muzx[c], dt = x.dt;
    double *G = x.G + (size_t)s * 5 * x.fs;
    G[0 * x.fs + c] += (double)dt * ((double)zz + xx) * ((double)e_xx + e_zz);
    G[1 * x.fs + c] += 2.0 * dt * ((double)zz * e_zz + (double)xx * e_xx);
    G[2 * x.fs + c] += (double)dt * zx * e_zx;
    float d0 = x.Dv[0][o] + dt * ((l + 2.f * u) * zz + l * xx);
    float d3 = x.Dv[3][o] + dt * (l * zz + (l + 2.f * u) * xx);
    float d1 = dt * mzx * zx, d2 = d1;
    if (x.pml_z) {
        const CpmlPoint p = x.zp[iz];
        float q = x.cv[0][o] + d0;
        d0 = d0 * p.ik + p.a * q;
        x.cv[0][o] = p.b * q;
        q = x.cv[1][o] + d1;
        d1 = d1 * p.ikh + p.ah * q;
        x.cv[1][o] = p.bh * q;
    }
    if (x.pml_x) {
        const CpmlPoint p = x.xp[ix];
        float q = x.cv[3][o] + d3;
        d3 = d3 * p.ik + p.a * q;
        x.cv[3][o] = p.b * q;
        q = x.cv[2][o] + d2;
        d2 = d2 * p.ikh + p.ah * q;
        x.cv[2][o] = p.bh * q;
    }
    x.Dv[0][o] = d0;
    x.Dv[1][o] = d1;
    x.Dv[2][o] = d2;
    x.Dv[3][o] = d3;
}

// velocity-derivative transposes gathered onto every cell; 0 outside the interior
__global__ void xa_velocity_gather(const XArgs x) {
    int s, iz, ix;
    if (!cell_of(x, (size_t)blockIdx.x * blockDim.x + threadIdx.x, s, iz, ix)) return;
    const size_t so = (size_t)s * x.fs, o = so + (size_t)iz * x.P + ix;
    if (!interior(x, iz, ix)) {
        x.bvz[o] = x.bvx[o] = 0.f;
        return;
    }
    const float *D0 = x.Dv[0] + so, *D1 = x.Dv[1] + so, *D2 = x.Dv[2] + so, *D3 = x.Dv[3] + so;
    float vz = 0.f, vx = 0.f;
    for (int k = 1; k <= x.N; ++k) {
        vz += x.cz[k - 1] * (at(D0, x, iz - k + 1, ix) - at(D0, x, iz + k, ix));
        vz += x.cx[k - 1] * (at(D2, x, iz, ix - k) - at(D2, x, iz, ix + k - 1));
        vx += x.cz[k - 1] * (at(D1, x, iz - k, ix) - at(D1, x, iz + k - 1, ix));
        vx += x.cx[k - 1] * (at(D3, x, iz, ix - k + 1) - at(D3, x, iz, ix + k));
    }
    x.bvz[o] += vz;
    x.bvx[o] += vx;
}

// the receivers' adjoint sources: A[r][t] = rc sum_{t' >= t} dt adj[r][t'] (the running
// sum's transpose; the loss is sum 0.5 dt adj^2, so its cotangent is dt adj)
__global__ void xk_revcumsum(const double *adj_z, const double *adj_x, double *A_z, double *A_x, int rows, int nt,
                             double dt, double rc) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double z = 0.0, xs = 0.0;
    for (int t = nt - 1; t >= 0; --t) {
        z += dt * adj_z[(size_t)r * nt + t];
        xs += dt * adj_x[(size_t)r * nt + t];
        A_z[(size_t)r * nt + t] = rc * z;
        A_x[(size_t)r * nt + t] = rc * xs;
    }
}

// the shots' media gradients summed and carried to lam, mu, rho through mat_av2's
// averages (harmonic mu_zx, mean rho), FP64 from the FP64 model
__global__ void xk_media_chain(const double *G, int nb, size_t fs, int P, const double *mu, const double *rho,
                               int nz, int nx, int z1, int z2, int x1, int x2, double *d_lam, double *d_mu,
                               double *d_rho) {
    const size_t n = (size_t)nz * nx;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(o / nx), ix = (int)(o - (size_t)iz * nx);
        if (iz < z1 || iz > z2 || ix < x1 || ix > x2) {
            d_lam[o] = d_mu[o] = d_rho[o] = 0.0;
            continue;
        }
        auto g = [&](int k, int z, int x_) {
            double v = 0.0;
            for (int s = 0; s < nb; ++s) v += G[((size_t)s * 5 + k) * fs + (size_t)z * P + x_];
            return v;
        };
        const double u = mu[o];
        double dm = g(1, iz, ix), dr = 0.0;
        for (int a = 0; a <= 1; ++a)
            for (int b = 0; b <= 1; ++b) {
                const int pz = iz - a, px = ix - b;
                if (pz < 0 || px < 0 || pz >= nz - 1 || px >= nx - 1 || u == 0.0) continue;
                const double m00 = mu[(size_t)pz * nx + px], m01 = mu[(size_t)pz * nx + px + 1];
                const double m10 = mu[(size_t)(pz + 1) * nx + px], m11 = mu[(size_t)(pz + 1) * nx + px + 1];
                if (m00 == 0.0 || m01 == 0.0 || m10 == 0.0 || m11 == 0.0) continue;
                const double q = 4.0 / (1.0 / m00 + 1.0 / m01 + 1.0 / m10 + 1.0 / m11);
                dm += g(2, pz, px) * q * q / (4.0 * u * u);
            }
        for (int a = 0; a <= 1; ++a) {
            const int pz = iz - a;
            if (pz >= 0 && pz < nz - 1 && ix < nx - 1) {
                const double q = 1.0 / (0.5 * (rho[(size_t)pz * nx + ix] + rho[(size_t)(pz + 1) * nx + ix]));
                dr -= 0.5 * g(3, pz, ix) * q * q;
            }
            const int px = ix - a;
            if (px >= 0 && px < nx - 1 && iz < nz - 1) {
                const double q = 1.0 / (0.5 * (rho[(size_t)iz * nx + px] + rho[(size_t)iz * nx + px + 1]));
                dr -= 0.5 * g(4, iz, px) * q * q;
            }
        }
        d_lam[o] = g(0, iz, ix);
        d_mu[o] = dm;
        d_rho[o] = dr;
    }
}

unsigned cells_grid(const XArgs &x) { return (unsigned)(((size_t)x.B * x.nz * x.nx + kXB - 1) / kXB); }

}  // namespace

void x_forward_step(const XArgs &x, int n, const int isurf[4], hipStream_t st) {
    hipLaunchKernelGGL(xk_stress, dim3(cells_grid(x)), dim3(kXB), 0, st, x, n);
    for (int side = 0; side < 4; ++side) {
        if (isurf[side] <= 0) continue;
        const int len = side < 2 ? x.nx - 2 * x.N : x.nz - 2 * x.N;
        hipLaunchKernelGGL(xk_surface, dim3((len + 127) / 128, x.B), dim3(128), 0, st, x, side, isurf[side]);
    }
    hipLaunchKernelGGL(xk_velocity, dim3(cells_grid(x)), dim3(kXB), 0, st, x, n);
}

void x_adjoint_step(const XArgs &x, int n, const int isurf[4], const SrcArgs *type0, hipStream_t st) {
    if (type0 && type0->ngroups)
        hipLaunchKernelGGL(xa_source_rho, dim3((type0->ngroups + 63) / 64), dim3(64), 0, st, x, *type0, n);
    hipLaunchKernelGGL(xa_velocity, dim3(cells_grid(x)), dim3(kXB), 0, st, x, n);
    hipLaunchKernelGGL(xa_stress_gather, dim3(cells_grid(x)), dim3(kXB), 0, st, x);
    for (int k = 0; k < 4; ++k) (void)hipMemsetAsync(x.Dv[k], 0, sizeof(float) * x.B * x.fs, st);
    for (int side = 3; side >= 0; --side) {
        if (isurf[side] <= 0) continue;
        const int len = side < 2 ? x.nx - 2 * x.N : x.nz - 2 * x.N;
        hipLaunchKernelGGL(xa_surface, dim3((len + 127) / 128, x.B), dim3(128), 0, st, x, side, isurf[side], n);
    }
    hipLaunchKernelGGL(xa_zero_outside, dim3(cells_grid(x)), dim3(kXB), 0, st, x);
    hipLaunchKernelGGL(xa_stress, dim3(cells_grid(x)), dim3(kXB), 0, st, x, n);
    hipLaunchKernelGGL(xa_velocity_gather, dim3(cells_grid(x)), dim3(kXB), 0, st, x);
}

void x_revcumsum(const double *adj_z, const double *adj_x, double *A_z, double *A_x, int rows, int nt, double dt,
                 double rc, hipStream_t st) {
    if (rows > 0)
        hipLaunchKernelGGL(xk_revcumsum, dim3((rows + 63) / 64), dim3(64), 0, st, adj_z, adj_x, A_z, A_x, rows, nt,
                           dt, rc);
}

void x_media_chain(const XArgs &x, int nb, const double *mu, const double *rho, int z1, int z2, int x1, int x2,
                   double *d_lam, double *d_mu, double *d_rho, hipStream_t st) {
    const size_t n = (size_t)x.nz * x.nx;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(xk_media_chain, dim3(g), dim3(256), 0, st, x.G, nb, x.fs, x.P, mu, rho, x.nz, x.nx, z1, z2,
                       x1, x2, d_lam, d_mu, d_rho);
}

}  // namespace psv
