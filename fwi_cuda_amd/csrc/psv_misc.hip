// psv_misc.hip -- per-model and per-shot kernels around the time loop.
//
// These run once per shot or per model, not per time step, so they favour
// exactness (FP64, deterministic reduction order) over speed:
//   mat_av2          src/cpu/n_kernel_lib_PSV.cpp:853-922
//   adjsrc2          :579-626
//   energy_weights2  :696-757    interpol_grad2 :628-693    scale_grad_E2 :760-783
//   taper2           :942-1010   update_mat2    :785-832
#include "psv_internal.h"

namespace psv {

namespace {

constexpr int kRB = 256;   // reduction block
constexpr double kPiRef = 3.14159265;   // the reference's truncated pi (include/h_globvar.hpp:13)

template <typename Op>
__device__ double block_reduce(double v, Op op, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] = op(sh[t], sh[t + w]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

struct OpSum { __device__ double operator()(double a, double b) const { return a + b; } };
struct OpMax { __device__ double operator()(double a, double b) const { return a > b ? a : b; } };

// mat_av2: harmonic mu at (i+1/2, j+1/2), inverse mean rho at (i+1/2, j) and
// (i, j+1/2); the rho_xp zero guard writes rho_zp as the reference does.
__global__ void k_mat_av(const double *lam, const double *mu, const double *rho, float *flam, float *fmu,
                         float *muzx, float *rhoz, float *rhox, double *part, int nz, int nx, int P) {
    __shared__ double sh[kRB];
    const size_t n = (size_t)nz * nx;
    double sl = 0, sm = 0, sr = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(i / nx), ix = (int)(i - (size_t)iz * nx);
        const size_t o = (size_t)iz * P + ix;
        flam[o] = (float)lam[i];
        fmu[o] = (float)mu[i];
        if (iz < nz - 1 && ix < nx - 1) {
            const double m00 = mu[i], m01 = mu[i + 1], m10 = mu[i + nx], m11 = mu[i + nx + 1];
            double v = 4.0 / ((1.0 / m00) + (1.0 / m01) + (1.0 / m10) + (1.0 / m11));
            if (m00 == 0.0 || m01 == 0.0 || m10 == 0.0 || m11 == 0.0) v = 0.0;
            const double r00 = rho[i], r10 = rho[i + nx], r01 = rho[i + 1];
            double rz = 1.0 / (0.5 * (r00 + r10));
            const double rx = 1.0 / (0.5 * (r00 + r01));
            if (r00 < 1e-4 && r10 < 1e-4) rz = 0.0;
            if (r00 < 1e-4 && r01 < 1e-4) rz = 0.0;
            muzx[o] = (float)v;
            rhoz[o] = (float)rz;
            rhox[o] = (float)rx;
            sl += lam[i];
            sm += mu[i];
            sr += rho[i];
        } else {
            muzx[o] = 0.f;
            rhoz[o] = 0.f;
            rhox[o] = 0.f;
        }
    }
    sl = block_reduce(sl, OpSum(), sh);
    sm = block_reduce(sm, OpSum(), sh);
    sr = block_reduce(sr, OpSum(), sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = sl;
        part[3 * blockIdx.x + 1] = sm;
        part[3 * blockIdx.x + 2] = sr;
    }
}

__global__ void k_sum_parts3(const double *part, int nparts, double *out, double ncell) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0, b = 0, c = 0;
    for (int i = 0; i < nparts; ++i) {
        a += part[3 * i];
        b += part[3 * i + 1];
        c += part[3 * i + 2];
    }
    out[0] = a / ncell;
    out[1] = b / ncell;
    out[2] = c / ncell;
}

// adjsrc2 for a batch: block s handles shot s's receivers.
__global__ void k_residual(const double *seis_z, const double *seis_x, const double *obs_z, const double *obs_x,
                           double *res_z, double *res_x, const int *rec_off, int nt, double dt, double *misfit) {
    __shared__ double sh[kRB];
    const int s = blockIdx.x;
    const size_t b = (size_t)rec_off[s] * nt, e = (size_t)rec_off[s + 1] * nt;
    double acc = 0;
    for (size_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const double rz = seis_z[i] - obs_z[i], rx = seis_x[i] - obs_x[i];
        res_z[i] = rz;
        res_x[i] = rx;
        acc += 0.5 * dt * (rz * rz) + 0.5 * dt * (rx * rx);
    }
    acc = block_reduce(acc, OpSum(), sh);
    if (threadIdx.x == 0) misfit[s] = acc;
}

// per-shot max(|a|, |b|) over the shot's receiver rows (block s = shot s)
__global__ void k_shot_absmax(const double *a, const double *b, const int *rec_off, int nt, double *out) {
    __shared__ double sh[kRB];
    const int s = blockIdx.x;
    const size_t lo = (size_t)rec_off[s] * nt, hi = (size_t)rec_off[s + 1] * nt;
    double m = 0;
    for (size_t i = lo + threadIdx.x; i < hi; i += blockDim.x) m = fmax(m, fmax(fabs(a[i]), fabs(b[i])));
    m = block_reduce(m, OpMax(), sh);
    if (threadIdx.x == 0) out[s] = m;
}

__global__ void k_absmax(const double *x, size_t n, double *part) {
    __shared__ double sh[kRB];
    double m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[i]));
    m = block_reduce(m, OpMax(), sh);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void k_max_parts(const double *part, int n, double *out) {
    __shared__ double sh[kRB];
    double m = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, part[i]);
    m = block_reduce(m, OpMax(), sh);
    if (threadIdx.x == 0) *out = m;
}

// energy_weights2, first pass: We = sqrt(We_fwd * We_adj) on [z1,z2) x [x1,x2), block maxima.
__global__ void k_energy(const float *We, const float *Wa, double wa_scale, double *Wc, double *part, int nx, int P,
                         int z1, int z2, int x1, int x2) {
    __shared__ double sh[kRB];
    const int bw = x2 - x1;
    const size_t n = (size_t)(z2 - z1) * bw;
    double m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int iz = z1 + (int)(i / bw), ix = x1 + (int)(i % bw);
        const size_t f = (size_t)iz * P + ix;
        const double w = sqrt((double)We[f] * ((double)Wa[f] * wa_scale));
        Wc[(size_t)iz * nx + ix] = w;
        m = fmax(m, w);
    }
    m = block_reduce(m, OpMax(), sh);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void k_energy_add(double *Wc, const double *wmax, int nx, int z1, int z2, int x1, int x2) {
    const int bw = x2 - x1;
    const size_t n = (size_t)(z2 - z1) * bw;
    const double add = 0.005 * (*wmax);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int iz = z1 + (int)(i / bw), ix = x1 + (int)(i % bw);
        Wc[(size_t)iz * nx + ix] += add;
    }
}

// interpol_grad2 phase 1: snap points onto a zeroed full grid
__global__ void k_interp_points(const double *g, double g_scale, double *T, int nx, int z1, int x1, int sdz,
                                int sdx, int snz, int snx) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= snz * snx) return;
    const int jz = q / snx, jx = q - jz * snx;
    T[(size_t)(z1 + jz * sdz) * nx + x1 + jx * sdx] = g[q] * g_scale;
}

// phase 2: along x on snap rows iz < z2 (one thread per row segment)
__global__ void k_interp_x(double *T, int nx, int z1, int z2, int x1, int x2, int sdz, int sdx) {
    const int nrow = (z2 - z1 + sdz - 1) / sdz, nseg = (x2 - x1 + sdx - 1) / sdx;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nrow * nseg) return;
    const int iz = z1 + (q / nseg) * sdz, ix = x1 + (q % nseg) * sdx;
    double *row = T + (size_t)iz * nx;
    const double t = (row[ix + sdx] - row[ix]) / sdx;
    for (int k = 1; k < sdx; ++k) row[ix + k] = row[ix] + t * k;
}

// phase 3: along z for every column x1 <= ix < x2 (one thread per segment column)
__global__ void k_interp_z(double *T, int nx, int z1, int z2, int x1, int x2, int sdz) {
    const int nseg = (z2 - z1 + sdz - 1) / sdz, bw = x2 - x1;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nseg * bw) return;
    const int iz = z1 + (q / bw) * sdz, ix = x1 + q % bw;
    const double t = (T[(size_t)(iz + sdz) * nx + ix] - T[(size_t)iz * nx + ix]) / sdz;
    for (int k = 1; k < sdz; ++k) T[(size_t)(iz + k) * nx + ix] = T[(size_t)iz * nx + ix] + t * k;
}

// scale_grad_E2: G += T / (We C^2) on [z1,z2) x [x1,x2)
__global__ void k_scale(double *G, const double *T, const double *Wc, double Cm, int nx, int z1, int z2, int x1,
                        int x2) {
    const int bw = x2 - x1;
    const size_t n = (size_t)(z2 - z1) * bw;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t o = (size_t)(z1 + (int)(i / bw)) * nx + x1 + (int)(i % bw);
        G[o] += T[o] / (Wc[o] * Cm * Cm);
    }
}

// taper2: x factor then z factor, per cell
__global__ void k_taper(double *A, int nz, int nx, int z1, int z2, int x1, int x2, int t1, int t2, int b1, int b2,
                        int l1, int l2, int r1, int r2) {
    const size_t n = (size_t)nz * nx;
    const int tl = l2 - l1, tr = r1 - r2, tt = t2 - t1, tb = b1 - b2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(i / nx), ix = (int)(i - (size_t)iz * nx);
        double a = A[i];
        if (ix >= x1 && ix < l1) a *= 0.0;
        else if (ix >= l1 && ix < l2) a *= 0.5 * (1.0 - cos(kPiRef * (ix - l1) / tl));
        else if (ix > r2 && ix < r1) a *= 0.5 * (1.0 - cos(kPiRef * (r1 - ix) / tr));
        else if (ix >= r1 && ix <= x2) a *= 0.0;
        if (iz >= z1 && iz < t1) a *= 0.0;
        else if (iz >= t1 && iz < t2) a *= 0.5 * (1.0 - cos(kPiRef * (iz - t1) / tt));
        else if (iz > b2 && iz < b1) a *= 0.5 * (1.0 - cos(kPiRef * (b1 - iz) / tb));
        else if (iz >= b1 && iz <= z2) a *= 0.0;
        A[i] = a;
    }
}

__global__ void k_update(double *mat, const double *old, const double *G, double f, double mmax, double mmin,
                         size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double v = old[i] + f * G[i];
        if (v > mmax) v = mmax;
        if (v < mmin) v = mmin;
        mat[i] = v;
    }
}

// cp = sqrt((lam + 2 mu) / rho), cs = sqrt(mu / rho)
__global__ void k_vel_of(const double *lam, const double *mu, const double *rho, double *cp, double *cs, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        cp[i] = sqrt((lam[i] + 2.0 * mu[i]) / rho[i]);
        cs[i] = sqrt(mu[i] / rho[i]);
    }
}

// chain rule through lam = rho (cp^2 - 2 cs^2), mu = rho cs^2 at fixed rho
__global__ void k_vel_grad(const double *Gl, const double *Gm, const double *rho, const double *cp, const double *cs,
                           double *gcp, double *gcs, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        gcp[i] = 2.0 * rho[i] * cp[i] * Gl[i];
        gcs[i] = 2.0 * rho[i] * cs[i] * (Gm[i] - 2.0 * Gl[i]);
    }
}

// update_mat2's clipped step on cp and cs, then lam and mu at the saved rho
__global__ void k_vel_update(double *lam, double *mu, double *rho, const double *rho_old, const double *cp_old,
                             const double *cs_old, const double *gcp, const double *gcs, double fcp, double fcs,
                             double cp_lo, double cp_hi, double cs_lo, double cs_hi, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double cp = cp_old[i] + fcp * gcp[i];
        if (cp > cp_hi) cp = cp_hi;
        if (cp < cp_lo) cp = cp_lo;
        double cs = cs_old[i] + fcs * gcs[i];
        if (cs > cs_hi) cs = cs_hi;
        if (cs < cs_lo) cs = cs_lo;
        const double r = rho_old[i];
        const double m = r * cs * cs;
        mu[i] = m;
        lam[i] = r * cp * cp - 2.0 * m;
        rho[i] = r;
    }
}

// One shot's five FP64 correlation totals (s1, s1 + s2, dt vz'.Vz, s3, dt vx'.Vx on the
// snapshot box) added into the sums over shots: raw[k] += g_k / (alpha * gain_k), the lam / mu
// terms carrying the gain.
__global__ void k_acc_raw64(double *raw, const double *g0, const double *g1, const double *g2, const double *g3,
                            const double *g4, size_t cells, const double *alpha, double gain) {
    const double *g[5] = {g0, g1, g2, g3, g4};
    const double al = *alpha;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += (size_t)gridDim.x * blockDim.x)
        for (int k = 0; k < 5; ++k) {
            const bool gained = k == 0 || k == 1 || k == 3;
            raw[k * cells + i] += g[k][i] / (al * (gained ? gain : 1.0));
        }
}

// FP32 correlation sums added into their FP64 totals and zeroed: g64[k][i] += g32[k][i]
__global__ void k_flush_sums(double *g64, float *g0, float *g1, float *g2, float *g3, float *g4, int nk, size_t n,
                             size_t cells) {
    float *g[5] = {g0, g1, g2, g3, g4};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        for (int k = 0; k < nk; ++k) {
            g64[k * cells + i] += (double)g[k][i];
            g[k][i] = 0.f;
        }
}

// The same flush inside a replayed graph, taken only after the adjoint step whose time
// index t = *base + dir * k leaves a multiple of `every` steps behind it ((nt - t) % every == 0)
__global__ void k_flush_sums_at(double *g64, float *g0, float *g1, float *g2, float *g3, float *g4, int nk, size_t n,
                                size_t cells, const int *base, int dir, int k, int nt, int every) {
    const int t = *base + dir * k;
    if ((nt - t) % every != 0) return;
    float *g[5] = {g0, g1, g2, g3, g4};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        for (int q = 0; q < nk; ++q) {
            g64[q * cells + i] += (double)g[q][i];
            g[q][i] = 0.f;
        }
}

// The staggered chain rule from the five sums to dJ/d(lam, mu, rho) on [nz][nx] (zero outside
// the box): mu_zx at (i, j) is the harmonic mean of mu on (i..i+1, j..j+1) (mat_av2,
// src/cpu/n_kernel_lib_PSV.cpp:875-880), so the shear sum of each of the four mu_zx points a
// cell enters, times mu_zx^2 / (4 mu^2) with the sum's own 1 / mu^2 undone; rho_zp and
// rho_xp average rho over (i, i+1) and (j, j+1) (:886-887), each density gathering half of
// its two points of each kind.
__global__ void k_misfit_chain(const double *raw, const double *mu, double *d_lam, double *d_mu, double *d_rho,
                               int nz, int nx, int z1, int x1, int snz, int snx, double sg) {
    const size_t n = (size_t)nz * nx, cells = (size_t)snz * snx;
    for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(o / nx), ix = (int)(o - (size_t)iz * nx);
        const int jz = iz - z1, jx = ix - x1;
        if (jz < 0 || jz >= snz || jx < 0 || jx >= snx) {
            d_lam[o] = 0.0;
            d_mu[o] = 0.0;
            d_rho[o] = 0.0;
            continue;
        }
        auto R = [&](int k, int qz, int qx) { return (qz < 0 || qx < 0) ? 0.0 : raw[k * cells + (size_t)qz * snx + qx]; };
        d_lam[o] = sg * R(0, jz, jx);
        double shear = 0.0;
        for (int dz = 0; dz <= 1; ++dz)
            for (int dx = 0; dx <= 1; ++dx) {
                if (jz - dz < 0 || jx - dx < 0) continue;
                const double mp = mu[(size_t)(iz - dz) * nx + (ix - dx)];
                shear += R(3, jz - dz, jx - dx) * mp * mp;
            }
        const double mk = mu[o];
        d_mu[o] = sg * (R(1, jz, jx) + (mk != 0.0 ? shear / (4.0 * mk * mk) : 0.0));
        d_rho[o] = sg * 0.5 * (R(2, jz, jx) + R(2, jz - 1, jx) + R(4, jz, jx) + R(4, jz, jx - 1));
    }
}

// a copy between device arrays that may come from the caller (no runtime copy engine involved)
__global__ void k_copy_d(double *dst, const double *src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// apply_gauss_filter (src/cpu/n_kernel_lib_PSV.cpp:1051-1084): the (2 hfs + 1)^2 filter over
// cells [hfs, nz - hfs - 1) x [hfs, nx - hfs - 1) of the copy src into dst, other cells copied;
// accumulated in double (the reference's accumulator is an int)
__global__ void k_gauss(const double *src, double *dst, const double *filt, int nz, int nx, int hfs) {
    const size_t n = (size_t)nz * nx;
    const int w = 2 * hfs + 1;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int iz = (int)(i / nx), ix = (int)(i - (size_t)iz * nx);
        if (iz < hfs || iz >= nz - hfs - 1 || ix < hfs || ix >= nx - hfs - 1) {
            dst[i] = src[i];
            continue;
        }
        double conv = 0.0;
        for (int jj = 0; jj < w; ++jj)
            for (int ii = 0; ii < w; ++ii)
                conv += src[(size_t)(iz + jj - hfs) * nx + ix + ii - hfs] * filt[jj * w + ii];
        dst[i] = conv;
    }
}

// block partials of sum g^2 and sum old^2 (PCG_PSV's Fletcher-Reeves beta), then old = -g
__global__ void k_pcg_sums(const double *g, double *old, size_t n, double *part) {
    __shared__ double sh[kRB];
    double a = 0, b = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        a += g[i] * g[i];
        b += old[i] * old[i];
        old[i] = -g[i];
    }
    a = block_reduce(a, OpSum(), sh);
    b = block_reduce(b, OpSum(), sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = a;
        part[2 * blockIdx.x + 1] = b;
    }
}

// beta = sum g^2 / sum old^2 from the partials, clipped at 0 (0 when old was 0: the first call)
__global__ void k_pcg_beta(const double *part, int nparts, double *beta) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0, b = 0;
    for (int i = 0; i < nparts; ++i) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    const double v = b > 0.0 ? a / b : 0.0;
    *beta = v > 0 ? v : 0.0;
}

// dir = -g + beta dir; g = dir
__global__ void k_pcg_dir(double *g, double *dir, const double *beta, size_t n) {
    const double bt = *beta;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double d = -g[i] + bt * dir[i];
        dir[i] = d;
        g[i] = d;
    }
}

__global__ void k_nonfinite(const float *x, size_t n, int *out) {
    int c = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        c += !isfinite(x[i]);
    if (c) atomicAdd(out, c);
}

__global__ void k_fill(float *p, float v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

inline int nblocks(size_t n, int cap = 1024) {
    size_t b = (n + kRB - 1) / kRB;
    return (int)(b < (size_t)cap ? (b ? b : 1) : cap);
}

}  // namespace

int mat_av_parts(int nz, int nx) { return nblocks((size_t)nz * nx, kMatAvMaxParts); }

void dev_mat_av_into(const double *lam, const double *mu, const double *rho, float *flam, float *fmu, float *muzx,
                     float *rhoz, float *rhox, double *part, double *means, int nz, int nx, int P, hipStream_t st) {
    const int nb = mat_av_parts(nz, nx);
    hipLaunchKernelGGL(k_mat_av, dim3(nb), dim3(kRB), 0, st, lam, mu, rho, flam, fmu, muzx, rhoz, rhox, part, nz,
                       nx, P);
    hipLaunchKernelGGL(k_sum_parts3, dim3(1), dim3(1), 0, st, part, nb, means, (double)((nz - 1) * (nx - 1)));
}

void dev_mat_av(const double *lam, const double *mu, const double *rho, float *flam, float *fmu, float *muzx,
                float *rhoz, float *rhox, double *sums3, int nz, int nx, int P, hipStream_t st) {
    // partial sums live right after the three means
    dev_mat_av_into(lam, mu, rho, flam, fmu, muzx, rhoz, rhox, sums3 + 4, sums3, nz, nx, P, st);
}

void dev_residual(const double *seis_z, const double *seis_x, const double *obs_z, const double *obs_x,
                  double *res_z, double *res_x, const int *rec_off_local, int nshots, int nt, double dt,
                  double *misfit, hipStream_t st) {
    hipLaunchKernelGGL(k_residual, dim3(nshots), dim3(kRB), 0, st, seis_z, seis_x, obs_z, obs_x, res_z, res_x,
                       rec_off_local, nt, dt, misfit);
}

void dev_absmax(const double *x, size_t n, double *out, hipStream_t st) {
    const int nb = nblocks(n, 512);
    double *part = out + 1;
    hipLaunchKernelGGL(k_absmax, dim3(nb), dim3(kRB), 0, st, x, n, part);
    hipLaunchKernelGGL(k_max_parts, dim3(1), dim3(kRB), 0, st, part, nb, out);
}

void dev_shot_absmax(const double *a, const double *b, const int *rec_off_local, int nshots, int nt, double *out,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_shot_absmax, dim3(nshots), dim3(kRB), 0, st, a, b, rec_off_local, nt, out);
}

void dev_precondition(const float *We, const float *Wa, double wa_scale, const double *const g[3],
                      const double g_scale[3], double *const G[3], const double C[3], double *tmp, double *wmax, int nz, int nx, int P,
                      int z1, int z2, int x1, int x2, int sdz, int sdx, int snz, int snx, hipStream_t st) {
    double *Wc = tmp, *T = tmp + (size_t)nz * nx;
    const size_t nbox = (size_t)(z2 - z1) * (x2 - x1);
    const int nb = nblocks(nbox, 512);
    double *part = wmax + 1;
    hipLaunchKernelGGL(k_energy, dim3(nb), dim3(kRB), 0, st, We, Wa, wa_scale, Wc, part, nx, P, z1, z2, x1, x2);
    hipLaunchKernelGGL(k_max_parts, dim3(1), dim3(kRB), 0, st, part, nb, wmax);
    hipLaunchKernelGGL(k_energy_add, dim3(nb), dim3(kRB), 0, st, Wc, wmax, nx, z1, z2, x1, x2);
    for (int p = 0; p < 3; ++p) {
        if (!(C[p] > 0)) continue;
        (void)hipMemsetAsync(T, 0, sizeof(double) * (size_t)nz * nx, st);
        hipLaunchKernelGGL(k_interp_points, dim3((snz * snx + 255) / 256), dim3(256), 0, st, g[p], g_scale[p], T, nx,
                           z1, x1, sdz, sdx, snz, snx);
        if (sdx > 1) {
            const int n = ((z2 - z1 + sdz - 1) / sdz) * ((x2 - x1 + sdx - 1) / sdx);
            hipLaunchKernelGGL(k_interp_x, dim3((n + 255) / 256), dim3(256), 0, st, T, nx, z1, z2, x1, x2, sdz, sdx);
        }
        if (sdz > 1) {
            const int n = ((z2 - z1 + sdz - 1) / sdz) * (x2 - x1);
            hipLaunchKernelGGL(k_interp_z, dim3((n + 255) / 256), dim3(256), 0, st, T, nx, z1, z2, x1, x2, sdz);
        }
        hipLaunchKernelGGL(k_scale, dim3(nb), dim3(kRB), 0, st, G[p], T, Wc, C[p], nx, z1, z2, x1, x2);
    }
}

void dev_taper(double *A, int nz, int nx, int z1, int z2, int x1, int x2, const int taper[8], hipStream_t st) {
    hipLaunchKernelGGL(k_taper, dim3(nblocks((size_t)nz * nx)), dim3(kRB), 0, st, A, nz, nx, z1, z2, x1, x2, taper[0],
                       taper[1], taper[2], taper[3], taper[4], taper[5], taper[6], taper[7]);
}

void dev_update(double *mat, const double *old, const double *G, double factor, double mat_max, double mat_min,
                size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_update, dim3(nblocks(n)), dim3(kRB), 0, st, mat, old, G, factor, mat_max, mat_min, n);
}

void dev_velocity_of(const double *lam, const double *mu, const double *rho, double *cp, double *cs, size_t n,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_vel_of, dim3(nblocks(n)), dim3(kRB), 0, st, lam, mu, rho, cp, cs, n);
}

void dev_velocity_gradient(const double *G_lam, const double *G_mu, const double *rho, const double *cp,
                           const double *cs, double *g_cp, double *g_cs, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_vel_grad, dim3(nblocks(n)), dim3(kRB), 0, st, G_lam, G_mu, rho, cp, cs, g_cp, g_cs, n);
}

void dev_velocity_update(double *lam, double *mu, double *rho, const double *rho_old, const double *cp_old,
                         const double *cs_old, const double *g_cp, const double *g_cs, double f_cp, double f_cs,
                         const double bounds[4], size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_vel_update, dim3(nblocks(n)), dim3(kRB), 0, st, lam, mu, rho, rho_old, cp_old, cs_old, g_cp,
                       g_cs, f_cp, f_cs, bounds[0], bounds[1], bounds[2], bounds[3], n);
}

void dev_acc_raw64(double *raw, const double *const g[5], size_t cells, const double *alpha, double gain,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_acc_raw64, dim3(nblocks(cells)), dim3(kRB), 0, st, raw, g[0], g[1], g[2], g[3], g[4], cells,
                       alpha, gain);
}

void dev_flush_sums(double *g64, float *const g32[5], int nk, size_t n, size_t cells, hipStream_t st) {
    hipLaunchKernelGGL(k_flush_sums, dim3(nblocks(n)), dim3(kRB), 0, st, g64, g32[0], g32[1], g32[2], g32[3], g32[4],
                       nk, n, cells);
}

void dev_flush_sums_at(double *g64, float *const g32[5], int nk, size_t n, size_t cells, const int *base, int dir,
                       int k, int nt, int every, hipStream_t st) {
    hipLaunchKernelGGL(k_flush_sums_at, dim3(nblocks(n)), dim3(kRB), 0, st, g64, g32[0], g32[1], g32[2], g32[3],
                       g32[4], nk, n, cells, base, dir, k, nt, every);
}

void dev_misfit_chain(const double *raw, const double *mu, double *d_lam, double *d_mu, double *d_rho, int nz,
                      int nx, int z1, int x1, int snz, int snx, double sg, hipStream_t st) {
    hipLaunchKernelGGL(k_misfit_chain, dim3(nblocks((size_t)nz * nx)), dim3(kRB), 0, st, raw, mu, d_lam, d_mu,
                       d_rho, nz, nx, z1, x1, snz, snx, sg);
}

void dev_copy(double *dst, const double *src, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_copy_d, dim3(nblocks(n)), dim3(kRB), 0, st, dst, src, n);
}

void dev_gauss(const double *src, double *dst, const double *filt, int nz, int nx, int hfs, hipStream_t st) {
    hipLaunchKernelGGL(k_gauss, dim3(nblocks((size_t)nz * nx, 4096)), dim3(kRB), 0, st, src, dst, filt, nz, nx, hfs);
}

void dev_pcg(double *g, double *dir, double *old, size_t n, double *scratch, hipStream_t st) {
    const int nb = nblocks(n, 512);
    double *beta = scratch, *part = scratch + 1;
    hipLaunchKernelGGL(k_pcg_sums, dim3(nb), dim3(kRB), 0, st, g, old, n, part);
    hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(1), 0, st, part, nb, beta);
    hipLaunchKernelGGL(k_pcg_dir, dim3(nblocks(n)), dim3(kRB), 0, st, g, dir, beta, n);
}

void dev_count_nonfinite(const float *x, size_t n, int *out, hipStream_t st) {
    hipLaunchKernelGGL(k_nonfinite, dim3(nblocks(n)), dim3(kRB), 0, st, x, n, out);
}

void dev_fill(float *p, float v, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kRB), 0, st, p, v, n);
}

}  // namespace psv
