/*
 * fwi_oracle.c -- FP64 CPU restatement of the reference's elastic P/SV path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; never shipped, never
 * measured as the product).  Each routine cites the reference function it
 * restates.  At fdorder = 2 the operation order of every floating-point
 * expression is the reference's, so results agree with the compiled reference
 * to the last bit except where the reference itself reduces in an
 * OpenMP-scheduled order (mat_av2's scalar means).
 *
 * Parity pinning (see tests/test_oracle_*.py):
 *   - golden vectors from the reference compiled in this container
 *     (tests/golden fixtures, made by tests/golden/make_golden.py);
 *   - closed forms: Taylor staggered coefficients, CPML profile formulas,
 *     wavelet, parabola vertex, per-cell gradient formulas;
 *   - invariants: zero source gives a zero field, polynomial exactness of the
 *     order-2N derivative, CPML identity outside the strips.
 */
#include "fwi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* The reference truncates pi (include/h_globvar.hpp:13, include/cpu/n_globvar.hpp:15);
 * the CPML alpha, the wavelet and the taper all use this value. */
#define OR_PI 3.14159265

#define AT(a, iz, ix) (a)[(size_t)(iz) * (size_t)nx + (size_t)(ix)]

/* ========================================================================== */
/* Host preprocessing                                                          */
/* ========================================================================== */

/* Holberg (1987) staggered-grid coefficients as tabulated for DENISE and used by
 * the reference (src/h_holbergcoeff.cpp:25-67): index [maxrelerror][order/2-1],
 * entries {points per wavelength, c1..c6}.  Row 0 is Taylor, rows 1-4 are the
 * 0.1 %, 0.5 %, 1 % and 3 % maximum group-velocity error designs. */
static const double holberg_table[5][6][7] = {
    {{23.0, 1.0, 0, 0, 0, 0, 0},
     {8.0, 9.0 / 8.0, -1.0 / 24.0, 0, 0, 0, 0},
     {6.0, 75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0, 0, 0, 0},
     {5.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0, 0, 0},
     {5.0, 19845.0 / 16384.0, -735.0 / 8192.0, 567.0 / 40960.0, -405.0 / 229376.0, 35.0 / 294912.0, 0},
     {4.0, 160083.0 / 131072.0, -12705.0 / 131072.0, 22869.0 / 1310720.0, -5445.0 / 1835008.0,
      847.0 / 2359296.0, -63.0 / 2883584.0}},
    {{49.7, 1.0010, 0, 0, 0, 0, 0},
     {8.32, 1.1382, -0.046414, 0, 0, 0, 0},
     {4.77, 1.1965, -0.078804, 0.0081781, 0, 0, 0},
     {3.69, 1.2257, -0.099537, 0.018063, -0.0026274, 0, 0},
     {3.19, 1.2415, -0.11231, 0.026191, -0.0064682, 0.001191, 0},
     {2.91, 1.2508, -0.12034, 0.032131, -0.010142, 0.0029857, -0.00066667}},
    {{22.2, 1.0050, 0, 0, 0, 0, 0},
     {5.65, 1.1534, -0.052806, 0, 0, 0, 0},
     {3.74, 1.2111, -0.088313, 0.011768, 0, 0, 0},
     {3.11, 1.2367, -0.10815, 0.023113, -0.0046905, 0, 0},
     {2.80, 1.2496, -0.11921, 0.031130, -0.0093272, 0.0025161, 0},
     {2.62, 1.2568, -0.12573, 0.036423, -0.013132, 0.0047484, -0.0015979}},
    {{15.8, 1.0100, 0, 0, 0, 0, 0},
     {4.80, 1.1640, -0.057991, 0, 0, 0, 0},
     {3.39, 1.2192, -0.094070, 0.014608, 0, 0, 0},
     {2.90, 1.2422, -0.11269, 0.026140, -0.0064054, 0, 0},
     {2.65, 1.2534, -0.12257, 0.033755, -0.011081, 0.0036784, 0},
     {2.51, 1.2596, -0.12825, 0.038550, -0.014763, 0.0058619, -0.0024538}},
    {{9.16, 1.0300, 0, 0, 0, 0, 0},
     {3.47, 1.1876, -0.072518, 0, 0, 0, 0},
     {2.91, 1.2341, -0.10569, 0.022589, 0, 0, 0},
     {2.61, 1.2516, -0.12085, 0.032236, -0.011459, 0, 0},
     {2.45, 1.2596, -0.12829, 0.038533, -0.014681, 0.0072580, 0},
     {2.36, 1.2640, -0.13239, 0.042217, -0.017803, 0.0081959, -0.0051848}},
};

/* holbergcoeff (src/h_holbergcoeff.cpp:21-80). */
void or_holbergcoeff(int32_t fdorder, int32_t maxrelerror, double hc[7])
{
    int nl = fdorder / 2 - 1;
    for (int i = 0; i < 7; i++) hc[i] = holberg_table[maxrelerror][nl][i];
}

/* Taylor staggered coefficients from the closed form (independent of the table):
 * with a_j = (2j-1)^2,  c_k = prod_{j != k} a_j / ((2k-1) prod_{j != k} (a_j - a_k)),
 * the solution of sum_k c_k (2k-1)^(2m+1) = delta_{m0}, m = 0..n-1. */
void or_taylorcoeff(int32_t n, double c[7])
{
    for (int i = 0; i < 7; i++) c[i] = 0.0;
    c[0] = 0.0;
    for (int k = 1; k <= n; k++) {
        double num = 1.0, den = (double)(2 * k - 1);
        double ak = (double)(2 * k - 1) * (2 * k - 1);
        for (int j = 1; j <= n; j++) {
            if (j == k) continue;
            double aj = (double)(2 * j - 1) * (2 * j - 1);
            num *= aj;
            den *= (aj - ak);
        }
        c[k] = num / den;
    }
}

/* cpml_PSV (src/h_preproc_PSV.cpp:112-263): CPML profiles d, K, alpha at the
 * integer and half-integer points of one axis, turned into the recursive
 * convolution coefficients b = exp(-(d/K + alpha) dt), a = d (b - 1) / (K (d + K alpha)). */
void or_cpml(double *a, double *b, double *K, double *a_half, double *b_half, double *K_half,
             double npower, double damp_v, double rcoef, double k_max, double freq,
             int32_t npml_h1, int32_t npml_h2, int32_t fpad, int32_t nh, double dt, double dh)
{
    const double alpha_max = 2.0 * OR_PI * (freq / 2.0);
    for (int ih = 0; ih < nh; ih++) {
        a[ih] = 0.0; a_half[ih] = 0.0;
        b[ih] = 1.0; b_half[ih] = 1.0;
        K[ih] = 1.0; K_half[ih] = 1.0;
    }
    for (int side = 0; side < 2; side++) {
        int npml = side == 0 ? npml_h1 : npml_h2;
        if (!npml) continue;
        double thick = npml * dh;
        double d0 = -(npower + 1) * damp_v * log(rcoef) / (2.0 * thick);
        int i0, i1;
        double origin;
        if (side == 0) { i0 = fpad; i1 = npml + fpad + 1; origin = thick; }
        else { i0 = nh - fpad - npml - 1; i1 = nh - fpad; origin = dh * (nh - fpad - npml - 1); }
        for (int i = i0; i < i1; i++) {
            double xn, d = 0, d_half = 0, alpha = 0.0, alpha_half = 0.0;
            double x = dh * (double)i;
            double absc = side == 0 ? origin - x : x - origin;
            if (absc >= 0) {
                xn = absc / thick;
                d = d0 * pow(xn, npower);
                K[i] = 1.0 + (k_max - 1.0) * pow(xn, npower);
                alpha = alpha_max * (1.0 - xn);
            }
            absc = side == 0 ? origin - (x + dh / 2.0) : (x + dh / 2.0) - origin;
            if (absc >= 0) {
                xn = absc / thick;
                d_half = d0 * pow(xn, npower);
                K_half[i] = 1.0 + (k_max - 1.0) * pow(xn, npower);
                alpha_half = alpha_max * (1.0 - xn);
            }
            if (alpha < 0.0) alpha = 0.0;
            if (alpha_half < 0.0) alpha_half = 0.0;
            b[i] = exp(-(d / K[i] + alpha) * dt);
            b_half[i] = exp(-(d_half / K_half[i] + alpha_half) * dt);
            if (fabs(d) > 1.0e-6) a[i] = d * (b[i] - 1.0) / (K[i] * (d + K[i] * alpha));
            if (fabs(d_half) > 1.0e-6)
                a_half[i] = d_half * (b_half[i] - 1.0) / (K_half[i] * (d_half + K_half[i] * alpha_half));
        }
    }
}

/* wavelet (src/h_preproc_PSV.cpp:267-286): the reference's Ricker-like pulse
 * amp (1 - 2 tau^2) exp(-2 tau^2), tau = pi (t - 1.5/fc - ts) / (1.5/fc). */
void or_wavelet(double *signal, int32_t nt, double dt, double amp, double fc, double ts, int32_t shape)
{
    const double fci = 1.0 / fc;
    if (shape != 1) return;
    for (int it = 0; it < nt; it++) {
        double t = it * dt;
        double tau = OR_PI * (t - 1.5 * fci - ts) / (1.5 * fci);
        signal[it] = amp * (1.0 - 2.0 * tau * tau) * exp(-2.0 * tau * tau);
    }
}

/* checkfd_ssg_elastic (src/h_checkfd_ssg_elastic.cpp:14-144), stability part:
 * dtstab = dh / (sqrt(2) * sum|hc_k| * cmax).  cmax and cmin scan rows 1..nz-fw
 * and columns 1+fw..nx-fw (clamped to the grid) over cells with mu != 0. */
double or_checkfd_dtstab(int32_t nz, int32_t nx, double dh, int32_t fw,
                         const double *rho, const double *lam, const double *mu, const double *hc,
                         double *cmax_out, double *cmin_out)
{
    double cmax = 0.0, cmin = 1e9;
    int i1 = 1 + fw, i2 = nx - fw; if (i2 > nx - 1) i2 = nx - 1;
    int j1 = 1, j2 = nz - fw; if (j2 > nz - 1) j2 = nz - 1;
    for (int i = i1; i <= i2; i++)
        for (int j = j1; j <= j2; j++) {
            double m = mu[(size_t)j * nx + i];
            if (fabs(m) < 1.0e-9) continue;
            double r = rho[(size_t)j * nx + i];
            double cs = sqrt(m / r), cp = sqrt((lam[(size_t)j * nx + i] + 2.0 * m) / r);
            if (cp > cmax) cmax = cp;
            if (cs > cmax) cmax = cs;
            if (cs < cmin) cmin = cs;
            if (cp < cmin) cmin = cp;
        }
    double gamma = fabs(hc[1]) + fabs(hc[2]) + fabs(hc[3]) + fabs(hc[4]) + fabs(hc[5]) + fabs(hc[6]);
    if (cmax_out) *cmax_out = cmax;
    if (cmin_out) *cmin_out = cmin;
    return dh / (sqrt(2) * gamma * cmax);
}

/* ========================================================================== */
/* Medium                                                                      */
/* ========================================================================== */

/* mat_av2 (src/cpu/n_kernel_lib_PSV.cpp:853-922): harmonic 4-point mu at
 * (i+1/2, j+1/2), inverse arithmetic 2-point rho at (i+1/2, j) and (i, j+1/2),
 * scalar means over (nz-1)(nx-1) cells.  The zero-density guard for rho_xp
 * writes rho_zp, as the reference does (:894-897).  The last row and column
 * are not written by the reference; here they are zeroed. */
void or_mat_av2(const double *lam, const double *mu, const double *rho,
                double *mu_zx, double *rho_zp, double *rho_xp, double C[3], int32_t nz, int32_t nx)
{
    double cl = 0.0, cm = 0.0, cr = 0.0;
    for (int iz = 0; iz < nz; iz++) {
        AT(mu_zx, iz, nx - 1) = 0.0; AT(rho_zp, iz, nx - 1) = 0.0; AT(rho_xp, iz, nx - 1) = 0.0;
    }
    for (int ix = 0; ix < nx; ix++) {
        AT(mu_zx, nz - 1, ix) = 0.0; AT(rho_zp, nz - 1, ix) = 0.0; AT(rho_xp, nz - 1, ix) = 0.0;
    }
#pragma omp parallel for collapse(2) reduction(+ : cl, cm, cr)
    for (int iz = 0; iz < nz - 1; iz++)
        for (int ix = 0; ix < nx - 1; ix++) {
            double m00 = AT(mu, iz, ix), m01 = AT(mu, iz, ix + 1);
            double m10 = AT(mu, iz + 1, ix), m11 = AT(mu, iz + 1, ix + 1);
            double v = 4.0 / ((1.0 / m00) + (1.0 / m01) + (1.0 / m10) + (1.0 / m11));
            if (m00 == 0.0 || m01 == 0.0 || m10 == 0.0 || m11 == 0.0) v = 0.0;
            AT(mu_zx, iz, ix) = v;
            double r00 = AT(rho, iz, ix), r10 = AT(rho, iz + 1, ix), r01 = AT(rho, iz, ix + 1);
            AT(rho_zp, iz, ix) = 1.0 / (0.5 * (r00 + r10));
            AT(rho_xp, iz, ix) = 1.0 / (0.5 * (r00 + r01));
            if (r00 < 1e-4 && r10 < 1e-4) AT(rho_zp, iz, ix) = 0.0;
            if (r00 < 1e-4 && r01 < 1e-4) AT(rho_zp, iz, ix) = 0.0;
            cl += AT(lam, iz, ix);
            cm += AT(mu, iz, ix);
            cr += AT(rho, iz, ix);
        }
    int ncell = (nz - 1) * (nx - 1);
    C[0] = cl / ncell;
    C[1] = cm / ncell;
    C[2] = cr / ncell;
}

/* copy_mat (src/cpu/n_kernel_lib_PSV.cpp:834-851), one array. */
void or_copy_mat(double *dst, const double *src, int32_t nz, int32_t nx)
{
    memcpy(dst, src, sizeof(double) * (size_t)nz * nx);
}

/* update_mat2 (src/cpu/n_kernel_lib_PSV.cpp:785-832):
 * f = max(max|m_old|, m_max) / max|G|;  m = clip(m_old + eps f G, m_min, m_max). */
void or_update_mat2(double *mat, const double *mat_old, const double *grad, double mat_max, double mat_min,
                    double step_length, int32_t nz, int32_t nx, double *step_factor_out)
{
    double grad_max = 0.0, mat_array_max = 0.0;
    size_t n = (size_t)nz * nx;
    for (size_t i = 0; i < n; i++) {
        if (fabs(grad[i]) > grad_max) grad_max = fabs(grad[i]);
        if (fabs(mat_old[i]) > mat_array_max) mat_array_max = fabs(mat_old[i]);
    }
    if (mat_array_max < mat_max) mat_array_max = mat_max;
    double step_factor = mat_array_max / grad_max;
    if (step_factor_out) *step_factor_out = step_factor;
    for (size_t i = 0; i < n; i++) {
        double v = mat_old[i] + step_length * step_factor * grad[i];
        if (v > mat_max) v = mat_max;
        if (v < mat_min) v = mat_min;
        mat[i] = v;
    }
}

/* ========================================================================== */
/* Time stepping (kernel_PSV and its kernel library)                           */
/* ========================================================================== */

/* reset_sv2 / reset_PML_memory2 (src/cpu/n_kernel_lib_PSV.cpp:19-69). */
void or_reset_state(const or_grid *g, or_state *s)
{
    size_t n = (size_t)g->nz * g->nx, b = sizeof(double) * n;
    memset(s->vz, 0, b); memset(s->vx, 0, b); memset(s->uz, 0, b); memset(s->ux, 0, b);
    memset(s->szz, 0, b); memset(s->szx, 0, b); memset(s->sxx, 0, b); memset(s->We, 0, b);
    if (g->pml_z) {
        memset(s->mem_vz_z, 0, b); memset(s->mem_vx_z, 0, b);
        memset(s->mem_szz_z, 0, b); memset(s->mem_szx_z, 0, b);
    }
    if (g->pml_x) {
        memset(s->mem_vz_x, 0, b); memset(s->mem_vx_x, 0, b);
        memset(s->mem_szx_x, 0, b); memset(s->mem_sxx_x, 0, b);
    }
}

/* vdiff2 (src/cpu/n_kernel_lib_PSV.cpp:100-131), order 2N. */
void or_vdiff(const or_grid *g, or_state *s)
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    const double dxi = 1.0 / g->dx, dzi = 1.0 / g->dz;
    const double *hc = g->hc;
    const double *vz = s->vz, *vx = s->vx;
#pragma omp parallel for collapse(2)
    for (int iz = nz1; iz < nz2; iz++)
        for (int ix = nx1; ix < nx2; ix++) {
            double a = dzi * hc[1] * (AT(vz, iz, ix) - AT(vz, iz - 1, ix));
            double b = dzi * hc[1] * (AT(vx, iz + 1, ix) - AT(vx, iz, ix));
            double c = dxi * hc[1] * (AT(vz, iz, ix + 1) - AT(vz, iz, ix));
            double d = dxi * hc[1] * (AT(vx, iz, ix) - AT(vx, iz, ix - 1));
            for (int k = 2; k <= N; k++) {
                a += dzi * hc[k] * (AT(vz, iz + k - 1, ix) - AT(vz, iz - k, ix));
                b += dzi * hc[k] * (AT(vx, iz + k, ix) - AT(vx, iz - k + 1, ix));
                c += dxi * hc[k] * (AT(vz, iz, ix + k) - AT(vz, iz, ix - k + 1));
                d += dxi * hc[k] * (AT(vx, iz, ix + k - 1) - AT(vx, iz, ix - k));
            }
            AT(s->dz_z, iz, ix) = a;
            AT(s->dx_z, iz, ix) = b;
            AT(s->dz_x, iz, ix) = c;
            AT(s->dx_x, iz, ix) = d;
        }
}

/* sdiff2 (src/cpu/n_kernel_lib_PSV.cpp:207-236), order 2N; outputs in the
 * reference's slots dz_z = szz_z, dx_z = szx_z, dz_x = szx_x, dx_x = sxx_x. */
void or_sdiff(const or_grid *g, or_state *s)
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    const double dxi = 1.0 / g->dx, dzi = 1.0 / g->dz;
    const double *hc = g->hc;
    const double *szz = s->szz, *szx = s->szx, *sxx = s->sxx;
#pragma omp parallel for collapse(2)
    for (int iz = nz1; iz < nz2; iz++)
        for (int ix = nx1; ix < nx2; ix++) {
            double a = dzi * hc[1] * (AT(szz, iz + 1, ix) - AT(szz, iz, ix));
            double b = dzi * hc[1] * (AT(szx, iz, ix) - AT(szx, iz - 1, ix));
            double c = dxi * hc[1] * (AT(szx, iz, ix) - AT(szx, iz, ix - 1));
            double d = dxi * hc[1] * (AT(sxx, iz, ix + 1) - AT(sxx, iz, ix));
            for (int k = 2; k <= N; k++) {
                a += dzi * hc[k] * (AT(szz, iz + k, ix) - AT(szz, iz - k + 1, ix));
                b += dzi * hc[k] * (AT(szx, iz + k - 1, ix) - AT(szx, iz - k, ix));
                c += dxi * hc[k] * (AT(szx, iz, ix + k - 1) - AT(szx, iz, ix - k));
                d += dxi * hc[k] * (AT(sxx, iz, ix + k) - AT(sxx, iz, ix - k + 1));
            }
            AT(s->dz_z, iz, ix) = a;
            AT(s->dx_z, iz, ix) = b;
            AT(s->dz_x, iz, ix) = c;
            AT(s->dx_x, iz, ix) = d;
        }
}

/* pml_diff2 (src/cpu/n_kernel_lib_PSV.cpp:134-178): psi = b psi + a d;
 * d = d / K + psi, over the whole interior.  Slot dz_z and dx_x take the
 * integer-point profiles, dx_z and dz_x the half-point ones, in both the
 * velocity and the stress pass (src/cpu/n_kernel_PSV.cpp:112-116,142-146). */
static void pml_diff(const or_grid *g, or_state *s, double *mem_z_z, double *mem_x_z,
                     double *mem_z_x, double *mem_x_x)
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    double *dz_z = s->dz_z, *dx_z = s->dx_z, *dz_x = s->dz_x, *dx_x = s->dx_x;
#pragma omp parallel for collapse(2)
    for (int iz = nz1; iz < nz2; iz++)
        for (int ix = nx1; ix < nx2; ix++) {
            if (g->pml_z) {
                AT(mem_z_z, iz, ix) = g->b_z[iz] * AT(mem_z_z, iz, ix) + g->a_z[iz] * AT(dz_z, iz, ix);
                AT(mem_x_z, iz, ix) = g->b_half_z[iz] * AT(mem_x_z, iz, ix) + g->a_half_z[iz] * AT(dx_z, iz, ix);
                AT(dz_z, iz, ix) = AT(dz_z, iz, ix) / g->K_z[iz] + AT(mem_z_z, iz, ix);
                AT(dx_z, iz, ix) = AT(dx_z, iz, ix) / g->K_half_z[iz] + AT(mem_x_z, iz, ix);
            }
            if (g->pml_x) {
                AT(mem_x_x, iz, ix) = g->b_x[ix] * AT(mem_x_x, iz, ix) + g->a_x[ix] * AT(dx_x, iz, ix);
                AT(mem_z_x, iz, ix) = g->b_half_x[ix] * AT(mem_z_x, iz, ix) + g->a_half_x[ix] * AT(dz_x, iz, ix);
                AT(dx_x, iz, ix) = AT(dx_x, iz, ix) / g->K_x[ix] + AT(mem_x_x, iz, ix);
                AT(dz_x, iz, ix) = AT(dz_x, iz, ix) / g->K_half_x[ix] + AT(mem_z_x, iz, ix);
            }
        }
}

/* update_s2 (src/cpu/n_kernel_lib_PSV.cpp:181-205). */
static void update_s(const or_grid *g, const or_media *m, or_state *s)
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    const double dt = g->dt;
#pragma omp parallel for collapse(2)
    for (int iz = nz1; iz < nz2; iz++)
        for (int ix = nx1; ix < nx2; ix++) {
            double vz_z = AT(s->dz_z, iz, ix), vx_z = AT(s->dx_z, iz, ix);
            double vz_x = AT(s->dz_x, iz, ix), vx_x = AT(s->dx_x, iz, ix);
            double lam = AT(m->lam, iz, ix), mu = AT(m->mu, iz, ix);
            AT(s->szz, iz, ix) += dt * (lam * (vx_x + vz_z) + (2.0 * mu * vz_z));
            AT(s->szx, iz, ix) += dt * AT(m->mu_zx, iz, ix) * (vz_x + vx_z);
            AT(s->sxx, iz, ix) += dt * (lam * (vx_x + vz_z) + (2.0 * mu * vx_x));
        }
}

/* surf_mirror (src/cpu/n_kernel_lib_PSV.cpp:272-395): free surface by stress
 * imaging.  Surface row/column: the normal and shear stresses are zeroed and the
 * tangential stress is set (not incremented) to 4 dt d_t v_t (lam mu + mu^2) / (lam + 2 mu);
 * then the stresses in between the interior edge and the surface are the
 * negated mirror images.  Sides are processed top, bottom, left, right. */
static void surf_mirror(const or_grid *g, const or_media *m, or_state *s)
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    const double dt = g->dt;
    double *szz = s->szz, *szx = s->szx, *sxx = s->sxx;
    const double *lam = m->lam, *mu = m->mu;
    int is;
    if ((is = g->isurf[0]) > 0) {
        for (int ix = nx1; ix < nx2; ix++) {
            double l = AT(lam, is, ix), u = AT(mu, is, ix);
            AT(szz, is, ix) = 0.0;
            AT(szx, is, ix) = 0.0;
            AT(sxx, is, ix) = 4.0 * dt * AT(s->dx_x, is, ix) * (l * u + u * u) / (l + 2.0 * u);
            for (int k = 1; k < is - nz1 + 1; k++) {
                AT(szx, is - k, ix) = -AT(szx, is + k, ix);
                AT(szz, is - k, ix) = -AT(szz, is + k, ix);
            }
        }
    }
    if ((is = g->isurf[1]) > 0) {
        for (int ix = nx1; ix < nx2; ix++) {
            double l = AT(lam, is, ix), u = AT(mu, is, ix);
            AT(szz, is, ix) = 0.0;
            AT(szx, is, ix) = 0.0;
            AT(sxx, is, ix) = 4.0 * dt * AT(s->dx_x, is, ix) * (l * u + u * u) / (l + 2.0 * u);
            for (int k = 1; k <= nz2 - is; k++) {
                AT(szx, is + k, ix) = -AT(szx, is - k, ix);
                AT(szz, is + k, ix) = -AT(szz, is - k, ix);
            }
        }
    }
    if ((is = g->isurf[2]) > 0) {
        for (int iz = nz1; iz < nz2; iz++) {
            double l = AT(lam, iz, is), u = AT(mu, iz, is);
            AT(sxx, iz, is) = 0.0;
            AT(szx, iz, is) = 0.0;
            AT(szz, iz, is) = 4.0 * dt * AT(s->dz_z, iz, is) * (l * u + u * u) / (l + 2.0 * u);
            for (int k = 1; k < is - nx1 + 1; k++) {
                AT(szx, iz, is - k) = -AT(szx, iz, is + k);
                AT(sxx, iz, is - k) = -AT(sxx, iz, is + k);
            }
        }
    }
    if ((is = g->isurf[3]) > 0) {
        for (int iz = nz1; iz < nz2; iz++) {
            double l = AT(lam, iz, is), u = AT(mu, iz, is);
            AT(sxx, iz, is) = 0.0;
            AT(szx, iz, is) = 0.0;
            AT(szz, iz, is) = 4.0 * dt * AT(s->dz_z, iz, is) * (l * u + u * u) / (l + 2.0 * u);
            for (int k = 1; k <= nx2 - is; k++) {
                AT(szx, iz, is + k) = -AT(szx, iz, is - k);
                AT(sxx, iz, is + k) = -AT(sxx, iz, is - k);
            }
        }
    }
}

/* update_v2 (src/cpu/n_kernel_lib_PSV.cpp:239-270): u += dt v (old v), v update,
 * We += vx^2 + vz^2 (new v). */
static void update_v(const or_grid *g, const or_media *m, or_state *s)
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    const double dt = g->dt;
#pragma omp parallel for collapse(2)
    for (int iz = nz1; iz < nz2; iz++)
        for (int ix = nx1; ix < nx2; ix++) {
            double vz = AT(s->vz, iz, ix), vx = AT(s->vx, iz, ix);
            AT(s->uz, iz, ix) += dt * vz;
            AT(s->ux, iz, ix) += dt * vx;
            vz += dt * AT(m->rho_zp, iz, ix) * (AT(s->dz_x, iz, ix) + AT(s->dz_z, iz, ix));
            vx += dt * AT(m->rho_xp, iz, ix) * (AT(s->dx_x, iz, ix) + AT(s->dx_z, iz, ix));
            AT(s->vz, iz, ix) = vz;
            AT(s->vx, iz, ix) = vx;
            AT(s->We, iz, ix) += vx * vx + vz * vz;
        }
}

/* vsrc2 (src/cpu/n_kernel_lib_PSV.cpp:487-536); sources added in index order. */
static void vsrc(const or_grid *g, const or_media *m, or_state *s, const or_sources *src, int ishot, int it)
{
    const int nx = g->nx, nt = g->nt;
    const double dt = g->dt, dz = g->dz, dx = g->dx;
    for (int is = 0; is < src->nsrc; is++) {
        if (src->shot_to_fire[is] != ishot) continue;
        int zs = src->z_src[is], xs = src->x_src[is];
        double sz = src->stf_z[(size_t)is * nt + it], sx = src->stf_x[(size_t)is * nt + it];
        if (src->stf_type == 0) {
            AT(s->vz, zs, xs) += dt * AT(m->rho_zp, zs, xs) * sz / (dz * dx);
            AT(s->vx, zs, xs) += dt * AT(m->rho_xp, zs, xs) * sx / (dz * dx);
        } else if (src->stf_type == 1) {
            AT(s->vz, zs, xs) += sz;
            AT(s->vx, zs, xs) += sx;
        }
    }
}

/* urec2 (src/cpu/n_kernel_lib_PSV.cpp:538-576): displacement seismogram as the
 * running sum of dt v / (dz dx), sampled after source injection. */
static void urec(const or_grid *g, const or_state *s, or_receivers *rec, int it)
{
    const int nx = g->nx, nt = g->nt;
    const double dt = g->dt, dz = g->dz, dx = g->dx;
    if (rec->rtf_type != 0) return;
    for (int ir = 0; ir < rec->nrec; ir++) {
        int zr = rec->z_rec[ir], xr = rec->x_rec[ir];
        double *uz = rec->rtf_uz + (size_t)ir * nt, *ux = rec->rtf_ux + (size_t)ir * nt;
        if (it == 0) {
            uz[it] = dt * AT(s->vz, zr, xr) / (dz * dx);
            ux[it] = dt * AT(s->vx, zr, xr) / (dz * dx);
        } else {
            uz[it] = uz[it - 1] + dt * AT(s->vz, zr, xr) / (dz * dx);
            ux[it] = ux[it - 1] + dt * AT(s->vx, zr, xr) / (dz * dx);
        }
    }
}

/* gard_fwd_storage2 (src/cpu/n_kernel_lib_PSV.cpp:398-436): stresses as they
 * are, velocities divided by dt, over the inclusive strided snap box. */
static void fwd_storage(const or_grid *g, const or_state *s, or_snaps *a, int tf)
{
    const int nx = g->nx;
    const int snz = 1 + (g->snap_z2 - g->snap_z1) / g->snap_dz;
    const int snx = 1 + (g->snap_x2 - g->snap_x1) / g->snap_dx;
    const double dt = g->dt;
    const size_t base = (size_t)tf * snz * snx;
#pragma omp parallel for
    for (int jz = 0; jz < snz; jz++) {
        int iz = g->snap_z1 + jz * g->snap_dz;
        for (int jx = 0; jx < snx; jx++) {
            int ix = g->snap_x1 + jx * g->snap_dx;
            size_t o = base + (size_t)jz * snx + jx;
            a->sxx[o] = AT(s->sxx, iz, ix);
            a->szx[o] = AT(s->szx, iz, ix);
            a->szz[o] = AT(s->szz, iz, ix);
            a->vx[o] = AT(s->vx, iz, ix) / dt;
            a->vz[o] = AT(s->vz, iz, ix) / dt;
        }
    }
}

/* fwi_grad2 (src/cpu/n_kernel_lib_PSV.cpp:439-484): zero-lag correlation of the
 * stored forward field with the adjoint stresses and adjoint displacement,
 * subtracted (the CPU sign) with weight snap_dt * dt. */
static void fwi_grad(const or_grid *g, const or_media *m, const or_state *s, const or_snaps *a, int tf,
                     double *g_lam, double *g_mu, double *g_rho)
{
    const int nx = g->nx;
    const int snz = 1 + (g->snap_z2 - g->snap_z1) / g->snap_dz;
    const int snx = 1 + (g->snap_x2 - g->snap_x1) / g->snap_dx;
    const double w = g->snap_dt * g->dt;
    const size_t base = (size_t)tf * snz * snx;
#pragma omp parallel for
    for (int jz = 0; jz < snz; jz++) {
        int iz = g->snap_z1 + jz * g->snap_dz;
        for (int jx = 0; jx < snx; jx++) {
            int ix = g->snap_x1 + jx * g->snap_dx;
            size_t o = base + (size_t)jz * snx + jx, q = (size_t)jz * snx + jx;
            double lam = AT(m->lam, iz, ix), mu = AT(m->mu, iz, ix);
            double szz = AT(s->szz, iz, ix), sxx = AT(s->sxx, iz, ix), szx = AT(s->szx, iz, ix);
            double s1 = 0.25 * (a->szz[o] + a->sxx[o]) * (szz + sxx) / ((lam + mu) * (lam + mu));
            double s2 = 0.25 * (a->szz[o] - a->sxx[o]) * (szz - sxx) / (mu * mu);
            double s3 = (a->szx[o] * szx) / (mu * mu);
            double s4 = AT(s->ux, iz, ix) * a->vx[o] + AT(s->uz, iz, ix) * a->vz[o];
            g_lam[q] -= w * s1;
            g_mu[q] -= w * (s3 + s1 + s2);
            g_rho[q] -= w * s4;
        }
    }
}

/* One time step of kernel_PSV (src/cpu/n_kernel_PSV.cpp:102-208). */
void or_step(const or_grid *g, const or_media *m, or_state *s, int32_t it, int32_t ishot,
             const or_sources *src, or_receivers *rec, int32_t record,
             int32_t accu, or_snaps *snap,
             int32_t grad, double *g_lam, double *g_mu, double *g_rho)
{
    or_vdiff(g, s);
    if (g->pml_z || g->pml_x) pml_diff(g, s, s->mem_vz_z, s->mem_vx_z, s->mem_vz_x, s->mem_vx_x);
    update_s(g, m, s);
    if (g->surf) surf_mirror(g, m, s);
    or_sdiff(g, s);
    if (g->pml_z || g->pml_x) pml_diff(g, s, s->mem_szz_z, s->mem_szx_z, s->mem_szx_x, s->mem_sxx_x);
    update_v(g, m, s);
    if (src && src->nsrc) vsrc(g, m, s, src, ishot, it);
    if (record && rec && rec->nrec) urec(g, s, rec, it);
    if (!(it % g->snap_dt)) {
        int tf = it / g->snap_dt;
        if (accu && snap) fwd_storage(g, s, snap, tf);
        if (grad) fwi_grad(g, m, s, snap, tf, g_lam, g_mu, g_rho);
    }
}

/* kernel_PSV (src/cpu/n_kernel_PSV.cpp:20-213): reset, then nt steps forward
 * (grad = 0) or time-reversed (grad = 1, receivers not recorded). */
void or_kernel_psv(const or_grid *g, const or_media *m, or_state *s, int32_t ishot,
                   int32_t grad, double *g_lam, double *g_mu, double *g_rho,
                   const or_sources *src, or_receivers *rec,
                   int32_t accu, or_snaps *snap)
{
    or_reset_state(g, s);
    if (grad) {
        const int snz = 1 + (g->snap_z2 - g->snap_z1) / g->snap_dz;
        const int snx = 1 + (g->snap_x2 - g->snap_x1) / g->snap_dx;
        size_t n = (size_t)snz * snx;
        memset(g_lam, 0, sizeof(double) * n);
        memset(g_mu, 0, sizeof(double) * n);
        memset(g_rho, 0, sizeof(double) * n);
    }
    for (int jt = 0; jt < g->nt; jt++) {
        int it = grad ? g->nt - jt - 1 : jt;
        or_step(g, m, s, it, ishot, src, rec, !grad, accu, snap, grad, g_lam, g_mu, g_rho);
    }
}

/* ========================================================================== */
/* Misfit, preconditioning, taper                                              */
/* ========================================================================== */

/* adjsrc2 (src/cpu/n_kernel_lib_PSV.cpp:579-626): residual d_mod - d_true into
 * a_*, and L2 = sum 0.5 dt (r_z^2 + r_x^2), summed serially in (receiver, time)
 * order.  a_* may alias mod_* (the reference writes the residual in place). */
double or_adjsrc2(double *a_uz, double *a_ux, const double *mod_uz, const double *mod_ux,
                  const double *true_uz, const double *true_ux, double dt, int32_t nseis, int32_t nt)
{
    double L2 = 0;
    size_t n = (size_t)nseis * nt;
    for (size_t i = 0; i < n; i++) {
        double rz = mod_uz[i] - true_uz[i];
        double rx = mod_ux[i] - true_ux[i];
        a_uz[i] = rz;
        a_ux[i] = rx;
        L2 += 0.5 * dt * (rz * rz) + 0.5 * dt * (rx * rx);
    }
    return L2;
}

/* energy_weights2 (src/cpu/n_kernel_lib_PSV.cpp:696-757): We = sqrt(We We_adj)
 * over the half-open snap box, then We += 0.005 max(We).  Returns max(We). */
double or_energy_weights2(double *We, const double *We_adj, int32_t nx,
                          int32_t snap_z1, int32_t snap_z2, int32_t snap_x1, int32_t snap_x2)
{
    double mx = 0.0;
    for (int iz = snap_z1; iz < snap_z2; iz++)
        for (int ix = snap_x1; ix < snap_x2; ix++) {
            double w = sqrt(AT(We, iz, ix) * AT(We_adj, iz, ix));
            AT(We, iz, ix) = w;
            if (w > mx) mx = w;
        }
    for (int iz = snap_z1; iz < snap_z2; iz++)
        for (int ix = snap_x1; ix < snap_x2; ix++) AT(We, iz, ix) += 0.005 * mx;
    return mx;
}

/* interpol_grad2 (src/cpu/n_kernel_lib_PSV.cpp:628-693): snapped gradient onto
 * the full grid: copy the snap points, then linear along x on the snap rows,
 * then linear along z for every column. */
void or_interpol_grad2(double *grad, const double *grad_shot, int32_t nx,
                       int32_t snap_z1, int32_t snap_z2, int32_t snap_x1, int32_t snap_x2,
                       int32_t snap_dz, int32_t snap_dx)
{
    const int snx = 1 + (snap_x2 - snap_x1) / snap_dx;
    int jz = 0;
    for (int iz = snap_z1; iz <= snap_z2; iz += snap_dz, jz++) {
        int jx = 0;
        for (int ix = snap_x1; ix <= snap_x2; ix += snap_dx, jx++)
            AT(grad, iz, ix) = grad_shot[(size_t)jz * snx + jx];
    }
    if (snap_dx > 1)
        for (int iz = snap_z1; iz < snap_z2; iz += snap_dz)
            for (int ix = snap_x1; ix < snap_x2; ix += snap_dx) {
                double t = (AT(grad, iz, ix + snap_dx) - AT(grad, iz, ix)) / snap_dx;
                for (int k = 1; k < snap_dx; k++) AT(grad, iz, ix + k) = AT(grad, iz, ix) + t * k;
            }
    if (snap_dz > 1)
        for (int iz = snap_z1; iz < snap_z2; iz += snap_dz)
            for (int ix = snap_x1; ix < snap_x2; ix++) {
                double t = (AT(grad, iz + snap_dz, ix) - AT(grad, iz, ix)) / snap_dz;
                for (int k = 1; k < snap_dz; k++) AT(grad, iz + k, ix) = AT(grad, iz, ix) + t * k;
            }
}

/* scale_grad_E2 (src/cpu/n_kernel_lib_PSV.cpp:760-783): G += G_shot / (We m^2). */
void or_scale_grad_E2(double *grad, const double *grad_shot, double mat_av, const double *We, int32_t nx,
                      int32_t snap_z1, int32_t snap_z2, int32_t snap_x1, int32_t snap_x2)
{
    if (!(mat_av > 0)) return;
    for (int iz = snap_z1; iz < snap_z2; iz++)
        for (int ix = snap_x1; ix < snap_x2; ix++)
            AT(grad, iz, ix) += AT(grad_shot, iz, ix) / (AT(We, iz, ix) * mat_av * mat_av);
}

/* taper2 (src/cpu/n_kernel_lib_PSV.cpp:942-1010): zero / Tukey ramps in x, then
 * in z.  taper = {t1, t2, b1, b2, l1, l2, r1, r2}. */
void or_taper2(double *A, int32_t nz, int32_t nx, int32_t snap_z1, int32_t snap_z2, int32_t snap_x1,
               int32_t snap_x2, const int32_t taper[8])
{
    const int t1 = taper[0], t2 = taper[1], b1 = taper[2], b2 = taper[3];
    const int l1 = taper[4], l2 = taper[5], r1 = taper[6], r2 = taper[7];
    const int tl = l2 - l1, tr = r1 - r2, tt = t2 - t1, tb = b1 - b2;
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            if (ix >= snap_x1 && ix < l1) AT(A, iz, ix) *= 0.0;
            else if (ix >= l1 && ix < l2) AT(A, iz, ix) *= 0.5 * (1.0 - cos(OR_PI * (ix - l1) / tl));
            else if (ix > r2 && ix < r1) AT(A, iz, ix) *= 0.5 * (1.0 - cos(OR_PI * (r1 - ix) / tr));
            else if (ix >= r1 && ix <= snap_x2) AT(A, iz, ix) *= 0.0;
        }
    for (int ix = 0; ix < nx; ix++)
        for (int iz = 0; iz < nz; iz++) {
            if (iz >= snap_z1 && iz < t1) AT(A, iz, ix) *= 0.0;
            else if (iz >= t1 && iz < t2) AT(A, iz, ix) *= 0.5 * (1.0 - cos(OR_PI * (iz - t1) / tt));
            else if (iz > b2 && iz < b1) AT(A, iz, ix) *= 0.5 * (1.0 - cos(OR_PI * (b1 - iz) / tb));
            else if (iz >= b1 && iz <= snap_z2) AT(A, iz, ix) *= 0.0;
        }
}

/* ========================================================================== */
/* Step length                                                                 */
/* ========================================================================== */

/* calc_opt_step (src/cpu/n_step_length_PSV.cpp:461-518): vertex of the parabola
 * through (sl_i, L2_i), then the reference's clamps.  The reference solves the
 * 3x3 Vandermonde system with a Gauss routine whose back-substitution runs out
 * of bounds (include/cpu/n_solvelin.hpp:37-42); here the vertex is the closed
 * form of the same interpolating parabola (Newton divided differences). */
double or_calc_opt_step(const double L2[3], const double sl[3])
{
    double d01 = (L2[1] - L2[0]) / (sl[1] - sl[0]);
    double d12 = (L2[2] - L2[1]) / (sl[2] - sl[1]);
    double x2 = (d12 - d01) / (sl[2] - sl[0]);
    double x1 = d01 - x2 * (sl[0] + sl[1]);
    double opteps = -x1 / (2.0 * x2);
    if (opteps > sl[2]) opteps = sl[2];
    if (opteps < 0.0) opteps = sl[1];
    if (L2[1] < L2[0] && L2[2] < L2[1]) opteps = sl[2];
    return opteps;
}

/* ========================================================================== */
/* Exact discrete adjoint (reverse mode through or_step)                        */
/* ========================================================================== */

/* Scatter of the transposed staggered derivatives.  vdiff forms, per interior
 * cell, dz_z = sum_k cz_k (vz[i+k-1] - vz[i-k]), dx_z = sum cz_k (vx[i+k] - vx[i-k+1]),
 * dz_x = sum cx_k (vz[j+k] - vz[j-k+1]), dx_x = sum cx_k (vx[j+k-1] - vx[j-k]); the
 * transpose adds each coefficient times the derivative's adjoint back onto the cells
 * it read.  sdiff (second function) uses the mirrored offsets. */
static void vdiff_T(const or_grid *g, const double *D[4], double *bvz, double *bvx)
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    const double dxi = 1.0 / g->dx, dzi = 1.0 / g->dz;
    for (int iz = nz1; iz < nz2; iz++)
        for (int ix = nx1; ix < nx2; ix++) {
            const double a = AT(D[0], iz, ix), b = AT(D[1], iz, ix), c = AT(D[2], iz, ix), d = AT(D[3], iz, ix);
            for (int k = 1; k <= N; k++) {
                const double cz = dzi * g->hc[k], cx = dxi * g->hc[k];
                AT(bvz, iz + k - 1, ix) += cz * a;
                AT(bvz, iz - k, ix) -= cz * a;
                AT(bvx, iz + k, ix) += cz * b;
                AT(bvx, iz - k + 1, ix) -= cz * b;
                AT(bvz, iz, ix + k) += cx * c;
                AT(bvz, iz, ix - k + 1) -= cx * c;
                AT(bvx, iz, ix + k - 1) += cx * d;
                AT(bvx, iz, ix - k) -= cx * d;
            }
        }
}

static void sdiff_T(const or_grid *g, const double *D[4], double *bzz, double *bzx, double *bxx)
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    const double dxi = 1.0 / g->dx, dzi = 1.0 / g->dz;
    for (int iz = nz1; iz < nz2; iz++)
        for (int ix = nx1; ix < nx2; ix++) {
            const double a = AT(D[0], iz, ix), b = AT(D[1], iz, ix), c = AT(D[2], iz, ix), d = AT(D[3], iz, ix);
            for (int k = 1; k <= N; k++) {
                const double cz = dzi * g->hc[k], cx = dxi * g->hc[k];
                AT(bzz, iz + k, ix) += cz * a;
                AT(bzz, iz - k + 1, ix) -= cz * a;
                AT(bzx, iz + k - 1, ix) += cz * b;
                AT(bzx, iz - k, ix) -= cz * b;
                AT(bzx, iz, ix + k - 1) += cx * c;
                AT(bzx, iz, ix - k) -= cx * c;
                AT(bxx, iz, ix + k) += cx * d;
                AT(bxx, iz, ix - k + 1) -= cx * d;
            }
        }
}

/* pml_diff reversed: the forward psi' = b psi + a d, d~ = d / K + psi' over the
 * interior (slots 0, 1 with the z profiles (integer, half), 2, 3 with the x profiles
 * (half, integer)).  In reverse, P = carry + adj(d~), adj(d) = adj(d~) / K + a P and the
 * carry passed to the previous step is b P. */
static void pml_T(const or_grid *g, double *D[4], double *carry[4])
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    for (int iz = nz1; iz < nz2; iz++)
        for (int ix = nx1; ix < nx2; ix++) {
            if (g->pml_z) {
                double P = AT(carry[0], iz, ix) + AT(D[0], iz, ix);
                AT(D[0], iz, ix) = AT(D[0], iz, ix) / g->K_z[iz] + g->a_z[iz] * P;
                AT(carry[0], iz, ix) = g->b_z[iz] * P;
                P = AT(carry[1], iz, ix) + AT(D[1], iz, ix);
                AT(D[1], iz, ix) = AT(D[1], iz, ix) / g->K_half_z[iz] + g->a_half_z[iz] * P;
                AT(carry[1], iz, ix) = g->b_half_z[iz] * P;
            }
            if (g->pml_x) {
                double P = AT(carry[3], iz, ix) + AT(D[3], iz, ix);
                AT(D[3], iz, ix) = AT(D[3], iz, ix) / g->K_x[ix] + g->a_x[ix] * P;
                AT(carry[3], iz, ix) = g->b_x[ix] * P;
                P = AT(carry[2], iz, ix) + AT(D[2], iz, ix);
                AT(D[2], iz, ix) = AT(D[2], iz, ix) / g->K_half_x[ix] + g->a_half_x[ix] * P;
                AT(carry[2], iz, ix) = g->b_half_x[ix] * P;
            }
        }
}

/* surf_mirror reversed, sides in reverse order (right, left, bottom, top); per line the
 * images are undone, then the assignments: the set stress s = 4 dt d (lam mu + mu^2) /
 * (lam + 2 mu) sends its adjoint to the tangential strain d (bD) and to lam, mu. */
static void surf_T(const or_grid *g, const or_media *m, const double *Ev[4], double *bzz, double *bzx,
                   double *bxx, double *bD[4], double *gl, double *gm)
{
    const int nx = g->nx, N = g->fdorder / 2;
    const int nz1 = N, nz2 = g->nz - N, nx1 = N, nx2 = g->nx - N;
    const double dt = g->dt;
    for (int side = 3; side >= 0; side--) {
        const int is = g->isurf[side];
        if (is <= 0) continue;
        const int n_lines = side < 2 ? nx2 - nx1 : nz2 - nz1;
        for (int l = 0; l < n_lines; l++) {
            int iz = side < 2 ? is : nz1 + l, ix = side < 2 ? nx1 + l : is;
            if (side == 0)
                for (int k = 1; k < is - nz1 + 1; k++) {
                    AT(bzx, is + k, ix) -= AT(bzx, is - k, ix); AT(bzx, is - k, ix) = 0.0;
                    AT(bzz, is + k, ix) -= AT(bzz, is - k, ix); AT(bzz, is - k, ix) = 0.0;
                }
            else if (side == 1)
                for (int k = 1; k <= nz2 - is; k++) {
                    AT(bzx, is - k, ix) -= AT(bzx, is + k, ix); AT(bzx, is + k, ix) = 0.0;
                    AT(bzz, is - k, ix) -= AT(bzz, is + k, ix); AT(bzz, is + k, ix) = 0.0;
                }
            else if (side == 2)
                for (int k = 1; k < is - nx1 + 1; k++) {
                    AT(bzx, iz, is + k) -= AT(bzx, iz, is - k); AT(bzx, iz, is - k) = 0.0;
                    AT(bxx, iz, is + k) -= AT(bxx, iz, is - k); AT(bxx, iz, is - k) = 0.0;
                }
            else
                for (int k = 1; k <= nx2 - is; k++) {
                    AT(bzx, iz, is - k) -= AT(bzx, iz, is + k); AT(bzx, iz, is + k) = 0.0;
                    AT(bxx, iz, is - k) -= AT(bxx, iz, is + k); AT(bxx, iz, is + k) = 0.0;
                }
            const double l_ = AT(m->lam, iz, ix), u = AT(m->mu, iz, ix), den = l_ + 2.0 * u;
            const double f = (l_ * u + u * u) / den;
            const double f_l = u * u / (den * den), f_u = (l_ * l_ + 2.0 * l_ * u + 2.0 * u * u) / (den * den);
            /* top / bottom set sxx from dx_x (slot 3); left / right set szz from dz_z (slot 0) */
            double *bs = side < 2 ? bxx : bzz;
            const int slot = side < 2 ? 3 : 0;
            const double a = AT(bs, iz, ix), d = AT(Ev[slot], iz, ix);
            AT(bD[slot], iz, ix) += a * 4.0 * dt * f;
            AT(gl, iz, ix) += a * 4.0 * dt * d * f_l;
            AT(gm, iz, ix) += a * 4.0 * dt * d * f_u;
            AT(bzz, iz, ix) = 0.0;
            AT(bzx, iz, ix) = 0.0;
            AT(bxx, iz, ix) = 0.0;
        }
    }
}

void or_adjoint_exact(const or_grid *g, const or_media *m, int32_t ishot, const or_sources *src,
                      int32_t nrec, const int32_t *z_rec, const int32_t *x_rec, const double *cot_uz,
                      const double *cot_ux, double *g_lam, double *g_mu, double *g_muzx, double *g_rhoz,
                      double *g_rhox)
{
    const int nz = g->nz, nx = g->nx, nt = g->nt, N = g->fdorder / 2;
    const int nz1 = N, nz2 = nz - N, nx1 = N, nx2 = nx - N;
    const size_t n = (size_t)nz * nx, fb = sizeof(double) * n;
    const double dt = g->dt, rc = dt / (g->dz * g->dx);
    /* forward state at the start of every step: vz vx szz szx sxx and the 8 memories */
    double *hist = malloc(fb * 13 * (size_t)nt);
    double *buf = calloc(n * 20, sizeof(double));
    or_state s = {buf, buf + n, buf + 2 * n, buf + 3 * n, buf + 4 * n, buf + 5 * n, buf + 6 * n, buf + 7 * n,
                  buf + 8 * n, buf + 9 * n, buf + 10 * n, buf + 11 * n,
                  buf + 12 * n, buf + 13 * n, buf + 14 * n, buf + 15 * n,
                  buf + 16 * n, buf + 17 * n, buf + 18 * n, buf + 19 * n};
    double **keep[13] = {&s.vz, &s.vx, &s.szz, &s.szx, &s.sxx, &s.mem_vz_z, &s.mem_vx_z, &s.mem_szz_z,
                         &s.mem_szx_z, &s.mem_vz_x, &s.mem_vx_x, &s.mem_szx_x, &s.mem_sxx_x};
    or_reset_state(g, &s);
    memset(s.mem_vz_z, 0, fb * 8);
    for (int it = 0; it < nt; it++) {
        for (int k = 0; k < 13; k++) memcpy(hist + ((size_t)it * 13 + k) * n, *keep[k], fb);
        or_step(g, m, &s, it, ishot, src, NULL, 0, 0, NULL, 0, NULL, NULL, NULL);
    }
    /* adjoint state and per-step intermediates */
    double *abuf = calloc(n * 25, sizeof(double));
    double *bvz = abuf, *bvx = abuf + n, *bzz = abuf + 2 * n, *bzx = abuf + 3 * n, *bxx = abuf + 4 * n;
    double *cv[4] = {abuf + 5 * n, abuf + 6 * n, abuf + 7 * n, abuf + 8 * n};     /* stress-pass carries */
    double *cs[4] = {abuf + 9 * n, abuf + 10 * n, abuf + 11 * n, abuf + 12 * n};  /* velocity-pass carries */
    double *Ev[4] = {abuf + 13 * n, abuf + 14 * n, abuf + 15 * n, abuf + 16 * n}; /* forward strains */
    double *Sz = abuf + 17 * n, *Sx = abuf + 18 * n;
    double *bD[4] = {abuf + 19 * n, abuf + 20 * n, abuf + 21 * n, abuf + 22 * n};
    double *acc = calloc(2 * (size_t)(nrec > 0 ? nrec : 1), sizeof(double));
    memset(g_lam, 0, fb); memset(g_mu, 0, fb); memset(g_muzx, 0, fb); memset(g_rhoz, 0, fb); memset(g_rhox, 0, fb);
    for (int it = nt - 1; it >= 0; it--) {
        /* replay step it up to its intermediates */
        for (int k = 0; k < 13; k++) memcpy(*keep[k], hist + ((size_t)it * 13 + k) * n, fb);
        or_vdiff(g, &s);
        if (g->pml_z || g->pml_x) pml_diff(g, &s, s.mem_vz_z, s.mem_vx_z, s.mem_vz_x, s.mem_vx_x);
        memcpy(Ev[0], s.dz_z, fb); memcpy(Ev[1], s.dx_z, fb); memcpy(Ev[2], s.dz_x, fb); memcpy(Ev[3], s.dx_x, fb);
        update_s(g, m, &s);
        if (g->surf) surf_mirror(g, m, &s);
        or_sdiff(g, &s);
        if (g->pml_z || g->pml_x) pml_diff(g, &s, s.mem_szz_z, s.mem_szx_z, s.mem_szx_x, s.mem_sxx_x);
        for (size_t i = 0; i < n; i++) {
            Sz[i] = s.dz_x[i] + s.dz_z[i];
            Sx[i] = s.dx_x[i] + s.dx_z[i];
        }
        /* urec: d[it] = d[it-1] + rc v; the running sum's adjoint carries sum_{t >= it} cot[t] */
        for (int ir = 0; ir < nrec; ir++) {
            acc[2 * ir] += cot_uz[(size_t)ir * nt + it];
            acc[2 * ir + 1] += cot_ux[(size_t)ir * nt + it];
            AT(bvz, z_rec[ir], x_rec[ir]) += rc * acc[2 * ir];
            AT(bvx, z_rec[ir], x_rec[ir]) += rc * acc[2 * ir + 1];
        }
        /* vsrc: type 0 sources scale by the inverse densities of their cells */
        if (src && src->stf_type == 0)
            for (int is = 0; is < src->nsrc; is++) {
                if (src->shot_to_fire[is] != ishot) continue;
                const int zs = src->z_src[is], xs = src->x_src[is];
                AT(g_rhoz, zs, xs) += AT(bvz, zs, xs) * dt * src->stf_z[(size_t)is * nt + it] / (g->dz * g->dx);
                AT(g_rhox, zs, xs) += AT(bvx, zs, xs) * dt * src->stf_x[(size_t)is * nt + it] / (g->dz * g->dx);
            }
        /* update_v: v += dt rho^-1 S */
        memset(bD[0], 0, fb * 4);
        for (int iz = nz1; iz < nz2; iz++)
            for (int ix = nx1; ix < nx2; ix++) {
                const double az = AT(bvz, iz, ix), ax = AT(bvx, iz, ix);
                AT(g_rhoz, iz, ix) += az * dt * AT(Sz, iz, ix);
                AT(g_rhox, iz, ix) += ax * dt * AT(Sx, iz, ix);
                const double sz = dt * AT(m->rho_zp, iz, ix) * az, sx = dt * AT(m->rho_xp, iz, ix) * ax;
                AT(bD[0], iz, ix) = sz;   /* szz_z */
                AT(bD[2], iz, ix) = sz;   /* szx_x */
                AT(bD[1], iz, ix) = sx;   /* szx_z */
                AT(bD[3], iz, ix) = sx;   /* sxx_x */
            }
        if (g->pml_z || g->pml_x) pml_T(g, bD, cs);
        sdiff_T(g, (const double **)bD, bzz, bzx, bxx);
        /* surf_mirror, then update_s */
        memset(bD[0], 0, fb * 4);
        if (g->surf) surf_T(g, m, (const double **)Ev, bzz, bzx, bxx, bD, g_lam, g_mu);
        for (int iz = nz1; iz < nz2; iz++)
            for (int ix = nx1; ix < nx2; ix++) {
                const double zz = AT(bzz, iz, ix), zx = AT(bzx, iz, ix), xx = AT(bxx, iz, ix);
                const double lam = AT(m->lam, iz, ix), mu = AT(m->mu, iz, ix), mzx = AT(m->mu_zx, iz, ix);
                const double e_zz = AT(Ev[0], iz, ix), e_xx = AT(Ev[3], iz, ix), e_zx = AT(Ev[2], iz, ix) + AT(Ev[1], iz, ix);
                AT(g_lam, iz, ix) += dt * (zz + xx) * (e_xx + e_zz);
                AT(g_mu, iz, ix) += dt * 2.0 * (zz * e_zz + xx * e_xx);
                AT(g_muzx, iz, ix) += dt * zx * e_zx;
                AT(bD[0], iz, ix) += dt * ((lam + 2.0 * mu) * zz + lam * xx);
                AT(bD[3], iz, ix) += dt * (lam * zz + (lam + 2.0 * mu) * xx);
                AT(bD[1], iz, ix) += dt * mzx * zx;
                AT(bD[2], iz, ix) += dt * mzx * zx;
            }
        if (g->pml_z || g->pml_x) pml_T(g, bD, cv);
        vdiff_T(g, (const double **)bD, bvz, bvx);
        /* variables outside the interior are constants of the forward: no adjoint */
        for (int iz = 0; iz < nz; iz++)
            for (int ix = 0; ix < nx; ix++)
                if (iz < nz1 || iz >= nz2 || ix < nx1 || ix >= nx2) {
                    AT(bvz, iz, ix) = AT(bvx, iz, ix) = 0.0;
                    AT(bzz, iz, ix) = AT(bzx, iz, ix) = AT(bxx, iz, ix) = 0.0;
                }
    }
    free(acc);
    free(abuf);
    free(buf);
    free(hist);
}

/* The media gradient carried to lam, mu, rho through mat_av2 (src/cpu/n_kernel_lib_PSV.cpp:
 * 853-922): mu_zx[i][j] = 4 / sum of 1/mu over (i..i+1, j..j+1), so d mu_zx / d mu = mu_zx^2 /
 * (4 mu^2); rho_zp[i][j] = 1 / (0.5 (rho[i][j] + rho[i+1][j])), so d rho_zp / d rho =
 * -0.5 rho_zp^2 for both, and rho_xp over (j, j+1) likewise; the last row and column of the
 * averaged media are not defined (0). */
void or_media_chain(const double *g_lam, const double *g_mu, const double *g_muzx, const double *g_rhoz,
                    const double *g_rhox, const double *mu, const double *mu_zx, const double *rho_zp,
                    const double *rho_xp, int32_t nz, int32_t nx, double *d_lam, double *d_mu, double *d_rho)
{
    for (int iz = 0; iz < nz; iz++)
        for (int ix = 0; ix < nx; ix++) {
            AT(d_lam, iz, ix) = AT(g_lam, iz, ix);
            double dm = AT(g_mu, iz, ix), dr = 0.0;
            const double u = AT(mu, iz, ix);
            for (int a = 0; a <= 1; a++)
                for (int b = 0; b <= 1; b++) {
                    const int pz = iz - a, px = ix - b;
                    if (pz < 0 || px < 0 || pz >= nz - 1 || px >= nx - 1 || u == 0.0) continue;
                    const double q = AT(mu_zx, pz, px);
                    dm += AT(g_muzx, pz, px) * q * q / (4.0 * u * u);
                }
            for (int a = 0; a <= 1; a++) {
                const int pz = iz - a;
                if (pz >= 0 && pz < nz - 1 && ix < nx - 1) dr -= 0.5 * AT(g_rhoz, pz, ix) * AT(rho_zp, pz, ix) * AT(rho_zp, pz, ix);
                const int px = ix - a;
                if (px >= 0 && px < nx - 1 && iz < nz - 1) dr -= 0.5 * AT(g_rhox, iz, px) * AT(rho_xp, iz, px) * AT(rho_xp, iz, px);
            }
            AT(d_mu, iz, ix) = dm;
            AT(d_rho, iz, ix) = dr;
        }
}

/* ========================================================================== */
/* Conjugate-gradient direction and Gaussian gradient smoothing (disabled in    */
/* the reference's loop, src/cpu/n_simulate_PSV.cpp:431-455)                    */
/* ========================================================================== */

/* PCG_PSV (src/cpu/n_PCG_PSV.cpp:13-44): Fletcher-Reeves beta = sum g^2 / sum old^2 with
 * old = -g of the previous call, clipped at 0; dir = -g + beta dir; old = -g; g = dir.  The
 * first call has old = 0, where the reference divides by zero; here beta is 0 then. */
void or_pcg_psv(double *dir, double *old, double *grad, int32_t nz, int32_t nx)
{
    const size_t n = (size_t)nz * nx;
    double bi = 0.0, bj = 0.0;
    for (size_t i = 0; i < n; i++) {
        bi += grad[i] * grad[i];
        bj += old[i] * old[i];
        old[i] = -grad[i];
    }
    double beta = bj > 0.0 ? bi / bj : 0.0;
    beta = beta > 0 ? beta : 0.0;
    for (size_t i = 0; i < n; i++) {
        dir[i] = -grad[i] + beta * dir[i];
        grad[i] = dir[i];
    }
}

/* gauss_filter_kernel (src/cpu/n_kernel_lib_PSV.cpp:1012-1049): sigma = hfs / 2 in integer
 * division, exp(-(i^2 + j^2) / (2 sigma^2)) over [-hfs, hfs]^2, normalised to sum 1. */
void or_gauss_filter_kernel(double *kernel, int32_t hfs)
{
    const int w = 2 * hfs + 1;
    const double sigma = hfs / 2, q = 2.0 * sigma * sigma;
    double sum = 0.0;
    for (int ii = -hfs; ii <= hfs; ii++)
        for (int jj = -hfs; jj <= hfs; jj++) {
            const double v = exp(-((ii * ii) / q) - ((jj * jj) / q));
            kernel[(jj + hfs) * w + ii + hfs] = v;
            sum += v;
        }
    for (int i = 0; i < w * w; i++) kernel[i] /= sum;
}

/* apply_gauss_filter (:1051-1084): the filter over cells [hfs, nz - hfs - 1) x [hfs, nx - hfs - 1)
 * of a copy, other cells kept.  The reference accumulates in an int (every partial sum
 * truncated); the restatement accumulates in double, as its `conv = 0.0` start intends. */
void or_apply_gauss_filter(double *mat, const double *filter, int32_t nz, int32_t nx, int32_t hfs)
{
    const int w = 2 * hfs + 1;
    double *tmp = malloc(sizeof(double) * (size_t)nz * nx);
    memcpy(tmp, mat, sizeof(double) * (size_t)nz * nx);
    for (int iz = hfs; iz < nz - hfs - 1; iz++)
        for (int ix = hfs; ix < nx - hfs - 1; ix++) {
            double conv = 0.0;
            for (int jj = 0; jj < w; jj++)
                for (int ii = 0; ii < w; ii++)
                    conv += tmp[(size_t)(iz + jj - hfs) * nx + ix + ii - hfs] * filter[jj * w + ii];
            mat[(size_t)iz * nx + ix] = conv;
        }
    free(tmp);
}
