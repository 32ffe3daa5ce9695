// psv_preproc.cpp -- host-side set-up of the P/SV path (no device needed).
//
//   fwi_holbergcoeff  holbergcoeff         src/h_holbergcoeff.cpp:21-80
//   fwi_cpml          cpml_PSV             src/h_preproc_PSV.cpp:112-263
//   fwi_wavelet       wavelet              src/h_preproc_PSV.cpp:267-286
//   fwi_dtstab        checkfd_ssg_elastic  src/h_checkfd_ssg_elastic.cpp:14-144
#include <array>
#include <cmath>

#include "fwi_psv.h"

namespace {

// The reference's pi (include/h_globvar.hpp:13); the CPML alpha and the wavelet use it.
constexpr double kPi = 3.14159265;

// Staggered-grid first-derivative coefficients: Taylor, then Holberg's optimised
// sets for 0.1 %, 0.5 %, 1 % and 3 % maximum group-velocity error, for half
// orders 1..6; entry 0 is the points-per-wavelength figure of the design.
// The values are the published DENISE table the reference ships.
using Row = std::array<double, 7>;
const std::array<std::array<Row, 6>, 5> kHolberg = {{
    {{{23.0, 1.0, 0, 0, 0, 0, 0},
      {8.0, 9.0 / 8.0, -1.0 / 24.0, 0, 0, 0, 0},
      {6.0, 75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0, 0, 0, 0},
      {5.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0, 0, 0},
      {5.0, 19845.0 / 16384.0, -735.0 / 8192.0, 567.0 / 40960.0, -405.0 / 229376.0, 35.0 / 294912.0, 0},
      {4.0, 160083.0 / 131072.0, -12705.0 / 131072.0, 22869.0 / 1310720.0, -5445.0 / 1835008.0,
       847.0 / 2359296.0, -63.0 / 2883584.0}}},
    {{{49.7, 1.0010, 0, 0, 0, 0, 0},
      {8.32, 1.1382, -0.046414, 0, 0, 0, 0},
      {4.77, 1.1965, -0.078804, 0.0081781, 0, 0, 0},
      {3.69, 1.2257, -0.099537, 0.018063, -0.0026274, 0, 0},
      {3.19, 1.2415, -0.11231, 0.026191, -0.0064682, 0.001191, 0},
      {2.91, 1.2508, -0.12034, 0.032131, -0.010142, 0.0029857, -0.00066667}}},
    {{{22.2, 1.0050, 0, 0, 0, 0, 0},
      {5.65, 1.1534, -0.052806, 0, 0, 0, 0},
      {3.74, 1.2111, -0.088313, 0.011768, 0, 0, 0},
      {3.11, 1.2367, -0.10815, 0.023113, -0.0046905, 0, 0},
      {2.80, 1.2496, -0.11921, 0.031130, -0.0093272, 0.0025161, 0},
      {2.62, 1.2568, -0.12573, 0.036423, -0.013132, 0.0047484, -0.0015979}}},
    {{{15.8, 1.0100, 0, 0, 0, 0, 0},
      {4.80, 1.1640, -0.057991, 0, 0, 0, 0},
      {3.39, 1.2192, -0.094070, 0.014608, 0, 0, 0},
      {2.90, 1.2422, -0.11269, 0.026140, -0.0064054, 0, 0},
      {2.65, 1.2534, -0.12257, 0.033755, -0.011081, 0.0036784, 0},
      {2.51, 1.2596, -0.12825, 0.038550, -0.014763, 0.0058619, -0.0024538}}},
    {{{9.16, 1.0300, 0, 0, 0, 0, 0},
      {3.47, 1.1876, -0.072518, 0, 0, 0, 0},
      {2.91, 1.2341, -0.10569, 0.022589, 0, 0, 0},
      {2.61, 1.2516, -0.12085, 0.032236, -0.011459, 0, 0},
      {2.45, 1.2596, -0.12829, 0.038533, -0.014681, 0.0072580, 0},
      {2.36, 1.2640, -0.13239, 0.042217, -0.017803, 0.0081959, -0.0051848}}},
}};

// One CPML side: the damping profile d = d0 (x/L)^N, K = 1 + (Kmax - 1)(x/L)^N and
// alpha = alpha_max (1 - x/L) at the integer and the half-integer points of
// indices [i0, i1), with x the distance into the layer measured from `origin`
// (towards lower indices when `inward_low`), turned into the recursive
// convolution coefficients b = exp(-(d/K + alpha) dt), a = d (b - 1) / (K (d + K alpha)).
struct Side {
    double *a, *b, *K, *ah, *bh, *Kh;
    double npower, d0, kmax, alpha_max, thick, dt, dh;

    void point(int i, double absc, double absc_half) const {
        double d = 0, dhf = 0, al = 0, alh = 0;
        if (absc >= 0) {
            const double xn = absc / thick, p = std::pow(xn, npower);
            d = d0 * p;
            K[i] = 1.0 + (kmax - 1.0) * p;
            al = alpha_max * (1.0 - xn);
        }
        if (absc_half >= 0) {
            const double xn = absc_half / thick, p = std::pow(xn, npower);
            dhf = d0 * p;
            Kh[i] = 1.0 + (kmax - 1.0) * p;
            alh = alpha_max * (1.0 - xn);
        }
        if (al < 0.0) al = 0.0;
        if (alh < 0.0) alh = 0.0;
        b[i] = std::exp(-(d / K[i] + al) * dt);
        bh[i] = std::exp(-(dhf / Kh[i] + alh) * dt);
        if (std::fabs(d) > 1.0e-6) a[i] = d * (b[i] - 1.0) / (K[i] * (d + K[i] * al));
        if (std::fabs(dhf) > 1.0e-6) ah[i] = dhf * (bh[i] - 1.0) / (Kh[i] * (dhf + Kh[i] * alh));
    }
};

}  // namespace

extern "C" int32_t fwi_holbergcoeff(int32_t fdorder, int32_t maxrelerror, double hc[7]) {
    if (!hc || fdorder < 2 || fdorder > 12 || fdorder % 2 || maxrelerror < 0 || maxrelerror > 4) return FWI_EINVAL;
    const Row &r = kHolberg[maxrelerror][fdorder / 2 - 1];
    for (int i = 0; i < 7; ++i) hc[i] = r[i];
    return FWI_OK;
}

extern "C" int32_t fwi_cpml(double *a, double *b, double *K, double *a_half, double *b_half, double *K_half,
                            double npower, double damp_v, double rcoef, double k_max, double freq, int32_t npml_h1,
                            int32_t npml_h2, int32_t fpad, int32_t nh, double dt, double dh) {
    if (!a || !b || !K || !a_half || !b_half || !K_half || nh <= 0 || npml_h1 < 0 || npml_h2 < 0 ||
        npml_h1 + fpad >= nh || nh - fpad - npml_h2 - 1 < 0)
        return FWI_EINVAL;
    for (int i = 0; i < nh; ++i) {
        a[i] = a_half[i] = 0.0;
        b[i] = b_half[i] = 1.0;
        K[i] = K_half[i] = 1.0;
    }
    const double alpha_max = 2.0 * kPi * (freq / 2.0);
    if (npml_h1) {
        const double L = npml_h1 * dh;
        Side sd{a, b, K, a_half, b_half, K_half, npower,
                -(npower + 1) * damp_v * std::log(rcoef) / (2.0 * L), k_max, alpha_max, L, dt, dh};
        for (int i = fpad; i <= npml_h1 + fpad; ++i) {
            const double x = dh * double(i);
            sd.point(i, L - x, L - (x + dh / 2.0));
        }
    }
    if (npml_h2) {
        const double L = npml_h2 * dh;
        const double origin = dh * (nh - fpad - npml_h2 - 1);
        Side sd{a, b, K, a_half, b_half, K_half, npower,
                -(npower + 1) * damp_v * std::log(rcoef) / (2.0 * L), k_max, alpha_max, L, dt, dh};
        for (int i = nh - fpad - npml_h2 - 1; i < nh - fpad; ++i) {
            const double x = dh * double(i);
            sd.point(i, x - origin, (x + dh / 2.0) - origin);
        }
    }
    return FWI_OK;
}

extern "C" int32_t fwi_wavelet(double *signal, int32_t nt, double dt, double amp, double fc, double ts,
                               int32_t shape) {
    if (!signal || nt < 0 || shape != 1 || !(fc > 0)) return FWI_EINVAL;
    const double fci = 1.0 / fc;
    for (int it = 0; it < nt; ++it) {
        const double t = it * dt;
        const double tau = kPi * (t - 1.5 * fci - ts) / (1.5 * fci);
        signal[it] = amp * (1.0 - 2.0 * tau * tau) * std::exp(-2.0 * tau * tau);
    }
    return FWI_OK;
}

extern "C" double fwi_dtstab(int32_t nz, int32_t nx, double dh, int32_t fw, const double *rho, const double *lam,
                             const double *mu, const double hc[7], double *cmax_out, double *cmin_out) {
    double cmax = 0.0, cmin = 1e9;
    const int i1 = 1 + fw, i2 = std::min(nx - fw, nx - 1);
    const int j1 = 1, j2 = std::min(nz - fw, nz - 1);
    for (int j = j1; j <= j2; ++j)
        for (int i = i1; i <= i2; ++i) {
            const size_t o = (size_t)j * nx + i;
            if (std::fabs(mu[o]) < 1.0e-9) continue;
            const double cs = std::sqrt(mu[o] / rho[o]);
            const double cp = std::sqrt((lam[o] + 2.0 * mu[o]) / rho[o]);
            cmax = std::max(cmax, std::max(cp, cs));
            cmin = std::min(cmin, std::min(cp, cs));
        }
    double gamma = 0;
    for (int k = 1; k <= 6; ++k) gamma += std::fabs(hc[k]);
    if (cmax_out) *cmax_out = cmax;
    if (cmin_out) *cmin_out = cmin;
    return dh / (std::sqrt(2.0) * gamma * cmax);
}
