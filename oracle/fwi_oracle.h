/*
 * fwi_oracle.h -- FP64 CPU restatement of the reference's elastic P/SV FWI path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (fwi_cuda_amd / libfwi_psv.so) never links or calls it.
 *
 * Layout: every 2-D field is a flat row-major [nz][nx] double array (x fastest,
 * z = depth = row, top row 0), the layout of include/h_contiguous_arrays.hpp:55-65
 * in the reference.  Snapshot cubes are [snap_nt][snap_nz][snap_nx].
 *
 * Stencil order: fdorder = 2N, N = 1..6.  At N = 1 every routine reduces to the
 * reference's order-2 code operation by operation (same FP64 operation order),
 * which the golden-vector tests check.  For N > 1 the derivative is the
 * standard staggered sum over k = 1..N of hc[k] * (f[i+k-1] - f[i-k]) (or the
 * forward-shifted form), the generalisation the reference's hc table is made
 * for but its kernels never apply (src/cpu/n_kernel_lib_PSV.cpp:125-128 use
 * hc[1] only).
 */
#ifndef FWI_ORACLE_H
#define FWI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid, stencil, boundaries and snapshot box: what simulate_PSV hands to
 * kernel_PSV (src/cpu/n_kernel_PSV.cpp:20-55). */
typedef struct {
    int32_t nt, nz, nx;
    double dt, dz, dx;
    int32_t fdorder;
    double hc[7];
    int32_t surf;
    int32_t isurf[4];
    int32_t pml_z, pml_x;
    const double *a_z, *b_z, *K_z, *a_half_z, *b_half_z, *K_half_z;
    const double *a_x, *b_x, *K_x, *a_half_x, *b_half_x, *K_half_x;
    int32_t snap_z1, snap_z2, snap_x1, snap_x2, snap_dt, snap_dz, snap_dx;
} or_grid;

/* Medium on the staggered grid (mat_av2 outputs). */
typedef struct {
    const double *lam, *mu, *mu_zx, *rho_zp, *rho_xp;
} or_media;

/* Wavefield state, derivative scratch and CPML memory variables
 * (the buffer inventory of src/cpu/n_alloc_PSV.cpp:129-165). */
typedef struct {
    double *vz, *vx, *uz, *ux, *szz, *szx, *sxx, *We;
    double *dz_z, *dx_z, *dz_x, *dx_x;
    double *mem_vz_z, *mem_vx_z, *mem_szz_z, *mem_szx_z;
    double *mem_vz_x, *mem_vx_x, *mem_szx_x, *mem_sxx_x;
} or_state;

/* Point sources (src/cpu/n_kernel_lib_PSV.cpp:487-536). */
typedef struct {
    int32_t nsrc, stf_type;
    const double *stf_z, *stf_x;     /* [nsrc][nt] */
    const int32_t *z_src, *x_src, *shot_to_fire;
} or_sources;

/* Receivers (src/cpu/n_kernel_lib_PSV.cpp:538-576). */
typedef struct {
    int32_t nrec, rtf_type;
    double *rtf_uz, *rtf_ux;         /* [nrec][nt] */
    const int32_t *z_rec, *x_rec;
} or_receivers;

/* Forward snapshots (accu_*) [snap_nt][snap_nz][snap_nx]. */
typedef struct {
    double *vz, *vx, *szz, *szx, *sxx;
} or_snaps;

/* ---- host preprocessing ---------------------------------------------------- */
void or_holbergcoeff(int32_t fdorder, int32_t maxrelerror, double hc[7]);
void or_taylorcoeff(int32_t n, double c[7]);
void or_cpml(double *a, double *b, double *K, double *a_half, double *b_half, double *K_half,
             double npower, double damp_v, double rcoef, double k_max, double freq,
             int32_t npml_h1, int32_t npml_h2, int32_t fpad, int32_t nh, double dt, double dh);
void or_wavelet(double *signal, int32_t nt, double dt, double amp, double fc, double ts, int32_t shape);
double or_checkfd_dtstab(int32_t nz, int32_t nx, double dh, int32_t fw,
                         const double *rho, const double *lam, const double *mu, const double *hc,
                         double *cmax_out, double *cmin_out);

/* ---- medium ------------------------------------------------------------------ */
void or_mat_av2(const double *lam, const double *mu, const double *rho,
                double *mu_zx, double *rho_zp, double *rho_xp, double C[3], int32_t nz, int32_t nx);
void or_copy_mat(double *dst, const double *src, int32_t nz, int32_t nx);
void or_update_mat2(double *mat, const double *mat_old, const double *grad, double mat_max, double mat_min,
                    double step_length, int32_t nz, int32_t nx, double *step_factor_out);

/* ---- time stepping ------------------------------------------------------------ */
void or_reset_state(const or_grid *g, or_state *s);
void or_vdiff(const or_grid *g, or_state *s);
void or_sdiff(const or_grid *g, or_state *s);
void or_step(const or_grid *g, const or_media *m, or_state *s, int32_t it, int32_t ishot,
             const or_sources *src, or_receivers *rec, int32_t record,
             int32_t accu, or_snaps *snap,
             int32_t grad, double *g_lam, double *g_mu, double *g_rho);
void or_kernel_psv(const or_grid *g, const or_media *m, or_state *s, int32_t ishot,
                   int32_t grad, double *g_lam, double *g_mu, double *g_rho,
                   const or_sources *src, or_receivers *rec,
                   int32_t accu, or_snaps *snap);

/* ---- misfit, preconditioning, taper --------------------------------------------- */
double or_adjsrc2(double *a_uz, double *a_ux, const double *mod_uz, const double *mod_ux,
                  const double *true_uz, const double *true_ux, double dt, int32_t nseis, int32_t nt);
double or_energy_weights2(double *We, const double *We_adj, int32_t nx,
                          int32_t snap_z1, int32_t snap_z2, int32_t snap_x1, int32_t snap_x2);
void or_interpol_grad2(double *grad, const double *grad_shot, int32_t nx,
                       int32_t snap_z1, int32_t snap_z2, int32_t snap_x1, int32_t snap_x2,
                       int32_t snap_dz, int32_t snap_dx);
void or_scale_grad_E2(double *grad, const double *grad_shot, double mat_av, const double *We, int32_t nx,
                      int32_t snap_z1, int32_t snap_z2, int32_t snap_x1, int32_t snap_x2);
void or_taper2(double *A, int32_t nz, int32_t nx, int32_t snap_z1, int32_t snap_z2, int32_t snap_x1,
               int32_t snap_x2, const int32_t taper[8]);

/* ---- exact discrete adjoint --------------------------------------------------------- */
/* Reverse mode through or_step for one shot: cot_uz / cot_ux [nrec][nt] are the
 * cotangents of the receivers' displacement seismograms (dL/d seismogram); out the
 * gradient of L with respect to the media lam, mu, mu_zx, rho_zp, rho_xp [nz][nx].  The
 * forward states of every step are kept (5 fields and 8 memories per step). */
void or_adjoint_exact(const or_grid *g, const or_media *m, int32_t ishot, const or_sources *src,
                      int32_t nrec, const int32_t *z_rec, const int32_t *x_rec, const double *cot_uz,
                      const double *cot_ux, double *g_lam, double *g_mu, double *g_muzx, double *g_rhoz,
                      double *g_rhox);
/* The media gradient carried to lam, mu, rho through mat_av2's averages. */
void or_media_chain(const double *g_lam, const double *g_mu, const double *g_muzx, const double *g_rhoz,
                    const double *g_rhox, const double *mu, const double *mu_zx, const double *rho_zp,
                    const double *rho_xp, int32_t nz, int32_t nx, double *d_lam, double *d_mu, double *d_rho);

/* ---- conjugate-gradient direction, Gaussian smoothing --------------------------------- */
void or_pcg_psv(double *dir, double *old, double *grad, int32_t nz, int32_t nx);
void or_gauss_filter_kernel(double *kernel, int32_t hfs);
void or_apply_gauss_filter(double *mat, const double *filter, int32_t nz, int32_t nx, int32_t hfs);

/* ---- step length --------------------------------------------------------------- */
double or_calc_opt_step(const double L2[3], const double sl[3]);

#ifdef __cplusplus
}
#endif
#endif
