// ref_harness.cpp -- C entry points onto the reference's own CPU sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's src/*.cpp and src/cpu/*.cpp (read in place under
// /root/reference, never copied) into oracle/_ref/libref_psv.so.  The entry
// points take flat row-major arrays, build the T** / T*** row tables the
// reference's functions expect (include/h_contiguous_arrays.hpp:37-65) and call
// the reference functions unchanged.  The argument structs are the oracle's
// (oracle/fwi_oracle.h), so one Python-side description of a case drives both
// the reference and the restatement.
#include <vector>
#include <cstddef>

#include "n_kernel_PSV.hpp"
#include "n_kernel_lib_PSV.hpp"
#include "n_step_length_PSV.hpp"
#include "n_PCG_PSV.hpp"
#include "h_preproc_PSV.hpp"
#include "h_holbergcoeff.hpp"

#include "fwi_oracle.h"

namespace {

// Row table over a contiguous [n1][n2] block.
struct Rows {
    std::vector<double *> p;
    double **ptr;
    Rows(double *base, int n1, int n2) : p(n1 > 0 ? n1 : 1, nullptr) {
        for (int i = 0; i < n1; i++) p[i] = base ? base + (size_t)i * n2 : nullptr;
        ptr = p.data();
    }
};

// Row tables over a contiguous [n1][n2][n3] block.
struct Cube {
    std::vector<double *> rows;
    std::vector<double **> planes;
    double ***ptr;
    Cube(double *base, int n1, int n2, int n3)
        : rows((size_t)(n1 > 0 ? n1 : 1) * (n2 > 0 ? n2 : 1), nullptr), planes(n1 > 0 ? n1 : 1, nullptr) {
        for (int i = 0; i < n1; i++) {
            planes[i] = rows.data() + (size_t)i * n2;
            for (int j = 0; j < n2; j++) rows[(size_t)i * n2 + j] = base ? base + ((size_t)i * n2 + j) * n3 : nullptr;
        }
        ptr = planes.data();
    }
};

inline double *mut(const double *p) { return const_cast<double *>(p); }
inline int *mut(const int32_t *p) { return const_cast<int *>(reinterpret_cast<const int *>(p)); }

}  // namespace

extern "C" {

void ref_holbergcoeff(int fdorder, int maxrelerror, double *hc7) { holbergcoeff(fdorder, maxrelerror, hc7); }

void ref_cpml(double *a, double *b, double *K, double *a_half, double *b_half, double *K_half, double npower,
              double damp_v, double rcoef, double k_max, double freq, int npml_h1, int npml_h2, int fpad, int nh,
              double dt, double dh) {
    cpml_PSV(a, b, K, a_half, b_half, K_half, npower, damp_v, rcoef, k_max, freq, npml_h1, npml_h2, fpad, nh, dt, dh);
}

void ref_wavelet(double *signal, int nt, double dt, double amp, double fc, double ts, int shape) {
    wavelet(signal, nt, dt, amp, fc, ts, shape);
}

void ref_mat_av2(double *lam, double *mu, double *rho, double *mu_zx, double *rho_zp, double *rho_xp, double *C,
                 int nz, int nx) {
    Rows L(lam, nz, nx), M(mu, nz, nx), R(rho, nz, nx), MZX(mu_zx, nz, nx), RZ(rho_zp, nz, nx), RX(rho_xp, nz, nx);
    mat_av2(L.ptr, M.ptr, R.ptr, MZX.ptr, RZ.ptr, RX.ptr, C[0], C[1], C[2], nz, nx);
}

void ref_kernel_psv(const or_grid *g, const or_media *m, or_state *s, int ishot, int grad, double *g_lam,
                    double *g_mu, double *g_rho, const or_sources *src, or_receivers *rec, int accu, or_snaps *snap) {
    const int nz = g->nz, nx = g->nx, nt = g->nt;
    const int snz = 1 + (g->snap_z2 - g->snap_z1) / g->snap_dz;
    const int snx = 1 + (g->snap_x2 - g->snap_x1) / g->snap_dx;
    const int snt = 1 + (nt - 1) / g->snap_dt;
    int *isurf = mut(g->isurf);
    double *hc = mut(g->hc);
    Rows vz(s->vz, nz, nx), vx(s->vx, nz, nx), uz(s->uz, nz, nx), ux(s->ux, nz, nx);
    Rows szz(s->szz, nz, nx), szx(s->szx, nz, nx), sxx(s->sxx, nz, nx), We(s->We, nz, nx);
    Rows dzz(s->dz_z, nz, nx), dxz(s->dx_z, nz, nx), dzx(s->dz_x, nz, nx), dxx(s->dx_x, nz, nx);
    Rows lam(mut(m->lam), nz, nx), mu(mut(m->mu), nz, nx), mu_zx(mut(m->mu_zx), nz, nx);
    Rows rho_zp(mut(m->rho_zp), nz, nx), rho_xp(mut(m->rho_xp), nz, nx);
    Rows gl(g_lam, snz, snx), gm(g_mu, snz, snx), gr(g_rho, snz, snx);
    double *a_z = mut(g->a_z), *b_z = mut(g->b_z), *K_z = mut(g->K_z);
    double *a_hz = mut(g->a_half_z), *b_hz = mut(g->b_half_z), *K_hz = mut(g->K_half_z);
    double *a_x = mut(g->a_x), *b_x = mut(g->b_x), *K_x = mut(g->K_x);
    double *a_hx = mut(g->a_half_x), *b_hx = mut(g->b_half_x), *K_hx = mut(g->K_half_x);
    Rows mvzz(s->mem_vz_z, nz, nx), mvxz(s->mem_vx_z, nz, nx), mszzz(s->mem_szz_z, nz, nx), mszxz(s->mem_szx_z, nz, nx);
    Rows mvzx(s->mem_vz_x, nz, nx), mvxx(s->mem_vx_x, nz, nx), mszxx(s->mem_szx_x, nz, nx), msxxx(s->mem_sxx_x, nz, nx);
    Rows stf_z(mut(src->stf_z), src->nsrc, nt), stf_x(mut(src->stf_x), src->nsrc, nt);
    int *z_src = mut(src->z_src), *x_src = mut(src->x_src), *fire = mut(src->shot_to_fire);
    Rows rtf_uz(rec->rtf_uz, rec->nrec, nt), rtf_ux(rec->rtf_ux, rec->nrec, nt);
    int *z_rec = mut(rec->z_rec), *x_rec = mut(rec->x_rec);
    Cube avz(snap ? snap->vz : nullptr, snt, snz, snx), avx(snap ? snap->vx : nullptr, snt, snz, snx);
    Cube aszz(snap ? snap->szz : nullptr, snt, snz, snx), aszx(snap ? snap->szx : nullptr, snt, snz, snx);
    Cube asxx(snap ? snap->sxx : nullptr, snt, snz, snx);
    kernel_PSV(ishot, nt, nz, nx, g->dt, g->dx, g->dz, g->surf != 0, isurf, hc, g->fdorder, vz.ptr, vx.ptr, uz.ptr,
               ux.ptr, szz.ptr, szx.ptr, sxx.ptr, We.ptr, dzz.ptr, dxz.ptr, dzx.ptr, dxx.ptr, lam.ptr, mu.ptr,
               mu_zx.ptr, rho_zp.ptr, rho_xp.ptr, grad != 0, gl.ptr, gm.ptr, gr.ptr, g->pml_z != 0, a_z, b_z, K_z,
               a_hz, b_hz, K_hz, g->pml_x != 0, a_x, b_x, K_x, a_hx, b_hx, K_hx, mvzz.ptr, mvxz.ptr, mszzz.ptr,
               mszxz.ptr, mvzx.ptr, mvxx.ptr, mszxx.ptr, msxxx.ptr, src->nsrc, src->stf_type, stf_z.ptr, stf_x.ptr,
               z_src, x_src, fire, rec->nrec, rec->rtf_type, rtf_uz.ptr, rtf_ux.ptr, z_rec, x_rec, accu != 0,
               avz.ptr, avx.ptr, aszz.ptr, aszx.ptr, asxx.ptr, g->snap_z1, g->snap_z2, g->snap_x1, g->snap_x2,
               g->snap_dt, g->snap_dz, g->snap_dx);
}

// adjsrc2 for one shot: true data [nrec][nt] of that shot, residual written in place into mod_*.
double ref_adjsrc2(double *mod_uz, double *mod_ux, double *true_uz, double *true_ux, double dt, int nseis, int nt) {
    Rows muz(mod_uz, nseis, nt), mux(mod_ux, nseis, nt);
    Cube tuz(true_uz, 1, nseis, nt), tux(true_ux, 1, nseis, nt);
    int *a_type = nullptr;
    return adjsrc2(0, a_type, muz.ptr, mux.ptr, 0, tuz.ptr, tux.ptr, muz.ptr, mux.ptr, dt, nseis, nt);
}

void ref_energy_weights2(double *We, double *We_adj, int nz, int nx, int z1, int z2, int x1, int x2) {
    Rows W(We, nz, nx), WA(We_adj, nz, nx);
    energy_weights2(W.ptr, WA.ptr, z1, z2, x1, x2);
}

void ref_interpol_grad2(double *grad, double *grad_shot, int nz, int nx, int z1, int z2, int x1, int x2, int sdz,
                        int sdx) {
    const int snz = 1 + (z2 - z1) / sdz, snx = 1 + (x2 - x1) / sdx;
    Rows G(grad, nz, nx), GS(grad_shot, snz, snx);
    interpol_grad2(G.ptr, GS.ptr, z1, z2, x1, x2, sdz, sdx);
}

void ref_scale_grad_E2(double *grad, double *grad_shot, double mat_av, double *We, int nz, int nx, int z1, int z2,
                       int x1, int x2) {
    Rows G(grad, nz, nx), GS(grad_shot, nz, nx), W(We, nz, nx);
    scale_grad_E2(G.ptr, GS.ptr, mat_av, W.ptr, z1, z2, x1, x2);
}

void ref_taper2(double *A, int nz, int nx, int z1, int z2, int x1, int x2, const int32_t *taper) {
    Rows R(A, nz, nx);
    int t[8];
    for (int i = 0; i < 8; i++) t[i] = taper[i];
    taper2(R.ptr, nz, nx, z1, z2, x1, x2, t[0], t[1], t[2], t[3], t[4], t[5], t[6], t[7]);
}

void ref_update_mat2(double *mat, double *mat_old, double *grad, double mat_max, double mat_min, double step,
                     int nz, int nx) {
    Rows M(mat, nz, nx), MO(mat_old, nz, nx), G(grad, nz, nx);
    update_mat2(M.ptr, MO.ptr, G.ptr, mat_max, mat_min, step, nz, nx);
}

// step_length_PSV over all shots (update_param 0): trial models are written into
// lam/mu/rho, observed data true_* are [nshot][nrec][nt].
double ref_step_length_psv(double est, double L2_0, int nshot, const or_grid *g, double *lam, double *mu,
                           double *rho, double *lam_copy, double *mu_copy, double *rho_copy, double *mu_zx,
                           double *rho_zp, double *rho_xp, double *g_lam, double *g_mu, double *g_rho, or_state *s,
                           const or_sources *src, or_receivers *rec, double *true_uz, double *true_ux,
                           or_snaps *snap) {
    const int nz = g->nz, nx = g->nx, nt = g->nt;
    const int snz = 1 + (g->snap_z2 - g->snap_z1) / g->snap_dz;
    const int snx = 1 + (g->snap_x2 - g->snap_x1) / g->snap_dx;
    const int snt = 1 + (nt - 1) / g->snap_dt;
    int *isurf = mut(g->isurf);
    double *hc = mut(g->hc);
    Rows vz(s->vz, nz, nx), vx(s->vx, nz, nx), uz(s->uz, nz, nx), ux(s->ux, nz, nx);
    Rows szz(s->szz, nz, nx), szx(s->szx, nz, nx), sxx(s->sxx, nz, nx), We(s->We, nz, nx);
    Rows dzz(s->dz_z, nz, nx), dxz(s->dx_z, nz, nx), dzx(s->dz_x, nz, nx), dxx(s->dx_x, nz, nx);
    Rows L(lam, nz, nx), M(mu, nz, nx), R(rho, nz, nx), LC(lam_copy, nz, nx), MC(mu_copy, nz, nx),
        RC(rho_copy, nz, nx);
    Rows MZX(mu_zx, nz, nx), RZ(rho_zp, nz, nx), RX(rho_xp, nz, nx);
    Rows GL(g_lam, nz, nx), GM(g_mu, nz, nx), GR(g_rho, nz, nx);
    double *a_z = mut(g->a_z), *b_z = mut(g->b_z), *K_z = mut(g->K_z);
    double *a_hz = mut(g->a_half_z), *b_hz = mut(g->b_half_z), *K_hz = mut(g->K_half_z);
    double *a_x = mut(g->a_x), *b_x = mut(g->b_x), *K_x = mut(g->K_x);
    double *a_hx = mut(g->a_half_x), *b_hx = mut(g->b_half_x), *K_hx = mut(g->K_half_x);
    Rows mvzz(s->mem_vz_z, nz, nx), mvxz(s->mem_vx_z, nz, nx), mszzz(s->mem_szz_z, nz, nx), mszxz(s->mem_szx_z, nz, nx);
    Rows mvzx(s->mem_vz_x, nz, nx), mvxx(s->mem_vx_x, nz, nx), mszxx(s->mem_szx_x, nz, nx), msxxx(s->mem_sxx_x, nz, nx);
    Rows stf_z(mut(src->stf_z), src->nsrc, nt), stf_x(mut(src->stf_x), src->nsrc, nt);
    int *z_src = mut(src->z_src), *x_src = mut(src->x_src), *fire = mut(src->shot_to_fire);
    Rows rtf_uz(rec->rtf_uz, rec->nrec, nt), rtf_ux(rec->rtf_ux, rec->nrec, nt);
    int *z_rec = mut(rec->z_rec), *x_rec = mut(rec->x_rec);
    Cube tuz(true_uz, nshot, rec->nrec, nt), tux(true_ux, nshot, rec->nrec, nt);
    Cube avz(snap->vz, snt, snz, snx), avx(snap->vx, snt, snz, snx), aszz(snap->szz, snt, snz, snx),
        aszx(snap->szx, snt, snz, snx), asxx(snap->sxx, snt, snz, snx);
    return step_length_PSV(est, L2_0, nshot, nt, nz, nx, g->dt, g->dx, g->dz, g->surf != 0, isurf, hc, g->fdorder,
                           vz.ptr, vx.ptr, uz.ptr, ux.ptr, szz.ptr, szx.ptr, sxx.ptr, We.ptr, dzz.ptr, dxz.ptr,
                           dzx.ptr, dxx.ptr, L.ptr, M.ptr, R.ptr, LC.ptr, MC.ptr, RC.ptr, MZX.ptr, RZ.ptr, RX.ptr,
                           false, GL.ptr, GM.ptr, GR.ptr, g->pml_z != 0, a_z, b_z, K_z, a_hz, b_hz, K_hz,
                           g->pml_x != 0, a_x, b_x, K_x, a_hx, b_hx, K_hx, mvzz.ptr, mvxz.ptr, mszzz.ptr, mszxz.ptr,
                           mvzx.ptr, mvxx.ptr, mszxx.ptr, msxxx.ptr, src->nsrc, src->stf_type, stf_z.ptr, stf_x.ptr,
                           z_src, x_src, fire, rec->nrec, rec->rtf_type, rtf_uz.ptr, rtf_ux.ptr, z_rec, x_rec,
                           tuz.ptr, tux.ptr, true, avz.ptr, avx.ptr, aszz.ptr, aszx.ptr, asxx.ptr, g->snap_z1,
                           g->snap_z2, g->snap_x1, g->snap_x2, g->snap_dt, g->snap_dz, g->snap_dx, 0);
}

// PCG_PSV (src/cpu/n_PCG_PSV.cpp:13-44) on one parameter's arrays [nz][nx]
void ref_pcg_psv(double *dir, double *old, double *grad, int nz, int nx) {
    Rows d(dir, nz, nx), o(old, nz, nx), g(grad, nz, nx);
    PCG_PSV(d.ptr, o.ptr, g.ptr, nz, nx);
}

// gauss_filter_kernel (src/cpu/n_kernel_lib_PSV.cpp:1012-1049): [2 hfs + 1][2 hfs + 1]
void ref_gauss_filter_kernel(double *kernel, int hfs) {
    Rows k(kernel, 2 * hfs + 1, 2 * hfs + 1);
    gauss_filter_kernel(k.ptr, hfs);
}

// apply_gauss_filter (:1051-1084) as written (its accumulator is an int)
void ref_apply_gauss_filter(double *mat, double *mat_temp, double *filter, int nz, int nx, int hfs) {
    Rows m(mat, nz, nx), t(mat_temp, nz, nx), f(filter, 2 * hfs + 1, 2 * hfs + 1);
    apply_gauss_filter(m.ptr, t.ptr, f.ptr, nz, nx, hfs);
}

}  // extern "C"
