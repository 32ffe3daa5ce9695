// psv_internal.h -- shared declarations of the gfx950 P/SV library.
//
// Device data layout (one context, a batch of B shots advanced together):
//   float fields [B][nz][P]      vz vx szz szx sxx uz ux We We_adj; P = nx + 1 rounded up to 64
//                                (column P - 1 is never read: stores of halo lanes land there)
//   float psi_z  [B][nzs][P]     CPML memory of the z-derivatives, only the nzs rows
//                                where the z profile is active (zmap[iz] >= 0)
//   float psi_x  [B][nz][nxs]    CPML memory of the x-derivatives, only the nxs active columns
//   float media  [nz][P]         lam mu mu_zx rho_zp rho_xp (shared by all shots)
//   float snaps  [B][snt][snz][snx] x5 (full storage) or a checkpoint segment buffer
//   float grad   [B][snz][snx] x3  raw correlation sums of the batch's shots
//   double model [nz][nx]        lam mu rho and the saved copies; global gradient x3
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "fwi_psv.h"

namespace psv {

constexpr int kPitchAlign = 64;   // floats: rows start on 256-byte boundaries
constexpr int kBX = 64;           // stencil block: one wavefront across x ...
constexpr int kBY = 4;            // ... times four z-segments

// CPML profile of one grid line, packed so one pointer serves the one-pass step:
// b, a, 1/K at the integer point and at the half point, and the strip index
// (row of psi_z / column of psi_x, -1 outside the strips).
struct CpmlPoint {
    float b, a, ik, bh, ah, ikh;
    int map, pad;
};

// CPML z profiles of the strip rows, passed by value to the variant-5 stencil
// kernels so that a wave-uniform row reads them as scalar loads.  Interior rows
// [N, nz - N) are outside the z strips for lo <= z < hi and inside them
// otherwise; strip row r (psi_z row) holds b, a, 1/K at the integer point, then
// at the half point.
constexpr int kZTabRows = 64;
struct ZTab {
    int lo, hi;
    float p[kZTabRows][6];
};

// What changes from one time step to the next, for launches captured in a hipGraph
// and replayed: indexed by the time index it, built once per context and pass.
struct StepInfo {
    int it;          // time index
    int grad;        // correlate on this step (it % snap_dt == 0, adjoint pass)
    int store;       // store a snapshot on this step (forward / recomputation passes)
    int pad;
    long long off;   // offset of this step's slab in the snapshot arrays
};

// Everything a stencil / injection kernel reads or writes, passed by value.
struct KArgs {
    int nz, nx, P, B;
    int order;                 // work order of the stencil launches (tuned)
    size_t fs;                 // field stride per shot = nz * P
    float dt;
    float cz[6], cx[6];        // hc[k] / dz, hc[k] / dx
    // wavefield of the batch
    float *vz, *vx, *szz, *szx, *sxx, *uz, *ux, *We;
    // media
    const float *lam, *mu, *muzx, *rhoz, *rhox;
    // CPML: maps and profiles (b, a, 1/K) for integer and half points
    const int *zmap, *xmap;
    int nzs, nxs;
    const float *bz, *az, *ikz, *bhz, *ahz, *ikhz;
    const float *bx, *ax, *ikx, *bhx, *ahx, *ikhx;
    const CpmlPoint *zp, *xp;                 // the same profiles and maps, packed [nz], [nx]
    float *pvz_z, *pvx_z, *pszz_z, *pszx_z;   // [B][nzs][P]
    float *pvx_x, *pvz_x, *psxx_x, *pszx_x;   // [B][nz][nxs]
    // velocity-kernel options
    int energy;                // We += vx^2 + vz^2
    int displ;                 // u += dt v (adjoint)
    int grad;                  // correlate with the snapshot slab this step
    float gw;                  // grad_sign * snap_dt * dt
    float gain_lm;             // power of two on the lam / mu terms, keeps FP32 sums off denormals
    const float *sn_vz, *sn_vx, *sn_szz, *sn_szx, *sn_sxx;  // slab of shot 0 at this step
    size_t sn_stride;          // between shots
    float *g_lam, *g_mu, *g_rho;  // [B][snz][snx]
    // gform 0: the reference's sums (fwi_grad2).  gform 1: the terms of the misfit
    // derivative kept apart for the staggered-grid chain rule on the host:
    // g_lam = s1, g_mu = s1 + s2, g_rho = dt vz'.Vz, g_ext = s3, g_ext + g_ext_stride = dt vx'.Vx
    int gform;
    float *g_ext;
    size_t g_ext_stride;
    int snap_z1, snap_z2, snap_x1, snap_x2, snap_dz, snap_dx, snz, snx;
    // one-pass step (k_step): fields and stress-side psi of the next time level go
    // to a second buffer set; the velocity-side psi is updated in place
    float *o_vz, *o_vx, *o_szz, *o_szx, *o_sxx;
    float *o_pvz_z, *o_pvx_z, *o_pvx_x, *o_pvz_x;
    int surf_top;              // free-surface row handled inside k_step (== N), or -1
    int surf5;                 // the same inside the stress kernel of variants 5, 10-13 (== N), or -1
    // snapshot of this step (gard_fwd_storage2) written by k_step and patched by k_inject
    int store;
    float *st_vz, *st_vx, *st_szz, *st_szx, *st_sxx;  // slab of shot 0
    size_t st_stride;          // between shots
    // Graph replay: with si set, the time index is *si_base + si_dir * si_k (si_base is
    // advanced between replays), and grad / store / the slab offsets come from si[it]
    // (the sn_* and st_* pointers then name the start of the snapshot arrays).
    const StepInfo *si;
    const int *si_base;
    int si_k, si_dir;
};

// Point injection / recording (vsrc2, urec2) for the batch.
struct SrcArgs {
    int shot0;                 // first shot of the batch (grp_shot is absolute)
    int ngroups;               // cells with at least one source, over the batch
    const int *grp_shot;       // shot of each group
    const int *grp_cell;       // iz * P + ix
    const int *grp_off;        // [ngroups + 1] into grp_idx
    const int *grp_idx;        // source rows into stf (absolute)
    const double *stf_z, *stf_x;  // [rows][nt]
    int nt, type;
    double dt, dzdx;
    const double *scale_shot;  // per-shot amplitude factor (adjoint rescaling), null = 1
};

struct RecArgs {
    int shot0;                 // first shot of the batch (rec_shot is absolute)
    int nrec;                  // receivers of the batch
    const int *rec_shot;       // shot of each receiver
    const int *rec_cell;       // iz * P + ix
    double *acc_z, *acc_x;     // running sums [nrec]
    double *seis_z, *seis_x;   // [nrec][nt] (rows of the batch's first receiver onwards)
    int nt;
    double dt, dzdx;
};

// ---- launchers (psv_kernels.hip) ------------------------------------------------
// var 1: x taps as shifted global loads; var 2: x taps from neighbouring lanes, next row
// prefetched; var 3: var 2 at six waves per SIMD; var 5: var 2 with wave-uniform addressing
// and ordered loads (needs zt)
void launch_stress(const KArgs &a, int N, int tz, int var, hipStream_t st, const ZTab *zt = nullptr,
                   int depth = 1);
void launch_velocity(const KArgs &a, int N, int tz, int var, hipStream_t st, const ZTab *zt = nullptr,
                     int depth = 1);
// Variant 6 (psv_stencil6.hip): every load issued `depth` (1 or 2) rows ahead.  Its
// rows per segment are one or two blocks of v6_block(N, depth) rows; v6_rows gives
// the segment height launch6 uses for a requested tz.  which: 0 stress, 1 velocity;
// false (nothing launched) for a velocity launch mode it has no kernel for.
constexpr int v6_block(int N, int D) { return ((2 * N + D + D) / (D + 1)) * (D + 1); }
int v6_rows(int N, int depth, int tz);
bool launch6(int which, const KArgs &a, const ZTab &zt, int N, int depth, int tz, hipStream_t st);
// Variant 7 (psv_step7.hip): the one-pass step with variant 6's addressing and row-ahead
// loads; segments of four or eight blocks of 2N + 2 rows (tz >= 8 (2N + 2) picks eight);
// xw sets a block's wavefronts side by side in x (variant 8) instead of one above the
// other in z.  Same conditions as launch_step (top free surface at row N only, interior
// sources); false for a launch mode it has no kernel for.
bool launch_step7(const KArgs &a, const ZTab &zt, int N, int tz, bool xw, hipStream_t st);
// launch_step7's kernels for N = 4, 5, 6 and for N = 5, 6 on a grid it has sized
// (psv_step7_b.hip, psv_step7_c.hip)
bool launch_step7_high(const KArgs &a, const ZTab &zt, int N, int M, bool xw, dim3 grid, int mult, hipStream_t st);
bool launch_step7_top(const KArgs &a, const ZTab &zt, int N, int M, bool xw, dim3 grid, int mult, hipStream_t st);
// A half-step (stress, or velocity) of stencil variant 5, 10, 11, 12 or 13 on the given grid
// (psv_v5.hip, psv_v11.hip, psv_v13.hip); false: a velocity launch mode it has no kernel for.
bool launch_stencil5_v5(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st);
bool launch_stencil5_v10(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st);
bool launch_stencil5_v11(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st);
bool launch_stencil5_v12(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st);
bool launch_stencil5_v13(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st);
// stress, top free surface and velocity of one time step in one pass (kernel variant 4);
// tz = rows per segment (32, 64), depth = rows loaded ahead (1 is the one instantiated)
void launch_step(const KArgs &a, int N, int tz, int depth, hipStream_t st);
void launch_surface(const KArgs &a, int N, const int isurf[4], hipStream_t st);
void launch_inject(const KArgs &a, const SrcArgs &s, int it, hipStream_t st);
void launch_record(const KArgs &a, const RecArgs &r, int it, hipStream_t st);
void launch_set_int(int *p, int v, hipStream_t st);   // *p = v, as a kernel
void launch_store(const KArgs &a, float *const dst[5], size_t dst_stride, float dt, hipStream_t st);

// plain reference-structured path (psv_plain.hip)
struct PlainBufs {
    float *dz_z, *dx_z, *dz_x, *dx_x;                 // [nz][P] derivative scratch
    float *mvz_z, *mvx_z, *mszz_z, *mszx_z;           // [nz][P] full-grid CPML memory
    float *mvz_x, *mvx_x, *mszx_x, *msxx_x;
    const float *fbz, *faz, *fikz, *fbhz, *fahz, *fikhz;   // full-length profiles
    const float *fbx, *fax, *fikx, *fbhx, *fahx, *fikhx;
};
void plain_step(const KArgs &a, const PlainBufs &p, int N, const int isurf[4], int pml_z, int pml_x,
                hipStream_t st);
void plain_grad(const KArgs &a, hipStream_t st);

// ---- exact discrete adjoint (psv_exact.hip) -------------------------------------------
// One batch of B shots, every array [B][nz][P] unless noted.  The forward kernels keep per
// step and cell e_zz, e_xx, e_zx (strain rates of the stress update) and S_z, S_x (stress
// divergences of the velocity update) in hist; the reverse kernels apply the transposed
// step and add the media gradient (lam, mu, mu_zx, rho_zp, rho_xp) into G in FP64.
struct XArgs {
    int nz, nx, P, B, N, pml_z, pml_x;
    size_t fs;                             // nz * P
    float dt;
    float cz[6], cx[6];                    // hc[k] / dz, hc[k] / dx
    const float *lam, *mu, *muzx, *rhoz, *rhox;   // [nz][P]
    const CpmlPoint *zp, *xp;
    float *vz, *vx, *szz, *szx, *sxx;      // forward state
    float *pv[4], *ps[4];                  // forward CPML memories, stress / velocity pass (slots
                                           // d_z vz|szz, d_z vx|szx, d_x vz|szx, d_x vx|sxx)
    float *hist;                           // [steps][B][5][nz][P]: step n at slot n - hist_base
    int hist_base;
    float *bvz, *bvx, *bzz, *bzx, *bxx;    // adjoint state
    float *cv[4], *cs[4];                  // adjoint CPML carries, stress / velocity pass
    float *E[4], *D[4], *Dv[4];            // scratch (may be one set): forward strain rates of a
                                           // step; velocity-pass and stress-pass derivative adjoints
    double *G;                             // [B][5][nz][P]
};
void x_forward_step(const XArgs &x, int n, const int isurf[4], hipStream_t st);
// type0: the shots' sources when they are type 0 (their amplitude carries rho^-1), else null
void x_adjoint_step(const XArgs &x, int n, const int isurf[4], const SrcArgs *type0, hipStream_t st);
// A = rc x the reverse cumulative sum of dt x adj, per receiver row [rows][nt] (may be in place)
void x_revcumsum(const double *adj_z, const double *adj_x, double *A_z, double *A_x, int rows, int nt, double dt,
                 double rc, hipStream_t st);
// G summed over nb shots and carried to lam, mu, rho on the box [z1, z2] x [x1, x2] (0 elsewhere)
void x_media_chain(const XArgs &x, int nb, const double *mu, const double *rho, int z1, int z2, int x1, int x2,
                   double *d_lam, double *d_mu, double *d_rho, hipStream_t st);

// ---- the FP64 propagation for seismograms (psv_fp64.hip) ------------------------------
// The forward step of the exact path's kernels in FP64 (one thread per cell), with its own
// FP64 media and CPML profiles, for the seismograms that the misfit and the adjoint
// sources are made of when fwi_opts.seis_fp64 is set.
struct DArgs {
    int nz, nx, P, B, N, pml_z, pml_x;
    size_t fs;                                        // nz * P
    double dt;
    double cz[6], cx[6];                              // hc[k] / dz, hc[k] / dx
    const double *lam, *mu, *muzx, *rhoz, *rhox;      // [nz][P]
    const double *zp, *xp;                            // [nz][6], [nx][6]: b, a, 1/K, then at the half point
    double *vz, *vx, *szz, *szx, *sxx;                // [B][nz][P]
    double *pv[4], *ps[4];                            // CPML memories, stress / velocity pass
    double *E0, *E3;                                  // d_z vz, d_x vx after CPML (the surfaces use them)
};
// media from the FP64 model: lam, mu as they are, mat_av2's mu_zx, rho_zp, rho_xp
void d_media(const double *lam, const double *mu, const double *rho, double *dlam, double *dmu, double *dmuzx,
             double *drhoz, double *drhox, int nz, int nx, int P, hipStream_t st);
void d_forward_step(const DArgs &d, const int isurf[4], hipStream_t st);
void d_inject(const DArgs &d, const SrcArgs &sa, int it, hipStream_t st);
void d_record(const DArgs &d, const RecArgs &ra, int it, hipStream_t st);

// ---- small kernels on the model and gradient (psv_misc.hip) ------------------------
// mat_av2: float lam, mu and averaged media [nz][P] from the FP64 model; the means of lam,
// mu, rho over (nz - 1)(nx - 1) cells into means[3], summed in a fixed order from
// mat_av_parts(nz, nx) <= kMatAvMaxParts block partial triples in part.
constexpr int kMatAvMaxParts = 512;
int mat_av_parts(int nz, int nx);
void dev_mat_av_into(const double *lam, const double *mu, const double *rho, float *flam, float *fmu, float *muzx,
                     float *rhoz, float *rhox, double *part, double *means, int nz, int nx, int P, hipStream_t st);
void dev_mat_av(const double *lam, const double *mu, const double *rho, float *flam, float *fmu,
                float *muzx, float *rhoz, float *rhox, double *sums3, int nz, int nx, int P,
                hipStream_t st);
void dev_residual(const double *seis_z, const double *seis_x, const double *obs_z, const double *obs_x,
                  double *res_z, double *res_x, const int *rec_off_local, int nshots, int nt, double dt,
                  double *misfit, hipStream_t st);
void dev_absmax(const double *x, size_t n, double *out, hipStream_t st);
void dev_shot_absmax(const double *a, const double *b, const int *rec_off_local, int nshots, int nt, double *out,
                     hipStream_t st);
void dev_precondition(const float *We, const float *Wa, double wa_scale, const double *const g[3],
                      const double g_scale[3], double *const G[3], const double C[3], double *tmp,
                      double *wmax, int nz, int nx, int P, int z1, int z2, int x1, int x2, int sdz,
                      int sdx, int snz, int snx, hipStream_t st);
void dev_taper(double *A, int nz, int nx, int z1, int z2, int x1, int x2, const int taper[8],
               hipStream_t st);
void dev_update(double *mat, const double *old, const double *G, double factor, double mat_max,
                double mat_min, size_t n, hipStream_t st);
// FWI_PARAM_VELOCITY: cp, cs of a model; the cp / cs gradient by the chain rule at fixed
// rho; update_mat2 on cp and cs (f_cp, f_cs = step x update factor) and the new lam, mu.
void dev_velocity_of(const double *lam, const double *mu, const double *rho, double *cp, double *cs, size_t n,
                     hipStream_t st);
void dev_velocity_gradient(const double *G_lam, const double *G_mu, const double *rho, const double *cp,
                           const double *cs, double *g_cp, double *g_cs, size_t n, hipStream_t st);
void dev_velocity_update(double *lam, double *mu, double *rho, const double *rho_old, const double *cp_old,
                         const double *cs_old, const double *g_cp, const double *g_cs, double f_cp, double f_cs,
                         const double bounds[4], size_t n, hipStream_t st);
// misfit derivative: one shot's five FP64 totals added into the sums over shots raw[5][snz][snx],
// divided by the shot's alpha (and the lam / mu terms by gain); then the staggered chain rule
// to dJ/d(lam, mu, rho) [nz][nx] with sign sg
void dev_acc_raw64(double *raw, const double *const g[5], size_t cells, const double *alpha, double gain,
                   hipStream_t st);
// the FP32 correlation sums g32[0..nk) ([n] active cells, stride cells in g64) into g64, zeroed
void dev_flush_sums(double *g64, float *const g32[5], int nk, size_t n, size_t cells, hipStream_t st);
void dev_flush_sums_at(double *g64, float *const g32[5], int nk, size_t n, size_t cells, const int *base, int dir,
                       int k, int nt, int every, hipStream_t st);
void dev_misfit_chain(const double *raw, const double *mu, double *d_lam, double *d_mu, double *d_rho, int nz,
                      int nx, int z1, int x1, int snz, int snx, double sg, hipStream_t st);
void dev_copy(double *dst, const double *src, size_t n, hipStream_t st);   // by a kernel
// apply_gauss_filter of src into dst with the host-built filter [2 hfs + 1]^2 (device)
void dev_gauss(const double *src, double *dst, const double *filt, int nz, int nx, int hfs, hipStream_t st);
// PCG_PSV on one parameter: beta from sum g^2 / sum old^2 (scratch: 1 + 2 x 512 doubles)
void dev_pcg(double *g, double *dir, double *old, size_t n, double *scratch, hipStream_t st);
void dev_count_nonfinite(const float *x, size_t n, int *out, hipStream_t st);
void dev_fill(float *p, float v, size_t n, hipStream_t st);

// the message fwi_last_error(NULL) returns (calls without a context)
void set_global_error(const char *msg);

}  // namespace psv
