/*
 * fwi_psv.h -- C ABI of the MI355X (gfx950) elastic P/SV FWI hot path.
 *
 * One library, libfwi_psv.so, built from the HIP sources in fwi_cuda_amd/csrc.  Plain C
 * types only: host pointers to row-major arrays, sizes, and an opaque context.
 *
 * What each entry point replaces in the reference (Mohammad-debug/FWI_CUDA):
 *
 *   fwi_holbergcoeff      holbergcoeff            src/h_holbergcoeff.cpp:21-80
 *   fwi_cpml              cpml_PSV                src/h_preproc_PSV.cpp:112-263
 *   fwi_wavelet           wavelet                 src/h_preproc_PSV.cpp:267-286
 *   fwi_dtstab            checkfd_ssg_elastic     src/h_checkfd_ssg_elastic.cpp:14-144
 *   fwi_create            alloc_varmain_PSV + the host->device copy
 *                                                 src/cpu/n_alloc_PSV.cpp:96-216,
 *                                                 src/cuda/d_preproc.cu:97-202
 *   fwi_set_model         mat_av2 on a new model  src/cpu/n_kernel_lib_PSV.cpp:853-922
 *   fwi_forward           kernel_PSV, forward     src/cpu/n_kernel_PSV.cpp:20-213 (grad = false)
 *                         (per shot as in simulate_fwd_PSV, src/cpu/n_simulate_PSV.cpp:105-147)
 *   fwi_gradient          one shot of the FWI loop: kernel_PSV forward, adjsrc2,
 *                         kernel_PSV adjoint, energy_weights2, interpol_grad2,
 *                         scale_grad_E2           src/cpu/n_simulate_PSV.cpp:270-418
 *   fwi_gradient_raw      kernel_PSV adjoint with caller-given adjoint sources,
 *                         fwi_grad2 sums          src/cpu/n_kernel_lib_PSV.cpp:439-484
 *   fwi_taper_gradient    taper2 x3               src/cpu/n_kernel_lib_PSV.cpp:942-1010
 *   fwi_save_model        copy_mat                src/cpu/n_kernel_lib_PSV.cpp:834-851
 *   fwi_update_model      update_mat2 x3          src/cpu/n_simulate_PSV.cpp:542-546
 *   fwi_misfit            trial forwards + adjsrc2 sum
 *                                                 src/cpu/n_step_length_PSV.cpp:97-122
 *   fwi_step_length       step_length_PSV + calc_opt_step
 *                                                 src/cpu/n_step_length_PSV.cpp:13-209,461-518
 *   fwi_iteration         one iteration of simulate_fwi_PSV
 *                                                 src/cpu/n_simulate_PSV.cpp:249-582
 *
 * Conventions (the reference has none: void functions that exit(0) on errors):
 *   - every call returns FWI_OK (0) or a negative FWI_E* code and never exits;
 *     fwi_last_error(ctx) gives the message of the last failure;
 *   - the context owns all device memory; the caller owns host buffers;
 *   - calls on one context run in order on the context's HIP stream and
 *     return when their host outputs are written; one context per host thread.
 *
 * Array layouts follow the reference: fields row-major [nz][nx] (z = depth =
 * row, x fastest, src/cpu/n_alloc_PSV.cpp:129-170), seismograms [nrec][nt]
 * per shot (src/h_io_PSV.cpp:292-317).  Shots' receivers are concatenated
 * (CSR over shots), so shots may have different receiver counts.
 */
#ifndef FWI_PSV_H
#define FWI_PSV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FWI_PSV_ABI_VERSION 6

enum fwi_status {
    FWI_OK = 0,
    FWI_EINVAL = -1,     /* bad argument or inconsistent sizes */
    FWI_ENOMEM = -2,     /* device or host allocation failed, or the budget is too small */
    FWI_EHIP = -3,       /* HIP runtime error */
    FWI_ENONFINITE = -4, /* NaN or Inf in a wavefield or a misfit */
    FWI_ECFL = -5,       /* dt above the stability limit of the stencil */
    FWI_ESTEP = -6,      /* step-length search found no decrease (the reference exits) */
    FWI_ESTATE = -7,     /* call out of order, e.g. a gradient before fwi_set_observed */
    FWI_ENODEV = -8      /* no gfx950 device */
};

/* Wavefield storage for the gradient. */
enum fwi_storage {
    FWI_STORE_AUTO = 0,        /* full snapshots if they fit the budget, else checkpoints */
    FWI_STORE_FULL = 1,        /* every snap_dt-th forward state kept (the reference's scheme) */
    FWI_STORE_CHECKPOINT = 2   /* checkpoints + forward recomputation per segment */
};

/* Model parameters that fwi_update_model and the step-length trials move. */
enum fwi_param {
    FWI_PARAM_LAME = 0,      /* lam, mu, rho, each by update_mat2 (src/cpu/n_simulate_PSV.cpp:542-546) */
    FWI_PARAM_VELOCITY = 1   /* cp and cs by update_mat2 with rho held fixed; the global lam / mu
                                gradient is carried to them by the chain rule through
                                lam = rho (cp^2 - 2 cs^2), mu = rho cs^2 at the saved model */
};

/* Kernel structure of a time step. */
enum fwi_form {
    FWI_FORM_FUSED = 0,  /* two fused half-step kernels, CPML in strips only */
    FWI_FORM_PLAIN = 1   /* the reference GPU path's structure: one kernel per sub-step,
                            CPML over the whole interior, a device sync after each launch
                            (src/cuda/d_kernel_PSV.cu:16-220) */
};

typedef struct fwi_ctx fwi_ctx;

/* Grid, stencil, boundaries: everything simulate_PSV derives from metaint /
 * metafloat before calling kernel_PSV (src/h_simulate_PSV.cpp:101-185). */
typedef struct {
    int32_t nt, nz, nx;
    double dt, dz, dx;
    int32_t fdorder;            /* 2, 4, ..., 12; the stencil uses hc[1..fdorder/2] */
    double hc[7];               /* fwi_holbergcoeff output */
    int32_t isurf[4];           /* free-surface index: top, bottom, left, right; 0 = none */
    int32_t pml_z, pml_x;       /* CPML present along z / x */
    const double *a_z, *b_z, *K_z, *a_half_z, *b_half_z, *K_half_z;  /* [nz] (fwi_cpml) */
    const double *a_x, *b_x, *K_x, *a_half_x, *b_half_x, *K_half_x;  /* [nx] */
    int32_t snap_z1, snap_z2, snap_x1, snap_x2;  /* gradient box (inclusive for storage) */
    int32_t snap_dt, snap_dz, snap_dx;           /* storage strides */
    int32_t taper[8];           /* t1 t2 b1 b2 l1 l2 r1 r2 (metaint[20..27]) */
} fwi_grid;

/* Acquisition as CSR lists over shots.  The reference's form (nsrc sources
 * with src_shot_to_fire, every shot recording the same nrec receivers,
 * src/h_io_PSV.cpp:123-153) is the special case rec_off[s] = s * nrec. */
typedef struct {
    int32_t nshot;
    int32_t stf_type;             /* 0: v += dt rho^-1 s / (dz dx); 1: v += s (vsrc2) */
    const int32_t *src_off;       /* [nshot + 1] */
    const int32_t *src_z, *src_x; /* [src_off[nshot]] grid indices */
    const double *stf_z, *stf_x;  /* [src_off[nshot]][nt] */
    const int32_t *rec_off;       /* [nshot + 1] */
    const int32_t *rec_z, *rec_x; /* [rec_off[nshot]] */
} fwi_acq;

typedef struct {
    int32_t batch;          /* shots advanced together in one launch; 0 = as many as memory allows.
                               With CHECKPOINT storage a smaller batch leaves memory for more kept
                               units (ckpt_keep); at 1001 x 2001, 8 shots per launch fill the GPU */
    int32_t storage;        /* enum fwi_storage */
    int32_t ckpt_interval;  /* time steps per checkpoint segment; 0 = chosen from memory */
    int64_t mem_budget;     /* device bytes the context may allocate; 0 = 90 % of free memory */
    int32_t form;           /* enum fwi_form */
    int32_t grad_sign;      /* -1 (default): the CPU path's "grad -= ..." (n_kernel_lib_PSV.cpp:477-479);
                               +1: the GPU path's "+=" (d_kernel_lib_PSV.cu:838-840) */
    int32_t use_graph;      /* capture chunks of time steps (a checkpoint segment, or 64 steps)
                               in a hipGraph once per propagation and replay it; per-kernel
                               times then only see the replays (FWI_K_GRAPH) */
    int32_t tile_tz;        /* rows per thread in the stencil kernels; 0 = tuned;
                               8, 16, 32, 64 (variants 1-3), 32, 64 (variant 4), 16, 32
                               (variants 5, 10-13), fdorder + 2 or 2 fdorder + 4 (variant 6),
                               4 fdorder + 8 or 8 fdorder + 16 (variants 7, 8), 1 or 2
                               (variant 9: waves of resident blocks) */
    int32_t device;         /* HIP device ordinal */
    int32_t check_finite;   /* scan the wavefield for NaN/Inf after each propagation */
    int32_t kernel_variant; /* stencil kernels: 0 = tuned, 1 = x taps as shifted loads,
                               2 = x taps from neighbouring lanes with next-row prefetch,
                               3 = variant 2 at six waves per SIMD,
                               4 = the whole step in one pass (needs no free surface other
                                   than the top one at row fdorder/2, and sources and
                                   receivers inside the interior),
                               5 = variant 2 with wave-uniform addressing and ordered loads,
                               6 = variant 5 with every load issued one row ahead,
                               7 = variant 4 built as variant 6 (variant 4's conditions),
                               8 = variant 7 with a block's wavefronts side by side in x,
                               9 = variant 8 launched as one or two waves of resident
                                   blocks with the rows shared evenly,
                               10 = variant 5 on full 64-column tiles from column 0 (rows
                                   aligned, the x halo from one extra load per row),
                               11 = variant 10 with a block's four wavefronts on four
                                   shots of one tile and segment (shared media rows),
                               12 = variant 10 with a block's four wavefronts on four
                                   adjacent tiles of one segment (1 KiB row runs),
                               13 = variant 12 with the non-temporal hint on the loads and
                                   stores of values used once per launch
                               (5 to 13 need z CPML strips of at most 64 rows, at the
                               top and bottom of the grid) */
    int32_t work_order;     /* order of the stencil tiles over the grid: 0 = tuned,
                               1 = shot fastest then z then x, 2 = shot fastest then x then z,
                               3 = x tiles fastest, shots outermost,
                               4 = x tiles fastest, then shots, z segments outermost */
    int32_t param;          /* enum fwi_param */
    double vel_bounds[4];   /* FWI_PARAM_VELOCITY: cp_min, cp_max, cs_min, cs_max in m/s, the
                               clip bounds and update scales of update_mat2 (all 0: 1000, 6000,
                               500, 3500) */
    int32_t pcg;            /* fwi_iteration: replace the global gradient by its conjugate-gradient
                               direction (fwi_pcg_direction) before the taper */
    int32_t smooth_hfs;     /* fwi_iteration: Gaussian smoothing of half width smooth_hfs (>= 2;
                               0 = none) after the taper (fwi_smooth_gradient) */
    int32_t ckpt_slots;     /* CHECKPOINT storage: at most this many checkpoint buffers (0 = as many
                               as the budget allows).  With fewer than the ceil(nt / ckpt_interval) - 1
                               segment boundaries, the reversal recomputes segments from held
                               checkpoints on a binomial (revolve-style) schedule */
    int32_t seis_fp64;      /* 1: the seismograms that fwi_forward returns and that misfits and
                               adjoint sources are made of come from an FP64 propagation (FP64
                               media, CPML and fields); the FP32 stencil still runs the wavefields
                               the gradient correlates (ABI 4) */
    int32_t ckpt_keep;      /* CHECKPOINT storage: the forward pass keeps the snapshots of the last
                               segments, which the reversal then runs without recomputing them;
                               0 = as many segments as the budget allows beside the batch and one
                               checkpoint per remaining boundary (at least the last one), k > 0 =
                               at most k (ABI 5) */
} fwi_opts;

/* Per-kernel timing collected while profiling is on: HIP events on the context's
 * stream around every launch of that kernel, or around every stride-th one
 * (fwi_profile_stride). */
typedef struct {
    int64_t launches;   /* launches while profiling was on */
    double total_ms;    /* summed duration of the timed ones */
    int64_t timed;      /* launches that carried events (ABI 6) */
} fwi_kernel_time;

enum fwi_kernel_id {
    FWI_K_STRESS = 0,     /* stress half-step (vdiff2 + pml_diff2 + update_s2) */
    FWI_K_VELOCITY = 1,   /* velocity half-step (sdiff2 + pml_diff2 + update_v2 [+ fwi_grad2]) */
    FWI_K_SURFACE = 2,    /* surf_mirror */
    FWI_K_SOURCE = 3,     /* vsrc2 */
    FWI_K_RECEIVER = 4,   /* urec2 */
    FWI_K_STORE = 5,      /* gard_fwd_storage2 / checkpoint copies */
    FWI_K_OTHER = 6,
    FWI_K_STEP = 7,       /* one-pass time step: stress, top free surface and velocity (kernel variant 4) */
    FWI_K_GRAPH = 8,      /* a replayed hipGraph of several whole time steps (use_graph) */
    FWI_K_COUNT = 9
};

/* ---- library ------------------------------------------------------------- */
int32_t fwi_abi_version(void);
const char *fwi_strerror(int32_t status);
const char *fwi_last_error(const fwi_ctx *ctx);
void fwi_default_opts(fwi_opts *opts);

/* ---- host preprocessing (no device needed) -------------------------------- */
int32_t fwi_holbergcoeff(int32_t fdorder, int32_t maxrelerror, double hc[7]);
int32_t fwi_cpml(double *a, double *b, double *K, double *a_half, double *b_half, double *K_half,
                 double npower, double damp_v, double rcoef, double k_max, double freq,
                 int32_t npml_h1, int32_t npml_h2, int32_t fpad, int32_t nh, double dt, double dh);
int32_t fwi_wavelet(double *signal, int32_t nt, double dt, double amp, double fc, double ts, int32_t shape);
/* Stability limit dh / (sqrt 2 sum|hc| cmax); cmax/cmin returned if non-null. */
double fwi_dtstab(int32_t nz, int32_t nx, double dh, int32_t fw, const double *rho, const double *lam,
                  const double *mu, const double hc[7], double *cmax, double *cmin);

/* ---- context ----------------------------------------------------------------- */
int32_t fwi_create(const fwi_grid *grid, const fwi_acq *acq, const fwi_opts *opts, fwi_ctx **out);
void fwi_destroy(fwi_ctx *ctx);
/* Batch size and storage the context chose, and its device memory in bytes. */
int32_t fwi_plan(const fwi_ctx *ctx, int32_t *batch, int32_t *storage, int32_t *ckpt_interval,
                 int64_t *device_bytes);

/* Checkpoint schedule of CHECKPOINT storage: units = segments of ckpt_interval steps,
 * slots = checkpoint buffers, recompute = forward steps run again per time step of the
 * adjoint (below 1 when the forward pass keeps the snapshots of the last segments; with a
 * checkpoint at every other boundary each remaining segment is recomputed once). */
int32_t fwi_checkpoint_plan(const fwi_ctx *ctx, int32_t *units, int32_t *slots, double *recompute);
/* Segments at the end of the time axis whose snapshots the forward pass keeps (ABI 5). */
int32_t fwi_checkpoint_kept(const fwi_ctx *ctx, int32_t *kept);

/* Model lam, mu, rho [nz][nx]; derives mu_zx, rho_zp, rho_xp and the scalar means
 * (mat_av2) on the device. */
int32_t fwi_set_model(fwi_ctx *ctx, const double *lam, const double *mu, const double *rho);
int32_t fwi_get_model(fwi_ctx *ctx, double *lam, double *mu, double *rho);
/* Averaged media as used by the stencil ([nz][nx]) and the means C[3] (lam, mu, rho). */
int32_t fwi_get_media(fwi_ctx *ctx, double *mu_zx, double *rho_zp, double *rho_xp, double C[3]);
/* Observed displacement seismograms, CSR over shots: [rec_off[nshot]][nt]. */
int32_t fwi_set_observed(fwi_ctx *ctx, const double *obs_z, const double *obs_x);

/* Forward modelling of shots [shot0, shot0 + nshots): receiver displacement
 * seismograms into seis_z / seis_x (CSR rows rec_off[shot0] .. rec_off[shot0 + nshots]). */
int32_t fwi_forward(fwi_ctx *ctx, int32_t shot0, int32_t nshots, double *seis_z, double *seis_x);
/* Final wavefield of one shot of the last propagation: field 0..8 = vz vx uz ux szz szx sxx We We_adj. */
int32_t fwi_get_wavefield(fwi_ctx *ctx, int32_t shot, int32_t field, double *out);
/* Forward of one shot keeping its snapshots (the accu_* cubes that write_accu saves,
 * src/h_io_PSV.cpp:246-288): each output is [snap_nt][snap_nz][snap_nx], velocities
 * divided by dt as gard_fwd_storage2 stores them.  Needs FULL storage. */
int32_t fwi_forward_snapshots(fwi_ctx *ctx, int32_t shot, double *seis_z, double *seis_x, double *vz,
                              double *vx, double *szz, double *szx, double *sxx);

/* Gradient of shots [shot0, shot0 + nshots) against the observed data:
 * misfit[i] (may be null) receives each shot's L2, and the energy-weighted,
 * mean-scaled shot gradients are added to the context's global gradient. */
int32_t fwi_gradient(fwi_ctx *ctx, int32_t shot0, int32_t nshots, double *misfit);
/* Raw correlation gradient sum over the shots (fwi_grad2 output, snap grid
 * [snap_nz][snap_nx]) for caller-given adjoint sources adj_z / adj_x
 * (CSR like the seismograms, injected as type-0 sources in reversed time). */
int32_t fwi_gradient_raw(fwi_ctx *ctx, int32_t shot0, int32_t nshots, const double *adj_z,
                         const double *adj_x, double *g_lam, double *g_mu, double *g_rho);
/* Derivative of J = sum over samples of 0.5 dt r^2 with respect to lam, mu and rho
 * ([nz][nx], zero outside the snapshot box) for adjoint sources r = adj_z / adj_x
 * (CSR like the seismograms): the backward pass of the modelling operator.  No
 * single reference function: it is fwi_grad2's correlation
 * (src/cpu/n_kernel_lib_PSV.cpp:439-484) with the density term taken as the adjoint
 * velocity times the forward velocity, and the shear and density terms carried
 * through mat_av2's staggered averages (:853-922) by the chain rule on the device.
 * Needs snap_dz = snap_dx = 1 and the fused form. */
int32_t fwi_misfit_gradient(fwi_ctx *ctx, int32_t shot0, int32_t nshots, const double *adj_z,
                            const double *adj_x, double *d_lam, double *d_mu, double *d_rho);

/* The exact derivative of the same J (no reference counterpart: the reference's gradient is
 * the correlation of fwi_grad2): the discrete adjoint of the time step -- every operation of
 * the forward step (stencils, CPML recursions, free surface, sources, the receivers'
 * running sum) transposed and applied in reverse order, the media gradient taken from the
 * forward's stored strain rates and stress divergences and carried to lam, mu, rho through
 * mat_av2's averages.  Zero outside the snapshot box.  It keeps five fields per time step
 * and shot (FP32) for the call, in batches of shots that fit; on_device: adj_* and the
 * outputs are device arrays (as in the _device calls below). */
int32_t fwi_misfit_gradient_exact(fwi_ctx *ctx, int32_t shot0, int32_t nshots, const double *adj_z,
                                  const double *adj_x, double *d_lam, double *d_mu, double *d_rho,
                                  int32_t on_device);

/* The same three calls with device arrays (HIP pointers of the runtime the library runs
 * on, e.g. torch tensors in a process that loaded torch first): model lam, mu, rho
 * [nz][nx] in, seismograms [rows][nt] out, adjoint sources in and the derivative out, all
 * FP64.  Their contents are read and written by kernels on the context's stream; each call
 * returns when its outputs are complete, and the caller makes its own writes to the inputs
 * complete before the call (e.g. a stream synchronisation).  The reference's entry points
 * take host arrays only (include/cpu/n_kernel_PSV.hpp:8-42). */
int32_t fwi_set_model_device(fwi_ctx *ctx, const double *lam, const double *mu, const double *rho);
int32_t fwi_forward_device(fwi_ctx *ctx, int32_t shot0, int32_t nshots, double *seis_z, double *seis_x);
int32_t fwi_misfit_gradient_device(fwi_ctx *ctx, int32_t shot0, int32_t nshots, const double *adj_z,
                                   const double *adj_x, double *d_lam, double *d_mu, double *d_rho);
int32_t fwi_zero_gradient(fwi_ctx *ctx);
int32_t fwi_get_gradient(fwi_ctx *ctx, double *g_lam, double *g_mu, double *g_rho);
int32_t fwi_set_gradient(fwi_ctx *ctx, const double *g_lam, const double *g_mu, const double *g_rho);
int32_t fwi_taper_gradient(fwi_ctx *ctx);
/* PCG_PSV on the three global gradients (src/cpu/n_PCG_PSV.cpp:13-44, disabled in the
 * reference's loop, src/cpu/n_simulate_PSV.cpp:431-435): Fletcher-Reeves beta = sum g^2 /
 * sum old^2 clipped at 0, dir = -g + beta dir, old = -g, g = dir; the context keeps dir and
 * old between calls.  The first call (old = 0) takes beta = 0 where the reference divides
 * by zero. */
int32_t fwi_pcg_direction(fwi_ctx *ctx);
/* apply_gauss_filter with gauss_filter_kernel on the three global gradients
 * (src/cpu/n_kernel_lib_PSV.cpp:1012-1084, disabled in the reference's loop): a
 * (2 hfs + 1)^2 Gaussian with sigma = hfs / 2 (integer division) over the cells
 * [hfs, n - hfs - 1) of each axis, accumulated in FP64 (the reference's accumulator is an
 * int, which truncates every partial sum).  hfs >= 2 (hfs = 1 makes sigma 0). */
int32_t fwi_smooth_gradient(fwi_ctx *ctx, int32_t hfs);

/* Sum of the shots' L2 misfits for the current model (forward only). */
int32_t fwi_misfit(fwi_ctx *ctx, int32_t shot0, int32_t nshots, double *L2);
int32_t fwi_save_model(fwi_ctx *ctx);
/* m = clip(m_old + eps f G): lam, mu with eps, rho with rho_factor * eps.  With
 * FWI_PARAM_VELOCITY: cp, cs with eps (f from the cp / cs gradient and vel_bounds),
 * rho = rho_old; rho_factor and rho_min are not used. */
int32_t fwi_update_model(fwi_ctx *ctx, double eps, double rho_factor, double rho_min);
/* step_length_PSV's bracketing from est (halve until a trial lowers the misfit, then grow)
 * and calc_opt_step's parabola vertex (src/cpu/n_step_length_PSV.cpp:13-209).  A trial
 * step below 2^-20 (a model change the FP32 media do not resolve) never counts as lowering
 * the misfit.  FWI_ESTEP where the reference prints "Steplength estimation failed!". */
int32_t fwi_step_length(fwi_ctx *ctx, double est, double L2_0, double *step_out, int32_t *ntrials);
/* One simulate_fwi_PSV iteration over all shots; returns its misfit and step. */
int32_t fwi_iteration(fwi_ctx *ctx, double *L2, double *step);

/* ---- allocation-free entry point -------------------------------------------------- */
/* No counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * mat_av2 (src/cpu/n_kernel_lib_PSV.cpp:853-922) as fwi_set_model runs it, on arrays
 * the caller owns, all device pointers: lam, mu, rho [nz][nx] FP64 in; lam_f, mu_f
 * (the model in FP32), mu_zx, rho_zp, rho_xp [nz][nx] FP32 and means[3] (mean lam, mu,
 * rho over the (nz - 1)(nx - 1) cells) out; the last row and column of the averaged
 * media are written as 0.  Asynchronous on `stream` (a hipStream_t, NULL = the null
 * stream).  The block partial sums of the means go to a fixed 12 KiB device array of
 * the library (one per GPU, held until the process ends) instead of a buffer the call
 * allocates: nothing is allocated or freed.  Calls that overlap on one GPU share that
 * array and race.  Returns FWI_EINVAL for nz or nx below 2 or more than
 * FWI_NOALLOC_MAX_CELLS cells (message in fwi_last_error(NULL)). */
#define FWI_NOALLOC_MAX_CELLS (1LL << 30)
int32_t fwi_mat_av2_noalloc(const double *lam, const double *mu, const double *rho, int32_t nz, int32_t nx,
                            float *lam_f, float *mu_f, float *mu_zx, float *rho_zp, float *rho_xp, double *means,
                            void *stream);

/* ---- profiling ------------------------------------------------------------------ */
/* Rows per thread the stencil kernels use (the autotuner's choice once a
 * propagation has run) and the tuner's timings as text. */
int32_t fwi_tuning(const fwi_ctx *ctx, int32_t *tile_tz, char *log, int32_t log_len);
int32_t fwi_profile(fwi_ctx *ctx, int32_t enable);
/* fwi_profile for the kernel classes whose bit (1u << FWI_K_*) is set in mask only (no
 * counterpart in the reference): each timed launch costs two event records, so timing
 * only the kernel whose duration matters disturbs the run less.  0 switches timing off. */
int32_t fwi_profile_kernels(fwi_ctx *ctx, uint32_t mask);
/* Events on every stride-th launch of each timed kernel class only (1 = every launch,
 * the default; no counterpart in the reference): the average duration is then sampled
 * uniformly over the launches while their event records cost 1/stride as much (ABI 6). */
int32_t fwi_profile_stride(fwi_ctx *ctx, int32_t stride);
int32_t fwi_kernel_times(fwi_ctx *ctx, fwi_kernel_time times[FWI_K_COUNT]);
int32_t fwi_sync(fwi_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* FWI_PSV_H */
