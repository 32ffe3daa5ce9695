// psv_api.hip -- context, memory plan, time loops and the C ABI of include/fwi_psv.h.
//
// Shots are processed in batches of B: every launch advances all B shots of a
// batch (blockIdx.z = shot), so the media and CPML profiles are read once per
// tile for B shots and small grids still fill the chip.  The reference loops
// shots serially (src/cpu/n_simulate_PSV.cpp:270, src/cuda/d_simulate_PSV.cu:443).
//
// Gradient storage.  FULL keeps every snap_dt-th forward state of the snap box
// (the reference's accu_* cubes, src/cpu/n_alloc_PSV.cpp:178-192).  CHECKPOINT
// keeps the complete propagation state (velocities, stresses, CPML memory)
// every S steps; the adjoint walks the segments backwards, recomputing each
// segment's forward states into a segment buffer before running the segment's
// adjoint steps.  The recomputation repeats the forward kernels on the same
// inputs, so the checkpointed gradient is bit-identical to the FULL one.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "psv_internal.h"

using namespace psv;

namespace {

struct EventPair {
    int kid;
    hipEvent_t a, b;
};

size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// gain on the lam / mu correlation terms (divided out in FP64 on the host side)
constexpr double kGainLM = 1152921504606846976.0;   // 2^60

}  // namespace

struct fwi_ctx {
    // ---- configuration
    fwi_grid g{};
    fwi_opts o{};
    int N = 1, P = 0, nzs = 0, nxs = 0, snt = 0, snz = 0, snx = 0;
    int nshot = 0, stf_type = 1, nrec_tot = 0, nsrc_tot = 0;
    int tz = 16;
    std::vector<int> src_off, rec_off, isurf;
    std::vector<double> prof[12];
    std::string err;
    hipStream_t st = nullptr;
    int have_model = 0, have_obs = 0;
    double C[3] = {0, 0, 0};
    double step_est = 0.01;

    // ---- plan
    int B = 1, storage = FWI_STORE_FULL, S = 0, nseg = 0;
    // CHECKPOINT: nseg units of S steps and `slots` checkpoint buffers; with slots < nseg - 1
    // the reversal advances from held states on a binomial schedule (ck_top / ck_rev: best
    // split of a chain of n units with s free slots, index n * (slots + 1) + s)
    // The forward pass keeps the snapshots of the last `keep` units (from time step t_keep):
    // the reversal runs their adjoint steps first, without recomputation, then reverses the
    // nrev = nseg - keep units before them from checkpoints.
    int slots = 0, keep = 0, nrev = 0, t_keep = 0;
    std::vector<int> ck_top, ck_rev;
    std::vector<int> spine_slot;   // [nseg]: slot the forward pass saves a unit's start state into, -1 none
    long long extra_units = 0;     // units advanced again beyond the forward pass and the leaves
    size_t bytes = 0;
    std::vector<void *> allocs;

    // ---- device: shared
    float *lam = nullptr, *mu = nullptr, *muzx = nullptr, *rhoz = nullptr, *rhox = nullptr;
    double *m_lam = nullptr, *m_mu = nullptr, *m_rho = nullptr, *o_lam = nullptr, *o_mu = nullptr, *o_rho = nullptr;
    double *G[3] = {nullptr, nullptr, nullptr};
    double *tmp = nullptr, *red = nullptr;  // precondition scratch, reduction scratch
    double *vel = nullptr;   // FWI_PARAM_VELOCITY: [4][nz][nx] cp, cs of the saved model, cp / cs gradient
    double *raw = nullptr;     // misfit derivative: FP64 sums over shots [5][snz][snx] (first use)
    double *mg_out = nullptr;  // misfit derivative for host callers [3][nz][nx] (first use)
    int adj_on_device = 0;     // adjoint sources of this call are device arrays
    double *pcg = nullptr;     // conjugate-gradient state [3 params][dir, old][nz][nx] (first use)
    double *gfilt = nullptr;   // Gaussian filter of the last smoothing (first use)
    int gfilt_hfs = -1;
    double vb[4] = {1000.0, 6000.0, 500.0, 3500.0};   // cp_min, cp_max, cs_min, cs_max
    int *zmap = nullptr, *xmap = nullptr, *nonfinite = nullptr;
    float *pf[12] = {};   // b a 1/K  b_half a_half 1/K_half, z then x
    CpmlPoint *zp = nullptr, *xp = nullptr;   // the same, packed per grid line
    // acquisition
    double *stf_z = nullptr, *stf_x = nullptr;
    int *sg_shot = nullptr, *sg_cell = nullptr, *sg_off = nullptr, *sg_idx = nullptr;   // source groups
    int *rg_shot = nullptr, *rg_cell = nullptr, *rg_off = nullptr, *rg_idx = nullptr;   // receiver groups (adjoint)
    std::vector<int> sg_first, rg_first;   // first group of each shot (+ total)
    int *rec_shot = nullptr, *rec_cell = nullptr, *rec_off_d = nullptr;
    double *syn_z = nullptr, *syn_x = nullptr, *obs_z = nullptr, *obs_x = nullptr, *res_z = nullptr,
           *res_x = nullptr, *acc = nullptr, *misfit = nullptr;
    double *alpha_d = nullptr, *shotmax = nullptr;   // per-shot adjoint source scale, residual maxima
    std::vector<double> alpha;
    // ---- device: batch
    float *f[9] = {};      // vz vx szz szx sxx uz ux We Wa   [B][nz][P]
    float *psi[8] = {};    // pvz_z pvx_z pszz_z pszx_z [B][nzs][P]; pvx_x pvz_x psxx_x pszx_x [B][nz][nxs]
    float *gr[3] = {};     // [B][snz][snx]
    float *gx = nullptr;   // [2][B][snz][snx]: the extra sums of the misfit derivative (gform 1)
    // FP64 totals of the five sums [5][B][snz][snx]: the FP32 sums above cover at most one
    // chunk of kFlushSteps steps (or one checkpoint segment) before they are added in here and
    // zeroed, so their rounding does not grow with nt over sums of oscillating terms
    double *gr64 = nullptr;
    int gform = 0;         // 1 while fwi_misfit_gradient runs
    // graph replay: per-step tables of the forward, recomputation and adjoint passes [nt],
    // and the base time index of the chunk being replayed
    StepInfo *si_fwd = nullptr, *si_rec = nullptr, *si_adj = nullptr;
    int *si_base = nullptr;
    float *snap[5] = {};   // FULL: [B][snt][snz][snx]; CHECKPOINT: [B][S/snap_dt][snz][snx]
    size_t snap_stride = 0;
    // one-pass step: second buffer set of the 5 fields and the stress-side psi
    // (pvz_z pvx_z pvx_x pvz_x); f / psi always name the current time level
    bool fusable = false;               // k_step supports this grid, surface and acquisition
    bool surf_top_only = false;         // the only free surface is the top one at row N
    ZTab ztab{};                        // z profiles of the strip rows for variant 5
    bool ztab_ok = false;               // the z strips are [N, lo) and [hi, nz - N), at most kZTabRows rows
    float *f2[5] = {}, *psi2[4] = {};
    // checkpoint mode
    float *rf[5] = {}, *rpsi[8] = {};   // recompute state
    float *rf2[5] = {}, *rpsi2[4] = {};
    float *ck = nullptr;                // [slots][13 arrays of the batch]
    float *rf_canon = nullptr;          // rf[0] as allocated: the recompute state's level order
    size_t ck_bytes_per_seg = 0;
    // seis_fp64: the FP64 propagation's fields [15][B][nz][P] (5 fields, 8 CPML memories, 2
    // strain rates), its media [5][nz][P] and CPML profiles [nz + nx][6]; media_ver counts
    // model changes, d64_ver the one its media were made from
    double *d64 = nullptr, *dmed = nullptr, *dprof = nullptr;
    long long media_ver = 0, d64_ver = -1;
    int last_s0 = -1, last_nb = 0;
    // plain form: full-grid derivative scratch and CPML memory (one shot)
    bool plain = false;
    PlainBufs pb{};
    bool tuned = false;
    int var = 2;   // stencil kernel variant (autotuned); 4 = one-pass step
    int order = 0; // stencil work order (autotuned)
    int depth = 1; // rows the one-pass step loads ahead (autotuned)
    std::string tune_log;

    // ---- profiling
    int prof_on = 0, prof_stride = 1;
    std::vector<hipEvent_t> ev_pool;
    std::vector<EventPair> ev_pending;
    fwi_kernel_time times[FWI_K_COUNT] = {};

    int fail(int code, const char *fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    template <class T>
    T *dalloc(size_t n) {
        if (n == 0) n = 1;
        void *p = nullptr;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
        allocs.push_back(p);
        bytes += n * sizeof(T);
        return static_cast<T *>(p);
    }
    size_t fs() const { return (size_t)g.nz * P; }
    size_t psiz() const { return (size_t)nzs * P; }
    size_t psix() const { return (size_t)g.nz * nxs; }
};

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) return c->fail(FWI_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define RETIF(x)                  \
    do {                          \
        int r_ = (x);             \
        if (r_ != FWI_OK) return r_; \
    } while (0)

namespace {

// ---------------------------------------------------------------------------
// profiling
// ---------------------------------------------------------------------------
hipEvent_t ev_get(fwi_ctx *c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hipEventCreate(&e);
    return e;
}

// Events of the first n pending pairs (all complete) into the per-kernel totals.
void prof_take(fwi_ctx *c, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        const EventPair &p = c->ev_pending[i];
        float ms = 0;
        hipEventElapsedTime(&ms, p.a, p.b);
        c->times[p.kid].timed += 1;
        c->times[p.kid].total_ms += ms;
        c->ev_pool.push_back(p.a);
        c->ev_pool.push_back(p.b);
    }
    c->ev_pending.erase(c->ev_pending.begin(), c->ev_pending.begin() + n);
}

// At most kMaxPending timed launches are outstanding: past that, the host waits
// for the older half (the stream keeps running the rest) and folds it in.
constexpr size_t kMaxPending = 2048;

struct Timer {
    fwi_ctx *c;
    int kid;
    hipEvent_t a = nullptr;
    Timer(fwi_ctx *c_, int k) : c(c_), kid(k) {
        if (c->prof_on & (1 << kid)) {   // every launch counts; every prof_stride-th is timed
            if (c->times[kid].launches++ % c->prof_stride == 0) {
                a = ev_get(c);
                hipEventRecord(a, c->st);
            }
        }
    }
    ~Timer() {
        if (a) {
            hipEvent_t b = ev_get(c);
            hipEventRecord(b, c->st);
            c->ev_pending.push_back({kid, a, b});
            if (c->ev_pending.size() >= kMaxPending) {
                const size_t half = kMaxPending / 2;
                hipEventSynchronize(c->ev_pending[half - 1].b);
                prof_take(c, half);
            }
        }
    }
};

void prof_collect(fwi_ctx *c) {
    if (c->ev_pending.empty()) return;
    hipStreamSynchronize(c->st);
    prof_take(c, c->ev_pending.size());
}

int sync(fwi_ctx *c) {
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return c->fail(FWI_EHIP, "stream: %s", hipGetErrorString(e));
    e = hipGetLastError();
    if (e != hipSuccess) return c->fail(FWI_EHIP, "launch: %s", hipGetErrorString(e));
    prof_collect(c);
    return FWI_OK;
}

template <class T>
int upload(fwi_ctx *c, T *dst, const T *src, size_t n) {
    if (n == 0) return FWI_OK;
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->st));
    return FWI_OK;
}

// ---------------------------------------------------------------------------
// propagation state
// ---------------------------------------------------------------------------
// The current time level (f[0..4], psi[0..7]) and, for the one-pass step, the
// buffers the next level is written to (f2, psi2 = pvz_z pvx_z pvx_x pvz_x);
// flip() swaps them in the owner's arrays, so f and psi always name the
// current level.
struct Wave {
    float **f, **psi, **f2, **psi2;
    void flip() const {
        for (int i = 0; i < 5; ++i) std::swap(f[i], f2[i]);
        std::swap(psi[0], psi2[0]);
        std::swap(psi[1], psi2[1]);
        std::swap(psi[4], psi2[2]);
        std::swap(psi[5], psi2[3]);
    }
};
Wave main_wave(fwi_ctx *c) { return {c->f, c->psi, c->f2, c->psi2}; }
Wave recompute_wave(fwi_ctx *c) { return {c->rf, c->rpsi, c->rf2, c->rpsi2}; }

// ---------------------------------------------------------------------------
// kernel arguments
// ---------------------------------------------------------------------------
void set_state(KArgs &a, const Wave &w) {
    a.vz = w.f[0];
    a.vx = w.f[1];
    a.szz = w.f[2];
    a.szx = w.f[3];
    a.sxx = w.f[4];
    a.pvz_z = w.psi[0]; a.pvx_z = w.psi[1]; a.pszz_z = w.psi[2]; a.pszx_z = w.psi[3];
    a.pvx_x = w.psi[4]; a.pvz_x = w.psi[5]; a.psxx_x = w.psi[6]; a.pszx_x = w.psi[7];
    a.o_vz = w.f2[0];
    a.o_vx = w.f2[1];
    a.o_szz = w.f2[2];
    a.o_szx = w.f2[3];
    a.o_sxx = w.f2[4];
    a.o_pvz_z = w.psi2[0]; a.o_pvx_z = w.psi2[1]; a.o_pvx_x = w.psi2[2]; a.o_pvz_x = w.psi2[3];
}

KArgs base_args(fwi_ctx *c, int nb, const Wave &w) {
    KArgs a{};
    a.nz = c->g.nz;
    a.nx = c->g.nx;
    a.P = c->P;
    a.B = nb;
    a.fs = c->fs();
    a.dt = (float)c->g.dt;
    for (int k = 0; k < 6; ++k) {
        a.cz[k] = k < c->N ? (float)(c->g.hc[k + 1] / c->g.dz) : 0.f;
        a.cx[k] = k < c->N ? (float)(c->g.hc[k + 1] / c->g.dx) : 0.f;
    }
    a.uz = c->f[5];
    a.ux = c->f[6];
    a.We = c->f[7];
    a.lam = c->lam;
    a.mu = c->mu;
    a.muzx = c->muzx;
    a.rhoz = c->rhoz;
    a.rhox = c->rhox;
    a.zmap = c->zmap;
    a.xmap = c->xmap;
    a.nzs = c->nzs;
    a.nxs = c->nxs;
    a.bz = c->pf[0]; a.az = c->pf[1]; a.ikz = c->pf[2];
    a.bhz = c->pf[3]; a.ahz = c->pf[4]; a.ikhz = c->pf[5];
    a.bx = c->pf[6]; a.ax = c->pf[7]; a.ikx = c->pf[8];
    a.bhx = c->pf[9]; a.ahx = c->pf[10]; a.ikhx = c->pf[11];
    a.zp = c->zp;
    a.xp = c->xp;
    set_state(a, w);
    a.surf_top = c->g.isurf[0] > 0 ? c->g.isurf[0] : -1;   // used by k_step only (fusable: row N)
    a.surf5 = -1;                                          // set per launch by half_steps
    a.snap_z1 = c->g.snap_z1;
    a.snap_z2 = c->g.snap_z2;
    a.snap_x1 = c->g.snap_x1;
    a.snap_x2 = c->g.snap_x2;
    a.snap_dz = c->g.snap_dz;
    a.snap_dx = c->g.snap_dx;
    a.snz = c->snz;
    a.snx = c->snx;
    a.g_lam = c->gr[0];
    a.g_mu = c->gr[1];
    a.g_rho = c->gr[2];
    a.gform = c->gform;
    a.g_ext = c->gx;
    a.g_ext_stride = (size_t)c->B * c->snz * c->snx;
    a.gw = (float)(c->o.grad_sign * (double)c->g.snap_dt * c->g.dt);
    a.gain_lm = (float)kGainLM;
    a.order = c->order;
    return a;
}

SrcArgs src_args(fwi_ctx *c, int s0, int nb) {
    SrcArgs s{};
    s.shot0 = s0;
    const int g0 = c->sg_first[s0], g1 = c->sg_first[s0 + nb];
    s.ngroups = g1 - g0;
    s.grp_shot = c->sg_shot + g0;
    s.grp_cell = c->sg_cell + g0;
    s.grp_off = c->sg_off + g0;
    s.grp_idx = c->sg_idx;
    s.stf_z = c->stf_z;
    s.stf_x = c->stf_x;
    s.nt = c->g.nt;
    s.type = c->stf_type;
    s.dt = c->g.dt;
    s.dzdx = c->g.dz * c->g.dx;
    s.scale_shot = nullptr;
    return s;
}

// adjoint sources: the receivers of each shot, residual rows as source functions
SrcArgs adj_args(fwi_ctx *c, int s0, int nb) {
    SrcArgs s{};
    s.shot0 = s0;
    const int g0 = c->rg_first[s0], g1 = c->rg_first[s0 + nb];
    s.ngroups = g1 - g0;
    s.grp_shot = c->rg_shot + g0;
    s.grp_cell = c->rg_cell + g0;
    s.grp_off = c->rg_off + g0;
    s.grp_idx = c->rg_idx;
    s.stf_z = c->res_z;
    s.stf_x = c->res_x;
    s.nt = c->g.nt;
    s.type = 0;
    s.dt = c->g.dt;
    s.dzdx = c->g.dz * c->g.dx;
    s.scale_shot = c->alpha_d;
    return s;
}

RecArgs rec_args(fwi_ctx *c, int s0, int nb) {
    RecArgs r{};
    r.shot0 = s0;
    const int r0 = c->rec_off[s0], r1 = c->rec_off[s0 + nb];
    r.nrec = r1 - r0;
    r.rec_shot = c->rec_shot + r0;
    r.rec_cell = c->rec_cell + r0;
    r.acc_z = c->acc;
    r.acc_x = c->acc + c->nrec_tot;
    r.seis_z = c->syn_z + (size_t)r0 * c->g.nt;
    r.seis_x = c->syn_x + (size_t)r0 * c->g.nt;
    r.nt = c->g.nt;
    r.dt = c->g.dt;
    r.dzdx = c->g.dz * c->g.dx;
    return r;
}

// variants that advance a whole time step in one pass over two time levels
bool one_pass(int var) { return var == 4 || var == 7 || var == 8 || var == 9; }

// one time step of a one-pass variant (variants 7 to 9 fall back to 4 for a launch mode
// they have no kernel for).  Variant 9 is variant 8's kernel in balanced mode: tz is the
// number of resident-block waves the launch holds (1 or 2), the rows shared evenly.
void launch_one_pass(fwi_ctx *c, const KArgs &a, int var, int tz, int depth) {
    if (var >= 7 && c->ztab_ok && launch_step7(a, c->ztab, c->N, var == 9 ? -tz : tz, var >= 8, c->st)) return;
    launch_step(a, c->N, var >= 7 ? 32 : tz, depth, c->st);
}

bool has_surface(const fwi_ctx *c) {
    for (int i = 0; i < 4; ++i)
        if (c->g.isurf[i] > 0) return true;
    return false;
}

// stress half-step, free surface, velocity half-step on the given state
void half_steps(fwi_ctx *c, const KArgs &a) {
    if (c->plain) {
        Timer t(c, FWI_K_STRESS);
        plain_step(a, c->pb, c->N, c->g.isurf, c->g.pml_z, c->g.pml_x, c->st);
        return;
    }
    const ZTab *zt = c->ztab_ok ? &c->ztab : nullptr;
    // variants 5 and 10-13 apply a free surface at the top row N (and none elsewhere) in
    // the stress kernel itself, as the one-pass step does
    const bool surf_in_stress = c->surf_top_only && zt && (c->var == 5 || c->var >= 10);
    {
        Timer t(c, FWI_K_STRESS);
        if (surf_in_stress) {
            KArgs b = a;
            b.surf5 = c->N;
            launch_stress(b, c->N, c->tz, c->var, c->st, zt, c->depth);
        } else {
            launch_stress(a, c->N, c->tz, c->var, c->st, zt, c->depth);
        }
    }
    if (has_surface(c) && !surf_in_stress) {
        Timer t(c, FWI_K_SURFACE);
        launch_surface(a, c->N, c->g.isurf, c->st);
    }
    {
        Timer t(c, FWI_K_VELOCITY);
        launch_velocity(a, c->N, c->tz, c->var, c->st, zt, c->depth);
    }
}

// One time step of state w (stress, free surface, velocity, then the sources of
// time index it).  With dst, the step's state is stored into the snapshot slab
// dst[5] (stride snap_stride between shots) as gard_fwd_storage2 stores it after
// the injection.  a follows w: on return it names the new current level.
void time_step(fwi_ctx *c, KArgs &a, const Wave &w, const SrcArgs &sa, int it, float *const *dst) {
    set_state(a, w);
    a.store = 0;
    if (dst && !c->plain) {   // the velocity (or one-pass) kernel writes the snapshot
        a.store = 1;
        a.st_vz = dst[0]; a.st_vx = dst[1]; a.st_szz = dst[2]; a.st_szx = dst[3]; a.st_sxx = dst[4];
        a.st_stride = c->snap_stride;
    }
    if (one_pass(c->var) && !c->plain) {
        {
            Timer t(c, FWI_K_STEP);
            launch_one_pass(c, a, c->var, c->tz, c->depth);
        }
        w.flip();
        set_state(a, w);
        Timer t(c, FWI_K_SOURCE);
        launch_inject(a, sa, it, c->st);   // also rewrites the stored velocities of source cells
        return;
    }
    half_steps(c, a);
    {
        Timer t(c, FWI_K_SOURCE);
        launch_inject(a, sa, it, c->st);   // with a.store, also rewrites the stored source cells
    }
    if (dst && c->plain) {   // the reference-structured form keeps its separate storage pass
        Timer t(c, FWI_K_STORE);
        launch_store(a, dst, c->snap_stride, (float)c->g.dt, c->st);
    }
}

// ---------------------------------------------------------------------------
// hipGraph replay of chunks of whole time steps
// ---------------------------------------------------------------------------
constexpr int kGraphChunk = 64;   // steps per replayed graph outside checkpoint segments

bool use_graphs(const fwi_ctx *c) { return c->o.use_graph && !c->plain; }

// The launches of one chunk (its steps built with si set, so that every per-step
// quantity comes from the step table and the base counter) are captured the first
// time and replayed for every later chunk of the same pass; only the base time
// index is written before each replay.  The chunk must leave the two time levels
// of the one-pass step where it found them (an even number of steps).
class ChunkGraph {
  public:
    explicit ChunkGraph(fwi_ctx *c) : c_(c) {}
    // the last replay may still run when the pass returns: wait for it before the
    // executable graph (and the kernel arguments it owns) is released
    ~ChunkGraph() {
        if (exec_) {
            (void)hipStreamSynchronize(c_->st);
            hipGraphExecDestroy(exec_);
        }
    }
    ChunkGraph(const ChunkGraph &) = delete;
    ChunkGraph &operator=(const ChunkGraph &) = delete;

    template <class F>
    int run(int base, F &&launch_chunk) {
        fwi_ctx *c = c_;
        launch_set_int(c->si_base, base, c->st);
        if (!exec_) {
            const int prof = c->prof_on;
            c->prof_on = 0;   // no event records inside the capture
            hipError_t e = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                launch_chunk();
                hipGraph_t g = nullptr;
                e = hipStreamEndCapture(c->st, &g);
                if (e == hipSuccess) e = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
                if (g) hipGraphDestroy(g);
            }
            c->prof_on = prof;
            if (e != hipSuccess) {
                exec_ = nullptr;
                return c->fail(FWI_EHIP, "graph capture: %s", hipGetErrorString(e));
            }
        }
        Timer t(c, FWI_K_GRAPH);
        HIPCHK(hipGraphLaunch(exec_, c->st));
        return FWI_OK;
    }

  private:
    fwi_ctx *c_;
    hipGraphExec_t exec_ = nullptr;
};

int zero_state(fwi_ctx *c, int nb, const Wave &w, bool u_we) {
    for (int i = 0; i < 5; ++i) HIPCHK(hipMemsetAsync(w.f[i], 0, sizeof(float) * nb * c->fs(), c->st));
    for (int i = 0; i < 4; ++i) HIPCHK(hipMemsetAsync(w.psi[i], 0, sizeof(float) * nb * c->psiz(), c->st));
    for (int i = 4; i < 8; ++i) HIPCHK(hipMemsetAsync(w.psi[i], 0, sizeof(float) * nb * c->psix(), c->st));
    if (c->fusable) {   // the one-pass step never writes cells outside the interior: keep them zero
        for (int i = 0; i < 5; ++i) HIPCHK(hipMemsetAsync(w.f2[i], 0, sizeof(float) * nb * c->fs(), c->st));
        for (int i = 0; i < 2; ++i) HIPCHK(hipMemsetAsync(w.psi2[i], 0, sizeof(float) * nb * c->psiz(), c->st));
        for (int i = 2; i < 4; ++i) HIPCHK(hipMemsetAsync(w.psi2[i], 0, sizeof(float) * nb * c->psix(), c->st));
    }
    if (c->plain) {
        float *m[8] = {c->pb.mvz_z, c->pb.mvx_z, c->pb.mszz_z, c->pb.mszx_z,
                       c->pb.mvz_x, c->pb.mvx_x, c->pb.mszx_x, c->pb.msxx_x};
        for (int i = 0; i < 8; ++i) HIPCHK(hipMemsetAsync(m[i], 0, sizeof(float) * c->fs(), c->st));
    }
    if (u_we)
        for (int i = 5; i < 9; ++i) HIPCHK(hipMemsetAsync(c->f[i], 0, sizeof(float) * nb * c->fs(), c->st));
    return FWI_OK;
}

// copy the 13-array propagation state of nb shots to / from checkpoint k
int ckpt_copy(fwi_ctx *c, int k, int nb, const Wave &w, bool save) {
    char *base = reinterpret_cast<char *>(c->ck) + (size_t)k * c->ck_bytes_per_seg;
    size_t off = 0;
    for (int i = 0; i < 13; ++i) {
        float *p = i < 5 ? w.f[i] : w.psi[i - 5];
        const size_t per = i < 5 ? c->fs() : (i < 9 ? c->psiz() : c->psix());
        const size_t nbytes = sizeof(float) * per * nb;
        void *dst = save ? (void *)(base + off) : (void *)p;
        const void *src = save ? (const void *)p : (const void *)(base + off);
        HIPCHK(hipMemcpyAsync(dst, src, nbytes, hipMemcpyDeviceToDevice, c->st));
        off += sizeof(float) * per * c->B;
    }
    return FWI_OK;
}

int check_finite(fwi_ctx *c, int nb) {
    if (!c->o.check_finite) return FWI_OK;
    HIPCHK(hipMemsetAsync(c->nonfinite, 0, sizeof(int), c->st));
    dev_count_nonfinite(c->f[0], nb * c->fs(), c->nonfinite, c->st);
    dev_count_nonfinite(c->f[1], nb * c->fs(), c->nonfinite, c->st);
    int h = 0;
    HIPCHK(hipMemcpyAsync(&h, c->nonfinite, sizeof(int), hipMemcpyDeviceToHost, c->st));
    RETIF(sync(c));
    if (h) return c->fail(FWI_ENONFINITE, "%d non-finite velocity values after propagation", h);
    return FWI_OK;
}

// Autotuner: kernel variant, rows per thread (TZ) and work order of the two
// stencil kernels, timed on the batch (zeroed state, a few steps per candidate);
// the best is kept per (grid, order, batch) for the process.
struct TuneKey {
    int nz, nx, N, B;
    bool operator<(const TuneKey &o) const {
        return std::tie(nz, nx, N, B) < std::tie(o.nz, o.nx, o.N, o.B);
    }
};
struct TuneVal {
    int var, tz, depth, order;
};
std::map<TuneKey, TuneVal> g_tuned;
std::mutex g_tune_mu;

void take(fwi_ctx *c, const TuneVal &v) {
    c->var = v.var;
    c->tz = v.tz;
    c->depth = v.depth;
    c->order = v.order;
    c->tuned = true;
}

int autotune(fwi_ctx *c, int nb) {
    if (c->plain || c->tuned) return FWI_OK;
    const TuneKey key{c->g.nz, c->g.nx, c->N, nb};
    const int fv = c->o.kernel_variant;   // 0 = any
    const bool order_forced = c->o.work_order >= 1 && c->o.work_order <= 4;
    // candidates (variant, rows per segment, prefetch depth); the one-pass step only where it applies
    std::vector<TuneVal> cand;
    for (int var = 1; var <= 13; ++var) {
        if ((fv && var != fv) || (one_pass(var) && !c->fusable) || (var >= 5 && !c->ztab_ok)) continue;
        if (var == 9) {   // balanced: one or two waves of resident blocks
            for (int m : {1, 2})
                if (c->o.tile_tz <= 0 || c->o.tile_tz == m) cand.push_back({9, m, 1, 2});
            continue;
        }
        if (var >= 6 && var <= 8) {   // loads one row ahead, segments of 1, 2 (variant 6) or 4, 8 (variants 7, 8) blocks
            for (int m : var == 6 ? std::vector<int>{1, 2} : std::vector<int>{4, 8}) {
                const int tz = v6_block(c->N, 1) * m;
                if (c->o.tile_tz <= 0 || c->o.tile_tz == tz) cand.push_back({var, tz, 1, 0});
            }
            continue;
        }
        const std::vector<int> tzs = var == 4   ? std::vector<int>{32, 64}
                                     : var == 5 || var >= 10 ? std::vector<int>{16, 32}
                                                : std::vector<int>{8, 16, 32, 64};
        for (int tz : tzs)
            if (c->o.tile_tz <= 0 || c->o.tile_tz == tz) cand.push_back({var, tz, 1, 0});
    }
    if (cand.empty())
        return c->fail(FWI_EINVAL, "no stencil kernel for kernel_variant %d, tile_tz %d%s", fv, c->o.tile_tz,
                       fv == 4   ? " (the one-pass step takes tile_tz 32 or 64, no free surface other than "
                                   "the top one at row fdorder/2, and sources and receivers inside the interior)"
                       : fv == 5 || fv >= 10
                           ? " (variants 5 and 10 to 13 take tile_tz 16 or 32, and z CPML strips of at most 64 rows "
                             "at the top and bottom of the grid)"
                       : fv == 6 ? " (variant 6 takes tile_tz 0, fdorder + 2 or 2 fdorder + 4, and z CPML "
                                   "strips of at most 64 rows at the top and bottom of the grid)"
                       : fv == 7 || fv == 8 ? " (variants 7 and 8 take tile_tz 0, 4 fdorder + 8 or 8 fdorder + 16, "
                                   "variant 4's conditions and variant 6's z CPML strips)"
                                 : "");
    if (cand.size() == 1 && order_forced) {   // nothing left to choose
        take(c, {cand[0].var, cand[0].tz, cand[0].depth, c->o.work_order - 1});
        return FWI_OK;
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end() && (!fv || fv == it->second.var) && (!one_pass(it->second.var) || c->fusable) &&
            (!order_forced || c->o.work_order - 1 == it->second.order) &&
            (c->o.tile_tz <= 0 || c->o.tile_tz == it->second.tz)) {
            take(c, it->second);
            return FWI_OK;
        }
    }
    const Wave w = main_wave(c);
    RETIF(zero_state(c, nb, w, true));
    KArgs a = base_args(c, nb, w);   // the one-pass step reads f and writes f2 on every repetition
    a.energy = 1;
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // A gradient's mix of launches over snap_dt steps (at most 4): snap_dt forward steps,
    // snap_dt recomputed steps of which one stores its snapshot and snap_dt adjoint steps
    // of which one correlates with it (slab 0 of the snapshot buffers; the state, sums
    // and slab are overwritten or zeroed before any pass).
    const int kmix = std::min(std::max(c->g.snap_dt, 1), 4);
    KArgs ar = a, as = a, ap = a, ag = a;
    ar.energy = 0;
    as.energy = 0;
    as.store = 1;
    as.st_vz = c->snap[0]; as.st_vx = c->snap[1]; as.st_szz = c->snap[2]; as.st_szx = c->snap[3];
    as.st_sxx = c->snap[4];
    as.st_stride = c->snap_stride;
    ag.We = c->f[8];
    ag.displ = 1;
    ag.grad = 1;
    ag.sn_vz = c->snap[0]; ag.sn_vx = c->snap[1]; ag.sn_szz = c->snap[2]; ag.sn_szx = c->snap[3];
    ag.sn_sxx = c->snap[4];
    ag.sn_stride = c->snap_stride;
    ap.We = c->f[8];
    ap.displ = 1;
    std::vector<KArgs *> mix;
    for (int k = 0; k < kmix; ++k) mix.push_back(&a);
    for (int k = 0; k < kmix; ++k) mix.push_back(k ? &ar : &as);
    for (int k = 0; k < kmix; ++k) mix.push_back(k ? &ap : &ag);
    auto run = [&](const TuneVal &v) {
        for (KArgs *x : mix) {
            x->order = a.order;
            if (one_pass(v.var)) {
                launch_one_pass(c, *x, v.var, v.tz, v.depth);
            } else {
                launch_stress(*x, c->N, v.tz, v.var, c->st, &c->ztab, v.depth);
                launch_velocity(*x, c->N, v.tz, v.var, c->st, &c->ztab, v.depth);
            }
        }
    };
    float best = 1e30f;
    const int o_lo = order_forced ? c->o.work_order - 1 : 0, o_hi = order_forced ? o_lo + 1 : 4;
    TuneVal bv = cand[0];
    bv.order = o_lo;
    for (int order = o_lo; order < o_hi; ++order) {
        a.order = order;
        for (const auto &vt : cand) {
            if (vt.var == 9 && order != o_lo) continue;   // balanced launches ignore the work order
            run(vt);
            HIPCHK(hipEventRecord(e0, c->st));
            for (int rep = 0; rep < 2; ++rep) run(vt);
            HIPCHK(hipEventRecord(e1, c->st));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            char buf[64];
            std::snprintf(buf, sizeof buf, "o%d/v%d/tz%d/d%d %.3f; ", order, vt.var, vt.tz, vt.depth, ms / 2);
            c->tune_log += buf;
            if (ms < best) {
                best = ms;
                bv = {vt.var, vt.tz, vt.depth, order};
            }
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    take(c, bv);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = bv;
    return FWI_OK;
}

// Forward pass (kernel_PSV with grad = false) for shots [s0, s0 + nb).
int run_forward(fwi_ctx *c, int s0, int nb, bool store, bool record, bool energy) {
    RETIF(autotune(c, nb));
    const Wave w = main_wave(c);
    RETIF(zero_state(c, nb, w, true));
    if (record) HIPCHK(hipMemsetAsync(c->acc, 0, sizeof(double) * 2 * c->nrec_tot, c->st));
    KArgs a = base_args(c, nb, w);
    a.energy = energy;
    const SrcArgs sa = src_args(c, s0, nb);
    const RecArgs ra = rec_args(c, s0, nb);
    const bool ckpt = store && c->storage == FWI_STORE_CHECKPOINT;
    const int nt = c->g.nt;
    // graph chunks: a checkpoint segment, else kGraphChunk steps (even: the one-pass
    // step's two time levels are back in place after each chunk)
    const int K = ckpt ? c->S : kGraphChunk;
    const bool graph = use_graphs(c) && K % 2 == 0;
    ChunkGraph cg(c), cgk(c);   // cgk: the kept units of a checkpointed pass (they store)
    float *snap0[5] = {c->snap[0], c->snap[1], c->snap[2], c->snap[3], c->snap[4]};
    // snapshots go to slab (it - tb) / snap_dt: every step with full storage, from t_keep
    // (a unit boundary) on with checkpoints
    const int tb = ckpt ? c->t_keep : 0;
    for (int it = 0; it < nt;) {
        if (ckpt && it % c->S == 0 && c->spine_slot[it / c->S] >= 0) {
            Timer t(c, FWI_K_STORE);
            RETIF(ckpt_copy(c, c->spine_slot[it / c->S], nb, w, true));
        }
        const bool keep = store && it >= tb;
        if (graph && it + K <= nt) {
            KArgs ag = a;
            ag.si = c->si_fwd;
            ag.si_base = c->si_base;
            ag.si_dir = 1;
            RETIF((keep && ckpt ? cgk : cg).run(it, [&]() {
                for (int k = 0; k < K; ++k) {
                    ag.si_k = k;
                    time_step(c, ag, w, sa, it + k, keep ? snap0 : nullptr);
                    if (record) launch_record(ag, ra, it + k, c->st);
                }
            }));
            it += K;
            continue;
        }
        const size_t off = keep ? (size_t)((it - tb) / c->g.snap_dt) * c->snz * c->snx : 0;
        float *dst[5] = {c->snap[0] + off, c->snap[1] + off, c->snap[2] + off, c->snap[3] + off, c->snap[4] + off};
        const bool snap_step = keep && it % c->g.snap_dt == 0;
        time_step(c, a, w, sa, it, snap_step ? dst : nullptr);
        if (record) {
            Timer t(c, FWI_K_RECEIVER);
            launch_record(a, ra, it, c->st);
        }
        ++it;
    }
    c->last_s0 = s0;
    c->last_nb = nb;
    return check_finite(c, nb);
}

// One adjoint step (time index it); on snapshot steps the velocity kernel
// correlates with the forward slab at slab_off (fwi_grad2 fused), or the plain
// form runs fwi_grad2 as its own pass after the injection, as the reference does.
void adj_step(fwi_ctx *c, KArgs &a, const Wave &w, const SrcArgs &sa, int it, size_t slab_off) {
    if (a.si) {   // replayed graph: the step table gates the correlation and places the slab
        a.sn_vz = c->snap[0];
        a.sn_vx = c->snap[1];
        a.sn_szz = c->snap[2];
        a.sn_szx = c->snap[3];
        a.sn_sxx = c->snap[4];
        a.sn_stride = c->snap_stride;
        a.grad = 1;
        time_step(c, a, w, sa, it, nullptr);
        return;
    }
    const bool gstep = it % c->g.snap_dt == 0;
    if (gstep) {
        a.sn_vz = c->snap[0] + slab_off;
        a.sn_vx = c->snap[1] + slab_off;
        a.sn_szz = c->snap[2] + slab_off;
        a.sn_szx = c->snap[3] + slab_off;
        a.sn_sxx = c->snap[4] + slab_off;
        a.sn_stride = c->snap_stride;
    }
    a.grad = gstep && !c->plain;
    time_step(c, a, w, sa, it, nullptr);
    if (gstep && c->plain) {
        Timer t(c, FWI_K_OTHER);
        plain_grad(a, c->st);
    }
}

// The FP32 correlation sums go into their FP64 totals after the adjoint step of every time
// index t with (nt - t) % kFlushSteps == 0 and once at the end: the same time steps whatever
// the storage, so full storage, checkpoints and graph replay group the FP32 partial sums alike.
constexpr int kFlushSteps = 64;
static_assert(kGraphChunk == kFlushSteps, "a full-storage graph chunk ends on a flush step");

bool flush_after(int nt, int it) { return (nt - it) % kFlushSteps == 0; }

// the batch's FP32 correlation sums added into their FP64 totals and zeroed
void flush_sums(fwi_ctx *c, int nb) {
    Timer t(c, FWI_K_OTHER);
    const size_t cells = (size_t)c->B * c->snz * c->snx;
    float *g[5] = {c->gr[0], c->gr[1], c->gr[2], c->gx, c->gx + cells};
    dev_flush_sums(c->gr64, g, c->gform ? 5 : 3, (size_t)nb * c->snz * c->snx, cells, c->st);
}

// the same inside a captured chunk of adjoint steps: step k of the chunk is time index
// *si_base - k, known only at replay, so the kernel itself decides whether to flush
void flush_sums_at(fwi_ctx *c, int nb, int k) {
    const size_t cells = (size_t)c->B * c->snz * c->snx;
    float *g[5] = {c->gr[0], c->gr[1], c->gr[2], c->gx, c->gx + cells};
    dev_flush_sums_at(c->gr64, g, c->gform ? 5 : 3, (size_t)nb * c->snz * c->snx, cells, c->si_base, -1, k,
                      c->g.nt, kFlushSteps, c->st);
}

// Adjoint pass (kernel_PSV with grad = true): residuals in res_* are injected at
// the receivers in reversed time, each shot's scaled by its alpha (a power of two).
int run_adjoint(fwi_ctx *c, int s0, int nb) {
    const Wave w = main_wave(c);
    RETIF(zero_state(c, nb, w, false));
    for (int i = 5; i < 7; ++i) HIPCHK(hipMemsetAsync(c->f[i], 0, sizeof(float) * nb * c->fs(), c->st));
    HIPCHK(hipMemsetAsync(c->f[8], 0, sizeof(float) * nb * c->fs(), c->st));
    for (int i = 0; i < 3; ++i)
        HIPCHK(hipMemsetAsync(c->gr[i], 0, sizeof(float) * nb * c->snz * c->snx, c->st));
    if (c->gform) HIPCHK(hipMemsetAsync(c->gx, 0, sizeof(float) * 2 * c->B * c->snz * c->snx, c->st));
    HIPCHK(hipMemsetAsync(c->gr64, 0, sizeof(double) * 5 * c->B * c->snz * c->snx, c->st));
    KArgs a = base_args(c, nb, w);
    a.We = c->f[8];
    a.energy = 1;
    a.displ = 1;
    const SrcArgs sa = adj_args(c, s0, nb);
    const int nt = c->g.nt, sdt = c->g.snap_dt;
    const size_t slab = (size_t)c->snz * c->snx;
    KArgs ag = a;   // adjoint steps of a replayed graph: time index base - k
    ag.si = c->si_adj;
    ag.si_base = c->si_base;
    ag.si_dir = -1;
    ChunkGraph cga(c);
    if (c->storage != FWI_STORE_CHECKPOINT) {
        const int K = kGraphChunk;
        const bool graph = use_graphs(c);
        for (int jt = 0; jt < nt;) {
            const int it = nt - 1 - jt;
            if (graph && jt + K <= nt) {
                RETIF(cga.run(it, [&]() {
                    for (int k = 0; k < K; ++k) {
                        ag.si_k = k;
                        adj_step(c, ag, w, sa, it - k, 0);
                    }
                }));
                jt += K;
                flush_sums(c, nb);   // jt is a multiple of K = kFlushSteps
                continue;
            }
            adj_step(c, a, w, sa, it, (size_t)(it / sdt) * slab);
            if (flush_after(nt, it)) flush_sums(c, nb);
            ++jt;
        }
        if (nt % kFlushSteps) flush_sums(c, nb);
    } else {
        const Wave rw = recompute_wave(c);
        RETIF(zero_state(c, nb, rw, false));   // cells outside the interior stay zero in both levels
        KArgs r = base_args(c, nb, rw);
        const SrcArgs fsa = src_args(c, s0, nb);
        KArgs rg = r;   // recomputed steps of a replayed graph: time index base + k
        rg.si = c->si_rec;
        rg.si_base = c->si_base;
        rg.si_dir = 1;
        ChunkGraph cgr(c), cgv(c);
        const bool graph = use_graphs(c) && c->S % 2 == 0;
        float *snap0[5] = {c->snap[0], c->snap[1], c->snap[2], c->snap[3], c->snap[4]};
        const int S = c->S, W = c->slots + 1;
        // the recompute state at the start of a unit: zero at unit 0, else a slot's copy;
        // the level order is reset first, so replayed graphs (captured in that order) and
        // a short last segment that left it swapped agree
        auto restore = [&](int hold) -> int {
            if (rw.f[0] != c->rf_canon) rw.flip();
            if (hold < 0) return zero_state(c, nb, rw, false);
            Timer t(c, FWI_K_STORE);
            return ckpt_copy(c, hold, nb, rw, false);
        };
        auto save = [&](int slot) -> int {
            Timer t(c, FWI_K_STORE);
            return ckpt_copy(c, slot, nb, rw, true);
        };
        // units [u0, u0 + k) forward from the state in rw, nothing stored (never the last unit)
        auto advance = [&](int u0, int k) -> int {
            for (int u = u0; u < u0 + k; ++u) {
                const int t0 = u * S;
                if (graph) {
                    RETIF(cgv.run(t0, [&]() {
                        for (int j = 0; j < S; ++j) {
                            rg.si_k = j;
                            time_step(c, rg, rw, fsa, t0 + j, nullptr);
                        }
                    }));
                    continue;
                }
                for (int it = t0; it < t0 + S; ++it) time_step(c, r, rw, fsa, it, nullptr);
            }
            return FWI_OK;
        };
        // the adjoint steps of unit u, its snapshots in the buffer from time step tb on
        auto adjoint_unit = [&](int u, int tb) -> int {
            const int t0 = u * S, t1 = std::min(t0 + S, nt);
            if (graph && t1 - t0 == S) {   // the step table places the slabs
                return cga.run(t1 - 1, [&]() {
                    for (int j = 0; j < S; ++j) {
                        ag.si_k = j;
                        adj_step(c, ag, w, sa, t1 - 1 - j, 0);
                        flush_sums_at(c, nb, j);
                    }
                });
            }
            for (int it = t1 - 1; it >= t0; --it) {
                adj_step(c, a, w, sa, it, (size_t)((it - tb) / sdt) * slab);
                if (flush_after(nt, it)) flush_sums(c, nb);
            }
            return FWI_OK;
        };
        // unit u recomputed from the state in rw into the segment buffer, then its adjoint steps
        auto leaf = [&](int u) -> int {
            const int t0 = u * S, t1 = std::min(t0 + S, nt);
            if (graph && t1 - t0 == S) {
                RETIF(cgr.run(t0, [&]() {
                    for (int j = 0; j < S; ++j) {
                        rg.si_k = j;
                        time_step(c, rg, rw, fsa, t0 + j, snap0);
                    }
                }));
                return adjoint_unit(u, t0);
            }
            for (int it = t0; it < t1; ++it) {
                const size_t off = (size_t)((it - t0) / sdt) * slab;
                float *dst[5] = {c->snap[0] + off, c->snap[1] + off, c->snap[2] + off, c->snap[3] + off,
                                 c->snap[4] + off};
                time_step(c, r, rw, fsa, it, it % sdt == 0 ? dst : nullptr);
            }
            return adjoint_unit(u, t0);
        };
        // the leaves of units [a0, a0 + n) in reverse order; the state at a0 is in slot hold
        // (-1: the zero state of unit 0) and slots [slots - s, slots) are free.  top() is the
        // chain the forward pass walked: its split states are already saved.
        std::function<int(int, int, int, int)> rev, top;
        auto restart_each = [&](int a0, int n, int hold) -> int {   // no free slot
            for (int j = n - 1; j >= 0; --j) {
                RETIF(restore(hold));
                RETIF(advance(a0, j));
                RETIF(leaf(a0 + j));
            }
            return FWI_OK;
        };
        rev = [&](int a0, int n, int sfree, int hold) -> int {
            if (n == 1) {
                RETIF(restore(hold));
                return leaf(a0);
            }
            if (sfree == 0) return restart_each(a0, n, hold);
            const int m = c->ck_rev[(size_t)n * W + sfree], slot = c->slots - sfree;
            RETIF(restore(hold));
            RETIF(advance(a0, m));
            RETIF(save(slot));
            RETIF(rev(a0 + m, n - m, sfree - 1, slot));
            return rev(a0, m, sfree, hold);
        };
        top = [&](int a0, int n, int sfree, int hold) -> int {
            if (n == 1) {
                RETIF(restore(hold));
                return leaf(a0);
            }
            if (sfree == 0) return restart_each(a0, n, hold);
            const int m = c->ck_top[(size_t)n * W + sfree], slot = c->slots - sfree;
            RETIF(top(a0 + m, n - m, sfree - 1, slot));
            return rev(a0, m, sfree, hold);
        };
        // the kept units first (the forward pass stored them), then the others' reversal
        for (int u = c->nseg - 1; u >= c->nrev; --u) RETIF(adjoint_unit(u, c->t_keep));
        if (c->nrev > 0) RETIF(top(0, c->nrev, c->slots, -1));
        if (nt % kFlushSteps) flush_sums(c, nb);
    }
    c->last_s0 = s0;
    c->last_nb = nb;
    return check_finite(c, nb);
}

// power of two bringing max|x| into [1, 2)
double pow2_scale(double m) {
    if (!(m > 0) || !std::isfinite(m)) return 1.0;
    int e;
    std::frexp(m, &e);
    return std::ldexp(1.0, 1 - e);
}

int absmax_host(fwi_ctx *c, const double *x, size_t n, double *out) {
    dev_absmax(x, n, c->red, c->st);
    HIPCHK(hipMemcpyAsync(out, c->red, sizeof(double), hipMemcpyDeviceToHost, c->st));
    return sync(c);
}

int batches(const fwi_ctx *c, int shot0, int nshots, std::vector<std::pair<int, int>> &out) {
    out.clear();
    if (shot0 < 0 || nshots < 0 || shot0 + nshots > c->nshot) return FWI_EINVAL;
    const int nbat = (nshots + c->B - 1) / c->B;
    for (int b = 0, s = shot0; b < nbat; ++b) {
        const int left = shot0 + nshots - s, rem = nbat - b;
        const int nb = (left + rem - 1) / rem;   // balanced batches
        out.push_back({s, nb});
        s += nb;
    }
    return FWI_OK;
}

// forward + residual + adjoint for one batch; leaves raw gradients in gr, each
// shot's scaled by alpha[shot] (and the lam / mu sums by kGainLM)
// The seismograms of shots [s0, s0 + nb) from the FP64 propagation into syn_* (seis_fp64).
int run_forward64(fwi_ctx *c, int s0, int nb) {
    const int nz = c->g.nz, nx = c->g.nx, P = c->P, nt = c->g.nt;
    const size_t fs = c->fs();
    if (!c->d64) return c->fail(FWI_ESTATE, "the FP64 propagation's buffers were not planned (seis_fp64)");
    double *med[5];
    for (int k = 0; k < 5; ++k) med[k] = c->dmed + (size_t)k * fs;
    if (c->d64_ver != c->media_ver) {
        d_media(c->m_lam, c->m_mu, c->m_rho, med[0], med[1], med[2], med[3], med[4], nz, nx, P, c->st);
        c->d64_ver = c->media_ver;
    }
    DArgs d{};
    d.nz = nz; d.nx = nx; d.P = P; d.B = nb; d.N = c->N;
    d.pml_z = c->g.pml_z; d.pml_x = c->g.pml_x;
    d.fs = fs;
    d.dt = c->g.dt;
    for (int k = 0; k < 6; ++k) {
        d.cz[k] = k < c->N ? c->g.hc[k + 1] / c->g.dz : 0.0;
        d.cx[k] = k < c->N ? c->g.hc[k + 1] / c->g.dx : 0.0;
    }
    d.lam = med[0]; d.mu = med[1]; d.muzx = med[2]; d.rhoz = med[3]; d.rhox = med[4];
    d.zp = c->dprof;
    d.xp = c->dprof + (size_t)6 * nz;
    double **slot[15] = {&d.vz, &d.vx, &d.szz, &d.szx, &d.sxx, &d.pv[0], &d.pv[1], &d.pv[2], &d.pv[3],
                         &d.ps[0], &d.ps[1], &d.ps[2], &d.ps[3], &d.E0, &d.E3};
    for (int k = 0; k < 15; ++k) *slot[k] = c->d64 + (size_t)k * c->B * fs;
    for (int k = 0; k < 13; ++k) HIPCHK(hipMemsetAsync(*slot[k], 0, sizeof(double) * nb * fs, c->st));
    HIPCHK(hipMemsetAsync(c->acc, 0, sizeof(double) * 2 * c->nrec_tot, c->st));
    const SrcArgs sa = src_args(c, s0, nb);
    const RecArgs ra = rec_args(c, s0, nb);
    int isurf[4];
    for (int k = 0; k < 4; ++k) isurf[k] = c->g.isurf[k];
    for (int it = 0; it < nt; ++it) {
        d_forward_step(d, isurf, c->st);
        d_inject(d, sa, it, c->st);
        d_record(d, ra, it, c->st);
    }
    if (hipGetLastError() != hipSuccess) return c->fail(FWI_EHIP, "FP64 propagation launch");
    return FWI_OK;
}

int batch_gradient(fwi_ctx *c, int s0, int nb, const double *adj_z, const double *adj_x, double *misfit_out) {
    const bool seis64 = c->o.seis_fp64 && adj_z == nullptr;   // residuals from the FP64 propagation
    RETIF(run_forward(c, s0, nb, true, adj_z == nullptr && !seis64, true));
    if (seis64) RETIF(run_forward64(c, s0, nb));
    const int r0 = c->rec_off[s0], r1 = c->rec_off[s0 + nb];
    const size_t rows = (size_t)(r1 - r0) * c->g.nt, off = (size_t)r0 * c->g.nt;
    if (adj_z && c->adj_on_device) {   // the caller's device arrays, read by a kernel
        dev_copy(c->res_z + off, adj_z, rows, c->st);
        dev_copy(c->res_x + off, adj_x, rows, c->st);
    } else if (adj_z) {
        HIPCHK(hipMemcpyAsync(c->res_z + off, adj_z, sizeof(double) * rows, hipMemcpyHostToDevice, c->st));
        HIPCHK(hipMemcpyAsync(c->res_x + off, adj_x, sizeof(double) * rows, hipMemcpyHostToDevice, c->st));
    } else {
        dev_residual(c->syn_z, c->syn_x, c->obs_z, c->obs_x, c->res_z, c->res_x, c->rec_off_d + s0, nb, c->g.nt,
                     c->g.dt, c->misfit, c->st);
        if (misfit_out)
            HIPCHK(hipMemcpyAsync(misfit_out, c->misfit, sizeof(double) * nb, hipMemcpyDeviceToHost, c->st));
    }
    (void)rows;
    std::vector<double> mx(nb);
    dev_shot_absmax(c->res_z, c->res_x, c->rec_off_d + s0, nb, c->g.nt, c->shotmax, c->st);
    HIPCHK(hipMemcpyAsync(mx.data(), c->shotmax, sizeof(double) * nb, hipMemcpyDeviceToHost, c->st));
    RETIF(sync(c));
    for (int s = 0; s < nb; ++s) c->alpha[s0 + s] = pow2_scale(mx[s]);
    RETIF(upload(c, c->alpha_d + s0, c->alpha.data() + s0, nb));
    return run_adjoint(c, s0, nb);
}

int refresh_media(fwi_ctx *c) {
    ++c->media_ver;
    dev_mat_av(c->m_lam, c->m_mu, c->m_rho, c->lam, c->mu, c->muzx, c->rhoz, c->rhox, c->red, c->g.nz, c->g.nx,
               c->P, c->st);
    HIPCHK(hipMemcpyAsync(c->C, c->red, sizeof(double) * 3, hipMemcpyDeviceToHost, c->st));
    return sync(c);
}

// ---------------------------------------------------------------------------
// planning and allocation
// ---------------------------------------------------------------------------
struct Need {
    size_t shared, per_shot, per_shot_fp64;
};

Need need_bytes(const fwi_ctx *c, int storage, int S, int slots, int keep = 1) {
    const size_t nz = c->g.nz, nx = c->g.nx, nt = c->g.nt, fs = c->fs();
    const size_t snapcell = (size_t)c->snz * c->snx;
    Need n{};
    n.shared = 5 * fs * 4 + 9 * nz * nx * 8 + 2 * nz * nx * 8 + (size_t)c->nrec_tot * nt * 8 * 6 +
               (size_t)c->nsrc_tot * nt * 8 * 2 + (64 << 20);
    if (c->o.param == FWI_PARAM_VELOCITY) n.shared += 4 * nz * nx * 8;
    if (c->o.seis_fp64) {   // the FP64 propagation: media and profiles, 15 fields per shot
        n.shared += 5 * fs * 8 + 6 * (nz + nx) * 8;
        n.per_shot_fp64 = 15 * fs * 8;
    }
    const size_t state_ck = 5 * fs * 4 + 4 * (c->psiz() + c->psix()) * 4;
    n.per_shot = 9 * fs * 4 + 8 * (c->psiz() + c->psix()) * 4 + 5 * snapcell * 4 + 5 * snapcell * 8 +
                 n.per_shot_fp64;
    if (c->plain) n.per_shot += 12 * fs * 4;
    // second time level of the one-pass step (5 fields, 4 stress-side psi)
    const size_t level2 = c->fusable ? (5 * fs + 2 * (c->psiz() + c->psix())) * 4 : 0;
    n.per_shot += level2;
    if (storage == FWI_STORE_FULL) {
        n.per_shot += 5 * (size_t)c->snt * snapcell * 4;
    } else {
        n.per_shot += state_ck + level2 + (size_t)slots * state_ck +
                      (size_t)std::max(keep, 1) * 5 * (size_t)(S / c->g.snap_dt) * snapcell * 4;
    }
    return n;
}

// Binomial checkpointing over units of S time steps (Griewank's revolve, the splits
// chosen by dynamic programming on a unit-advance cost).  rev(n, s): units advanced to
// run the leaves of a chain of n units in reverse order with the state at its first
// unit held and s free slots: a single unit needs none; with no free slot every leaf
// restarts from the held state (n (n - 1) / 2); otherwise advance m units, save, reverse
// the right part with s - 1 slots, then the left part with s.  top(n, s): the chain the
// forward pass walks, whose advances are that pass itself.  Returns top(n, s) and fills
// the best splits (single level, s >= n - 1: every boundary saved, nothing advanced).
long long plan_schedule(int n, int s, std::vector<int> &top_m, std::vector<int> &rev_m) {
    const int W = s + 1;
    top_m.assign((size_t)(n + 1) * W, 1);
    rev_m.assign((size_t)(n + 1) * W, 1);
    if (s >= n - 1) return 0;
    std::vector<long long> rev((size_t)(n + 1) * W, 0), top((size_t)(n + 1) * W, 0);
    auto at = [W](int k, int j) { return (size_t)k * W + j; };
    for (int k = 2; k <= n; ++k) {
        rev[at(k, 0)] = top[at(k, 0)] = (long long)k * (k - 1) / 2;
        for (int j = 1; j <= s; ++j) {
            long long br = -1, bt = -1;
            int mr = 1, mt = 1;
            for (int m = 1; m < k; ++m) {
                const long long r = m + rev[at(k - m, j - 1)] + rev[at(m, j)];
                const long long t = top[at(k - m, j - 1)] + rev[at(m, j)];
                if (br < 0 || r < br) br = r, mr = m;
                if (bt < 0 || t < bt) bt = t, mt = m;
            }
            rev[at(k, j)] = br;
            top[at(k, j)] = bt;
            rev_m[at(k, j)] = mr;
            top_m[at(k, j)] = mt;
        }
    }
    return top[at(n, s)];
}

constexpr int kMaxUnits = 2048;   // the schedule's dynamic programme is O(units^2 slots)

int plan(fwi_ctx *c) {
    size_t freeb = 0, total = 0;
    HIPCHK(hipMemGetInfo(&freeb, &total));
    const size_t budget = c->o.mem_budget > 0 ? (size_t)c->o.mem_budget : (size_t)(0.9 * (double)freeb);
    const int sdt = c->g.snap_dt, nt = c->g.nt;
    const bool even = c->o.use_graph != 0;   // replayed segments need an even step count
    auto leaf_len = [&](int S_) {            // a multiple of snap_dt, at most the time axis rounded up
        S_ = std::max(sdt, (S_ + sdt - 1) / sdt * sdt);
        S_ = std::min(S_, (nt + sdt - 1) / sdt * sdt);
        if (even && S_ % 2) S_ += sdt;       // (snap_dt is then odd)
        return S_;
    };
    auto units_of = [&](int S_) { return (nt + S_ - 1) / S_; };
    auto slots_cap = [&](int S_) {           // every unit boundary, or fewer as asked
        const int full = units_of(S_) - 1;
        return c->o.ckpt_slots > 0 ? std::min(full, c->o.ckpt_slots) : full;
    };
    // segment length: minimise nseg * state + S/snap_dt * snapshot
    int S = c->o.ckpt_interval;
    if (S <= 0) {
        const double state_ck = 5.0 * c->fs() + 4.0 * (c->psiz() + c->psix());
        const double snapb = 5.0 * c->snz * c->snx;
        S = (int)std::lround(std::sqrt((double)nt * state_ck * sdt / snapb));
    }
    S = leaf_len(S);
    int slots = slots_cap(S);
    int storage = c->o.storage;
    if (c->plain) {
        if (storage == FWI_STORE_CHECKPOINT) return c->fail(FWI_EINVAL, "the plain form stores full snapshots");
        storage = FWI_STORE_FULL;
    }
    auto max_batch = [&](int st_, int S_, int sl_) {
        Need n = need_bytes(c, st_, S_, sl_);
        return n.shared + n.per_shot > budget ? 0 : (int)std::min<size_t>((budget - n.shared) / n.per_shot, c->nshot);
    };
    if (storage == FWI_STORE_AUTO) {
        // Full storage moves 40 B per cell-step (store + load) against the ~100 B of a
        // recomputed step, but its footprint caps the batch; batching is worth more
        // than the recomputation, so full storage is kept only if it still allows
        // the batch checkpointing would (or every shot).
        const int bf = max_batch(FWI_STORE_FULL, S, slots), bc = max_batch(FWI_STORE_CHECKPOINT, S, slots);
        const int want = c->o.batch > 0 ? std::min(c->o.batch, c->nshot) : c->nshot;
        storage = (bf >= std::min(want, bc) && bf > 0) ? FWI_STORE_FULL : FWI_STORE_CHECKPOINT;
    }
    if (storage == FWI_STORE_CHECKPOINT && max_batch(storage, S, slots) == 0) {
        // One shot with a checkpoint per segment does not fit: fewer slots and a binomial
        // schedule, over the segment lengths that leave room for at least one slot, the
        // one that recomputes the fewest steps.
        const Need n0 = need_bytes(c, storage, S, 0), n1 = need_bytes(c, storage, S, 1);
        const size_t state_ck = n1.per_shot - n0.per_shot;
        long long best = -1;
        int bS = 0, bsl = 0;
        for (int div = 1; div <= 16; div *= 2) {
            const int S_ = leaf_len(S / div);
            const int n_ = units_of(S_);
            const Need z = need_bytes(c, storage, S_, 0);
            if (n_ > kMaxUnits || z.shared + z.per_shot > budget) continue;
            const int sl = (int)std::min<size_t>((budget - z.shared - z.per_shot) / state_ck, (size_t)slots_cap(S_));
            std::vector<int> tm, rm;
            const long long cost = ((long long)n_ + plan_schedule(n_, sl, tm, rm)) * S_;
            if (best < 0 || cost < best) best = cost, bS = S_, bsl = sl;
            if (S_ == sdt) break;
        }
        if (best < 0) {
            const Need n = need_bytes(c, storage, S, 0);
            return c->fail(FWI_ENOMEM, "one shot needs %.2f GB even without checkpoints, budget %.2f GB",
                           (n.shared + n.per_shot) / 1e9, budget / 1e9);
        }
        S = bS;
        slots = bsl;
    }
    Need n = need_bytes(c, storage, S, slots);
    if (n.shared + n.per_shot > budget)
        return c->fail(FWI_ENOMEM, "one shot needs %.2f GB, budget %.2f GB", (n.shared + n.per_shot) / 1e9,
                       budget / 1e9);
    int Bmax = (int)std::min<size_t>((budget - n.shared) / n.per_shot, (size_t)c->nshot);
    if (c->o.batch > 0) Bmax = std::min(Bmax, c->o.batch);
    if (c->plain) Bmax = 1;
    Bmax = std::max(1, Bmax);
    const int nbat = (c->nshot + Bmax - 1) / Bmax;
    c->B = (c->nshot + nbat - 1) / nbat;
    c->storage = storage;
    c->S = S;
    c->nseg = units_of(S);
    c->slots = 0;
    c->keep = 0;
    c->nrev = 0;
    c->t_keep = nt;
    c->spine_slot.assign(c->nseg, -1);
    if (storage == FWI_STORE_CHECKPOINT) {
        // Units whose snapshots the forward pass keeps: the last one always (the segment
        // buffer holds it); with a checkpoint at every boundary, more while the budget left
        // beside the batch allows, each costing one unit of snapshots less one checkpoint.
        int keep = 1;
        const int cap = c->o.ckpt_keep > 0 ? std::min(c->o.ckpt_keep, c->nseg) : c->nseg;
        if (cap > 1 && slots >= c->nseg - 1) {
            const Need k1 = need_bytes(c, storage, S, slots, 1), k2 = need_bytes(c, storage, S, slots - 1, 2);
            const size_t used = k1.shared + (size_t)c->B * k1.per_shot;
            if (k2.per_shot <= k1.per_shot) {
                keep = cap;
            } else if (used <= budget) {
                const size_t more = (budget - used) / ((size_t)c->B * (k2.per_shot - k1.per_shot));
                keep = (int)std::min<size_t>((size_t)cap, 1 + more);
            }
        }
        c->keep = keep;
        c->nrev = c->nseg - keep;
        c->t_keep = c->nrev * S;
        c->slots = std::min(slots, std::max(c->nrev - 1, 0));
        if (c->nrev > kMaxUnits && c->slots < c->nrev - 1)
            return c->fail(FWI_EINVAL, "%d segments with %d checkpoints: more than %d segments need one each",
                           c->nrev, c->slots, kMaxUnits);
        c->extra_units = plan_schedule(c->nrev, c->slots, c->ck_top, c->ck_rev);
        // the states the forward pass saves: the splits along top()'s first chain
        for (int a0 = 0, k = c->nrev, j = c->slots; k > 1 && j > 0; --j) {
            const int m = c->slots >= c->nrev - 1 ? 1 : c->ck_top[(size_t)k * (c->slots + 1) + j];
            c->spine_slot[a0 + m] = c->slots - j;
            a0 += m;
            k -= m;
        }
    }
    return FWI_OK;
}

int allocate(fwi_ctx *c) {
    const int nz = c->g.nz, nx = c->g.nx, nt = c->g.nt, B = c->B;
    const size_t fs = c->fs(), nn = (size_t)nz * nx, snapcell = (size_t)c->snz * c->snx;
    bool ok = true;
    auto F = [&](size_t n) { float *p = c->dalloc<float>(n); ok = ok && p; return p; };
    auto D = [&](size_t n) { double *p = c->dalloc<double>(n); ok = ok && p; return p; };
    auto I = [&](size_t n) { int *p = c->dalloc<int>(n); ok = ok && p; return p; };
    c->lam = F(fs); c->mu = F(fs); c->muzx = F(fs); c->rhoz = F(fs); c->rhox = F(fs);
    c->m_lam = D(nn); c->m_mu = D(nn); c->m_rho = D(nn);
    c->o_lam = D(nn); c->o_mu = D(nn); c->o_rho = D(nn);
    for (int i = 0; i < 3; ++i) c->G[i] = D(nn);
    c->tmp = D(2 * nn);
    c->red = D(4 + 3 * 1024);
    if (c->o.param == FWI_PARAM_VELOCITY) c->vel = D(4 * nn);
    c->nonfinite = I(1);
    c->zmap = I(nz);
    c->xmap = I(nx);
    for (int i = 0; i < 6; ++i) c->pf[i] = F(nz);
    for (int i = 6; i < 12; ++i) c->pf[i] = F(nx);
    c->zp = c->dalloc<CpmlPoint>(nz);
    c->xp = c->dalloc<CpmlPoint>(nx);
    c->si_fwd = c->dalloc<StepInfo>(nt);
    c->si_rec = c->dalloc<StepInfo>(nt);
    c->si_adj = c->dalloc<StepInfo>(nt);
    c->si_base = c->dalloc<int>(1);
    ok = ok && c->zp && c->xp && c->si_fwd && c->si_rec && c->si_adj && c->si_base;
    const size_t recrows = (size_t)c->nrec_tot * nt;
    c->syn_z = D(recrows); c->syn_x = D(recrows);
    c->obs_z = D(recrows); c->obs_x = D(recrows);
    c->res_z = D(recrows); c->res_x = D(recrows);
    c->acc = D(2 * (size_t)c->nrec_tot);
    c->misfit = D(c->nshot);
    c->alpha_d = D(c->nshot);
    c->shotmax = D(c->nshot);
    c->alpha.assign(c->nshot, 1.0);
    for (int i = 0; i < 9; ++i) c->f[i] = F(B * fs);
    for (int i = 0; i < 4; ++i) c->psi[i] = F(B * c->psiz());
    for (int i = 4; i < 8; ++i) c->psi[i] = F(B * c->psix());
    for (int i = 0; i < 3; ++i) c->gr[i] = F(B * snapcell);
    c->gx = F(2 * B * snapcell);
    c->gr64 = D(5 * (size_t)B * snapcell);
    auto level2 = [&](float **f2, float **psi2) {
        for (int i = 0; i < 5; ++i) f2[i] = F(B * fs);
        for (int i = 0; i < 2; ++i) psi2[i] = F(B * c->psiz());
        for (int i = 2; i < 4; ++i) psi2[i] = F(B * c->psix());
    };
    if (c->fusable) level2(c->f2, c->psi2);
    if (c->storage == FWI_STORE_FULL) {
        c->snap_stride = (size_t)c->snt * snapcell;
    } else {
        c->snap_stride = (size_t)c->keep * (c->S / c->g.snap_dt) * snapcell;   // >= one unit
        for (int i = 0; i < 5; ++i) c->rf[i] = F(B * fs);
        for (int i = 0; i < 4; ++i) c->rpsi[i] = F(B * c->psiz());
        for (int i = 4; i < 8; ++i) c->rpsi[i] = F(B * c->psix());
        if (c->fusable) level2(c->rf2, c->rpsi2);
        c->ck_bytes_per_seg = sizeof(float) * B * (5 * fs + 4 * (c->psiz() + c->psix()));
        c->ck = F(c->ck_bytes_per_seg / sizeof(float) * std::max(c->slots, 1));
        c->rf_canon = c->rf[0];
    }
    for (int i = 0; i < 5; ++i) c->snap[i] = F(B * c->snap_stride);
    if (c->o.seis_fp64) {
        c->d64 = D(15 * (size_t)B * fs);
        c->dmed = D(5 * fs);
        c->dprof = D(6 * ((size_t)c->g.nz + c->g.nx));
    }
    if (c->plain) {
        PlainBufs &p = c->pb;
        float **m[12] = {&p.dz_z, &p.dx_z, &p.dz_x, &p.dx_x, &p.mvz_z, &p.mvx_z,
                         &p.mszz_z, &p.mszx_z, &p.mvz_x, &p.mvx_x, &p.mszx_x, &p.msxx_x};
        for (int i = 0; i < 12; ++i) *m[i] = F(fs);
        p.fbz = c->pf[0]; p.faz = c->pf[1]; p.fikz = c->pf[2];
        p.fbhz = c->pf[3]; p.fahz = c->pf[4]; p.fikhz = c->pf[5];
        p.fbx = c->pf[6]; p.fax = c->pf[7]; p.fikx = c->pf[8];
        p.fbhx = c->pf[9]; p.fahx = c->pf[10]; p.fikhx = c->pf[11];
    }
    if (!ok) return c->fail(FWI_ENOMEM, "device allocation failed after %.2f GB", c->bytes / 1e9);
    return FWI_OK;
}

void free_all(fwi_ctx *c) {
    for (void *p : c->allocs) hipFree(p);
    c->allocs.clear();
}

// groups of (shot, cell) for deterministic injection in source order
void build_groups(const std::vector<int> &off, const int32_t *z, const int32_t *x, int P, std::vector<int> &gshot,
                  std::vector<int> &gcell, std::vector<int> &goff, std::vector<int> &gidx, std::vector<int> &first) {
    const int nshot = (int)off.size() - 1;
    first.assign(nshot + 1, 0);
    goff.clear();
    for (int s = 0; s < nshot; ++s) {
        first[s] = (int)gshot.size();
        std::vector<std::pair<int, int>> cells;   // (cell, index) in source order
        for (int i = off[s]; i < off[s + 1]; ++i) cells.push_back({z[i] * P + x[i], i});
        std::vector<char> used(cells.size(), 0);
        for (size_t a = 0; a < cells.size(); ++a) {
            if (used[a]) continue;
            gshot.push_back(s);
            gcell.push_back(cells[a].first);
            goff.push_back((int)gidx.size());
            for (size_t b = a; b < cells.size(); ++b)
                if (!used[b] && cells[b].first == cells[a].first) {
                    used[b] = 1;
                    gidx.push_back(cells[b].second);
                }
        }
    }
    first[nshot] = (int)gshot.size();
    goff.push_back((int)gidx.size());
}

int setup_device(fwi_ctx *c, const fwi_acq *q) {
    const int nz = c->g.nz, nx = c->g.nx, N = c->N;
    // CPML active rows / columns inside the interior
    std::vector<int> zmap(nz, -1), xmap(nx, -1);
    auto active = [](const fwi_ctx *c_, int base, int i) {
        const auto &p = c_->prof;
        return p[base + 0][i] != 0.0 || p[base + 3][i] != 0.0 || p[base + 2][i] != 1.0 || p[base + 5][i] != 1.0;
    };
    c->nzs = 0;
    if (c->g.pml_z)
        for (int i = N; i < nz - N; ++i)
            if (active(c, 0, i)) zmap[i] = c->nzs++;
    c->nxs = 0;
    if (c->g.pml_x)
        for (int i = N; i < nx - N; ++i)
            if (active(c, 6, i)) xmap[i] = c->nxs++;
    // variant 5 takes the z strips as the row ranges [N, lo) and [hi, nz - N)
    {
        int lo = N, hi = nz - N;
        while (lo < nz - N && zmap[lo] >= 0) ++lo;
        while (hi > lo && zmap[hi - 1] >= 0) --hi;
        bool ok = c->nzs <= kZTabRows;
        for (int i = lo; ok && i < hi; ++i) ok = zmap[i] < 0;
        c->ztab_ok = ok;
        c->ztab.lo = lo;
        c->ztab.hi = hi;
    }
    // The one-pass step handles a free surface only at the top row N (no image rows),
    // and injects only into interior cells (it never writes the others).
    auto interior = [&](int z, int x) { return z >= N && z < nz - N && x >= N && x < nx - N; };
    c->fusable = !c->plain && c->g.isurf[1] == 0 && c->g.isurf[2] == 0 && c->g.isurf[3] == 0 &&
                 (c->g.isurf[0] == 0 || c->g.isurf[0] == N);
    c->surf_top_only = c->g.isurf[0] == N && c->g.isurf[1] == 0 && c->g.isurf[2] == 0 && c->g.isurf[3] == 0;
    for (int i = 0; c->fusable && i < c->nsrc_tot; ++i) c->fusable = interior(q->src_z[i], q->src_x[i]);
    for (int i = 0; c->fusable && i < c->nrec_tot; ++i) c->fusable = interior(q->rec_z[i], q->rec_x[i]);
    RETIF(plan(c));
    RETIF(allocate(c));
    // snapshot cells outside the interior are never written by the one-pass step: zero once
    for (int i = 0; i < 5; ++i)
        HIPCHK(hipMemsetAsync(c->snap[i], 0, sizeof(float) * c->B * c->snap_stride, c->st));
    {   // per-step tables for graph replay (what time_step / adj_step derive from it)
        const int nt = c->g.nt, sdt = c->g.snap_dt;
        const long long slab = (long long)c->snz * c->snx;
        // full storage: slab it / snap_dt; checkpoints: from the unit's start in a recomputed
        // unit, from t_keep in the kept ones (which the forward pass stores)
        const bool full = c->storage != FWI_STORE_CHECKPOINT;
        std::vector<StepInfo> f(nt), r(nt), a(nt);
        for (int it = 0; it < nt; ++it) {
            const bool kept = full || it >= c->t_keep;
            const int tu = full ? 0 : (it / c->S) * c->S, t0 = full ? 0 : kept ? c->t_keep : tu;
            const long long off = (long long)((it - t0) / sdt) * slab, roff = (long long)((it - tu) / sdt) * slab;
            const int snap = it % sdt == 0;
            f[it] = {it, 0, kept && snap, 0, kept ? off : 0};
            r[it] = {it, 0, snap, 0, roff};
            a[it] = {it, snap, 0, 0, off};
        }
        RETIF(upload(c, c->si_fwd, f.data(), nt));
        RETIF(upload(c, c->si_rec, r.data(), nt));
        RETIF(upload(c, c->si_adj, a.data(), nt));
    }
    RETIF(upload(c, c->zmap, zmap.data(), nz));
    RETIF(upload(c, c->xmap, xmap.data(), nx));
    // profiles as b, a, 1/K (integer then half), z then x
    std::vector<float> hz(6 * nz), hx(6 * nx);
    for (int i = 0; i < nz; ++i) {
        hz[0 * nz + i] = (float)c->prof[1][i];
        hz[1 * nz + i] = (float)c->prof[0][i];
        hz[2 * nz + i] = (float)(1.0 / c->prof[2][i]);
        hz[3 * nz + i] = (float)c->prof[4][i];
        hz[4 * nz + i] = (float)c->prof[3][i];
        hz[5 * nz + i] = (float)(1.0 / c->prof[5][i]);
    }
    for (int i = 0; i < nx; ++i) {
        hx[0 * nx + i] = (float)c->prof[7][i];
        hx[1 * nx + i] = (float)c->prof[6][i];
        hx[2 * nx + i] = (float)(1.0 / c->prof[8][i]);
        hx[3 * nx + i] = (float)c->prof[10][i];
        hx[4 * nx + i] = (float)c->prof[9][i];
        hx[5 * nx + i] = (float)(1.0 / c->prof[11][i]);
    }
    for (int k = 0; k < 6; ++k) RETIF(upload(c, c->pf[k], hz.data() + k * nz, nz));
    for (int k = 0; k < 6; ++k) RETIF(upload(c, c->pf[6 + k], hx.data() + k * nx, nx));
    std::vector<CpmlPoint> pz(nz), px(nx);
    for (int i = 0; i < nz; ++i)
        pz[i] = {hz[i], hz[nz + i], hz[2 * nz + i], hz[3 * nz + i], hz[4 * nz + i], hz[5 * nz + i], zmap[i], 0};
    for (int i = 0; i < nx; ++i)
        px[i] = {hx[i], hx[nx + i], hx[2 * nx + i], hx[3 * nx + i], hx[4 * nx + i], hx[5 * nx + i], xmap[i], 0};
    RETIF(upload(c, c->zp, pz.data(), nz));
    RETIF(upload(c, c->xp, px.data(), nx));
    if (c->dprof) {   // the FP64 propagation's profiles: b, a, 1/K, then at the half point, z then x
        std::vector<double> dp(6 * ((size_t)nz + nx));
        for (int i = 0; i < nz + nx; ++i) {
            const int base = i < nz ? 0 : 6, j = i < nz ? i : i - nz;
            const std::vector<double> *pr = c->prof + base;
            const double v[6] = {pr[1][j], pr[0][j], 1.0 / pr[2][j], pr[4][j], pr[3][j], 1.0 / pr[5][j]};
            for (int k = 0; k < 6; ++k) dp[(size_t)i * 6 + k] = v[k];
        }
        RETIF(upload(c, c->dprof, dp.data(), dp.size()));
    }
    if (c->ztab_ok)
        for (int i = 0; i < nz; ++i)
            if (zmap[i] >= 0)
                for (int k = 0; k < 6; ++k) c->ztab.p[zmap[i]][k] = hz[k * nz + i];
    // sources
    std::vector<int> gs, gc, go, gi;
    build_groups(c->src_off, q->src_z, q->src_x, c->P, gs, gc, go, gi, c->sg_first);
    bool ok = true;
    auto I = [&](size_t n) { int *p = c->dalloc<int>(n); ok = ok && p; return p; };
    c->sg_shot = I(gs.size()); c->sg_cell = I(gc.size()); c->sg_off = I(go.size()); c->sg_idx = I(gi.size());
    c->stf_z = c->dalloc<double>((size_t)c->nsrc_tot * c->g.nt);
    c->stf_x = c->dalloc<double>((size_t)c->nsrc_tot * c->g.nt);
    if (!ok || !c->stf_z || !c->stf_x) return c->fail(FWI_ENOMEM, "source tables");
    RETIF(upload(c, c->sg_shot, gs.data(), gs.size()));
    RETIF(upload(c, c->sg_cell, gc.data(), gc.size()));
    RETIF(upload(c, c->sg_off, go.data(), go.size()));
    RETIF(upload(c, c->sg_idx, gi.data(), gi.size()));
    RETIF(upload(c, c->stf_z, q->stf_z, (size_t)c->nsrc_tot * c->g.nt));
    RETIF(upload(c, c->stf_x, q->stf_x, (size_t)c->nsrc_tot * c->g.nt));
    // receivers: recording list and adjoint-injection groups
    std::vector<int> rshot(c->nrec_tot), rcell(c->nrec_tot);
    for (int s = 0; s < c->nshot; ++s)
        for (int r = c->rec_off[s]; r < c->rec_off[s + 1]; ++r) {
            rshot[r] = s;
            rcell[r] = q->rec_z[r] * c->P + q->rec_x[r];
        }
    c->rec_shot = I(c->nrec_tot);
    c->rec_cell = I(c->nrec_tot);
    c->rec_off_d = I(c->nshot + 1);
    gs.clear(); gc.clear(); go.clear(); gi.clear();
    build_groups(c->rec_off, q->rec_z, q->rec_x, c->P, gs, gc, go, gi, c->rg_first);
    c->rg_shot = I(gs.size()); c->rg_cell = I(gc.size()); c->rg_off = I(go.size()); c->rg_idx = I(gi.size());
    if (!ok) return c->fail(FWI_ENOMEM, "receiver tables");
    RETIF(upload(c, c->rec_shot, rshot.data(), rshot.size()));
    RETIF(upload(c, c->rec_cell, rcell.data(), rcell.size()));
    RETIF(upload(c, c->rec_off_d, c->rec_off.data(), c->rec_off.size()));
    RETIF(upload(c, c->rg_shot, gs.data(), gs.size()));
    RETIF(upload(c, c->rg_cell, gc.data(), gc.size()));
    RETIF(upload(c, c->rg_off, go.data(), go.size()));
    RETIF(upload(c, c->rg_idx, gi.data(), gi.size()));
    for (int i = 0; i < 3; ++i) HIPCHK(hipMemsetAsync(c->G[i], 0, sizeof(double) * (size_t)nz * nx, c->st));
    return sync(c);
}

int validate(fwi_ctx *c, const fwi_grid *g, const fwi_acq *q) {
    if (g->nt <= 0 || g->nz <= 0 || g->nx <= 0) return c->fail(FWI_EINVAL, "empty grid or time axis");
    if (g->fdorder < 2 || g->fdorder > 12 || g->fdorder % 2) return c->fail(FWI_EINVAL, "fdorder %d", g->fdorder);
    const int N = g->fdorder / 2;
    if (g->nz < 2 * N + 1 || g->nx < 2 * N + 1) return c->fail(FWI_EINVAL, "grid smaller than the stencil");
    if (!(g->dt > 0) || !(g->dz > 0) || !(g->dx > 0)) return c->fail(FWI_EINVAL, "dt, dz, dx must be positive");
    const int tt = c->o.tile_tz;
    if (tt < 0 || (c->o.kernel_variant < 6 && tt != 0 && tt != 8 && tt != 16 && tt != 32 && tt != 64))
        return c->fail(FWI_EINVAL,
                       "tile_tz %d (0, 8, 16, 32 or 64; variant 6: 0 or its row block times 1 or 2; "
                       "variants 7, 8: 0 or their row block times 4 or 8)", tt);
    if (c->o.kernel_variant < 0 || c->o.kernel_variant > 13)
        return c->fail(FWI_EINVAL, "kernel_variant %d (0..13)", c->o.kernel_variant);
    if (c->o.kernel_variant == 9 && tt > 2) return c->fail(FWI_EINVAL, "variant 9 takes tile_tz 0, 1 or 2");
    if (c->o.work_order < 0 || c->o.work_order > 4)
        return c->fail(FWI_EINVAL, "work_order %d (0..4)", c->o.work_order);
    if (c->o.ckpt_slots < 0) return c->fail(FWI_EINVAL, "ckpt_slots %d < 0", c->o.ckpt_slots);
    if (c->o.ckpt_keep < 0) return c->fail(FWI_EINVAL, "ckpt_keep %d < 0", c->o.ckpt_keep);
    if (c->o.seis_fp64 != 0 && c->o.seis_fp64 != 1) return c->fail(FWI_EINVAL, "seis_fp64 %d (0 or 1)", c->o.seis_fp64);
    if (c->o.smooth_hfs != 0 && (c->o.smooth_hfs < 2 || c->o.smooth_hfs > 64))
        return c->fail(FWI_EINVAL, "smooth_hfs %d (0, or 2..64)", c->o.smooth_hfs);
    if (c->o.param != FWI_PARAM_LAME && c->o.param != FWI_PARAM_VELOCITY)
        return c->fail(FWI_EINVAL, "param %d (0 lam/mu/rho, 1 cp/cs)", c->o.param);
    if (c->o.param == FWI_PARAM_VELOCITY) {
        const double *b = c->o.vel_bounds;
        if (b[0] != 0.0 || b[1] != 0.0 || b[2] != 0.0 || b[3] != 0.0) {
            if (!(b[0] >= 0.0 && b[0] < b[1] && b[2] >= 0.0 && b[2] < b[3]))
                return c->fail(FWI_EINVAL, "vel_bounds need 0 <= cp_min < cp_max and 0 <= cs_min < cs_max");
            std::memcpy(c->vb, b, sizeof c->vb);
        }
    }
    if (g->snap_dt < 1 || g->snap_dz < 1 || g->snap_dx < 1) return c->fail(FWI_EINVAL, "snap strides < 1");
    if (g->snap_z1 < 0 || g->snap_z2 >= g->nz || g->snap_z1 > g->snap_z2 || g->snap_x1 < 0 ||
        g->snap_x2 >= g->nx || g->snap_x1 > g->snap_x2)
        return c->fail(FWI_EINVAL, "snap box outside the grid");
    if ((g->snap_dx > 1 && (g->snap_x2 - g->snap_x1) % g->snap_dx) ||
        (g->snap_dz > 1 && (g->snap_z2 - g->snap_z1) % g->snap_dz))
        return c->fail(FWI_EINVAL, "snap box must span whole snap strides");
    for (int s = 0; s < 4; ++s) {
        const int v = g->isurf[s];
        if (v < 0) return c->fail(FWI_EINVAL, "isurf[%d] < 0", s);
        if (v > 0 && (s < 2 ? (v < N || v >= g->nz - N) : (v < N || v >= g->nx - N)))
            return c->fail(FWI_EINVAL, "free surface %d at %d is outside the interior", s, v);
        // the image rows / columns (surf_mirror) must stay inside the grid
        const int n = s < 2 ? g->nz : g->nx;
        if (v > 0 && ((s % 2 == 0 && 2 * v - N > n - 1) || (s % 2 == 1 && 2 * v - n + N < 0)))
            return c->fail(FWI_EINVAL, "free surface %d at %d mirrors beyond the grid", s, v);
    }
    if (g->pml_z && (!g->a_z || !g->b_z || !g->K_z || !g->a_half_z || !g->b_half_z || !g->K_half_z))
        return c->fail(FWI_EINVAL, "missing z CPML profiles");
    if (g->pml_x && (!g->a_x || !g->b_x || !g->K_x || !g->a_half_x || !g->b_half_x || !g->K_half_x))
        return c->fail(FWI_EINVAL, "missing x CPML profiles");
    if (!q || q->nshot <= 0 || !q->src_off || !q->rec_off) return c->fail(FWI_EINVAL, "no shots");
    if (q->stf_type != 0 && q->stf_type != 1) return c->fail(FWI_EINVAL, "stf_type %d", q->stf_type);
    for (int s = 0; s < q->nshot; ++s)
        if (q->src_off[s + 1] < q->src_off[s] || q->rec_off[s + 1] < q->rec_off[s])
            return c->fail(FWI_EINVAL, "CSR offsets decrease at shot %d", s);
    if (q->src_off[0] != 0 || q->rec_off[0] != 0) return c->fail(FWI_EINVAL, "CSR offsets must start at 0");
    const int ns = q->src_off[q->nshot], nr = q->rec_off[q->nshot];
    if ((ns && (!q->src_z || !q->src_x || !q->stf_z || !q->stf_x)) || (nr && (!q->rec_z || !q->rec_x)))
        return c->fail(FWI_EINVAL, "missing source or receiver arrays");
    for (int i = 0; i < ns; ++i)
        if (q->src_z[i] < 0 || q->src_z[i] >= g->nz || q->src_x[i] < 0 || q->src_x[i] >= g->nx)
            return c->fail(FWI_EINVAL, "source %d at (%d, %d) outside the grid", i, q->src_z[i], q->src_x[i]);
    for (int i = 0; i < nr; ++i)
        if (q->rec_z[i] < 0 || q->rec_z[i] >= g->nz || q->rec_x[i] < 0 || q->rec_x[i] >= g->nx)
            return c->fail(FWI_EINVAL, "receiver %d at (%d, %d) outside the grid", i, q->rec_z[i], q->rec_x[i]);
    return FWI_OK;
}

thread_local std::string g_create_error;

}  // namespace

void psv::set_global_error(const char *msg) { g_create_error = msg ? msg : ""; }

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int32_t fwi_abi_version(void) { return FWI_PSV_ABI_VERSION; }

const char *fwi_strerror(int32_t s) {
    switch (s) {
        case FWI_OK: return "ok";
        case FWI_EINVAL: return "invalid argument";
        case FWI_ENOMEM: return "out of memory";
        case FWI_EHIP: return "HIP runtime error";
        case FWI_ENONFINITE: return "non-finite values";
        case FWI_ECFL: return "time step above the stability limit";
        case FWI_ESTEP: return "step-length search failed";
        case FWI_ESTATE: return "call out of order";
        case FWI_ENODEV: return "no gfx950 device";
        default: return "unknown status";
    }
}

const char *fwi_last_error(const fwi_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

void fwi_default_opts(fwi_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->storage = FWI_STORE_AUTO;
    o->form = FWI_FORM_FUSED;
    o->grad_sign = -1;
    o->check_finite = 1;
}

int32_t fwi_create(const fwi_grid *g, const fwi_acq *q, const fwi_opts *opts, fwi_ctx **out) {
    if (!out || !g) return FWI_EINVAL;
    *out = nullptr;
    fwi_ctx *c = new (std::nothrow) fwi_ctx();
    if (!c) return FWI_ENOMEM;
    fwi_default_opts(&c->o);
    if (opts) c->o = *opts;
    if (c->o.grad_sign == 0) c->o.grad_sign = -1;
    int r = validate(c, g, q);
    if (r == FWI_OK) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= c->o.device)
            r = c->fail(FWI_ENODEV, "no HIP device %d", c->o.device);
        else if (hipSetDevice(c->o.device) != hipSuccess)
            r = c->fail(FWI_ENODEV, "cannot select device %d", c->o.device);
    }
    if (r == FWI_OK) {
        c->g = *g;
        c->N = g->fdorder / 2;
        // at least one pad column: variant 5's halo lanes store into column P - 1
        c->P = (int)round_up(g->nx + 1, kPitchAlign);
        c->snt = 1 + (g->nt - 1) / g->snap_dt;
        c->snz = 1 + (g->snap_z2 - g->snap_z1) / g->snap_dz;
        c->snx = 1 + (g->snap_x2 - g->snap_x1) / g->snap_dx;
        c->nshot = q->nshot;
        c->stf_type = q->stf_type;
        c->src_off.assign(q->src_off, q->src_off + q->nshot + 1);
        c->rec_off.assign(q->rec_off, q->rec_off + q->nshot + 1);
        c->nsrc_tot = c->src_off.back();
        c->nrec_tot = c->rec_off.back();
        c->tz = c->o.tile_tz > 0 ? c->o.tile_tz : 16;
        c->plain = c->o.form == FWI_FORM_PLAIN;
        const double *pz[6] = {g->a_z, g->b_z, g->K_z, g->a_half_z, g->b_half_z, g->K_half_z};
        const double *px[6] = {g->a_x, g->b_x, g->K_x, g->a_half_x, g->b_half_x, g->K_half_x};
        for (int k = 0; k < 6; ++k) {
            c->prof[k].assign(g->nz, k % 3 == 1 || k % 3 == 2 ? 1.0 : 0.0);
            c->prof[6 + k].assign(g->nx, k % 3 == 1 || k % 3 == 2 ? 1.0 : 0.0);
            if (g->pml_z) c->prof[k].assign(pz[k], pz[k] + g->nz);
            if (g->pml_x) c->prof[6 + k].assign(px[k], px[k] + g->nx);
        }
        // the context keeps its own copies; the caller's profile pointers are not retained
        c->g.a_z = c->g.b_z = c->g.K_z = c->g.a_half_z = c->g.b_half_z = c->g.K_half_z = nullptr;
        c->g.a_x = c->g.b_x = c->g.K_x = c->g.a_half_x = c->g.b_half_x = c->g.K_half_x = nullptr;
        if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess)
            r = c->fail(FWI_EHIP, "stream creation failed");
    }
    if (r == FWI_OK) r = setup_device(c, q);
    if (r != FWI_OK) {
        g_create_error = c->err;
        fwi_destroy(c);
        return r;
    }
    *out = c;
    return FWI_OK;
}

void fwi_destroy(fwi_ctx *c) {
    if (!c) return;
    if (c->st) hipStreamSynchronize(c->st);
    for (auto &p : c->ev_pending) {
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    for (auto e : c->ev_pool) hipEventDestroy(e);
    free_all(c);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
}

int32_t fwi_plan(const fwi_ctx *c, int32_t *batch, int32_t *storage, int32_t *ckpt_interval, int64_t *bytes) {
    if (!c) return FWI_EINVAL;
    if (batch) *batch = c->B;
    if (storage) *storage = c->storage;
    if (ckpt_interval) *ckpt_interval = c->storage == FWI_STORE_CHECKPOINT ? c->S : 0;
    if (bytes) *bytes = (int64_t)c->bytes;
    return FWI_OK;
}

int32_t fwi_checkpoint_plan(const fwi_ctx *c, int32_t *units, int32_t *slots, double *recompute) {
    if (!c) return FWI_EINVAL;
    const bool ck = c->storage == FWI_STORE_CHECKPOINT;
    if (units) *units = ck ? c->nseg : 0;
    if (slots) *slots = ck ? c->slots : 0;
    // the leaves recompute each unit before t_keep once; advances are whole units
    if (recompute) *recompute = ck ? ((double)c->nrev + (double)c->extra_units) * c->S / c->g.nt : 0.0;
    return FWI_OK;
}

int32_t fwi_checkpoint_kept(const fwi_ctx *c, int32_t *kept) {
    if (!c) return FWI_EINVAL;
    if (kept) *kept = c->storage == FWI_STORE_CHECKPOINT ? c->keep : 0;
    return FWI_OK;
}

int32_t fwi_set_model(fwi_ctx *c, const double *lam, const double *mu, const double *rho) {
    if (!c || !lam || !mu || !rho) return FWI_EINVAL;
    const size_t n = (size_t)c->g.nz * c->g.nx;
    RETIF(upload(c, c->m_lam, lam, n));
    RETIF(upload(c, c->m_mu, mu, n));
    RETIF(upload(c, c->m_rho, rho, n));
    RETIF(refresh_media(c));
    c->have_model = 1;
    return FWI_OK;
}

int32_t fwi_set_model_device(fwi_ctx *c, const double *lam, const double *mu, const double *rho) {
    if (!c || !lam || !mu || !rho) return FWI_EINVAL;
    const size_t n = (size_t)c->g.nz * c->g.nx;
    dev_copy(c->m_lam, lam, n, c->st);
    dev_copy(c->m_mu, mu, n, c->st);
    dev_copy(c->m_rho, rho, n, c->st);
    RETIF(refresh_media(c));
    c->have_model = 1;
    return FWI_OK;
}

int32_t fwi_get_model(fwi_ctx *c, double *lam, double *mu, double *rho) {
    if (!c || !c->have_model) return c ? c->fail(FWI_ESTATE, "no model") : FWI_EINVAL;
    const size_t n = (size_t)c->g.nz * c->g.nx;
    if (lam) HIPCHK(hipMemcpyAsync(lam, c->m_lam, n * 8, hipMemcpyDeviceToHost, c->st));
    if (mu) HIPCHK(hipMemcpyAsync(mu, c->m_mu, n * 8, hipMemcpyDeviceToHost, c->st));
    if (rho) HIPCHK(hipMemcpyAsync(rho, c->m_rho, n * 8, hipMemcpyDeviceToHost, c->st));
    return sync(c);
}

int32_t fwi_get_media(fwi_ctx *c, double *mu_zx, double *rho_zp, double *rho_xp, double C[3]) {
    if (!c || !c->have_model) return c ? c->fail(FWI_ESTATE, "no model") : FWI_EINVAL;
    const int nz = c->g.nz, nx = c->g.nx, P = c->P;
    std::vector<float> h((size_t)nz * P);
    const float *src[3] = {c->muzx, c->rhoz, c->rhox};
    double *dst[3] = {mu_zx, rho_zp, rho_xp};
    for (int k = 0; k < 3; ++k) {
        if (!dst[k]) continue;
        HIPCHK(hipMemcpyAsync(h.data(), src[k], h.size() * 4, hipMemcpyDeviceToHost, c->st));
        RETIF(sync(c));
        for (int iz = 0; iz < nz; ++iz)
            for (int ix = 0; ix < nx; ++ix) dst[k][(size_t)iz * nx + ix] = h[(size_t)iz * P + ix];
    }
    if (C) std::memcpy(C, c->C, sizeof(double) * 3);
    return FWI_OK;
}

int32_t fwi_set_observed(fwi_ctx *c, const double *oz, const double *ox) {
    if (!c || !oz || !ox) return FWI_EINVAL;
    const size_t n = (size_t)c->nrec_tot * c->g.nt;
    RETIF(upload(c, c->obs_z, oz, n));
    RETIF(upload(c, c->obs_x, ox, n));
    RETIF(sync(c));
    c->have_obs = 1;
    return FWI_OK;
}

int32_t fwi_forward(fwi_ctx *c, int32_t shot0, int32_t nshots, double *seis_z, double *seis_x) {
    if (!c) return FWI_EINVAL;
    if (!c->have_model) return c->fail(FWI_ESTATE, "fwi_forward before fwi_set_model");
    std::vector<std::pair<int, int>> bs;
    if (batches(c, shot0, nshots, bs) != FWI_OK) return c->fail(FWI_EINVAL, "shot range");
    const int nt = c->g.nt;
    for (auto [s0, nb] : bs) {
        RETIF(run_forward(c, s0, nb, false, !c->o.seis_fp64, true));
        if (c->o.seis_fp64) RETIF(run_forward64(c, s0, nb));
        const int r0 = c->rec_off[s0], r1 = c->rec_off[s0 + nb], rb = c->rec_off[shot0];
        const size_t n = (size_t)(r1 - r0) * nt;
        if (seis_z)
            HIPCHK(hipMemcpyAsync(seis_z + (size_t)(r0 - rb) * nt, c->syn_z + (size_t)r0 * nt, n * 8,
                                  hipMemcpyDeviceToHost, c->st));
        if (seis_x)
            HIPCHK(hipMemcpyAsync(seis_x + (size_t)(r0 - rb) * nt, c->syn_x + (size_t)r0 * nt, n * 8,
                                  hipMemcpyDeviceToHost, c->st));
        RETIF(sync(c));
    }
    return FWI_OK;
}

int32_t fwi_forward_device(fwi_ctx *c, int32_t shot0, int32_t nshots, double *seis_z, double *seis_x) {
    if (!c) return FWI_EINVAL;
    if (!c->have_model) return c->fail(FWI_ESTATE, "fwi_forward_device before a model");
    std::vector<std::pair<int, int>> bs;
    if (batches(c, shot0, nshots, bs) != FWI_OK) return c->fail(FWI_EINVAL, "shot range");
    const int nt = c->g.nt;
    for (auto [s0, nb] : bs) {
        RETIF(run_forward(c, s0, nb, false, !c->o.seis_fp64, true));
        if (c->o.seis_fp64) RETIF(run_forward64(c, s0, nb));
        const int r0 = c->rec_off[s0], r1 = c->rec_off[s0 + nb], rb = c->rec_off[shot0];
        const size_t n = (size_t)(r1 - r0) * nt;
        if (seis_z) dev_copy(seis_z + (size_t)(r0 - rb) * nt, c->syn_z + (size_t)r0 * nt, n, c->st);
        if (seis_x) dev_copy(seis_x + (size_t)(r0 - rb) * nt, c->syn_x + (size_t)r0 * nt, n, c->st);
        RETIF(sync(c));
    }
    return FWI_OK;
}

int32_t fwi_get_wavefield(fwi_ctx *c, int32_t shot, int32_t field, double *out) {
    if (!c || !out || field < 0 || field > 8) return FWI_EINVAL;
    if (shot < c->last_s0 || shot >= c->last_s0 + c->last_nb)
        return c->fail(FWI_ESTATE, "shot %d is not in the last propagated batch", shot);
    const int nz = c->g.nz, nx = c->g.nx, P = c->P;
    std::vector<float> h((size_t)nz * P);
    // ABI order vz vx uz ux szz szx sxx We We_adj -> device arrays vz vx szz szx sxx uz ux We Wa
    static const int dev_index[9] = {0, 1, 5, 6, 2, 3, 4, 7, 8};
    const float *src = c->f[dev_index[field]] + (size_t)(shot - c->last_s0) * c->fs();
    HIPCHK(hipMemcpyAsync(h.data(), src, h.size() * 4, hipMemcpyDeviceToHost, c->st));
    RETIF(sync(c));
    for (int iz = 0; iz < nz; ++iz)
        for (int ix = 0; ix < nx; ++ix) out[(size_t)iz * nx + ix] = h[(size_t)iz * P + ix];
    return FWI_OK;
}

int32_t fwi_forward_snapshots(fwi_ctx *c, int32_t shot, double *seis_z, double *seis_x, double *vz, double *vx,
                              double *szz, double *szx, double *sxx) {
    if (!c) return FWI_EINVAL;
    if (!c->have_model) return c->fail(FWI_ESTATE, "fwi_forward_snapshots before fwi_set_model");
    if (shot < 0 || shot >= c->nshot) return c->fail(FWI_EINVAL, "shot %d", shot);
    if (c->storage != FWI_STORE_FULL) return c->fail(FWI_ESTATE, "snapshots need FULL storage");
    RETIF(run_forward(c, shot, 1, true, true, true));
    const int nt = c->g.nt, r0 = c->rec_off[shot], r1 = c->rec_off[shot + 1];
    const size_t n = (size_t)(r1 - r0) * nt;
    if (seis_z) HIPCHK(hipMemcpyAsync(seis_z, c->syn_z + (size_t)r0 * nt, n * 8, hipMemcpyDeviceToHost, c->st));
    if (seis_x) HIPCHK(hipMemcpyAsync(seis_x, c->syn_x + (size_t)r0 * nt, n * 8, hipMemcpyDeviceToHost, c->st));
    double *outs[5] = {vz, vx, szz, szx, sxx};
    std::vector<float> h(c->snap_stride);
    for (int k = 0; k < 5; ++k) {
        if (!outs[k]) continue;
        HIPCHK(hipMemcpyAsync(h.data(), c->snap[k], h.size() * 4, hipMemcpyDeviceToHost, c->st));
        RETIF(sync(c));
        for (size_t i = 0; i < h.size(); ++i) outs[k][i] = h[i];
    }
    return sync(c);
}

int32_t fwi_gradient(fwi_ctx *c, int32_t shot0, int32_t nshots, double *misfit) {
    if (!c) return FWI_EINVAL;
    if (!c->have_model || !c->have_obs) return c->fail(FWI_ESTATE, "fwi_gradient needs a model and observed data");
    std::vector<std::pair<int, int>> bs;
    if (batches(c, shot0, nshots, bs) != FWI_OK) return c->fail(FWI_EINVAL, "shot range");
    const int nz = c->g.nz, nx = c->g.nx;
    for (auto [s0, nb] : bs) {
        RETIF(batch_gradient(c, s0, nb, nullptr, nullptr, misfit ? misfit + (s0 - shot0) : nullptr));
        const size_t snapcell = (size_t)c->snz * c->snx;
        for (int s = 0; s < nb; ++s) {
            const double alpha = c->alpha[s0 + s];
            const size_t cells = (size_t)c->B * snapcell;
            const double *gs[3] = {c->gr64 + s * snapcell, c->gr64 + cells + s * snapcell,
                                   c->gr64 + 2 * cells + s * snapcell};
            const double gsc[3] = {1.0 / (alpha * kGainLM), 1.0 / (alpha * kGainLM), 1.0 / alpha};
            dev_precondition(c->f[7] + s * c->fs(), c->f[8] + s * c->fs(), 1.0 / (alpha * alpha), gs, gsc,
                             c->G, c->C, c->tmp, c->red, nz, nx, c->P, c->g.snap_z1, c->g.snap_z2, c->g.snap_x1,
                             c->g.snap_x2, c->g.snap_dz, c->g.snap_dx, c->snz, c->snx, c->st);
        }
        RETIF(sync(c));
    }
    return FWI_OK;
}

int32_t fwi_gradient_raw(fwi_ctx *c, int32_t shot0, int32_t nshots, const double *adj_z, const double *adj_x,
                         double *g_lam, double *g_mu, double *g_rho) {
    if (!c || !adj_z || !adj_x) return FWI_EINVAL;
    if (!c->have_model) return c->fail(FWI_ESTATE, "fwi_gradient_raw before fwi_set_model");
    std::vector<std::pair<int, int>> bs;
    if (batches(c, shot0, nshots, bs) != FWI_OK) return c->fail(FWI_EINVAL, "shot range");
    const size_t snapcell = (size_t)c->snz * c->snx;
    double *outs[3] = {g_lam, g_mu, g_rho};
    for (int k = 0; k < 3; ++k)
        if (outs[k]) std::memset(outs[k], 0, snapcell * 8);
    std::vector<double> h(snapcell * c->B);
    const int nt = c->g.nt, rb = c->rec_off[shot0];
    for (auto [s0, nb] : bs) {
        const size_t off = (size_t)(c->rec_off[s0] - rb) * nt;
        RETIF(batch_gradient(c, s0, nb, adj_z + off, adj_x + off, nullptr));
        for (int k = 0; k < 3; ++k) {
            if (!outs[k]) continue;
            HIPCHK(hipMemcpyAsync(h.data(), c->gr64 + (size_t)k * c->B * snapcell, snapcell * nb * 8,
                                  hipMemcpyDeviceToHost, c->st));
            RETIF(sync(c));
            for (int s = 0; s < nb; ++s) {
                const double den = c->alpha[s0 + s] * (k < 2 ? kGainLM : 1.0);
                for (size_t i = 0; i < snapcell; ++i) outs[k][i] += (double)h[s * snapcell + i] / den;
            }
        }
    }
    return FWI_OK;
}

// dJ/d(lam, mu, rho) into the context's scratch [3][nz][nx] (c->tmp holds 2 of them, the
// third goes to the device array d3): adjoint sources adj_z / adj_x are device arrays
// [rows][nt] (the batch's rows are copied into res_* by a kernel).
int misfit_gradient_dev(fwi_ctx *c, int shot0, int nshots, const double *adj_z, const double *adj_x,
                        double *d_lam, double *d_mu, double *d_rho) {
    if (c->plain) return c->fail(FWI_EINVAL, "the plain form computes the reference's sums only");
    if (c->g.snap_dz != 1 || c->g.snap_dx != 1)
        return c->fail(FWI_EINVAL, "the misfit derivative needs snap_dz = snap_dx = 1 (staggered chain rule)");
    std::vector<std::pair<int, int>> bs;
    if (batches(c, shot0, nshots, bs) != FWI_OK) return c->fail(FWI_EINVAL, "shot range");
    const size_t snapcell = (size_t)c->snz * c->snx, cells = (size_t)c->B * snapcell;
    if (!c->raw) {   // FP64 totals of the five sums over the shots, allocated on first use
        c->raw = c->dalloc<double>(5 * snapcell);
        if (!c->raw) return c->fail(FWI_ENOMEM, "misfit-derivative sums");
    }
    HIPCHK(hipMemsetAsync(c->raw, 0, sizeof(double) * 5 * snapcell, c->st));
    const int nt = c->g.nt, rb = c->rec_off[shot0];
    c->gform = 1;
    for (auto [s0, nb] : bs) {
        const size_t off = (size_t)(c->rec_off[s0] - rb) * nt;
        const int r = batch_gradient(c, s0, nb, adj_z + off, adj_x + off, nullptr);
        if (r != FWI_OK) {
            c->gform = 0;
            return r;
        }
        // per-shot totals of this batch (gr64 [k][B][cells]) into the sums over shots
        for (int s = 0; s < nb; ++s) {
            const double *g[5];
            for (int k = 0; k < 5; ++k) g[k] = c->gr64 + k * cells + s * snapcell;
            dev_acc_raw64(c->raw, g, snapcell, c->alpha_d + s0 + s, kGainLM, c->st);
        }
    }
    c->gform = 0;
    // The sums were taken with weight grad_sign * snap_dt * dt; with the CPU sign (-1)
    // they are +dJ/dm of J = sum 0.5 dt r^2 for adjoint sources r.
    const double sg = -1.0 / c->o.grad_sign;
    dev_misfit_chain(c->raw, c->m_mu, d_lam, d_mu, d_rho, c->g.nz, c->g.nx, c->g.snap_z1, c->g.snap_x1, c->snz,
                     c->snx, sg, c->st);
    return sync(c);
}

// The exact derivative of J = sum 0.5 dt r^2 (adjoint sources r = adj_z / adj_x, host or
// device arrays [rows][nt]) with respect to lam, mu, rho into device arrays [nz][nx] (zero
// outside the snapshot box): the forward run storing its intermediates for every step, the
// reverse sweep of psv_exact.hip, the media gradient carried through mat_av2.  Its buffers
// are allocated for the call and freed before it returns; shots go in batches that fit.
int exact_gradient(fwi_ctx *c, int shot0, int nshots, const double *adj_z, const double *adj_x, bool on_device,
                   double *d_lam, double *d_mu, double *d_rho) {
    if (shot0 < 0 || nshots <= 0 || shot0 + nshots > c->nshot) return c->fail(FWI_EINVAL, "shot range");
    const int nz = c->g.nz, nx = c->g.nx, nt = c->g.nt;
    const size_t fs = c->fs();
    // Per shot: 30 float fields (state, memories, adjoint state, carries, scratch) and the
    // intermediates of S steps (5 S fields); with S < nt also a checkpoint of the forward
    // state (5 fields, 8 memories) at the start of each of the ceil(nt / S) segments, from
    // which each segment's forward is recomputed before its reverse sweep.  S = nt (no
    // recomputation) when that fits, else S ~ sqrt(13 nt / 5), the smallest footprint;
    // storage FWI_STORE_CHECKPOINT with ckpt_interval > 0 asks for that S, FWI_STORE_FULL
    // for S = nt.
    auto per_shot_of = [&](int S_) {
        const size_t nseg = (size_t)(nt + S_ - 1) / S_;
        return fs * 4 * (30 + 5 * (size_t)S_ + (S_ < nt ? 13 * nseg : 0));
    };
    size_t freeb = 0, total = 0;
    HIPCHK(hipMemGetInfo(&freeb, &total));
    const size_t avail = (size_t)(0.85 * (double)freeb);
    const size_t g_all = fs * 8 * 5 * (size_t)nshots;   // the media gradients of every shot
    int S = nt;
    if (c->o.storage == FWI_STORE_CHECKPOINT && c->o.ckpt_interval > 0)
        S = std::min(c->o.ckpt_interval, nt);
    else if (c->o.storage != FWI_STORE_FULL && per_shot_of(nt) + g_all > avail)
        S = std::max(1, std::min(nt, (int)std::lround(std::sqrt(13.0 * nt / 5.0))));
    const size_t per_shot = per_shot_of(S);
    if (per_shot + g_all > avail)
        return c->fail(FWI_ENOMEM, "the exact adjoint keeps %.2f GB per shot (%d steps of intermediates), %.2f GB free",
                       per_shot / 1e9, S, avail / 1e9);
    const int Bx = (int)std::min<size_t>((size_t)nshots, (avail - g_all) / per_shot);
    const int nseg = (nt + S - 1) / S;
    std::vector<void *> mem;
    auto A = [&](size_t bytes) -> void * {
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        mem.push_back(p);
        return p;
    };
    auto release = [&]() {
        (void)hipStreamSynchronize(c->st);
        for (void *p : mem) (void)hipFree(p);
        mem.clear();
    };
    XArgs x{};
    x.nz = nz; x.nx = nx; x.P = c->P; x.B = Bx; x.N = c->N;
    x.pml_z = c->g.pml_z; x.pml_x = c->g.pml_x;
    x.fs = fs;
    x.dt = (float)c->g.dt;
    for (int k = 0; k < 6; ++k) {
        x.cz[k] = k < c->N ? (float)(c->g.hc[k + 1] / c->g.dz) : 0.f;
        x.cx[k] = k < c->N ? (float)(c->g.hc[k + 1] / c->g.dx) : 0.f;
    }
    x.lam = c->lam; x.mu = c->mu; x.muzx = c->muzx; x.rhoz = c->rhoz; x.rhox = c->rhox;
    x.zp = c->zp; x.xp = c->xp;
    float *fl = static_cast<float *>(A(sizeof(float) * fs * Bx * 30));
    float *hist = static_cast<float *>(A(sizeof(float) * fs * Bx * 5 * (size_t)S));
    double *G = static_cast<double *>(A(sizeof(double) * fs * 5 * nshots));
    float *ckp = S < nt ? static_cast<float *>(A(sizeof(float) * fs * Bx * 13 * (size_t)nseg)) : nullptr;
    if (!fl || !hist || !G || (S < nt && !ckp)) {
        release();
        return c->fail(FWI_ENOMEM, "exact adjoint buffers");
    }
    float **slots[] = {&x.vz, &x.vx, &x.szz, &x.szx, &x.sxx, &x.pv[0], &x.pv[1], &x.pv[2], &x.pv[3],
                       &x.ps[0], &x.ps[1], &x.ps[2], &x.ps[3], &x.bvz, &x.bvx, &x.bzz, &x.bzx, &x.bxx,
                       &x.cv[0], &x.cv[1], &x.cv[2], &x.cv[3], &x.cs[0], &x.cs[1], &x.cs[2], &x.cs[3],
                       &x.E[0], &x.E[1], &x.E[2], &x.E[3]};
    for (int k = 0; k < 30; ++k) *slots[k] = fl + (size_t)k * fs * Bx;
    for (int k = 0; k < 4; ++k) x.D[k] = x.Dv[k] = x.E[k];   // one scratch set serves all three
    x.hist = hist;
    x.hist_base = 0;
    const int rb = c->rec_off[shot0], rows = c->rec_off[shot0 + nshots] - rb;
    const size_t nrow = (size_t)rows * nt, off = (size_t)rb * nt;
    // the receivers' adjoint sources, in res_*: rc x reverse cumulative sum of dt adj
    if (on_device) {
        dev_copy(c->res_z + off, adj_z, nrow, c->st);
        dev_copy(c->res_x + off, adj_x, nrow, c->st);
    } else {
        RETIF(upload(c, c->res_z + off, adj_z, nrow));
        RETIF(upload(c, c->res_x + off, adj_x, nrow));
    }
    x_revcumsum(c->res_z + off, c->res_x + off, c->res_z + off, c->res_x + off, rows, nt, c->g.dt,
                c->g.dt / (c->g.dz * c->g.dx), c->st);
    int isurf[4];
    for (int k = 0; k < 4; ++k) isurf[k] = c->g.isurf[k];
    HIPCHK(hipMemsetAsync(G, 0, sizeof(double) * fs * 5 * nshots, c->st));
    int r = FWI_OK;
    for (int s0 = shot0; s0 < shot0 + nshots && r == FWI_OK; s0 += Bx) {
        const int nb = std::min(Bx, shot0 + nshots - s0);
        x.B = nb;
        x.G = G + (size_t)(s0 - shot0) * 5 * fs;
        (void)hipMemsetAsync(fl, 0, sizeof(float) * fs * Bx * 30, c->st);
        KArgs ka{};   // the injection kernel's view of the forward / adjoint velocities
        ka.P = c->P; ka.fs = fs; ka.dt = x.dt; ka.rhoz = c->rhoz; ka.rhox = c->rhox;
        ka.vz = x.vz; ka.vx = x.vx;
        const SrcArgs fsa = src_args(c, s0, nb);
        // the forward state of segment k (5 fields, 8 memories of nb shots) to or from ckp
        float *const state[13] = {x.vz, x.vx, x.szz, x.szx, x.sxx, x.pv[0], x.pv[1], x.pv[2], x.pv[3],
                                  x.ps[0], x.ps[1], x.ps[2], x.ps[3]};
        auto ckpt = [&](int k, bool save) {
            for (int i = 0; i < 13; ++i) {
                float *slot = ckp + ((size_t)k * 13 + i) * fs * Bx;
                (void)hipMemcpyAsync(save ? slot : state[i], save ? state[i] : slot, sizeof(float) * fs * nb,
                                     hipMemcpyDeviceToDevice, c->st);
            }
        };
        // forward over all steps (each step's intermediates at its slot of the current
        // segment), saving the state at each segment start; the last segment's
        // intermediates are then in place for its reverse sweep
        for (int n = 0; n < nt; ++n) {
            if (n % S == 0) {
                x.hist_base = n;
                if (S < nt) ckpt(n / S, true);
            }
            x_forward_step(x, n, isurf, c->st);
            launch_inject(ka, fsa, n, c->st);
        }
        SrcArgs asa = adj_args(c, s0, nb);   // receiver groups, rows of res_*
        asa.type = 1;                        // v' += A (the running sum's transpose)
        asa.scale_shot = nullptr;
        KArgs kb = ka;
        kb.vz = x.bvz;
        kb.vx = x.bvx;
        const SrcArgs *t0 = c->stf_type == 0 ? &fsa : nullptr;
        for (int k = nseg - 1; k >= 0; --k) {
            const int n0 = k * S, n1 = std::min(n0 + S, nt);
            x.hist_base = n0;
            if (k < nseg - 1) {   // recompute the segment's intermediates from its checkpoint
                ckpt(k, false);
                for (int n = n0; n < n1; ++n) {
                    x_forward_step(x, n, isurf, c->st);
                    launch_inject(ka, fsa, n, c->st);
                }
            }
            for (int n = n1 - 1; n >= n0; --n) {
                launch_inject(kb, asa, n, c->st);
                x_adjoint_step(x, n, isurf, t0, c->st);
            }
        }
        if (hipGetLastError() != hipSuccess) r = c->fail(FWI_EHIP, "exact adjoint launch");
    }
    if (r == FWI_OK) {
        x.G = G;
        x_media_chain(x, nshots, c->m_mu, c->m_rho, c->g.snap_z1, c->g.snap_z2, c->g.snap_x1, c->g.snap_x2, d_lam,
                      d_mu, d_rho, c->st);
        const hipError_t e = hipStreamSynchronize(c->st);
        if (e != hipSuccess) r = c->fail(FWI_EHIP, "exact adjoint: %s", hipGetErrorString(e));
    }
    release();
    return r;
}

int32_t fwi_misfit_gradient_exact(fwi_ctx *c, int32_t shot0, int32_t nshots, const double *adj_z,
                                  const double *adj_x, double *d_lam, double *d_mu, double *d_rho,
                                  int32_t on_device) {
    if (!c || !adj_z || !adj_x || !d_lam || !d_mu || !d_rho) return FWI_EINVAL;
    if (!c->have_model) return c->fail(FWI_ESTATE, "fwi_misfit_gradient_exact before a model");
    if (on_device) return exact_gradient(c, shot0, nshots, adj_z, adj_x, true, d_lam, d_mu, d_rho);
    const size_t n = (size_t)c->g.nz * c->g.nx;
    if (!c->mg_out) {
        c->mg_out = c->dalloc<double>(3 * n);
        if (!c->mg_out) return c->fail(FWI_ENOMEM, "misfit-derivative output");
    }
    RETIF(exact_gradient(c, shot0, nshots, adj_z, adj_x, false, c->mg_out, c->mg_out + n, c->mg_out + 2 * n));
    double *outs[3] = {d_lam, d_mu, d_rho};
    for (int k = 0; k < 3; ++k)
        HIPCHK(hipMemcpyAsync(outs[k], c->mg_out + k * n, n * 8, hipMemcpyDeviceToHost, c->st));
    return sync(c);
}

int32_t fwi_misfit_gradient(fwi_ctx *c, int32_t shot0, int32_t nshots, const double *adj_z, const double *adj_x,
                            double *d_lam, double *d_mu, double *d_rho) {
    if (!c || !adj_z || !adj_x || !d_lam || !d_mu || !d_rho) return FWI_EINVAL;
    if (!c->have_model) return c->fail(FWI_ESTATE, "fwi_misfit_gradient before fwi_set_model");
    // host arrays: through the context's own device buffers
    const size_t n = (size_t)c->g.nz * c->g.nx;
    if (!c->mg_out) {
        c->mg_out = c->dalloc<double>(3 * n);
        if (!c->mg_out) return c->fail(FWI_ENOMEM, "misfit-derivative output");
    }
    RETIF(misfit_gradient_dev(c, shot0, nshots, adj_z, adj_x, c->mg_out, c->mg_out + n, c->mg_out + 2 * n));
    double *outs[3] = {d_lam, d_mu, d_rho};
    for (int k = 0; k < 3; ++k)
        HIPCHK(hipMemcpyAsync(outs[k], c->mg_out + k * n, n * 8, hipMemcpyDeviceToHost, c->st));
    return sync(c);
}

int32_t fwi_misfit_gradient_device(fwi_ctx *c, int32_t shot0, int32_t nshots, const double *adj_z,
                                   const double *adj_x, double *d_lam, double *d_mu, double *d_rho) {
    if (!c || !adj_z || !adj_x || !d_lam || !d_mu || !d_rho) return FWI_EINVAL;
    if (!c->have_model) return c->fail(FWI_ESTATE, "fwi_misfit_gradient_device before a model");
    c->adj_on_device = 1;
    const int r = misfit_gradient_dev(c, shot0, nshots, adj_z, adj_x, d_lam, d_mu, d_rho);
    c->adj_on_device = 0;
    return r;
}

int32_t fwi_zero_gradient(fwi_ctx *c) {
    if (!c) return FWI_EINVAL;
    for (int i = 0; i < 3; ++i)
        HIPCHK(hipMemsetAsync(c->G[i], 0, sizeof(double) * (size_t)c->g.nz * c->g.nx, c->st));
    return sync(c);
}

int32_t fwi_get_gradient(fwi_ctx *c, double *a, double *b, double *d) {
    if (!c) return FWI_EINVAL;
    const size_t n = (size_t)c->g.nz * c->g.nx;
    double *o[3] = {a, b, d};
    for (int i = 0; i < 3; ++i)
        if (o[i]) HIPCHK(hipMemcpyAsync(o[i], c->G[i], n * 8, hipMemcpyDeviceToHost, c->st));
    return sync(c);
}

int32_t fwi_set_gradient(fwi_ctx *c, const double *a, const double *b, const double *d) {
    if (!c || !a || !b || !d) return FWI_EINVAL;
    const size_t n = (size_t)c->g.nz * c->g.nx;
    RETIF(upload(c, c->G[0], a, n));
    RETIF(upload(c, c->G[1], b, n));
    RETIF(upload(c, c->G[2], d, n));
    return sync(c);
}

int32_t fwi_taper_gradient(fwi_ctx *c) {
    if (!c) return FWI_EINVAL;
    for (int i = 0; i < 3; ++i)
        dev_taper(c->G[i], c->g.nz, c->g.nx, c->g.snap_z1, c->g.snap_z2, c->g.snap_x1, c->g.snap_x2, c->g.taper,
                  c->st);
    return sync(c);
}

int32_t fwi_pcg_direction(fwi_ctx *c) {
    if (!c) return FWI_EINVAL;
    const size_t n = (size_t)c->g.nz * c->g.nx;
    if (!c->pcg) {
        c->pcg = c->dalloc<double>(6 * n);
        if (!c->pcg) return c->fail(FWI_ENOMEM, "conjugate-gradient state");
        HIPCHK(hipMemsetAsync(c->pcg, 0, sizeof(double) * 6 * n, c->st));
    }
    for (int p = 0; p < 3; ++p) dev_pcg(c->G[p], c->pcg + 2 * p * n, c->pcg + (2 * p + 1) * n, n, c->red, c->st);
    return sync(c);
}

int32_t fwi_smooth_gradient(fwi_ctx *c, int32_t hfs) {
    if (!c) return FWI_EINVAL;
    if (hfs < 2 || hfs > 64) return c->fail(FWI_EINVAL, "smoothing half width %d (2..64)", hfs);
    const int w = 2 * hfs + 1;
    if (c->gfilt_hfs != hfs) {   // gauss_filter_kernel (src/cpu/n_kernel_lib_PSV.cpp:1012-1049)
        if (!c->gfilt) {
            c->gfilt = c->dalloc<double>(129 * 129);
            if (!c->gfilt) return c->fail(FWI_ENOMEM, "Gaussian filter");
        }
        std::vector<double> k((size_t)w * w);
        const double sigma = hfs / 2, q = 2.0 * sigma * sigma;
        double sum = 0.0;
        for (int ii = -hfs; ii <= hfs; ++ii)
            for (int jj = -hfs; jj <= hfs; ++jj) {
                const double v = std::exp(-((ii * ii) / q) - ((jj * jj) / q));
                k[(size_t)(jj + hfs) * w + ii + hfs] = v;
                sum += v;
            }
        for (double &v : k) v /= sum;
        RETIF(upload(c, c->gfilt, k.data(), k.size()));
        c->gfilt_hfs = hfs;
    }
    const size_t n = (size_t)c->g.nz * c->g.nx;
    for (int p = 0; p < 3; ++p) {
        HIPCHK(hipMemcpyAsync(c->tmp, c->G[p], n * 8, hipMemcpyDeviceToDevice, c->st));
        dev_gauss(c->tmp, c->G[p], c->gfilt, c->g.nz, c->g.nx, hfs, c->st);
    }
    return sync(c);
}

int32_t fwi_misfit(fwi_ctx *c, int32_t shot0, int32_t nshots, double *L2) {
    if (!c || !L2) return FWI_EINVAL;
    if (!c->have_model || !c->have_obs) return c->fail(FWI_ESTATE, "fwi_misfit needs a model and observed data");
    std::vector<std::pair<int, int>> bs;
    if (batches(c, shot0, nshots, bs) != FWI_OK) return c->fail(FWI_EINVAL, "shot range");
    double total = 0;
    std::vector<double> m(c->B);
    for (auto [s0, nb] : bs) {
        if (c->o.seis_fp64)
            RETIF(run_forward64(c, s0, nb));
        else
            RETIF(run_forward(c, s0, nb, false, true, false));
        dev_residual(c->syn_z, c->syn_x, c->obs_z, c->obs_x, c->res_z, c->res_x, c->rec_off_d + s0, nb, c->g.nt,
                     c->g.dt, c->misfit, c->st);
        HIPCHK(hipMemcpyAsync(m.data(), c->misfit, sizeof(double) * nb, hipMemcpyDeviceToHost, c->st));
        RETIF(sync(c));
        for (int s = 0; s < nb; ++s) total += m[s];
    }
    if (!std::isfinite(total)) return c->fail(FWI_ENONFINITE, "misfit is not finite");
    *L2 = total;
    return FWI_OK;
}

int32_t fwi_save_model(fwi_ctx *c) {
    if (!c || !c->have_model) return c ? c->fail(FWI_ESTATE, "no model") : FWI_EINVAL;
    const size_t n = (size_t)c->g.nz * c->g.nx * 8;
    HIPCHK(hipMemcpyAsync(c->o_lam, c->m_lam, n, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->o_mu, c->m_mu, n, hipMemcpyDeviceToDevice, c->st));
    HIPCHK(hipMemcpyAsync(c->o_rho, c->m_rho, n, hipMemcpyDeviceToDevice, c->st));
    if (c->o.param == FWI_PARAM_VELOCITY) {   // cp, cs of the saved model: the update's starting point
        const size_t nn = (size_t)c->g.nz * c->g.nx;
        dev_velocity_of(c->o_lam, c->o_mu, c->o_rho, c->vel, c->vel + nn, nn, c->st);
    }
    return sync(c);
}

// update_mat2 for the three parameters (caps and floors of src/cpu/n_simulate_PSV.cpp:542-546)
int32_t fwi_update_model(fwi_ctx *c, double eps, double rho_factor, double rho_min) {
    if (!c || !c->have_model) return c ? c->fail(FWI_ESTATE, "no model") : FWI_EINVAL;
    const size_t n = (size_t)c->g.nz * c->g.nx;
    if (c->o.param == FWI_PARAM_VELOCITY) {
        // cp, cs gradient at the saved model, then update_mat2's factor and clipped step
        // on each (src/cpu/n_kernel_lib_PSV.cpp:785-832) and lam, mu at the saved rho
        double *cp = c->vel, *cs = c->vel + n, *gcp = c->vel + 2 * n, *gcs = c->vel + 3 * n;
        dev_velocity_gradient(c->G[0], c->G[1], c->o_rho, cp, cs, gcp, gcs, n, c->st);
        double f[2];
        const double *vv[2] = {cp, cs}, *gg[2] = {gcp, gcs};
        for (int p = 0; p < 2; ++p) {
            double gmax = 0, am = 0;
            RETIF(absmax_host(c, gg[p], n, &gmax));
            RETIF(absmax_host(c, vv[p], n, &am));
            if (am < c->vb[2 * p + 1]) am = c->vb[2 * p + 1];
            f[p] = eps * (am / gmax);
        }
        dev_velocity_update(c->m_lam, c->m_mu, c->m_rho, c->o_rho, cp, cs, gcp, gcs, f[0], f[1], c->vb, n, c->st);
        return refresh_media(c);
    }
    double *m[3] = {c->m_lam, c->m_mu, c->m_rho};
    const double *old[3] = {c->o_lam, c->o_mu, c->o_rho};
    const double mmax[3] = {4.8e10, 2.7e10, 3000.0}, mmin[3] = {0.0, 0.0, rho_min};
    const double step[3] = {eps, eps, rho_factor * eps};
    for (int p = 0; p < 3; ++p) {
        double gmax = 0, am = 0;
        RETIF(absmax_host(c, c->G[p], n, &gmax));
        RETIF(absmax_host(c, old[p], n, &am));
        if (am < mmax[p]) am = mmax[p];
        const double factor = am / gmax;
        dev_update(m[p], old[p], c->G[p], step[p] * factor, mmax[p], mmin[p], n, c->st);
    }
    return refresh_media(c);
}

// The smallest trial step whose model change the FP32 media resolve: a step eps moves a
// parameter by at most eps times the larger of its largest value and its clip, i.e. by
// ~eps of the values themselves, and below 2^-20 (~16 FP32 ulps) a trial's misfit differs
// from the current one by rounding only.  Such a trial never counts as a decrease (the
// FP64 reference sees the true sign there: along an ascent direction it halves on to its
// "Steplength estimation failed" exit, and so does this search).
constexpr double kResolvedStep = 0x1p-20;

// step_length_PSV (src/cpu/n_step_length_PSV.cpp:13-209) with update_param 0
int32_t fwi_step_length(fwi_ctx *c, double est, double L2_0, double *step_out, int32_t *ntrials) {
    if (!c || !step_out) return FWI_EINVAL;
    const double scalefac = 2.0;
    const int stepmax_1 = 30, stepmax_2 = 4;
    double L2t[3] = {L2_0, 0, 0}, sl[3] = {0, 0, 0};
    int step1 = 0, step2 = 0, itests = 1, iteste = 1, countstep = 0, trials = 0;
    while (step2 != 1 || step1 != 1) {
        for (int itest = itests; itest <= iteste; ++itest) {
            RETIF(fwi_update_model(c, est, 0.5, 1.5));
            double L2 = 0;
            RETIF(fwi_misfit(c, 0, c->nshot, &L2));
            ++trials;
            L2t[itest] = L2;
            sl[itest] = est;
            sl[0] = 0.0;
        }
        const bool lower = L2t[0] > L2t[1] && sl[1] >= kResolvedStep;   // the first trial's test
        if (step1 == 0 && !lower) {
            est = est / scalefac;
            countstep++;
        }
        if (step1 == 1 && L2t[1] < L2t[2]) {
            sl[2] = est;
            step2 = 1;
        }
        if (step1 == 1 && L2t[1] >= L2t[2]) {
            sl[2] = est;
            est = est + est / scalefac;
            countstep++;
        }
        if (step1 == 0 && lower) {
            sl[1] = est;
            step1 = 1;
            iteste = itests = 2;
            countstep = 0;
            est = est + est / scalefac;
        }
        if (step1 == 0 && countstep > stepmax_1) {
            if (ntrials) *ntrials = trials;
            return c->fail(FWI_ESTEP, "no step in %d halvings lowered the misfit", stepmax_1);
        }
        if (step1 == 1 && countstep > stepmax_2) step1 = step2 = 1;
    }
    // calc_opt_step: vertex of the parabola through the three (step, misfit) pairs
    const double d01 = (L2t[1] - L2t[0]) / (sl[1] - sl[0]);
    const double d12 = (L2t[2] - L2t[1]) / (sl[2] - sl[1]);
    const double x2 = (d12 - d01) / (sl[2] - sl[0]);
    const double x1 = d01 - x2 * (sl[0] + sl[1]);
    double opt = -x1 / (2.0 * x2);
    if (opt > sl[2]) opt = sl[2];
    if (opt < 0.0) opt = sl[1];
    if (L2t[1] < L2t[0] && L2t[2] < L2t[1]) opt = sl[2];
    *step_out = opt;
    if (ntrials) *ntrials = trials;
    return FWI_OK;
}

int32_t fwi_iteration(fwi_ctx *c, double *L2_out, double *step_out) {
    if (!c) return FWI_EINVAL;
    RETIF(fwi_save_model(c));
    RETIF(refresh_media(c));
    std::vector<double> m(c->nshot);
    RETIF(fwi_gradient(c, 0, c->nshot, m.data()));
    double L2 = 0;
    for (double v : m) L2 += v;
    // the reference's order (src/cpu/n_simulate_PSV.cpp:431-455): conjugate direction, taper,
    // smoothing -- the first and last disabled there, options here
    if (c->o.pcg) RETIF(fwi_pcg_direction(c));
    RETIF(fwi_taper_gradient(c));
    if (c->o.smooth_hfs) RETIF(fwi_smooth_gradient(c, c->o.smooth_hfs));
    double step = 0;
    RETIF(fwi_step_length(c, c->step_est, L2, &step, nullptr));
    RETIF(fwi_update_model(c, step, 0.5, 1.25));
    c->step_est = step;
    if (L2_out) *L2_out = L2;
    if (step_out) *step_out = step;
    return FWI_OK;
}

int32_t fwi_tuning(const fwi_ctx *c, int32_t *tile_tz, char *log, int32_t log_len) {
    if (!c) return FWI_EINVAL;
    if (tile_tz) *tile_tz = c->tz;
    if (log && log_len > 0) {
        char head[64];
        std::snprintf(head, sizeof head, "chosen o%d/v%d/tz%d/d%d; ", c->order, c->var, c->tz, c->depth);
        const std::string s = (c->tuned ? std::string(head) : std::string()) + c->tune_log;
        std::strncpy(log, s.c_str(), (size_t)log_len - 1);
        log[log_len - 1] = 0;
    }
    return FWI_OK;
}

int32_t fwi_profile(fwi_ctx *c, int32_t enable) {
    return fwi_profile_kernels(c, enable ? (1u << FWI_K_COUNT) - 1 : 0u);
}

int32_t fwi_profile_kernels(fwi_ctx *c, uint32_t mask) {
    if (!c) return FWI_EINVAL;
    RETIF(sync(c));
    c->prof_on = (int)(mask & ((1u << FWI_K_COUNT) - 1));
    if (c->prof_on) std::memset(c->times, 0, sizeof c->times);
    return FWI_OK;
}

int32_t fwi_profile_stride(fwi_ctx *c, int32_t stride) {
    if (!c) return FWI_EINVAL;
    if (stride < 1) return c->fail(FWI_EINVAL, "profile stride %d < 1", stride);
    RETIF(sync(c));
    c->prof_stride = stride;
    return FWI_OK;
}

int32_t fwi_kernel_times(fwi_ctx *c, fwi_kernel_time t[FWI_K_COUNT]) {
    if (!c || !t) return FWI_EINVAL;
    RETIF(sync(c));
    std::memcpy(t, c->times, sizeof c->times);
    return FWI_OK;
}

int32_t fwi_sync(fwi_ctx *c) { return c ? sync(c) : FWI_EINVAL; }

}  // extern "C"
