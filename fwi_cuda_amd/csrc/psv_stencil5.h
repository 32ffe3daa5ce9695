// psv_stencil5.h -- the two-kernel stencil with wave-uniform addressing: variant 5 and
// its full-tile forms 10 to 13 (device bodies, kernels and the launch of one variant).
// Each translation unit that includes it instantiates the kernels of the variants it
// launches (psv_v5.hip: 5 and 10, psv_v11.hip: 11 and 12, psv_v13.hip: 13), so that
// they compile in parallel.
#pragma once

#include "psv_device.h"

namespace psv {
namespace {

// ---------------------------------------------------------------------------
// Variant 5: variant 2's stencil with wave-uniform addressing and ordered loads.
//
// The wavefront's row segment is made wave-uniform (readfirstlane), so the row
// loop, the z-strip tests and the snapshot-lattice tests run on the scalar unit,
// and the z profiles of the strip rows come from the kernel arguments (ZTab) as
// scalar loads.  Every array is addressed as a wave-uniform base pointer (per
// shot and row offset, in SGPRs) plus one 32-bit byte offset per lane, the
// saddr + voffset form of the global loads: a lane keeps a few offset registers
// instead of a 64-bit address per field.
//
// Vector-memory loads retire in order (one vmcnt counter), so waiting for a load
// also waits for every load issued before it.  Each row therefore issues, in
// this order: the loads row z uses only after its derivatives (fields to update,
// CPML memory, displacement, energy, snapshot and gradient sums), then the loads
// of row z + 1's derivative inputs (the queue heads, and sxx for the velocity
// kernel).  The derivatives of row z wait only for what row z - 1 prefetched,
// the update waits only for row z's own loads, and the next row's loads stay in
// flight across both.  All lanes load (halo lanes read clamped, valid cells) and
// only owned lanes store.  The velocity kernel's launch modes are template flags.
// ---------------------------------------------------------------------------

// FULL (variant 10): every lane of a 64-column tile starting at column 64 k owns its
// column (rows start on 256-byte boundaries, no lane is spent on the halo); the N halo
// columns each side that the x taps need come from one extra load per field and row
// into a halo register, read through lane_off_h.
// SB (variant 11): a block's kBY wavefronts are kBY shots of one tile and segment.
// XW (variant 12): a block's kBY wavefronts are kBY adjacent tiles of one segment.
// NT (variant 13): the fields updated in place are read and written with the
// non-temporal hint (ld_once / st_once).
template <int N, int TZ, bool FULL, bool SB = false, bool XW = false, bool NT = false>
__device__ __forceinline__ void stress5_body(const KArgs &a, const ZTab &zt) {
    constexpr int OUTW = FULL ? kBX : kBX - 2 * N;
    int tbx, tby, s;
    tile_of<TZ, OUTW, XW, FULL, SB>(a, N, tbx, tby, s);
    const int lane = threadIdx.x;
    if constexpr (SB) {
        s = s * kBY + wave_uniform((int)threadIdx.y);
        if (s >= a.B) return;
    }
    if constexpr (XW) {   // past the last tile: nothing to do
        tbx = tbx * kBY + wave_uniform((int)threadIdx.y);
        if (tbx * OUTW >= (FULL ? a.nx - N : a.nx - 2 * N)) return;
    }
    const int z0 = N + (SB || XW ? tby : tby * kBY + wave_uniform((int)threadIdx.y)) * TZ;
    const int nzi = a.nz - N;
    if (z0 >= nzi) return;
    const int zend = min(z0 + TZ, nzi);
    const int ix = tbx * OUTW + lane;
    const bool out = (FULL ? ix >= N : (lane >= N && lane < kBX - N)) && ix < a.nx - N;
    const int cx = min(ix, a.nx - 1);   // halo lanes past the grid read a valid column
    const uint32_t c4 = (uint32_t)cx * 4u;
    const uint32_t h4 = FULL ? (uint32_t)halo_col(tbx * OUTW, lane, N, a.nx) * 4u : 0u;
    const uint32_t P4 = (uint32_t)a.P * 4u;
    const size_t P = a.P, fs = (size_t)s * a.fs;
    // base of row z0 + d of a field of this shot, of a media array
    auto fld = [&](float *f, int d) { return f + fs + (size_t)(z0 + d) * P; };
    auto med = [&](const float *f, int d) { return f + (size_t)(z0 + d) * P; };
    const float dt = a.dt;

    // CPML: the lane's x profile; psi_z of this shot (strip row r at r * P);
    // psi_x of this shot from row z0 (row z at (z - z0) * nxs)
    const CpmlPoint xq = a.xp[cx];
    const int cxm = out ? xq.map : -1;
    const uint32_t xm4 = (uint32_t)max(cxm, 0) * 4u, nxs4 = (uint32_t)a.nxs * 4u;
    float *pz1 = a.pvz_z + (size_t)s * a.nzs * P, *pz2 = a.pvx_z + (size_t)s * a.nzs * P;
    float *px1 = a.pvx_x + ((size_t)s * a.nz + z0) * a.nxs, *px2 = a.pvz_x + ((size_t)s * a.nz + z0) * a.nxs;
    // stores of [..][P] rows: owned lanes at their column, halo lanes at pad column P - 1
    const uint32_t dtr = out ? 0u : (uint32_t)(a.P - 1) * 4u - c4;

    // qz[j] = vz[z - N + j], qx[j] = vx[z - N + 1 + j]
    float qz[2 * N], qx[2 * N];
#pragma unroll
    for (int j = 0; j < 2 * N - 1; ++j) {
        qz[j + 1] = ld_at(fld(a.vz, j - N), c4);
        qx[j + 1] = ld_at(fld(a.vx, j - N + 1), c4);
    }
    const float *vzq = fld(a.vz, N - 1), *vxq = fld(a.vx, N);   // queue heads vz[z + N - 1], vx[z + N]
    float *szz = fld(a.szz, 0), *szx = fld(a.szx, 0), *sxx = fld(a.sxx, 0);
    const float *lam = med(a.lam, 0), *mu = med(a.mu, 0), *muzx = med(a.muzx, 0);
    // Queue heads two rows ahead, in two register pairs that alternate between even
    // and odd rows (the loop is unrolled by two, so no in-flight load is copied);
    // the last rows reload row zend - 1.
    const uint32_t vlast = (uint32_t)(zend - 1 - z0) * P4 + c4;
    float hzA = ld_at(vzq, c4), hxA = ld_at(vxq, c4);
    float hzB = ld_at(vzq, min(c4 + P4, vlast)), hxB = ld_at(vxq, min(c4 + P4, vlast));
    // FULL: halo columns of vz, vx at rows z0 and z0 + 1 (byte offsets from row z0 at h4 - c4)
    const float *vzc_row = fld(a.vz, 0), *vxc_row = fld(a.vx, 0);
    const uint32_t hoff = h4 - c4, hlast = (uint32_t)(zend - 1 - z0) * P4 + h4;
    float gzA = 0.f, gxA = 0.f, gzB = 0.f, gxB = 0.f;
    if constexpr (FULL) {
        gzA = ld_at(vzc_row, h4);
        gxA = ld_at(vxc_row, h4);
        gzB = ld_at(vzc_row, min(h4 + P4, hlast));
        gxB = ld_at(vxc_row, min(h4 + P4, hlast));
    }
    // row z (byte offset vo of (z, cx) from row z0); hz, hx hold its heads on entry and
    // row z + 2's on return, gz, gx (FULL) its halo columns of vz, vx and row z + 2's
    auto row = [&](int z, uint32_t vo, float &hz, float &hx, float &gz, float &gx) {
#pragma unroll
        for (int j = 0; j < 2 * N - 1; ++j) {
            qz[j] = qz[j + 1];
            qx[j] = qx[j + 1];
        }
        qz[2 * N - 1] = hz;
        qx[2 * N - 1] = hx;
        // row z's update inputs, then row z + 2's queue heads; nothing else is loaded
        // outside the rare CPML strips, so the waits below count exactly these
        const float c_zz = ld_once<NT>(szz, vo), c_zx = ld_once<NT>(szx, vo), c_xx = ld_once<NT>(sxx, vo);
        const float l = ld_at(lam, vo), m = ld_at(mu, vo), mzx = ld_at(muzx, vo);
        const uint32_t v2 = min(vo + 2 * P4, vlast);
        hz = ld_at(vzq, v2);
        hx = ld_at(vxq, v2);
        const float vzc = qz[N], vxc = qx[N - 1];
        float vzh = 0.f, vxh = 0.f;
        if constexpr (FULL) {   // this row's halo, then row z + 2's
            vzh = gz;
            vxh = gx;
            const uint32_t g2 = min(vo + hoff + 2 * P4, hlast);
            gz = ld_at(vzc_row, g2);
            gx = ld_at(vxc_row, g2);
        }
        float vz_z = 0.f, vx_z = 0.f, vz_x = 0.f, vx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            vz_z += a.cz[k - 1] * (qz[N + k - 1] - qz[N - k]);
            vx_z += a.cz[k - 1] * (qx[N + k - 1] - qx[N - k]);
            if constexpr (FULL) {
                vz_x += a.cx[k - 1] * (lane_off_h(vzc, vzh, lane, k) - lane_off_h(vzc, vzh, lane, 1 - k));
                vx_x += a.cx[k - 1] * (lane_off_h(vxc, vxh, lane, k - 1) - lane_off_h(vxc, vxh, lane, -k));
            } else {
                vz_x += a.cx[k - 1] * (lane_off(vzc, lane, k) - lane_off(vzc, lane, 1 - k));
                vx_x += a.cx[k - 1] * (lane_off(vxc, lane, k - 1) - lane_off(vxc, lane, -k));
            }
        }
        if (in_zstrip(zt, z)) {   // wave-uniform; b, a, 1/K at the integer point, then at the half point
            const int zr = strip_row(zt, N, z);
            const float *pf = zt.p[zr];
            const uint32_t zo = (uint32_t)zr * P4 + c4;
            const float p1 = pf[0] * ld_at(pz1, zo) + pf[1] * vz_z;
            const float p2 = pf[3] * ld_at(pz2, zo) + pf[4] * vx_z;
            st_at(pz1, zo + dtr, p1);
            st_at(pz2, zo + dtr, p2);
            vz_z = vz_z * pf[2] + p1;
            vx_z = vx_z * pf[5] + p2;
        }
        if (cxm >= 0) {
            const uint32_t xo = (uint32_t)(z - z0) * nxs4 + xm4;
            const float p1 = xq.b * ld_at(px1, xo) + xq.a * vx_x;
            const float p2 = xq.bh * ld_at(px2, xo) + xq.ah * vz_x;
            st_at(px1, xo, p1);
            st_at(px2, xo, p2);
            vx_x = vx_x * xq.ik + p1;
            vz_x = vz_x * xq.ikh + p2;
        }
        const float div = l * (vx_x + vz_z);
        float nzz = c_zz + dt * (div + 2.f * m * vz_z);
        float nzx = c_zx + dt * mzx * (vz_x + vx_z);
        float nxx = c_xx + dt * (div + 2.f * m * vx_x);
        if (z == a.surf5) {   // wave-uniform; surf_mirror at row N: no image rows (src/cpu/n_kernel_lib_PSV.cpp:297-301)
            nzz = 0.f;
            nzx = 0.f;
            nxx = 4.f * dt * vx_x * (l * m + m * m) / (l + 2.f * m);
        }
        const uint32_t so = vo + dtr;
        st_once<NT>(szz, so, nzz);
        st_once<NT>(szx, so, nzx);
        st_once<NT>(sxx, so, nxx);
    };
    int z = z0;
    uint32_t vo = c4;
    for (; z + 1 < zend; z += 2, vo += 2 * P4) {
        row(z, vo, hzA, hxA, gzA, gxA);
        row(z + 1, vo + P4, hzB, hxB, gzB, gxB);
    }
    if (z < zend) row(z, vo, hzA, hxA, gzA, gxA);
}

template <int N, int TZ>
__global__ __launch_bounds__(kBX *kBY) void k_stress5(const KArgs a, const ZTab zt) {
    stress5_body<N, TZ, false>(a, zt);
}

template <int N, int TZ>
__global__ __launch_bounds__(kBX *kBY) void k_stress10(const KArgs a, const ZTab zt) {
    stress5_body<N, TZ, true>(a, zt);
}

template <int N, int TZ>
__global__ __launch_bounds__(kBX *kBY) void k_stress11(const KArgs a, const ZTab zt) {
    stress5_body<N, TZ, true, true>(a, zt);
}

template <int N, int TZ>
__global__ __launch_bounds__(kBX *kBY) void k_stress12(const KArgs a, const ZTab zt) {
    stress5_body<N, TZ, true, false, true>(a, zt);
}

template <int N, int TZ>
__global__ __launch_bounds__(kBX *kBY) void k_stress13(const KArgs a, const ZTab zt) {
    stress5_body<N, TZ, true, false, true, true>(a, zt);
}

// The velocity rows of one wavefront; GRAD: this step correlates with the snapshot slab;
// FULL: full 64-column tiles with halo registers (variant 10, as in stress5_body).
template <int N, int TZ, int MODE, bool GRAD, bool FULL, bool SB, bool XW, bool NT>
__device__ __forceinline__ void velocity5_rows(const KArgs &a, const ZTab &zt, const StepView &sv) {
    constexpr int OUTW = FULL ? kBX : kBX - 2 * N;
    constexpr bool displ = (MODE & kVelAdjoint) != 0;   // u += dt v
    int tbx, tby, s;
    tile_of<TZ, OUTW, XW, FULL, SB>(a, N, tbx, tby, s);
    const int lane = threadIdx.x;
    if constexpr (SB) {
        s = s * kBY + wave_uniform((int)threadIdx.y);
        if (s >= a.B) return;
    }
    if constexpr (XW) {   // past the last tile: nothing to do
        tbx = tbx * kBY + wave_uniform((int)threadIdx.y);
        if (tbx * OUTW >= (FULL ? a.nx - N : a.nx - 2 * N)) return;
    }
    const int z0 = N + (SB || XW ? tby : tby * kBY + wave_uniform((int)threadIdx.y)) * TZ;
    const int nzi = a.nz - N;
    if (z0 >= nzi) return;
    const int zend = min(z0 + TZ, nzi);
    const int ix = tbx * OUTW + lane;
    const bool out = (FULL ? ix >= N : (lane >= N && lane < kBX - N)) && ix < a.nx - N;
    const int cx = min(ix, a.nx - 1);
    const uint32_t c4 = (uint32_t)cx * 4u;
    const uint32_t h4 = FULL ? (uint32_t)halo_col(tbx * OUTW, lane, N, a.nx) * 4u : 0u;
    const uint32_t P4 = (uint32_t)a.P * 4u;
    const size_t P = a.P, fs = (size_t)s * a.fs;
    auto fld = [&](float *f, int d) { return f + fs + (size_t)(z0 + d) * P; };
    auto med = [&](const float *f, int d) { return f + (size_t)(z0 + d) * P; };
    const float dt = a.dt;

    const CpmlPoint xq = a.xp[cx];
    const int cxm = out ? xq.map : -1;
    const uint32_t xm4 = (uint32_t)max(cxm, 0) * 4u, nxs4 = (uint32_t)a.nxs * 4u;
    float *pz1 = a.pszz_z + (size_t)s * a.nzs * P, *pz2 = a.pszx_z + (size_t)s * a.nzs * P;
    float *px1 = a.psxx_x + ((size_t)s * a.nz + z0) * a.nxs, *px2 = a.pszx_x + ((size_t)s * a.nz + z0) * a.nxs;
    // stores of [..][P] rows: owned lanes at their column, halo lanes at pad column P - 1
    const uint32_t dtr = out ? 0u : (uint32_t)(a.P - 1) * 4u - c4;

    // snapshot lattice: this step's switches, the lane's column, the next lattice row
    const bool store = (MODE & kVelStore) && sv.store;
    const bool gcol = out && (GRAD || store) && ix_on_lattice(a, ix);
    const uint32_t j4 = gcol ? (uint32_t)((ix - a.snap_x1) / a.snap_dx) * 4u : 0u;
    int zl = a.snap_z1, jz = 0;
    if (z0 > zl) {
        jz = (z0 - zl + a.snap_dz - 1) / a.snap_dz;
        zl += jz * a.snap_dz;
    }
    const size_t slab = (size_t)a.snz * a.snx;
    const size_t soff = (size_t)sv.off;

    // qzz[j] = szz[z - N + 1 + j], qzx[j] = szx[z - N + j]
    float qzz[2 * N], qzx[2 * N];
#pragma unroll
    for (int j = 0; j < 2 * N - 1; ++j) {
        qzz[j + 1] = ld_at(fld(a.szz, j - N + 1), c4);
        qzx[j + 1] = ld_at(fld(a.szx, j - N), c4);
    }
    const float *szzq = fld(a.szz, N), *szxq = fld(a.szx, N - 1);   // queue heads szz[z + N], szx[z + N - 1]
    const float *sxx = fld(a.sxx, 0);
    float *vz = fld(a.vz, 0), *vx = fld(a.vx, 0);
    const float *rhoz = med(a.rhoz, 0), *rhox = med(a.rhox, 0);
    // Derivative inputs two rows ahead, in register sets that alternate between even
    // and odd rows (the loop is unrolled by two); the last rows reload row zend - 1.
    const uint32_t vlast = (uint32_t)(zend - 1 - z0) * P4 + c4;
    float hzzA = ld_at(szzq, c4), hzxA = ld_at(szxq, c4), hxxA = ld_at(sxx, c4);
    const uint32_t v1 = min(c4 + P4, vlast);
    float hzzB = ld_at(szzq, v1), hzxB = ld_at(szxq, v1), hxxB = ld_at(sxx, v1);
    // FULL: halo columns of szx, sxx at rows z0 and z0 + 1
    const float *szxc_row = fld(a.szx, 0);
    const uint32_t hoff = h4 - c4, hlast = (uint32_t)(zend - 1 - z0) * P4 + h4;
    float gzxA = 0.f, gxxA = 0.f, gzxB = 0.f, gxxB = 0.f;
    if constexpr (FULL) {
        gzxA = ld_at(szxc_row, h4);
        gxxA = ld_at(sxx, h4);
        gzxB = ld_at(szxc_row, min(h4 + P4, hlast));
        gxxB = ld_at(sxx, min(h4 + P4, hlast));
    }
    // row z (byte offset vo of (z, cx) from row z0); hzz, hzx, hxx hold its derivative
    // inputs on entry and row z + 2's on return, gzx, gxx (FULL) its halo columns of
    // szx, sxx and row z + 2's
    auto row = [&](int z, uint32_t vo, float &hzz, float &hzx, float &hxx, float &gzx, float &gxx) {
#pragma unroll
        for (int j = 0; j < 2 * N - 1; ++j) {
            qzz[j] = qzz[j + 1];
            qzx[j] = qzx[j + 1];
        }
        qzz[2 * N - 1] = hzz;
        qzx[2 * N - 1] = hzx;
        const float sxxc = hxx;
        // row z's update inputs (with the snapshot and gradient sums of a correlating
        // step: read on every row, a clamped lattice row off the lattice), then row z + 2's
        // derivative inputs; nothing else is loaded outside the rare CPML strips, so the
        // waits below count exactly these
        const bool lrow = (GRAD || store) && z == zl && zl <= a.snap_z2;   // wave-uniform
        const size_t qrow = (size_t)s * slab + (size_t)min(jz, a.snz - 1) * a.snx;
        const float vz0 = ld_once<NT>(vz, vo), vx0 = ld_once<NT>(vx, vo), rz_ = ld_at(rhoz, vo),
                    rx_ = ld_at(rhox, vo);
        float uz0 = 0.f, ux0 = 0.f, we0 = 0.f;
        if constexpr (displ) {
            uz0 = ld_once<NT>(fld(a.uz, 0), vo);
            ux0 = ld_once<NT>(fld(a.ux, 0), vo);
        }
        if constexpr ((MODE & kVelEnergy) != 0) we0 = ld_once<NT>(fld(a.We, 0), vo);
        float Szz = 0.f, Sxx = 0.f, Szx = 0.f, Vz = 0.f, Vx = 0.f, gl = 0.f, gm = 0.f;
        float g0 = 0.f, g1 = 0.f, g2 = 0.f, g3 = 0.f, g4 = 0.f;
        if constexpr (GRAD) {
            const size_t so = soff + (size_t)s * a.sn_stride + (size_t)min(jz, a.snz - 1) * a.snx;
            Szz = ld_once<NT>(a.sn_szz + so, j4); Sxx = ld_once<NT>(a.sn_sxx + so, j4);
            Szx = ld_once<NT>(a.sn_szx + so, j4);
            Vz = ld_once<NT>(a.sn_vz + so, j4); Vx = ld_once<NT>(a.sn_vx + so, j4);
            gl = ld_at(med(a.lam, 0), vo); gm = ld_at(med(a.mu, 0), vo);
            g0 = ld_once<NT>(a.g_lam + qrow, j4); g1 = ld_once<NT>(a.g_mu + qrow, j4);
            g2 = ld_once<NT>(a.g_rho + qrow, j4);
            if constexpr ((MODE & kVelMisfit) != 0) {
                g3 = ld_once<NT>(a.g_ext + qrow, j4);
                g4 = ld_once<NT>(a.g_ext + a.g_ext_stride + qrow, j4);
            }
        }
        const uint32_t v2 = min(vo + 2 * P4, vlast);
        hzz = ld_at(szzq, v2);
        hzx = ld_at(szxq, v2);
        hxx = ld_at(sxx, v2);
        const float szxc = qzx[N];
        float szxh = 0.f, sxxh = 0.f;
        if constexpr (FULL) {   // this row's halo, then row z + 2's
            szxh = gzx;
            sxxh = gxx;
            const uint32_t g2 = min(vo + hoff + 2 * P4, hlast);
            gzx = ld_at(szxc_row, g2);
            gxx = ld_at(sxx, g2);
        }
        float szz_z = 0.f, szx_z = 0.f, szx_x = 0.f, sxx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            szz_z += a.cz[k - 1] * (qzz[N + k - 1] - qzz[N - k]);
            szx_z += a.cz[k - 1] * (qzx[N + k - 1] - qzx[N - k]);
            if constexpr (FULL) {
                szx_x += a.cx[k - 1] * (lane_off_h(szxc, szxh, lane, k - 1) - lane_off_h(szxc, szxh, lane, -k));
                sxx_x += a.cx[k - 1] * (lane_off_h(sxxc, sxxh, lane, k) - lane_off_h(sxxc, sxxh, lane, 1 - k));
            } else {
                szx_x += a.cx[k - 1] * (lane_off(szxc, lane, k - 1) - lane_off(szxc, lane, -k));
                sxx_x += a.cx[k - 1] * (lane_off(sxxc, lane, k) - lane_off(sxxc, lane, 1 - k));
            }
        }
        // integer-point profiles for szz_z and sxx_x, half-point for szx_z and szx_x,
        // the reference's slot mapping (src/cpu/n_kernel_PSV.cpp:142-146)
        if (in_zstrip(zt, z)) {   // wave-uniform
            const int zr = strip_row(zt, N, z);
            const float *pf = zt.p[zr];
            const uint32_t zo = (uint32_t)zr * P4 + c4;
            const float p1 = pf[0] * ld_at(pz1, zo) + pf[1] * szz_z;
            const float p2 = pf[3] * ld_at(pz2, zo) + pf[4] * szx_z;
            st_at(pz1, zo + dtr, p1);
            st_at(pz2, zo + dtr, p2);
            szz_z = szz_z * pf[2] + p1;
            szx_z = szx_z * pf[5] + p2;
        }
        if (cxm >= 0) {
            const uint32_t xo = (uint32_t)(z - z0) * nxs4 + xm4;
            const float p1 = xq.b * ld_at(px1, xo) + xq.a * sxx_x;
            const float p2 = xq.bh * ld_at(px2, xo) + xq.ah * szx_x;
            st_at(px1, xo, p1);
            st_at(px2, xo, p2);
            sxx_x = sxx_x * xq.ik + p1;
            szx_x = szx_x * xq.ikh + p2;
        }
        const float vz1 = vz0 + dt * rz_ * (szx_x + szz_z);
        const float vx1 = vx0 + dt * rx_ * (sxx_x + szx_z);
        const uint32_t so = vo + dtr;
        st_once<NT>(vz, so, vz1);
        st_once<NT>(vx, so, vx1);
        float uz = 0.f, ux = 0.f;
        if constexpr (displ) {
            uz = uz0 + dt * vz0;
            ux = ux0 + dt * vx0;
            st_once<NT>(fld(a.uz, 0), so, uz);
            st_once<NT>(fld(a.ux, 0), so, ux);
        }
        if constexpr ((MODE & kVelEnergy) != 0) st_once<NT>(fld(a.We, 0), so, we0 + (vx1 * vx1 + vz1 * vz1));
        if (lrow) {
            const float c_zz = qzz[N - 1], c_xx = sxxc;
            if (store && gcol) {   // gard_fwd_storage2; k_inject rewrites source cells after injection
                const size_t o = soff + (size_t)s * a.st_stride + (size_t)jz * a.snx;
                st_once<NT>(a.st_vz + o, j4, vz1 / dt);
                st_once<NT>(a.st_vx + o, j4, vx1 / dt);
                st_once<NT>(a.st_szz + o, j4, c_zz);
                st_once<NT>(a.st_szx + o, j4, szxc);
                st_once<NT>(a.st_sxx + o, j4, c_xx);
            }
            if (GRAD && gcol) {
                const float lm = gl + gm;
                const float il = a.gain_lm / (lm * lm), im = a.gain_lm / (gm * gm);
                const float s1 = 0.25f * il * (Szz + Sxx) * (c_zz + c_xx);
                const float s2 = 0.25f * im * (Szz - Sxx) * (c_zz - c_xx);
                const float s3 = im * (Szx * szxc);
                st_once<NT>(a.g_lam + qrow, j4, g0 + a.gw * s1);
                if (!((MODE & kVelMisfit) != 0 && a.gform)) {
                    st_once<NT>(a.g_mu + qrow, j4, g1 + a.gw * (s3 + s1 + s2));
                    st_once<NT>(a.g_rho + qrow, j4, g2 + a.gw * (ux * Vx + uz * Vz));
                } else {
                    st_once<NT>(a.g_mu + qrow, j4, g1 + a.gw * (s1 + s2));
                    st_once<NT>(a.g_rho + qrow, j4, g2 + a.gw * (dt * (vz1 * Vz)));
                    st_once<NT>(a.g_ext + qrow, j4, g3 + a.gw * s3);
                    st_once<NT>(a.g_ext + a.g_ext_stride + qrow, j4, g4 + a.gw * (dt * (vx1 * Vx)));
                }
            }
            zl += a.snap_dz;
            ++jz;
        }
    };
    int z = z0;
    uint32_t vo = c4;
    for (; z + 1 < zend; z += 2, vo += 2 * P4) {
        row(z, vo, hzzA, hzxA, hxxA, gzxA, gxxA);
        row(z + 1, vo + P4, hzzB, hzxB, hxxB, gzxB, gxxB);
    }
    if (z < zend) row(z, vo, hzzA, hzxA, hxxA, gzxA, gxxA);
}

// kVelAdjoint launches branch once, on the step's correlation switch, to a loop
// whose loads are all unconditional
template <int N, int TZ, int MODE, bool FULL, bool SB = false, bool XW = false, bool NT = false>
__device__ __forceinline__ void velocity5_body(const KArgs &a, const ZTab &zt) {
    const StepView sv = step_view(a, 0);
    if constexpr ((MODE & kVelAdjoint) != 0) {
        if (sv.grad) {
            velocity5_rows<N, TZ, MODE, true, FULL, SB, XW, NT>(a, zt, sv);
            return;
        }
    }
    velocity5_rows<N, TZ, MODE, false, FULL, SB, XW, NT>(a, zt, sv);
}

template <int N, int TZ, int MODE>
__global__ __launch_bounds__(kBX *kBY) void k_velocity5(const KArgs a, const ZTab zt) {
    velocity5_body<N, TZ, MODE, false>(a, zt);
}

template <int N, int TZ, int MODE>
__global__ __launch_bounds__(kBX *kBY) void k_velocity10(const KArgs a, const ZTab zt) {
    velocity5_body<N, TZ, MODE, true>(a, zt);
}

template <int N, int TZ, int MODE>
__global__ __launch_bounds__(kBX *kBY) void k_velocity11(const KArgs a, const ZTab zt) {
    velocity5_body<N, TZ, MODE, true, true>(a, zt);
}

template <int N, int TZ, int MODE>
__global__ __launch_bounds__(kBX *kBY) void k_velocity12(const KArgs a, const ZTab zt) {
    velocity5_body<N, TZ, MODE, true, false, true>(a, zt);
}

template <int N, int TZ, int MODE>
__global__ __launch_bounds__(kBX *kBY) void k_velocity13(const KArgs a, const ZTab zt) {
    velocity5_body<N, TZ, MODE, true, false, true, true>(a, zt);
}

#define PSV_TZ5_CASES(KER, N, ...)                                                            \
    if (tz == 32)                                                                             \
        hipLaunchKernelGGL((KER<N, 32 __VA_ARGS__>), grid, block, 0, st, a, *zt);             \
    else                                                                                      \
        hipLaunchKernelGGL((KER<N, 16 __VA_ARGS__>), grid, block, 0, st, a, *zt);

#define PSV_N5_CASES(KER, ...)                             \
    switch (N) {                                         \
        case 1: PSV_TZ5_CASES(KER, 1, __VA_ARGS__) break;  \
        case 2: PSV_TZ5_CASES(KER, 2, __VA_ARGS__) break;  \
        case 3: PSV_TZ5_CASES(KER, 3, __VA_ARGS__) break;  \
        case 4: PSV_TZ5_CASES(KER, 4, __VA_ARGS__) break;  \
        case 5: PSV_TZ5_CASES(KER, 5, __VA_ARGS__) break;  \
        default: PSV_TZ5_CASES(KER, 6, __VA_ARGS__) break; \
    }

// the velocity launch modes of variant V: the ones the passes run get their own code
// (false: a launch mode it has no kernel for)
template <int V>
bool launch_velocity5(const KArgs &a, const ZTab *zt, int N, int tz, dim3 grid, dim3 block, hipStream_t st) {
    if constexpr (V == 13) {
        switch (vel_mode(a)) {
            case 0: PSV_N5_CASES(k_velocity13, , 0) return true;
            case kVelEnergy: PSV_N5_CASES(k_velocity13, , kVelEnergy) return true;
            case kVelStore: PSV_N5_CASES(k_velocity13, , kVelStore) return true;
            case kVelEnergy | kVelStore: PSV_N5_CASES(k_velocity13, , kVelEnergy | kVelStore) return true;
            case kVelEnergy | kVelAdjoint: PSV_N5_CASES(k_velocity13, , kVelEnergy | kVelAdjoint) return true;
            case kVelEnergy | kVelAdjoint | kVelMisfit:
                PSV_N5_CASES(k_velocity13, , kVelEnergy | kVelAdjoint | kVelMisfit) return true;
            default: return false;
        }
    } else if constexpr (V == 12) {
        switch (vel_mode(a)) {
            case 0: PSV_N5_CASES(k_velocity12, , 0) return true;
            case kVelEnergy: PSV_N5_CASES(k_velocity12, , kVelEnergy) return true;
            case kVelStore: PSV_N5_CASES(k_velocity12, , kVelStore) return true;
            case kVelEnergy | kVelStore: PSV_N5_CASES(k_velocity12, , kVelEnergy | kVelStore) return true;
            case kVelEnergy | kVelAdjoint: PSV_N5_CASES(k_velocity12, , kVelEnergy | kVelAdjoint) return true;
            case kVelEnergy | kVelAdjoint | kVelMisfit:
                PSV_N5_CASES(k_velocity12, , kVelEnergy | kVelAdjoint | kVelMisfit) return true;
            default: return false;
        }
    } else if constexpr (V == 11) {
        switch (vel_mode(a)) {
            case 0: PSV_N5_CASES(k_velocity11, , 0) return true;
            case kVelEnergy: PSV_N5_CASES(k_velocity11, , kVelEnergy) return true;
            case kVelStore: PSV_N5_CASES(k_velocity11, , kVelStore) return true;
            case kVelEnergy | kVelStore: PSV_N5_CASES(k_velocity11, , kVelEnergy | kVelStore) return true;
            case kVelEnergy | kVelAdjoint: PSV_N5_CASES(k_velocity11, , kVelEnergy | kVelAdjoint) return true;
            case kVelEnergy | kVelAdjoint | kVelMisfit:
                PSV_N5_CASES(k_velocity11, , kVelEnergy | kVelAdjoint | kVelMisfit) return true;
            default: return false;
        }
    } else if constexpr (V == 10) {
        switch (vel_mode(a)) {
            case 0: PSV_N5_CASES(k_velocity10, , 0) return true;
            case kVelEnergy: PSV_N5_CASES(k_velocity10, , kVelEnergy) return true;
            case kVelStore: PSV_N5_CASES(k_velocity10, , kVelStore) return true;
            case kVelEnergy | kVelStore: PSV_N5_CASES(k_velocity10, , kVelEnergy | kVelStore) return true;
            case kVelEnergy | kVelAdjoint: PSV_N5_CASES(k_velocity10, , kVelEnergy | kVelAdjoint) return true;
            case kVelEnergy | kVelAdjoint | kVelMisfit:
                PSV_N5_CASES(k_velocity10, , kVelEnergy | kVelAdjoint | kVelMisfit) return true;
            default: return false;
        }
    } else {
        switch (vel_mode(a)) {
            case 0: PSV_N5_CASES(k_velocity5, , 0) return true;                                   // recomputation
            case kVelEnergy: PSV_N5_CASES(k_velocity5, , kVelEnergy) return true;                 // forward
            case kVelStore: PSV_N5_CASES(k_velocity5, , kVelStore) return true;                   // recomputation, stored
            case kVelEnergy | kVelStore: PSV_N5_CASES(k_velocity5, , kVelEnergy | kVelStore) return true;
            case kVelEnergy | kVelAdjoint: PSV_N5_CASES(k_velocity5, , kVelEnergy | kVelAdjoint) return true;
            case kVelEnergy | kVelAdjoint | kVelMisfit:
                PSV_N5_CASES(k_velocity5, , kVelEnergy | kVelAdjoint | kVelMisfit) return true;
            default: return false;
        }
    }
}

// the stress half-step of variant V
template <int V>
void launch_stress5(const KArgs &a, const ZTab *zt, int N, int tz, dim3 grid, dim3 block, hipStream_t st) {
    if constexpr (V == 13) {
        PSV_N5_CASES(k_stress13, )
    } else if constexpr (V == 12) {
        PSV_N5_CASES(k_stress12, )
    } else if constexpr (V == 11) {
        PSV_N5_CASES(k_stress11, )
    } else if constexpr (V == 10) {
        PSV_N5_CASES(k_stress10, )
    } else {
        PSV_N5_CASES(k_stress5, )
    }
}

// one half-step of variant V (false: a velocity launch mode it has no kernel for)
template <int V>
bool launch5(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st) {
    const dim3 block(kBX, kBY);
    if (!velocity) {
        launch_stress5<V>(a, &zt, N, tz, grid, block, st);
        return true;
    }
    return launch_velocity5<V>(a, &zt, N, tz, grid, block, st);
}

}  // namespace
}  // namespace psv
