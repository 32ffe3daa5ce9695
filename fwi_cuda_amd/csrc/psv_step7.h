// psv_step7.h -- variant 7: the one-pass time step with ordered, row-ahead loads.
//
// Variant 4's scheme -- stress, top free surface and velocity of a time step in a
// single pass over two time levels, the stresses recomputed on an x and z halo --
// with variant 6's machinery: wave-uniform addressing (a base pointer per array in
// SGPRs, one 32-bit offset per lane), the z-strip test from the kernel arguments
// (the strip rows' profiles are read inside their blocks: a table indexed in the
// kernel arguments would be copied to scratch in a kernel this size), every load of
// input row j + 1 issued while row j is computed,
// register rings with compile-time slots (the loop is unrolled over a block of
// R = 2N + 2 rows, so no in-flight value is copied), CPML-strip loads inside their
// rare blocks, and stores of lanes or rows that own no output sent to the pad
// column P - 1 instead of predicated.
//
// Input row j runs from z0 - N to z1 + N - 1.  It computes the new stresses of row
// j from the old velocities (vz rows j - N .. j + N - 1, vx rows j - N + 1 .. j + N,
// in rings), and, from row z0 + N on, the new velocities of row z = j - N from the
// new stresses of rows z - N .. z + N (rings) and the old vz, vx of row z (in the
// velocity rings).  A wavefront loads 64 columns, computes stresses on the inner
// 64 - 2N and writes stresses and velocities on the inner 64 - 4N.  Fields and the
// stress-side CPML memory are read at the old level and written at the new one
// (the host swaps the levels every step); the velocity-side memory, u and We are
// updated in place by their owner.  Compulsory traffic per cell-step: 68 B forward
// (R+W five fields, R five media, R+W We): each array read and written once.
//
// A block's kBY wavefronts are kBY consecutive z segments of one column tile, or, in
// variant 8 (XW), kBY consecutive column tiles of one z segment: the block then reads
// kBY * 256 contiguous bytes of each row it touches.
//
//   stress   vdiff2 + pml_diff2 + update_s2 (src/cpu/n_kernel_lib_PSV.cpp:100-131, 134-178,
//            181-205); surf_mirror's top row (:297-301)
//   velocity sdiff2 + pml_diff2 + update_v2 (:207-270), fwi_grad2 (:439-484) on correlating
//            steps, gard_fwd_storage2 (:398-436) on storing steps
//
// Included by psv_step7.hip (orders 2, 4, 6 and the launch), psv_step7_b.hip (order 8)
// and psv_step7_c.hip (orders 10, 12), which instantiate the kernels of their orders in
// parallel.
#pragma once

#include <algorithm>

#include "psv_device.h"

namespace psv {
namespace {

// Rows [z0, z1) of shot s for the wavefront's column tile tx (z0 >= N, z1 <= nz - N).
template <int N, int MODE, bool GRAD>
__device__ __forceinline__ void step7_rows(const KArgs &a, const ZTab &zt, const StepView &sv, int s, int tx, int z0,
                                           int z1) {
    constexpr int R = 2 * N + 2, RS = N + 1, OUTW = kBX - 4 * N;
    constexpr bool adj = (MODE & kVelAdjoint) != 0, energy = (MODE & kVelEnergy) != 0;
    const int lane = threadIdx.x;
    const int x = tx * OUTW + lane - N;         // lane 2N holds the first owned column N + tx * OUTW
    const int nzi = a.nz - N;
    if (z0 >= z1 || tx * OUTW >= a.nx - 2 * N) return;
    const int cx = min(max(x, 0), a.nx - 1);   // lanes off the grid read a valid column
    // stresses are updated where the lane has its x taps and the column is interior;
    // lanes [2N, 64 - 2N) own their column
    const bool sint = lane >= N && lane < kBX - N && x >= N && x < a.nx - N;
    const bool own = lane >= 2 * N && lane < kBX - 2 * N && x < a.nx - N;
    const uint32_t c4 = (uint32_t)cx * 4u, P4 = (uint32_t)a.P * 4u;
    const size_t P = a.P, fs = (size_t)s * a.fs;
    const float dt = a.dt;
    // [..][nz][P] arrays from row rb (the lowest row read); offsets clamp to row nz - 1
    const int rb = max(z0 - 2 * N, 0);
    auto fld = [&](float *f) { return f + fs + (size_t)rb * P; };
    auto med = [&](const float *f) { return f + (size_t)rb * P; };
    const uint32_t olast = (uint32_t)(a.nz - 1 - rb) * P4 + c4;
    auto orow = [&](int r) { return min((uint32_t)(r - rb) * P4 + c4, olast); };   // rows < rb wrap and clamp
    auto clampo = [&](uint32_t o) { return min(o, olast); };
    const uint32_t dpad = (uint32_t)(a.P - 1) * 4u - c4;   // offset + dpad: the row's pad column

    const CpmlPoint xq = a.xp[cx];
    const int cxs = sint ? xq.map : -1, cxv = own ? xq.map : -1;
    const uint32_t xs4 = (uint32_t)max(cxs, 0) * 4u, xv4 = (uint32_t)max(cxv, 0) * 4u;
    const uint32_t nxs4 = (uint32_t)a.nxs * 4u;
    const size_t zso = (size_t)s * a.nzs * P, xso = ((size_t)s * a.nz + rb) * a.nxs;   // psi_z, psi_x of this shot

    // snapshot lattice: the step's store switch, the lane's column, the next lattice
    // row of the computed velocity rows (zl, jz) and of the loaded rows (zl_l, jz_l)
    const bool store = (MODE & kVelStore) && sv.store;
    const bool gcol = own && (GRAD || store) && ix_on_lattice(a, x);
    const uint32_t j4 = gcol ? (uint32_t)((x - a.snap_x1) / a.snap_dx) * 4u : 0u;
    int zl = a.snap_z1, jz = 0;
    if (z0 > zl) {
        jz = (z0 - zl + a.snap_dz - 1) / a.snap_dz;
        zl += jz * a.snap_dz;
    }
    int zl_l = zl, jz_l = jz;
    const size_t slab = (size_t)a.snz * a.snx;
    const size_t soff = (size_t)sv.off;

    // old level, read
    const float *vzb = fld(a.vz), *vxb = fld(a.vx);
    const float *szz = fld(a.szz), *szx = fld(a.szx), *sxx = fld(a.sxx);
    const float *lam = med(a.lam), *mu = med(a.mu), *muzx = med(a.muzx), *rhoz = med(a.rhoz), *rhox = med(a.rhox);
    // new level, written
    float *oszz = fld(a.o_szz), *oszx = fld(a.o_szx), *osxx = fld(a.o_sxx), *ovz = fld(a.o_vz), *ovx = fld(a.o_vx);

    // Input row i = j - (z0 - N).  Velocity rings: vz, vx of row j + d at slot (i + N + d) % R.
    // New-stress rings: row j + d at slot (i + d) % R (szz, szx) and (i + d) % RS (sxx).
    // Per-row inputs of input row i (stress inputs of row j, velocity inputs of row
    // j - N) at slot i % 2.
    float qz[R], qx[R], nzz[R], nzx[R], nxx[RS];
    float r_zz[2], r_zx[2], r_xx[2], r_l[2], r_m[2], r_mzx[2];
    float r_rz[2] = {}, r_rx[2] = {}, r_uz[2] = {}, r_ux[2] = {}, r_we[2] = {};
    float r_S[5][2] = {}, r_gl[2] = {}, r_gm[2] = {}, r_g[5][2] = {};

    // stress inputs of row j (offset o) into slot k
    auto load_stress_in = [&](auto slot, uint32_t o) {
        constexpr int k = decltype(slot)::value;
        o = clampo(o);
        r_zz[k] = ld_at(szz, o); r_zx[k] = ld_at(szx, o); r_xx[k] = ld_at(sxx, o);
        r_l[k] = ld_at(lam, o); r_m[k] = ld_at(mu, o); r_mzx[k] = ld_at(muzx, o);
    };
    // velocity inputs of row z (offset o) into slot k
    auto load_vel_in = [&](auto slot, int z, uint32_t o) {
        constexpr int k = decltype(slot)::value;
        o = clampo(o);
        r_rz[k] = ld_at(rhoz, o); r_rx[k] = ld_at(rhox, o);
        if constexpr (adj) {
            r_uz[k] = ld_at(fld(a.uz), o);
            r_ux[k] = ld_at(fld(a.ux), o);
        }
        if constexpr (energy) r_we[k] = ld_at(fld(a.We), o);
        if constexpr (GRAD) {   // read on every row: a clamped lattice row off the lattice
            const int jq = min(jz_l, a.snz - 1);
            const size_t so = soff + (size_t)s * a.sn_stride + (size_t)jq * a.snx;
            const size_t q = (size_t)s * slab + (size_t)jq * a.snx;
            r_S[0][k] = ld_at(a.sn_szz + so, j4); r_S[1][k] = ld_at(a.sn_sxx + so, j4);
            r_S[2][k] = ld_at(a.sn_szx + so, j4); r_S[3][k] = ld_at(a.sn_vz + so, j4);
            r_S[4][k] = ld_at(a.sn_vx + so, j4);
            r_gl[k] = ld_at(lam, o); r_gm[k] = ld_at(mu, o);
            r_g[0][k] = ld_at(a.g_lam + q, j4); r_g[1][k] = ld_at(a.g_mu + q, j4);
            r_g[2][k] = ld_at(a.g_rho + q, j4);
            if constexpr ((MODE & kVelMisfit) != 0) {
                r_g[3][k] = ld_at(a.g_ext + q, j4);
                r_g[4][k] = ld_at(a.g_ext + a.g_ext_stride + q, j4);
            }
            if (z == zl_l && zl_l <= a.snap_z2) {
                zl_l += a.snap_dz;
                ++jz_l;
            }
        }
    };

    // Input row i (compile-time I = i, or i's phase in a block: only I % R, I % RS and
    // I % 2 matter), row j, offset vo of row j.  VEL: update the velocities of row
    // j - N; NVEL: the next input row updates velocities (load their inputs).
    auto row = [&](auto ci, auto vel, auto nvel, int j, uint32_t vo) {
        constexpr int I = decltype(ci)::value;
        constexpr bool VEL = decltype(vel)::value, NVEL = decltype(nvel)::value;
        constexpr int c = I % 2;
        auto sv_ = [](int d) { return ((I + N + d) % R + R) % R; };   // velocity ring slot of row j + d
        auto ss_ = [](int d) { return ((I + d) % R + R) % R; };       // stress ring slot of row j + d
        // ---- loads of input row i + 1: queue rows, then per-row inputs
        qz[sv_(N)] = ld_at(vzb, clampo(vo + (uint32_t)N * P4));
        qx[sv_(N + 1)] = ld_at(vxb, clampo(vo + (uint32_t)(N + 1) * P4));
        load_stress_in(std::integral_constant<int, (I + 1) % 2>{}, vo + P4);
        if constexpr (NVEL) load_vel_in(std::integral_constant<int, (I + 1) % 2>{}, j + 1 - N, vo + P4 - (uint32_t)N * P4);

        // ---- stress half-step on row j
        const float vzc = qz[sv_(0)], vxc = qx[sv_(0)];
        float vz_z = 0.f, vx_z = 0.f, vz_x = 0.f, vx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            vz_z += a.cz[k - 1] * (qz[sv_(k - 1)] - qz[sv_(-k)]);
            vx_z += a.cz[k - 1] * (qx[sv_(k)] - qx[sv_(1 - k)]);
            vz_x += a.cx[k - 1] * (lane_off(vzc, lane, k) - lane_off(vzc, lane, 1 - k));
            vx_x += a.cx[k - 1] * (lane_off(vxc, lane, k - 1) - lane_off(vxc, lane, -k));
        }
        const bool jint = j >= N && j < nzi;               // wave-uniform
        const bool wrow = j >= z0 && j < z1;               // this segment's row
        const uint32_t so = (own && wrow) ? vo : vo + dpad;   // new stresses (and stress-side psi_z)
        if (jint && in_zstrip(zt, j)) {   // wave-uniform; the row's packed profile and strip row
            const KSeg ks = karg_seg();
            const CpmlPoint pz = PSV_KARG(ks, zp)[j];
            const uint32_t zo = (uint32_t)pz.map * P4 + c4, zs = (own && wrow) ? zo : zo + dpad;
            const float p1 = pz.b * ld_at(PSV_KARG(ks, pvz_z) + zso, zo) + pz.a * vz_z;
            const float p2 = pz.bh * ld_at(PSV_KARG(ks, pvx_z) + zso, zo) + pz.ah * vx_z;
            st_at(PSV_KARG(ks, o_pvz_z) + zso, zs, p1);
            st_at(PSV_KARG(ks, o_pvx_z) + zso, zs, p2);
            vz_z = vz_z * pz.ik + p1;
            vx_z = vx_z * pz.ikh + p2;
        }
        if (jint && cxs >= 0) {
            const KSeg ks = karg_seg();
            const uint32_t xo = (uint32_t)(j - rb) * nxs4 + xs4;
            const float p1 = xq.b * ld_at(PSV_KARG(ks, pvx_x) + xso, xo) + xq.a * vx_x;
            const float p2 = xq.bh * ld_at(PSV_KARG(ks, pvz_x) + xso, xo) + xq.ah * vz_x;
            if (own && wrow) {
                st_at(PSV_KARG(ks, o_pvx_x) + xso, xo, p1);
                st_at(PSV_KARG(ks, o_pvz_x) + xso, xo, p2);
            }
            vx_x = vx_x * xq.ik + p1;
            vz_x = vz_x * xq.ikh + p2;
        }
        const float l = r_l[c], m = r_m[c];
        const float div = l * (vx_x + vz_z);
        float tzz = r_zz[c] + dt * (div + 2.f * m * vz_z);
        float tzx = r_zx[c] + dt * r_mzx[c] * (vz_x + vx_z);
        float txx = r_xx[c] + dt * (div + 2.f * m * vx_x);
        if (j == a.surf_top) {   // surf_mirror at row N: no image rows (src/cpu/n_kernel_lib_PSV.cpp:297-301)
            tzz = 0.f;
            tzx = 0.f;
            txx = 4.f * dt * vx_x * (l * m + m * m) / (l + 2.f * m);
        }
        const bool upd = sint && jint;   // rows and columns outside the interior keep their values
        const float n_zz = upd ? tzz : r_zz[c], n_zx = upd ? tzx : r_zx[c], n_xx = upd ? txx : r_xx[c];
        st_at(oszz, so, n_zz);
        st_at(oszx, so, n_zx);
        st_at(osxx, so, n_xx);
        nzz[ss_(0)] = n_zz;
        nzx[ss_(0)] = n_zx;
        nxx[((I % RS) + RS) % RS] = n_xx;
        if constexpr (!VEL) return;

        // ---- velocity half-step on row z = j - N
        const int z = j - N;
        const uint32_t vz_o = vo - (uint32_t)N * P4;            // offset of row z
        const uint32_t svo = own ? vz_o : vz_o + dpad;          // its stores
        const float szxc = nzx[ss_(-N)], sxxc = nxx[(((I - N) % RS) + RS) % RS];
        float szz_z = 0.f, szx_z = 0.f, szx_x = 0.f, sxx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            szz_z += a.cz[k - 1] * (nzz[ss_(k - N)] - nzz[ss_(1 - N - k)]);
            szx_z += a.cz[k - 1] * (nzx[ss_(k - 1 - N)] - nzx[ss_(-N - k)]);
            szx_x += a.cx[k - 1] * (lane_off(szxc, lane, k - 1) - lane_off(szxc, lane, -k));
            sxx_x += a.cx[k - 1] * (lane_off(sxxc, lane, k) - lane_off(sxxc, lane, 1 - k));
        }
        // integer-point profiles for szz_z and sxx_x, half-point for szx_z and szx_x,
        // the reference's slot mapping (src/cpu/n_kernel_PSV.cpp:142-146)
        if (in_zstrip(zt, z)) {
            const KSeg ks = karg_seg();
            const CpmlPoint pz = PSV_KARG(ks, zp)[z];
            const uint32_t zo = (uint32_t)pz.map * P4 + c4, zs = own ? zo : zo + dpad;
            const float p1 = pz.b * ld_at(PSV_KARG(ks, pszz_z) + zso, zo) + pz.a * szz_z;
            const float p2 = pz.bh * ld_at(PSV_KARG(ks, pszx_z) + zso, zo) + pz.ah * szx_z;
            st_at(PSV_KARG(ks, pszz_z) + zso, zs, p1);
            st_at(PSV_KARG(ks, pszx_z) + zso, zs, p2);
            szz_z = szz_z * pz.ik + p1;
            szx_z = szx_z * pz.ikh + p2;
        }
        if (cxv >= 0) {
            const KSeg ks = karg_seg();
            const uint32_t xo = (uint32_t)(z - rb) * nxs4 + xv4;
            const float p1 = xq.b * ld_at(PSV_KARG(ks, psxx_x) + xso, xo) + xq.a * sxx_x;
            const float p2 = xq.bh * ld_at(PSV_KARG(ks, pszx_x) + xso, xo) + xq.ah * szx_x;
            st_at(PSV_KARG(ks, psxx_x) + xso, xo, p1);
            st_at(PSV_KARG(ks, pszx_x) + xso, xo, p2);
            sxx_x = sxx_x * xq.ik + p1;
            szx_x = szx_x * xq.ikh + p2;
        }
        const float vz0 = qz[sv_(-N)], vx0 = qx[sv_(-N)];   // vz, vx of row z at the old level
        const float vz1 = vz0 + dt * r_rz[c] * (szx_x + szz_z);
        const float vx1 = vx0 + dt * r_rx[c] * (sxx_x + szx_z);
        st_at(ovz, svo, vz1);
        st_at(ovx, svo, vx1);
        float uz = 0.f, ux = 0.f;
        if constexpr (adj) {
            uz = r_uz[c] + dt * vz0;
            ux = r_ux[c] + dt * vx0;
            st_at(fld(a.uz), svo, uz);
            st_at(fld(a.ux), svo, ux);
        }
        if constexpr (energy) st_at(fld(a.We), svo, r_we[c] + (vx1 * vx1 + vz1 * vz1));
        if ((GRAD || store) && z == zl && zl <= a.snap_z2) {   // a lattice row (wave-uniform)
            const float c_zz = nzz[ss_(-N)], c_xx = sxxc;
            if (store && gcol) {   // gard_fwd_storage2; k_inject rewrites source cells after injection
                const size_t o = soff + (size_t)s * a.st_stride + (size_t)jz * a.snx;
                st_at(a.st_vz + o, j4, vz1 / dt);
                st_at(a.st_vx + o, j4, vx1 / dt);
                st_at(a.st_szz + o, j4, c_zz);
                st_at(a.st_szx + o, j4, szxc);
                st_at(a.st_sxx + o, j4, c_xx);
            }
            if (GRAD && gcol) {
                const float gl = r_gl[c], gm = r_gm[c];
                const float Szz = r_S[0][c], Sxx = r_S[1][c], Szx = r_S[2][c], Vz = r_S[3][c], Vx = r_S[4][c];
                const float lm = gl + gm;
                const float il = a.gain_lm / (lm * lm), im = a.gain_lm / (gm * gm);
                const float s1 = 0.25f * il * (Szz + Sxx) * (c_zz + c_xx);
                const float s2 = 0.25f * im * (Szz - Sxx) * (c_zz - c_xx);
                const float s3 = im * (Szx * szxc);
                const size_t q = (size_t)s * slab + (size_t)jz * a.snx;
                st_at(a.g_lam + q, j4, r_g[0][c] + a.gw * s1);
                if (!((MODE & kVelMisfit) != 0 && a.gform)) {
                    st_at(a.g_mu + q, j4, r_g[1][c] + a.gw * (s3 + s1 + s2));
                    st_at(a.g_rho + q, j4, r_g[2][c] + a.gw * (ux * Vx + uz * Vz));
                } else {
                    st_at(a.g_mu + q, j4, r_g[1][c] + a.gw * (s1 + s2));
                    st_at(a.g_rho + q, j4, r_g[2][c] + a.gw * (dt * (vz1 * Vz)));
                    st_at(a.g_ext + q, j4, r_g[3][c] + a.gw * s3);
                    st_at(a.g_ext + a.g_ext_stride + q, j4, r_g[4][c] + a.gw * (dt * (vx1 * Vx)));
                }
            }
            zl += a.snap_dz;
            ++jz;
        }
    };

    // the first input row's queues (vz rows j - N .. j + N - 1, vx rows j - N + 1 .. j + N
    // of j = z0 - N) and stress inputs
    const int j0 = z0 - N;
    static_for<2 * N>([&](auto k) {
        qz[((k + 0) % R)] = ld_at(vzb, orow(j0 - N + k));
        qx[((k + 1) % R)] = ld_at(vxb, orow(j0 - N + 1 + k));
    });
    load_stress_in(std::integral_constant<int, 0>{}, orow(j0));
    // 2N stress-only rows fill the stress rings; the last of them loads the first
    // velocity row's inputs
    const uint32_t vo0 = orow(j0);
    static_for<2 * N>([&](auto i) {
        row(i, std::false_type{}, std::integral_constant<bool, i == 2 * N - 1>{}, j0 + i, vo0 + (uint32_t)i * P4);
    });
    // rows j = z0 + N .. z1 + N - 1: stress and velocity, in blocks of R rows
    const int nrow = z1 - z0, nfull = nrow / R, rem = nrow - nfull * R;
    int jb = z0 + N;
    uint32_t vb = vo0 + (uint32_t)(2 * N) * P4;
    for (int b = 0; b < nfull; ++b, jb += R, vb += (uint32_t)R * P4)
        static_for<R>([&](auto p) {
            row(std::integral_constant<int, 2 * N + p>{}, std::true_type{}, std::true_type{}, jb + p,
                vb + (uint32_t)p * P4);
        });
    if (rem > 0)   // the last segment of the grid only
        static_for<R>([&](auto p) {
            if (p < rem)
                row(std::integral_constant<int, 2 * N + p>{}, std::true_type{}, std::true_type{}, jb + p,
                    vb + (uint32_t)p * P4);
        });
}

// One wavefront's rows, with the correlation switch of kVelAdjoint launches taken once.
template <int N, int MODE>
__device__ __forceinline__ void step7_run(const KArgs &a, const ZTab &zt, const StepView &sv, int s, int tx, int z0,
                                          int z1) {
    if constexpr ((MODE & kVelAdjoint) != 0) {
        if (sv.grad) {
            step7_rows<N, MODE, true>(a, zt, sv, s, tx, z0, z1);
            return;
        }
    }
    step7_rows<N, MODE, false>(a, zt, sv, s, tx, z0, z1);
}

// Tile mode (rows == 0): a block is kBY segments of TZ = (2N + 2) M rows of one column tile,
// or (XW) kBY adjacent column tiles of one segment.  Balanced mode (rows > 0, XW only): the
// launch holds as many blocks as the GPU keeps resident, and block k takes rows
// [k rows, (k + 1) rows) of the sequence of every interior row of every block column
// (kBY tiles wide) of every shot, x fastest, then shots; so every block gets the same
// number of rows, in at most two pieces, and no partly filled last wave of blocks is left.
// KArgs stays the first kernel argument: the CPML blocks read its pointers back at segment
// offset 0.
template <int N, int M, int MODE, bool XW>
__global__ __launch_bounds__(kBX *kBY) void k_step7(const KArgs a, const ZTab zt, const int rows) {
    constexpr int TZ = (2 * N + 2) * M, OUTW = kBX - 4 * N;
    const StepView sv = step_view(a, 0);
    const int wy = wave_uniform((int)threadIdx.y);
    if (XW && rows > 0) {
        const int gx = ((a.nx - 2 * N + OUTW - 1) / OUTW + kBY - 1) / kBY;
        const int nzi = a.nz - 2 * N;
        const int n = gridDim.x, L = blockIdx.x, xcd = L & 7, loc = L >> 3, q = n >> 3, r = n & 7;
        const long long W = xcd * q + min(xcd, r) + loc;   // each XCD a contiguous range of blocks
        const long long total = (long long)a.B * gx * nzi;
        long long r0 = W * rows;
        const long long r1 = min(r0 + rows, total);
        while (r0 < r1) {
            const int u = (int)(r0 / nzi), zs = (int)(r0 - (long long)u * nzi);
            const int ze = (int)min((long long)nzi, r1 - (long long)u * nzi);
            const int s = u / gx, bx = u - s * gx;
            step7_run<N, MODE>(a, zt, sv, s, bx * kBY + wy, N + zs, N + ze);
            r0 = (long long)u * nzi + ze;
        }
        return;
    }
    int tbx, tby, s;
    tile_of<TZ, OUTW, XW>(a, N, tbx, tby, s);
    const int tx = XW ? tbx * kBY + wy : tbx;   // the wavefront's column tile
    const int z0 = N + (XW ? tby : tby * kBY + wy) * TZ;
    step7_run<N, MODE>(a, zt, sv, s, tx, z0, min(z0 + TZ, a.nz - N));
}

int compute_units() {
    static int cus = 0;
    if (!cus) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
            cus = v;
        if (cus <= 0) cus = 256;
    }
    return cus;
}

// Blocks of an instance the GPU keeps resident at once (balanced mode's launch size),
// from its register and LDS use; asked once per instance.
template <int N, int M, int MODE, bool XW>
int resident_blocks() {
    static int per_cu = 0;
    if (!per_cu) {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_step7<N, M, MODE, XW>, kBX * kBY, 0) != hipSuccess ||
            v <= 0)
            v = 1;
        per_cu = v;
    }
    return compute_units() * per_cu;
}

template <int N, int M, int MODE, bool XW>
void step7_go(const KArgs &a, const ZTab &zt, dim3 grid, int mult, hipStream_t st) {
    auto *k = k_step7<N, M, MODE, XW>;
    int rows = 0;
    if (XW && mult > 0) {   // balanced: mult x the resident blocks, the rows shared evenly
        constexpr int OUTW = kBX - 4 * N;
        const int gx = ((a.nx - 2 * N + OUTW - 1) / OUTW + kBY - 1) / kBY;
        const long long total = (long long)a.B * gx * (a.nz - 2 * N);
        const long long nblk = std::min<long long>((long long)mult * resident_blocks<N, M, MODE, XW>(), total);
        rows = (int)((total + nblk - 1) / nblk);
        grid = dim3((unsigned)((total + rows - 1) / rows));
    }
    hipLaunchKernelGGL(k, grid, dim3(kBX, kBY), 0, st, a, zt, rows);
}

template <int N, int M, bool XW>
bool step7_m(const KArgs &a, const ZTab &zt, dim3 grid, int mult, hipStream_t st) {
    switch (vel_mode(a)) {
        case 0: step7_go<N, M, 0, XW>(a, zt, grid, mult, st); return true;
        case kVelEnergy: step7_go<N, M, kVelEnergy, XW>(a, zt, grid, mult, st); return true;
        case kVelStore: step7_go<N, M, kVelStore, XW>(a, zt, grid, mult, st); return true;
        case kVelEnergy | kVelStore: step7_go<N, M, kVelEnergy | kVelStore, XW>(a, zt, grid, mult, st); return true;
        case kVelEnergy | kVelAdjoint: step7_go<N, M, kVelEnergy | kVelAdjoint, XW>(a, zt, grid, mult, st); return true;
        case kVelEnergy | kVelAdjoint | kVelMisfit:
            step7_go<N, M, kVelEnergy | kVelAdjoint | kVelMisfit, XW>(a, zt, grid, mult, st);
            return true;
        default: return false;
    }
}

// Segments of four or eight blocks: every segment recomputes 2N halo rows of stresses,
// so one block (2N + 2 rows) alone would spend almost half its stress work on halos.
template <int N>
bool step7_n(const KArgs &a, const ZTab &zt, int M, bool xw, dim3 grid, int mult, hipStream_t st) {
    if (xw)
        return M == 8 ? step7_m<N, 8, true>(a, zt, grid, mult, st) : step7_m<N, 4, true>(a, zt, grid, mult, st);
    return M == 8 ? step7_m<N, 8, false>(a, zt, grid, mult, st) : step7_m<N, 4, false>(a, zt, grid, mult, st);
}

}  // namespace
}  // namespace psv
