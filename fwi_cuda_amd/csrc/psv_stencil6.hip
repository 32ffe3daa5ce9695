// psv_stencil6.hip -- stencil variant 6: every load of a row issued D rows ahead.
//
// The stress and velocity half-steps of variant 5 (the same arithmetic, the
// same wave-uniform addressing, strip profiles from the kernel arguments and
// halo stores into the pad column), restructured so that a wavefront never
// waits on a load it issued in the same row.  While row z is computed, the
// loads of row z + D are in flight: the z-queue rows it brings in and every
// per-row input (fields to update, media, displacement, energy, snapshot and
// gradient sums).  With a memory latency Lt and a row's arithmetic C, a
// wavefront then advances a row every (Lt + C) / (D + 1) instead of every
// Lt + C, and each wavefront keeps D rows of loads in flight.
//
// The registers are rings with compile-time slots: a z queue holds the rows
// z0 - N + k at slot k mod R, the per-row inputs of row z0 + k sit at slot
// k mod (D + 1), and the row loop is unrolled over a block of R rows (R a
// multiple of D + 1, at least 2N + D), so no value is ever copied between
// registers and no load is waited on before its first use.  A segment is
// M blocks (TZ = M R rows); a shorter last segment of the grid runs its
// remainder rows through a guarded copy of the block.
//
//   k_stress6    vdiff2 + pml_diff2 + update_s2
//                (src/cpu/n_kernel_lib_PSV.cpp:100-131, 134-178, 181-205)
//   k_velocity6  sdiff2 + pml_diff2 + update_v2 (:207-270), fwi_grad2 (:439-484)
//                on correlating steps, gard_fwd_storage2 (:398-436) on storing steps
#include "psv_device.h"

namespace psv {
namespace {

template <int N, int D, int M>
__global__ __launch_bounds__(kBX *kBY) void k_stress6(const KArgs a, const ZTab zt) {
    constexpr int R = v6_block(N, D), L = D + 1, TZ = R * M, OUTW = kBX - 2 * N;
    int tbx, tby, s;
    tile_of<TZ, OUTW>(a, N, tbx, tby, s);
    const int lane = threadIdx.x;
    const int z0 = N + (tby * kBY + wave_uniform((int)threadIdx.y)) * TZ;
    const int nzi = a.nz - N;
    if (z0 >= nzi) return;
    const int zend = min(z0 + TZ, nzi);
    const int ix = tbx * OUTW + lane;
    const bool out = lane >= N && lane < kBX - N && ix < a.nx - N;
    const int cx = min(ix, a.nx - 1);   // halo lanes past the grid read a valid column
    const uint32_t c4 = (uint32_t)cx * 4u;
    const uint32_t P4 = (uint32_t)a.P * 4u;
    const size_t P = a.P, fs = (size_t)s * a.fs;
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

    // z queues: qz[k % R] = vz[z0 - N + k], qx[k % R] = vx[z0 - N + 1 + k]; rows past
    // the segment's last need (vz row zend + N - 2) reload that row
    const float *vzb = fld(a.vz, -N), *vxb = fld(a.vx, 1 - N);
    const uint32_t qlast = (uint32_t)(zend - z0 + 2 * N - 2) * P4 + c4;
    float qz[R], qx[R];
    // per-row inputs of row z0 + k at slot k % L; rows past zend - 1 reload that row
    float *szz = fld(a.szz, 0), *szx = fld(a.szx, 0), *sxx = fld(a.sxx, 0);
    const float *lam = med(a.lam, 0), *mu = med(a.mu, 0), *muzx = med(a.muzx, 0);
    const uint32_t vlast = (uint32_t)(zend - 1 - z0) * P4 + c4;
    float r_zz[L], r_zx[L], r_xx[L], r_l[L], r_m[L], r_mzx[L];
    auto load_row = [&](auto slot, uint32_t o) {   // o: byte offset of the row from row z0
        constexpr int k = decltype(slot)::value;
        o = min(o, vlast);
        r_zz[k] = ld_at(szz, o); r_zx[k] = ld_at(szx, o); r_xx[k] = ld_at(sxx, o);
        r_l[k] = ld_at(lam, o); r_m[k] = ld_at(mu, o); r_mzx[k] = ld_at(muzx, o);
    };
    static_for<2 * N + D - 1>([&](auto k) {
        const uint32_t o = min((uint32_t)k * P4 + c4, qlast);
        qz[k] = ld_at(vzb, o);
        qx[k] = ld_at(vxb, o);
    });
    static_for<D>([&](auto k) { load_row(k, (uint32_t)k * P4 + c4); });

    // row z = z0 + (block) + p, at byte offset vo from row z0
    auto row = [&](auto phase, int z, uint32_t vo) {
        constexpr int p = decltype(phase)::value;
        // loads of row z + D: its new queue rows, then its per-row inputs
        constexpr int qn = (p + 2 * N + D - 1) % R;
        const uint32_t oq = min(vo + (uint32_t)(2 * N + D - 1) * P4, qlast);
        qz[qn] = ld_at(vzb, oq);
        qx[qn] = ld_at(vxb, oq);
        load_row(std::integral_constant<int, (p + D) % L>{}, vo + (uint32_t)D * P4);
        // row z: its queue rows z - N .. z + N - 1 at slots (p + j) % R
        constexpr int c = p % L;
        const float vzc = qz[(p + N) % R], vxc = qx[(p + N - 1) % R];
        float vz_z = 0.f, vx_z = 0.f, vz_x = 0.f, vx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            vz_z += a.cz[k - 1] * (qz[(p + N + k - 1) % R] - qz[(p + N - k) % R]);
            vx_z += a.cz[k - 1] * (qx[(p + N + k - 1) % R] - qx[(p + N - k) % R]);
            vz_x += a.cx[k - 1] * (lane_off(vzc, lane, k) - lane_off(vzc, lane, 1 - k));
            vx_x += a.cx[k - 1] * (lane_off(vxc, lane, k - 1) - lane_off(vxc, lane, -k));
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
        const float div = r_l[c] * (vx_x + vz_z);
        const uint32_t so = vo + dtr;
        st_at(szz, so, r_zz[c] + dt * (div + 2.f * r_m[c] * vz_z));
        st_at(szx, so, r_zx[c] + dt * r_mzx[c] * (vz_x + vx_z));
        st_at(sxx, so, r_xx[c] + dt * (div + 2.f * r_m[c] * vx_x));
    };
    const int nrow = zend - z0, nfull = nrow / R, rem = nrow - nfull * R;
    int zb = z0;
    uint32_t vb = c4;
    for (int b = 0; b < nfull; ++b, zb += R, vb += (uint32_t)R * P4)
        static_for<R>([&](auto p) { row(p, zb + p, vb + (uint32_t)p * P4); });
    if (rem > 0)   // the last segment of the grid only
        static_for<R>([&](auto p) {
            if (p < rem) row(p, zb + p, vb + (uint32_t)p * P4);
        });
}

// The velocity rows of one wavefront; GRAD: this step correlates with the snapshot slab.
template <int N, int D, int M, int MODE, bool GRAD>
__device__ __forceinline__ void velocity6_rows(const KArgs &a, const ZTab &zt, const StepView &sv) {
    constexpr int R = v6_block(N, D), L = D + 1, TZ = R * M, OUTW = kBX - 2 * N;
    constexpr bool displ = (MODE & kVelAdjoint) != 0;   // u += dt v
    constexpr bool energy = (MODE & kVelEnergy) != 0;   // We += v^2
    int tbx, tby, s;
    tile_of<TZ, OUTW>(a, N, tbx, tby, s);
    const int lane = threadIdx.x;
    const int z0 = N + (tby * kBY + wave_uniform((int)threadIdx.y)) * TZ;
    const int nzi = a.nz - N;
    if (z0 >= nzi) return;
    const int zend = min(z0 + TZ, nzi);
    const int ix = tbx * OUTW + lane;
    const bool out = lane >= N && lane < kBX - N && ix < a.nx - N;
    const int cx = min(ix, a.nx - 1);
    const uint32_t c4 = (uint32_t)cx * 4u;
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
    const uint32_t dtr = out ? 0u : (uint32_t)(a.P - 1) * 4u - c4;

    // snapshot lattice: the step's store switch, the lane's column, the next lattice
    // row of the computed rows (zl, jz) and of the loaded rows (zl_l, jz_l)
    const bool store = (MODE & kVelStore) && sv.store;
    const bool gcol = out && (GRAD || store) && ix_on_lattice(a, ix);
    const uint32_t j4 = gcol ? (uint32_t)((ix - a.snap_x1) / a.snap_dx) * 4u : 0u;
    int zl = a.snap_z1, jz = 0;
    if (z0 > zl) {
        jz = (z0 - zl + a.snap_dz - 1) / a.snap_dz;
        zl += jz * a.snap_dz;
    }
    int zl_l = zl, jz_l = jz;
    const size_t slab = (size_t)a.snz * a.snx;
    const size_t soff = (size_t)sv.off;

    // z queues: qzz[k % R] = szz[z0 - N + 1 + k], qzx[k % R] = szx[z0 - N + k]
    const float *szzb = fld(a.szz, 1 - N), *szxb = fld(a.szx, -N);
    const uint32_t qlast = (uint32_t)(zend - z0 + 2 * N - 2) * P4 + c4;
    float qzz[R], qzx[R];
    // per-row inputs of row z0 + k at slot k % L
    const float *sxx = fld(a.sxx, 0);
    float *vz = fld(a.vz, 0), *vx = fld(a.vx, 0);
    const float *rhoz = med(a.rhoz, 0), *rhox = med(a.rhox, 0);
    const uint32_t vlast = (uint32_t)(zend - 1 - z0) * P4 + c4;
    float r_xx[L], r_vz[L], r_vx[L], r_rz[L], r_rx[L];
    float r_uz[L] = {}, r_ux[L] = {}, r_we[L] = {};
    float r_S[5][L] = {}, r_gl[L] = {}, r_gm[L] = {}, r_g[5][L] = {};
    // loads of row r (byte offset o from row z0) into slot k
    auto load_row = [&](auto slot, int r, uint32_t o) {
        constexpr int k = decltype(slot)::value;
        o = min(o, vlast);
        r_xx[k] = ld_at(sxx, o);
        r_vz[k] = ld_at(vz, o); r_vx[k] = ld_at(vx, o); r_rz[k] = ld_at(rhoz, o); r_rx[k] = ld_at(rhox, o);
        if constexpr (displ) {
            r_uz[k] = ld_at(fld(a.uz, 0), o);
            r_ux[k] = ld_at(fld(a.ux, 0), o);
        }
        if constexpr (energy) r_we[k] = ld_at(fld(a.We, 0), o);
        if constexpr (GRAD) {   // read on every row: a clamped lattice row off the lattice
            const int jq = min(jz_l, a.snz - 1);
            const size_t so = soff + (size_t)s * a.sn_stride + (size_t)jq * a.snx;
            const size_t q = (size_t)s * slab + (size_t)jq * a.snx;
            r_S[0][k] = ld_at(a.sn_szz + so, j4); r_S[1][k] = ld_at(a.sn_sxx + so, j4);
            r_S[2][k] = ld_at(a.sn_szx + so, j4); r_S[3][k] = ld_at(a.sn_vz + so, j4);
            r_S[4][k] = ld_at(a.sn_vx + so, j4);
            r_gl[k] = ld_at(med(a.lam, 0), o);
            r_gm[k] = ld_at(med(a.mu, 0), o);
            r_g[0][k] = ld_at(a.g_lam + q, j4); r_g[1][k] = ld_at(a.g_mu + q, j4);
            r_g[2][k] = ld_at(a.g_rho + q, j4);
            if constexpr ((MODE & kVelMisfit) != 0) {
                r_g[3][k] = ld_at(a.g_ext + q, j4);
                r_g[4][k] = ld_at(a.g_ext + a.g_ext_stride + q, j4);
            }
            if (r == zl_l && zl_l <= a.snap_z2) {
                zl_l += a.snap_dz;
                ++jz_l;
            }
        }
    };
    static_for<2 * N + D - 1>([&](auto k) {
        const uint32_t o = min((uint32_t)k * P4 + c4, qlast);
        qzz[k] = ld_at(szzb, o);
        qzx[k] = ld_at(szxb, o);
    });
    static_for<D>([&](auto k) { load_row(k, z0 + k, (uint32_t)k * P4 + c4); });

    auto row = [&](auto phase, int z, uint32_t vo) {
        constexpr int p = decltype(phase)::value;
        constexpr int qn = (p + 2 * N + D - 1) % R;
        const uint32_t oq = min(vo + (uint32_t)(2 * N + D - 1) * P4, qlast);
        qzz[qn] = ld_at(szzb, oq);
        qzx[qn] = ld_at(szxb, oq);
        load_row(std::integral_constant<int, (p + D) % L>{}, z + D, vo + (uint32_t)D * P4);
        constexpr int c = p % L;
        const float sxxc = r_xx[c], szxc = qzx[(p + N) % R];
        float szz_z = 0.f, szx_z = 0.f, szx_x = 0.f, sxx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            szz_z += a.cz[k - 1] * (qzz[(p + N + k - 1) % R] - qzz[(p + N - k) % R]);
            szx_z += a.cz[k - 1] * (qzx[(p + N + k - 1) % R] - qzx[(p + N - k) % R]);
            szx_x += a.cx[k - 1] * (lane_off(szxc, lane, k - 1) - lane_off(szxc, lane, -k));
            sxx_x += a.cx[k - 1] * (lane_off(sxxc, lane, k) - lane_off(sxxc, lane, 1 - k));
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
        const float vz0 = r_vz[c], vx0 = r_vx[c];
        const float vz1 = vz0 + dt * r_rz[c] * (szx_x + szz_z);
        const float vx1 = vx0 + dt * r_rx[c] * (sxx_x + szx_z);
        const uint32_t so = vo + dtr;
        st_at(vz, so, vz1);
        st_at(vx, so, vx1);
        float uz = 0.f, ux = 0.f;
        if constexpr (displ) {
            uz = r_uz[c] + dt * vz0;
            ux = r_ux[c] + dt * vx0;
            st_at(fld(a.uz, 0), so, uz);
            st_at(fld(a.ux, 0), so, ux);
        }
        if constexpr (energy) st_at(fld(a.We, 0), so, r_we[c] + (vx1 * vx1 + vz1 * vz1));
        if ((GRAD || store) && z == zl && zl <= a.snap_z2) {   // a lattice row (wave-uniform)
            const float c_zz = qzz[(p + N - 1) % R], c_xx = sxxc;
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
    const int nrow = zend - z0, nfull = nrow / R, rem = nrow - nfull * R;
    int zb = z0;
    uint32_t vb = c4;
    for (int b = 0; b < nfull; ++b, zb += R, vb += (uint32_t)R * P4)
        static_for<R>([&](auto p) { row(p, zb + p, vb + (uint32_t)p * P4); });
    if (rem > 0)
        static_for<R>([&](auto p) {
            if (p < rem) row(p, zb + p, vb + (uint32_t)p * P4);
        });
}

// kVelAdjoint launches branch once, on the step's correlation switch
template <int N, int D, int M, int MODE>
__global__ __launch_bounds__(kBX *kBY) void k_velocity6(const KArgs a, const ZTab zt) {
    const StepView sv = step_view(a, 0);
    if constexpr ((MODE & kVelAdjoint) != 0) {
        if (sv.grad) {
            velocity6_rows<N, D, M, MODE, true>(a, zt, sv);
            return;
        }
    }
    velocity6_rows<N, D, M, MODE, false>(a, zt, sv);
}

template <int N, int D, int M>
void stress6_m(const KArgs &a, const ZTab &zt, dim3 grid, dim3 block, hipStream_t st) {
    hipLaunchKernelGGL((k_stress6<N, D, M>), grid, block, 0, st, a, zt);
}

// the launch modes the passes run (forward, recomputation, either storing
// snapshots, adjoint, adjoint for the misfit derivative); false for any other
template <int N, int D, int M>
bool velocity6_m(const KArgs &a, const ZTab &zt, dim3 grid, dim3 block, hipStream_t st) {
    switch (vel_mode(a)) {
        case 0: hipLaunchKernelGGL((k_velocity6<N, D, M, 0>), grid, block, 0, st, a, zt); return true;
        case kVelEnergy: hipLaunchKernelGGL((k_velocity6<N, D, M, kVelEnergy>), grid, block, 0, st, a, zt); return true;
        case kVelStore: hipLaunchKernelGGL((k_velocity6<N, D, M, kVelStore>), grid, block, 0, st, a, zt); return true;
        case kVelEnergy | kVelStore:
            hipLaunchKernelGGL((k_velocity6<N, D, M, kVelEnergy | kVelStore>), grid, block, 0, st, a, zt);
            return true;
        case kVelEnergy | kVelAdjoint:
            hipLaunchKernelGGL((k_velocity6<N, D, M, kVelEnergy | kVelAdjoint>), grid, block, 0, st, a, zt);
            return true;
        case kVelEnergy | kVelAdjoint | kVelMisfit:
            hipLaunchKernelGGL((k_velocity6<N, D, M, kVelEnergy | kVelAdjoint | kVelMisfit>), grid, block, 0, st, a,
                               zt);
            return true;
        default: return false;
    }
}

// half-step `which` (0 stress, 1 velocity) at order 2N, depth D, M blocks per segment
template <int N, int D, int M>
bool launch6_ndm(int which, const KArgs &a, const ZTab &zt, dim3 grid, dim3 block, hipStream_t st) {
    if (which != 0) return velocity6_m<N, D, M>(a, zt, grid, block, st);
    stress6_m<N, D, M>(a, zt, grid, block, st);
    return true;
}

// Only depth 1 is instantiated: at order 8 the adjoint launch of depth 2 needs more
// than 256 VGPRs of rings and spills.
template <int N>
bool launch6_n(int which, const KArgs &a, const ZTab &zt, int M, dim3 grid, dim3 block, hipStream_t st) {
    return M == 2 ? launch6_ndm<N, 1, 2>(which, a, zt, grid, block, st)
                  : launch6_ndm<N, 1, 1>(which, a, zt, grid, block, st);
}

}  // namespace

int v6_rows(int N, int depth, int tz) {
    (void)depth;
    const int R = v6_block(N, 1);
    return tz >= 2 * R ? 2 * R : R;
}

bool launch6(int which, const KArgs &a, const ZTab &zt, int N, int depth, int tz, hipStream_t st) {
    (void)depth;
    const int D = 1, R = v6_block(N, D);
    const int M = tz >= 2 * R ? 2 : 1;
    const int nxi = a.nx - 2 * N, nzi = a.nz - 2 * N, outw = kBX - 2 * N, TZ = R * M;
    const int segs = (nzi + TZ - 1) / TZ;
    const dim3 grid(((nxi + outw - 1) / outw) * ((segs + kBY - 1) / kBY) * a.B), block(kBX, kBY);
    switch (N) {
        case 1: return launch6_n<1>(which, a, zt, M, grid, block, st);
        case 2: return launch6_n<2>(which, a, zt, M, grid, block, st);
        case 3: return launch6_n<3>(which, a, zt, M, grid, block, st);
        case 4: return launch6_n<4>(which, a, zt, M, grid, block, st);
        case 5: return launch6_n<5>(which, a, zt, M, grid, block, st);
        default: return launch6_n<6>(which, a, zt, M, grid, block, st);
    }
}

}  // namespace psv
