// psv_kernels.hip -- gfx950 time-step kernels of the elastic P/SV path.
//
// One time step is two fused half-step kernels plus O(nx + nrec) epilogues:
//
//   k_stress    vdiff2 + pml_diff2 (velocity pass) + update_s2
//               (src/cpu/n_kernel_lib_PSV.cpp:100-131, 134-178, 181-205)
//   k_surface   surf_mirror (:272-395), one launch per side in the reference's order
//   k_velocity  sdiff2 + pml_diff2 (stress pass) + update_v2 (:207-270), and on
//               snapshot steps of the adjoint the fwi_grad2 correlation (:439-484)
//   k_inject    vsrc2 (:487-536)       k_record  urec2 (:538-576)
//   k_store     gard_fwd_storage2 (:398-436)
//
// Stencil kernels: a thread owns one column ix and marches down TZ rows.  The
// z-derivative taps live in a register queue of 2N values per field, so each
// row of vz, vx (or szz, szx) is fetched once per segment; the x taps are
// neighbouring lanes of the same wavefront (one 256-byte row segment per load,
// served by L1/L2).  CPML memory variables exist only in the active strips
// (zmap / xmap >= 0); elsewhere the CPML update is the identity, as it is in
// the reference where a = 0, b = 1, K = 1 outside the strips.  Variant 5 and its
// full-tile forms 10 to 13 live in psv_stencil5.h (compiled in psv_v5.hip, psv_v11.hip
// and psv_v13.hip).
#include "psv_device.h"

namespace psv {

namespace {

__device__ __forceinline__ size_t fidx(const KArgs &a, int s, int z, int x) {
    return (size_t)s * a.fs + (size_t)z * a.P + x;
}

// Gradient sums at snapshot cell q (offset in the batch's [B][snz][snx] arrays).
// s1, s2, s3: the lam+mu, mu and shear correlations of fwi_grad2
// (src/cpu/n_kernel_lib_PSV.cpp:468-479); s4 = u'.V its density term.  With
// gform 1 the density terms are the adjoint velocity times the forward velocity
// per component (rz, rx) and s3 is kept apart, for the host's chain rule.
__device__ __forceinline__ void grad_accumulate(const KArgs &a, size_t q, float s1, float s2, float s3, float s4,
                                                float rz, float rx) {
    if (!a.gform) {
        a.g_lam[q] += a.gw * s1;
        a.g_mu[q] += a.gw * (s3 + s1 + s2);
        a.g_rho[q] += a.gw * s4;
    } else {
        a.g_lam[q] += a.gw * s1;
        a.g_mu[q] += a.gw * (s1 + s2);
        a.g_rho[q] += a.gw * rz;
        a.g_ext[q] += a.gw * s3;
        a.g_ext[a.g_ext_stride + q] += a.gw * rx;
    }
}


template <int N, int TZ>
__global__ __launch_bounds__(kBX *kBY) void k_stress(const KArgs a) {
    int tbx, tby, s;
    tile_of<TZ>(a, N, tbx, tby, s);
    const int ix = N + tbx * kBX + threadIdx.x;
    const int z0 = N + (tby * kBY + threadIdx.y) * TZ;
    const int nzi = a.nz - N;
    if (ix >= a.nx - N || z0 >= nzi) return;
    const int zend = min(z0 + TZ, nzi);
    const int P = a.P;
    const size_t fo = (size_t)s * a.fs;
    const float *__restrict__ vz = a.vz + fo;
    const float *__restrict__ vx = a.vx + fo;
    float *__restrict__ szz = a.szz + fo;
    float *__restrict__ szx = a.szx + fo;
    float *__restrict__ sxx = a.sxx + fo;
    const float dt = a.dt;

    const int cxm = a.xmap[ix];
    float bx = 1.f, ax = 0.f, ikx = 1.f, bhx = 1.f, ahx = 0.f, ikhx = 1.f;
    if (cxm >= 0) {
        bx = a.bx[ix]; ax = a.ax[ix]; ikx = a.ikx[ix];
        bhx = a.bhx[ix]; ahx = a.ahx[ix]; ikhx = a.ikhx[ix];
    }

    // qz[j] = vz[z - N + j], qx[j] = vx[z - N + 1 + j] for the current row z
    float qz[2 * N], qx[2 * N];
#pragma unroll
    for (int j = 0; j < 2 * N - 1; ++j) {
        qz[j + 1] = vz[(size_t)(z0 - N + j) * P + ix];
        qx[j + 1] = vx[(size_t)(z0 - N + 1 + j) * P + ix];
    }
    for (int z = z0; z < zend; ++z) {
#pragma unroll
        for (int j = 0; j < 2 * N - 1; ++j) {
            qz[j] = qz[j + 1];
            qx[j] = qx[j + 1];
        }
        qz[2 * N - 1] = vz[(size_t)(z + N - 1) * P + ix];
        qx[2 * N - 1] = vx[(size_t)(z + N) * P + ix];
        const size_t r = (size_t)z * P + ix;
        float vz_z = 0.f, vx_z = 0.f, vz_x = 0.f, vx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            vz_z += a.cz[k - 1] * (qz[N + k - 1] - qz[N - k]);
            vx_z += a.cz[k - 1] * (qx[N + k - 1] - qx[N - k]);
            vz_x += a.cx[k - 1] * (vz[r + k] - vz[r - k + 1]);
            vx_x += a.cx[k - 1] * (vx[r + k - 1] - vx[r - k]);
        }
        const int rz = a.zmap[z];
        if (rz >= 0) {
            const size_t pi = ((size_t)s * a.nzs + rz) * P + ix;
            const float p1 = a.bz[z] * a.pvz_z[pi] + a.az[z] * vz_z;
            const float p2 = a.bhz[z] * a.pvx_z[pi] + a.ahz[z] * vx_z;
            a.pvz_z[pi] = p1;
            a.pvx_z[pi] = p2;
            vz_z = vz_z * a.ikz[z] + p1;
            vx_z = vx_z * a.ikhz[z] + p2;
        }
        if (cxm >= 0) {
            const size_t pi = ((size_t)s * a.nz + z) * a.nxs + cxm;
            const float p1 = bx * a.pvx_x[pi] + ax * vx_x;
            const float p2 = bhx * a.pvz_x[pi] + ahx * vz_x;
            a.pvx_x[pi] = p1;
            a.pvz_x[pi] = p2;
            vx_x = vx_x * ikx + p1;
            vz_x = vz_x * ikhx + p2;
        }
        const float lam = a.lam[r], mu = a.mu[r], muzx = a.muzx[r];
        const float div = lam * (vx_x + vz_z);
        szz[r] += dt * (div + 2.f * mu * vz_z);
        szx[r] += dt * muzx * (vz_x + vx_z);
        sxx[r] += dt * (div + 2.f * mu * vx_x);
    }
}

template <int N, int TZ>
__global__ __launch_bounds__(kBX *kBY) void k_velocity(const KArgs a) {
    int tbx, tby, s;
    tile_of<TZ>(a, N, tbx, tby, s);
    const int ix = N + tbx * kBX + threadIdx.x;
    const int z0 = N + (tby * kBY + threadIdx.y) * TZ;
    const int nzi = a.nz - N;
    if (ix >= a.nx - N || z0 >= nzi) return;
    const int zend = min(z0 + TZ, nzi);
    const int P = a.P;
    const size_t fo = (size_t)s * a.fs;
    float *__restrict__ vz = a.vz + fo;
    float *__restrict__ vx = a.vx + fo;
    const float *__restrict__ szz = a.szz + fo;
    const float *__restrict__ szx = a.szx + fo;
    const float *__restrict__ sxx = a.sxx + fo;
    const float dt = a.dt;

    const int cxm = a.xmap[ix];
    float bx = 1.f, ax = 0.f, ikx = 1.f, bhx = 1.f, ahx = 0.f, ikhx = 1.f;
    if (cxm >= 0) {
        bx = a.bx[ix]; ax = a.ax[ix]; ikx = a.ikx[ix];
        bhx = a.bhx[ix]; ahx = a.ahx[ix]; ikhx = a.ikhx[ix];
    }
    // snapshot lattice column: gradient correlation (adjoint) or snapshot store (forward)
    const StepView sv = step_view(a, 0);
    const bool gcol = (sv.grad || sv.store) && ix_on_lattice(a, ix);
    const int jx = gcol ? (ix - a.snap_x1) / a.snap_dx : 0;

    // qzz[j] = szz[z - N + 1 + j], qzx[j] = szx[z - N + j]
    float qzz[2 * N], qzx[2 * N];
#pragma unroll
    for (int j = 0; j < 2 * N - 1; ++j) {
        qzz[j + 1] = szz[(size_t)(z0 - N + 1 + j) * P + ix];
        qzx[j + 1] = szx[(size_t)(z0 - N + j) * P + ix];
    }
    for (int z = z0; z < zend; ++z) {
#pragma unroll
        for (int j = 0; j < 2 * N - 1; ++j) {
            qzz[j] = qzz[j + 1];
            qzx[j] = qzx[j + 1];
        }
        qzz[2 * N - 1] = szz[(size_t)(z + N) * P + ix];
        qzx[2 * N - 1] = szx[(size_t)(z + N - 1) * P + ix];
        const size_t r = (size_t)z * P + ix;
        float szz_z = 0.f, szx_z = 0.f, szx_x = 0.f, sxx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            szz_z += a.cz[k - 1] * (qzz[N + k - 1] - qzz[N - k]);
            szx_z += a.cz[k - 1] * (qzx[N + k - 1] - qzx[N - k]);
            szx_x += a.cx[k - 1] * (szx[r + k - 1] - szx[r - k]);
            sxx_x += a.cx[k - 1] * (sxx[r + k] - sxx[r - k + 1]);
        }
        // the reference reuses the velocity pass's slot-to-profile mapping here
        // (src/cpu/n_kernel_PSV.cpp:142-146): szz_z and sxx_x take the integer-point
        // profiles, szx_z and szx_x the half-point ones.
        const int rz = a.zmap[z];
        if (rz >= 0) {
            const size_t pi = ((size_t)s * a.nzs + rz) * P + ix;
            const float p1 = a.bz[z] * a.pszz_z[pi] + a.az[z] * szz_z;
            const float p2 = a.bhz[z] * a.pszx_z[pi] + a.ahz[z] * szx_z;
            a.pszz_z[pi] = p1;
            a.pszx_z[pi] = p2;
            szz_z = szz_z * a.ikz[z] + p1;
            szx_z = szx_z * a.ikhz[z] + p2;
        }
        if (cxm >= 0) {
            const size_t pi = ((size_t)s * a.nz + z) * a.nxs + cxm;
            const float p1 = bx * a.psxx_x[pi] + ax * sxx_x;
            const float p2 = bhx * a.pszx_x[pi] + ahx * szx_x;
            a.psxx_x[pi] = p1;
            a.pszx_x[pi] = p2;
            sxx_x = sxx_x * ikx + p1;
            szx_x = szx_x * ikhx + p2;
        }
        const float vz0 = vz[r], vx0 = vx[r];
        const float vz1 = vz0 + dt * a.rhoz[r] * (szx_x + szz_z);
        const float vx1 = vx0 + dt * a.rhox[r] * (sxx_x + szx_z);
        vz[r] = vz1;
        vx[r] = vx1;
        float uz = 0.f, ux = 0.f;
        if (a.displ) {
            uz = a.uz[fo + r] + dt * vz0;
            ux = a.ux[fo + r] + dt * vx0;
            a.uz[fo + r] = uz;
            a.ux[fo + r] = ux;
        }
        if (a.energy) a.We[fo + r] += vx1 * vx1 + vz1 * vz1;
        if (gcol && z >= a.snap_z1 && z <= a.snap_z2 && (z - a.snap_z1) % a.snap_dz == 0) {
            const int jz = (z - a.snap_z1) / a.snap_dz;
            const float c_zz = qzz[N - 1], c_zx = qzx[N], c_xx = sxx[r];
            if (sv.grad) {
                const size_t q = (size_t)s * a.snz * a.snx + (size_t)jz * a.snx + jx;
                const size_t o = sv.off + (size_t)s * a.sn_stride + (size_t)jz * a.snx + jx;
                const float Szz = a.sn_szz[o], Sxx = a.sn_sxx[o], Szx = a.sn_szx[o];
                const float Vz = a.sn_vz[o], Vx = a.sn_vx[o];
                const float lam = a.lam[r], mu = a.mu[r];
                const float lm = lam + mu;
                // the 1/(lam+mu)^2 and 1/mu^2 factors (~1e-20) carry the gain so the
                // products stay in the FP32 normal range; the host divides it out in FP64
                const float il = a.gain_lm / (lm * lm), im = a.gain_lm / (mu * mu);
                const float s1 = 0.25f * il * (Szz + Sxx) * (c_zz + c_xx);
                const float s2 = 0.25f * im * (Szz - Sxx) * (c_zz - c_xx);
                const float s3 = im * (Szx * c_zx);
                grad_accumulate(a, q, s1, s2, s3, ux * Vx + uz * Vz, dt * (vz1 * Vz), dt * (vx1 * Vx));
            }
            if (sv.store) {   // gard_fwd_storage2; k_inject rewrites source cells after injection
                const size_t o = sv.off + (size_t)s * a.st_stride + (size_t)jz * a.snx + jx;
                a.st_vz[o] = vz1 / dt;
                a.st_vx[o] = vx1 / dt;
                a.st_szz[o] = c_zz;
                a.st_szx[o] = c_zx;
                a.st_sxx[o] = c_xx;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Variant 2: x taps from neighbouring lanes, next row prefetched.
//
// A wavefront covers 64 consecutive columns; its N lanes at each end are halo
// lanes that only load, so OUTW = 64 - 2N columns are written per wavefront.
// The z-queue registers already hold every lane's value of the current row,
// so the x-derivative taps are cross-lane reads of those registers
// (ds_bpermute) instead of 2N shifted global loads per field.  The global
// loads of row z + 1 (queue entries, stresses, media) are issued before row z
// is computed, which doubles the loads in flight per wavefront.
// ---------------------------------------------------------------------------

template <int N, int TZ, int MINW = 1>
__global__ __launch_bounds__(kBX *kBY, MINW) void k_stress2(const KArgs a) {
    constexpr int OUTW = kBX - 2 * N;
    int tbx, tby, s;
    tile_of<TZ, OUTW>(a, N, tbx, tby, s);
    const int lane = threadIdx.x;
    const int ix = tbx * OUTW + lane;   // lane N holds the first output column N + tbx * OUTW
    const int z0 = N + (tby * kBY + threadIdx.y) * TZ;
    const int nzi = a.nz - N;
    if (z0 >= nzi) return;              // uniform over the wavefront
    const int zend = min(z0 + TZ, nzi);
    const bool out = lane >= N && lane < kBX - N && ix < a.nx - N;
    const int cx = min(ix, a.nx - 1);   // halo lanes past the grid read a valid column
    const int P = a.P;
    const size_t fo = (size_t)s * a.fs;
    const float *__restrict__ vz = a.vz + fo;
    const float *__restrict__ vx = a.vx + fo;
    float *__restrict__ szz = a.szz + fo;
    float *__restrict__ szx = a.szx + fo;
    float *__restrict__ sxx = a.sxx + fo;
    const float dt = a.dt;

    const int cxm = out ? a.xmap[ix] : -1;
    float bx = 1.f, ax = 0.f, ikx = 1.f, bhx = 1.f, ahx = 0.f, ikhx = 1.f;
    if (cxm >= 0) {
        bx = a.bx[ix]; ax = a.ax[ix]; ikx = a.ikx[ix];
        bhx = a.bhx[ix]; ahx = a.ahx[ix]; ikhx = a.ikhx[ix];
    }
    // qz[j] = vz[z - N + j], qx[j] = vx[z - N + 1 + j]
    float qz[2 * N], qx[2 * N];
#pragma unroll
    for (int j = 0; j < 2 * N - 1; ++j) {
        qz[j + 1] = vz[(size_t)(z0 - N + j) * P + cx];
        qx[j + 1] = vx[(size_t)(z0 - N + 1 + j) * P + cx];
    }
    // prefetched row
    float nvz = vz[(size_t)(z0 + N - 1) * P + cx], nvx = vx[(size_t)(z0 + N) * P + cx];
    float n_zz = 0.f, n_zx = 0.f, n_xx = 0.f, n_l = 0.f, n_m = 0.f, n_mzx = 0.f;
    if (out) {
        const size_t r = (size_t)z0 * P + ix;
        n_zz = szz[r]; n_zx = szx[r]; n_xx = sxx[r];
        n_l = a.lam[r]; n_m = a.mu[r]; n_mzx = a.muzx[r];
    }
    for (int z = z0; z < zend; ++z) {
#pragma unroll
        for (int j = 0; j < 2 * N - 1; ++j) {
            qz[j] = qz[j + 1];
            qx[j] = qx[j + 1];
        }
        qz[2 * N - 1] = nvz;
        qx[2 * N - 1] = nvx;
        const float c_zz = n_zz, c_zx = n_zx, c_xx = n_xx, lam = n_l, mu = n_m, muzx = n_mzx;
        if (z + 1 < zend) {
            nvz = vz[(size_t)(z + N) * P + cx];
            nvx = vx[(size_t)(z + N + 1) * P + cx];
            if (out) {
                const size_t r1 = (size_t)(z + 1) * P + ix;
                n_zz = szz[r1]; n_zx = szx[r1]; n_xx = sxx[r1];
                n_l = a.lam[r1]; n_m = a.mu[r1]; n_mzx = a.muzx[r1];
            }
        }
        const float vzc = qz[N], vxc = qx[N - 1];   // this lane's vz, vx on row z
        float vz_z = 0.f, vx_z = 0.f, vz_x = 0.f, vx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            vz_z += a.cz[k - 1] * (qz[N + k - 1] - qz[N - k]);
            vx_z += a.cz[k - 1] * (qx[N + k - 1] - qx[N - k]);
            vz_x += a.cx[k - 1] * (lane_off(vzc, lane, k) - lane_off(vzc, lane, 1 - k));
            vx_x += a.cx[k - 1] * (lane_off(vxc, lane, k - 1) - lane_off(vxc, lane, -k));
        }
        if (!out) continue;
        const size_t r = (size_t)z * P + ix;
        const int rz = a.zmap[z];
        if (rz >= 0) {
            const size_t pi = ((size_t)s * a.nzs + rz) * P + ix;
            const float p1 = a.bz[z] * a.pvz_z[pi] + a.az[z] * vz_z;
            const float p2 = a.bhz[z] * a.pvx_z[pi] + a.ahz[z] * vx_z;
            a.pvz_z[pi] = p1;
            a.pvx_z[pi] = p2;
            vz_z = vz_z * a.ikz[z] + p1;
            vx_z = vx_z * a.ikhz[z] + p2;
        }
        if (cxm >= 0) {
            const size_t pi = ((size_t)s * a.nz + z) * a.nxs + cxm;
            const float p1 = bx * a.pvx_x[pi] + ax * vx_x;
            const float p2 = bhx * a.pvz_x[pi] + ahx * vz_x;
            a.pvx_x[pi] = p1;
            a.pvz_x[pi] = p2;
            vx_x = vx_x * ikx + p1;
            vz_x = vz_x * ikhx + p2;
        }
        const float div = lam * (vx_x + vz_z);
        szz[r] = c_zz + dt * (div + 2.f * mu * vz_z);
        szx[r] = c_zx + dt * muzx * (vz_x + vx_z);
        sxx[r] = c_xx + dt * (div + 2.f * mu * vx_x);
    }
}

template <int N, int TZ, int MINW = 1>
__global__ __launch_bounds__(kBX *kBY, MINW) void k_velocity2(const KArgs a) {
    constexpr int OUTW = kBX - 2 * N;
    int tbx, tby, s;
    tile_of<TZ, OUTW>(a, N, tbx, tby, s);
    const int lane = threadIdx.x;
    const int ix = tbx * OUTW + lane;
    const int z0 = N + (tby * kBY + threadIdx.y) * TZ;
    const int nzi = a.nz - N;
    if (z0 >= nzi) return;
    const int zend = min(z0 + TZ, nzi);
    const bool out = lane >= N && lane < kBX - N && ix < a.nx - N;
    const int cx = min(ix, a.nx - 1);
    const int P = a.P;
    const size_t fo = (size_t)s * a.fs;
    float *__restrict__ vz = a.vz + fo;
    float *__restrict__ vx = a.vx + fo;
    const float *__restrict__ szz = a.szz + fo;
    const float *__restrict__ szx = a.szx + fo;
    const float *__restrict__ sxx = a.sxx + fo;
    const float dt = a.dt;

    const int cxm = out ? a.xmap[ix] : -1;
    float bx = 1.f, ax = 0.f, ikx = 1.f, bhx = 1.f, ahx = 0.f, ikhx = 1.f;
    if (cxm >= 0) {
        bx = a.bx[ix]; ax = a.ax[ix]; ikx = a.ikx[ix];
        bhx = a.bhx[ix]; ahx = a.ahx[ix]; ikhx = a.ikhx[ix];
    }
    const StepView sv = step_view(a, 0);
    const bool gcol = out && (sv.grad || sv.store) && ix_on_lattice(a, ix);
    const int jx = gcol ? (ix - a.snap_x1) / a.snap_dx : 0;

    // qzz[j] = szz[z - N + 1 + j], qzx[j] = szx[z - N + j]
    float qzz[2 * N], qzx[2 * N];
#pragma unroll
    for (int j = 0; j < 2 * N - 1; ++j) {
        qzz[j + 1] = szz[(size_t)(z0 - N + 1 + j) * P + cx];
        qzx[j + 1] = szx[(size_t)(z0 - N + j) * P + cx];
    }
    float n_szz = szz[(size_t)(z0 + N) * P + cx], n_szx = szx[(size_t)(z0 + N - 1) * P + cx];
    float n_sxx = sxx[(size_t)z0 * P + cx];
    float n_vz = 0.f, n_vx = 0.f, n_rz = 0.f, n_rx = 0.f;
    if (out) {
        const size_t r = (size_t)z0 * P + ix;
        n_vz = vz[r]; n_vx = vx[r]; n_rz = a.rhoz[r]; n_rx = a.rhox[r];
    }
    for (int z = z0; z < zend; ++z) {
#pragma unroll
        for (int j = 0; j < 2 * N - 1; ++j) {
            qzz[j] = qzz[j + 1];
            qzx[j] = qzx[j + 1];
        }
        qzz[2 * N - 1] = n_szz;
        qzx[2 * N - 1] = n_szx;
        const float sxxc = n_sxx, vz0 = n_vz, vx0 = n_vx, rz_ = n_rz, rx_ = n_rx;
        if (z + 1 < zend) {
            n_szz = szz[(size_t)(z + 1 + N) * P + cx];
            n_szx = szx[(size_t)(z + N) * P + cx];
            n_sxx = sxx[(size_t)(z + 1) * P + cx];
            if (out) {
                const size_t r1 = (size_t)(z + 1) * P + ix;
                n_vz = vz[r1]; n_vx = vx[r1]; n_rz = a.rhoz[r1]; n_rx = a.rhox[r1];
            }
        }
        const float szxc = qzx[N];   // this lane's szx on row z
        float szz_z = 0.f, szx_z = 0.f, szx_x = 0.f, sxx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            szz_z += a.cz[k - 1] * (qzz[N + k - 1] - qzz[N - k]);
            szx_z += a.cz[k - 1] * (qzx[N + k - 1] - qzx[N - k]);
            szx_x += a.cx[k - 1] * (lane_off(szxc, lane, k - 1) - lane_off(szxc, lane, -k));
            sxx_x += a.cx[k - 1] * (lane_off(sxxc, lane, k) - lane_off(sxxc, lane, 1 - k));
        }
        if (!out) continue;
        const size_t r = (size_t)z * P + ix;
        // integer-point profiles for szz_z and sxx_x, half-point for szx_z and szx_x,
        // the reference's slot mapping (src/cpu/n_kernel_PSV.cpp:142-146)
        const int rzm = a.zmap[z];
        if (rzm >= 0) {
            const size_t pi = ((size_t)s * a.nzs + rzm) * P + ix;
            const float p1 = a.bz[z] * a.pszz_z[pi] + a.az[z] * szz_z;
            const float p2 = a.bhz[z] * a.pszx_z[pi] + a.ahz[z] * szx_z;
            a.pszz_z[pi] = p1;
            a.pszx_z[pi] = p2;
            szz_z = szz_z * a.ikz[z] + p1;
            szx_z = szx_z * a.ikhz[z] + p2;
        }
        if (cxm >= 0) {
            const size_t pi = ((size_t)s * a.nz + z) * a.nxs + cxm;
            const float p1 = bx * a.psxx_x[pi] + ax * sxx_x;
            const float p2 = bhx * a.pszx_x[pi] + ahx * szx_x;
            a.psxx_x[pi] = p1;
            a.pszx_x[pi] = p2;
            sxx_x = sxx_x * ikx + p1;
            szx_x = szx_x * ikhx + p2;
        }
        const float vz1 = vz0 + dt * rz_ * (szx_x + szz_z);
        const float vx1 = vx0 + dt * rx_ * (sxx_x + szx_z);
        vz[r] = vz1;
        vx[r] = vx1;
        float uz = 0.f, ux = 0.f;
        if (a.displ) {
            uz = a.uz[fo + r] + dt * vz0;
            ux = a.ux[fo + r] + dt * vx0;
            a.uz[fo + r] = uz;
            a.ux[fo + r] = ux;
        }
        if (a.energy) a.We[fo + r] += vx1 * vx1 + vz1 * vz1;
        if (gcol && z >= a.snap_z1 && z <= a.snap_z2 && (z - a.snap_z1) % a.snap_dz == 0) {
            const int jz = (z - a.snap_z1) / a.snap_dz;
            const float c_zz = qzz[N - 1], c_xx = sxxc;
            if (sv.store) {   // gard_fwd_storage2; k_inject rewrites source cells after injection
                const size_t o = sv.off + (size_t)s * a.st_stride + (size_t)jz * a.snx + jx;
                a.st_vz[o] = vz1 / dt;
                a.st_vx[o] = vx1 / dt;
                a.st_szz[o] = c_zz;
                a.st_szx[o] = szxc;
                a.st_sxx[o] = c_xx;
            }
            if (!sv.grad) continue;
            const size_t q = (size_t)s * a.snz * a.snx + (size_t)jz * a.snx + jx;
            const size_t o = sv.off + (size_t)s * a.sn_stride + (size_t)jz * a.snx + jx;
            const float Szz = a.sn_szz[o], Sxx = a.sn_sxx[o], Szx = a.sn_szx[o];
            const float Vz = a.sn_vz[o], Vx = a.sn_vx[o];
            const float lam = a.lam[r], mu = a.mu[r];
            const float lm = lam + mu;
            const float il = a.gain_lm / (lm * lm), im = a.gain_lm / (mu * mu);
            const float s1 = 0.25f * il * (Szz + Sxx) * (c_zz + c_xx);
            const float s2 = 0.25f * im * (Szz - Sxx) * (c_zz - c_xx);
            const float s3 = im * (Szx * szxc);
            grad_accumulate(a, q, s1, s2, s3, ux * Vx + uz * Vz, dt * (vz1 * Vz), dt * (vx1 * Vx));
        }
    }
}

// ---------------------------------------------------------------------------
// Variant 4: the whole time step in one pass (k_step).
//
// The wavefront streams down its segment of rows [z0, z1).  At input row j it
// computes the new stresses of row j (and applies the top free surface when
// j is that row), pushes them into a stress queue, and then updates the
// velocities of row z = j - N, whose z taps (szz rows z-N+1 .. z+N, szx rows
// z-N .. z+N-1) are that queue.  The x taps of both half-steps are
// neighbouring lanes: a wavefront loads 64 columns, computes stresses on the
// inner 64 - 2N and writes velocities (and stresses) on the inner 64 - 4N.
// Stresses the velocity taps need outside the owned block (N columns each
// side, N rows above and below the segment) are recomputed, so every input is
// read at the previous time level: fields and the stress-side CPML memory are
// read from one buffer set and written to the other (the host swaps them each
// step).  The velocity-side memory is only touched by the owning lane and is
// updated in place.  Compulsory traffic per cell-step: R+W five fields, R five
// media, R+W We (68 B): each array read and written once.
// ---------------------------------------------------------------------------
// MODE bits (compile time, so each launch mode keeps only the pointers it uses):
// 1 = accumulate We, 2 = adjoint (u += dt v, gradient on snapshot steps), 4 = store
// snapshots; kModeGeneral reads the three flags from the arguments at run time.
constexpr int kModeEnergy = 1, kModeAdjoint = 2, kModeStore = 4, kModeGeneral = 8;

// D = prefetch depth: the loads of input row j + D are issued while row j is computed.
template <int N, int TZ, int MODE, int D>
__global__ __launch_bounds__(kBX *kBY) void k_step(const KArgs a) {
    constexpr bool kGeneral = MODE == kModeGeneral;
    const bool kEnergy = kGeneral ? a.energy != 0 : (MODE & kModeEnergy) != 0;
    const bool kAdj = kGeneral ? a.displ != 0 : (MODE & kModeAdjoint) != 0;
    const bool kStore = kGeneral ? a.store != 0 : (MODE & kModeStore) != 0;
    const StepView sv = step_view(a, 0);
    const bool doStore = kStore && sv.store, doGrad = kAdj && sv.grad;
    constexpr int OUTW = kBX - 4 * N;
    int tbx, tby, s;
    tile_of<TZ, OUTW>(a, N, tbx, tby, s);
    const int lane = threadIdx.x;
    const int x = tbx * OUTW + lane - N;   // lane 2N holds the first owned column N + tbx * OUTW
    const int z0 = N + (tby * kBY + threadIdx.y) * TZ;
    const int nzi = a.nz - N;
    if (z0 >= nzi) return;                 // uniform over the wavefront
    const int z1 = min(z0 + TZ, nzi);
    const int cx = min(max(x, 0), a.nx - 1);   // lanes off the grid read a valid column
    // stresses are computed where the lane has its 2N x taps and the cell is interior;
    // lanes [2N, 64 - 2N) own their column
    const bool sint = lane >= N && lane < kBX - N && x >= N && x < a.nx - N;
    const bool own = lane >= 2 * N && lane < kBX - 2 * N && x < a.nx - N;
    const int P = a.P;
    // Per-shot base pointers are wavefront-uniform; element offsets within one
    // shot's arrays are 32-bit (the host keeps nz * P below 2^29).
    const size_t fo = (size_t)s * a.fs;
    const float *__restrict__ vz = a.vz + fo;
    const float *__restrict__ vx = a.vx + fo;
    const float *__restrict__ szz = a.szz + fo;
    const float *__restrict__ szx = a.szx + fo;
    const float *__restrict__ sxx = a.sxx + fo;
    float *__restrict__ ovz = a.o_vz + fo;
    float *__restrict__ ovx = a.o_vx + fo;
    float *__restrict__ oszz = a.o_szz + fo;
    float *__restrict__ oszx = a.o_szx + fo;
    float *__restrict__ osxx = a.o_sxx + fo;
    const size_t pzo = (size_t)s * a.nzs * P, pxo = (size_t)s * a.nz * a.nxs;
    const float dt = a.dt;

    CpmlPoint px{1.f, 0.f, 1.f, 1.f, 0.f, 1.f, -1, 0};
    if (sint) px = a.xp[x];
    const int cxm = px.map;
    const float bx = px.b, ax = px.a, ikx = px.ik, bhx = px.bh, ahx = px.ah, ikhx = px.ikh;
    const bool lcol = own && ix_on_lattice(a, x);
    const int jx = lcol ? (x - a.snap_x1) / a.snap_dx : 0;
    // unsigned offsets let the loads use a scalar base plus a 32-bit vector offset
    auto row = [&](int r) { return (unsigned)(min(max(r, 0), a.nz - 1) * P + cx); };

    // Rotating register queues of Q slots.  The row loop below is unrolled Q
    // times, so with t the unrolled position of input row j every slot index is
    // a compile-time constant and the queues advance by renaming registers
    // instead of moving them:
    //   vz, vx of row j + d            in slot VQ(d) = (t + N + d) mod Q, d = -N .. N + D
    //   new szz, szx, sxx of row j + d in slot SQ(d) = (t + 2N + d) mod Q, d = -2N .. 0
    //   old stresses and media of row j + d in slot (t + d) mod R, d = 0 .. D
    // Q is a multiple of R = D + 1 and holds the 2N + 1 + D rows of vx in flight.
    constexpr int R = D + 1;
    constexpr int Q = ((2 * N + 1 + D + R - 1) / R) * R;
#define VQ(d) ((t + N + (d) + 2 * Q) % Q)
#define SQ(d) ((t + 2 * N + (d) + 2 * Q) % Q)
    float qz[Q], qx[Q], qzz[Q], qzx[Q], qxx[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        qz[i] = qx[i] = 0.f;
        qzz[i] = qzx[i] = qxx[i] = 0.f;
    }
    const int jb = z0 - N, je = z1 + N;
    // t = 0 at row jb: vz rows jb-N .. jb+N-2+D in slots 0 .., vx rows jb-N .. jb+N-1+D
#pragma unroll
    for (int i = 0; i < 2 * N - 1 + D; ++i) qz[i] = vz[row(jb - N + i)];
#pragma unroll
    for (int i = 0; i < 2 * N + D; ++i) qx[i] = vx[row(jb - N + i)];
    float n_zz[R], n_zx[R], n_xx[R], n_l[R], n_m[R], n_mzx[R];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const unsigned r0 = row(jb + k);
        n_zz[k] = szz[r0]; n_zx[k] = szx[r0]; n_xx[k] = sxx[r0];
        n_l[k] = a.lam[r0]; n_m[k] = a.mu[r0]; n_mzx[k] = a.muzx[r0];
    }

    for (int j0 = jb; j0 < je; j0 += Q) {
#pragma unroll
    for (int t = 0; t < Q; ++t) {
        const int j = j0 + t;
        if (j >= je) break;                    // uniform over the wavefront
        const int cur = t % R, ahead = (t + D) % R;
        const float c_zz = n_zz[cur], c_zx = n_zx[cur], c_xx = n_xx[cur];
        const float lam = n_l[cur], mu = n_m[cur], muzx = n_mzx[cur];
        if (j + D < je) {                      // input row j + D, into slots no longer read
            qz[VQ(N - 1 + D)] = vz[row(j + N - 1 + D)];
            qx[VQ(N + D)] = vx[row(j + N + D)];
            const unsigned r1 = row(j + D);
            n_zz[ahead] = szz[r1]; n_zx[ahead] = szx[r1]; n_xx[ahead] = sxx[r1];
            n_l[ahead] = a.lam[r1]; n_m[ahead] = a.mu[r1]; n_mzx[ahead] = a.muzx[r1];
        }
        // ---- stress half-step on row j
        const float vzc = qz[VQ(0)], vxc = qx[VQ(0)];   // this lane's vz, vx on row j
        float vz_z = 0.f, vx_z = 0.f, vz_x = 0.f, vx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            vz_z += a.cz[k - 1] * (qz[VQ(k - 1)] - qz[VQ(-k)]);
            vx_z += a.cz[k - 1] * (qx[VQ(k)] - qx[VQ(1 - k)]);
            vz_x += a.cx[k - 1] * (lane_off(vzc, lane, k) - lane_off(vzc, lane, 1 - k));
            vx_x += a.cx[k - 1] * (lane_off(vxc, lane, k - 1) - lane_off(vxc, lane, -k));
        }
        float nzz = c_zz, nzx = c_zx, nxx = c_xx;   // rows and columns outside the interior keep their values
        if (sint && j >= N && j < nzi) {
            const bool wr = own && j >= z0 && j < z1;
            const CpmlPoint pz = a.zp[j];
            if (pz.map >= 0) {
                const unsigned pi = pz.map * P + x;
                const float p1 = pz.b * a.pvz_z[pzo + pi] + pz.a * vz_z;
                const float p2 = pz.bh * a.pvx_z[pzo + pi] + pz.ah * vx_z;
                if (wr) {
                    a.o_pvz_z[pzo + pi] = p1;
                    a.o_pvx_z[pzo + pi] = p2;
                }
                vz_z = vz_z * pz.ik + p1;
                vx_z = vx_z * pz.ikh + p2;
            }
            if (cxm >= 0) {
                const unsigned pi = j * a.nxs + cxm;
                const float p1 = bx * a.pvx_x[pxo + pi] + ax * vx_x;
                const float p2 = bhx * a.pvz_x[pxo + pi] + ahx * vz_x;
                if (wr) {
                    a.o_pvx_x[pxo + pi] = p1;
                    a.o_pvz_x[pxo + pi] = p2;
                }
                vx_x = vx_x * ikx + p1;
                vz_x = vz_x * ikhx + p2;
            }
            const float div = lam * (vx_x + vz_z);
            nzz = c_zz + dt * (div + 2.f * mu * vz_z);
            nzx = c_zx + dt * muzx * (vz_x + vx_z);
            nxx = c_xx + dt * (div + 2.f * mu * vx_x);
            if (j == a.surf_top) {   // surf_mirror at row N: no image rows (src/cpu/n_kernel_lib_PSV.cpp:297-301)
                nzz = 0.f;
                nzx = 0.f;
                nxx = 4.f * dt * vx_x * (lam * mu + mu * mu) / (lam + 2.f * mu);
            }
            if (wr) {
                const unsigned r = j * P + x;
                oszz[r] = nzz;
                oszx[r] = nzx;
                osxx[r] = nxx;
            }
        }
        qzz[SQ(0)] = nzz;
        qzx[SQ(0)] = nzx;
        qxx[SQ(0)] = nxx;

        // ---- velocity half-step on row z = j - N
        const int z = j - N;
        if (z < z0) continue;                  // the first 2N input rows only fill the queues
        const float szxc = qzx[SQ(-N)], sxxc = qxx[SQ(-N)];   // this lane's new szx, sxx on row z
        float szz_z = 0.f, szx_z = 0.f, szx_x = 0.f, sxx_x = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) {
            szz_z += a.cz[k - 1] * (qzz[SQ(k - N)] - qzz[SQ(1 - N - k)]);
            szx_z += a.cz[k - 1] * (qzx[SQ(k - 1 - N)] - qzx[SQ(-N - k)]);
            szx_x += a.cx[k - 1] * (lane_off(szxc, lane, k - 1) - lane_off(szxc, lane, -k));
            sxx_x += a.cx[k - 1] * (lane_off(sxxc, lane, k) - lane_off(sxxc, lane, 1 - k));
        }
        if (!own) continue;
        const unsigned r = z * P + x;
        // integer-point profiles for szz_z and sxx_x, half-point for szx_z and szx_x,
        // the reference's slot mapping (src/cpu/n_kernel_PSV.cpp:142-146)
        const CpmlPoint pz = a.zp[z];
        if (pz.map >= 0) {
            const unsigned pi = pz.map * P + x;
            const float p1 = pz.b * a.pszz_z[pzo + pi] + pz.a * szz_z;
            const float p2 = pz.bh * a.pszx_z[pzo + pi] + pz.ah * szx_z;
            a.pszz_z[pzo + pi] = p1;
            a.pszx_z[pzo + pi] = p2;
            szz_z = szz_z * pz.ik + p1;
            szx_z = szx_z * pz.ikh + p2;
        }
        if (cxm >= 0) {
            const unsigned pi = z * a.nxs + cxm;
            const float p1 = bx * a.psxx_x[pxo + pi] + ax * sxx_x;
            const float p2 = bhx * a.pszx_x[pxo + pi] + ahx * szx_x;
            a.psxx_x[pxo + pi] = p1;
            a.pszx_x[pxo + pi] = p2;
            sxx_x = sxx_x * ikx + p1;
            szx_x = szx_x * ikhx + p2;
        }
        const float vz0 = qz[VQ(-N)], vx0 = qx[VQ(-N)];   // vz, vx of row z at the previous level
        const float vz1 = vz0 + dt * a.rhoz[r] * (szx_x + szz_z);
        const float vx1 = vx0 + dt * a.rhox[r] * (sxx_x + szx_z);
        ovz[r] = vz1;
        ovx[r] = vx1;
        float uz = 0.f, ux = 0.f;
        if (kAdj) {
            float *__restrict__ uzs = a.uz + fo;
            float *__restrict__ uxs = a.ux + fo;
            uz = uzs[r] + dt * vz0;
            ux = uxs[r] + dt * vx0;
            uzs[r] = uz;
            uxs[r] = ux;
        }
        if (kEnergy) {
            float *__restrict__ we = a.We + fo;
            we[r] += vx1 * vx1 + vz1 * vz1;
        }
        if (!kAdj && !kStore) continue;
        if (lcol && z >= a.snap_z1 && z <= a.snap_z2 && (z - a.snap_z1) % a.snap_dz == 0) {
            const int jz = (z - a.snap_z1) / a.snap_dz;
            const unsigned q = jz * a.snx + jx;   // offset in one shot's snapshot slab
            const float c_zz2 = qzz[SQ(-N)];   // new szz on row z
            if (doGrad) {
                const size_t go = (size_t)s * a.snz * a.snx, so = sv.off + (size_t)s * a.sn_stride;
                const float Szz = a.sn_szz[so + q], Sxx = a.sn_sxx[so + q], Szx = a.sn_szx[so + q];
                const float Vz = a.sn_vz[so + q], Vx = a.sn_vx[so + q];
                const float glam = a.lam[r], gmu = a.mu[r];
                const float lm = glam + gmu;
                const float il = a.gain_lm / (lm * lm), im = a.gain_lm / (gmu * gmu);
                const float s1 = 0.25f * il * (Szz + Sxx) * (c_zz2 + sxxc);
                const float s2 = 0.25f * im * (Szz - Sxx) * (c_zz2 - sxxc);
                const float s3 = im * (Szx * szxc);
                grad_accumulate(a, go + q, s1, s2, s3, ux * Vx + uz * Vz, dt * (vz1 * Vz), dt * (vx1 * Vx));
            }
            if (doStore) {   // gard_fwd_storage2; k_inject rewrites source cells after injection
                const size_t o = sv.off + (size_t)s * a.st_stride + q;
                a.st_vz[o] = vz1 / dt;
                a.st_vx[o] = vx1 / dt;
                a.st_szz[o] = c_zz2;
                a.st_szx[o] = szxc;
                a.st_sxx[o] = sxxc;
            }
        }
    }
    }
#undef VQ
#undef SQ
}

// surf_mirror, one side per launch.  side 0/1: top/bottom row, one thread per
// column; side 2/3: left/right column, one thread per row.  The tangential
// velocity derivative is recomputed with the CPML correction the stress kernel
// applied (its memory variable already holds this step's value).
template <int N>
__global__ void k_surface(const KArgs a, int side, int is) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    const size_t fo = (size_t)s * a.fs;
    const int P = a.P;
    const float dt = a.dt;
    if (side < 2) {
        const int ix = N + t;
        if (ix >= a.nx - N) return;
        const size_t r = (size_t)is * P + ix;
        float d = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k) d += a.cx[k - 1] * (a.vx[fo + r + k - 1] - a.vx[fo + r - k]);
        const int cxm = a.xmap[ix];
        if (cxm >= 0) d = d * a.ikx[ix] + a.pvx_x[((size_t)s * a.nz + is) * a.nxs + cxm];
        const float l = a.lam[r], u = a.mu[r];
        a.szz[fo + r] = 0.f;
        a.szx[fo + r] = 0.f;
        a.sxx[fo + r] = 4.f * dt * d * (l * u + u * u) / (l + 2.f * u);
        if (side == 0) {
            for (int k = 1; k < is - N + 1; ++k) {
                a.szx[fo + r - (size_t)k * P] = -a.szx[fo + r + (size_t)k * P];
                a.szz[fo + r - (size_t)k * P] = -a.szz[fo + r + (size_t)k * P];
            }
        } else {
            for (int k = 1; k <= a.nz - N - is; ++k) {
                a.szx[fo + r + (size_t)k * P] = -a.szx[fo + r - (size_t)k * P];
                a.szz[fo + r + (size_t)k * P] = -a.szz[fo + r - (size_t)k * P];
            }
        }
    } else {
        const int iz = N + t;
        if (iz >= a.nz - N) return;
        const size_t r = (size_t)iz * P + is;
        float d = 0.f;
#pragma unroll
        for (int k = 1; k <= N; ++k)
            d += a.cz[k - 1] * (a.vz[fo + r + (size_t)(k - 1) * P] - a.vz[fo + r - (size_t)k * P]);
        const int rz = a.zmap[iz];
        if (rz >= 0) d = d * a.ikz[iz] + a.pvz_z[((size_t)s * a.nzs + rz) * P + is];
        const float l = a.lam[r], u = a.mu[r];
        a.sxx[fo + r] = 0.f;
        a.szx[fo + r] = 0.f;
        a.szz[fo + r] = 4.f * dt * d * (l * u + u * u) / (l + 2.f * u);
        if (side == 2) {
            for (int k = 1; k < is - N + 1; ++k) {
                a.szx[fo + r - k] = -a.szx[fo + r + k];
                a.sxx[fo + r - k] = -a.sxx[fo + r + k];
            }
        } else {
            for (int k = 1; k <= a.nx - N - is; ++k) {
                a.szx[fo + r + k] = -a.szx[fo + r - k];
                a.sxx[fo + r + k] = -a.sxx[fo + r - k];
            }
        }
    }
}

// vsrc2: one thread per (shot, cell) group; a group's sources are added in
// source order, as the reference's serial loop does for sources sharing a cell.
__global__ void k_inject(const KArgs a, const SrcArgs sa, int it_arg) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= sa.ngroups) return;
    const StepView sv = step_view(a, it_arg);
    const int it = sv.it;
    const int s = sa.grp_shot[g] - sa.shot0;
    const int cell = sa.grp_cell[g];
    const size_t o = (size_t)s * a.fs + cell;
    float vz = a.vz[o], vx = a.vx[o];
    const double sc = sa.scale_shot ? sa.scale_shot[sa.grp_shot[g]] : 1.0;
    for (int j = sa.grp_off[g]; j < sa.grp_off[g + 1]; ++j) {
        const size_t row = (size_t)sa.grp_idx[j] * sa.nt + it;
        const double fz = sa.stf_z[row] * sc, fx = sa.stf_x[row] * sc;
        if (sa.type == 0) {
            vz += (float)(sa.dt * (double)a.rhoz[cell] * fz / sa.dzdx);
            vx += (float)(sa.dt * (double)a.rhox[cell] * fx / sa.dzdx);
        } else {
            vz += (float)fz;
            vx += (float)fx;
        }
    }
    a.vz[o] = vz;
    a.vx[o] = vx;
    if (sv.store) {   // the stencil kernel stored this cell's velocities before the injection
        const int iz = cell / a.P, ix = cell - iz * a.P;
        if (iz_on_lattice(a, iz) && ix_on_lattice(a, ix)) {
            const size_t q = sv.off + (size_t)s * a.st_stride + (size_t)((iz - a.snap_z1) / a.snap_dz) * a.snx +
                             (ix - a.snap_x1) / a.snap_dx;
            a.st_vz[q] = vz / a.dt;
            a.st_vx[q] = vx / a.dt;
        }
    }
}

// the base time index of a replayed graph (a kernel, not a memset: it may sit in a graph)
__global__ void k_set_int(int *p, int v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

// urec2: displacement as the FP64 running sum of dt v / (dz dx).
__global__ void k_record(const KArgs a, const RecArgs ra, int it_arg) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ra.nrec) return;
    const int it = step_view(a, it_arg).it;
    const size_t o = (size_t)(ra.rec_shot[r] - ra.shot0) * a.fs + ra.rec_cell[r];
    const double uz = ra.acc_z[r] + ra.dt * (double)a.vz[o] / ra.dzdx;
    const double ux = ra.acc_x[r] + ra.dt * (double)a.vx[o] / ra.dzdx;
    ra.acc_z[r] = uz;
    ra.acc_x[r] = ux;
    ra.seis_z[(size_t)r * ra.nt + it] = uz;
    ra.seis_x[(size_t)r * ra.nt + it] = ux;
}

// gard_fwd_storage2: stresses as they are, velocities / dt, on the strided box.
__global__ void k_store(const KArgs a, float *d0, float *d1, float *d2, float *d3, float *d4, size_t dst_stride,
                        float dt) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (q >= a.snz * a.snx) return;
    const int jz = q / a.snx, jx = q - jz * a.snx;
    const int iz = a.snap_z1 + jz * a.snap_dz, ix = a.snap_x1 + jx * a.snap_dx;
    const size_t i = fidx(a, s, iz, ix);
    const size_t o = (size_t)s * dst_stride + q;
    d0[o] = a.vz[i] / dt;
    d1[o] = a.vx[i] / dt;
    d2[o] = a.szz[i];
    d3[o] = a.szx[i];
    d4[o] = a.sxx[i];
}

}  // namespace

// KER<N, TZ, EXTRA...>: EXTRA is empty or the min-waves bound of variant 3
#define PSV_TZ_CASES(KER, N, ...)                                                             \
    switch (tz) {                                                                             \
        case 8: hipLaunchKernelGGL((KER<N, 8 __VA_ARGS__>), grid, block, 0, st, a); break;   \
        case 32: hipLaunchKernelGGL((KER<N, 32 __VA_ARGS__>), grid, block, 0, st, a); break; \
        case 64: hipLaunchKernelGGL((KER<N, 64 __VA_ARGS__>), grid, block, 0, st, a); break; \
        default: hipLaunchKernelGGL((KER<N, 16 __VA_ARGS__>), grid, block, 0, st, a); break; \
    }

#define PSV_N_CASES(KER, ...)                            \
    switch (N) {                                         \
        case 1: PSV_TZ_CASES(KER, 1, __VA_ARGS__) break; \
        case 2: PSV_TZ_CASES(KER, 2, __VA_ARGS__) break; \
        case 3: PSV_TZ_CASES(KER, 3, __VA_ARGS__) break; \
        case 4: PSV_TZ_CASES(KER, 4, __VA_ARGS__) break; \
        case 5: PSV_TZ_CASES(KER, 5, __VA_ARGS__) break; \
        default: PSV_TZ_CASES(KER, 6, __VA_ARGS__) break; \
    }

// one-dimensional launch; tile_of() maps the index to (tile, shot).  Variant 1
// writes 64 columns per wavefront, variants 2 and 3 write 64 - 2N.  Variant 3 is
// variant 2 compiled for at least 6 wavefronts per SIMD (a register cap).
static dim3 stencil_grid(const KArgs &a, int N, int tz, int var) {
    const bool full = var >= 10;   // tiles of 64 columns from column 0
    const int nxi = full ? a.nx - N : a.nx - 2 * N, nzi = a.nz - 2 * N;
    const int segs = (nzi + tz - 1) / tz;
    const int outw = var >= 2 && !full ? kBX - 2 * N : kBX;
    const int tiles = (nxi + outw - 1) / outw;
    if (var == 11)   // a block: kBY shots of one tile and segment
        return dim3(tiles * segs * ((a.B + kBY - 1) / kBY));
    if (var >= 12)   // a block: kBY adjacent tiles of one segment
        return dim3((tiles + kBY - 1) / kBY * segs * a.B);
    return dim3(tiles * ((segs + kBY - 1) / kBY) * a.B);
}

// the two-kernel form is instantiated for TZ 8, 16, 32, 64 (16 for anything else);
// variants 5 and 10 for TZ 16 and 32
static int two_kernel_tz(int tz, int var) {
    if (var == 5 || var >= 10) return tz == 32 ? 32 : 16;
    return (tz == 8 || tz == 32 || tz == 64) ? tz : 16;
}

// variant 5 and its full-tile forms, each compiled in its own translation unit
static bool launch_family5(int var, bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid,
                           hipStream_t st) {
    switch (var) {
        case 13: return launch_stencil5_v13(velocity, a, zt, N, tz, grid, st);
        case 12: return launch_stencil5_v12(velocity, a, zt, N, tz, grid, st);
        case 11: return launch_stencil5_v11(velocity, a, zt, N, tz, grid, st);
        case 10: return launch_stencil5_v10(velocity, a, zt, N, tz, grid, st);
        default: return launch_stencil5_v5(velocity, a, zt, N, tz, grid, st);
    }
}

void launch_stress(const KArgs &a, int N, int tz, int var, hipStream_t st, const ZTab *zt, int depth) {
    if (var >= 5 && !zt) var = 2;   // variants 5, 6 and 10 need the strip-row table
    if (var == 6) {
        launch6(0, a, *zt, N, depth, tz, st);
        return;
    }
    tz = two_kernel_tz(tz, var);
    const dim3 grid = stencil_grid(a, N, tz, var), block(kBX, kBY);
    if (var == 5 || var >= 10) {
        launch_family5(var, false, a, *zt, N, tz, grid, st);
    } else if (var == 3) {
        PSV_N_CASES(k_stress2, , 6)
    } else if (var == 2) {
        PSV_N_CASES(k_stress2, )
    } else {
        PSV_N_CASES(k_stress, )
    }
}

void launch_velocity(const KArgs &a, int N, int tz, int var, hipStream_t st, const ZTab *zt, int depth) {
    // variants 5, 6 and 10 need the strip-row table and correlate in adjoint launches only
    if (var >= 5 && (!zt || (a.grad && !a.displ))) var = 2;
    if (var == 6) {
        if (launch6(1, a, *zt, N, depth, tz, st)) return;
        var = 2;   // a launch mode variant 6 has no kernel for
    }
    if (var == 5 || var >= 10) {
        const int tz5 = two_kernel_tz(tz, var);
        const dim3 g5 = stencil_grid(a, N, tz5, var);
        if (launch_family5(var, true, a, *zt, N, tz5, g5, st)) return;
        var = 2;   // a launch mode variants 5 and 10 to 13 have no kernel for
    }
    tz = two_kernel_tz(tz, var);
    const dim3 grid = stencil_grid(a, N, tz, var), block(kBX, kBY);
    if (var == 3) {
        PSV_N_CASES(k_velocity2, , 6)
    } else if (var == 2) {
        PSV_N_CASES(k_velocity2, )
    } else {
        PSV_N_CASES(k_velocity, )
    }
}

// the launch modes the host runs get their own code; anything else takes the general one
template <int N, int TZ, int D>
static void step_modes(const KArgs &a, dim3 grid, dim3 block, hipStream_t st) {
    const int mode = (a.energy ? kModeEnergy : 0) | (a.displ ? kModeAdjoint : 0) | (a.store ? kModeStore : 0);
    switch (mode) {
        case 0: hipLaunchKernelGGL((k_step<N, TZ, 0, D>), grid, block, 0, st, a); break;            // recompute
        case kModeEnergy: hipLaunchKernelGGL((k_step<N, TZ, kModeEnergy, D>), grid, block, 0, st, a); break;
        case kModeStore: hipLaunchKernelGGL((k_step<N, TZ, kModeStore, D>), grid, block, 0, st, a); break;
        case kModeEnergy | kModeStore:
            hipLaunchKernelGGL((k_step<N, TZ, kModeEnergy | kModeStore, D>), grid, block, 0, st, a);
            break;
        case kModeEnergy | kModeAdjoint:
            hipLaunchKernelGGL((k_step<N, TZ, kModeEnergy | kModeAdjoint, D>), grid, block, 0, st, a);
            break;
        default: hipLaunchKernelGGL((k_step<N, TZ, kModeGeneral, D>), grid, block, 0, st, a); break;
    }
}

// Only depth 1 is instantiated: on gfx950 at order 8, depths 2 and 3 cost a wave per
// SIMD (136-146 VGPRs) and measured 13-15 % slower at 1001 x 2001 x 32 shots.
template <int N, int TZ>
static void step_depth(const KArgs &a, int depth, dim3 grid, dim3 block, hipStream_t st) {
    (void)depth;
    step_modes<N, TZ, 1>(a, grid, block, st);
}

template <int N>
static void step_tz(const KArgs &a, int tz, int depth, dim3 grid, dim3 block, hipStream_t st) {
    if (tz == 32)
        step_depth<N, 32>(a, depth, grid, block, st);
    else
        step_depth<N, 64>(a, depth, grid, block, st);
}

// k_step writes 64 - 4N columns per wavefront; tz is 32 or 64, depth 1 .. 3
void launch_step(const KArgs &a, int N, int tz, int depth, hipStream_t st) {
    if (tz != 32) tz = 64;   // the grid must match the instantiated TZ
    const int nxi = a.nx - 2 * N, nzi = a.nz - 2 * N;
    const int outw = kBX - 4 * N;
    const int segs = (nzi + tz - 1) / tz;
    const dim3 grid(((nxi + outw - 1) / outw) * ((segs + kBY - 1) / kBY) * a.B), block(kBX, kBY);
    switch (N) {
        case 1: step_tz<1>(a, tz, depth, grid, block, st); break;
        case 2: step_tz<2>(a, tz, depth, grid, block, st); break;
        case 3: step_tz<3>(a, tz, depth, grid, block, st); break;
        case 4: step_tz<4>(a, tz, depth, grid, block, st); break;
        case 5: step_tz<5>(a, tz, depth, grid, block, st); break;
        default: step_tz<6>(a, tz, depth, grid, block, st); break;
    }
}

void launch_surface(const KArgs &a, int N, const int isurf[4], hipStream_t st) {
    for (int side = 0; side < 4; ++side) {
        if (isurf[side] <= 0) continue;
        const int n = side < 2 ? a.nx - 2 * N : a.nz - 2 * N;
        const dim3 grid((n + 255) / 256, a.B), block(256);
        switch (N) {
            case 1: hipLaunchKernelGGL(k_surface<1>, grid, block, 0, st, a, side, isurf[side]); break;
            case 2: hipLaunchKernelGGL(k_surface<2>, grid, block, 0, st, a, side, isurf[side]); break;
            case 3: hipLaunchKernelGGL(k_surface<3>, grid, block, 0, st, a, side, isurf[side]); break;
            case 4: hipLaunchKernelGGL(k_surface<4>, grid, block, 0, st, a, side, isurf[side]); break;
            case 5: hipLaunchKernelGGL(k_surface<5>, grid, block, 0, st, a, side, isurf[side]); break;
            default: hipLaunchKernelGGL(k_surface<6>, grid, block, 0, st, a, side, isurf[side]); break;
        }
    }
}

void launch_inject(const KArgs &a, const SrcArgs &s, int it, hipStream_t st) {
    if (s.ngroups <= 0) return;
    hipLaunchKernelGGL(k_inject, dim3((s.ngroups + 127) / 128), dim3(128), 0, st, a, s, it);
}

void launch_set_int(int *p, int v, hipStream_t st) { hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, p, v); }

void launch_record(const KArgs &a, const RecArgs &r, int it, hipStream_t st) {
    if (r.nrec <= 0) return;
    hipLaunchKernelGGL(k_record, dim3((r.nrec + 127) / 128), dim3(128), 0, st, a, r, it);
}

void launch_store(const KArgs &a, float *const dst[5], size_t dst_stride, float dt, hipStream_t st) {
    const int n = a.snz * a.snx;
    hipLaunchKernelGGL(k_store, dim3((n + 255) / 256, a.B), dim3(256), 0, st, a, dst[0], dst[1], dst[2], dst[3],
                       dst[4], dst_stride, dt);
}

}  // namespace psv
