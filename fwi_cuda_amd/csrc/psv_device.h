// psv_device.h -- device helpers shared by the stencil translation units.
#pragma once

#include <type_traits>
#include <utility>

#include "psv_internal.h"

namespace psv {
namespace {

template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
// f(std::integral_constant<int, 0>{}) ... f(std::integral_constant<int, n - 1>{}): loop
// positions that index register arrays must be compile-time constants
template <int n, class F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, n>{});
}

// column / row on the strided snapshot box
__device__ __forceinline__ bool ix_on_lattice(const KArgs &a, int x) {
    return x >= a.snap_x1 && x <= a.snap_x2 && (x - a.snap_x1) % a.snap_dx == 0;
}
__device__ __forceinline__ bool iz_on_lattice(const KArgs &a, int z) {
    return z >= a.snap_z1 && z <= a.snap_z2 && (z - a.snap_z1) % a.snap_dz == 0;
}

// Per-step switches of a launch: from the arguments, or, in a replayed graph,
// from the step table at the time index the graph's base counter gives.
struct StepView {
    int it, grad, store;
    long long off;   // slab offset in the snapshot arrays
};
__device__ __forceinline__ StepView step_view(const KArgs &a, int it) {
    if (!a.si) return {it, a.grad, a.store, 0};
    const int t = *a.si_base + a.si_dir * a.si_k;
    const StepInfo d = a.si[t];
    return {t, a.grad && d.grad, a.store && d.store, d.off};
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).
// The 1-D launch index is remapped so that every XCD owns one contiguous range
// of work, ordered tile-major and shot-minor: the B shots of a tile run back to
// back on one XCD (the tile's media and CPML profiles stay in that L2), then the
// vertically adjacent segment (its z halo rows are the previous tile's rows).
// With XW the kBY wavefronts of a block sit side by side in x (a tile is kBY * OUTW
// columns of one z segment), otherwise one above the other in z (kBY segments of one
// OUTW-column tile).
// FULL: tiles of OUTW = kBX columns starting at column 0 (aligned rows), covering the
// interior columns [N, nx - N); otherwise tiles of OUTW output columns from column N.
// SB: the kBY wavefronts of a block are kBY shots of one tile and one z segment (s
// returns the group of kBY shots, by the segment): they read the same media rows at
// nearly the same time, so those loads are served once per block from the CU's cache.
template <int TZ, int OUTW = kBX, bool XW = false, bool FULL = false, bool SB = false>
__device__ __forceinline__ void tile_of(const KArgs &a, int N, int &bx, int &by, int &s) {
    const int xspan = FULL ? a.nx - N : a.nx - 2 * N;
    const int gx = XW ? ((xspan + OUTW - 1) / OUTW + kBY - 1) / kBY : (xspan + OUTW - 1) / OUTW;
    const int gy = XW || SB ? (a.nz - 2 * N + TZ - 1) / TZ : ((a.nz - 2 * N + TZ - 1) / TZ + kBY - 1) / kBY;
    const int ns = SB ? (a.B + kBY - 1) / kBY : a.B;
    const int n = gx * gy * ns;
    const int L = blockIdx.x, xcd = L & 7, loc = L >> 3;
    const int q = n >> 3, r = n & 7;
    const int W = xcd * q + min(xcd, r) + loc;
    if (a.order == 0) {          // shot fastest, then z segments, then x tiles
        s = W % ns;
        const int t = W / ns;
        by = t % gy;
        bx = t / gy;
    } else if (a.order == 1) {   // shot fastest, then x tiles, then z segments
        s = W % ns;
        const int t = W / ns;
        bx = t % gx;
        by = t / gx;
    } else if (a.order == 2) {   // x tiles fastest, then z segments, shots outermost
        const int t = W % (gx * gy);
        s = W / (gx * gy);
        bx = t % gx;
        by = t / gx;
    } else {                     // x tiles fastest, then shots, z segments outermost
        bx = W % gx;
        const int t = W / gx;
        s = t % ns;
        by = t / ns;
    }
}

// value of v in lane (lane + d); d is a compile-time constant after unrolling
__device__ __forceinline__ float lane_off(float v, int lane, int d) {
    return d == 0 ? v : __shfl(v, lane + d);
}

// Column (lane + d) of a full kBX-column tile, |d| <= the halo width H: v holds the
// lane's own column; h holds the halo, the H columns right of the tile in lanes
// 0..H-1 and the H columns left of it in lanes kBX-H..kBX-1.  One rotation by d: the
// source lanes that only wrapped destinations read offer their halo value instead.
__device__ __forceinline__ float lane_off_h(float v, float h, int lane, int d) {
    if (d == 0) return v;
    const bool from_h = d > 0 ? lane < d : lane >= kBX + d;
    return __shfl(from_h ? h : v, (lane + d) & (kBX - 1));
}

// The column a lane loads into the halo register of a full tile starting at column
// x0 (see lane_off_h), clamped into [0, nx): lanes H..kBX-H-1 repeat lane 0's column
// so that the load touches only the halo's two lines.
__device__ __forceinline__ int halo_col(int x0, int lane, int H, int nx) {
    const int c = lane < H ? x0 + kBX + lane : lane >= kBX - H ? x0 - kBX + lane : x0 + kBX;
    return min(max(c, 0), nx - 1);
}

// Wave-uniform base pointer plus a 32-bit byte offset per lane: the global
// load / store takes its saddr + voffset form.
template <class T>
__device__ __forceinline__ T ld_at(const T *base, uint32_t off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <class T>
__device__ __forceinline__ void st_at(T *base, uint32_t off, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ld_at / st_at of a value used once per launch: with NT the access carries the
// non-temporal hint, so its line is not kept in the caches for reuse (fields updated in
// place, sums, snapshots; not the z-queue rows and halo columns that neighbouring
// segments and tiles read again, nor the media that every shot reads)
template <bool NT, class T>
__device__ __forceinline__ T ld_once(const T *base, uint32_t off) {
    const T *p = reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
    if constexpr (NT) return __builtin_nontemporal_load(p);
    return *p;
}
template <bool NT, class T>
__device__ __forceinline__ void st_once(T *base, uint32_t off, T v) {
    T *p = reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off);
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Fields of the launch's KArgs loaded again from the kernel-argument segment where they
// are used, for blocks that run on a few rows only, such as the CPML strips: they then
// keep no base pointer live in SGPRs across the row loop.  The kernels that use it take
// their KArgs as the first argument (segment offset 0).  karg_seg() passes the segment
// address through an empty asm statement, so the scalar loads from it stay in the block
// that calls it, and issue there together.
using KSeg = const __attribute__((address_space(4))) char *;
__device__ __forceinline__ KSeg karg_seg() {
    KSeg p = (KSeg)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return p;
}
template <class T>
__device__ __forceinline__ T karg_at(KSeg seg, uint32_t off) {
    return *(const __attribute__((address_space(4))) T *)(seg + off);
}
#define PSV_KARG(seg, field) (karg_at<decltype(KArgs::field)>(seg, offsetof(KArgs, field)))
static_assert(std::is_standard_layout<KArgs>::value && std::is_trivially_copyable<KArgs>::value,
              "KArgs is passed by value and read back at offsetof() positions");

// z-strip test and strip row (the psi_z row) of a wave-uniform row
__device__ __forceinline__ bool in_zstrip(const ZTab &zt, int z) { return z < zt.lo || z >= zt.hi; }
__device__ __forceinline__ int strip_row(const ZTab &zt, int N, int z) {
    return z < zt.lo ? z - N : zt.lo - N + (z - zt.hi);
}

// velocity-kernel launch modes of variants 5 and 6 (template flags): We += v^2,
// adjoint (u += dt v, correlation on the step's switch), snapshot store, and the
// misfit-derivative sums (read and written when the arguments' gform is 1; without
// the flag gform must be 0)
constexpr int kVelEnergy = 1, kVelAdjoint = 2, kVelStore = 4, kVelMisfit = 8;

// the velocity launch mode of a set of arguments
__host__ __device__ inline int vel_mode(const KArgs &a) {
    return (a.energy ? kVelEnergy : 0) | (a.displ ? kVelAdjoint : 0) | (a.store ? kVelStore : 0) |
           (a.displ && a.gform ? kVelMisfit : 0);
}

}  // namespace
}  // namespace psv
