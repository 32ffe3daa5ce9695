// stream_probe.hip -- the bandwidth a kernel with the stencil's memory access shape
// reaches on this GPU, with no arithmetic: the ceiling the stencil kernels are set
// beside in bench.py (no counterpart in the reference).
//
// Shape: 10 arrays read and 5 written, [B][nz][P] floats each (the batch's wavefield
// layout); a wavefront owns 64 consecutive columns and marches down TZ rows, every
// load of row z + 1 issued while row z is stored; a block's four wavefronts take
// four adjacent column tiles of one z segment (layout 0) or four z segments of one
// column tile (layout 1); blocks run x fastest, then z, shots outermost, each XCD
// owning a contiguous range (the stencil launches' layout).
//
//     stream_probe <nz> <P> <B> [layout [TZ]]     layout 0 (default) or 1, TZ 40 (default)
//                                                 or 16; prints {"shape_GBps": ..., ...}
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

constexpr int kIn = 10, kOut = 5, kWaves = 4;

struct ProbeArgs {
    const float *in[kIn];
    float *out[kOut];
    int nz, P, B;
};

// blocks of the launch: column tiles x z segments x shots
template <int TZ, int LAYOUT>
int probe_blocks(const ProbeArgs &a) {
    const int seg = LAYOUT == 0 ? TZ : TZ * kWaves;
    return (LAYOUT == 0 ? a.P / (64 * kWaves) : a.P / 64) * ((a.nz + seg - 1) / seg) * a.B;
}

template <int TZ, int LAYOUT>
__global__ __launch_bounds__(64 * kWaves) void k_probe(ProbeArgs a) {
    const int seg = LAYOUT == 0 ? TZ : TZ * kWaves;
    const int gx = LAYOUT == 0 ? a.P / (64 * kWaves) : a.P / 64, gy = (a.nz + seg - 1) / seg;
    const int n = gx * gy * a.B;
    const int L = blockIdx.x, xcd = L & 7, loc = L >> 3, q = n >> 3, r = n & 7;
    int t = xcd * q + min(xcd, r) + loc;
    const int bx = t % gx;
    t /= gx;
    const int by = t % gy, s = t / gy;
    const int w = (int)(threadIdx.x >> 6);
    const int col = (LAYOUT == 0 ? bx * kWaves + w : bx) * 64;
    const int z0 = by * seg + (LAYOUT == 0 ? 0 : w * TZ), z1 = min(z0 + TZ, a.nz);
    if (z0 >= z1) return;   // a z segment past the last row
    const uint32_t lane4 = (threadIdx.x & 63) * 4u;
    // byte offset of row z of this wavefront's tile; every array is below 4 GiB
    auto off = [&](int z) {
        return (uint32_t)__builtin_amdgcn_readfirstlane(
                   (int)(uint32_t)((((size_t)s * a.nz + z) * a.P + col) * 4)) + lane4;
    };
    float cur[kIn], nxt[kIn];
    uint32_t o = off(z0);
#pragma unroll
    for (int f = 0; f < kIn; ++f) cur[f] = *(const float *)((const char *)a.in[f] + o);
    for (int z = z0; z < z1; ++z) {
        const uint32_t on = off(min(z + 1, z1 - 1));
#pragma unroll
        for (int f = 0; f < kIn; ++f) nxt[f] = *(const float *)((const char *)a.in[f] + on);
#pragma unroll
        for (int k = 0; k < kOut; ++k) *(float *)((char *)a.out[k] + o) = cur[2 * k] + cur[2 * k + 1];
#pragma unroll
        for (int f = 0; f < kIn; ++f) cur[f] = nxt[f];
        o = on;
    }
}

int fail(const char *what, hipError_t e) {
    std::fprintf(stderr, "stream_probe: %s: %s\n", what, hipGetErrorString(e));
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 4 || argc > 6) {
        std::fprintf(stderr, "usage: stream_probe <nz> <P> <B> [layout 0|1 [TZ 40|16]]\n");
        return 2;
    }
    ProbeArgs a{};
    a.nz = std::atoi(argv[1]);
    a.P = std::atoi(argv[2]);
    a.B = std::atoi(argv[3]);
    const int layout = argc > 4 ? std::atoi(argv[4]) : 0, tz = argc > 5 ? std::atoi(argv[5]) : 40;
    if ((layout != 0 && layout != 1) || (tz != 40 && tz != 16)) {
        std::fprintf(stderr, "stream_probe: layout 0 or 1, TZ 40 or 16\n");
        return 2;
    }
    const size_t bytes = (size_t)a.nz * a.P * a.B * 4;
    if (a.nz < 1 || a.B < 1 || a.P < 64 * kWaves || a.P % (64 * kWaves) || bytes >= (1ull << 31)) {
        std::fprintf(stderr, "stream_probe: need P a multiple of %d and arrays below 2 GiB\n", 64 * kWaves);
        return 2;
    }
    std::vector<void *> bufs;
    for (int i = 0; i < kIn + kOut; ++i) {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return fail("hipMalloc", e);
        e = hipMemset(p, 0, bytes);
        if (e != hipSuccess) return fail("hipMemset", e);
        bufs.push_back(p);
        if (i < kIn)
            a.in[i] = static_cast<const float *>(p);
        else
            a.out[i - kIn] = static_cast<float *>(p);
    }
    using Kernel = void (*)(ProbeArgs);
    Kernel k = layout == 0 ? (tz == 16 ? k_probe<16, 0> : k_probe<40, 0>) : (tz == 16 ? k_probe<16, 1> : k_probe<40, 1>);
    const int grid = layout == 0 ? (tz == 16 ? probe_blocks<16, 0>(a) : probe_blocks<40, 0>(a))
                                 : (tz == 16 ? probe_blocks<16, 1>(a) : probe_blocks<40, 1>(a));
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k, dim3(grid), dim3(64 * kWaves), 0, 0, a);   // warm-up
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail("launch", e);
    const int reps = 10;
    (void)hipEventRecord(e0);
    for (int i = 0; i < reps; ++i) hipLaunchKernelGGL(k, dim3(grid), dim3(64 * kWaves), 0, 0, a);
    (void)hipEventRecord(e1);
    e = hipEventSynchronize(e1);
    if (e != hipSuccess) return fail("kernel", e);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double moved = (double)bytes * (kIn + kOut);
    std::printf("{\"shape_GBps\": %.1f, \"bytes_per_launch\": %.0f, \"avg_launch_ms\": %.4f, \"layout\": %d, "
                "\"tz\": %d}\n",
                moved / (ms / reps / 1e3) / 1e9, moved, ms / reps, layout, tz);
    for (void *p : bufs) (void)hipFree(p);
    return 0;
}
