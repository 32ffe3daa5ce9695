// psv_step7.hip -- variant 7, the one-pass time step with ordered, row-ahead loads
// (psv_step7.h): the launch, and the kernels of orders 2, 4 and 6.  Order 8 is
// instantiated in psv_step7_b.hip and orders 10 and 12 in psv_step7_c.hip, so that the
// three compile in parallel.
#include "psv_step7.h"

namespace psv {

bool launch_step7(const KArgs &a, const ZTab &zt, int N, int tz, bool xw, hipStream_t st) {
    // tz < 0: balanced mode (variant 8's layout) with -tz times the resident blocks
    const int mult = tz < 0 ? -tz : 0;
    if (mult) xw = true;
    const int R = 2 * N + 2, M = tz >= 8 * R ? 8 : 4, TZ = R * M;
    const int nxi = a.nx - 2 * N, nzi = a.nz - 2 * N, outw = kBX - 4 * N;
    const int segs = (nzi + TZ - 1) / TZ, tiles = (nxi + outw - 1) / outw;
    // grid as tile_of<TZ, OUTW, XW> counts it
    const int nblk = xw ? ((tiles + kBY - 1) / kBY) * segs : tiles * ((segs + kBY - 1) / kBY);
    const dim3 grid(nblk * a.B);
    switch (N) {
        case 1: return step7_n<1>(a, zt, M, xw, grid, mult, st);
        case 2: return step7_n<2>(a, zt, M, xw, grid, mult, st);
        case 3: return step7_n<3>(a, zt, M, xw, grid, mult, st);
        default: return launch_step7_high(a, zt, N, M, xw, grid, mult, st);   // psv_step7_b/c.hip
    }
}

}  // namespace psv
