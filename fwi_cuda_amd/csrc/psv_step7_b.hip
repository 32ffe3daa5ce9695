// psv_step7_b.hip -- the kernels of variant 7's one-pass step (psv_step7.h) for order 8
// (N = 4), in a translation unit of their own; orders 10 and 12 are in psv_step7_c.hip.
#include "psv_step7.h"

namespace psv {

bool launch_step7_high(const KArgs &a, const ZTab &zt, int N, int M, bool xw, dim3 grid, int mult, hipStream_t st) {
    if (N == 4) return step7_n<4>(a, zt, M, xw, grid, mult, st);
    return launch_step7_top(a, zt, N, M, xw, grid, mult, st);
}

}  // namespace psv
