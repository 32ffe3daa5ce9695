// psv_step7_c.hip -- the kernels of variant 7's one-pass step (psv_step7.h) for orders
// 10 and 12 (N = 5, 6), in a translation unit of their own.
#include "psv_step7.h"

namespace psv {

bool launch_step7_top(const KArgs &a, const ZTab &zt, int N, int M, bool xw, dim3 grid, int mult, hipStream_t st) {
    if (N == 5) return step7_n<5>(a, zt, M, xw, grid, mult, st);
    return step7_n<6>(a, zt, M, xw, grid, mult, st);
}

}  // namespace psv
