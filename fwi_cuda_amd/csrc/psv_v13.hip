// psv_v13.hip -- the kernels of stencil variant 13 (psv_stencil5.h), in a
// translation unit of their own so that the variants compile in parallel.
#include "psv_stencil5.h"

namespace psv {

bool launch_stencil5_v13(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st) {
    return launch5<13>(velocity, a, zt, N, tz, grid, st);
}

}  // namespace psv
