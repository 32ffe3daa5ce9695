// psv_v5.hip -- the kernels of stencil variants 5 and 10 (psv_stencil5.h), in a
// translation unit of their own so that the variants compile in parallel.
#include "psv_stencil5.h"

namespace psv {

bool launch_stencil5_v5(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st) {
    return launch5<5>(velocity, a, zt, N, tz, grid, st);
}

bool launch_stencil5_v10(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st) {
    return launch5<10>(velocity, a, zt, N, tz, grid, st);
}

}  // namespace psv
