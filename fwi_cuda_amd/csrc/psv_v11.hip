// psv_v11.hip -- the kernels of stencil variants 11 and 12 (psv_stencil5.h), in a
// translation unit of their own so that the variants compile in parallel.
#include "psv_stencil5.h"

namespace psv {

bool launch_stencil5_v11(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st) {
    return launch5<11>(velocity, a, zt, N, tz, grid, st);
}

bool launch_stencil5_v12(bool velocity, const KArgs &a, const ZTab &zt, int N, int tz, dim3 grid, hipStream_t st) {
    return launch5<12>(velocity, a, zt, N, tz, grid, st);
}

}  // namespace psv
