// Fused evaluation launcher: register-class and all-code kernel instances,
// and dispatch to the single-code instances (nw_fused_code.hip).
#include "nw_fused_dev.h"

namespace nw {

__global__ void k_reset_counter(unsigned long long *counter, int n)
{
    for (int k = threadIdx.x; k < n; k += blockDim.x) counter[k] = 0ull;
}

void launch_reset_counters(hipStream_t st, unsigned long long *counters, int n)
{
    if (n > 0) hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(256), 0, st, counters, n);
}

void launch_eval_fused(hipStream_t st, int sel, bool dyn, const BatchTile *tiles, long long ntiles,
                       const BatchSys *sys, unsigned long long *counter, int grid, bool reset)
{
    if (ntiles <= 0) return;
    if (reset && (dyn || sel >= 0)) launch_reset_counters(st, counter, 1);
    if ((long long)grid > ntiles) grid = (int)ntiles;
    if (sel >= 0) {  // single-code kernels exist in the dynamic-schedule form only
        launch_eval_fused_code(st, sel, tiles, ntiles, sys, counter, grid);
        return;
    }
    dim3 g((unsigned)grid), b(256);
#define NW_FC(C)                                                                                           \
    case C:                                                                                                \
        if (dyn) hipLaunchKernelGGL((k_eval_fused<C, true>), g, b, 0, st, tiles, ntiles, sys, counter);   \
        else hipLaunchKernelGGL((k_eval_fused<C, false>), g, b, 0, st, tiles, ntiles, sys, counter);      \
        return;
    switch (sel) {
        NW_FC(-1) NW_FC(-2) NW_FC(-3) NW_FC(FUSE_ALL)
    default: return;
    }
#undef NW_FC
}

}  // namespace nw
