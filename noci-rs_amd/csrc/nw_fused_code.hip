// Single-code instances of the fused evaluation kernel (dynamic schedule).
#include "nw_fused_dev.h"

namespace nw {

void launch_eval_fused_code(hipStream_t st, int code, const BatchTile *tiles, long long ntiles, const BatchSys *sys,
                            unsigned long long *counter, int grid)
{
    dim3 g((unsigned)grid), b(256);
#define NW_FD(C)                                                                                           \
    case C:                                                                                                \
        hipLaunchKernelGGL((k_eval_fused<C, true>), g, b, 0, st, tiles, ntiles, sys, counter);            \
        return;
#define NW_FB(A, B) NW_FD((((A) * 5 + (B)) << 1)) NW_FD(((((A) * 5 + (B)) << 1) | 1))
    switch (code) {
        NW_BIN_CASES(NW_FB)
    default: return;
    }
#undef NW_FB
#undef NW_FD
}

}  // namespace nw
