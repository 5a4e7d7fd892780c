// Multi-system fused evaluation kernels (nw_batch_*): persistent workgroups
// over a tile list that interleaves the determinant pairs of many systems.
// One work item = one tile of up to 256 determinant pairs of one system, one
// ordered reference pair and one contraction-rank code; the tile is TI x TJ =
// 256 pairs with TJ = 1 << shape.  A persistent workgroup walks its item list
// (static stride or an atomic work counter) and dispatches on the item's
// code.  SEL picks which codes a kernel instance contains: SEL >= 0 one code,
// SEL = -1 - c register class c (light bins keep their occupancy), SEL =
// FUSE_ALL every code.
#pragma once
#include "nw_eval_dev.h"

namespace nw {

template <int SEL, int CODE>
__host__ __device__ constexpr bool sel_has()
{
    return SEL == CODE || SEL == FUSE_ALL || (SEL < 0 && fused_class(CODE) == -1 - SEL);
}

template <int SEL>
__device__ inline void eval_code(int code, const PairInfo &P, slab_t slab, const PackedDet &bra,
                                 const PackedDet &ket, const BatchSys &S, double &v0, double &v1)
{
    switch (code) {
#define NW_F(A, B)                                                                                   \
    case ((((A) * 5 + (B)) << 1) | 1):                                                               \
        if constexpr (sel_has<SEL, ((((A) * 5 + (B)) << 1) | 1)>())                                  \
            hs_m0<A, B>(P, slab, bra, ket, S.nmo, S.enuc, v0, v1);                                   \
        break;                                                                                       \
    case (((A) * 5 + (B)) << 1):                                                                     \
        if constexpr (sel_has<SEL, (((A) * 5 + (B)) << 1)>())                                        \
            hs_gen<A, B>(P, slab, bra, ket, S.nmo, S.enuc, S.tol, S.mode, v0, v1);                   \
        break;
        NW_BIN_CASES(NW_F)
#undef NW_F
    default: break;
    }
}

template <int SEL, bool DYN>
__global__ __launch_bounds__(256) void k_eval_fused(const BatchTile *__restrict__ tiles, long long ntiles,
                                                    const BatchSys *__restrict__ sys,
                                                    unsigned long long *__restrict__ counter)
{
    __shared__ long long wsh;
    long long w = blockIdx.x;
    for (;;) {
        if constexpr (DYN) {
            if (threadIdx.x == 0) wsh = (long long)atomicAdd(counter, 1ull);
            __syncthreads();
            w = wsh;
            __syncthreads();
        }
        if (w >= ntiles) break;
        const BatchTile T = tiles[w];
        const int sh = T.shape;
        const int li = (int)threadIdx.x >> sh, lj = (int)threadIdx.x & ((1 << sh) - 1);
        const long long i = T.i0 + li, j = T.j0 + lj;
        if (li < T.ni && lj < T.nj && !(T.tri && j < i)) {
            const BatchSys &S = sys[T.sys];
            const PairInfo &P = S.pairs[T.pair];
            double v0 = 0.0, v1 = 0.0;
            eval_code<SEL>(T.code, P, (slab_t)(S.slab + P.base), S.dets[i], S.dets[j], S, v0, v1);
            S.H[i * S.ldo + j] = v0;
            S.S[i * S.ldo + j] = v1;
            if (i != j) {
                S.H[j * S.ldo + i] = v0;
                S.S[j * S.ldo + i] = v1;
            }
        }
        if constexpr (!DYN) w += gridDim.x;
    }
}

void launch_eval_fused_code(hipStream_t st, int code, const BatchTile *tiles, long long ntiles, const BatchSys *sys,
                            unsigned long long *counter, int grid);

}  // namespace nw
