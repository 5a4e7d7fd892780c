// Shared definitions for the nonorthogonal-Wick engine (host and device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nw {

// Slab offsets of one ordered reference pair, in the order of
// assign_offsets (src/nonorthogonalwicks/layout.rs:21-189).  Per spin:
// x[2], y[2], fh[2][2], ff[2][2], v[2][2][2], j[10], hcol0 (31 entries);
// then vab[2][2][2], vba[2][2][2], iiab[2][2][2][2]; the last entry is the
// block length.  -1 = not allocated.
enum : int {
    O_X = 0, O_Y = 2, O_FH = 4, O_FF = 8, O_V = 12, O_J = 20, O_HCOL0 = 30, O_SAME = 31,
    O_AA = 0, O_BB = 31, O_VAB = 62, O_VBA = 70, O_II = 78, O_TOTAL = 94, O_LEN = 95
};

// Scalar meta of one ordered pair: per spin (aa at 0, bb at 11)
// tsp, phase, m, nocc, f0h[2], f0f[2], v0[3]; vab0[2][2] at 22; vba0[2][2] at 26.
enum : int {
    M_TSP = 0, M_PH = 1, M_M = 2, M_NOCC = 3, M_F0H = 4, M_F0F = 6, M_V0 = 8,
    M_AA = 0, M_BB = 11, M_VAB0 = 22, M_VBA0 = 26, M_LEN = 32
};

// SAME_SPIN_J_BRANCHES (build.rs:19-30)
__host__ __device__ inline int jbranch(int slot, int k)
{
    const int t[10][4] = {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}, {0, 0, 1, 1}, {0, 1, 0, 1},
                          {0, 1, 1, 0}, {0, 1, 1, 1}, {1, 0, 1, 0}, {1, 0, 1, 1}, {1, 1, 1, 1}};
    return t[slot][k];
}

// Per ordered reference pair record on the device (one per x*nref + w).
struct alignas(16) PairInfo {
    long long base;        // slab offset of the block
    long long off[O_LEN];  // relative offsets
    double meta[M_LEN];
    int built, ma, mb, bidx;  // bidx: position of the pair in the last build
    double smin;              // smallest singular value above the zero threshold (both spins)
    int fp32, pad;            // 1: four-index transforms of this pair ran in FP32
};

// Compact determinant for evaluation kernels (ExcitationCache, src/lib.rs:128-167):
// per spin the rank and the first four holes then four particles (ascending).
struct alignas(8) PackedDet {
    uint8_t ra, rb;
    uint16_t parent;
    int8_t pha, phb;        // excitation phases per spin, +-1
    uint8_t pad[2];
    uint8_t ia[8], ib[8];
};

constexpr int MAX_NOCC = 64;
constexpr int MAX_NMO = 128;

}  // namespace nw
