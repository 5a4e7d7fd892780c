// Rank-4 slab blocks on demand from Cholesky-type integral factors.
//
// With (ab|cd) = sum_K L_K[a,b] L_K[c,d], every rank-4 tensor of a reference
// pair is a sum of products of MO-transformed factors:
//   II[p,q,r,s]  = sum_K A^a_K[p,q] A^b_K[r,s]                  (build.rs:877-917)
//   J[p,q,r,s]   = sum_K A_K[p,q] A'_K[r,s] - sum_K E_K[p,r] F_K[q,s]
//                                                               (build.rs:261-313)
// with A = cx^T L_K xc, E = cx^T L_K cx, F = xc^T L_K xc for the coefficient
// matrices of the J slot or II block, each held as [row][column][K] (K
// contiguous).  The evaluators of pairs with m_alpha = m_beta = 0 read J
// (slot 0 only) as J[p,q,r,s] - J[p,s,r,q] (the same-spin two-electron term
// of preparehamiltonianoverlap.rs, and every fixed and run-time-rank kernel
// here), and that slot's F_K is symmetric, so the exchange products cancel
// in what they read: their blocks hold the direct products only.  Pairs
// with zero-overlap orbitals read single J entries, and get the exchange
// products as well.
//
// A determinant pair reads its tensors at (row, row) positions
// (p, r) = (rows[eta], rows[xi]) of its contraction labels
// (construct_determinant_indices, eval/prepare.rs:555-595) and, within the
// block T[p][:][r][:], at columns (q, s) of a bra hole (q < nocc or
// s < nocc) or two ket particles of one ket determinant.  The (p, r) blocks
// an evaluation call reads form rectangles (every bra row of one parent x
// every ket row of the other) plus a few single blocks (two rows of one
// determinant).  Each rectangle's bra-hole columns and ket-particle pairs,
// and the single blocks whole (grouped by row p), are row-gathered A B^T
// products on MFMA (k_r4_gemm).  Only what a call reads is built,
// into the reference slab layout; a later call builds what it adds.
// Exports complete the tensors with the dense transforms.
#include <algorithm>
#include <cmath>
#include <map>
#include <thread>
#include <cstdio>
#include <vector>

#include "nw_internal.h"

namespace nw {

typedef double r4_d4 __attribute__((ext_vector_type(4)));
typedef float r4_f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct R4Mfma;
template <>
struct R4Mfma<double> {
    typedef r4_d4 acc;
    static __device__ inline acc mma(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of the 16x16 result held in register r of lane group lk = lane / 16
    static __device__ inline int row(int lk, int r) { return 4 * r + lk; }
};
template <>
struct R4Mfma<float> {
    typedef r4_f4 acc;
    static __device__ inline acc mma(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lk, int r) { return 4 * lk + r; }
};

// Index pairs (u, v) of a product's rows or columns: element i is
// (u[i], v[i]) when paired, else (u[i / nv], v[i % nv]) or (u[i / nv],
// v0 + i % nv) without a v list.
struct R4Axis {
    const int *u, *v;
    int nu, nv, v0, paired;
};
__device__ inline void axis_at(const R4Axis &a, int i, int &u, int &v)
{
    if (a.paired) {
        u = a.u[i];
        v = a.v[i];
    } else {
        u = a.u[i / a.nv];
        v = a.v ? a.v[i % a.nv] : a.v0 + i % a.nv;
    }
}
// One product C[i][j] = sum_K A[row_i][K] B[col_j][K] with A rows
// (u, v) = A[u * nm + v] and B rows (u', v') = B[u' * nm + v'] (factor
// tensors [nm][nm][naux]); mode 0 stores C at out[((u nm + v) nm + u') nm + v']
// (direct products: u = p, v = q, u' = r, v' = s), mode 1 subtracts it at
// out[((u nm + u') nm + v) nm + v'] (exchange products: rows (p, r) of E,
// columns (q, s) of F), mode 2 stores it at out[((u' nm + v') nm + u) nm + v]
// (direct products with rows (r, s) and columns (p, q)), mode 3 as mode 0
// and also at out[((u' nm + v') nm + u) nm + v] (blocks (p, r) and (r, p) of
// a product whose two factor tensors are one: J[r,s,p,q] = J[p,q,r,s]).
struct R4Gemm {
    const void *A, *B;
    void *out;
    R4Axis row, col;
    int M, N, mode, pad;
};
struct R4GTile {
    int g, m0, n0, pad;
};

// Row-gathered A B^T products in BM x BN tiles: four waves in a 2 x 2
// layout, each (BM/2) x (BN/2) of 16 x 16 MFMA tiles; K staged through LDS
// KC at a time ([row][k], rows gathered from the factor tensors, 16-byte
// loads), the next chunk held in registers during the products.  With GL
// the loads go through global-address-space pointers, rows past the tile's
// end and K past naux read the tensor's first entries and are replaced by
// zeros (no branches around the loads); without, they are skipped.  Which
// is faster depends on the tile shape (launch_r4g).  Factor-tensor rows are
// naux values with naux a multiple of 4 (the factors are zero-padded), so
// every row starts 16-byte aligned.  Tiles are dealt to the eight XCDs in
// contiguous chunks, so the tiles of one product, which share rows, meet in
// one L2.
template <typename T>
struct R4Vec;
template <>
struct R4Vec<float> {
    typedef float t __attribute__((ext_vector_type(4)));
};
template <>
struct R4Vec<double> {
    typedef double t __attribute__((ext_vector_type(2)));
};

// The tile's row and column offsets into the factor tensors (ra, rb; -1 past
// the product's edge) and into the output (ro + co; mro + mco for the mirror
// of mode 3), in LDS.
template <int BM, int BN>
__device__ inline void r4_tile_index(const R4Gemm &g, const R4GTile &tl, int nm, int naux, long long *ra,
                                     long long *rb, long long *ro, long long *co, long long *mro, long long *mco)
{
    const long long nm2 = (long long)nm * nm;
    const int mloc = min(BM, g.M - tl.m0), nloc = min(BN, g.N - tl.n0);
    for (int t = threadIdx.x; t < BM + BN; t += blockDim.x) {
        int u = 0, v = 0;
        if (t < BM) {
            const bool ok = t < mloc;
            if (ok) axis_at(g.row, tl.m0 + t, u, v);
            ra[t] = ok ? ((long long)u * nm + v) * naux : -1;
            ro[t] = g.mode == 0 || g.mode == 3 ? ((long long)u * nm + v) * nm2
                    : g.mode == 1              ? (long long)u * nm * nm2 + (long long)v * nm
                                               : (long long)u * nm + v;
            mro[t] = (long long)u * nm + v;
        } else {
            const int j = t - BM;
            const bool ok = j < nloc;
            if (ok) axis_at(g.col, tl.n0 + j, u, v);
            rb[j] = ok ? ((long long)u * nm + v) * naux : -1;
            co[j] = g.mode == 0 || g.mode == 3 ? (long long)u * nm + v
                    : g.mode == 1              ? (long long)u * nm2 + v
                                               : ((long long)u * nm + v) * nm2;
            mco[j] = ((long long)u * nm + v) * nm2;
        }
    }
}

// Column-band width of the tile order (see r4_tile_at).
constexpr int R4_BAND = 8;

// Tile t of a launch over products gs: tstart[g] (nprod + 1 entries) counts
// the tiles before product g; within a product the tiles go in column bands
// of R4_BAND tiles, every row tile of a band before the next band, so an
// XCD's run of tiles (its resident workgroups) shares a few A and B row
// blocks in L2.  The host passes only tstart; no tile list is built.
__device__ inline R4GTile r4_tile_at(const long long *__restrict__ tstart, int nprod, const R4Gemm *__restrict__ gs,
                                     long long t, int bm, int bn)
{
    int lo = 0, hi = nprod - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tstart[mid] <= t) lo = mid; else hi = mid - 1;
    }
    const R4Gemm &g = gs[lo];
    const long long l = t - tstart[lo];
    const int mt = (g.M + bm - 1) / bm, nt = (g.N + bn - 1) / bn, full = nt / R4_BAND;
    const long long fb = (long long)R4_BAND * mt;
    int m, n;
    if (l < full * fb) {
        const long long rem = l % fb;
        m = (int)(rem / R4_BAND);
        n = (int)(l / fb) * R4_BAND + (int)(rem % R4_BAND);
    } else {
        const int w = nt - full * R4_BAND;
        const long long rem = l - full * fb;
        m = (int)(rem / w);
        n = full * R4_BAND + (int)(rem % w);
    }
    return R4GTile{lo, m * bm, n * bn, 0};
}

// Writes a wave's TM x TN 16 x 16 result tiles (register r4 of lane group lk
// holds row L::row(lk, r4), lane & 15 the column) by the product's mode.
template <typename T, typename L, typename Acc, int TM, int TN, int WM, int WN>
__device__ inline void r4_store(const R4Gemm &g, const Acc (&acc)[TM][TN], int wm, int wn, int mloc, int nloc,
                                const long long *ro, const long long *co, const long long *mro, const long long *mco)
{
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    T *out = (T *)g.out;
#pragma unroll
    for (int b = 0; b < TN; ++b) {
        const int j = wn * WN + b * 16 + lr;
        if (j >= nloc) continue;
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int i = wm * WM + a * 16 + L::row(lk, r4);
                if (i >= mloc) continue;
                T *o = out + ro[i] + co[j];
                if (g.mode == 1) {
                    *o -= acc[a][b][r4];
                } else {
                    *o = acc[a][b][r4];
                    if (g.mode == 3) out[mco[j] + mro[i]] = acc[a][b][r4];
                }
            }
    }
}

// MFMA step kk of lane group lk takes k = lk * (KC/4) + kk (the dot product
// does not depend on the order of k), so each lane's fragments of a chunk
// are contiguous in LDS and read 16 bytes at a time.
template <typename T, int BM, int BN, int KC, int OCC, bool GL>
__global__ __launch_bounds__(256, OCC) void k_r4_gemm(const R4Gemm *__restrict__ gs, const long long *__restrict__ tstart,
                                                      int nprod, long long ntiles, int nm, int naux)
{
    typedef typename R4Vec<T>::t VT;
    typedef const T __attribute__((address_space(1))) *gptr;
    constexpr int VE = 16 / sizeof(T), KV = KC / VE, LDK = KC + VE, NT = 256;
    constexpr int PER = (BM + BN) * KV / NT;  // 16-byte vectors per thread and chunk
    static_assert((BM + BN) * KV % NT == 0 && KC % (4 * VE) == 0, "tile shape");
    constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16, KQ = KC / 4;
    __shared__ __attribute__((aligned(16))) T As[BM][LDK];
    __shared__ __attribute__((aligned(16))) T Bs[BN][LDK];
    __shared__ long long ra[BM], rb[BN], ro[BM], co[BN], mro[BM], mco[BN];
    using M = R4Mfma<T>;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int wm = wave >> 1, wn = wave & 1;
    const int gper = gridDim.x / 8, chunk = blockIdx.x % 8, lidx = blockIdx.x / 8;
    const long long per = (ntiles + 7) / 8;
    const long long t0 = chunk * per, t1 = min(ntiles, t0 + per);
    for (long long it = t0 + lidx; it < t1; it += gper) {
        const R4GTile tl = r4_tile_at(tstart, nprod, gs, it, BM, BN);
        const R4Gemm g = gs[tl.g];
        const int mloc = min(BM, g.M - tl.m0), nloc = min(BN, g.N - tl.n0);
        __syncthreads();  // the previous tile's indices are no longer read
        r4_tile_index<BM, BN>(g, tl, nm, naux, ra, rb, ro, co, mro, mco);
        __syncthreads();
        const T *A = (const T *)g.A;
        const T *B = (const T *)g.B;
        // each thread loads the same rows and k offsets in every chunk
        gptr src[PER];
        bool okr[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = tid + j * NT, r = e / KV, kv = e % KV;
            const long long b = r < BM ? ra[r] : rb[r - BM];
            okr[j] = b >= 0;
            src[j] = (gptr)((r < BM ? A : B) + (b < 0 ? 0 : b) + kv * VE);
        }
        VT buf[PER];
        auto gload = [&](int k0) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int kv = (tid + j * NT) % KV;
                const bool ok = okr[j] && k0 + kv * VE < naux;
                if constexpr (GL) {
                    const VT v = *reinterpret_cast<const VT __attribute__((address_space(1))) *>(src[j] + (ok ? k0 : 0));
                    buf[j] = ok ? v : VT{};
                } else {
                    buf[j] = ok ? *reinterpret_cast<const VT *>((const T *)src[j] + k0) : VT{};
                }
            }
        };
        typename M::acc acc[TM][TN];
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b) acc[a][b] = typename M::acc{0, 0, 0, 0};
        gload(0);
        for (int k0 = 0; k0 < naux; k0 += KC) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int e = tid + j * NT, r = e / KV, kv = e % KV;
                T *dst = r < BM ? &As[r][kv * VE] : &Bs[r - BM][kv * VE];
                *reinterpret_cast<VT *>(dst) = buf[j];
            }
            __syncthreads();
            if (k0 + KC < naux) gload(k0 + KC);  // next chunk in flight during the products
#pragma unroll
            for (int q0 = 0; q0 < KQ; q0 += VE) {
                VT af[TM], bf[TN];
#pragma unroll
                for (int a = 0; a < TM; ++a)
                    af[a] = *reinterpret_cast<const VT *>(&As[wm * WM + a * 16 + lr][lk * KQ + q0]);
#pragma unroll
                for (int b = 0; b < TN; ++b)
                    bf[b] = *reinterpret_cast<const VT *>(&Bs[wn * WN + b * 16 + lr][lk * KQ + q0]);
#pragma unroll
                for (int e = 0; e < VE; ++e)
#pragma unroll
                    for (int a = 0; a < TM; ++a)
#pragma unroll
                        for (int b = 0; b < TN; ++b) acc[a][b] = M::mma(af[a][e], bf[b][e], acc[a][b]);
            }
        }
        r4_store<T, M, typename M::acc, TM, TN, WM, WN>(g, acc, wm, wn, mloc, nloc, ro, co, mro, mco);
    }
}

// Staging map of the BF16-split kernels: element e (a 16-byte piece of
// four FP32 values, KV = 8 pieces per row of a K = 32 chunk) goes to row r
// and piece kv so that each wave's 64 lanes hold 16 rows x 4 pieces, lane
// pairs on adjacent pieces of one row: the 8-byte writes of one plane then
// cover 16-byte slots of consecutive rows, which fall on distinct LDS banks
// at a row pitch of 80 bytes (as do the 16-byte fragment reads).
__device__ inline void bf3_piece(int e, int &r, int &kv)
{
    const int g = e >> 6, l = e & 63;
    r = (g >> 1) * 16 + ((l >> 1) & 15);
    kv = (g & 1) * 4 + 2 * (l >> 5) + (l & 1);
}

// The same FP32 products on BF16 MFMA: each FP32 operand v is split into
// three BF16 parts h + m + l (h = bf16(v), m = bf16(v - h), l = bf16(v - h -
// m), round to nearest even), which hold its 24-bit significand, as the
// chunk is written to LDS (three planes [row][k]); a chunk of K = 32 is one
// v_mfma_f32_16x16x32_bf16 per pair of parts, and the six pairs whose
// product is above 2^-24 of h_a h_b (hh, hm, mh, hl, lh, mm) accumulate in
// FP32, smallest first.  The dropped ml, lm and ll terms are below FP32
// rounding, so the result carries FP32-level error.  Lane group lk holds
// k = 8 lk .. 8 lk + 7 of the chunk (one 16-byte LDS read per plane).
template <int BM, int BN, int OCC, bool GL>
__global__ __launch_bounds__(256, OCC) void k_r4_bf3(const R4Gemm *__restrict__ gs, const long long *__restrict__ tstart,
                                                     int nprod, long long ntiles, int nm, int naux)
{
    typedef float gf4 __attribute__((ext_vector_type(4)));
    typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
    typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
    typedef const float __attribute__((address_space(1))) *gptr;
    constexpr int KC = 32, VE = 4, KV = KC / VE, LDK = KC + 8, NT = 256, ROWS = BM + BN;
    constexpr int PER = ROWS * KV / NT;
    static_assert(ROWS * KV % NT == 0, "tile shape");
    static_assert(KV == 8 && ROWS % 16 == 0, "bf3_piece layout");
    constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
    __shared__ __attribute__((aligned(16))) __bf16 Xs[3][ROWS][LDK];  // planes h, m, l; rows A then B
    __shared__ long long ra[BM], rb[BN], ro[BM], co[BN], mro[BM], mco[BN];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int wm = wave >> 1, wn = wave & 1;
    const int gper = gridDim.x / 8, chunk = blockIdx.x % 8, lidx = blockIdx.x / 8;
    const long long per = (ntiles + 7) / 8;
    const long long t0 = chunk * per, t1 = min(ntiles, t0 + per);
    for (long long it = t0 + lidx; it < t1; it += gper) {
        const R4GTile tl = r4_tile_at(tstart, nprod, gs, it, BM, BN);
        const R4Gemm g = gs[tl.g];
        const int mloc = min(BM, g.M - tl.m0), nloc = min(BN, g.N - tl.n0);
        __syncthreads();
        r4_tile_index<BM, BN>(g, tl, nm, naux, ra, rb, ro, co, mro, mco);
        __syncthreads();
        const float *A = (const float *)g.A;
        const float *B = (const float *)g.B;
        gptr src[PER];
        bool okr[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            int r, kv;
            bf3_piece(tid + j * NT, r, kv);
            const long long b = r < BM ? ra[r] : rb[r - BM];
            okr[j] = b >= 0;
            src[j] = (gptr)((r < BM ? A : B) + (b < 0 ? 0 : b) + kv * VE);
        }
        gf4 buf[PER];
        auto gload = [&](int k0) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                int r, kv;
                bf3_piece(tid + j * NT, r, kv);
                const bool ok = okr[j] && k0 + kv * VE < naux;
                if constexpr (GL) {
                    const gf4 v = *reinterpret_cast<const gf4 __attribute__((address_space(1))) *>(src[j] + (ok ? k0 : 0));
                    buf[j] = ok ? v : gf4{0, 0, 0, 0};
                } else {
                    buf[j] = ok ? *reinterpret_cast<const gf4 *>((const float *)src[j] + k0) : gf4{0, 0, 0, 0};
                }
            }
        };
        r4_f4 acc[TM][TN];
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b) acc[a][b] = r4_f4{0, 0, 0, 0};
        gload(0);
        for (int k0 = 0; k0 < naux; k0 += KC) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                int r, kv;
                bf3_piece(tid + j * NT, r, kv);
                bf4 h, m, l;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float x = buf[j][i];
                    const __bf16 hh = (__bf16)x;
                    const float rem = x - (float)hh;
                    const __bf16 mm = (__bf16)rem;
                    h[i] = hh;
                    m[i] = mm;
                    l[i] = (__bf16)(rem - (float)mm);
                }
                *reinterpret_cast<bf4 *>(&Xs[0][r][kv * VE]) = h;
                *reinterpret_cast<bf4 *>(&Xs[1][r][kv * VE]) = m;
                *reinterpret_cast<bf4 *>(&Xs[2][r][kv * VE]) = l;
            }
            __syncthreads();
            if (k0 + KC < naux) gload(k0 + KC);
            bf8 ah[TM], am[TM], al[TM];
#pragma unroll
            for (int a = 0; a < TM; ++a) {
                const int r = wm * WM + a * 16 + lr;
                ah[a] = *reinterpret_cast<const bf8 *>(&Xs[0][r][lk * 8]);
                am[a] = *reinterpret_cast<const bf8 *>(&Xs[1][r][lk * 8]);
                al[a] = *reinterpret_cast<const bf8 *>(&Xs[2][r][lk * 8]);
            }
#pragma unroll
            for (int b = 0; b < TN; ++b) {
                const int r = BM + wn * WN + b * 16 + lr;
                const bf8 bh = *reinterpret_cast<const bf8 *>(&Xs[0][r][lk * 8]);
                const bf8 bm = *reinterpret_cast<const bf8 *>(&Xs[1][r][lk * 8]);
                const bf8 bl = *reinterpret_cast<const bf8 *>(&Xs[2][r][lk * 8]);
                // each part pair over the TM row tiles in turn, so that
                // consecutive MFMAs write different accumulators
#pragma unroll
                for (int a = 0; a < TM; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh, acc[a][b], 0, 0, 0);
#pragma unroll
                for (int a = 0; a < TM; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl, acc[a][b], 0, 0, 0);
#pragma unroll
                for (int a = 0; a < TM; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[a], bm, acc[a][b], 0, 0, 0);
#pragma unroll
                for (int a = 0; a < TM; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[a], bh, acc[a][b], 0, 0, 0);
#pragma unroll
                for (int a = 0; a < TM; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bm, acc[a][b], 0, 0, 0);
#pragma unroll
                for (int a = 0; a < TM; ++a) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh, acc[a][b], 0, 0, 0);
            }
        }
        r4_store<float, R4Mfma<float>, r4_f4, TM, TN, WM, WN>(g, acc, wm, wn, mloc, nloc, ro, co, mro, mco);
    }
}

// out[b][j][i] = in[b][i][j] for rows x cols matrices (rounded to To).
template <typename To>
__global__ void k_transpose_batched(const double *const *__restrict__ in, To *const *__restrict__ out, int rows,
                                    int cols)
{
    const double *a = in[blockIdx.y];
    To *o = out[blockIdx.y];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < rows * cols; e += gridDim.x * blockDim.x) {
        const int i = e / cols, j = e % cols;
        o[(size_t)j * rows + i] = (To)a[e];
    }
}

// One product of the factor transforms: C[i][j] = sum_k A[i][k] B[j][k].
struct FacGemm {
    const void *A, *B;
    void *C;
};

// Factor-transform products: few rows (M = nmo, one tile of 16 TM rows per
// row block, blockIdx.z), few k (K = n_bas) and many columns (naux n_bas or
// naux nmo), batched over blockIdx.y.  64 TN columns per workgroup (wave w:
// 16 TN of them), K through LDS 16 at a time with 16-byte loads (lda and ldb
// multiples of 4, 16-byte aligned rows), the next chunk held in registers;
// lane group lk's MFMA steps take k = lk * 4 + kk, so its fragments are
// contiguous in LDS.  Every B row is read once; the A tile is shared by the
// workgroups of a product through L2.
template <typename T, int TM, int TN>
__global__ __launch_bounds__(256) void k_fac_gemm(const FacGemm *__restrict__ descs, int M, int N, int K, int lda,
                                                  int ldb, int ldc)
{
    typedef typename R4Vec<T>::t VT;
    constexpr int KC = 16, VE = 16 / sizeof(T), KV = KC / VE, LDK = KC + VE, BM = 16 * TM, BN = 64 * TN, KQ = KC / 4;
    constexpr int NV = (BM + BN) * KV, PER = (NV + 255) / 256;
    static_assert(KQ % VE == 0, "chunk");
    __shared__ __attribute__((aligned(16))) T As[BM][LDK];
    __shared__ __attribute__((aligned(16))) T Bs[BN][LDK];
    using Mf = R4Mfma<T>;
    const FacGemm g = descs[blockIdx.y];
    const T *__restrict__ A = (const T *)g.A;
    const T *__restrict__ B = (const T *)g.B;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int n0 = blockIdx.x * BN, m0 = blockIdx.z * BM, mloc = min(BM, M - m0);
    const T *src[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int e = tid + j * 256, r = e / KV, kv = e % KV;
        const T *q = nullptr;
        if (e < NV) {
            if (r < BM) {
                if (r < mloc) q = A + (size_t)(m0 + r) * lda + kv * VE;
            } else if (n0 + r - BM < N) {
                q = B + (size_t)(n0 + r - BM) * ldb + kv * VE;
            }
        }
        src[j] = q;
    }
    VT buf[PER];
    auto gload = [&](int k0) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int kv = (tid + j * 256) % KV;
            buf[j] = src[j] && k0 + kv * VE < K ? *reinterpret_cast<const VT *>(src[j] + k0) : VT{};
        }
    };
    typename Mf::acc acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = typename Mf::acc{0, 0, 0, 0};
    gload(0);
    for (int k0 = 0; k0 < K; k0 += KC) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = tid + j * 256, r = e / KV, kv = e % KV;
            if (e < NV) *reinterpret_cast<VT *>(r < BM ? &As[r][kv * VE] : &Bs[r - BM][kv * VE]) = buf[j];
        }
        __syncthreads();
        if (k0 + KC < K) gload(k0 + KC);  // next chunk in flight during the products
#pragma unroll
        for (int q0 = 0; q0 < KQ; q0 += VE) {
            VT bf[TN];
#pragma unroll
            for (int b = 0; b < TN; ++b)
                bf[b] = *reinterpret_cast<const VT *>(&Bs[wave * 16 * TN + b * 16 + lr][lk * KQ + q0]);
#pragma unroll
            for (int a = 0; a < TM; ++a) {
                const VT af = *reinterpret_cast<const VT *>(&As[a * 16 + lr][lk * KQ + q0]);
#pragma unroll
                for (int e = 0; e < VE; ++e)
#pragma unroll
                    for (int b = 0; b < TN; ++b)
                        acc[a][b] = Mf::mma(reinterpret_cast<const T *>(&af)[e], reinterpret_cast<const T *>(&bf[b])[e],
                                            acc[a][b]);
            }
        }
    }
    T *C = (T *)g.C;
#pragma unroll
    for (int b = 0; b < TN; ++b) {
        const int col = n0 + wave * 16 * TN + b * 16 + lr;
        if (col >= N) continue;
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = a * 16 + Mf::row(lk, r);
                if (row < mloc) C[(size_t)(m0 + row) * ldc + col] = acc[a][b][r];
            }
    }
}

// Largest |sum_K L[K,a,b] L[K,c,d] - eri[a,b,c,d]| and |L[K,a,b] - L[K,b,a]|
// over a fixed pseudo-random sample of index tuples (as the bits of a
// nonnegative double, which order like the values).
__global__ void k_factor_check(const double *__restrict__ L, const double *__restrict__ eri, int n, int naux,
                               int nsample, unsigned long long *__restrict__ out)
{
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nsample; s += gridDim.x * blockDim.x) {
        unsigned long long h = 0x9E3779B97F4A7C15ull * (unsigned long long)(s + 1);
        h ^= h >> 29;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 32;
        const int a = (int)(h % n), b = (int)((h >> 8) % n), c = (int)((h >> 16) % n), d = (int)((h >> 24) % n);
        const size_t nn = (size_t)n * n;
        double v = 0.0;
        for (int k = 0; k < naux; ++k) v += L[k * nn + (size_t)a * n + b] * L[k * nn + (size_t)c * n + d];
        double dev = fabs(v - eri[((size_t)a * n + b) * nn + (size_t)c * n + d]);
        const int k = s % naux;
        dev = fmax(dev, fabs(L[k * nn + (size_t)a * n + b] - L[k * nn + (size_t)b * n + a]));
        if (!(dev == dev)) dev = INFINITY;
        atomicMax(out, (unsigned long long)__double_as_longlong(dev));
    }
}

}  // namespace nw

namespace nwi {
using namespace nw;

namespace {

// Tensor classes whose blocks share one coverage map: alpha J slots, beta J
// slots, II blocks.
enum { CL_JA = 0, CL_JB = 1, CL_II = 2 };

// Factor tensors [nm][nm][naux] of one pair, keyed fkey(kind, spin, a, b):
// kind FA = cx_a^T L_K xc_b, FE = cx_a^T L_K cx_b, FF = xc_a^T L_K xc_b.
enum { FA = 0, FE = 1, FF = 2 };
inline int fkey(int kind, int sp, int a, int b) { return ((kind * 2 + sp) * 2 + a) * 2 + b; }

struct CholPair {
    bool lazy = false;
    std::vector<uint8_t> cov[3];    // nm x nm: block (p, r) built
    std::vector<uint8_t> kkcov[3];  // nm x nm: ket-particle column pairs (q, s) the built blocks hold
    std::map<int, double *> fac;  // FP64 factor tensors
    std::map<int, float *> fac32; // their FP32 copies (pairs built in reduced precision)
};

}  // namespace

struct CholState {
    double *d_L = nullptr;  // [naux][n][n]
    int naux = 0;           // factors held: the given ones, zero-padded to a multiple of 4
    int naux_in = 0;        // factors given
    std::vector<CholPair> pairs;  // x * nref + w, for the context's current build
    std::vector<std::pair<char *, size_t>> chunks;  // factor-tensor arena, reset by every build
    size_t ci = 0, used = 0;
    DBuf t1, trans, ptrs, items, tensors, t1f, trans32;
    float *d_L32 = nullptr;  // FP32 copy of d_L, made on first FP32 use
    double ms_blocks = 0.0;
    // last chol_ensure: device ms of the factor products, rectangle products,
    // single blocks and exchange products; their algorithmic flops; single blocks
    double stats[10] = {};
    // the block products by launch class (FP64 large tiles, FP64 narrow
    // tiles, FP32 pairs' large tiles (FP32 or BF16 MFMA), FP32 pairs' narrow
    // tiles): device ms, then algorithmic flops
    double kstats[8] = {};
    std::vector<hipEvent_t> evs;  // stage boundaries of the last build_blocks
    std::vector<int> ev_stage;    // stage, plus 16 (launch class + 1) for block products
    bool pending = false;  // stage times not yet read from the events
    hipEvent_t mark(int stage)
    {
        if (evs.size() == ev_stage.size()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            evs.push_back(e);
        }
        ev_stage.push_back(stage);
        return evs[ev_stage.size() - 1];
    }

    char *arena(size_t bytes)
    {
        bytes = (bytes + 255) & ~size_t(255);
        while (ci < chunks.size() && used + bytes > chunks[ci].second) {
            ++ci;
            used = 0;
        }
        if (ci == chunks.size()) {
            const size_t sz = std::max<size_t>(bytes, size_t(1) << 30);
            chunks.push_back({dalloc<char>(sz), sz});
            used = 0;
        }
        char *p = chunks[ci].first + used;
        used += bytes;
        return p;
    }
    void reset_arena()
    {
        ci = 0;
        used = 0;
    }
    ~CholState()
    {
        for (hipEvent_t e : evs) (void)hipEventDestroy(e);
        if (d_L) (void)hipFree(d_L);
        if (d_L32) (void)hipFree(d_L32);
        for (auto &ch : chunks) (void)hipFree(ch.first);
    }
};

void destroy_chol(nw_ctx *c)
{
    delete c->chol;
    c->chol = nullptr;
}

bool chol_route(nw_ctx *c) { return c->chol && c->chol->d_L && c->chol->naux > 0; }

void chol_on_build(nw_ctx *c, const std::vector<int> &keys, const std::vector<uint8_t> &lazy)
{
    CholState *s = c->chol;
    if (!s) return;
    s->pairs.assign((size_t)c->nref * c->nref, CholPair{});
    s->reset_arena();
    const size_t nm2 = (size_t)c->nmo * c->nmo;
    for (size_t k = 0; k < keys.size(); ++k) {
        if (!lazy[k]) continue;
        CholPair &cp = s->pairs[keys[k]];
        cp.lazy = true;
        for (auto &v : cp.cov) v.assign(nm2, 0);
        for (auto &v : cp.kkcov) v.assign(nm2, 0);
    }
}

void chol_forget(nw_ctx *c)
{
    if (c->chol) {
        c->chol->pairs.clear();
        c->chol->reset_arena();
    }
}

namespace {

int r4_ncu()
{
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        ncu = std::max(1, ncu);
    }
    return ncu;
}

// Workgroups of a product launch: up to eight per CU, in multiples of eight
// (the XCD chunks); each runs its chunk's tiles in a grid-stride loop.
int r4_grid(long long nitems) { return (int)std::max(8ll, std::min((nitems + 7) / 8 * 8, 8ll * r4_ncu())); }

size_t r4_tiles(const std::vector<R4Gemm> &gs, int bm, int bn)
{
    size_t n = 0;
    for (const R4Gemm &g : gs) n += (size_t)((g.M + bm - 1) / bm) * ((g.N + bn - 1) / bn);
    return n;
}

// Tile offsets of each product (r4_tile_at maps a tile index to its product
// and position on the device).
std::vector<long long> r4_tstart(const std::vector<R4Gemm> &gs, int bm, int bn)
{
    std::vector<long long> ts(gs.size() + 1, 0);
    for (size_t i = 0; i < gs.size(); ++i)
        ts[i + 1] = ts[i] + (long long)((gs[i].M + bm - 1) / bm) * ((gs[i].N + bn - 1) / bn);
    return ts;
}

template <typename T, int BM, int BN, int KC, int OCC, bool GL>
void launch_r4_shape(nw_ctx *c, Uploader &up, const R4Gemm *dg, const std::vector<R4Gemm> &gs, int naux)
{
    const std::vector<long long> ts = r4_tstart(gs, BM, BN);
    const long long nt = ts.back();
    if (nt == 0) return;
    hipLaunchKernelGGL((k_r4_gemm<T, BM, BN, KC, OCC, GL>), dim3(r4_grid(nt)), dim3(256), 0, c->st, dg, up.put(ts),
                       (int)gs.size(), nt, c->nmo, naux);
}

template <int BM, int BN, int OCC, bool GL>
void launch_r4_bf3_shape(nw_ctx *c, Uploader &up, const R4Gemm *dg, const std::vector<R4Gemm> &gs, int naux)
{
    const std::vector<long long> ts = r4_tstart(gs, BM, BN);
    const long long nt = ts.back();
    if (nt == 0) return;
    hipLaunchKernelGGL((k_r4_bf3<BM, BN, OCC, GL>), dim3(r4_grid(nt)), dim3(256), 0, c->st, dg, up.put(ts),
                       (int)gs.size(), nt, c->nmo, naux);
}

// FP32 pairs of an NW_PREC_BF16X3 context take their block products on BF16
// MFMA (k_r4_bf3), those of NW_PREC_MIXED on FP32 MFMA.
bool r4_bf3(const nw_ctx *c) { return c->precision == NW_PREC_BF16X3; }

// One launch over a list of products: 128 x 128 output tiles in FP32 and
// 128 x 64 in FP64 (measured 1.3-1.6x the rate of 64 x 64 tiles on the
// n_bas = 100 products) when there are enough of them to fill the chip
// (four per CU), else 64 x 64 tiles.  The BF16 form runs 128 x 128 tiles at
// two workgroups per CU (72 KiB of LDS each): 4.95 ms against 6.4 ms of
// FP32 MFMA for the C2 step's products.  These tile shapes load through
// generic pointers (GL = false): global-address loads measured slower on
// them (FP64 128 x 64: 16.0 against 12.7 ms per step), faster or equal on
// the narrow tiles.
template <typename T>
void launch_r4g(nw_ctx *c, Uploader &up, const std::vector<R4Gemm> &gs, int naux)
{
    if (gs.empty()) return;
    const R4Gemm *dg = up.put(gs);
    constexpr bool f32 = sizeof(T) == 4;
    const bool big = r4_tiles(gs, 128, f32 ? 128 : 64) >= (size_t)4 * r4_ncu();
    if constexpr (f32) {
        if (r4_bf3(c)) {
            if (big) launch_r4_bf3_shape<128, 128, 2, false>(c, up, dg, gs, naux);
            else launch_r4_bf3_shape<64, 64, 2, false>(c, up, dg, gs, naux);
            return;
        }
    }
    if (big) {
        if constexpr (f32) launch_r4_shape<float, 128, 128, 32, 1, false>(c, up, dg, gs, naux);
        else launch_r4_shape<double, 128, 64, 16, 2, false>(c, up, dg, gs, naux);
    } else {
        launch_r4_shape<T, 64, 64, 32, 2, false>(c, up, dg, gs, naux);
    }
}

// Products with a narrow side (at most a few dozen columns: ket-particle
// pairs) in 128 x 32 tiles, on FP32 (or FP64) MFMA in either reduced
// precision: the BF16 form measured no faster on them (1.60 ms against
// 1.2-1.6 ms per C2 step).
void launch_r4g_narrow(nw_ctx *c, Uploader &up, const std::vector<R4Gemm> &gs, int naux, bool f32)
{
    if (gs.empty()) return;
    const R4Gemm *dg = up.put(gs);
    if (f32) launch_r4_shape<float, 128, 32, 32, 2, true>(c, up, dg, gs, naux);
    else launch_r4_shape<double, 128, 32, 16, 2, true>(c, up, dg, gs, naux);
}

// Row index sets of one parent's determinants per spin, as bra (rows
// part - nocc) and as ket (rows nvirt + hole), and the row pairs that occur
// within one determinant (both orders; the mixed-spin pairs as (alpha, beta)).
struct ParentRows {
    bool present = false;
    std::vector<uint8_t> bra[2], ket[2];            // nm
    std::vector<uint8_t> bra2[2], ket2[2];          // nm x nm, same-spin pairs of one determinant
    std::vector<uint8_t> braab, ketab;              // nm x nm, (alpha row, beta row) of one determinant
    std::vector<uint8_t> kcol2[2], kcolab;          // nm x nm, ket particle (column label) pairs of one determinant
    // the same sets as lists (each entry once, in first-seen order)
    std::vector<int> bral[2], ketl[2], bra2l[2], ket2l[2], braabl, ketabl, kcol2l[2], kcolabl;
};

std::vector<ParentRows> parent_rows(nw_ctx *c, const std::vector<PackedDet> &h)
{
    const int nm = c->nmo, nref = c->nref;
    const size_t nm2 = (size_t)nm * nm;
    std::vector<ParentRows> pr(nref);
    for (const PackedDet &d : h) {
        if (d.parent >= nref) continue;
        ParentRows &r = pr[d.parent];
        if (!r.present) {
            r.present = true;
            for (int sp = 0; sp < 2; ++sp) {
                r.bra[sp].assign(nm, 0);
                r.ket[sp].assign(nm, 0);
                r.bra2[sp].assign(nm2, 0);
                r.ket2[sp].assign(nm2, 0);
            }
            r.braab.assign(nm2, 0);
            r.ketab.assign(nm2, 0);
            r.kcol2[0].assign(nm2, 0);
            r.kcol2[1].assign(nm2, 0);
            r.kcolab.assign(nm2, 0);
        }
        auto add = [](std::vector<uint8_t> &set, std::vector<int> &list, size_t k) {
            if (!set[k]) {
                set[k] = 1;
                list.push_back((int)k);
            }
        };
        int br[2][4], kr[2][4], kc[2][4], rk[2] = {d.ra, d.rb};
        for (int sp = 0; sp < 2; ++sp) {
            const uint8_t *ix = sp ? d.ib : d.ia;
            const int no = c->nocc[sp], nv = nm - no;
            for (int i = 0; i < rk[sp]; ++i) {
                br[sp][i] = std::min(nm - 1, std::max(0, (int)ix[4 + i] - no));
                kr[sp][i] = std::min(nm - 1, std::max(0, nv + (int)ix[i]));
                kc[sp][i] = std::min(nm - 1, (int)ix[4 + i]);  // column label of a ket particle
                add(r.bra[sp], r.bral[sp], br[sp][i]);
                add(r.ket[sp], r.ketl[sp], kr[sp][i]);
            }
            for (int i = 0; i < rk[sp]; ++i)
                for (int j = 0; j < rk[sp]; ++j)
                    if (i != j) {
                        add(r.bra2[sp], r.bra2l[sp], (size_t)br[sp][i] * nm + br[sp][j]);
                        add(r.ket2[sp], r.ket2l[sp], (size_t)kr[sp][i] * nm + kr[sp][j]);
                        add(r.kcol2[sp], r.kcol2l[sp], (size_t)kc[sp][i] * nm + kc[sp][j]);
                    }
        }
        for (int i = 0; i < rk[0]; ++i)
            for (int j = 0; j < rk[1]; ++j) {
                add(r.braab, r.braabl, (size_t)br[0][i] * nm + br[1][j]);
                add(r.ketab, r.ketabl, (size_t)kr[0][i] * nm + kr[1][j]);
                add(r.kcolab, r.kcolabl, (size_t)kc[0][i] * nm + kc[1][j]);
            }
    }
    return pr;
}

// Blocks (p, r) of pair (x, w) that its determinant pairs read, per class:
// J_sigma at (rows[eta], rows[xi]) -- both bra rows of one determinant, a
// bra row then a ket row, or both ket rows of one determinant; II at
// (alpha row, beta row) -- both of one bra determinant, bra alpha x ket
// beta, ket alpha x bra beta, or both of one ket determinant.  The
// evaluators of pairs with zero-overlap orbitals (both = true) read J rows
// in either order (the distribution sums and the slot swaps of
// preparehamiltonianoverlap.rs:20550-20966).  Within a block they read the
// columns (q, s) of the contraction labels: a bra hole (q < nocc or
// s < nocc, dense) or two ket particles of one ket determinant (kk, sparse).

// Work of one pair: per class, the blocks to (re)build and the union of the
// ket-particle column pairs they must hold (a later call whose kets add
// pairs rebuilds the class's built blocks with the union).
struct ClassWork {
    std::vector<int> blocks;   // p * nm + r: every block this call builds
    std::vector<uint8_t> kk;   // nm x nm: column pairs (q, s) to hold
    std::vector<std::pair<std::vector<int>, std::vector<int>>> rects;  // (p rows, r rows)
    std::vector<int> single;   // blocks outside the rectangles
};
struct PairWork {
    int key;
    ClassWork cls[3];
};

// Lazy pairs touched by h and their work.
bool plan_missing(nw_ctx *c, const std::vector<PackedDet> &h, std::vector<PairWork> &work,
                  const std::vector<uint8_t> *reads)
{
    CholState *s = c->chol;
    if (!s || s->pairs.empty()) return false;
    bool any_lazy = false;
    for (const CholPair &cp : s->pairs) any_lazy |= cp.lazy;
    if (!any_lazy) return false;
    const int nref = c->nref, nm = c->nmo;
    const size_t nm2 = (size_t)nm * nm;
    std::vector<ParentRows> pr = parent_rows(c, h);
    std::vector<int> keys;
    for (int x = 0; x < nref; ++x)
        for (int w = 0; w < nref; ++w) {
            const int key = x * nref + w;
            if (pr[x].present && pr[w].present && s->pairs[key].lazy && c->pinfo[key].built &&
                (!reads || (*reads)[key]))
                keys.push_back(key);
        }
    if (keys.empty()) return false;
    // the pairs' plans are independent: a few host threads, each with its own scratch
    std::vector<PairWork> res(keys.size());
    std::vector<uint8_t> has(keys.size(), 0);
    auto plan_range = [&](size_t t0, size_t step) {
        std::vector<unsigned> stamp(nm2, 0);  // block k seen in the current class when stamp[k] == stamp_gen
        unsigned stamp_gen = 0;
        std::vector<uint8_t> todo(nm2, 0), done(nm2, 0);  // cleared entry by entry after each class
        std::vector<int> pcount(nm + 1), sorted;
        for (size_t ki = t0; ki < keys.size(); ki += step) {
            const int key = keys[ki], x = key / nref, w = key % nref;
            const CholPair &cp = s->pairs[key];
            const PairInfo &pi = c->pinfo[key];
            const bool both = pi.ma + pi.mb > 0;
            PairWork pw;
            pw.key = key;
            bool any = false;
            for (int cl = 0; cl < 3; ++cl) {
                const std::vector<int> &need_kk = cl < 2 ? pr[w].kcol2l[cl] : pr[w].kcolabl;
                const std::vector<uint8_t> &cov = cp.cov[cl];
                bool grow = false;
                for (int k : need_kk) grow = grow || !cp.kkcov[cl][k];
                ClassWork &cw = pw.cls[cl];
                // candidates: the blocks the determinant pairs read (above), from the
                // parents' row lists; with new ket-particle pairs every built block is
                // rebuilt as well
                ++stamp_gen;
                auto see = [&](int p, int r) {
                    const size_t k = (size_t)p * nm + r;
                    if (stamp[k] == stamp_gen) return;
                    stamp[k] = stamp_gen;
                    if (!cov[k] || grow) cw.blocks.push_back((int)k);
                };
                const ParentRows &X = pr[x], &W = pr[w];
                if (cl < 2) {
                    for (int k : X.bra2l[cl]) see(k / nm, k % nm);
                    for (int k : W.ket2l[cl]) see(k / nm, k % nm);
                    for (int p : X.bral[cl])
                        for (int r : W.ketl[cl]) see(p, r);
                    if (both)
                        for (int p : W.ketl[cl])
                            for (int r : X.bral[cl]) see(p, r);
                } else {
                    for (int k : X.braabl) see(k / nm, k % nm);
                    for (int k : W.ketabl) see(k / nm, k % nm);
                    for (int p : X.bral[0])
                        for (int r : W.ketl[1]) see(p, r);
                    for (int p : W.ketl[0])
                        for (int r : X.bral[1]) see(p, r);
                }
                if (grow)
                    for (size_t k = 0; k < nm2; ++k)
                        if (cov[k] && stamp[k] != stamp_gen) cw.blocks.push_back((int)k);
                if (cw.blocks.empty()) continue;
                {  // order by p (counting sort): single blocks are grouped by p
                    std::fill(pcount.begin(), pcount.end(), 0);
                    for (int k : cw.blocks) ++pcount[k / nm + 1];
                    for (int p = 0; p < nm; ++p) pcount[p + 1] += pcount[p];
                    sorted.resize(cw.blocks.size());
                    for (int k : cw.blocks) sorted[pcount[k / nm]++] = k;
                    cw.blocks.swap(sorted);
                }
                cw.kk = cp.kkcov[cl];
                for (int k : need_kk) cw.kk[k] = 1;
                // rectangles: bra rows of x against ket rows of w (both orders
                // for J of pairs with zero-overlap orbitals; II: alpha bra x beta
                // ket and alpha ket x beta bra), cut to the rows with work
                for (int k : cw.blocks) todo[k] = 1;
                std::vector<std::pair<const std::vector<int> *, const std::vector<int> *>> rc;
                if (cl < 2) {
                    rc.push_back({&X.bral[cl], &W.ketl[cl]});
                    if (both) rc.push_back({&W.ketl[cl], &X.bral[cl]});
                } else {
                    rc.push_back({&X.bral[0], &W.ketl[1]});
                    rc.push_back({&W.ketl[0], &X.bral[1]});
                }
                std::vector<int> extra;  // rectangle blocks not in the list
                for (auto &cd : rc) {
                    std::vector<int> P, R;
                    for (int p : *cd.first) {
                        bool any_r = false;
                        for (int r : *cd.second) any_r = any_r || todo[(size_t)p * nm + r];
                        if (any_r) P.push_back(p);
                    }
                    for (int r : *cd.second) {
                        bool any_p = false;
                        for (int p : P) any_p = any_p || todo[(size_t)p * nm + r];
                        if (any_p) R.push_back(r);
                    }
                    if (P.empty() || R.empty()) continue;
                    std::sort(P.begin(), P.end());
                    std::sort(R.begin(), R.end());
                    for (int p : P)
                        for (int r : R) {
                            const size_t k = (size_t)p * nm + r;
                            if (!done[k] && !todo[k]) extra.push_back((int)k);
                            done[k] = 1;
                        }
                    cw.rects.push_back({std::move(P), std::move(R)});
                }
                for (int k : cw.blocks)
                    if (!done[k]) cw.single.push_back(k);
                for (int k : cw.blocks) todo[k] = done[k] = 0;
                for (int k : extra) done[k] = 0;
                cw.blocks.insert(cw.blocks.end(), extra.begin(), extra.end());
                any = true;
            }
            if (any) {
                res[ki] = std::move(pw);
                has[ki] = 1;
            }
        }
    };
    const size_t nt = std::min<size_t>({8, keys.size(), std::max(1u, std::thread::hardware_concurrency())});
    if (nt <= 1) {
        plan_range(0, 1);
    } else {
        std::vector<std::thread> th;
        for (size_t t = 1; t < nt; ++t) th.emplace_back(plan_range, t, nt);
        plan_range(0, nt);
        for (auto &t : th) t.join();
    }
    for (size_t ki = 0; ki < keys.size(); ++ki)
        if (has[ki]) work.push_back(std::move(res[ki]));
    return !work.empty();
}

// The FP32 factor products on BF16 MFMA with three-way split operands (the
// split and the six products of k_r4_bf3): K staged 32 at a time as three
// bf16 planes, one v_mfma_f32_16x16x32_bf16 per pair of parts; tiles of
// 16 TM rows x 128 columns, each wave 16 TM x 32.  K must be a multiple of
// 4 (16-byte loads; K past its end reads zeros).
template <int TM>
__global__ __launch_bounds__(256) void k_fac_gemm_bf3(const FacGemm *__restrict__ descs, int M, int N, int K, int lda,
                                                      int ldb, int ldc)
{
    typedef float gf4 __attribute__((ext_vector_type(4)));
    typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
    typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
    constexpr int TN = 2, KC = 32, VE = 4, KV = KC / VE, LDK = KC + 8, BM = 16 * TM, BN = 64 * TN;
    constexpr int NV = (BM + BN) * KV, PER = (NV + 255) / 256;
    static_assert(KV == 8 && (BM + BN) % 16 == 0, "bf3_piece layout");
    __shared__ __attribute__((aligned(16))) __bf16 Xs[3][BM + BN][LDK];
    const FacGemm g = descs[blockIdx.y];
    const float *__restrict__ A = (const float *)g.A;
    const float *__restrict__ B = (const float *)g.B;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int n0 = blockIdx.x * BN, m0 = blockIdx.z * BM, mloc = min(BM, M - m0);
    const float *src[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int e = tid + j * 256;
        int r, kv;
        bf3_piece(e, r, kv);
        const float *q = nullptr;
        if (e < NV) {
            if (r < BM) {
                if (r < mloc) q = A + (size_t)(m0 + r) * lda + kv * VE;
            } else if (n0 + r - BM < N) {
                q = B + (size_t)(n0 + r - BM) * ldb + kv * VE;
            }
        }
        src[j] = q;
    }
    gf4 buf[PER];
    auto gload = [&](int k0) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            int r, kv;
            bf3_piece(tid + j * 256, r, kv);
            buf[j] = src[j] && k0 + kv * VE < K ? *reinterpret_cast<const gf4 *>(src[j] + k0) : gf4{0, 0, 0, 0};
        }
    };
    r4_f4 acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = r4_f4{0, 0, 0, 0};
    gload(0);
    for (int k0 = 0; k0 < K; k0 += KC) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e = tid + j * 256;
            if (e >= NV) continue;
            int r, kv;
            bf3_piece(e, r, kv);
            bf4 h, m, l;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float x = buf[j][i];
                const __bf16 hh = (__bf16)x;
                const float rem = x - (float)hh;
                const __bf16 mm = (__bf16)rem;
                h[i] = hh;
                m[i] = mm;
                l[i] = (__bf16)(rem - (float)mm);
            }
            *reinterpret_cast<bf4 *>(&Xs[0][r][kv * VE]) = h;
            *reinterpret_cast<bf4 *>(&Xs[1][r][kv * VE]) = m;
            *reinterpret_cast<bf4 *>(&Xs[2][r][kv * VE]) = l;
        }
        __syncthreads();
        if (k0 + KC < K) gload(k0 + KC);  // next chunk in flight during the products
        bf8 bh[TN], bm[TN], bl[TN];
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            const int r = BM + wave * 16 * TN + b * 16 + lr;
            bh[b] = *reinterpret_cast<const bf8 *>(&Xs[0][r][lk * 8]);
            bm[b] = *reinterpret_cast<const bf8 *>(&Xs[1][r][lk * 8]);
            bl[b] = *reinterpret_cast<const bf8 *>(&Xs[2][r][lk * 8]);
        }
#pragma unroll
        for (int a = 0; a < TM; ++a) {
            const int r = a * 16 + lr;
            const bf8 ah = *reinterpret_cast<const bf8 *>(&Xs[0][r][lk * 8]);
            const bf8 am = *reinterpret_cast<const bf8 *>(&Xs[1][r][lk * 8]);
            const bf8 al = *reinterpret_cast<const bf8 *>(&Xs[2][r][lk * 8]);
#pragma unroll
            for (int b = 0; b < TN; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[b], acc[a][b], 0, 0, 0);
            }
        }
    }
    float *C = (float *)g.C;
#pragma unroll
    for (int b = 0; b < TN; ++b) {
        const int col = n0 + wave * 16 * TN + b * 16 + lr;
        if (col >= N) continue;
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = a * 16 + R4Mfma<float>::row(lk, r);
                if (row < mloc) C[(size_t)(m0 + row) * ldc + col] = acc[a][b][r];
            }
    }
}

void launch_fac_gemm_bf3(hipStream_t st, int M, int N, int K, int lda, int ldb, int ldc, const FacGemm *d, int nb)
{
    const int tm = std::min(8, (M + 15) / 16);
    const unsigned mblk = (unsigned)((M + 16 * tm - 1) / (16 * tm));
    for (int z0 = 0; z0 < nb; z0 += 65535) {
        const dim3 grid((unsigned)((N + 127) / 128), (unsigned)std::min(65535, nb - z0), mblk);
        const FacGemm *dz = d + z0;
        switch (tm) {
            case 1: hipLaunchKernelGGL((k_fac_gemm_bf3<1>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 2: hipLaunchKernelGGL((k_fac_gemm_bf3<2>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 3: hipLaunchKernelGGL((k_fac_gemm_bf3<3>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 4: hipLaunchKernelGGL((k_fac_gemm_bf3<4>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 5: hipLaunchKernelGGL((k_fac_gemm_bf3<5>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 6: hipLaunchKernelGGL((k_fac_gemm_bf3<6>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 7: hipLaunchKernelGGL((k_fac_gemm_bf3<7>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            default: hipLaunchKernelGGL((k_fac_gemm_bf3<8>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
        }
    }
}

template <typename T, int TN>
void launch_fac_gemm(hipStream_t st, int M, int N, int K, int lda, int ldb, int ldc, const FacGemm *d, int nb)
{
    const int tm = std::min(8, (M + 15) / 16);
    const unsigned mblk = (unsigned)((M + 16 * tm - 1) / (16 * tm));
    for (int z0 = 0; z0 < nb; z0 += 65535) {
        const dim3 grid((unsigned)((N + 64 * TN - 1) / (64 * TN)), (unsigned)std::min(65535, nb - z0), mblk);
        const FacGemm *dz = d + z0;
        switch (tm) {
            case 1: hipLaunchKernelGGL((k_fac_gemm<T, 1, TN>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 2: hipLaunchKernelGGL((k_fac_gemm<T, 2, TN>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 3: hipLaunchKernelGGL((k_fac_gemm<T, 3, TN>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 4: hipLaunchKernelGGL((k_fac_gemm<T, 4, TN>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 5: hipLaunchKernelGGL((k_fac_gemm<T, 5, TN>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 6: hipLaunchKernelGGL((k_fac_gemm<T, 6, TN>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            case 7: hipLaunchKernelGGL((k_fac_gemm<T, 7, TN>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
            default: hipLaunchKernelGGL((k_fac_gemm<T, 8, TN>), grid, dim3(256), 0, st, dz, M, N, K, lda, ldb, ldc); break;
        }
    }
}

// A_K = X^T L_K Y for every job (X, Y: n x nm row-major), as [p][q][K]:
// T1[q][K][a] = sum_b Y[b][q] L[K][a][b] (one product over the stacked
// factors per distinct Y), then out[p][(q, K)] = sum_a X[a][p] T1[(q, K)][a].
// Jobs of FP32 pairs run in FP32 (FP32-rounded L, X and Y; FP32 MFMA),
// the others in FP64.  n_bas a multiple of 4: k_fac_gemm (64 and 128
// columns per workgroup in FP64 and FP32, measured best at n_bas = 100);
// otherwise the generic FP64 product kernel, rounded to FP32 afterwards for
// FP32 jobs.
struct FacJob {
    const double *X, *Y;
    void *out;
    bool f32;
};

template <typename T>
void factor_run(nw_ctx *c, const std::vector<FacJob> &jobs, Uploader &up)
{
    CholState *s = c->chol;
    const int n = c->n, nm = c->nmo, naux = s->naux;
    const size_t nnm = (size_t)n * nm, t1sz = (size_t)nm * naux * n;
    constexpr bool f32 = sizeof(T) == 4;
    const bool fast = n % 4 == 0;
    // transposes of every distinct X and Y
    std::vector<const double *> mats;
    std::vector<int> xi(jobs.size()), yi(jobs.size());
    auto idx_of = [&](const double *m) {
        for (size_t i = 0; i < mats.size(); ++i)
            if (mats[i] == m) return (int)i;
        mats.push_back(m);
        return (int)mats.size() - 1;
    };
    for (size_t j = 0; j < jobs.size(); ++j) {
        xi[j] = idx_of(jobs[j].X);
        yi[j] = idx_of(jobs[j].Y);
    }
    T *tr = (f32 ? s->trans32 : s->trans).get<T>(mats.size() * nnm);
    std::vector<T *> trp(mats.size());
    for (size_t i = 0; i < mats.size(); ++i) trp[i] = tr + i * nnm;
    hipLaunchKernelGGL(k_transpose_batched<T>, dim3((unsigned)std::min<size_t>(64, (nnm + 255) / 256), (unsigned)mats.size()),
                       dim3(256), 0, c->st, up.put(mats), up.put(trp), n, nm);
    const T *L = (const T *)s->d_L;
    if (f32) {
        if (!s->d_L32) {
            s->d_L32 = dalloc<float>((size_t)naux * n * n);
            launch_to_f32(c->st, s->d_L, s->d_L32, (long long)naux * n * n);
        }
        L = (const T *)s->d_L32;
    }
    // distinct Y in groups whose T1 scratch stays within 1 GiB
    std::vector<int> ys;
    for (int y : yi)
        if (std::find(ys.begin(), ys.end(), y) == ys.end()) ys.push_back(y);
    const size_t gmax = std::max<size_t>(1, (size_t(1) << 30) / (t1sz * sizeof(T)));
    for (size_t g0 = 0; g0 < ys.size(); g0 += gmax) {
        const size_t gn = std::min(gmax, ys.size() - g0);
        T *t1 = (f32 ? s->t1f : s->t1).get<T>(gn * t1sz);
        std::vector<FacGemm> g1, g2;
        for (size_t k = 0; k < gn; ++k) g1.push_back({trp[ys[g0 + k]], L, t1 + k * t1sz});
        for (size_t j = 0; j < jobs.size(); ++j) {
            const auto pos = std::find(ys.begin() + g0, ys.begin() + g0 + gn, yi[j]);
            if (pos == ys.begin() + g0 + gn) continue;
            g2.push_back({trp[xi[j]], t1 + (size_t)(pos - (ys.begin() + g0)) * t1sz, jobs[j].out});
        }
        if (fast) {
            if (f32 && r4_bf3(c)) {  // NW_PREC_BF16X3: the FP32 jobs' products on BF16 MFMA
                launch_fac_gemm_bf3(c->st, nm, naux * n, n, n, n, naux * n, up.put(g1), (int)g1.size());
                launch_fac_gemm_bf3(c->st, nm, nm * naux, n, n, n, nm * naux, up.put(g2), (int)g2.size());
            } else {
                launch_fac_gemm<T, f32 ? 2 : 1>(c->st, nm, naux * n, n, n, n, naux * n, up.put(g1), (int)g1.size());
                launch_fac_gemm<T, f32 ? 2 : 1>(c->st, nm, nm * naux, n, n, n, nm * naux, up.put(g2), (int)g2.size());
            }
        } else if constexpr (!f32) {
            auto gd = [](const std::vector<FacGemm> &v) {
                std::vector<GemmDesc> o;
                for (const FacGemm &f : v) o.push_back({(const double *)f.A, (const double *)f.B, (double *)f.C});
                return o;
            };
            launch_gemm_nt_mfma(c->st, nm, naux * n, n, n, n, naux * n, up.put(gd(g1)), (int)g1.size());
            launch_gemm_nt_mfma(c->st, nm, nm * naux, n, n, n, nm * naux, up.put(gd(g2)), (int)g2.size());
        }
    }
    HIPCHK(hipGetLastError());
}

void factor_tensors(nw_ctx *c, const std::vector<FacJob> &jobs)
{
    CholState *s = c->chol;
    const size_t ften = (size_t)s->naux * c->nmo * c->nmo;
    std::vector<FacJob> j64, j32;
    std::vector<std::pair<const double *, float *>> lower;  // FP32 jobs without the fast kernel
    for (const FacJob &j : jobs) {
        if (!j.f32) j64.push_back(j);
        else if (c->n % 4 == 0) j32.push_back(j);
        else {
            double *t = (double *)s->arena(ften * sizeof(double));
            j64.push_back({j.X, j.Y, t, false});
            lower.push_back({t, (float *)j.out});
        }
    }
    Uploader up(c, (jobs.size() * 3 + 16) * (sizeof(FacGemm) + 2 * sizeof(void *)) + 16384);
    if (!j64.empty()) factor_run<double>(c, j64, up);
    if (!j32.empty()) factor_run<float>(c, j32, up);
    for (auto &l : lower) launch_to_f32(c->st, l.first, l.second, (long long)ften);
    HIPCHK(hipGetLastError());
}

// Build the planned blocks and mark them built.  Per pair and class, every
// tensor of the class (the spin's J slots, the II blocks) gets the same
// blocks:
//   J slot (i, j, k, l) = jbranch: sum_K A^{kl}[p,q] A^{ij}[r,s] (the slab block
//     of the transform of (ab||cd) with cx_k, xc_l, cx_i, xc_j, build.rs:261-313),
//     less sum_K E^{ki}[p,r] F^{lj}[q,s] for pairs with zero-overlap orbitals;
//   II (a0, aj, b0, bj): sum_K A_alpha^{a0 aj}[p,q] A_beta^{b0 bj}[r,s].
// Within a block, the columns of bra holes and the class's ket-particle
// pairs.  Per rectangle P x R three product families: rows (p, q < oq) x
// columns (r, every s); rows (p, q >= oq) x columns (r, s < os); and per ket
// particle q its pairs S(q): rows (p, q) x columns (r, s in S(q)).  Single
// blocks are whole: rows (p, every q) x columns (r in R(p), every s).
// Exchange products take every built block as a row and the read columns as
// columns.
// The rank-4 tensors of a pair's classes (bit cl of clmask) and the factor
// tensors of each: J slot (i, j, k, l) direct P = A^{kl}, Q = A^{ij} and,
// for pairs with zero-overlap orbitals, exchange E = E^{ki}, F = F^{lj};
// II (a0, aj, b0, bj) P = A_alpha^{a0 aj}, Q = A_beta^{b0 bj}.
struct TSpec {
    int wi, cl;
    int P, Q, E, F;  // factor keys (E, F < 0: no exchange)
    long long off;
};
void pair_specs(const nw_ctx *c, int wi, int key, unsigned clmask, std::vector<TSpec> &specs)
{
    const PairInfo &pi = c->pinfo[key];
    const bool exch = pi.ma + pi.mb > 0;
    for (int sp = 0; sp < 2; ++sp) {
        if (!(clmask >> sp & 1)) continue;
        const long long *o = pi.off + (sp ? O_BB : O_AA);
        for (int sl = 0; sl < 10; ++sl) {
            if (o[O_J + sl] < 0) continue;
            const int ii = jbranch(sl, 0), jj = jbranch(sl, 1), kk = jbranch(sl, 2), ll = jbranch(sl, 3);
            specs.push_back({wi, sp, fkey(FA, sp, kk, ll), fkey(FA, sp, ii, jj), exch ? fkey(FE, sp, kk, ii) : -1,
                             exch ? fkey(FF, sp, ll, jj) : -1, o[O_J + sl]});
        }
    }
    if (clmask >> CL_II & 1)
        for (int a0 = 0; a0 < 2; ++a0)
            for (int aj = 0; aj < 2; ++aj)
                for (int b0 = 0; b0 < 2; ++b0)
                    for (int bj = 0; bj < 2; ++bj) {
                        const long long oo = pi.off[O_II + ((a0 * 2 + aj) * 2 + b0) * 2 + bj];
                        if (oo < 0) continue;
                        specs.push_back({wi, CL_II, fkey(FA, 0, a0, aj), fkey(FA, 1, b0, bj), -1, -1, oo});
                    }
}

// Launch the factor products of every factor tensor the given (pair,
// class mask) entries read and the pairs do not hold yet, in each pair's
// precision (FacJob), as stage 1.
void make_factors(nw_ctx *c, const std::vector<std::pair<int, unsigned>> &pairs)
{
    CholState *s = c->chol;
    const int nm = c->nmo;
    const size_t nnm = (size_t)c->n * nm, ften = (size_t)s->naux * nm * nm;
    double *cx = (double *)c->keepCX.p, *xc = (double *)c->keepXC.p;
    std::vector<TSpec> specs;
    for (size_t i = 0; i < pairs.size(); ++i) pair_specs(c, (int)i, pairs[i].first, pairs[i].second, specs);
    std::vector<FacJob> fj;
    auto need = [&](int i, int key) {
        if (key < 0) return;
        CholPair &cp = s->pairs[pairs[i].first];
        const PairInfo &pi = c->pinfo[pairs[i].first];
        if (pi.fp32 ? cp.fac32.count(key) : cp.fac.count(key)) return;
        const int kind = key >> 3, sp = (key >> 2) & 1, a = (key >> 1) & 1, b = key & 1;
        const int pr = pi.bidx * 2 + sp;
        const double *ca = cx + (size_t)(pr * 2 + a) * nnm, *cb = cx + (size_t)(pr * 2 + b) * nnm;
        const double *xa = xc + (size_t)(pr * 2 + a) * nnm, *xb = xc + (size_t)(pr * 2 + b) * nnm;
        void *t;
        if (pi.fp32) t = cp.fac32[key] = (float *)s->arena(ften * sizeof(float));
        else t = cp.fac[key] = (double *)s->arena(ften * sizeof(double));
        const bool f = pi.fp32 != 0;
        if (kind == FA) fj.push_back({ca, xb, t, f});
        else if (kind == FE) fj.push_back({ca, cb, t, f});
        else fj.push_back({xa, xb, t, f});
    };
    for (const TSpec &t : specs) {
        need(t.wi, t.P);
        need(t.wi, t.Q);
        need(t.wi, t.E);
        need(t.wi, t.F);
    }
    if (fj.empty()) return;
    if (s->ev_stage.empty()) HIPCHK(hipEventRecord(s->mark(0), c->st));  // first launch of the call
    HIPCHK(hipEventRecord(s->mark(1), c->st));
    factor_tensors(c, fj);
    // per distinct (Y, precision) nm x (naux n) x n, per job nm x (nm naux) x n
    std::vector<std::pair<const double *, bool>> ys;
    for (const FacJob &j : fj)
        if (std::find(ys.begin(), ys.end(), std::make_pair(j.Y, j.f32)) == ys.end()) ys.push_back({j.Y, j.f32});
    s->stats[8] += 2.0 * nm * (double)s->naux_in * c->n * ((double)ys.size() * c->n + (double)fj.size() * nm);
}

void build_blocks(nw_ctx *c, std::vector<PairWork> &work)
{
    CholState *s = c->chol;
    const int nm = c->nmo, naux = s->naux;
    const size_t nm2 = (size_t)nm * nm;
    auto occ_q = [&](int cl) { return cl < 2 ? c->nocc[cl] : c->nocc[0]; };
    auto occ_s = [&](int cl) { return cl < 2 ? c->nocc[cl] : c->nocc[1]; };
    if (s->ev_stage.empty()) HIPCHK(hipEventRecord(s->mark(0), c->st));  // first launch of the call
    std::vector<TSpec> specs;
    std::vector<std::pair<int, unsigned>> pairs;
    for (size_t i = 0; i < work.size(); ++i) {
        unsigned mask = 0;
        for (int cl = 0; cl < 3; ++cl)
            if (!work[i].cls[cl].blocks.empty()) mask |= 1u << cl;
        pair_specs(c, (int)i, work[i].key, mask, specs);
        pairs.push_back({work[i].key, mask});
    }
    make_factors(c, pairs);  // those not made before the planning
    // index lists shared by the products of a (pair, class), in one upload:
    // rectangle rows, per-q ket-particle partners, sparse lists of single
    // blocks, exchange rows and columns.  Offsets are patched after upload.
    struct ClassLists {
        std::vector<std::pair<size_t, size_t>> rect;  // offsets of P and R lists
        std::vector<std::pair<int, size_t>> kq;       // (q, offset of S(q)) with |S(q)| in kqn
        std::vector<int> kqn;
        std::vector<std::pair<size_t, size_t>> sgl;   // single blocks grouped by p: offsets of {p} and R(p)
        std::vector<int> sgln;
        // the same blocks for products whose two factor tensors are one: (p, r)
        // with p < r built with its mirror (r, p) (sglm), the unpaired ones (sglp)
        std::vector<std::pair<size_t, size_t>> sglm, sglp;
        std::vector<int> sglmn, sglpn;
        size_t xp = 0, xr = 0, xq = 0, xs = 0;        // exchange rows (p, r) and columns (q, s)
        int nxrow = 0, nxcol = 0;
    };
    std::vector<ClassLists> cls(work.size() * 3);
    // The (pair, class) lists are independent: a few host threads, each
    // appending to an index array of its own; the arrays are then joined
    // and each list's offsets moved by its array's start.
    const size_t ntask = work.size() * 3;
    const size_t nth = std::min<size_t>({8, ntask, std::max(1u, std::thread::hardware_concurrency())});
    std::vector<std::vector<int>> tints(nth);
    auto lists_range = [&](size_t t0) {
        std::vector<int> &ints = tints[t0];
        ints.reserve(1 << 18);
        auto put = [&](const std::vector<int> &v) {
            const size_t o = ints.size();
            ints.insert(ints.end(), v.begin(), v.end());
            return o;
        };
        std::vector<uint8_t> is_single(nm2, 0);  // single blocks of the current J class (cleared after it)
        for (size_t task = t0; task < ntask; task += nth) {
            const size_t i = task / 3;
            const int cl = (int)(task % 3);
            const ClassWork &cw = work[i].cls[cl];
            if (cw.blocks.empty()) continue;
            ClassLists &L = cls[i * 3 + cl];
            const int oq = occ_q(cl), os = occ_s(cl);
            for (auto &rc : cw.rects) L.rect.push_back({put(rc.first), put(rc.second)});
            for (int q = oq; q < nm; ++q) {  // S(q), appended in place (sparse rows: 8 bytes at a time)
                const size_t o = ints.size();
                const uint8_t *row = cw.kk.data() + (size_t)q * nm;
                int t = os;
                for (; t + 8 <= nm; t += 8) {
                    uint64_t w8;
                    std::memcpy(&w8, row + t, 8);
                    if (!w8) continue;
                    for (int u = 0; u < 8; ++u)
                        if (row[t + u]) ints.push_back(t + u);
                }
                for (; t < nm; ++t)
                    if (row[t]) ints.push_back(t);
                if (ints.size() > o) {
                    L.kq.push_back({q, o});
                    L.kqn.push_back((int)(ints.size() - o));
                }
            }
            if (cl < 2)
                for (int k : cw.single) is_single[k] = 1;
            for (size_t k = 0; k < cw.single.size();) {  // cw.single is ordered by p
                const int p0 = cw.single[k] / nm;
                const size_t k0 = k;
                for (; k < cw.single.size() && cw.single[k] / nm == p0; ++k) {
                }
                const size_t po = ints.size();
                ints.push_back(p0);
                const size_t ro = ints.size();
                for (size_t e = k0; e < k; ++e) ints.push_back(cw.single[e] % nm);
                L.sgl.push_back({po, ro});
                L.sgln.push_back((int)(k - k0));
                if (cl >= 2) continue;
                // (p0, r) with p0 < r and (r, p0) also single comes with its mirror
                const size_t mo = ints.size();
                for (size_t e = k0; e < k; ++e) {
                    const int r = cw.single[e] % nm;
                    if (r != p0 && is_single[(size_t)r * nm + p0] && p0 < r) ints.push_back(r);
                }
                if (ints.size() > mo) {
                    L.sglm.push_back({po, mo});
                    L.sglmn.push_back((int)(ints.size() - mo));
                }
                const size_t uo = ints.size();
                for (size_t e = k0; e < k; ++e) {
                    const int r = cw.single[e] % nm;
                    if (!(r != p0 && is_single[(size_t)r * nm + p0])) ints.push_back(r);
                }
                if (ints.size() > uo) {
                    L.sglp.push_back({po, uo});
                    L.sglpn.push_back((int)(ints.size() - uo));
                }
            }
            if (cl < 2)
                for (int k : cw.single) is_single[k] = 0;
            if (c->pinfo[work[i].key].ma + c->pinfo[work[i].key].mb > 0 && cl < 2) {
                std::vector<int> rp, rr, cq, cs;
                for (int k : cw.blocks) {
                    rp.push_back(k / nm);
                    rr.push_back(k % nm);
                }
                for (int q = 0; q < nm; ++q) {  // q < oq: every s; else s < os and the row's kk pairs
                    const uint8_t *row = cw.kk.data() + (size_t)q * nm;
                    for (int t = 0; t < nm; ++t)
                        if (q < oq || t < os || row[t]) {
                            cq.push_back(q);
                            cs.push_back(t);
                        }
                }
                L.xp = put(rp);
                L.xr = put(rr);
                L.xq = put(cq);
                L.xs = put(cs);
                L.nxrow = (int)rp.size();
                L.nxcol = (int)cq.size();
            }
        }
    };
    if (nth > 1) {
        std::vector<std::thread> th;
        for (size_t t = 1; t < nth; ++t) th.emplace_back(lists_range, t);
        lists_range(0);
        for (auto &t : th) t.join();
    } else {
        lists_range(0);
    }
    std::vector<int> ints;
    std::vector<size_t> tbase(nth);
    {
        size_t tot = 0;
        for (const auto &v : tints) tot += v.size();
        ints.reserve(tot);
        for (size_t t = 0; t < nth; ++t) {
            tbase[t] = ints.size();
            ints.insert(ints.end(), tints[t].begin(), tints[t].end());
        }
    }
    for (size_t task = 0; task < ntask; ++task) {
        ClassLists &L = cls[task];
        const size_t b = tbase[task % nth];
        for (auto &x : L.rect) x = {x.first + b, x.second + b};
        for (auto &x : L.kq) x.second += b;
        for (auto *v : {&L.sgl, &L.sglm, &L.sglp})
            for (auto &x : *v) x = {x.first + b, x.second + b};
        L.xp += b;
        L.xr += b;
        L.xq += b;
        L.xs += b;
    }
    // List pointers hold (offset + 1) into `ints` until the upload; 0 = none.
    const int *dints = nullptr;
    auto at = [](size_t off) { return (const int *)(uintptr_t)(off + 1); };
    std::vector<R4Gemm> gd[2], gx[2];
    std::vector<R4Gemm> gk[2];   // ket-particle pair products (narrow)
    std::vector<R4Gemm> gs1[2];  // single blocks, whole
    double nsingle[2] = {0.0, 0.0};
    for (int f32 = 0; f32 < 2; ++f32) {
        for (const TSpec &t : specs) {
            const PairWork &pw = work[t.wi];
            const PairInfo &pi = c->pinfo[pw.key];
            if ((pi.fp32 != 0) != (f32 == 1)) continue;
            CholPair &cp = s->pairs[pw.key];
            auto ptr = [&](int key) -> const void * {
                return f32 ? (const void *)cp.fac32.at(key) : (const void *)cp.fac.at(key);
            };
            void *out = (void *)(c->d_slab + pi.base + t.off);
            const ClassWork &cw = pw.cls[t.cl];
            const ClassLists &L = cls[t.wi * 3 + t.cl];
            const int oq = occ_q(t.cl), os = occ_s(t.cl);
            const void *PA = ptr(t.P), *QB = ptr(t.Q);
            for (size_t ri = 0; ri < cw.rects.size(); ++ri) {
                const int np = (int)cw.rects[ri].first.size(), nr = (int)cw.rects[ri].second.size();
                const int *Pl = at(L.rect[ri].first), *Rl = at(L.rect[ri].second);
                if (oq > 0) {  // rows (p, q < oq) x columns (r, every s)
                    R4Gemm g{PA, QB, out, {Pl, nullptr, np, oq, 0, 0}, {Rl, nullptr, nr, nm, 0, 0}, np * oq, nr * nm, 0, 0};
                    gd[f32].push_back(g);
                }
                if (os > 0 && oq < nm) {  // rows (p, q >= oq) x columns (r, s < os)
                    R4Gemm g{PA, QB, out, {Pl, nullptr, np, nm - oq, oq, 0}, {Rl, nullptr, nr, os, 0, 0},
                             np * (nm - oq), nr * os, 0, 0};
                    gd[f32].push_back(g);
                }
                for (size_t k = 0; k < L.kq.size(); ++k) {  // rows (p, q) x columns (r, s in S(q)), or transposed
                    const int q = L.kq[k].first, ns = L.kqn[k];
                    const R4Axis pq{Pl, nullptr, np, 1, q, 0}, rs{Rl, at(L.kq[k].second), nr, ns, 0, 0};
                    if (np >= nr * ns) gk[f32].push_back(R4Gemm{PA, QB, out, pq, rs, np, nr * ns, 0, 0});
                    else gk[f32].push_back(R4Gemm{QB, PA, out, rs, pq, nr * ns, np, 2, 0});
                }
            }
            // whole blocks: rows (p, every q) x columns (r, every s); with one
            // factor tensor (J slots (i, j, i, j)) a pair of blocks (p, r), (r, p)
            // comes from one product
            auto singles = [&](const std::vector<std::pair<size_t, size_t>> &lst, const std::vector<int> &cnt,
                               int mode) {
                for (size_t k = 0; k < lst.size(); ++k) {
                    R4Gemm g{PA, QB, out, {at(lst[k].first), nullptr, 1, nm, 0, 0},
                             {at(lst[k].second), nullptr, cnt[k], nm, 0, 0}, nm, cnt[k] * nm, mode, 0};
                    gs1[f32].push_back(g);
                    nsingle[f32] += mode == 3 ? 2 * cnt[k] : cnt[k];
                }
            };
            if (t.cl < 2 && t.P == t.Q) {
                singles(L.sglm, L.sglmn, 3);
                singles(L.sglp, L.sglpn, 0);
            } else {
                singles(L.sgl, L.sgln, 0);
            }
            if (t.E >= 0 && L.nxrow > 0 && L.nxcol > 0) {
                R4Gemm g{ptr(t.E), ptr(t.F), out, {at(L.xp), at(L.xr), L.nxrow, 1, 0, 1},
                         {at(L.xq), at(L.xs), L.nxcol, 1, 0, 1}, L.nxrow, L.nxcol, 1, 0};
                gx[f32].push_back(g);
            }
        }
    }
    // one upload of the lists and every descriptor, then the launches
    size_t bytes = ints.size() * sizeof(int) + 8192;
    for (int f = 0; f < 2; ++f)  // descriptors, and tiles of at most 64 x 64 elements
        for (auto *v : {&gd[f], &gk[f], &gx[f], &gs1[f]})
            bytes += v->size() * (sizeof(R4Gemm) + 2 * sizeof(long long)) + 2048;
    Uploader up(c, bytes);
    dints = up.put(ints);
    auto fix = [&](const int *p) { return p ? dints + ((uintptr_t)p - 1) : nullptr; };
    for (int f = 0; f < 2; ++f) {
        for (auto *v : {&gd[f], &gk[f], &gx[f], &gs1[f]})
            for (R4Gemm &g : *v) {
                g.row.u = fix(g.row.u);
                g.row.v = fix(g.row.v);
                g.col.u = fix(g.col.u);
                g.col.v = fix(g.col.v);
            }
    }
    for (int f = 0; f < 2; ++f) {
        for (const R4Gemm &g : gd[f]) s->stats[5] += 2.0 * g.M * (double)g.N * s->naux_in;
        for (const R4Gemm &g : gk[f]) s->stats[5] += 2.0 * g.M * (double)g.N * s->naux_in;
        for (const R4Gemm &g : gx[f]) s->stats[7] += 2.0 * g.M * (double)g.N * s->naux_in;
        for (const R4Gemm &g : gs1[f]) s->stats[6] += 2.0 * g.M * (double)g.N * s->naux_in;
        s->stats[9] += nsingle[f];
        for (auto *v : {&gd[f], &gx[f], &gs1[f]})
            for (const R4Gemm &g : *v) s->kstats[4 + 2 * f] += 2.0 * g.M * (double)g.N * s->naux_in;
        for (const R4Gemm &g : gk[f]) s->kstats[5 + 2 * f] += 2.0 * g.M * (double)g.N * s->naux_in;
    }
    for (int f32 = 0; f32 < 2; ++f32) {
        // the exchange products run after the direct products of the same blocks (stream order)
        const int stage[4] = {2, 2, 3, 4};
        std::vector<R4Gemm> *lists[4] = {&gd[f32], &gk[f32], &gs1[f32], &gx[f32]};
        for (int k = 0; k < 4; ++k) {
            if (lists[k]->empty()) continue;
            const int cls = 2 * f32 + (k == 1 ? 1 : 0);
            HIPCHK(hipEventRecord(s->mark(stage[k] + 16 * (cls + 1)), c->st));
            if (k == 1) launch_r4g_narrow(c, up, *lists[k], naux, f32 != 0);
            else if (f32) launch_r4g<float>(c, up, *lists[k], naux);
            else launch_r4g<double>(c, up, *lists[k], naux);
        }
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(s->mark(0), c->st));
    for (PairWork &pw : work) {
        CholPair &cp = s->pairs[pw.key];
        for (int cl = 0; cl < 3; ++cl) {
            if (pw.cls[cl].blocks.empty()) continue;
            for (int k : pw.cls[cl].blocks) cp.cov[cl][k] = 1;
            cp.kkcov[cl] = pw.cls[cl].kk;
        }
    }
}

}  // namespace

void chol_read_stages(CholState *s);

std::vector<uint8_t> triangle_pairs(const nw_ctx *c, const std::vector<PackedDet> &h)
{
    // runs of equal (parent, ranks) in list order; the upper triangle pairs
    // runs r1 <= r2, so it reads the ordered reference pairs (parent(r1), parent(r2))
    std::vector<int> runp;
    for (size_t i = 0; i < h.size(); ++i)
        if (i == 0 || h[i].parent != h[i - 1].parent || h[i].ra != h[i - 1].ra || h[i].rb != h[i - 1].rb)
            runp.push_back(h[i].parent);
    std::vector<uint8_t> seen((size_t)c->nref, 0), out((size_t)c->nref * c->nref, 0);
    for (size_t r2 = 0; r2 < runp.size(); ++r2) {
        seen[runp[r2]] = 1;
        for (int x = 0; x < c->nref; ++x)
            if (seen[x]) out[(size_t)x * c->nref + runp[r2]] = 1;
    }
    return out;
}

void chol_ensure(nw_ctx *c, const std::vector<PackedDet> &h, bool accumulate, const std::vector<uint8_t> *reads)
{
    CholState *s = c->chol;
    if (!s) return;
    if (accumulate) chol_read_stages(s);
    else {
        for (double &v : s->stats) v = 0.0;
        for (double &v : s->kstats) v = 0.0;
    }
    s->ms_blocks = 0.0;
    s->ev_stage.clear();
    s->pending = false;
    // Factor products first, for every pair this call may build (lazy,
    // built, read, both parents among h; every class), so that they run while
    // the host plans the blocks; those of classes left without blocks are
    // unused but kept with the pair.
    std::vector<uint8_t> present((size_t)c->nref, 0);
    for (const PackedDet &d : h)
        if (d.parent < c->nref) present[d.parent] = 1;
    std::vector<std::pair<int, unsigned>> pre;
    for (size_t k = 0; k < s->pairs.size(); ++k) {
        const int x = (int)k / c->nref, w = (int)k % c->nref;
        if (s->pairs[k].lazy && c->pinfo[k].built && present[x] && present[w] && (!reads || (*reads)[k]))
            pre.push_back({(int)k, 7u});
    }
    if (pre.empty()) return;
    make_factors(c, pre);
    std::vector<PairWork> work;
    if (plan_missing(c, h, work, reads)) build_blocks(c, work);
    else if (!s->ev_stage.empty()) HIPCHK(hipEventRecord(s->mark(0), c->st));
    s->pending = !s->ev_stage.empty();  // the caller's evaluation is planned while the blocks build
}

// Stage times of the last chol_ensure, once its last event has completed.
void chol_read_stages(CholState *s)
{
    if (!s || !s->pending) return;
    s->pending = false;
    HIPCHK(hipEventSynchronize(s->evs[s->ev_stage.size() - 1]));
    // stage k runs from its mark to the next mark
    for (size_t k = 0; k + 1 < s->ev_stage.size(); ++k) {
        const int st = s->ev_stage[k] & 15, cls = (s->ev_stage[k] >> 4) - 1;
        if (st == 0) continue;
        const double ms = elapsed(s->evs[k], s->evs[k + 1]);
        s->stats[st] += ms;
        if (cls >= 0) s->kstats[cls] += ms;
    }
    s->ms_blocks = elapsed(s->evs[0], s->evs[s->ev_stage.size() - 1]);
    s->stats[0] += s->ms_blocks;
}

bool chol_missing(nw_ctx *c, const std::vector<PackedDet> &h)
{
    std::vector<PairWork> work;
    return plan_missing(c, h, work, nullptr);
}

void chol_complete(nw_ctx *c)
{
    CholState *s = c->chol;
    if (!s) return;
    std::vector<int> keys;
    for (size_t k = 0; k < s->pairs.size(); ++k)
        if (s->pairs[k].lazy && c->pinfo[k].built) keys.push_back((int)k);
    if (keys.empty()) return;
    run_rank4_transforms(c, keys);
    HIPCHK(hipStreamSynchronize(c->st));
    for (int k : keys) {
        CholPair &cp = s->pairs[k];
        cp.lazy = false;
        for (auto &v : cp.cov) std::fill(v.begin(), v.end(), 1);
        for (auto &v : cp.kkcov) std::fill(v.begin(), v.end(), 1);
    }
}

double chol_last_ms(nw_ctx *c)
{
    if (!c->chol) return 0.0;
    chol_read_stages(c->chol);
    return c->chol->ms_blocks;
}

bool chol_any_lazy(nw_ctx *c)
{
    if (!c->chol) return false;
    for (const CholPair &cp : c->chol->pairs)
        if (cp.lazy) return true;
    return false;
}

std::vector<PackedDet> host_pack(const nw_det *dets, uint64_t ndet)
{
    std::vector<PackedDet> h(ndet);
    auto bits = [](const uint64_t *m, uint8_t *out, int cap) {
        int k = 0;
        for (int w = 0; w < 2; ++w)
            for (int b = 0; b < 64; ++b)
                if ((m[w] >> b) & 1ull) {
                    if (k < cap) out[k] = (uint8_t)(w * 64 + b);
                    ++k;
                }
        return std::min(k, cap);
    };
    for (uint64_t i = 0; i < ndet; ++i) {
        const nw_det &d = dets[i];
        PackedDet &p = h[i];
        std::memset(&p, 0, sizeof(p));
        p.parent = (uint16_t)std::min<uint32_t>(d.parent, 0xffffu);
        const int ha = bits(d.holes_a, p.ia, 4), pa = bits(d.parts_a, p.ia + 4, 4);
        const int hb = bits(d.holes_b, p.ib, 4), pb = bits(d.parts_b, p.ib + 4, 4);
        p.ra = (uint8_t)std::min(ha, pa);
        p.rb = (uint8_t)std::min(hb, pb);
    }
    return h;
}

}  // namespace nwi

using namespace nwi;

extern "C" {

int nw_set_eri_factors(nw_ctx *c, const double *L, int naux, double check_tol, double *max_dev)
{
    return guarded(c, [&]() -> int {
        if (max_dev) *max_dev = 0.0;
        if (naux < 0) return fail(c, NW_E_ARG, "naux must be >= 0");
        if (naux > 0 && !L) return fail(c, NW_E_ARG, "null factors");
        if (!(check_tol >= 0.0)) return fail(c, NW_E_ARG, "bad check tolerance");
        if (!c->chol) c->chol = new CholState();
        CholState *s = c->chol;
        HIPCHK(hipStreamSynchronize(c->st));
        if (s->d_L) HIPCHK(hipFree(s->d_L));
        if (s->d_L32) HIPCHK(hipFree(s->d_L32));
        s->d_L = nullptr;
        s->d_L32 = nullptr;
        s->naux = s->naux_in = 0;
        s->pairs.clear();
        c->built = false;  // pairs built before take effect only after the next nw_build
        if (naux == 0) return NW_OK;
        const int n = c->n;
        const int npad = (naux + 3) & ~3;  // zero factors keep every factor-tensor row 16-byte aligned
        const size_t cnt = (size_t)naux * n * n, cpad = (size_t)npad * n * n;
        double *dL = dalloc<double>(cpad);
        HIPCHK(hipMemcpyAsync(dL, L, cnt * sizeof(double), hipMemcpyDefault, c->st));
        if (cpad > cnt) HIPCHK(hipMemsetAsync(dL + cnt, 0, (cpad - cnt) * sizeof(double), c->st));
        unsigned long long *dmax = c->flag.get<unsigned long long>(1);
        HIPCHK(hipMemsetAsync(dmax, 0, sizeof(unsigned long long), c->st));
        const int nsample = 16384;
        hipLaunchKernelGGL(k_factor_check, dim3(64), dim3(256), 0, c->st, dL, c->d_eri, n, naux, nsample, dmax);
        HIPCHK(hipGetLastError());
        unsigned long long bits = 0;
        HIPCHK(hipMemcpyAsync(&bits, dmax, sizeof(bits), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        double dev;
        std::memcpy(&dev, &bits, sizeof(dev));
        if (max_dev) *max_dev = dev;
        if (!(dev <= check_tol)) {
            (void)hipFree(dL);
            char msg[160];
            std::snprintf(msg, sizeof msg, "integral factors do not reproduce the context's ERIs: deviation %.3e > %.3e",
                          dev, check_tol);
            return fail(c, NW_E_ARG, msg);
        }
        s->d_L = dL;
        s->naux = npad;
        s->naux_in = naux;
        return NW_OK;
    });
}

int nw_rank4_stats(nw_ctx *c, double *out10)
{
    return guarded(c, [&]() -> int {
        if (!out10) return fail(c, NW_E_ARG, "null output");
        if (c->chol) chol_read_stages(c->chol);
        for (int k = 0; k < 10; ++k) out10[k] = c->chol ? c->chol->stats[k] : 0.0;
        return NW_OK;
    });
}

int nw_rank4_kernel_stats(nw_ctx *c, double *out8)
{
    return guarded(c, [&]() -> int {
        if (!out8) return fail(c, NW_E_ARG, "null output");
        if (c->chol) chol_read_stages(c->chol);
        for (int k = 0; k < 8; ++k) out8[k] = c->chol ? c->chol->kstats[k] : 0.0;
        return NW_OK;
    });
}

int nw_rank4_coverage(nw_ctx *c, int x, int w, double *frac3)
{
    return guarded(c, [&]() -> int {
        if (!frac3) return fail(c, NW_E_ARG, "null output");
        if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
        if (x < 0 || w < 0 || x >= c->nref || w >= c->nref) return fail(c, NW_E_ARG, "pair out of range");
        for (int k = 0; k < 3; ++k) frac3[k] = 1.0;
        const size_t key = (size_t)x * c->nref + w;
        if (!c->chol || key >= c->chol->pairs.size() || !c->chol->pairs[key].lazy) return NW_OK;
        const CholPair &cp = c->chol->pairs[key];
        for (int k = 0; k < 3; ++k) {
            size_t on = 0;
            for (uint8_t v : cp.cov[k]) on += v;
            frac3[k] = cp.cov[k].empty() ? 1.0 : (double)on / (double)cp.cov[k].size();
        }
        return NW_OK;
    });
}

}  // extern "C"
