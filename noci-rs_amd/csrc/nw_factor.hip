// Spin-factorised one-body and overlap operators over a determinant list.
//
// Every determinant I is (P, a_I, b_I): its parent reference P and its alpha
// and beta excitation strings relative to P, numbered per parent
// (SpinFactorisation, src/noci/factorise/mod.rs:55-95, 185-410).  A one-body
// operator and the overlap factorise per spin,
//   F_IJ = F^a_{aI aJ} S^b_{bI bJ} + S^a_{aI aJ} F^b_{bI bJ},   S_IJ = S^a S^b,
// with per-spin factors between the components of the target parent Q and the
// source parent P (build_spin_one_body_factor_row,
// src/noci/factorise/onebody.rs:1300-1569): the per-spin Wick overlap and
// generalised-Fock sums of the ordered pair whose bra is the parent with the
// earlier first determinant (ordered_parent_pair, mod.rs:163-183), times the
// pair's spin prefactor and both excitation phases.
//
// Tables are double2 {S, F} indexed [source component][target component]
// (target contiguous) and held in device memory for every ordered parent
// pair, or regenerated per application (the reference's RAM and None
// storage).  y = (F + lambda S) x = sum over parent pairs of
// sample_{(a_I, b_I)} (Tab_a^T X_P Tab_b), X_P the sparse (alpha x beta)
// component matrix of the source entries (apply_one_body_a_first / _b_first,
// onebody.rs:491-648).  The reference picks, per parent pair, whether to sum
// the source over alpha first (rows grouped by beta component) or beta first
// (select_one_body_contraction, :1583-1606).  Here every source entry is
// assigned to a group keyed by its beta or its alpha component through a
// greedy vertex cover of the bipartite (alpha, beta) entry graph, so that an
// arrow-shaped X_P (reference row and column full, a sparse block) needs only
// a few hundred rows where either single order needs thousands:
//   stage 1  U_g = sum_{e in g} x_e Tab_a[a_e]  (group keyed by beta b_g),
//            V_g = sum_{e in g} x_e Tab_b[b_e]  (group keyed by alpha a_g);
//   stage 2  y_I += sum_g U_g[a_I] . Tab_b[b_g][b_I]  or  Tab_a[a_g][a_I] . V_g[b_I],
//            + x_e Tab_a[a_e][a_I] . Tab_b[b_e][b_I] for one-entry groups,
// with u . v = u.F v.S + u.S (v.F + lambda v.S); ys = S x from the same pass.
// Stage 1 streams each table row about once; stage 2 is one lane per target
// entry over the group rows.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

#include "nw_eval_dev.h"
#include "nw_internal.h"

using namespace nw;
using namespace nwi;

namespace {

// One spin string of a determinant relative to its parent: rank, phase, the
// holes (i[0..3]) and particles (i[4..7]) ascending, as in PackedDet.
struct SpinComp {
    uint8_t r;
    int8_t ph;
    uint8_t pad[2];
    uint8_t i[8];
};
static_assert(sizeof(SpinComp) == 12, "SpinComp layout");

// Stage-1 work item: out[t] = sum_{k in [e0, e1)} x[mdet[k]] tab[midx[k] nt + t], t < nt.
// Tables and materialised rows hold {S, F} pairs as double2 (FP64 tables) or
// float2 (FP32 tables); the kernels take the element type as a template.
// NV input vectors are applied together (x[i NV + v], v < NV; NV = 2 for the
// real and imaginary parts of a complex vector): each table row is read once
// for all of them, and a materialised row holds NV interleaved entries per
// target component (out[t NV + v]).
struct S1Item {
    const void *tab;
    void *out;
    int nt, e0, e1, pad;
};
// Stage-2 row: alpha-side row u (over target alpha components), beta-side row
// v (over target beta components), weight x[wdet] (wdet >= 0) or 1; umat /
// vmat mark a materialised row (NV entries per component) on that side.
struct S2Row {
    const void *u, *v;
    int wdet;
    uint8_t umat, vmat, pad[2];
};
static_assert(sizeof(S2Row) == 24, "S2Row layout");

// Pointers read from item records are generic; loads through these
// global-address-space views compile to global (not flat) loads.
template <class T>
using gptr = const T __attribute__((address_space(1))) *;
// Native two-element vectors with the layout of double2 / float2.
template <class T2>
struct raw2;
template <>
struct raw2<double2> {
    typedef double __attribute__((ext_vector_type(2))) type;
};
template <>
struct raw2<float2> {
    typedef float __attribute__((ext_vector_type(2))) type;
};

template <class T2>
struct elem_of;
template <>
struct elem_of<double2> {
    typedef double type;
};
template <>
struct elem_of<float2> {
    typedef float type;
};

template <class T2>
__device__ inline T2 pair2(double a, double b);
template <>
__device__ inline double2 pair2<double2>(double a, double b) { return make_double2(a, b); }
template <>
__device__ inline float2 pair2<float2>(double a, double b) { return make_float2((float)a, (float)b); }

// Dual stage-1 item: one pass over the table Tab(q, p) of one spin (ns rows =
// source components of p, nt columns = components of q) gives both
//   forward  out_f[t] = sum_s x[fdet[s]] Tab[s][t]   (a group of p, target q) and
//   reverse  out_r[s] = sum_t Tab[s][t] x[rdet[t]]   (a group of q, target p),
// since Tab(p, q) = Tab(q, p)^T.  Partial sums per row chunk (forward) and per
// column strip (reverse) are reduced in a fixed order by k_fac_dual_reduce.
// sym: a same-parent table (symmetric, fdet = rdet, one output): only its
// upper triangle is read, s <= t forward and s < t reverse, both summed into
// out_f.
struct DualItem {
    const void *tab;
    void *out_f, *out_r;
    const int *fdet, *rdet;  // determinant per row / column, -1 = no weight
    double2 *fpart, *rpart;  // [nchunk][nt][NV], [nstrip][ns][NV]
    int ns, nt, nchunk, nstrip, sym, pad;
};
#ifndef NW_DUAL_ROWS
#define NW_DUAL_ROWS 256
#endif
#ifndef NW_DUAL_RB
#define NW_DUAL_RB 4
#endif
constexpr int DUAL_ROWS = NW_DUAL_ROWS;  // rows per dual block
constexpr int DUAL_COLS = 1024;  // columns per dual block: 4 waves x 64 lanes x 4

constexpr int S1_CHUNK = 128;  // entries per stage-1 item (pieces of a large group)
constexpr int S2_BATCH = 256;  // rows staged in LDS at a time

template <int L>
__device__ inline void spin_fs(slab_t slab, const long long *o, const SpinScal &sc, int nmo,
                               const SpinComp &bra, const SpinComp &ket, double tol, int mode, double &s, double &f)
{
    constexpr int L1 = L > 0 ? L : 1;
    int rows[L1] = {}, cols[L1] = {};
    double d0[L1 * L1] = {}, d1[L1 * L1] = {};
    if constexpr (L > 0) {
        labels<L>(bra.i, bra.r, ket.i, sc.nocc, nmo - sc.nocc, rows, cols);
        build_d<L>(slab + o[O_X], slab + o[O_Y], nmo, rows, cols, d0);
        if (sc.m > 0) build_d<L>(slab + o[O_X + 1], slab + o[O_Y + 1], nmo, rows, cols, d1);
    }
    fs_spin<L>(slab, o, sc, nmo, rows, cols, d0, d1, tol, mode, s, f);
}

// Factor table of one ordered parent pair and spin: tab[s][t] = phase * {S, F}
// with the bra = target when target_left, else the source; 16 x 16 tiles.
// The table of the reverse pair is the transpose (the same Wick element with
// the same bra), so when tabT is not null the tile is also written there,
// tabT[t][s], through LDS.  Elements are evaluated in FP64 and stored as T2
// (float2: rounded to nearest FP32).
template <class T2>
__global__ __launch_bounds__(256) void k_fac_table(const PairInfo *__restrict__ pairs, const double *__restrict__ slab0,
                                                   int pair, int spin, const SpinComp *__restrict__ tc, int nt,
                                                   const SpinComp *__restrict__ scomp, int ns, int target_left, int nmo,
                                                   double tol, int mode, T2 *__restrict__ tab, T2 *__restrict__ tabT)
{
    __shared__ T2 sh[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int t = blockIdx.x * 16 + tx, s = blockIdx.y * 16 + ty;
    T2 out = pair2<T2>(0.0, 0.0);
    if (t < nt && s < ns) {
        const PairInfo &P = pairs[pair];
        const slab_t slab = (slab_t)(slab0 + P.base);
        const SpinScal sc = load_spin(P, spin);
        const long long *o = P.off + (spin ? O_BB : O_AA);
        const SpinComp T = tc[t], S = scomp[s];
        const SpinComp &bra = target_left ? T : S;
        const SpinComp &ket = target_left ? S : T;
        double vs = 0.0, vf = 0.0;
        switch (bra.r + ket.r) {
            case 0: spin_fs<0>(slab, o, sc, nmo, bra, ket, tol, mode, vs, vf); break;
            case 1: spin_fs<1>(slab, o, sc, nmo, bra, ket, tol, mode, vs, vf); break;
            case 2: spin_fs<2>(slab, o, sc, nmo, bra, ket, tol, mode, vs, vf); break;
            case 3: spin_fs<3>(slab, o, sc, nmo, bra, ket, tol, mode, vs, vf); break;
            case 4: spin_fs<4>(slab, o, sc, nmo, bra, ket, tol, mode, vs, vf); break;
            default: break;  // excluded on the host
        }
        const double ph = (double)(T.ph * S.ph) * sc.pref;
        out = pair2<T2>(ph * vs, ph * vf);
        tab[(size_t)s * nt + t] = out;
    }
    if (!tabT) return;  // uniform over the launch
    sh[ty][tx] = out;
    __syncthreads();
    const int t2 = blockIdx.x * 16 + ty, s2 = blockIdx.y * 16 + tx;
    if (t2 < nt && s2 < ns) tabT[(size_t)t2 * ns + s2] = sh[tx][ty];
}

// Stage 1: one item per workgroup row of the grid (x values and table row
// indices staged in LDS); blockIdx.x is the 256-wide target column block.
// The host pads grid.x to a multiple of 8, so a column block always runs on
// the same XCD (workgroups are dealt round-robin over the 8 XCDs) and the
// table rows that several groups share are re-read from that XCD's L2.
// Sums run in FP64 whatever the table type.
template <class T2, int NV>
__global__ __launch_bounds__(256) void k_fac_mat(const S1Item *__restrict__ items, const uint32_t *__restrict__ midx,
                                                 const uint32_t *__restrict__ mdet, const double *__restrict__ x)
{
    __shared__ int s_r[S1_CHUNK];
    __shared__ double s_x[NV][S1_CHUNK];
    const S1Item it = items[blockIdx.y];
    const int ne = it.e1 - it.e0;
    if ((int)threadIdx.x < ne) {
        s_r[threadIdx.x] = (int)midx[it.e0 + threadIdx.x];
        const size_t d = mdet[it.e0 + threadIdx.x];
#pragma unroll
        for (int v = 0; v < NV; ++v) s_x[v][threadIdx.x] = x[d * NV + v];
    }
    __syncthreads();
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= it.nt) return;
    typedef typename raw2<T2>::type R2;
    const gptr<R2> col = (gptr<R2>)it.tab + t;
    double s[NV], f[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = f[v] = 0.0;
#pragma unroll 8
    for (int k = 0; k < ne; ++k) {
        const R2 e = col[(size_t)s_r[k] * it.nt];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            s[v] = fma(s_x[v][k], (double)e.x, s[v]);
            f[v] = fma(s_x[v][k], (double)e.y, f[v]);
        }
    }
    T2 *o = (T2 *)it.out + (size_t)t * NV;
#pragma unroll
    for (int v = 0; v < NV; ++v) o[v] = pair2<T2>(s[v], f[v]);
}

// Stage 2 over the rows [k chunk, (k+1) chunk): one lane per target entry;
// FP64 sums whatever the table type.
template <class T2, int NV>
__global__ __launch_bounds__(256) void k_fac_rows(const uint32_t *__restrict__ ta, const uint32_t *__restrict__ tb,
                                                  int nt, const S2Row *__restrict__ rows, int nrows, int chunk,
                                                  const double *__restrict__ x, double lambda,
                                                  double2 *__restrict__ part2)
{
    __shared__ const T2 *s_u[S2_BATCH];
    __shared__ const T2 *s_v[S2_BATCH];
    __shared__ double s_w[NV][S2_BATCH];
    __shared__ int s_f[S2_BATCH];  // bit 0: u materialised, bit 1: v materialised
    const int I = blockIdx.x * 256 + threadIdx.x;
    const bool live = I < nt;
    const int r0 = blockIdx.y * chunk, r1 = min(nrows, r0 + chunk);
    const uint32_t a = live ? ta[I] : 0, b = live ? tb[I] : 0;
    double m[NV], s[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) m[v] = s[v] = 0.0;
    for (int base = r0; base < r1; base += S2_BATCH) {
        const int nb = min(S2_BATCH, r1 - base);
        __syncthreads();
        if ((int)threadIdx.x < nb) {
            const S2Row r = rows[base + threadIdx.x];
            s_u[threadIdx.x] = (const T2 *)r.u;
            s_v[threadIdx.x] = (const T2 *)r.v;
            s_f[threadIdx.x] = (r.umat ? 1 : 0) | (r.vmat ? 2 : 0);
#pragma unroll
            for (int v = 0; v < NV; ++v) s_w[v][threadIdx.x] = r.wdet >= 0 ? x[(size_t)r.wdet * NV + v] : 1.0;
        }
        __syncthreads();
        if (live) {
#pragma unroll 8
            for (int k = 0; k < nb; ++k) {
                const int fl = NV > 1 ? s_f[k] : 0;
                typedef typename raw2<T2>::type R2;
                const gptr<R2> U = (gptr<R2>)s_u[k] + ((fl & 1) ? (size_t)a * NV : (size_t)a);  // alpha side {S, F}
                const gptr<R2> V = (gptr<R2>)s_v[k] + ((fl & 2) ? (size_t)b * NV : (size_t)b);  // beta side {S, F}
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const R2 u = U[(fl & 1) ? v : 0];
                    const R2 e = V[(fl & 2) ? v : 0];
                    const double w = s_w[v][k];
                    const double us = w * (double)u.x, uf = w * (double)u.y, vs = (double)e.x;
                    m[v] = fma(uf, vs, m[v]);
                    m[v] = fma(us, fma(lambda, vs, (double)e.y), m[v]);
                    s[v] = fma(us, vs, s[v]);
                }
            }
        }
    }
    if (live) {
#pragma unroll
        for (int v = 0; v < NV; ++v) part2[((size_t)blockIdx.y * nt + I) * NV + v] = make_double2(m[v], s[v]);
    }
}

// Stage 1 for a block of groups of one source parent and side whose entries
// index few components (K <= BLK_KMAX): U = W T on FP64 MFMA, with W the
// groups x components matrix of the entries' x weights (zero where a group
// has no entry) and T the components' table rows, read once per block
// instead of once per entry.  Output rows as k_fac_mat's: group g's row at
// out + g nt NV, {S, F} per target component and vector.
struct BlkItem {
    const void *tab;        // table [source component][target component]
    const uint32_t *rows;   // the block's K source components (table rows)
    const double *w;        // W [NV][G][Kp]
    void *out;
    int nt, G, K, Kp;
};
// Dense weights of one block, per apply: w[(v G + g) Kp + k] = x[d[g Kp + k] NV + v] (0 where d < 0).
struct BlkDense {
    const int *d;
    double *w;
    int G, Kp;
};
constexpr int BLK_GMAX = 128;  // groups per block (8 row blocks of 16)
constexpr int BLK_KMAX = 256;  // components per block
constexpr int BLK_KC = 16;     // components per LDS stage
constexpr int BLK_NC = 128;    // real columns ({S, F} of 64 target components) per workgroup
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int NV>
__global__ void k_fac_densify(const BlkDense *__restrict__ blks, const double *__restrict__ x)
{
    const BlkDense b = blks[blockIdx.y];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= b.G * b.Kp) return;
    const int d = b.d[e], g = e / b.Kp, k = e % b.Kp;
#pragma unroll
    for (int v = 0; v < NV; ++v) b.w[((size_t)v * b.G + g) * b.Kp + k] = d >= 0 ? x[(size_t)d * NV + v] : 0.0;
}

// One workgroup: all G groups x BLK_NC real columns of one block; each of the
// 4 waves owns 32 columns (two 16-column MFMA blocks) and every row block,
// v_mfma_f64_16x16x4f64 over K in LDS stages of BLK_KC; the next stage is
// loaded into registers while the current one is multiplied.  FP64 sums
// whatever the table type.  Used by one-vector applies: with two vectors the
// per-group sums read each row once for both and measured faster.
template <class T2, int NV>
__global__ __launch_bounds__(256, NV == 1 ? 2 : 1) void k_fac_block(const BlkItem *__restrict__ items)
{
    constexpr int NCOMP = BLK_NC / 2, TPT = BLK_KC * NCOMP / 256, WPT = BLK_GMAX * BLK_KC / 256;
    __shared__ double Ws[NV][BLK_GMAX][BLK_KC + 1];
    __shared__ double Ts[BLK_KC][BLK_NC + 2];
    const BlkItem it = items[blockIdx.y];
    const int ncol = 2 * it.nt, n0 = blockIdx.x * BLK_NC;
    if (n0 >= ncol) return;  // uniform over the workgroup, before any barrier
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int nms = (it.G + 15) >> 4;
    typedef typename raw2<T2>::type R2;
    const gptr<R2> tab = (gptr<R2>)it.tab;
    R2 tr[TPT];
    double wr[NV][WPT];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int j = 0; j < TPT; ++j) {
            const int e = tid + j * 256, k = e / NCOMP, t = n0 / 2 + e % NCOMP;
            tr[j] = (k0 + k < it.K && t < it.nt) ? tab[(size_t)it.rows[k0 + k] * it.nt + t] : R2{0, 0};
        }
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int j = 0; j < WPT; ++j) {
                const int e = tid + j * 256, g = e / BLK_KC, k = e % BLK_KC;
                wr[v][j] = (g < it.G && k0 + k < it.K) ? it.w[((size_t)v * it.G + g) * it.Kp + k0 + k] : 0.0;
            }
    };
    f64x4 acc[NV][BLK_GMAX / 16][2];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int ms = 0; ms < BLK_GMAX / 16; ++ms) acc[v][ms][0] = acc[v][ms][1] = f64x4{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    for (int k0 = 0; k0 < it.K; k0 += BLK_KC) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < TPT; ++j) {
            const int e = tid + j * 256, k = e / NCOMP, cc = e % NCOMP;
            Ts[k][2 * cc] = (double)tr[j].x;
            Ts[k][2 * cc + 1] = (double)tr[j].y;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int j = 0; j < WPT; ++j) {
                const int e = tid + j * 256;
                Ws[v][e / BLK_KC][e % BLK_KC] = wr[v][j];
            }
        __syncthreads();
        if (k0 + BLK_KC < it.K) fetch(k0 + BLK_KC);  // the next stage in flight during the products
#pragma unroll
        for (int kk = 0; kk < BLK_KC / 4; ++kk) {
            const double b0 = Ts[kk * 4 + lk][wave * 32 + lr], b1 = Ts[kk * 4 + lk][wave * 32 + 16 + lr];
#pragma unroll
            for (int ms = 0; ms < BLK_GMAX / 16; ++ms)
                if (ms < nms) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const double a = Ws[v][ms * 16 + lr][kk * 4 + lk];
                        acc[v][ms][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, acc[v][ms][0], 0, 0, 0);
                        acc[v][ms][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, acc[v][ms][1], 0, 0, 0);
                    }
                }
        }
    }
    typedef typename elem_of<T2>::type T;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        const int col = n0 + wave * 32 + nb * 16 + lr;
        if (col >= ncol) continue;
        const int t = col >> 1, part = col & 1;
#pragma unroll
        for (int ms = 0; ms < BLK_GMAX / 16; ++ms)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int g = ms * 16 + 4 * r + lk;
                if (g < it.G)
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        T *o = (T *)((T2 *)it.out + ((size_t)g * it.nt + t) * NV + v);
                        o[part] = (T)acc[v][ms][nb][r];
                    }
            }
    }
}

// y[det_I NV + v] += sum_k part2[k][I][v] (chunks in order); the same for ys.
template <int NV>
__global__ void k_fac_scatter(const uint32_t *__restrict__ tdet, int nt, int nk, const double2 *__restrict__ part2,
                              double *__restrict__ y, double *__restrict__ ys)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nt) return;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double m = 0.0, s = 0.0;
        for (int k = 0; k < nk; ++k) {
            const double2 e = part2[((size_t)k * nt + I) * NV + v];
            m += e.x;
            s += e.y;
        }
        const size_t d = (size_t)tdet[I] * NV + v;
        y[d] += m;
        if (ys) ys[d] += s;
    }
}

// Sums a[k] over the 64 lanes of the wave for every k < V (V a power of two,
// 2 <= V <= 64) by recursive halving: at offset o a lane keeps the upper half
// of its values if lane & o, else the lower half, and adds its partner's copy
// of that half; then the remaining offsets sum the one value left.  On return
// a[0] holds the total of value rs_index<V>(lane), the same in every lane with
// equal bits 5 .. 6 - log2(V).
template <int V, int L = V, int O = 32>
__device__ inline void wave_reduce_scatter(double (&a)[V], int lane)
{
    if constexpr (L > 1) {
        const bool up = (lane & O) != 0;
#pragma unroll
        for (int k = 0; k < L / 2; ++k) {
            const double send = up ? a[k] : a[k + L / 2];
            const double keep = up ? a[k + L / 2] : a[k];
            a[k] = keep + __shfl_xor(send, O);
        }
        wave_reduce_scatter<V, L / 2, O / 2>(a, lane);
    } else {
#pragma unroll
        for (int o = O; o > 0; o >>= 1) a[0] += __shfl_xor(a[0], o);
    }
}

template <int V>
__device__ inline int rs_index(int lane)
{
    int k = 0;
#pragma unroll
    for (int L = V, o = 32; L > 1; L >>= 1, o >>= 1)
        if (lane & o) k += L / 2;
    return k;
}

// Dual stage 1: workgroup = DUAL_ROWS rows x DUAL_COLS columns of one table;
// lane l of wave w owns 4 columns, VEC adjacent ones per 16-byte load (VEC =
// 1 for double2, 2 for float2) with the lanes contiguous, so that each load
// instruction reads 1 KB in one piece (forward sums over the rows in
// registers); each wave reduces a row's reverse
// products over its 256 columns, and the 4 waves' row partials are summed in
// order through LDS.
template <class T2, int NV>
__global__ __launch_bounds__(256) void k_fac_dual(const DualItem *__restrict__ items, const double *__restrict__ x)
{
    __shared__ double s_wf[NV][DUAL_ROWS];
    __shared__ double s_rev[4][DUAL_ROWS][2 * NV];
    const DualItem it = items[blockIdx.y];
    if ((int)blockIdx.x >= it.nchunk * it.nstrip) return;  // uniform over the workgroup
    const int chunk = blockIdx.x / it.nstrip, strip = blockIdx.x % it.nstrip;
    const int r0 = chunk * DUAL_ROWS, nr = min(it.ns - r0, DUAL_ROWS);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int VEC = sizeof(T2) == 8 ? 2 : 1;
    const int cbase = strip * DUAL_COLS + wave * 256 + lane * VEC;
    auto col = [&](int j) { return cbase + (j / VEC) * 64 * VEC + j % VEC; };
    if (it.sym && r0 > strip * DUAL_COLS + DUAL_COLS - 1) {
        // wholly below the diagonal of a symmetric table: zero partials
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (col(j) < it.nt) {
#pragma unroll
                for (int v = 0; v < NV; ++v) it.fpart[((size_t)chunk * it.nt + col(j)) * NV + v] = make_double2(0.0, 0.0);
            }
        for (int i = threadIdx.x; i < nr; i += 256) {
#pragma unroll
            for (int v = 0; v < NV; ++v) it.rpart[((size_t)strip * it.ns + r0 + i) * NV + v] = make_double2(0.0, 0.0);
        }
        return;  // uniform over the workgroup
    }
    // a symmetric table's elements count forward for s <= t, reverse for s < t
    const bool sym = it.sym != 0;
    for (int i = threadIdx.x; i < nr; i += 256) {
        const int d = it.fdet[r0 + i];
#pragma unroll
        for (int v = 0; v < NV; ++v) s_wf[v][i] = d >= 0 ? x[(size_t)d * NV + v] : 0.0;
    }
    double wr[4][NV];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int d = col(j) < it.nt ? it.rdet[col(j)] : -1;
#pragma unroll
        for (int v = 0; v < NV; ++v) wr[j][v] = d >= 0 ? x[(size_t)d * NV + v] : 0.0;
    }
    __syncthreads();
    double fS[4][NV], fF[4][NV];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < NV; ++v) fS[j][v] = fF[j][v] = 0.0;
    typedef typename raw2<T2>::type R2;
    const gptr<R2> tab = (gptr<R2>)it.tab;
    // in-range columns and rows for every load (no branches around the loads):
    // a clamped column has reverse weight 0 and its forward sums are not
    // stored, a clamped row has forward weight 0 and its reverse sum is not
    // stored
    int colc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) colc[j] = min(col(j), it.nt - 1);
    // rows in batches of DUAL_RB: reverse products a[(r NV + v) 2 + {S, F}]
    // reduced over the wave together
    constexpr int DUAL_RB = NW_DUAL_RB, V = 2 * NV * DUAL_RB;
    const int k_mine = rs_index<V>(lane);
    const bool writer = (lane & (64 / V - 1)) == 0;
    for (int i0 = 0; i0 < nr; i0 += DUAL_RB) {
        double a[V];
#pragma unroll
        for (int r = 0; r < DUAL_RB; ++r) {
            const int i = i0 + r;
            const gptr<R2> row = tab + (size_t)(r0 + min(i, nr - 1)) * it.nt;
            R2 e[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) e[j] = row[colc[j]];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double w = i < nr ? s_wf[v][i] : 0.0;
                double pS = 0.0, pF = 0.0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const double es = (double)e[j].x, ef = (double)e[j].y;
                    const double wf = (sym && r0 + i > col(j)) ? 0.0 : w;
                    const double wrj = (sym && r0 + i >= col(j)) ? 0.0 : wr[j][v];
                    fS[j][v] = fma(wf, es, fS[j][v]);
                    fF[j][v] = fma(wf, ef, fF[j][v]);
                    pS = fma(es, wrj, pS);
                    pF = fma(ef, wrj, pF);
                }
                a[(r * NV + v) * 2] = pS;
                a[(r * NV + v) * 2 + 1] = pF;
            }
        }
        wave_reduce_scatter<V>(a, lane);
        const int r = k_mine / (2 * NV);
        if (writer && i0 + r < nr) s_rev[wave][i0 + r][k_mine % (2 * NV)] = a[0];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (col(j) < it.nt) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
                it.fpart[((size_t)chunk * it.nt + col(j)) * NV + v] = make_double2(fS[j][v], fF[j][v]);
        }
    __syncthreads();
    for (int i = threadIdx.x; i < nr; i += 256) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double S = 0.0, F = 0.0;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                S += s_rev[w][i][2 * v];
                F += s_rev[w][i][2 * v + 1];
            }
            it.rpart[((size_t)strip * it.ns + r0 + i) * NV + v] = make_double2(S, F);
        }
    }
}

// Materialised rows of the dual items: forward over nt, reverse over ns,
// partials summed in chunk / strip order.
template <class T2, int NV>
__global__ void k_fac_dual_reduce(const DualItem *__restrict__ items)
{
    const DualItem it = items[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < it.nt) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double S = 0.0, F = 0.0;
            for (int k = 0; k < it.nchunk; ++k) {
                const double2 e = it.fpart[((size_t)k * it.nt + i) * NV + v];
                S += e.x;
                F += e.y;
            }
            if (it.sym)  // the lower triangle's share, from the transposed reads
                for (int k = 0; k < it.nstrip; ++k) {
                    const double2 e = it.rpart[((size_t)k * it.ns + i) * NV + v];
                    S += e.x;
                    F += e.y;
                }
            ((T2 *)it.out_f)[(size_t)i * NV + v] = pair2<T2>(S, F);
        }
    } else if (!it.sym && i < it.nt + it.ns) {
        const int r = i - it.nt;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double S = 0.0, F = 0.0;
            for (int k = 0; k < it.nstrip; ++k) {
                const double2 e = it.rpart[((size_t)k * it.ns + r) * NV + v];
                S += e.x;
                F += e.y;
            }
            ((T2 *)it.out_r)[(size_t)r * NV + v] = pair2<T2>(S, F);
        }
    }
}

// Diagonals from the same-parent tables (one_body_diagonals, onebody.rs:358-405;
// fill_one_body_diagonal_block :1068).
__global__ void k_fac_diag(const uint32_t *__restrict__ tdet, const uint32_t *__restrict__ ta,
                           const uint32_t *__restrict__ tb, int nt, const double2 *__restrict__ taba, int na,
                           const double2 *__restrict__ tabb, int nb, double lambda, double *__restrict__ md,
                           double *__restrict__ sd)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nt) return;
    const double2 A = taba[(size_t)ta[I] * na + ta[I]];
    const double2 B = tabb[(size_t)tb[I] * nb + tb[I]];
    const double s = A.x * B.x;
    if (md) md[tdet[I]] = A.y * B.x + A.x * B.y + lambda * s;
    if (sd) sd[tdet[I]] = s;
}

__global__ void k_fac_zero(double *__restrict__ p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// FP64 tables rounded to nearest FP32 (the FP32 table copies).
__global__ void k_fac_lower(const double2 *__restrict__ in, float2 *__restrict__ out, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double2 v = in[i];
        out[i] = make_float2((float)v.x, (float)v.y);
    }
}

__global__ void k_fac_widen(const float2 *__restrict__ in, double2 *__restrict__ out, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double2((double)in[i].x, (double)in[i].y);
}

unsigned blocks_of(long long n, int b) { return (unsigned)((n + b - 1) / b); }

// Owning device allocation.
template <class T>
struct DVec {
    T *p = nullptr;
    size_t n = 0;
    DVec() = default;
    DVec(const DVec &) = delete;
    DVec &operator=(const DVec &) = delete;
    DVec(DVec &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    ~DVec() { if (p) (void)hipFree(p); }
    void alloc(size_t cnt)
    {
        if (p) HIPCHK(hipFree(p));
        p = dalloc<T>(cnt);
        n = cnt;
    }
    void upload(const std::vector<T> &v)
    {
        alloc(v.size());
        if (!v.empty()) HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    }
};

// A group of source entries of one parent sharing one component: keyed by the
// beta component (side 1: alpha rows summed) or the alpha component (side 0:
// beta rows summed).  Entries [m0, m1) of the parent's materialisation list;
// a one-entry group keeps its entry in `single` instead.
struct Group {
    int side, key, m0, m1;
    int single_a, single_b;
    uint32_t single_det;
};

struct ParentSpace {
    int first_det = -1, nt = 0;
    int ncomp[2] = {0, 0};
    long long comp_off[2] = {0, 0};
    std::vector<uint32_t> tdet, ta, tb;  // target entries in determinant order
    DVec<uint32_t> d_tdet, d_ta, d_tb;
    int maxrank[2] = {0, 0};
    std::vector<Group> groups;
    long long m_off = 0;  // offset of this parent's materialisation entries in the global list
    int ngroup_a = 0, ngroup_b = 0, nsingle = 0;
};

// Greedy vertex cover of the bipartite (alpha component, beta component)
// graph whose edges are the parent's entries: repeatedly take the component
// with the most uncovered entries and group those entries under it.
void cover_groups(ParentSpace &P, std::vector<uint32_t> &midx, std::vector<uint32_t> &mdet)
{
    const int na = P.ncomp[0], nb = P.ncomp[1], n = P.nt;
    std::vector<std::vector<int>> by[2];
    by[0].resize(na);
    by[1].resize(nb);
    for (int k = 0; k < n; ++k) {
        by[0][P.ta[k]].push_back(k);
        by[1][P.tb[k]].push_back(k);
    }
    std::vector<int> deg[2];
    deg[0].resize(na);
    deg[1].resize(nb);
    std::vector<std::tuple<int, int, int>> heap;  // (degree, -side, -key): max-heap pops the largest degree
    for (int sd = 0; sd < 2; ++sd)
        for (int c = 0; c < (sd ? nb : na); ++c) {
            deg[sd][c] = (int)by[sd][c].size();
            heap.emplace_back(deg[sd][c], -sd, -c);
        }
    std::make_heap(heap.begin(), heap.end());
    std::vector<char> alive(n, 1);
    P.m_off = (long long)midx.size();
    while (!heap.empty()) {
        std::pop_heap(heap.begin(), heap.end());
        auto [d, nsd, nc] = heap.back();
        heap.pop_back();
        const int sd = -nsd, c = -nc;
        if (d == 0 || d != deg[sd][c]) continue;  // stale entry
        std::vector<int> mem;
        for (int k : by[sd][c])
            if (alive[k]) mem.push_back(k);
        deg[sd][c] = 0;
        for (int k : mem) {
            alive[k] = 0;
            const int osd = 1 - sd, oc = (int)(osd ? P.tb[k] : P.ta[k]);
            if (--deg[osd][oc] > 0) {
                heap.emplace_back(deg[osd][oc], -osd, -oc);
                std::push_heap(heap.begin(), heap.end());
            }
        }
        Group g{};
        g.side = sd == 0 ? 0 : 1;
        g.key = c;
        if (mem.size() == 1) {
            g.m0 = g.m1 = -1;
            g.single_a = (int)P.ta[mem[0]];
            g.single_b = (int)P.tb[mem[0]];
            g.single_det = P.tdet[mem[0]];
            ++P.nsingle;
        } else {
            // entries ordered by the summed component (table rows in ascending order)
            std::sort(mem.begin(), mem.end(), [&](int u, int v) {
                return sd == 1 ? P.ta[u] < P.ta[v] : P.tb[u] < P.tb[v];
            });
            g.m0 = (int)(midx.size() - P.m_off);
            for (int k : mem) {
                midx.push_back(sd == 1 ? P.ta[k] : P.tb[k]);
                mdet.push_back(P.tdet[k]);
            }
            g.m1 = (int)(midx.size() - P.m_off);
            (sd == 1 ? P.ngroup_a : P.ngroup_b)++;
        }
        P.groups.push_back(g);
    }
}

uint64_t spin_key(const PackedDet &d, int sp)
{
    const int r = sp ? d.rb : d.ra;
    const uint8_t *ix = sp ? d.ib : d.ia;
    const int ph = sp ? d.phb : d.pha;
    uint64_t k = (uint64_t)r << 57;
    for (int q = 0; q < 4; ++q) {
        const uint64_t h = q < r ? ix[q] : 0, p = q < r ? ix[4 + q] : 0;
        k |= h << (50 - 7 * q);
        k |= p << (22 - 7 * q);
    }
    return k | (ph < 0 ? 1u : 0u);
}

SpinComp spin_comp(const PackedDet &d, int sp)
{
    SpinComp c{};
    c.r = sp ? d.rb : d.ra;
    c.ph = sp ? d.phb : d.pha;
    const uint8_t *ix = sp ? d.ib : d.ia;
    for (int q = 0; q < 4; ++q) {
        c.i[q] = q < c.r ? ix[q] : 0;
        c.i[4 + q] = q < c.r ? ix[4 + q] : 0;
    }
    return c;
}

// A dual pass chosen for one unordered parent pair and spin: the table
// Tab(q, p) (rows = components of p), the forward group gp of p (target q)
// and the reverse group gq of q (target p); det_off: its fdet (ns entries)
// then rdet (nt entries) in the uploaded determinant-index array.
struct DualSel {
    int q, p, sp, gp, gq;  // q == p: a same-parent (symmetric) table, gq = gp
    long long det_off;
};

// Per ordered parent pair: its stage-1 items and stage-2 rows (ranges into the
// uploaded arrays) and its materialisation footprint.
struct PairPlan {
    int q = -1, p = -1;
    long long item0 = 0, item1 = 0, row0 = 0, row1 = 0, blk0 = 0, blk1 = 0;
    size_t mat_elems = 0;
};

// A block of one source parent's groups (one side) whose stage 1 is a dense
// product (k_fac_block): the groups (indices into P.groups), the sorted
// components their entries index, and offsets of its component list, entry
// determinant matrix (G x Kp, -1 = none) and dense weights (2 x G x Kp).
struct BlkSrc {
    int p = -1, side = 0, G = 0, K = 0, Kp = 0;
    std::vector<int> groups;
    long long row_off = 0, d_off = 0, w_off = 0;
};

}  // namespace

struct nw_factor {
    nw_ctx *c = nullptr;
    int nref = 0, mode = 0, storage = 0;
    uint64_t ndet = 0;
    unsigned long long build_gen = 0, fock_gen = 0;
    std::vector<ParentSpace> par;
    std::vector<uint32_t> aid, bid;  // per determinant
    DVec<SpinComp> d_comp[2];
    // tables: per (Q * nref + P) * 2 + spin, offset in elements into d_tab / d_tab32 (-1 = none)
    std::vector<long long> tab_off;
    DVec<double2> d_tab;    // FP64 tables (resident), or one pair's tables (regenerated)
    DVec<float2> d_tab32;   // FP32 copies of the resident tables, made on first FP32 use
    int prec = NW_PREC_FP64;  // tables nw_factor_apply reads: NW_PREC_FP64 or NW_PREC_MIXED (FP32)
    size_t tab_elems = 0;
    // plan: stage-1 items and stage-2 rows per table type (0 FP64, 1 FP32)
    // and vector count (1, 2)
    DVec<uint32_t> d_midx, d_mdet;
    DVec<S1Item> d_items[2][2];
    DVec<S2Row> d_rows[2][2];
    bool planned[2][2] = {{false, false}, {false, false}};
    std::vector<PairPlan> plans[2][2];  // per plan (lo, nv): ordered by target, then source
    std::vector<uint32_t> h_midx;  // host copy of the summed-component indices
    size_t mat_elems = 0;          // materialisation entries of one vector
    DVec<double2> d_mat, d_part2;  // d_mat holds double2 or float2 rows
    // dual passes (resident tables): selections, their row / column
    // determinants, per-plan items and the partial-sum buffer
    std::vector<DualSel> duals;
    std::map<long long, std::pair<int, int>> dual_of;  // (target, source, group) -> (dual, 0 fwd / 1 rev)
    // dense-product blocks of source groups: per parent, group -> (block, position) or (-1, -1)
    std::vector<BlkSrc> blks;
    std::vector<std::vector<std::pair<int, int>>> blk_of;
    DVec<uint32_t> d_brow;
    DVec<int> d_bD;
    DVec<double> d_bW;
    DVec<BlkDense> d_bdense;
    DVec<BlkItem> d_blk[2][2];
    int blk_gmax = 0, blk_kmax = 0;
    DVec<int> d_dualdet;
    DVec<DualItem> d_dual[2][2];
    DVec<double2> d_dpart;
    size_t dual_mat = 0, dual_part = 0;  // per vector: materialised entries, partial entries
    uint64_t apply_bytes[2] = {0, 0};    // table bytes one apply reads (FP64, FP32 tables)
    int maxnt = 0;
    double t_ms[2] = {0, 0};  // last apply: total, table regeneration
};

namespace {

int comps_of(const nw_factor *f, int p, int sp) { return f->par[p].ncomp[sp]; }

// Table element count of one ordered pair and spin ([source][target]).
size_t table_elems(const nw_factor *f, int q, int p, int sp)
{
    return (size_t)comps_of(f, p, sp) * (size_t)comps_of(f, q, sp);
}

size_t elem_bytes(int lo) { return lo ? sizeof(float2) : sizeof(double2); }

long long dual_key(const nw_factor *f, int target, int source, int group)
{
    return ((long long)(target * f->nref + source) << 32) | (long long)group;
}

// Table (q target, p source, spin sp) into out (float2 when lo, else
// double2); when outT is not null also its transpose, the table (p target,
// q source).
void launch_table(nw_factor *f, int q, int p, int sp, void *out, void *outT = nullptr, int lo = 0)
{
    nw_ctx *c = f->c;
    const ParentSpace &Q = f->par[q], &P = f->par[p];
    const bool target_left = Q.first_det <= P.first_det;
    const int pair = target_left ? q * c->nref + p : p * c->nref + q;
    const int nt = Q.ncomp[sp], ns = P.ncomp[sp];
    if (nt == 0 || ns == 0) return;
    dim3 grid(blocks_of(nt, 16), blocks_of(ns, 16));
    if (lo)
        hipLaunchKernelGGL(k_fac_table<float2>, grid, dim3(256), 0, c->st, c->d_pinfo, c->d_slab, pair, sp,
                           f->d_comp[sp].p + Q.comp_off[sp], nt, f->d_comp[sp].p + P.comp_off[sp], ns,
                           target_left ? 1 : 0, c->nmo, c->tol, f->mode, (float2 *)out, (float2 *)outT);
    else
        hipLaunchKernelGGL(k_fac_table<double2>, grid, dim3(256), 0, c->st, c->d_pinfo, c->d_slab, pair, sp,
                           f->d_comp[sp].p + Q.comp_off[sp], nt, f->d_comp[sp].p + P.comp_off[sp], ns,
                           target_left ? 1 : 0, c->nmo, c->tol, f->mode, (double2 *)out, (double2 *)outT);
}

// FP32 copies of the resident FP64 tables.
void lower_tables(nw_factor *f)
{
    if (f->tab_elems == 0) return;
    const long long n = (long long)f->tab_elems;
    const unsigned g = (unsigned)std::min<long long>(8192, blocks_of(n, 256));
    hipLaunchKernelGGL(k_fac_lower, dim3(g), dim3(256), 0, f->c->st, f->d_tab.p, f->d_tab32.p, n);
}

void build_tables(nw_factor *f)
{
    const int nref = f->nref;
    // each unordered pair once, with the earlier parent as target (bra); the
    // reverse pair's table is written as the transpose
    for (int q = 0; q < nref; ++q)
        for (int p = 0; p < nref; ++p)
            for (int sp = 0; sp < 2; ++sp) {
                const long long off = f->tab_off[(q * nref + p) * 2 + sp];
                if (off < 0) continue;
                if (q == p) launch_table(f, q, p, sp, f->d_tab.p + off);
                else if (f->par[q].first_det < f->par[p].first_det)
                    launch_table(f, q, p, sp, f->d_tab.p + off, f->d_tab.p + f->tab_off[(p * nref + q) * 2 + sp]);
            }
    if (f->d_tab32.p) lower_tables(f);
    HIPCHK(hipGetLastError());
    f->build_gen = f->c->build_gen;
    f->fock_gen = f->c->fock_gen;
}

void check_current(const nw_factor *f)
{
    if (!f->c->built) throw NwError(NW_E_STATE, "nw_build first");
    if (f->storage == NW_FACTOR_DEVICE && (f->build_gen != f->c->build_gen || f->fock_gen != f->c->fock_gen))
        throw NwError(NW_E_STATE, "factor tables are stale (the context was rebuilt or its Fock updated): "
                                  "call nw_factor_refresh");
    if (f->storage == NW_FACTOR_TRANSIENT && f->build_gen != f->c->build_gen)
        throw NwError(NW_E_STATE, "the context was rebuilt since nw_factor_create");
}

// Pointer of the FP64 (q, p, sp) table: resident, or regenerated into `scratch`.
double2 *table_ptr(nw_factor *f, int q, int p, int sp, double2 *scratch)
{
    if (f->storage == NW_FACTOR_DEVICE) return f->d_tab.p + f->tab_off[(q * f->nref + p) * 2 + sp];
    launch_table(f, q, p, sp, scratch);
    return scratch;
}

// Stage-1 items and stage-2 rows of every ordered pair.  Resident tables:
// the materialisation buffer holds one target's pairs at a time (pairs of a
// target are contiguous); regenerated tables: one pair at a time, both
// tables in the scratch table slot.  lo: pointers into the FP32 tables and
// float2 materialised rows.  nv: vectors per apply (materialised rows of nv
// interleaved entries).
void make_plans(nw_factor *f, int lo, int nv)
{
    const int nref = f->nref;
    std::vector<S1Item> items;
    std::vector<S2Row> rows;
    size_t mat_need = 0;
    const bool dev = f->storage == NW_FACTOR_DEVICE;
    const size_t es = elem_bytes(lo);
    const char *tbase = dev && lo ? (const char *)f->d_tab32.p : (const char *)f->d_tab.p;
    char *mbase = (char *)f->d_mat.p;
    std::vector<BlkItem> blk_items;
    // dense-product blocks in one-vector plans; two-vector plans keep the per-group sums
    const bool use_blk = nv == 1;
    std::vector<PairPlan> &plans = f->plans[lo][nv - 1];
    plans.clear();
    // dual passes: their materialised rows open the buffer, then the rows of
    // one target at a time
    std::vector<DualItem> dual_items;
    std::vector<std::array<char *, 2>> dual_out(f->duals.size());
    size_t mo_d = 0, po = 0;
    for (size_t d = 0; d < f->duals.size(); ++d) {
        const DualSel &D = f->duals[d];
        const int ns = f->par[D.p].ncomp[D.sp], nt = f->par[D.q].ncomp[D.sp];
        DualItem it{};
        it.sym = D.q == D.p ? 1 : 0;
        it.tab = tbase + es * (size_t)f->tab_off[(D.q * nref + D.p) * 2 + D.sp];
        it.out_f = mbase + es * mo_d;
        mo_d += (size_t)nt * nv;
        it.out_r = it.sym ? it.out_f : mbase + es * mo_d;
        if (!it.sym) mo_d += (size_t)ns * nv;
        it.fdet = f->d_dualdet.p + D.det_off;
        it.rdet = it.sym ? it.fdet : it.fdet + ns;
        it.ns = ns;
        it.nt = nt;
        it.nchunk = (ns + DUAL_ROWS - 1) / DUAL_ROWS;
        it.nstrip = (nt + DUAL_COLS - 1) / DUAL_COLS;
        it.fpart = f->d_dpart.p + po;
        po += (size_t)it.nchunk * nt * nv;
        it.rpart = f->d_dpart.p + po;
        po += (size_t)it.nstrip * ns * nv;
        dual_out[d] = {(char *)it.out_f, (char *)it.out_r};
        dual_items.push_back(it);
    }
    if (po > f->d_dpart.n) throw NwError(NW_E_STATE, "factor plan exceeds its partial-sum buffer");
    // table rows an apply touches: whole dual tables, and every other row that
    // stage 1 sums or stage 2 gathers from (each counted once per table)
    std::vector<std::vector<char>> touched((size_t)nref * nref * 2);
    std::vector<char> whole((size_t)nref * nref * 2, 0);  // tables a dual pass reads whole
    for (const DualSel &D : f->duals) whole[((size_t)D.q * nref + D.p) * 2 + D.sp] = 1;
    auto touch = [&](int tq, int tp, int sp, int row) {
        if (whole[((size_t)tq * nref + tp) * 2 + sp]) return;
        auto &t = touched[((size_t)tq * nref + tp) * 2 + sp];
        if (t.empty()) t.assign(f->par[tp].ncomp[sp], 0);
        t[row] = 1;
    };
    uint64_t bytes = 0;
    for (const DualSel &D : f->duals)
        bytes += (D.q == D.p ? (uint64_t)f->par[D.p].ncomp[D.sp] * (f->par[D.p].ncomp[D.sp] + 1ull) / 2
                             : (uint64_t)f->par[D.p].ncomp[D.sp] * (uint64_t)f->par[D.q].ncomp[D.sp]) * es;
    for (int q = 0; q < nref; ++q) {
        const ParentSpace &Q = f->par[q];
        if (Q.nt == 0) continue;
        size_t mat_q = mo_d;
        for (int p = 0; p < nref; ++p) {
            const ParentSpace &P = f->par[p];
            if (P.nt == 0) continue;
            PairPlan pl;
            pl.q = q;
            pl.p = p;
            const int nta = Q.ncomp[0], ntb = Q.ncomp[1];
            const char *taba, *tabb;
            if (dev) {
                taba = tbase + es * (size_t)f->tab_off[(q * nref + p) * 2 + 0];
                tabb = tbase + es * (size_t)f->tab_off[(q * nref + p) * 2 + 1];
            } else {
                taba = tbase;
                tabb = tbase + es * table_elems(f, q, p, 0);
            }
            size_t mo = dev ? mat_q : 0;
            pl.item0 = (long long)items.size();
            pl.row0 = (long long)rows.size();
            // dense-product blocks of p: their groups' rows, contiguous per block
            pl.blk0 = (long long)blk_items.size();
            std::vector<char *> blk_out;
            for (const BlkSrc &b : f->blks) {
                if (!use_blk || b.p != p) {
                    blk_out.push_back(nullptr);
                    continue;
                }
                const int len = b.side == 1 ? nta : ntb;
                char *out = mbase + es * mo;
                blk_out.push_back(out);
                blk_items.push_back({b.side == 1 ? taba : tabb, f->d_brow.p + b.row_off, f->d_bW.p + b.w_off, out, len,
                                     b.G, b.K, b.Kp});
                mo += (size_t)b.G * len * nv;
            }
            pl.blk1 = (long long)blk_items.size();
            for (int gi = 0; gi < (int)P.groups.size(); ++gi) {
                const Group &g = P.groups[gi];
                const auto du = dev ? f->dual_of.find(dual_key(f, q, p, gi)) : f->dual_of.end();
                if (g.m0 < 0) {
                    touch(q, p, 0, g.single_a);
                    touch(q, p, 1, g.single_b);
                } else {
                    touch(q, p, g.side == 1 ? 1 : 0, g.key);
                    if (du == f->dual_of.end())
                        for (int e = g.m0; e < g.m1; ++e) touch(q, p, g.side == 1 ? 0 : 1, (int)f->h_midx[P.m_off + e]);
                }
                if (du != f->dual_of.end()) {
                    char *out = dual_out[du->second.first][du->second.second];
                    if (g.side == 1) rows.push_back({out, tabb + es * (size_t)g.key * ntb, -1, 1, 0, {0, 0}});
                    else rows.push_back({taba + es * (size_t)g.key * nta, out, -1, 0, 1, {0, 0}});
                    continue;
                }
                if (g.m0 < 0) {
                    rows.push_back({taba + es * (size_t)g.single_a * nta, tabb + es * (size_t)g.single_b * ntb,
                                    (int)g.single_det, 0, 0, {0, 0}});
                    continue;
                }
                const int len = g.side == 1 ? nta : ntb;
                const auto bo = f->blk_of[p][gi];
                if (use_blk && bo.first >= 0) {
                    char *out = blk_out[bo.first] + es * ((size_t)bo.second * len * nv);
                    if (g.side == 1) rows.push_back({out, tabb + es * (size_t)g.key * ntb, -1, 1, 0, {0, 0}});
                    else rows.push_back({taba + es * (size_t)g.key * nta, out, -1, 0, 1, {0, 0}});
                    continue;
                }
                for (int e = g.m0; e < g.m1; e += S1_CHUNK) {
                    char *out = mbase + es * mo;
                    items.push_back({g.side == 1 ? taba : tabb, out, len, (int)(P.m_off + e),
                                     (int)(P.m_off + std::min(g.m1, e + S1_CHUNK)), 0});
                    if (g.side == 1) rows.push_back({out, tabb + es * (size_t)g.key * ntb, -1, 1, 0, {0, 0}});
                    else rows.push_back({taba + es * (size_t)g.key * nta, out, -1, 0, 1, {0, 0}});
                    mo += (size_t)len * nv;
                }
            }
            pl.item1 = (long long)items.size();
            pl.row1 = (long long)rows.size();
            pl.mat_elems = mo - (dev ? mat_q : 0);
            if (dev) mat_q = mo;
            else mat_need = std::max(mat_need, pl.mat_elems);
            plans.push_back(pl);
        }
        if (dev) mat_need = std::max(mat_need, mat_q);
    }
    // the float2 rows of an FP32 plan take half of the double2 buffer
    if (mat_need > f->d_mat.n) throw NwError(NW_E_STATE, "factor plan exceeds its materialisation buffer");
    for (int q = 0; q < nref; ++q)
        for (int p = 0; p < nref; ++p)
            for (int sp = 0; sp < 2; ++sp) {
                const auto &t = touched[((size_t)q * nref + p) * 2 + sp];
                bytes += (uint64_t)std::count(t.begin(), t.end(), 1) * (uint64_t)f->par[q].ncomp[sp] * es;
            }
    f->apply_bytes[lo] = bytes;
    f->d_items[lo][nv - 1].upload(items);
    f->d_rows[lo][nv - 1].upload(rows);
    f->d_blk[lo][nv - 1].upload(blk_items);
    f->d_dual[lo][nv - 1].upload(dual_items);
    f->planned[lo][nv - 1] = true;
}

// The plan of (lo, nv), made on first use; a larger materialisation buffer
// for more vectors invalidates the plans made so far.
void ensure_plan(nw_factor *f, int lo, int nv)
{
    if (f->planned[lo][nv - 1]) return;
    const size_t need = std::max<size_t>(1, f->mat_elems * (size_t)nv);
    const size_t need_part = std::max<size_t>(1, f->dual_part * (size_t)nv);
    if (need > f->d_mat.n || need_part > f->d_dpart.n) {
        HIPCHK(hipStreamSynchronize(f->c->st));
        if (need > f->d_mat.n) f->d_mat.alloc(need);
        if (need_part > f->d_dpart.n) f->d_dpart.alloc(need_part);
        for (auto &pl : f->planned) pl[0] = pl[1] = false;
    }
    make_plans(f, lo, nv);
}

size_t mat_elems_needed(const nw_factor *f)
{
    // the same walk as make_plans without pointers (one vector)
    size_t dual = 0;
    for (const DualSel &D : f->duals)
        dual += (size_t)f->par[D.q].ncomp[D.sp] + (D.q == D.p ? 0 : (size_t)f->par[D.p].ncomp[D.sp]);
    size_t need = dual;
    for (int q = 0; q < f->nref; ++q) {
        const ParentSpace &Q = f->par[q];
        if (Q.nt == 0) continue;
        size_t mq = dual;
        for (int p = 0; p < f->nref; ++p) {
            const ParentSpace &P = f->par[p];
            if (P.nt == 0) continue;
            size_t mp = 0;
            for (int gi = 0; gi < (int)P.groups.size(); ++gi) {
                const Group &g = P.groups[gi];
                if (g.m0 < 0 || f->dual_of.count(dual_key(f, q, p, gi))) continue;
                const size_t len = g.side == 1 ? Q.ncomp[0] : Q.ncomp[1];
                mp += len * (size_t)((g.m1 - g.m0 + S1_CHUNK - 1) / S1_CHUNK);
            }
            if (f->storage == NW_FACTOR_DEVICE) mq += mp;
            else need = std::max(need, mp);
        }
        need = std::max(need, mq);
    }
    return need;
}

// Dual passes over the resident tables: for each unordered pair of distinct
// parents and each spin, the largest group of either parent that sums rows of
// that spin's tables, when each covers at least a quarter of the components
// it indexes and repeats none.
void select_duals(nw_factor *f, const std::vector<uint32_t> &midx, const std::vector<uint32_t> &mdet,
                  std::vector<int> &dets)
{
    auto largest = [](const ParentSpace &P, int side) {
        int best = -1, bn = 0;
        for (int g = 0; g < (int)P.groups.size(); ++g) {
            const Group &G = P.groups[g];
            if (G.m0 >= 0 && G.side == side && G.m1 - G.m0 > bn) {
                bn = G.m1 - G.m0;
                best = g;
            }
        }
        return best;
    };
    // same-parent tables: the dominant group reads the symmetric table once,
    // its upper triangle only
    for (int p = 0; p < f->nref; ++p) {
        const ParentSpace &P = f->par[p];
        if (P.nt == 0) continue;
        for (int sp = 0; sp < 2; ++sp) {
            const int gp = largest(P, sp == 0 ? 1 : 0);
            if (gp < 0) continue;
            const Group &GP = P.groups[gp];
            const int ns = P.ncomp[sp];
            if (4 * (GP.m1 - GP.m0) < ns) continue;
            std::vector<int> fd(ns, -1);
            bool ok = true;
            for (int k = GP.m0; k < GP.m1 && ok; ++k) {
                int &slot = fd[midx[P.m_off + k]];
                ok = slot < 0;
                slot = (int)mdet[P.m_off + k];
            }
            if (!ok) continue;
            const int idx = (int)f->duals.size();
            f->duals.push_back({p, p, sp, gp, gp, (long long)dets.size()});
            dets.insert(dets.end(), fd.begin(), fd.end());
            f->dual_of[dual_key(f, p, p, gp)] = {idx, 0};
            f->dual_part += (size_t)((ns + DUAL_ROWS - 1) / DUAL_ROWS) * ns +
                            (size_t)((ns + DUAL_COLS - 1) / DUAL_COLS) * ns;
        }
    }
    for (int q = 0; q < f->nref; ++q)
        for (int p = q + 1; p < f->nref; ++p) {
            const ParentSpace &Q = f->par[q], &P = f->par[p];
            if (Q.nt == 0 || P.nt == 0) continue;
            for (int sp = 0; sp < 2; ++sp) {
                const int side = sp == 0 ? 1 : 0;  // alpha rows are summed by groups keyed by a beta component
                const int gp = largest(P, side), gq = largest(Q, side);
                if (gp < 0 || gq < 0) continue;
                const Group &GP = P.groups[gp], &GQ = Q.groups[gq];
                const int ns = P.ncomp[sp], nt = Q.ncomp[sp];
                if (4 * (GP.m1 - GP.m0) < ns || 4 * (GQ.m1 - GQ.m0) < nt) continue;
                std::vector<int> fd(ns, -1), rd(nt, -1);
                bool ok = true;
                for (int k = GP.m0; k < GP.m1 && ok; ++k) {
                    int &slot = fd[midx[P.m_off + k]];
                    ok = slot < 0;
                    slot = (int)mdet[P.m_off + k];
                }
                for (int k = GQ.m0; k < GQ.m1 && ok; ++k) {
                    int &slot = rd[midx[Q.m_off + k]];
                    ok = slot < 0;
                    slot = (int)mdet[Q.m_off + k];
                }
                if (!ok) continue;
                const int idx = (int)f->duals.size();
                f->duals.push_back({q, p, sp, gp, gq, (long long)dets.size()});
                dets.insert(dets.end(), fd.begin(), fd.end());
                dets.insert(dets.end(), rd.begin(), rd.end());
                f->dual_of[dual_key(f, q, p, gp)] = {idx, 0};
                f->dual_of[dual_key(f, p, q, gq)] = {idx, 1};
                f->dual_part += (size_t)((ns + DUAL_ROWS - 1) / DUAL_ROWS) * nt +
                                (size_t)((nt + DUAL_COLS - 1) / DUAL_COLS) * ns;
            }
        }
}

// Dense-product blocks: per source parent and side, the multi-entry groups of
// at most BLK_KMAX / 2 entries that no dual pass serves, in chunks of at most
// BLK_GMAX groups, when a chunk indexes at most BLK_KMAX components, holds at
// least NW_FAC_BLOCK_MIN_G groups (default 16) and fills at least 15% of its
// groups x components matrix (the single x single block of an arrow-shaped
// PT2 space).
// NW_FAC_BLOCK=0 keeps every group on the per-group sums.
void select_blocks(nw_factor *f, const std::vector<uint32_t> &midx, const std::vector<uint32_t> &mdet)
{
    f->blk_of.assign(f->nref, {});
    for (int p = 0; p < f->nref; ++p) f->blk_of[p].assign(f->par[p].groups.size(), {-1, -1});
    const char *on = std::getenv("NW_FAC_BLOCK");
    if (on && std::atoi(on) == 0) return;
    const char *mg = std::getenv("NW_FAC_BLOCK_MIN_G");
    const int min_g = std::max(1, mg ? std::atoi(mg) : 16);
    std::vector<uint32_t> rows;
    std::vector<int> dmat;
    for (int p = 0; p < f->nref; ++p) {
        const ParentSpace &P = f->par[p];
        if (P.nt == 0) continue;
        for (int side = 0; side < 2; ++side) {
            std::vector<int> cand;
            for (int gi = 0; gi < (int)P.groups.size(); ++gi) {
                const Group &g = P.groups[gi];
                // large groups (a reference row or column) stay on the per-group sums
                if (g.m0 < 0 || g.side != side || g.m1 - g.m0 > BLK_KMAX / 2) continue;
                bool dual = false;
                for (int q = 0; q < f->nref && !dual; ++q) dual = f->dual_of.count(dual_key(f, q, p, gi)) > 0;
                if (!dual) cand.push_back(gi);
            }
            for (size_t c0 = 0; c0 < cand.size(); c0 += BLK_GMAX) {
                BlkSrc b;
                b.p = p;
                b.side = side;
                b.groups.assign(cand.begin() + c0, cand.begin() + std::min(cand.size(), c0 + BLK_GMAX));
                std::vector<uint32_t> comp;
                long long nnz = 0;
                for (int gi : b.groups) {
                    const Group &g = P.groups[gi];
                    nnz += g.m1 - g.m0;
                    for (int e = g.m0; e < g.m1; ++e) comp.push_back(midx[P.m_off + e]);
                }
                std::sort(comp.begin(), comp.end());
                comp.erase(std::unique(comp.begin(), comp.end()), comp.end());
                b.G = (int)b.groups.size();
                b.K = (int)comp.size();
                if (b.G < min_g || b.K > BLK_KMAX || 20 * nnz < 3ll * b.G * b.K) continue;
                b.Kp = (b.K + 1) & ~1;
                b.row_off = (long long)rows.size();
                rows.insert(rows.end(), comp.begin(), comp.end());
                b.d_off = (long long)dmat.size();
                dmat.resize(dmat.size() + (size_t)b.G * b.Kp, -1);
                for (int j = 0; j < b.G; ++j) {
                    const Group &g = P.groups[b.groups[j]];
                    for (int e = g.m0; e < g.m1; ++e) {
                        const int k = (int)(std::lower_bound(comp.begin(), comp.end(), midx[P.m_off + e]) - comp.begin());
                        dmat[b.d_off + (size_t)j * b.Kp + k] = (int)mdet[P.m_off + e];
                    }
                    f->blk_of[p][b.groups[j]] = {(int)f->blks.size(), j};
                }
                f->blk_gmax = std::max(f->blk_gmax, b.G);
                f->blk_kmax = std::max(f->blk_kmax, b.Kp);
                f->blks.push_back(std::move(b));
            }
        }
    }
    if (f->blks.empty()) return;
    size_t wn = 0;
    for (BlkSrc &b : f->blks) {
        b.w_off = (long long)wn;
        wn += 2 * (size_t)b.G * b.Kp;  // room for two vectors
    }
    f->d_brow.upload(rows);
    f->d_bD.upload(dmat);
    f->d_bW.alloc(wn);
    std::vector<BlkDense> dense;
    for (const BlkSrc &b : f->blks) dense.push_back({f->d_bD.p + b.d_off, f->d_bW.p + b.w_off, b.G, b.Kp});
    f->d_bdense.upload(dense);
}

}  // namespace

extern "C" int nw_factor_create(nw_ctx *c, const nw_det *dets, uint64_t ndet, int mode, int storage,
                                nw_factor **out)
{
    if (out) *out = nullptr;
    return guarded(c, [&]() -> int {
        if (!out || !dets || ndet == 0) return fail(c, NW_E_ARG, "null arguments");
        if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
        if (storage != NW_FACTOR_DEVICE && storage != NW_FACTOR_TRANSIENT)
            return fail(c, NW_E_ARG, "storage must be NW_FACTOR_DEVICE or NW_FACTOR_TRANSIENT");
        if (ndet >= (1ull << 31)) return fail(c, NW_E_SHAPE, "too many determinants");
        PackedDet *dp = nullptr;
        const std::vector<PackedDet> h = pack(c, dets, ndet, dp);
        auto f = std::make_unique<nw_factor>();
        f->c = c;
        f->nref = c->nref;
        f->mode = mode;
        f->storage = storage;
        f->ndet = ndet;
        f->par.resize(c->nref);
        f->aid.assign(ndet, 0);
        f->bid.assign(ndet, 0);
        // ---- components: per parent and spin, unique strings sorted by (rank, holes, particles, phase)
        std::vector<std::vector<uint32_t>> members(c->nref);
        for (uint32_t i = 0; i < (uint32_t)ndet; ++i) members[h[i].parent].push_back(i);
        std::vector<SpinComp> comps[2];
        for (int p = 0; p < c->nref; ++p) {
            ParentSpace &P = f->par[p];
            const auto &mem = members[p];
            if (mem.empty()) continue;
            P.first_det = (int)mem.front();
            for (int sp = 0; sp < 2; ++sp) {
                std::vector<std::pair<uint64_t, uint32_t>> keyed;
                keyed.reserve(mem.size());
                for (uint32_t i : mem) keyed.push_back({spin_key(h[i], sp), i});
                std::sort(keyed.begin(), keyed.end());
                P.comp_off[sp] = (long long)comps[sp].size();
                int id = -1;
                for (size_t k = 0; k < keyed.size(); ++k) {
                    if (k == 0 || keyed[k].first != keyed[k - 1].first) {
                        ++id;
                        comps[sp].push_back(spin_comp(h[keyed[k].second], sp));
                        P.maxrank[sp] = std::max<int>(P.maxrank[sp], comps[sp].back().r);
                    }
                    (sp ? f->bid : f->aid)[keyed[k].second] = (uint32_t)id;
                }
                P.ncomp[sp] = id + 1;
            }
            P.nt = (int)mem.size();
            P.tdet = mem;
            for (uint32_t i : mem) {
                P.ta.push_back(f->aid[i]);
                P.tb.push_back(f->bid[i]);
            }
        }
        for (int q = 0; q < c->nref; ++q)
            for (int p = 0; p < c->nref; ++p) {
                const ParentSpace &Q = f->par[q], &P = f->par[p];
                if (Q.nt == 0 || P.nt == 0) continue;
                for (int sp = 0; sp < 2; ++sp)
                    if (Q.maxrank[sp] + P.maxrank[sp] > 4)
                        return fail(c, NW_E_RANK, "per-spin contraction rank above 4 between parents " +
                                                      std::to_string(q) + " and " + std::to_string(p));
                const bool tl = Q.first_det <= P.first_det;
                const int pair = tl ? q * c->nref + p : p * c->nref + q;
                if (!c->pinfo[pair].built)
                    return fail(c, NW_E_STATE, "reference pair (" + std::to_string(tl ? q : p) + "," +
                                                   std::to_string(tl ? p : q) + ") was not built");
            }
        for (const ParentSpace &P : f->par)
            if (P.ncomp[0] > 262140 || P.ncomp[1] > 262140)
                return fail(c, NW_E_SHAPE, "more than 262140 spin components in one parent");
        for (int sp = 0; sp < 2; ++sp) f->d_comp[sp].upload(comps[sp]);
        // ---- source groups (vertex cover) and target lists
        std::vector<uint32_t> midx, mdet;
        for (ParentSpace &P : f->par) {
            if (P.nt == 0) continue;
            P.d_tdet.upload(P.tdet);
            P.d_ta.upload(P.ta);
            P.d_tb.upload(P.tb);
            cover_groups(P, midx, mdet);
            f->maxnt = std::max(f->maxnt, P.nt);
        }
        if (midx.size() >= (1ull << 31)) return fail(c, NW_E_SHAPE, "too many source entries");
        f->d_midx.upload(midx);
        f->d_mdet.upload(mdet);
        f->h_midx = midx;
        if (storage == NW_FACTOR_DEVICE) {
            std::vector<int> dd;
            select_duals(f.get(), midx, mdet, dd);
            f->d_dualdet.upload(dd);
        }
        select_blocks(f.get(), midx, mdet);
        // ---- tables
        const int nref = c->nref;
        f->tab_off.assign((size_t)nref * nref * 2, -1);
        size_t off = 0, maxpair = 0;
        for (int q = 0; q < nref; ++q)
            for (int p = 0; p < nref; ++p) {
                if (f->par[q].nt == 0 || f->par[p].nt == 0) continue;
                size_t pair_elems = 0;
                for (int sp = 0; sp < 2; ++sp) {
                    f->tab_off[(q * nref + p) * 2 + sp] = (long long)off;
                    off += table_elems(f.get(), q, p, sp);
                    pair_elems += table_elems(f.get(), q, p, sp);
                }
                maxpair = std::max(maxpair, pair_elems);
            }
        f->tab_elems = off;
        f->d_tab.alloc(std::max<size_t>(1, storage == NW_FACTOR_DEVICE ? off : maxpair));
        f->mat_elems = mat_elems_needed(f.get());
        f->d_mat.alloc(std::max<size_t>(1, f->mat_elems));
        f->d_dpart.alloc(std::max<size_t>(1, f->dual_part));
        f->d_part2.alloc((size_t)64 * 2 * (size_t)std::max(1, f->maxnt));
        make_plans(f.get(), 0, 1);
        if (storage == NW_FACTOR_DEVICE) build_tables(f.get());
        else {
            f->build_gen = c->build_gen;
            f->fock_gen = c->fock_gen;
        }
        HIPCHK(hipStreamSynchronize(c->st));
        *out = f.release();
        c->factors.insert(*out);
        return NW_OK;
    });
}

extern "C" int nw_factor_refresh(nw_factor *f)
{
    if (!f) return NW_E_ARG;
    return guarded(f->c, [&]() -> int {
        if (!f->c->built) return fail(f->c, NW_E_STATE, "nw_build first");
        if (f->build_gen != f->c->build_gen)
            return fail(f->c, NW_E_STATE, "the context was rebuilt since nw_factor_create");
        if (f->storage == NW_FACTOR_DEVICE) build_tables(f);
        else f->fock_gen = f->c->fock_gen;
        HIPCHK(hipStreamSynchronize(f->c->st));
        return NW_OK;
    });
}

namespace {

// Stage 1 over items [i0, i1), stage 2 over rows [r0, r1) for the targets of
// parent q, on tables of type T2 (plan lo) and NV vectors.
template <class T2, int NV>
void run_range(nw_factor *f, int lo, int q, long long i0, long long i1, long long b0, long long b1, long long r0,
               long long r1, const double *dx, double lambda, double *dy, double *dys)
{
    nw_ctx *c = f->c;
    const ParentSpace &Q = f->par[q];
    const int maxlen = std::max(Q.ncomp[0], Q.ncomp[1]);
    const unsigned gx = (blocks_of(maxlen, 256) + 7u) & ~7u;
    for (long long b = i0; b < i1; b += 65535) {
        const unsigned ni = (unsigned)std::min<long long>(65535, i1 - b);
        hipLaunchKernelGGL((k_fac_mat<T2, NV>), dim3(gx, ni), dim3(256), 0, c->st, f->d_items[lo][NV - 1].p + b,
                           f->d_midx.p, f->d_mdet.p, dx);
    }
    if constexpr (NV == 1)
        for (long long b = b0; b < b1; b += 65535) {
            const unsigned ni = (unsigned)std::min<long long>(65535, b1 - b);
            hipLaunchKernelGGL((k_fac_block<T2, NV>), dim3(blocks_of(2ll * maxlen, BLK_NC), ni), dim3(256), 0, c->st,
                               f->d_blk[lo][NV - 1].p + b);
        }
    const int nr = (int)(r1 - r0);
    if (nr == 0) return;
    const long long waves = (Q.nt + 63) / 64;
    int nk = (int)std::min<long long>(64, std::max<long long>(1, (4096 + waves - 1) / waves));
    nk = std::min(nk, std::max(1, nr / 16));
    const int chunk = (nr + nk - 1) / nk;
    nk = (nr + chunk - 1) / chunk;
    hipLaunchKernelGGL((k_fac_rows<T2, NV>), dim3(blocks_of(Q.nt, 256), (unsigned)nk), dim3(256), 0, c->st,
                       Q.d_ta.p, Q.d_tb.p, Q.nt, f->d_rows[lo][NV - 1].p + r0, nr, chunk, dx, lambda, f->d_part2.p);
    hipLaunchKernelGGL(k_fac_scatter<NV>, dim3(blocks_of(Q.nt, 256)), dim3(256), 0, c->st, Q.d_tdet.p, Q.nt, nk,
                       f->d_part2.p, dy, dys);
}

template <class T2, int NV>
void run_duals(nw_factor *f, int lo, const double *dx)
{
    nw_ctx *c = f->c;
    const int nd = (int)f->duals.size();
    if (nd == 0) return;
    int maxblk = 0, maxlen = 0;
    for (const DualSel &D : f->duals) {
        const int ns = f->par[D.p].ncomp[D.sp], nt = f->par[D.q].ncomp[D.sp];
        maxblk = std::max(maxblk, ((ns + DUAL_ROWS - 1) / DUAL_ROWS) * ((nt + DUAL_COLS - 1) / DUAL_COLS));
        maxlen = std::max(maxlen, ns + nt);
    }
    const DualItem *items = f->d_dual[lo][NV - 1].p;
    hipLaunchKernelGGL((k_fac_dual<T2, NV>), dim3((unsigned)maxblk, (unsigned)nd), dim3(256), 0, c->st, items, dx);
    hipLaunchKernelGGL((k_fac_dual_reduce<T2, NV>), dim3(blocks_of(maxlen, 256), (unsigned)nd), dim3(256), 0, c->st,
                       items);
}

void run_duals_any(nw_factor *f, int lo, int nv, const double *dx)
{
    if (lo && nv == 2) run_duals<float2, 2>(f, 1, dx);
    else if (lo) run_duals<float2, 1>(f, 1, dx);
    else if (nv == 2) run_duals<double2, 2>(f, 0, dx);
    else run_duals<double2, 1>(f, 0, dx);
}

void run_range_any(nw_factor *f, int lo, int nv, int q, long long i0, long long i1, long long b0, long long b1,
                   long long r0, long long r1, const double *dx, double lambda, double *dy, double *dys)
{
    if (lo && nv == 2) run_range<float2, 2>(f, 1, q, i0, i1, b0, b1, r0, r1, dx, lambda, dy, dys);
    else if (lo) run_range<float2, 1>(f, 1, q, i0, i1, b0, b1, r0, r1, dx, lambda, dy, dys);
    else if (nv == 2) run_range<double2, 2>(f, 0, q, i0, i1, b0, b1, r0, r1, dx, lambda, dy, dys);
    else run_range<double2, 1>(f, 0, q, i0, i1, b0, b1, r0, r1, dx, lambda, dy, dys);
}

// Dense block weights from x, once per apply.
void run_densify(nw_factor *f, int nv, const double *dx)
{
    if (f->blks.empty() || nv != 1) return;
    const dim3 grid(blocks_of((long long)f->blk_gmax * f->blk_kmax, 256), (unsigned)f->blks.size());
    if (nv == 2) hipLaunchKernelGGL(k_fac_densify<2>, grid, dim3(256), 0, f->c->st, f->d_bdense.p, dx);
    else hipLaunchKernelGGL(k_fac_densify<1>, grid, dim3(256), 0, f->c->st, f->d_bdense.p, dx);
}

}  // namespace

namespace {

// y = (F + lambda S) x for nv interleaved vectors (x[i nv + v]); ys = S x if not null.
int apply_nv(nw_factor *f, int nv, double lambda, const double *x, double *y, double *ys)
{
    nw_ctx *c = f->c;
    return guarded(c, [&]() -> int {
        if (!x || !y) return fail(c, NW_E_ARG, "null arguments");
        check_current(f);
        const int lo = f->prec == NW_PREC_MIXED ? 1 : 0;
        ensure_plan(f, lo, nv);
        const long long n = (long long)f->ndet * nv;
        HIPCHK(hipEventRecord(c->ev[0], c->st));
        const double *dx = (const double *)to_device(c, c->outH, x, n * 8);
        const bool ydev = is_device_ptr(y), ysdev = ys && is_device_ptr(ys);
        double *dy = ydev ? y : c->outS.get<double>(2 * n);
        double *dys = ys ? (ysdev ? ys : c->outS.get<double>(2 * n) + n) : nullptr;
        hipLaunchKernelGGL(k_fac_zero, dim3(blocks_of(n, 256)), dim3(256), 0, c->st, dy, n);
        if (dys) hipLaunchKernelGGL(k_fac_zero, dim3(blocks_of(n, 256)), dim3(256), 0, c->st, dys, n);
        float regen = 0.f;
        run_densify(f, nv, dx);
        if (f->storage == NW_FACTOR_DEVICE) {
            run_duals_any(f, lo, nv, dx);
            // the pairs of one target are contiguous in the plan
            const std::vector<PairPlan> &plans = f->plans[lo][nv - 1];
            for (size_t k = 0; k < plans.size();) {
                size_t j = k;
                while (j < plans.size() && plans[j].q == plans[k].q) ++j;
                run_range_any(f, lo, nv, plans[k].q, plans[k].item0, plans[j - 1].item1, plans[k].blk0,
                              plans[j - 1].blk1, plans[k].row0, plans[j - 1].row1, dx, lambda, dy, dys);
                k = j;
            }
        } else {
            for (const PairPlan &pl : f->plans[lo][nv - 1]) {
                HIPCHK(hipEventRecord(c->ev[2], c->st));
                char *slot = (char *)f->d_tab.p;
                launch_table(f, pl.q, pl.p, 0, slot, nullptr, lo);
                launch_table(f, pl.q, pl.p, 1, slot + elem_bytes(lo) * table_elems(f, pl.q, pl.p, 0), nullptr, lo);
                HIPCHK(hipEventRecord(c->ev[3], c->st));
                run_range_any(f, lo, nv, pl.q, pl.item0, pl.item1, pl.blk0, pl.blk1, pl.row0, pl.row1, dx, lambda, dy,
                              dys);
                HIPCHK(hipStreamSynchronize(c->st));
                regen += elapsed(c->ev[2], c->ev[3]);
            }
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[1], c->st));
        if (!ydev) HIPCHK(hipMemcpyAsync(y, dy, n * 8, hipMemcpyDeviceToHost, c->st));
        if (ys && !ysdev) HIPCHK(hipMemcpyAsync(ys, dys, n * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        f->t_ms[0] = elapsed(c->ev[0], c->ev[1]);
        f->t_ms[1] = regen;
        return NW_OK;
    });
}

}  // namespace

extern "C" int nw_factor_apply(nw_factor *f, double lambda, const double *x, double *y, double *ys)
{
    if (!f) return NW_E_ARG;
    return apply_nv(f, 1, lambda, x, y, ys);
}

extern "C" int nw_factor_apply_pair(nw_factor *f, double lambda, const double *x2, double *y2, double *ys2)
{
    if (!f) return NW_E_ARG;
    return apply_nv(f, 2, lambda, x2, y2, ys2);
}

extern "C" int nw_factor_diagonals(nw_factor *f, double lambda, double *mdiag, double *sdiag)
{
    if (!f) return NW_E_ARG;
    nw_ctx *c = f->c;
    return guarded(c, [&]() -> int {
        if (!mdiag && !sdiag) return fail(c, NW_E_ARG, "null arguments");
        check_current(f);
        const long long n = (long long)f->ndet;
        const bool mdev = mdiag && is_device_ptr(mdiag), sdev = sdiag && is_device_ptr(sdiag);
        double *dm = mdiag ? (mdev ? mdiag : c->outS.get<double>(2 * n)) : nullptr;
        double *ds = sdiag ? (sdev ? sdiag : c->outS.get<double>(2 * n) + n) : nullptr;
        for (int p = 0; p < f->nref; ++p) {
            const ParentSpace &P = f->par[p];
            if (P.nt == 0) continue;
            double2 *ta = table_ptr(f, p, p, 0, f->d_tab.p);
            double2 *tb = table_ptr(f, p, p, 1, f->d_tab.p + table_elems(f, p, p, 0));
            hipLaunchKernelGGL(k_fac_diag, dim3(blocks_of(P.nt, 256)), dim3(256), 0, c->st, P.d_tdet.p, P.d_ta.p,
                               P.d_tb.p, P.nt, ta, P.ncomp[0], tb, P.ncomp[1], lambda, dm, ds);
        }
        HIPCHK(hipGetLastError());
        if (mdiag && !mdev) HIPCHK(hipMemcpyAsync(mdiag, dm, n * 8, hipMemcpyDeviceToHost, c->st));
        if (sdiag && !sdev) HIPCHK(hipMemcpyAsync(sdiag, ds, n * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        return NW_OK;
    });
}

extern "C" int nw_factor_info(const nw_factor *f, int parent, uint64_t *ncomp_a, uint64_t *ncomp_b,
                              uint64_t *table_bytes)
{
    if (!f) return NW_E_ARG;
    if (parent >= 0) {
        if (parent >= f->nref) return NW_E_ARG;
        if (ncomp_a) *ncomp_a = (uint64_t)f->par[parent].ncomp[0];
        if (ncomp_b) *ncomp_b = (uint64_t)f->par[parent].ncomp[1];
    }
    if (table_bytes) *table_bytes = (uint64_t)f->tab_elems * elem_bytes(f->prec == NW_PREC_MIXED);
    return NW_OK;
}

extern "C" int nw_factor_block_groups(const nw_factor *f, uint64_t *groups, uint64_t *blocks)
{
    if (!f) return NW_E_ARG;
    uint64_t g = 0;
    for (const BlkSrc &b : f->blks) g += (uint64_t)b.G;
    if (groups) *groups = g;
    if (blocks) *blocks = (uint64_t)f->blks.size();
    return NW_OK;
}

extern "C" int nw_factor_apply_bytes(nw_factor *f, uint64_t *bytes)
{
    if (!f || !bytes) return NW_E_ARG;
    return guarded(f->c, [&]() -> int {
        const int lo = f->prec == NW_PREC_MIXED ? 1 : 0;
        ensure_plan(f, lo, 1);
        *bytes = f->apply_bytes[lo];
        return NW_OK;
    });
}

extern "C" int nw_factor_set_precision(nw_factor *f, int precision)
{
    if (!f) return NW_E_ARG;
    nw_ctx *c = f->c;
    return guarded(c, [&]() -> int {
        if (precision != NW_PREC_FP64 && precision != NW_PREC_MIXED)
            return fail(c, NW_E_ARG, "factor precision must be NW_PREC_FP64 or NW_PREC_MIXED");
        if (precision == NW_PREC_MIXED) {
            check_current(f);
            if (f->storage == NW_FACTOR_DEVICE && !f->d_tab32.p) {
                f->d_tab32.alloc(std::max<size_t>(1, f->tab_elems));
                lower_tables(f);
                HIPCHK(hipGetLastError());
                HIPCHK(hipStreamSynchronize(c->st));
            }
            ensure_plan(f, 1, 1);
        }
        f->prec = precision;
        return NW_OK;
    });
}

extern "C" int nw_factor_ids(const nw_factor *f, uint32_t *aids, uint32_t *bids)
{
    if (!f) return NW_E_ARG;
    if (aids) std::copy(f->aid.begin(), f->aid.end(), aids);
    if (bids) std::copy(f->bid.begin(), f->bid.end(), bids);
    return NW_OK;
}

extern "C" int nw_factor_table(nw_factor *f, int target, int source, int spin, double *out)
{
    if (!f) return NW_E_ARG;
    nw_ctx *c = f->c;
    return guarded(c, [&]() -> int {
        if (target < 0 || source < 0 || target >= f->nref || source >= f->nref || spin < 0 || spin > 1 || !out)
            return fail(c, NW_E_ARG, "bad table index");
        check_current(f);
        const size_t ne = table_elems(f, target, source, spin);
        if (ne == 0) return NW_OK;
        if (f->prec == NW_PREC_MIXED) {
            // the FP32 values apply reads, widened to FP64
            const float2 *src32;
            if (f->storage == NW_FACTOR_DEVICE) src32 = f->d_tab32.p + f->tab_off[(target * f->nref + source) * 2 + spin];
            else {
                launch_table(f, target, source, spin, f->d_tab.p, nullptr, 1);
                src32 = (const float2 *)f->d_tab.p;
            }
            DVec<double2> wide;
            wide.alloc(ne);
            hipLaunchKernelGGL(k_fac_widen, dim3(blocks_of((long long)ne, 256)), dim3(256), 0, c->st, src32, wide.p,
                               (long long)ne);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(out, wide.p, ne * sizeof(double2), hipMemcpyDefault, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            return NW_OK;
        }
        const double2 *src = table_ptr(f, target, source, spin, f->d_tab.p);
        HIPCHK(hipMemcpyAsync(out, src, ne * sizeof(double2), hipMemcpyDefault, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        return NW_OK;
    });
}

extern "C" int nw_factor_last_ms(const nw_factor *f, double *ms2)
{
    if (!f || !ms2) return NW_E_ARG;
    ms2[0] = f->t_ms[0];
    ms2[1] = f->t_ms[1];
    return NW_OK;
}

extern "C" void nw_factor_destroy(nw_factor *f)
{
    if (!f) return;
    nw_ctx *c = f->c;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    c->factors.erase(f);
    delete f;
}

void nwi::destroy_factors(nw_ctx *c)
{
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->st);
    for (nw_factor *f : c->factors) delete f;
    c->factors.clear();
}
