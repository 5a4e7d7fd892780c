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
// columns (q, s) of F).
struct R4Gemm {
    const void *A, *B;
    void *out;
    R4Axis row, col;
    int M, N, mode, pad;
};
struct R4GTile {
    int g, m0, n0, pad;
};

// Row-gathered A B^T products in 64 x 64 tiles: four waves of 16 rows x 4
// column tiles each, K staged through LDS 32 at a time ([row][k], rows
// gathered from the factor tensors, 256 contiguous bytes per row and chunk
// in FP64), the next chunk prefetched into registers during the products.
// Tiles are dealt to the eight XCDs in contiguous chunks, so the tiles of
// one product, which share rows, meet in one L2.
template <typename T>
__global__ __launch_bounds__(256) void k_r4_gemm(const R4Gemm *__restrict__ gs, const R4GTile *__restrict__ tiles,
                                                 int ntiles, int nm, int naux)
{
    constexpr int BM = 64, BN = 64, KC = 32, LDK = KC + 2, NT = 256, PER = (BM + BN) * KC / NT;
    __shared__ T As[BM][LDK];
    __shared__ T Bs[BN][LDK];
    __shared__ long long ra[BM], rb[BN], ro[BM], co[BN];
    using M = R4Mfma<T>;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const int gper = gridDim.x / 8, chunk = blockIdx.x % 8, lidx = blockIdx.x / 8;
    const int per = (ntiles + 7) / 8;
    const int t0 = chunk * per, t1 = min(ntiles, t0 + per);
    const long long nm2 = (long long)nm * nm;
    for (int it = t0 + lidx; it < t1; it += gper) {
        const R4GTile tl = tiles[it];
        const R4Gemm g = gs[tl.g];
        const int mloc = min(BM, g.M - tl.m0), nloc = min(BN, g.N - tl.n0);
        __syncthreads();  // the previous tile's indices are no longer read
        if (tid < BM) {
            int u = 0, v = 0;
            if (tid < mloc) axis_at(g.row, tl.m0 + tid, u, v);
            ra[tid] = ((long long)u * nm + v) * naux;
            ro[tid] = g.mode == 0 ? ((long long)u * nm + v) * nm2 : (long long)u * nm * nm2 + (long long)v * nm;
        } else if (tid < BM + BN) {
            const int j = tid - BM;
            int u = 0, v = 0;
            if (j < nloc) axis_at(g.col, tl.n0 + j, u, v);
            rb[j] = ((long long)u * nm + v) * naux;
            co[j] = g.mode == 0 ? (long long)u * nm + v : (long long)u * nm2 + v;
        }
        __syncthreads();
        const T *__restrict__ A = (const T *)g.A;
        const T *__restrict__ B = (const T *)g.B;
        T buf[PER];
        auto gload = [&](int k0) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int e = tid + j * NT;
                const bool second = e >= BM * KC;
                const int e2 = second ? e - BM * KC : e;
                const int r = e2 / KC, k = e2 % KC;
                const bool ok = (second ? r < nloc : r < mloc) && k0 + k < naux;
                buf[j] = ok ? (second ? B[rb[r] + k0 + k] : A[ra[r] + k0 + k]) : T(0);
            }
        };
        typename M::acc acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = typename M::acc{0, 0, 0, 0};
        gload(0);
        for (int k0 = 0; k0 < naux; k0 += KC) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int e = tid + j * NT;
                const bool second = e >= BM * KC;
                const int e2 = second ? e - BM * KC : e;
                const int r = e2 / KC, k = e2 % KC;
                if (second) Bs[r][k] = buf[j];
                else As[r][k] = buf[j];
            }
            __syncthreads();
            if (k0 + KC < naux) gload(k0 + KC);
#pragma unroll
            for (int kk = 0; kk < KC / 4; ++kk) {
                const T a = As[wave * 16 + lr][kk * 4 + lk];
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t] = M::mma(a, Bs[t * 16 + lr][kk * 4 + lk], acc[t]);
            }
        }
        T *out = (T *)g.out;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int j = t * 16 + lr;
            if (j >= nloc) continue;
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int i = wave * 16 + M::row(lk, r4);
                if (i >= mloc) continue;
                T *o = out + ro[i] + co[j];
                if (g.mode == 0) *o = acc[t][r4];
                else *o -= acc[t][r4];
            }
        }
    }
}

// out[b][j][i] = in[b][i][j] for rows x cols matrices.
__global__ void k_transpose_batched(const double *const *__restrict__ in, double *const *__restrict__ out, int rows,
                                    int cols)
{
    const double *a = in[blockIdx.y];
    double *o = out[blockIdx.y];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < rows * cols; e += gridDim.x * blockDim.x) {
        const int i = e / cols, j = e % cols;
        o[(size_t)j * rows + i] = a[e];
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
    int naux = 0;
    std::vector<CholPair> pairs;  // x * nref + w, for the context's current build
    std::vector<std::pair<char *, size_t>> chunks;  // factor-tensor arena, reset by every build
    size_t ci = 0, used = 0;
    DBuf t1, trans, ptrs, items, tensors;
    double ms_blocks = 0.0;
    // last chol_ensure: device ms of the factor products, rectangle products,
    // single blocks and exchange products; their algorithmic flops; single blocks
    double stats[10] = {};
    std::vector<hipEvent_t> evs;  // stage boundaries of the last build_blocks
    std::vector<int> ev_stage;
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

// Workgroups of a product launch: up to eight per CU, in multiples of eight
// (the XCD chunks); each runs its chunk's tiles in a grid-stride loop.
int r4_grid(int nitems)
{
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        ncu = std::max(1, ncu);
    }
    return std::max(8, std::min((nitems + 7) / 8 * 8, 8 * ncu));
}

template <typename T>
void launch_r4g(hipStream_t st, const R4Gemm *gs, const R4GTile *tiles, int ntiles, int nm, int naux)
{
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(k_r4_gemm<T>, dim3(r4_grid(ntiles)), dim3(256), 0, st, gs, tiles, ntiles, nm, naux);
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
        int br[2][4], kr[2][4], kc[2][4], rk[2] = {d.ra, d.rb};
        for (int sp = 0; sp < 2; ++sp) {
            const uint8_t *ix = sp ? d.ib : d.ia;
            const int no = c->nocc[sp], nv = nm - no;
            for (int i = 0; i < rk[sp]; ++i) {
                br[sp][i] = std::min(nm - 1, std::max(0, (int)ix[4 + i] - no));
                kr[sp][i] = std::min(nm - 1, std::max(0, nv + (int)ix[i]));
                kc[sp][i] = std::min(nm - 1, (int)ix[4 + i]);  // column label of a ket particle
                r.bra[sp][br[sp][i]] = 1;
                r.ket[sp][kr[sp][i]] = 1;
            }
            for (int i = 0; i < rk[sp]; ++i)
                for (int j = 0; j < rk[sp]; ++j)
                    if (i != j) {
                        r.bra2[sp][(size_t)br[sp][i] * nm + br[sp][j]] = 1;
                        r.ket2[sp][(size_t)kr[sp][i] * nm + kr[sp][j]] = 1;
                        r.kcol2[sp][(size_t)kc[sp][i] * nm + kc[sp][j]] = 1;
                    }
        }
        for (int i = 0; i < rk[0]; ++i)
            for (int j = 0; j < rk[1]; ++j) {
                r.braab[(size_t)br[0][i] * nm + br[1][j]] = 1;
                r.ketab[(size_t)kr[0][i] * nm + kr[1][j]] = 1;
                r.kcolab[(size_t)kc[0][i] * nm + kc[1][j]] = 1;
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
bool block_needed(int cl, const ParentRows &X, const ParentRows &W, bool both, int p, int r, int nm)
{
    const size_t k = (size_t)p * nm + r;
    if (cl < 2)
        return X.bra2[cl][k] || W.ket2[cl][k] || (X.bra[cl][p] && W.ket[cl][r]) || (both && X.bra[cl][r] && W.ket[cl][p]);
    return X.braab[k] || W.ketab[k] || (X.bra[0][p] && W.ket[1][r]) || (W.ket[0][p] && X.bra[1][r]);
}

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
bool plan_missing(nw_ctx *c, const std::vector<PackedDet> &h, std::vector<PairWork> &work)
{
    CholState *s = c->chol;
    if (!s || s->pairs.empty()) return false;
    bool any_lazy = false;
    for (const CholPair &cp : s->pairs) any_lazy |= cp.lazy;
    if (!any_lazy) return false;
    const int nref = c->nref, nm = c->nmo;
    const size_t nm2 = (size_t)nm * nm;
    std::vector<ParentRows> pr = parent_rows(c, h);
    for (int x = 0; x < nref; ++x) {
        if (!pr[x].present) continue;
        for (int w = 0; w < nref; ++w) {
            if (!pr[w].present) continue;
            const int key = x * nref + w;
            const CholPair &cp = s->pairs[key];
            const PairInfo &pi = c->pinfo[key];
            if (!cp.lazy || !pi.built) continue;
            const bool both = pi.ma + pi.mb > 0;
            PairWork pw;
            pw.key = key;
            bool any = false;
            for (int cl = 0; cl < 3; ++cl) {
                const std::vector<uint8_t> &need_kk = cl < 2 ? pr[w].kcol2[cl] : pr[w].kcolab;
                bool grow = false;
                for (size_t k = 0; k < nm2 && !grow; ++k) grow = need_kk[k] && !cp.kkcov[cl][k];
                ClassWork &cw = pw.cls[cl];
                for (int p = 0; p < nm; ++p)
                    for (int r = 0; r < nm; ++r) {
                        const size_t k = (size_t)p * nm + r;
                        const bool built = cp.cov[cl][k] != 0;
                        if (built ? grow : block_needed(cl, pr[x], pr[w], both, p, r, nm)) cw.blocks.push_back((int)k);
                    }
                if (cw.blocks.empty()) continue;
                cw.kk = cp.kkcov[cl];
                for (size_t k = 0; k < nm2; ++k) cw.kk[k] |= need_kk[k];
                // rectangles: bra rows of x against ket rows of w (both orders
                // for J of pairs with zero-overlap orbitals; II: alpha bra x beta
                // ket and alpha ket x beta bra), cut to the rows with work
                std::vector<uint8_t> todo(nm2, 0), done(nm2, 0);
                for (int k : cw.blocks) todo[k] = 1;
                std::vector<std::pair<const std::vector<uint8_t> *, const std::vector<uint8_t> *>> cand;
                if (cl < 2) {
                    cand.push_back({&pr[x].bra[cl], &pr[w].ket[cl]});
                    if (both) cand.push_back({&pr[w].ket[cl], &pr[x].bra[cl]});
                } else {
                    cand.push_back({&pr[x].bra[0], &pr[w].ket[1]});
                    cand.push_back({&pr[w].ket[0], &pr[x].bra[1]});
                }
                for (auto &cd : cand) {
                    std::vector<int> P, R;
                    for (int p = 0; p < nm; ++p) {
                        if (!(*cd.first)[p]) continue;
                        bool any_r = false;
                        for (int r = 0; r < nm && !any_r; ++r) any_r = (*cd.second)[r] && todo[(size_t)p * nm + r];
                        if (any_r) P.push_back(p);
                    }
                    for (int r = 0; r < nm; ++r) {
                        if (!(*cd.second)[r]) continue;
                        bool any_p = false;
                        for (int p : P) any_p = any_p || todo[(size_t)p * nm + r];
                        if (any_p) R.push_back(r);
                    }
                    if (P.empty() || R.empty()) continue;
                    for (int p : P)
                        for (int r : R) done[(size_t)p * nm + r] = 1;
                    cw.rects.push_back({std::move(P), std::move(R)});
                }
                for (int k : cw.blocks)
                    if (!done[k]) cw.single.push_back(k);
                cw.blocks.clear();
                for (size_t k = 0; k < nm2; ++k)
                    if (todo[k] || done[k]) cw.blocks.push_back((int)k);
                any = true;
            }
            if (any) work.push_back(std::move(pw));
        }
    }
    return !work.empty();
}

// A_K = X^T L_K Y for every job (X, Y: n x nm row-major), as [p][q][K]:
// T1[q][K][a] = sum_b Y[b][q] L[K][a][b] (one product over the stacked
// factors per distinct Y), then out[p][(q, K)] = sum_a X[a][p] T1[(q, K)][a];
// both as A B^T products on FP64 MFMA.
struct FacJob {
    const double *X, *Y;
    double *out;
};

void factor_tensors(nw_ctx *c, const std::vector<FacJob> &jobs)
{
    CholState *s = c->chol;
    const int n = c->n, nm = c->nmo, naux = s->naux;
    const size_t nnm = (size_t)n * nm, t1sz = (size_t)nm * naux * n;
    // transposes of every distinct X and Y
    std::vector<const double *> mats;
    auto idx_of = [&](const double *m) {
        for (size_t i = 0; i < mats.size(); ++i)
            if (mats[i] == m) return (int)i;
        mats.push_back(m);
        return (int)mats.size() - 1;
    };
    std::vector<int> xi(jobs.size()), yi(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) {
        xi[j] = idx_of(jobs[j].X);
        yi[j] = idx_of(jobs[j].Y);
    }
    double *tr = s->trans.get<double>(mats.size() * nnm);
    std::vector<double *> trp(mats.size());
    for (size_t i = 0; i < mats.size(); ++i) trp[i] = tr + i * nnm;
    Uploader up(c, 16u << 20);
    {
        const double *const *din = up.put(mats);
        double *const *dout = up.put(trp);
        hipLaunchKernelGGL(k_transpose_batched, dim3((unsigned)std::min<size_t>(64, (nnm + 255) / 256), (unsigned)mats.size()),
                           dim3(256), 0, c->st, din, dout, n, nm);
    }
    // distinct Y in groups whose T1 scratch stays within 1 GiB
    std::vector<int> ys;
    for (int y : yi)
        if (std::find(ys.begin(), ys.end(), y) == ys.end()) ys.push_back(y);
    const size_t gmax = std::max<size_t>(1, (size_t(1) << 30) / (t1sz * sizeof(double)));
    for (size_t g0 = 0; g0 < ys.size(); g0 += gmax) {
        const size_t gn = std::min(gmax, ys.size() - g0);
        double *t1 = s->t1.get<double>(gn * t1sz);
        std::vector<GemmDesc> g1;
        for (size_t k = 0; k < gn; ++k) g1.push_back({trp[ys[g0 + k]], s->d_L, t1 + k * t1sz});
        launch_gemm_nt_mfma(c->st, nm, naux * n, n, n, n, naux * n, up.put(g1), (int)gn);
        std::vector<GemmDesc> g2;
        for (size_t j = 0; j < jobs.size(); ++j) {
            const auto pos = std::find(ys.begin() + g0, ys.begin() + g0 + gn, yi[j]);
            if (pos == ys.begin() + g0 + gn) continue;
            const double *t1y = t1 + (size_t)(pos - (ys.begin() + g0)) * t1sz;
            g2.push_back({trp[xi[j]], t1y, jobs[j].out});
        }
        launch_gemm_nt_mfma(c->st, nm, nm * naux, n, n, n, nm * naux, up.put(g2), (int)g2.size());
    }
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
void build_blocks(nw_ctx *c, std::vector<PairWork> &work)
{
    CholState *s = c->chol;
    const int nm = c->nmo, naux = s->naux;
    const size_t nm2 = (size_t)nm * nm, nnm = (size_t)c->n * nm, ften = (size_t)naux * nm2;
    double *cx = (double *)c->keepCX.p, *xc = (double *)c->keepXC.p;
    auto occ_q = [&](int cl) { return cl < 2 ? c->nocc[cl] : c->nocc[0]; };
    auto occ_s = [&](int cl) { return cl < 2 ? c->nocc[cl] : c->nocc[1]; };
    struct TSpec {
        int wi, cl;
        int P, Q, E, F;  // factor keys (E, F < 0: no exchange)
        long long off;
    };
    std::vector<TSpec> specs;
    for (size_t i = 0; i < work.size(); ++i) {
        const PairInfo &pi = c->pinfo[work[i].key];
        const bool exch = pi.ma + pi.mb > 0;
        for (int sp = 0; sp < 2; ++sp) {
            if (work[i].cls[sp].blocks.empty()) continue;
            const long long *o = pi.off + (sp ? O_BB : O_AA);
            for (int sl = 0; sl < 10; ++sl) {
                if (o[O_J + sl] < 0) continue;
                const int ii = jbranch(sl, 0), jj = jbranch(sl, 1), kk = jbranch(sl, 2), ll = jbranch(sl, 3);
                specs.push_back({(int)i, sp, fkey(FA, sp, kk, ll), fkey(FA, sp, ii, jj),
                                 exch ? fkey(FE, sp, kk, ii) : -1, exch ? fkey(FF, sp, ll, jj) : -1, o[O_J + sl]});
            }
        }
        if (!work[i].cls[CL_II].blocks.empty())
            for (int a0 = 0; a0 < 2; ++a0)
                for (int aj = 0; aj < 2; ++aj)
                    for (int b0 = 0; b0 < 2; ++b0)
                        for (int bj = 0; bj < 2; ++bj) {
                            const long long oo = pi.off[O_II + ((a0 * 2 + aj) * 2 + b0) * 2 + bj];
                            if (oo < 0) continue;
                            specs.push_back({(int)i, CL_II, fkey(FA, 0, a0, aj), fkey(FA, 1, b0, bj), -1, -1, oo});
                        }
    }
    // factor tensors not yet made (FP64), then FP32 copies for reduced-precision pairs
    std::vector<FacJob> fj;
    auto need = [&](int i, int key) {
        if (key < 0) return;
        CholPair &cp = s->pairs[work[i].key];
        if (cp.fac.count(key)) return;
        const PairInfo &pi = c->pinfo[work[i].key];
        const int kind = key >> 3, sp = (key >> 2) & 1, a = (key >> 1) & 1, b = key & 1;
        const int pr = pi.bidx * 2 + sp;
        const double *ca = cx + (size_t)(pr * 2 + a) * nnm, *cb = cx + (size_t)(pr * 2 + b) * nnm;
        const double *xa = xc + (size_t)(pr * 2 + a) * nnm, *xb = xc + (size_t)(pr * 2 + b) * nnm;
        double *t = (double *)s->arena(ften * sizeof(double));
        cp.fac[key] = t;
        if (kind == FA) fj.push_back({ca, xb, t});
        else if (kind == FE) fj.push_back({ca, cb, t});
        else fj.push_back({xa, xb, t});
    };
    for (const TSpec &t : specs) {
        need(t.wi, t.P);
        need(t.wi, t.Q);
        need(t.wi, t.E);
        need(t.wi, t.F);
    }
    HIPCHK(hipEventRecord(s->mark(1), c->st));
    if (!fj.empty()) {
        factor_tensors(c, fj);
        std::vector<const double *> ys;
        for (const FacJob &j : fj)
            if (std::find(ys.begin(), ys.end(), j.Y) == ys.end()) ys.push_back(j.Y);
        s->stats[8] += 2.0 * nm * (double)naux * c->n * c->n * (double)(ys.size() + fj.size());
    }
    auto f32of = [&](int i, int key) -> const void * {
        CholPair &cp = s->pairs[work[i].key];
        auto it = cp.fac32.find(key);
        if (it != cp.fac32.end()) return it->second;
        float *t = (float *)s->arena(ften * sizeof(float));
        launch_to_f32(c->st, cp.fac.at(key), t, (long long)ften);
        cp.fac32[key] = t;
        return t;
    };
    // index lists shared by the products of a (pair, class), in one upload:
    // rectangle rows, per-q ket-particle partners, sparse lists of single
    // blocks, exchange rows and columns.  Offsets are patched after upload.
    std::vector<int> ints;
    auto put = [&](const std::vector<int> &v) {
        const size_t o = ints.size();
        ints.insert(ints.end(), v.begin(), v.end());
        return o;
    };
    struct ClassLists {
        std::vector<std::pair<size_t, size_t>> rect;  // offsets of P and R lists
        std::vector<std::pair<int, size_t>> kq;       // (q, offset of S(q)) with |S(q)| in kqn
        std::vector<int> kqn;
        std::vector<std::pair<size_t, size_t>> sgl;   // single blocks grouped by p: offsets of {p} and R(p)
        std::vector<int> sgln;
        size_t xp = 0, xr = 0, xq = 0, xs = 0;        // exchange rows (p, r) and columns (q, s)
        int nxrow = 0, nxcol = 0;
    };
    std::vector<ClassLists> cls(work.size() * 3);
    for (size_t i = 0; i < work.size(); ++i)
        for (int cl = 0; cl < 3; ++cl) {
            const ClassWork &cw = work[i].cls[cl];
            if (cw.blocks.empty()) continue;
            ClassLists &L = cls[i * 3 + cl];
            const int oq = occ_q(cl), os = occ_s(cl);
            for (auto &rc : cw.rects) L.rect.push_back({put(rc.first), put(rc.second)});
            for (int q = oq; q < nm; ++q) {
                std::vector<int> S;
                for (int t = os; t < nm; ++t)
                    if (cw.kk[(size_t)q * nm + t]) S.push_back(t);
                if (!S.empty()) {
                    L.kq.push_back({q, put(S)});
                    L.kqn.push_back((int)S.size());
                }
            }
            for (size_t k = 0; k < cw.single.size();) {  // cw.single is ordered by p
                const int p0 = cw.single[k] / nm;
                std::vector<int> R;
                for (; k < cw.single.size() && cw.single[k] / nm == p0; ++k) R.push_back(cw.single[k] % nm);
                const size_t po = put(std::vector<int>{p0});
                L.sgl.push_back({po, put(R)});
                L.sgln.push_back((int)R.size());
            }
            if (c->pinfo[work[i].key].ma + c->pinfo[work[i].key].mb > 0 && cl < 2) {
                std::vector<int> rp, rr, cq, cs;
                for (int k : cw.blocks) {
                    rp.push_back(k / nm);
                    rr.push_back(k % nm);
                }
                for (int q = 0; q < nm; ++q)
                    for (int t = 0; t < nm; ++t)
                        if (q < oq || t < os || cw.kk[(size_t)q * nm + t]) {
                            cq.push_back(q);
                            cs.push_back(t);
                        }
                L.xp = put(rp);
                L.xr = put(rr);
                L.xq = put(cq);
                L.xs = put(cs);
                L.nxrow = (int)rp.size();
                L.nxcol = (int)cq.size();
            }
        }
    // List pointers hold (offset + 1) into `ints` until the upload; 0 = none.
    const int *dints = nullptr;
    auto at = [](size_t off) { return (const int *)(uintptr_t)(off + 1); };
    std::vector<R4Gemm> gd[2], gx[2];
    std::vector<R4GTile> td[2], tx[2];
    std::vector<R4Gemm> gs1[2];  // single blocks, whole
    std::vector<R4GTile> ts1[2];
    double nsingle[2] = {0.0, 0.0};
    for (int f32 = 0; f32 < 2; ++f32) {
        auto tiles_of = [&](std::vector<R4GTile> &tl, int g, int M, int N) {
            for (int m0 = 0; m0 < M; m0 += 64)
                for (int n0 = 0; n0 < N; n0 += 64) tl.push_back({g, m0, n0, 0});
        };
        for (const TSpec &t : specs) {
            const PairWork &pw = work[t.wi];
            const PairInfo &pi = c->pinfo[pw.key];
            if ((pi.fp32 != 0) != (f32 == 1)) continue;
            CholPair &cp = s->pairs[pw.key];
            auto ptr = [&](int key) -> const void * { return f32 ? f32of(t.wi, key) : (const void *)cp.fac.at(key); };
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
                    tiles_of(td[f32], (int)gd[f32].size(), g.M, g.N);
                    gd[f32].push_back(g);
                }
                if (os > 0 && oq < nm) {  // rows (p, q >= oq) x columns (r, s < os)
                    R4Gemm g{PA, QB, out, {Pl, nullptr, np, nm - oq, oq, 0}, {Rl, nullptr, nr, os, 0, 0},
                             np * (nm - oq), nr * os, 0, 0};
                    tiles_of(td[f32], (int)gd[f32].size(), g.M, g.N);
                    gd[f32].push_back(g);
                }
                for (size_t k = 0; k < L.kq.size(); ++k) {  // rows (p, q) x columns (r, s in S(q))
                    const int q = L.kq[k].first, ns = L.kqn[k];
                    R4Gemm g{PA, QB, out, {Pl, nullptr, np, 1, q, 0}, {Rl, at(L.kq[k].second), nr, ns, 0, 0},
                             np, nr * ns, 0, 0};
                    tiles_of(td[f32], (int)gd[f32].size(), g.M, g.N);
                    gd[f32].push_back(g);
                }
            }
            for (size_t k = 0; k < L.sgl.size(); ++k) {  // whole blocks: rows (p, every q) x columns (r, every s)
                R4Gemm g{PA, QB, out, {at(L.sgl[k].first), nullptr, 1, nm, 0, 0},
                         {at(L.sgl[k].second), nullptr, L.sgln[k], nm, 0, 0}, nm, L.sgln[k] * nm, 0, 0};
                tiles_of(ts1[f32], (int)gs1[f32].size(), g.M, g.N);
                gs1[f32].push_back(g);
                nsingle[f32] += L.sgln[k];
            }
            if (t.E >= 0 && L.nxrow > 0 && L.nxcol > 0) {
                R4Gemm g{ptr(t.E), ptr(t.F), out, {at(L.xp), at(L.xr), L.nxrow, 1, 0, 1},
                         {at(L.xq), at(L.xs), L.nxcol, 1, 0, 1}, L.nxrow, L.nxcol, 1, 0};
                tiles_of(tx[f32], (int)gx[f32].size(), g.M, g.N);
                gx[f32].push_back(g);
            }
        }
    }
    // one upload of the lists and every descriptor, then the launches
    size_t bytes = ints.size() * sizeof(int) + 8192;
    for (int f = 0; f < 2; ++f)
        bytes += (gd[f].size() + gx[f].size() + gs1[f].size()) * sizeof(R4Gemm) +
                 (td[f].size() + tx[f].size() + ts1[f].size()) * sizeof(R4GTile) + 3072;
    Uploader up(c, bytes);
    dints = up.put(ints);
    auto fix = [&](const int *p) { return p ? dints + ((uintptr_t)p - 1) : nullptr; };
    for (int f = 0; f < 2; ++f) {
        for (auto *v : {&gd[f], &gx[f], &gs1[f]})
            for (R4Gemm &g : *v) {
                g.row.u = fix(g.row.u);
                g.row.v = fix(g.row.v);
                g.col.u = fix(g.col.u);
                g.col.v = fix(g.col.v);
            }
    }
    for (int f = 0; f < 2; ++f) {
        for (const R4Gemm &g : gd[f]) s->stats[5] += 2.0 * g.M * (double)g.N * naux;
        for (const R4Gemm &g : gx[f]) s->stats[7] += 2.0 * g.M * (double)g.N * naux;
        for (const R4Gemm &g : gs1[f]) s->stats[6] += 2.0 * g.M * (double)g.N * naux;
        s->stats[9] += nsingle[f];
    }
    for (int f32 = 0; f32 < 2; ++f32) {
        if (!td[f32].empty()) {
            HIPCHK(hipEventRecord(s->mark(2), c->st));
            const R4Gemm *dg = up.put(gd[f32]);
            const R4GTile *dt = up.put(td[f32]);
            if (f32) launch_r4g<float>(c->st, dg, dt, (int)td[f32].size(), nm, naux);
            else launch_r4g<double>(c->st, dg, dt, (int)td[f32].size(), nm, naux);
        }
        if (!ts1[f32].empty()) {
            HIPCHK(hipEventRecord(s->mark(3), c->st));
            const R4Gemm *dg = up.put(gs1[f32]);
            const R4GTile *dt = up.put(ts1[f32]);
            if (f32) launch_r4g<float>(c->st, dg, dt, (int)ts1[f32].size(), nm, naux);
            else launch_r4g<double>(c->st, dg, dt, (int)ts1[f32].size(), nm, naux);
        }
        if (!tx[f32].empty()) {  // after the direct products of the same blocks (stream order)
            HIPCHK(hipEventRecord(s->mark(4), c->st));
            const R4Gemm *dg = up.put(gx[f32]);
            const R4GTile *dt = up.put(tx[f32]);
            if (f32) launch_r4g<float>(c->st, dg, dt, (int)tx[f32].size(), nm, naux);
            else launch_r4g<double>(c->st, dg, dt, (int)tx[f32].size(), nm, naux);
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

void chol_ensure(nw_ctx *c, const std::vector<PackedDet> &h)
{
    CholState *s = c->chol;
    if (!s) return;
    s->ms_blocks = 0.0;
    for (double &v : s->stats) v = 0.0;
    s->ev_stage.clear();
    std::vector<PairWork> work;
    if (!plan_missing(c, h, work)) return;
    HIPCHK(hipEventRecord(s->mark(0), c->st));
    build_blocks(c, work);
    HIPCHK(hipEventSynchronize(s->evs[s->ev_stage.size() - 1]));
    // stage k runs from its mark to the next mark
    for (size_t k = 0; k + 1 < s->ev_stage.size(); ++k)
        s->stats[s->ev_stage[k]] += elapsed(s->evs[k], s->evs[k + 1]);
    s->ms_blocks = elapsed(s->evs[0], s->evs[s->ev_stage.size() - 1]);
    s->stats[0] = s->ms_blocks;
}

bool chol_missing(nw_ctx *c, const std::vector<PackedDet> &h)
{
    std::vector<PairWork> work;
    return plan_missing(c, h, work);
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

double chol_last_ms(nw_ctx *c) { return c->chol ? c->chol->ms_blocks : 0.0; }

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
        s->d_L = nullptr;
        s->naux = 0;
        s->pairs.clear();
        c->built = false;  // pairs built before take effect only after the next nw_build
        if (naux == 0) return NW_OK;
        const int n = c->n;
        const size_t cnt = (size_t)naux * n * n;
        double *dL = dalloc<double>(cnt);
        HIPCHK(hipMemcpyAsync(dL, L, cnt * sizeof(double), hipMemcpyDefault, c->st));
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
        s->naux = naux;
        return NW_OK;
    });
}

int nw_rank4_stats(nw_ctx *c, double *out10)
{
    return guarded(c, [&]() -> int {
        if (!out10) return fail(c, NW_E_ARG, "null output");
        for (int k = 0; k < 10; ++k) out10[k] = c->chol ? c->chol->stats[k] : 0.0;
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
