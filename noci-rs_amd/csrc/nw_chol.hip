// Rank-4 slab blocks on demand from Cholesky-type integral factors.
//
// With (ab|cd) = sum_K L_K[a,b] L_K[c,d], every rank-4 tensor of a reference
// pair is a sum of products of MO-transformed factors:
//   II[p,q,r,s]  = sum_K A^a_K[p,q] A^b_K[r,s]                  (build.rs:877-917)
//   J[p,q,r,s]   = sum_K A_K[p,q] A'_K[r,s] - sum_K E_K[p,r] F_K[q,s]
//                                                               (build.rs:261-313)
// with A = cx^T L_K xc, E = cx^T L_K cx, F = xc^T L_K xc for the coefficient
// matrices of the J slot or II block.  The evaluators of pairs with
// m_alpha = m_beta = 0 read J (slot 0 only) as J[p,q,r,s] - J[p,s,r,q] (the
// same-spin two-electron term of preparehamiltonianoverlap.rs, and every
// fixed and run-time-rank kernel here), and that slot's F_K is symmetric, so
// the exchange products cancel in what they read: their blocks hold the
// direct products only.  Pairs with zero-overlap orbitals read single J
// entries, and get the exchange products as well (k_r4_exch).
//
// A determinant pair reads its tensors at (row, row) positions
// (p, r) = (rows[eta], rows[xi]) of its contraction labels
// (construct_determinant_indices, eval/prepare.rs:555-595), and every such
// block T[p][:][r][:] is one nm x nm product over K.  Only the blocks the
// determinants of an evaluation call read are built, each as a 16-row-tile
// MFMA product, into the reference slab layout; a later call builds what it
// adds.  Exports complete the tensors with the dense transforms.
#include <algorithm>
#include <cmath>
#include <map>
#include <cstdio>
#include <vector>

#include "nw_internal.h"

namespace nw {

struct R4Tensor {
    const void *P, *Q;  // [naux][nm][nm] factor tensors (double or float)
    void *out;          // rank-4 slab block [nm][nm][nm][nm] (double or float)
};
struct R4Item {
    int tensor, p, r, pad;
};

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

// One (row, row) block per workgroup iteration: out[p][q][r][s] =
// sum_K P[K][p][q] Q[K][r][s] for all q, s < nm.  NW waves; wave w owns the
// 16 output rows q in [16w, 16w + 16) and all NW column tiles s.  K is staged
// through LDS 32 at a time ([k][column] for both operands, so the A and B
// fragments are lane-contiguous rows), the next chunk prefetched into
// registers while the current one is multiplied.  Items are dealt to the
// eight XCDs in contiguous chunks (workgroup b runs chunk b mod 8), so
// neighbouring blocks of one tensor, which share factor rows, meet in one L2.
template <typename T, int NW>
__global__ __launch_bounds__(64 * NW) void k_r4_block(const R4Tensor *__restrict__ ts,
                                                      const R4Item *__restrict__ items, int nitems, int nm, int naux)
{
    // K chunk: 32 (16 for FP64 at NW = 8, to stay within 64 KB of LDS)
    constexpr int NP = 16 * NW, KC = (sizeof(T) == 8 && NW == 8) ? 16 : 32, NT = 64 * NW, LDW = NP + 2;
    constexpr int PER = 2 * KC * NP / NT;  // elements of the two operand chunks per thread
    static_assert(PER * NT == 2 * KC * NP, "operand chunks divide over the workgroup");
    __shared__ T Ps[KC][LDW];
    __shared__ T Qs[KC][LDW];
    using M = R4Mfma<T>;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const size_t nm2 = (size_t)nm * nm;
    const int gper = gridDim.x / 8, chunk = blockIdx.x % 8, lidx = blockIdx.x / 8;
    const int per = (nitems + 7) / 8;
    const int i0 = chunk * per, i1 = min(nitems, i0 + per);
    for (int it = i0 + lidx; it < i1; it += gper) {
        const R4Item w = items[it];
        const R4Tensor tt = ts[w.tensor];
        const T *__restrict__ P = (const T *)tt.P + (size_t)w.p * nm;
        const T *__restrict__ Q = (const T *)tt.Q + (size_t)w.r * nm;
        T buf[PER];
        auto gload = [&](int k0) {
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int e = tid + j * NT;
                const int e2 = j < PER / 2 ? e : e - KC * NP;
                const int k = e2 / NP, col = e2 % NP;
                const T *src = j < PER / 2 ? P : Q;
                buf[j] = (k0 + k < naux && col < nm) ? src[(size_t)(k0 + k) * nm2 + col] : T(0);
            }
        };
        typename M::acc acc[NW];
#pragma unroll
        for (int t = 0; t < NW; ++t) acc[t] = typename M::acc{0, 0, 0, 0};
        gload(0);
        for (int k0 = 0; k0 < naux; k0 += KC) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int e = tid + j * NT;
                const int e2 = j < PER / 2 ? e : e - KC * NP;
                const int k = e2 / NP, col = e2 % NP;
                if (j < PER / 2) Ps[k][col] = buf[j];
                else Qs[k][col] = buf[j];
            }
            __syncthreads();
            if (k0 + KC < naux) gload(k0 + KC);
#pragma unroll
            for (int kk = 0; kk < KC / 4; ++kk) {
                const T a = Ps[kk * 4 + lk][wave * 16 + lr];
#pragma unroll
                for (int t = 0; t < NW; ++t) acc[t] = M::mma(a, Qs[kk * 4 + lk][t * 16 + lr], acc[t]);
            }
        }
        T *out = (T *)tt.out;
#pragma unroll
        for (int t = 0; t < NW; ++t) {
            const int s = t * 16 + lr;
            if (s >= nm) continue;
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int q = wave * 16 + M::row(lk, r4);
                if (q < nm) out[(((size_t)w.p * nm + q) * nm + w.r) * nm + s] = acc[t][r4];
            }
        }
    }
}

// Exchange products of the J slots of pairs with zero-overlap orbitals
// (whose evaluators read single J entries): out[p][q][r][s] -=
// sum_K E[K][p][r] F[K][q][s] for every listed block (p, r), as one product
// per tensor with the blocks as rows: X[b][(q, s)] = sum_K Esel[b][K]
// F[K][(q, s)], 64 blocks x 64 columns per workgroup (4 waves of 16 blocks x
// 4 column tiles), K staged through LDS 32 at a time.  Runs after the direct
// products of the same blocks (stream order) and subtracts from them.
struct R4XTile {
    int tensor, i0, n, pad;  // items i0 .. i0 + n - 1 (n <= 64) of one tensor
};
template <typename T>
__global__ __launch_bounds__(256) void k_r4_exch(const R4Tensor *__restrict__ ts, const R4Item *__restrict__ items,
                                                 const R4XTile *__restrict__ tiles, int nm, int naux)
{
    constexpr int KC = 32, BT = 64, CT = 64, LDW = 64 + 2;
    __shared__ T Es[KC][LDW];
    __shared__ T Fs[KC][LDW];
    __shared__ int sp_[BT], sr_[BT];
    using M = R4Mfma<T>;
    const R4XTile tl = tiles[blockIdx.x];
    const R4Tensor tt = ts[tl.tensor];  // P = E, Q = F
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    const size_t nm2 = (size_t)nm * nm;
    const long long c0 = (long long)blockIdx.y * CT;
    if (tid < BT) {
        const R4Item w = tid < tl.n ? items[tl.i0 + tid] : R4Item{0, 0, 0, 0};
        sp_[tid] = w.p;
        sr_[tid] = w.r;
    }
    __syncthreads();
    const T *__restrict__ E = (const T *)tt.P;
    const T *__restrict__ F = (const T *)tt.Q;
    typename M::acc acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = typename M::acc{0, 0, 0, 0};
    for (int k0 = 0; k0 < naux; k0 += KC) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < KC * BT / 256; ++j) {
            const int e = tid + j * 256, k = e / BT, b = e % BT;
            Es[k][b] = (k0 + k < naux && b < tl.n) ? E[(size_t)(k0 + k) * nm2 + (size_t)sp_[b] * nm + sr_[b]] : T(0);
            const long long col = c0 + b;
            Fs[k][b] = (k0 + k < naux && col < (long long)nm2) ? F[(size_t)(k0 + k) * nm2 + col] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KC / 4; ++kk) {
            const T a = Es[kk * 4 + lk][wave * 16 + lr];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = M::mma(a, Fs[kk * 4 + lk][t * 16 + lr], acc[t]);
        }
    }
    T *out = (T *)tt.out;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const long long col = c0 + t * 16 + lr;
        if (col >= (long long)nm2) continue;
        const int q = (int)(col / nm), s = (int)(col % nm);
#pragma unroll
        for (int r4 = 0; r4 < 4; ++r4) {
            const int b = wave * 16 + M::row(lk, r4);
            if (b < tl.n) out[(((size_t)sp_[b] * nm + q) * nm + sr_[b]) * nm + s] -= acc[t][r4];
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

// Factor tensors [naux][nm][nm] of one pair, keyed fkey(kind, spin, a, b):
// kind FA = cx_a^T L_K xc_b, FE = cx_a^T L_K cx_b, FF = xc_a^T L_K xc_b.
enum { FA = 0, FE = 1, FF = 2 };
inline int fkey(int kind, int sp, int a, int b) { return ((kind * 2 + sp) * 2 + a) * 2 + b; }

struct CholPair {
    bool lazy = false;
    std::vector<uint8_t> cov[3];  // nm x nm: block (p, r) built
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

template <typename T, int NW>
int r4_grid()
{
    static int g = 0;
    if (!g) {
        int dev = 0, ncu = 0, occ = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_r4_block<T, NW>, 64 * NW, 0));
        g = std::max(8, (std::max(1, occ) * ncu) / 8 * 8);
    }
    return g;
}

template <typename T>
void launch_r4(hipStream_t st, int nw, const R4Tensor *ts, const R4Item *items, int nitems, int nm, int naux)
{
    if (nitems <= 0) return;
#define NW_R4_CASE(W)                                                                                     \
    case W: {                                                                                             \
        const int g = std::min(r4_grid<T, W>(), ((nitems + 7) / 8) * 8);                                  \
        hipLaunchKernelGGL((k_r4_block<T, W>), dim3(g), dim3(64 * W), 0, st, ts, items, nitems, nm, naux); \
        break;                                                                                            \
    }
    switch (nw) {
        NW_R4_CASE(1) NW_R4_CASE(2) NW_R4_CASE(3) NW_R4_CASE(4) NW_R4_CASE(5) NW_R4_CASE(6) NW_R4_CASE(7)
        NW_R4_CASE(8)
        default: throw NwError(NW_E_SHAPE, "on-demand rank-4 blocks need nmo <= 128");
    }
#undef NW_R4_CASE
}

// Row index sets of one parent's determinants per spin, as bra (rows
// part - nocc) and as ket (rows nvirt + hole), and the row pairs that occur
// within one determinant (both orders; the mixed-spin pairs as (alpha, beta)).
struct ParentRows {
    bool present = false;
    std::vector<uint8_t> bra[2], ket[2];            // nm
    std::vector<uint8_t> bra2[2], ket2[2];          // nm x nm, same-spin pairs of one determinant
    std::vector<uint8_t> braab, ketab;              // nm x nm, (alpha row, beta row) of one determinant
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
        }
        int br[2][4], kr[2][4], rk[2] = {d.ra, d.rb};
        for (int sp = 0; sp < 2; ++sp) {
            const uint8_t *ix = sp ? d.ib : d.ia;
            const int no = c->nocc[sp], nv = nm - no;
            for (int i = 0; i < rk[sp]; ++i) {
                br[sp][i] = std::min(nm - 1, std::max(0, (int)ix[4 + i] - no));
                kr[sp][i] = std::min(nm - 1, std::max(0, nv + (int)ix[i]));
                r.bra[sp][br[sp][i]] = 1;
                r.ket[sp][kr[sp][i]] = 1;
            }
            for (int i = 0; i < rk[sp]; ++i)
                for (int j = 0; j < rk[sp]; ++j)
                    if (i != j) {
                        r.bra2[sp][(size_t)br[sp][i] * nm + br[sp][j]] = 1;
                        r.ket2[sp][(size_t)kr[sp][i] * nm + kr[sp][j]] = 1;
                    }
        }
        for (int i = 0; i < rk[0]; ++i)
            for (int j = 0; j < rk[1]; ++j) {
                r.braab[(size_t)br[0][i] * nm + br[1][j]] = 1;
                r.ketab[(size_t)kr[0][i] * nm + kr[1][j]] = 1;
            }
    }
    return pr;
}

// Blocks (p, r) of pair (x, w) that its determinant pairs read, per class,
// not yet built: J_sigma at (rows[eta], rows[xi]) -- both bra rows of one
// determinant, a bra row then a ket row, or both ket rows of one
// determinant; II at (alpha row, beta row) -- both of one bra determinant,
// bra alpha x ket beta, ket alpha x bra beta, or both of one ket determinant.
// The evaluators of pairs with zero-overlap orbitals (both = true) read J
// rows in either order (the distribution sums and the slot swaps of
// preparehamiltonianoverlap.rs:20550-20966).
void missing_blocks(nw_ctx *c, const CholPair &cp, const ParentRows &X, const ParentRows &W, bool both,
                    std::vector<int> miss[3])
{
    const int nm = c->nmo;
    for (int sp = 0; sp < 2; ++sp) {
        const std::vector<uint8_t> &cov = cp.cov[sp];
        for (int p = 0; p < nm; ++p)
            for (int r = 0; r < nm; ++r) {
                const size_t k = (size_t)p * nm + r;
                if (cov[k]) continue;
                if (X.bra2[sp][k] || W.ket2[sp][k] || (X.bra[sp][p] && W.ket[sp][r]) ||
                    (both && X.bra[sp][r] && W.ket[sp][p]))
                    miss[sp].push_back((int)k);
            }
    }
    const std::vector<uint8_t> &cov = cp.cov[CL_II];
    for (int p = 0; p < nm; ++p)
        for (int r = 0; r < nm; ++r) {
            const size_t k = (size_t)p * nm + r;
            if (cov[k]) continue;
            if (X.braab[k] || W.ketab[k] || (X.bra[0][p] && W.ket[1][r]) || (W.ket[0][p] && X.bra[1][r]))
                miss[CL_II].push_back((int)k);
        }
}

// A_K = X^T L_K Y for every job (X, Y: n x nm row-major), [naux][nm][nm]:
// T1[q][K][a] = sum_b Y[b][q] L[K][a][b] (one product over the stacked
// factors per distinct Y), then per K A_K[p][q] = sum_a X[a][p] T1[q][K][a];
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
            for (int K = 0; K < naux; ++K)
                g2.push_back({trp[xi[j]], t1y + (size_t)K * n, jobs[j].out + (size_t)K * nm * nm});
        }
        launch_gemm_nt_mfma(c->st, nm, nm, n, n, naux * n, nm, up.put(g2), (int)g2.size());
    }
    HIPCHK(hipGetLastError());
}

// Build the listed missing blocks of lazy pairs (per pair, per class: block
// keys p * nm + r) and mark them built.  Per pair and class, every tensor of
// the class (the spin's J slots, the II blocks) gets the same blocks:
//   J slot (i, j, k, l) = jbranch: sum_K A^{kl}[p,q] A^{ij}[r,s] (the slab block
//     of the transform of (ab||cd) with cx_k, xc_l, cx_i, xc_j, build.rs:261-313),
//     less sum_K E^{ki}[p,r] F^{lj}[q,s] for pairs with zero-overlap orbitals;
//   II (a0, aj, b0, bj): sum_K A_alpha^{a0 aj}[p,q] A_beta^{b0 bj}[r,s].
void build_blocks(nw_ctx *c, const std::vector<int> &pkeys, std::vector<std::vector<int>> miss[3])
{
    CholState *s = c->chol;
    const int nm = c->nmo, naux = s->naux;
    const size_t nm2 = (size_t)nm * nm, nnm = (size_t)c->n * nm, ften = (size_t)naux * nm2;
    double *cx = (double *)c->keepCX.p, *xc = (double *)c->keepXC.p;
    struct TSpec {
        int pair_i, cl;
        int P, Q, E, F;  // factor keys (E, F < 0: no exchange)
        long long off;
    };
    std::vector<TSpec> specs;
    for (size_t i = 0; i < pkeys.size(); ++i) {
        const PairInfo &pi = c->pinfo[pkeys[i]];
        const bool exch = pi.ma + pi.mb > 0;
        for (int sp = 0; sp < 2; ++sp) {
            if (miss[sp][i].empty()) continue;
            const long long *o = pi.off + (sp ? O_BB : O_AA);
            for (int sl = 0; sl < 10; ++sl) {
                if (o[O_J + sl] < 0) continue;
                const int ii = jbranch(sl, 0), jj = jbranch(sl, 1), kk = jbranch(sl, 2), ll = jbranch(sl, 3);
                specs.push_back({(int)i, sp, fkey(FA, sp, kk, ll), fkey(FA, sp, ii, jj),
                                 exch ? fkey(FE, sp, kk, ii) : -1, exch ? fkey(FF, sp, ll, jj) : -1, o[O_J + sl]});
            }
        }
        if (!miss[CL_II][i].empty())
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
        CholPair &cp = s->pairs[pkeys[i]];
        if (cp.fac.count(key)) return;
        const PairInfo &pi = c->pinfo[pkeys[i]];
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
        need(t.pair_i, t.P);
        need(t.pair_i, t.Q);
        need(t.pair_i, t.E);
        need(t.pair_i, t.F);
    }
    if (!fj.empty()) factor_tensors(c, fj);
    auto f32of = [&](int i, int key) -> const void * {
        CholPair &cp = s->pairs[pkeys[i]];
        auto it = cp.fac32.find(key);
        if (it != cp.fac32.end()) return it->second;
        float *t = (float *)s->arena(ften * sizeof(float));
        launch_to_f32(c->st, cp.fac.at(key), t, (long long)ften);
        cp.fac32[key] = t;
        return t;
    };
    // direct products, then the exchange products of the same blocks, per precision
    for (int f32 = 0; f32 < 2; ++f32) {
        std::vector<R4Tensor> ts, tx;
        std::vector<R4Item> its, itx;
        std::vector<R4XTile> tiles;
        for (const TSpec &t : specs) {
            const PairInfo &pi = c->pinfo[pkeys[t.pair_i]];
            if ((pi.fp32 != 0) != (f32 == 1)) continue;
            CholPair &cp = s->pairs[pkeys[t.pair_i]];
            auto ptr = [&](int key) -> const void * {
                return f32 ? f32of(t.pair_i, key) : (const void *)cp.fac.at(key);
            };
            void *out = (void *)(c->d_slab + pi.base + t.off);
            const std::vector<int> &blocks = miss[t.cl][t.pair_i];
            const int ti = (int)ts.size();
            ts.push_back({ptr(t.P), ptr(t.Q), out});
            for (int k : blocks) its.push_back({ti, k / nm, k % nm, 0});
            if (t.E >= 0) {
                const int xi = (int)tx.size();
                tx.push_back({ptr(t.E), ptr(t.F), out});
                const int i0 = (int)itx.size();
                for (int k : blocks) itx.push_back({xi, k / nm, k % nm, 0});
                for (int b = 0; b < (int)blocks.size(); b += 64)
                    tiles.push_back({xi, i0 + b, std::min(64, (int)blocks.size() - b), 0});
            }
        }
        if (its.empty()) continue;
        const size_t bytes = (its.size() + itx.size()) * sizeof(R4Item) + (ts.size() + tx.size()) * sizeof(R4Tensor) +
                             tiles.size() * sizeof(R4XTile) + 8192;
        Uploader up(c, std::max<size_t>(1u << 20, bytes));
        const R4Tensor *dts = up.put(ts);
        const R4Item *dit = up.put(its);
        const int nw = (nm + 15) / 16;
        if (f32) launch_r4<float>(c->st, nw, dts, dit, (int)its.size(), nm, naux);
        else launch_r4<double>(c->st, nw, dts, dit, (int)its.size(), nm, naux);
        if (!tiles.empty()) {
            const R4Tensor *dtx = up.put(tx);
            const R4Item *ditx = up.put(itx);
            const R4XTile *dtl = up.put(tiles);
            const unsigned ncol = (unsigned)((nm2 + 63) / 64);
            for (size_t t0 = 0; t0 < tiles.size(); t0 += 65535) {
                const unsigned nt = (unsigned)std::min<size_t>(65535, tiles.size() - t0);
                if (f32)
                    hipLaunchKernelGGL(k_r4_exch<float>, dim3(nt, ncol), dim3(256), 0, c->st, dtx, ditx, dtl + t0, nm, naux);
                else
                    hipLaunchKernelGGL(k_r4_exch<double>, dim3(nt, ncol), dim3(256), 0, c->st, dtx, ditx, dtl + t0, nm,
                                       naux);
            }
        }
        HIPCHK(hipGetLastError());
    }
    for (size_t i = 0; i < pkeys.size(); ++i) {
        CholPair &cp = s->pairs[pkeys[i]];
        for (int cl = 0; cl < 3; ++cl)
            for (int k : miss[cl][i]) cp.cov[cl][k] = 1;
    }
}

// Lazy pairs touched by h and the blocks they still need.
bool plan_missing(nw_ctx *c, const std::vector<PackedDet> &h, std::vector<int> &pkeys,
                  std::vector<std::vector<int>> miss[3])
{
    CholState *s = c->chol;
    if (!s || s->pairs.empty()) return false;
    bool any_lazy = false;
    for (const CholPair &cp : s->pairs) any_lazy |= cp.lazy;
    if (!any_lazy) return false;
    const int nref = c->nref;
    std::vector<ParentRows> pr = parent_rows(c, h);
    bool any = false;
    for (int x = 0; x < nref; ++x) {
        if (!pr[x].present) continue;
        for (int w = 0; w < nref; ++w) {
            if (!pr[w].present) continue;
            const int key = x * nref + w;
            const CholPair &cp = s->pairs[key];
            if (!cp.lazy || !c->pinfo[key].built) continue;
            std::vector<int> m[3];
            const PairInfo &pi = c->pinfo[key];
            missing_blocks(c, cp, pr[x], pr[w], pi.ma + pi.mb > 0, m);
            if (m[0].empty() && m[1].empty() && m[2].empty()) continue;
            pkeys.push_back(key);
            for (int cl = 0; cl < 3; ++cl) miss[cl].push_back(std::move(m[cl]));
            any = true;
        }
    }
    return any;
}

}  // namespace

void chol_ensure(nw_ctx *c, const std::vector<PackedDet> &h)
{
    CholState *s = c->chol;
    if (!s) return;
    s->ms_blocks = 0.0;
    std::vector<int> pkeys;
    std::vector<std::vector<int>> miss[3];
    if (!plan_missing(c, h, pkeys, miss)) return;
    HIPCHK(hipEventRecord(c->ev[3], c->st));
    build_blocks(c, pkeys, miss);
    hipEvent_t e1;
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e1, c->st));
    HIPCHK(hipEventSynchronize(e1));
    s->ms_blocks = elapsed(c->ev[3], e1);
    (void)hipEventDestroy(e1);
}

bool chol_missing(nw_ctx *c, const std::vector<PackedDet> &h)
{
    std::vector<int> pkeys;
    std::vector<std::vector<int>> miss[3];
    return plan_missing(c, h, pkeys, miss);
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
