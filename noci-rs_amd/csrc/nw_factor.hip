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
// storage).  y = (F + lambda S) x is then, per ordered pair and with
// (1, 2) = (alpha, beta) or (beta, alpha) picked by the reference's cost model
// (select_one_body_contraction, onebody.rs:1583-1606):
//   stage 1  T[s2][t1] = sum_{e in P, c2_e = s2} x_e Tab1[c1_e][t1]           (sparse x dense)
//   stage 2  y_I += sum_s2 T.F[s2][t1_I] S2[s2][t2_I]
//                  + T.S[s2][t1_I] (F2 + lambda S2)[s2][t2_I]                   (sampled dense x dense)
// (apply_one_body_a_first / _b_first, onebody.rs:491-648); ys = S x comes
// from the same pass.  Stage 1 streams Tab1 once per pair with coalesced
// rows; stage 2 streams Tab2 and T with the target entries on the lanes.
#include <algorithm>
#include <cmath>
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

struct S1Piece { int e0, e1, row, part; };  // entry range -> output row (part >= 0: partial row)
struct S1Unit { int p0, p1; };              // pieces of one workgroup
struct S1Red { int row, p0, p1; };          // partial rows [p0, p1) summed into row

constexpr int S1_CHUNK = 64;   // entries per piece of a split group
constexpr int S1_UNIT = 64;    // entries per workgroup (packing target)

template <int L>
__device__ inline void spin_fs(const double *slab, const long long *o, const SpinScal &sc, int nmo,
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
// with the bra = target when target_left, else the source.
__global__ __launch_bounds__(256) void k_fac_table(const PairInfo *__restrict__ pairs, const double *__restrict__ slab0,
                                                   int pair, int spin, const SpinComp *__restrict__ tc, int nt,
                                                   const SpinComp *__restrict__ scomp, int ns, int target_left, int nmo,
                                                   double tol, int mode, double2 *__restrict__ tab)
{
    const int t = blockIdx.x * 64 + (threadIdx.x & 63);
    const int s = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (t >= nt || s >= ns) return;
    const PairInfo &P = pairs[pair];
    const double *slab = slab0 + P.base;
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
    tab[(size_t)s * nt + t] = make_double2(ph * vs, ph * vf);
}

// xs[e] = x[det[e]] for every source entry of every parent and order.
__global__ void k_fac_gather(const uint32_t *__restrict__ det, long long n, const double *__restrict__ x,
                             double *__restrict__ xs)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) xs[e] = x[det[e]];
}

// Stage 1: T[row][t] (or a partial row) = sum_e xs[e] Tab1[c1[e]][t].
__global__ __launch_bounds__(256) void k_fac_spmm(const S1Unit *__restrict__ units, const S1Piece *__restrict__ pieces,
                                                  const uint32_t *__restrict__ c1, const double *__restrict__ xs,
                                                  const double2 *__restrict__ tab1, int nt1, double2 *__restrict__ tt,
                                                  double2 *__restrict__ part)
{
    const S1Unit u = units[blockIdx.x];
    const int t = blockIdx.y * 256 + threadIdx.x;
    if (t >= nt1) return;
    for (int p = u.p0; p < u.p1; ++p) {
        const S1Piece pc = pieces[p];
        double s = 0.0, f = 0.0;
#pragma unroll 4
        for (int e = pc.e0; e < pc.e1; ++e) {
            const double xe = xs[e];
            const double2 v = tab1[(size_t)c1[e] * nt1 + t];
            s = fma(xe, v.x, s);
            f = fma(xe, v.y, f);
        }
        double2 *out = pc.part < 0 ? tt + (size_t)pc.row * nt1 : part + (size_t)pc.part * nt1;
        out[t] = make_double2(s, f);
    }
}

__global__ __launch_bounds__(256) void k_fac_spmm_reduce(const S1Red *__restrict__ reds, const double2 *__restrict__ part,
                                                         int nt1, double2 *__restrict__ tt)
{
    const S1Red r = reds[blockIdx.x];
    const int t = blockIdx.y * 256 + threadIdx.x;
    if (t >= nt1) return;
    double s = 0.0, f = 0.0;
    for (int p = r.p0; p < r.p1; ++p) {
        const double2 v = part[(size_t)p * nt1 + t];
        s += v.x;
        f += v.y;
    }
    tt[(size_t)r.row * nt1 + t] = make_double2(s, f);
}

// Stage 2 over the source components [k chunk, (k+1) chunk): one lane per target entry.
__global__ __launch_bounds__(256) void k_fac_sddmm(const uint32_t *__restrict__ t1i, const uint32_t *__restrict__ t2i,
                                                   int nt, const double2 *__restrict__ tt, int nt1,
                                                   const double2 *__restrict__ tab2, int nt2, int ns2, int chunk,
                                                   double lambda, double2 *__restrict__ part2)
{
    const int I = blockIdx.x * 256 + threadIdx.x;
    if (I >= nt) return;
    const int s0 = blockIdx.y * chunk, s1 = min(ns2, s0 + chunk);
    const double2 *pa = tt + t1i[I];
    const double2 *pb = tab2 + t2i[I];
    double m = 0.0, s = 0.0;
#pragma unroll 4
    for (int k = s0; k < s1; ++k) {
        const double2 u = pa[(size_t)k * nt1];  // {T.S, T.F}
        const double2 v = pb[(size_t)k * nt2];  // {S2, F2}
        m = fma(u.y, v.x, m);
        m = fma(u.x, fma(lambda, v.x, v.y), m);
        s = fma(u.x, v.x, s);
    }
    part2[(size_t)blockIdx.y * nt + I] = make_double2(m, s);
}

// y[det_I] += sum_k part2[k][I] (chunks in order); the same for ys.
__global__ void k_fac_scatter(const uint32_t *__restrict__ tdet, int nt, int nk, const double2 *__restrict__ part2,
                              double *__restrict__ y, double *__restrict__ ys)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= nt) return;
    double m = 0.0, s = 0.0;
    for (int k = 0; k < nk; ++k) {
        const double2 v = part2[(size_t)k * nt + I];
        m += v.x;
        s += v.y;
    }
    y[tdet[I]] += m;
    if (ys) ys[tdet[I]] += s;
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

// Source entries of one parent grouped by the second spin's component.
struct SrcOrder {
    std::vector<uint32_t> c1, det;  // host copies
    DVec<uint32_t> d_c1;
    DVec<S1Piece> d_pieces;
    DVec<S1Unit> d_units;
    DVec<S1Red> d_reds;
    long long xs_off = 0;  // offset of this order's entries in the gathered x
    int nunits = 0, nreds = 0, npart = 0;
};

struct ParentSpace {
    int first_det = -1, nt = 0;
    int ncomp[2] = {0, 0};
    long long comp_off[2] = {0, 0};
    std::vector<uint32_t> tdet, ta, tb;  // target entries in determinant order
    DVec<uint32_t> d_tdet, d_ta, d_tb;
    int maxrank[2] = {0, 0};
    SrcOrder ord[2];  // 0: grouped by beta (alpha contracted first), 1: grouped by alpha
};

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

}  // namespace

struct nw_factor {
    nw_ctx *c = nullptr;
    int nref = 0, mode = 0, storage = 0;
    uint64_t ndet = 0;
    unsigned long long build_gen = 0, fock_gen = 0;
    std::vector<ParentSpace> par;
    std::vector<uint32_t> aid, bid;  // per determinant
    DVec<SpinComp> d_comp[2];
    // tables: per (Q * nref + P) * 2 + spin, offset into d_tab (-1 = none)
    std::vector<long long> tab_off;
    std::vector<int> order;  // per (Q, P): 0 alpha first, 1 beta first
    DVec<double2> d_tab;
    size_t tab_elems = 0;
    // scratch
    DVec<uint32_t> d_srcdet;
    DVec<double> d_xs, d_y;
    DVec<double2> d_tt, d_part1, d_part2;
    size_t tt_elems = 0, part1_elems = 0, part2_elems = 0;
    long long nsrc = 0;
    double t_ms[2] = {0, 0};  // last apply: total, table regeneration
};

namespace {

int comps_of(const nw_factor *f, int p, int sp) { return f->par[p].ncomp[sp]; }

// Table element count of one ordered pair and spin ([source][target]).
size_t table_elems(const nw_factor *f, int q, int p, int sp)
{
    return (size_t)comps_of(f, p, sp) * (size_t)comps_of(f, q, sp);
}

void launch_table(nw_factor *f, int q, int p, int sp, double2 *out)
{
    nw_ctx *c = f->c;
    const ParentSpace &Q = f->par[q], &P = f->par[p];
    const bool target_left = Q.first_det <= P.first_det;
    const int pair = target_left ? q * c->nref + p : p * c->nref + q;
    const int nt = Q.ncomp[sp], ns = P.ncomp[sp];
    if (nt == 0 || ns == 0) return;
    dim3 grid(blocks_of(nt, 64), blocks_of(ns, 4));
    hipLaunchKernelGGL(k_fac_table, grid, dim3(256), 0, c->st, c->d_pinfo, c->d_slab, pair, sp,
                       f->d_comp[sp].p + Q.comp_off[sp], nt, f->d_comp[sp].p + P.comp_off[sp], ns,
                       target_left ? 1 : 0, c->nmo, c->tol, f->mode, out);
}

void build_tables(nw_factor *f)
{
    const int nref = f->nref;
    for (int q = 0; q < nref; ++q)
        for (int p = 0; p < nref; ++p)
            for (int sp = 0; sp < 2; ++sp) {
                const long long off = f->tab_off[(q * nref + p) * 2 + sp];
                if (off >= 0) launch_table(f, q, p, sp, f->d_tab.p + off);
            }
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

// Pointer of the (q, p, sp) table: resident, or regenerated into the scratch slot `slot`.
double2 *table_ptr(nw_factor *f, int q, int p, int sp, double2 *scratch)
{
    if (f->storage == NW_FACTOR_DEVICE) return f->d_tab.p + f->tab_off[(q * f->nref + p) * 2 + sp];
    launch_table(f, q, p, sp, scratch);
    return scratch;
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
        // ---- source orders: entries grouped by the second spin, pieces, units, reductions
        std::vector<uint32_t> srcdet;
        for (int p = 0; p < c->nref; ++p) {
            ParentSpace &P = f->par[p];
            if (P.nt == 0) continue;
            P.d_tdet.upload(P.tdet);
            P.d_ta.upload(P.ta);
            P.d_tb.upload(P.tb);
            for (int o = 0; o < 2; ++o) {
                SrcOrder &S = P.ord[o];
                const std::vector<uint32_t> &c1v = o == 0 ? P.ta : P.tb, &c2v = o == 0 ? P.tb : P.ta;
                std::vector<int> idx(P.nt);
                std::iota(idx.begin(), idx.end(), 0);
                std::stable_sort(idx.begin(), idx.end(), [&](int u, int v) {
                    return c2v[u] != c2v[v] ? c2v[u] < c2v[v] : c1v[u] < c1v[v];
                });
                for (int k : idx) {
                    S.c1.push_back(c1v[k]);
                    S.det.push_back(P.tdet[k]);
                }
                std::vector<S1Piece> pieces;
                std::vector<S1Red> reds;
                int npart = 0;
                for (int e = 0; e < P.nt;) {
                    int g = e;
                    while (g < P.nt && c2v[idx[g]] == c2v[idx[e]]) ++g;
                    const int row = (int)c2v[idx[e]];
                    if (g - e <= S1_CHUNK) {
                        pieces.push_back({e, g, row, -1});
                    } else {
                        const int p0 = npart;
                        for (int q0 = e; q0 < g; q0 += S1_CHUNK)
                            pieces.push_back({q0, std::min(g, q0 + S1_CHUNK), row, npart++});
                        reds.push_back({row, p0, npart});
                    }
                    e = g;
                }
                std::vector<S1Unit> units;
                for (int a = 0; a < (int)pieces.size();) {
                    int b = a, cnt = 0;
                    while (b < (int)pieces.size() && (b == a || cnt + pieces[b].e1 - pieces[b].e0 <= S1_UNIT)) {
                        cnt += pieces[b].e1 - pieces[b].e0;
                        ++b;
                    }
                    units.push_back({a, b});
                    a = b;
                }
                S.nunits = (int)units.size();
                S.nreds = (int)reds.size();
                S.npart = npart;
                S.xs_off = (long long)srcdet.size();
                srcdet.insert(srcdet.end(), S.det.begin(), S.det.end());
                S.d_c1.upload(S.c1);
                S.d_pieces.upload(pieces);
                S.d_units.upload(units);
                S.d_reds.upload(reds);
            }
        }
        // part1 rows are indexed by the target's first-spin components: size for the largest parent
        int maxcomp = 0, maxnt = 0;
        for (const ParentSpace &P : f->par) {
            maxcomp = std::max({maxcomp, P.ncomp[0], P.ncomp[1]});
            maxnt = std::max(maxnt, P.nt);
        }
        size_t maxpart = 0;
        for (const ParentSpace &P : f->par)
            for (int o = 0; o < 2; ++o) maxpart = std::max(maxpart, (size_t)P.ord[o].npart);
        f->part1_elems = std::max<size_t>(1, maxpart * (size_t)maxcomp);
        f->nsrc = (long long)srcdet.size();
        f->d_srcdet.upload(srcdet);
        f->d_xs.alloc(std::max<long long>(1, f->nsrc));
        // ---- contraction order per pair and table placement
        const int nref = c->nref;
        f->order.assign((size_t)nref * nref, 0);
        f->tab_off.assign((size_t)nref * nref * 2, -1);
        size_t off = 0, maxpair = 0;
        for (int q = 0; q < nref; ++q)
            for (int p = 0; p < nref; ++p) {
                const ParentSpace &Q = f->par[q], &P = f->par[p];
                if (Q.nt == 0 || P.nt == 0) continue;
                const double nta = Q.ncomp[0], ntb = Q.ncomp[1], nsa = P.ncomp[0], nsb = P.ncomp[1];
                const double ca = 2.0 * nta * P.nt + 2.0 * Q.nt * nsb;
                const double cb = 2.0 * ntb * P.nt + 2.0 * Q.nt * nsa;
                const int o = ca <= cb ? 0 : 1;
                f->order[q * nref + p] = o;
                // stage-1 output: second-spin source components x first-spin target components
                const size_t tt = o == 0 ? (size_t)nsb * (size_t)nta : (size_t)nsa * (size_t)ntb;
                f->tt_elems = std::max(f->tt_elems, tt);
                size_t pair_elems = 0;
                for (int sp = 0; sp < 2; ++sp) {
                    f->tab_off[(q * nref + p) * 2 + sp] = (long long)off;
                    off += table_elems(f.get(), q, p, sp);
                    pair_elems += table_elems(f.get(), q, p, sp);
                }
                maxpair = std::max(maxpair, pair_elems);
            }
        f->tab_elems = off;
        if (storage == NW_FACTOR_DEVICE) {
            f->d_tab.alloc(std::max<size_t>(1, off));
        } else {
            // one pair's two tables, regenerated per use
            f->d_tab.alloc(std::max<size_t>(1, maxpair));
        }
        f->d_tt.alloc(std::max<size_t>(1, f->tt_elems));
        f->d_part1.alloc(f->part1_elems);
        f->part2_elems = std::max<size_t>(1, (size_t)64 * (size_t)std::max(1, maxnt));
        f->d_part2.alloc(f->part2_elems);
        if (storage == NW_FACTOR_DEVICE) build_tables(f.get());
        else {
            f->build_gen = c->build_gen;
            f->fock_gen = c->fock_gen;
        }
        HIPCHK(hipStreamSynchronize(c->st));
        *out = f.release();
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

extern "C" int nw_factor_apply(nw_factor *f, double lambda, const double *x, double *y, double *ys)
{
    if (!f) return NW_E_ARG;
    nw_ctx *c = f->c;
    return guarded(c, [&]() -> int {
        if (!x || !y) return fail(c, NW_E_ARG, "null arguments");
        check_current(f);
        const long long n = (long long)f->ndet;
        HIPCHK(hipEventRecord(c->ev[0], c->st));
        const double *dx = (const double *)to_device(c, c->outH, x, n * 8);
        const bool ydev = is_device_ptr(y), ysdev = ys && is_device_ptr(ys);
        double *dy = ydev ? y : c->outS.get<double>(2 * n);
        double *dys = ys ? (ysdev ? ys : c->outS.get<double>(2 * n) + n) : nullptr;
        hipLaunchKernelGGL(k_fac_zero, dim3(blocks_of(n, 256)), dim3(256), 0, c->st, dy, n);
        if (dys) hipLaunchKernelGGL(k_fac_zero, dim3(blocks_of(n, 256)), dim3(256), 0, c->st, dys, n);
        if (f->nsrc > 0)
            hipLaunchKernelGGL(k_fac_gather, dim3(blocks_of(f->nsrc, 256)), dim3(256), 0, c->st, f->d_srcdet.p,
                               f->nsrc, dx, f->d_xs.p);
        const int nref = f->nref;
        float regen = 0.f;
        for (int q = 0; q < nref; ++q) {
            const ParentSpace &Q = f->par[q];
            if (Q.nt == 0) continue;
            for (int p = 0; p < nref; ++p) {
                const ParentSpace &P = f->par[p];
                if (P.nt == 0) continue;
                const int o = f->order[q * nref + p];
                const int s1 = o == 0 ? 0 : 1, s2 = 1 - s1;  // spin contracted in stage 1 / stage 2
                double2 *tab1, *tab2;
                if (f->storage == NW_FACTOR_DEVICE) {
                    tab1 = f->d_tab.p + f->tab_off[(q * nref + p) * 2 + s1];
                    tab2 = f->d_tab.p + f->tab_off[(q * nref + p) * 2 + s2];
                } else {
                    HIPCHK(hipEventRecord(c->ev[2], c->st));
                    tab1 = f->d_tab.p;
                    tab2 = f->d_tab.p + table_elems(f, q, p, s1);
                    launch_table(f, q, p, s1, tab1);
                    launch_table(f, q, p, s2, tab2);
                    HIPCHK(hipEventRecord(c->ev[3], c->st));
                }
                const SrcOrder &S = P.ord[o];
                const int nt1 = Q.ncomp[s1], nt2 = Q.ncomp[s2], ns2 = P.ncomp[s2];
                // stage 1
                hipLaunchKernelGGL(k_fac_spmm, dim3((unsigned)S.nunits, blocks_of(nt1, 256)), dim3(256), 0, c->st,
                                   S.d_units.p, S.d_pieces.p, S.d_c1.p, f->d_xs.p + S.xs_off, tab1, nt1, f->d_tt.p,
                                   f->d_part1.p);
                if (S.nreds > 0)
                    hipLaunchKernelGGL(k_fac_spmm_reduce, dim3((unsigned)S.nreds, blocks_of(nt1, 256)), dim3(256), 0,
                                       c->st, S.d_reds.p, f->d_part1.p, nt1, f->d_tt.p);
                // stage 2: split the source components so the grid covers the chip
                const long long waves = (Q.nt + 63) / 64;
                int nk = (int)std::min<long long>(64, std::max<long long>(1, (4096 + waves - 1) / waves));
                nk = std::min(nk, std::max(1, ns2 / 32));
                const int chunk = (ns2 + nk - 1) / nk;
                nk = (ns2 + chunk - 1) / chunk;
                const uint32_t *t1i = o == 0 ? Q.d_ta.p : Q.d_tb.p, *t2i = o == 0 ? Q.d_tb.p : Q.d_ta.p;
                hipLaunchKernelGGL(k_fac_sddmm, dim3(blocks_of(Q.nt, 256), (unsigned)nk), dim3(256), 0, c->st, t1i,
                                   t2i, Q.nt, f->d_tt.p, nt1, tab2, nt2, ns2, chunk, lambda, f->d_part2.p);
                hipLaunchKernelGGL(k_fac_scatter, dim3(blocks_of(Q.nt, 256)), dim3(256), 0, c->st, Q.d_tdet.p, Q.nt,
                                   nk, f->d_part2.p, dy, dys);
                if (f->storage == NW_FACTOR_TRANSIENT) {
                    HIPCHK(hipStreamSynchronize(c->st));
                    regen += elapsed(c->ev[2], c->ev[3]);
                }
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
    if (table_bytes) *table_bytes = (uint64_t)f->tab_elems * sizeof(double2);
    return NW_OK;
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
    {
        std::lock_guard<std::mutex> lk(c->mu);
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->st);
    }
    delete f;
}
