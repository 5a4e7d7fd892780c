// Device-side determinant-pair evaluators shared by the evaluation kernels
// (per-bin launches and the multi-system fused launch).  Arithmetic follows
// src/nonorthogonalwicks/eval/preparehamiltonianoverlap.rs (m = 0: the fixed
// kernels :380-19544 / generic m0 path :20304-20533; m > 0: :20550-20966),
// eval/helpers.rs and the one-body paths of eval/prepareonebodyoverlap.rs.
#pragma once
#include "nw_kernels.h"

namespace nw {

// Slab data through a global-address-space pointer: the evaluators read it
// with global (not flat) loads also where the slab base was read from
// memory (multi-system batches).
typedef const double __attribute__((address_space(1))) *slab_t;

// ---------------------------------------------------------------------------
// Small dense helpers, fully unrolled for compile-time sizes.
// ---------------------------------------------------------------------------
template <int N>
__device__ inline double det_n(const double *a)
{
    if constexpr (N == 0) return 1.0;
    else if constexpr (N == 1) return a[0];
    else if constexpr (N == 2) return a[0] * a[3] - a[1] * a[2];
    else if constexpr (N == 3)
        return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
               a[2] * (a[3] * a[7] - a[4] * a[6]);
    else {
        // Laplace along row 0
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < N; ++c) {
            double m[(N - 1) * (N - 1)];
#pragma unroll
            for (int i = 1; i < N; ++i) {
                int k = 0;
#pragma unroll
                for (int j = 0; j < N; ++j)
                    if (j != c) m[(i - 1) * (N - 1) + (k++)] = a[i * N + j];
            }
            double v = a[c] * det_n<N - 1>(m);
            acc += (c & 1) ? -v : v;
        }
        return acc;
    }
}

template <int N>
__device__ inline void minor_of(const double *a, int r, int c, double *m)
{
    int k = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i == r) continue;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (j == c) continue;
            m[k++] = a[i * N + j];
        }
    }
}

// cofactor matrix cof[r][c] = (-1)^{r+c} det(minor(r,c)); returns det
// (adjugate_transpose, src/maths/wick.rs:1923-2330).
template <int N>
__device__ inline double cofactors(const double *a, double *cof)
{
    if constexpr (N == 0) {
        return 1.0;
    } else if constexpr (N == 1) {
        cof[0] = 1.0;
        return a[0];
    } else {
#pragma unroll
        for (int r = 0; r < N; ++r)
#pragma unroll
            for (int c = 0; c < N; ++c) {
                double m[(N - 1) * (N - 1)];
                minor_of<N>(a, r, c, m);
                double v = det_n<N - 1>(m);
                cof[r * N + c] = ((r + c) & 1) ? -v : v;
            }
        double d = 0.0;
#pragma unroll
        for (int c = 0; c < N; ++c) d += a[c] * cof[c];
        return d;
    }
}

__device__ inline long long idx4(int n, int a, int b, int c, int d)
{
    return (((long long)a * n + b) * n + c) * n + d;
}

struct SpinScal {
    double f0h[2], f0f[2], v0[3];
    double pref;  // phase * tilde_s_prod
    int m, nocc;
    int r4f;      // rank-4 blocks (J slots, II) stored in FP32
};

// A pair whose four-index transforms ran in reduced precision (FP32
// accumulation) keeps its rank-4 blocks as FP32 values at the start of each
// FP64-sized block: the rounding to FP32 is the one the accumulator already
// had, so reading them back widened gives the values an FP64 store of the
// same accumulator would hold.
typedef const float __attribute__((address_space(1))) *slab32_t;
__device__ inline double r4(slab_t B, long long i, int f32) { return f32 ? (double)((slab32_t)B)[i] : B[i]; }

__device__ inline SpinScal load_spin(const PairInfo &P, int sp)
{
    const double *mt = P.meta + (sp ? M_BB : M_AA);
    SpinScal s;
    s.f0h[0] = mt[M_F0H];
    s.f0h[1] = mt[M_F0H + 1];
    s.f0f[0] = mt[M_F0F];
    s.f0f[1] = mt[M_F0F + 1];
    s.v0[0] = mt[M_V0];
    s.v0[1] = mt[M_V0 + 1];
    s.v0[2] = mt[M_V0 + 2];
    s.pref = mt[M_PH] * mt[M_TSP];
    s.m = (int)mt[M_M];
    s.nocc = (int)mt[M_NOCC];
    s.r4f = P.fp32;
    return s;
}

// Contraction-determinant labels (construct_determinant_indices, eval/prepare.rs:555-595):
// bra (x) pairs first -> rows = part - nocc, cols = hole; then ket (w) pairs
// -> rows = nvirt + hole, cols = part.
template <int L>
__device__ inline void labels(const uint8_t *bi, int brank, const uint8_t *ki, int nocc, int nvirt, int *rows,
                              int *cols)
{
#pragma unroll
    for (int i = 0; i < L; ++i) {
        if (i < brank) {
            rows[i] = (int)bi[4 + i] - nocc;
            cols[i] = bi[i];
        } else {
            int k = i - brank;
            rows[i] = nvirt + (int)ki[k];
            cols[i] = ki[4 + k];
        }
    }
}

// D_{ij} = X[r_i, c_j] (i >= j), Y[r_i, c_j] (i < j)   (build_d, src/maths/wick.rs:59)
template <int L>
__device__ inline void build_d(slab_t X, slab_t Y, int n, const int *rows, const int *cols, double *d)
{
#pragma unroll
    for (int i = 0; i < L; ++i)
#pragma unroll
        for (int j = 0; j < L; ++j) d[i * L + j] = (i >= j) ? X[rows[i] * n + cols[j]] : Y[rows[i] * n + cols[j]];
}

// ---------------------------------------------------------------------------
// m_alpha = m_beta = 0 element.
// ---------------------------------------------------------------------------
template <int L>
__device__ inline double same_m0(slab_t slab, const long long *o, const SpinScal &sc, int n, const int *rows,
                                 const int *cols, const double *d, const double *cof, double det)
{
    double acc = (sc.f0h[0] + 0.5 * sc.v0[0]) * det;
    if constexpr (L > 0) {
        slab_t hcol0 = slab + o[O_HCOL0];
        double repl = 0.0;
#pragma unroll
        for (int z = 0; z < L; ++z)
#pragma unroll
            for (int e = 0; e < L; ++e) repl += cof[e * L + z] * hcol0[cols[z] * n + rows[e]];
        acc -= repl;
    }
    if constexpr (L >= 2) {
        slab_t J = slab + o[O_J + 0];
#pragma unroll
        for (int eta = 0; eta < L; ++eta)
#pragma unroll
            for (int xi = eta + 1; xi < L; ++xi)
#pragma unroll
                for (int z = 0; z < L; ++z)
#pragma unroll
                    for (int y = z + 1; y < L; ++y) {
                        double mnr[(L - 2) * (L - 2) > 0 ? (L - 2) * (L - 2) : 1];
                        int ii = 0;
#pragma unroll
                        for (int r = 0; r < L; ++r) {
                            if (r == eta || r == xi) continue;
#pragma unroll
                            for (int c = 0; c < L; ++c) {
                                if (c == z || c == y) continue;
                                mnr[ii++] = d[r * L + c];
                            }
                        }
                        double second = det_n<L - 2>(mnr);
                        double direct = r4(J, idx4(n, rows[eta], cols[z], rows[xi], cols[y]), sc.r4f);
                        double exch = r4(J, idx4(n, rows[eta], cols[y], rows[xi], cols[z]), sc.r4f);
                        double term = second * (direct - exch);
                        acc += ((eta + xi + z + y) & 1) ? -term : term;
                    }
    }
    return acc;
}

template <int LA, int LB>
__device__ inline void hs_m0(const PairInfo &P, slab_t slab, const PackedDet &bra, const PackedDet &ket,
                             int nmo, double enuc, double &H, double &S)
{
    constexpr int LA1 = LA > 0 ? LA : 1, LB1 = LB > 0 ? LB : 1;
    const SpinScal sa = load_spin(P, 0), sb = load_spin(P, 1);
    const long long *oa = P.off + O_AA, *ob = P.off + O_BB;
    int ra[LA1], ca[LA1], rb[LB1], cb[LB1];
    double da[LA1 * LA1], db[LB1 * LB1], cofa[LA1 * LA1], cofb[LB1 * LB1];
    double deta = 1.0, detb = 1.0;
    if constexpr (LA > 0) {
        labels<LA>(bra.ia, bra.ra, ket.ia, sa.nocc, nmo - sa.nocc, ra, ca);
        build_d<LA>(slab + oa[O_X], slab + oa[O_Y], nmo, ra, ca, da);
        deta = cofactors<LA>(da, cofa);
    }
    if constexpr (LB > 0) {
        labels<LB>(bra.ib, bra.rb, ket.ib, sb.nocc, nmo - sb.nocc, rb, cb);
        build_d<LB>(slab + ob[O_X], slab + ob[O_Y], nmo, rb, cb, db);
        detb = cofactors<LB>(db, cofb);
    }
    double same_a = same_m0<LA>(slab, oa, sa, nmo, ra, ca, da, cofa, deta);
    double same_b = same_m0<LB>(slab, ob, sb, nmo, rb, cb, db, cofb, detb);
    double h2ab = P.meta[M_VAB0] * deta * detb;
    if constexpr (LA > 0 && LB > 0) {
        slab_t II = slab + P.off[O_II];
#pragma unroll
        for (int z = 0; z < LA; ++z)
#pragma unroll
            for (int e = 0; e < LA; ++e) {
                double inner = 0.0;
#pragma unroll
                for (int y = 0; y < LB; ++y)
#pragma unroll
                    for (int x = 0; x < LB; ++x) inner += cofb[x * LB + y] * r4(II, idx4(nmo, ra[e], ca[z], rb[x], cb[y]), P.fp32);
                h2ab += cofa[e * LA + z] * inner;
            }
    }
    double pref = (double)(bra.pha * bra.phb * ket.pha * ket.phb) * sa.pref * sb.pref;
    double s = pref * deta * detb;
    S = s;
    H = enuc * s + pref * (same_a * detb + same_b * deta + h2ab);
}

// ---------------------------------------------------------------------------
// m > 0 element: distribution sums (mode 0 = the reference's arithmetic,
// mode 1 = corrected; see nw_mode in include/nwick.h).
// ---------------------------------------------------------------------------
__device__ inline int jslot(int mi, int mj, int mk, int ml, int &swap)
{
    // (helpers.rs:165)
    const int code = ((mi * 2 + mj) * 2 + mk) * 2 + ml;
    const int slot[16] = {0, 1, 2, 3, 1, 4, 5, 6, 2, 5, 7, 8, 3, 6, 8, 9};
    swap = (code == 4 || code == 8 || code == 9 || code == 12 || code == 13 || code == 14) ? 1 : 0;
    return slot[code];
}

__device__ inline int bitk(unsigned b, int k) { return (b >> k) & 1; }

template <int L>
__device__ inline void same_gen(slab_t slab, const long long *o, const SpinScal &sc, int n, const int *rows,
                                const int *cols, const double *d0, const double *d1, double tol, int mode, double &s_out,
                                double &h1_out, double &h2_out)
{
    constexpr int L1 = L > 0 ? L : 1;
    double s = 0.0, h1 = 0.0, h2 = 0.0;
    const int nb = L + 2;
    for (unsigned bits = 0; bits < (1u << nb); ++bits) {
        if (__popc(bits) != sc.m) continue;
        const unsigned cbits = bits >> 2;
        const int m1 = bitk(bits, 0), m2 = bitk(bits, 1);
        double mix[L1 * L1], cof[L1 * L1];
#pragma unroll
        for (int r = 0; r < L; ++r)
#pragma unroll
            for (int c = 0; c < L; ++c) mix[r * L + c] = bitk(cbits, c) ? d1[r * L + c] : d0[r * L + c];
        double dd = 1.0;
        bool ok = true;
        if constexpr (L > 0) {
            dd = cofactors<L>(mix, cof);
            if (mode == 0 && !(fabs(dd) > tol)) ok = false;
        }
        if (!ok) {
            if (m1 == 0 && m2 == 0) s += dd;
            continue;
        }
        if (m1 == 0 && m2 == 0) s += dd;
        if (m2 == 0) {
            double contrib = dd * sc.f0h[m1];
#pragma unroll
            for (int k = 0; k < L; ++k) {
                const int mk = bitk(cbits, k);
                slab_t fsl = slab + o[O_FH + m1 * 2 + mk];
                double corr = 0.0;
#pragma unroll
                for (int r = 0; r < L; ++r) corr += (fsl[cols[k] * n + rows[r]] - mix[r * L + k]) * cof[r * L + k];
                contrib -= dd + corr;
            }
            h1 += contrib;
        }
        double v0 = sc.v0[m1 + m2];
        if (mode == 1 && m1 + m2 == 1) v0 *= 0.5;
        double contrib = v0 * dd;
#pragma unroll
        for (int k = 0; k < L; ++k) {
            const int mk = bitk(cbits, k);
            const int vi = mode == 0 ? (m1 * 2 + m2) * 2 + mk : (m1 * 2 + mk) * 2 + m2;
            slab_t vsl = slab + o[O_V + vi];
            double corr = 0.0;
#pragma unroll
            for (int r = 0; r < L; ++r) corr += (vsl[cols[k] * n + rows[r]] - mix[r * L + k]) * cof[r * L + k];
            contrib -= 2.0 * (dd + corr);
        }
        if constexpr (L >= 2) {
            constexpr int LM = L - 1;
#pragma unroll
            for (int i = 0; i < L; ++i)
#pragma unroll
                for (int j = 0; j < L; ++j) {
                    const double ph = ((i + j) & 1) ? -1.0 : 1.0;
                    const int ri = rows[i], cj = cols[j];
                    const int mj = bitk(cbits, j);
                    double mnr[LM * LM], mcof[LM * LM];
                    minor_of<L>(mix, i, j, mnr);
                    double dm = cofactors<LM>(mnr, mcof);
                    if (mode == 0 && !(fabs(dm) > tol)) continue;
#pragma unroll
                    for (int k2 = 0; k2 < LM; ++k2) {
                        const int kf = k2 < j ? k2 : k2 + 1;
                        const int mk = bitk(cbits, kf);
                        int swap;
                        const int slot = mode == 0 ? jslot(m1, m2, mk, mj, swap) : jslot(m1, mk, m2, mj, swap);
                        slab_t J = slab + o[O_J + slot];
                        double v = 0.0;
#pragma unroll
                        for (int r = 0; r < LM; ++r) {
                            const int rf = r < i ? r : r + 1;
                            const int rr = rows[rf], cc = cols[kf];
                            const double jv = r4(J, swap ? idx4(n, rr, cc, ri, cj) : idx4(n, ri, cj, rr, cc), sc.r4f);
                            v += jv * mcof[r * LM + k2];
                        }
                        contrib += ph * v;
                    }
                }
        }
        h2 += contrib;
    }
    s_out = s;
    h1_out = h1;
    h2_out = h2;
}

template <int LA, int LB>
__device__ inline void hs_gen(const PairInfo &P, slab_t slab, const PackedDet &bra, const PackedDet &ket,
                              int nmo, double enuc, double tol, int mode, double &H, double &S)
{
    constexpr int LA1 = LA > 0 ? LA : 1, LB1 = LB > 0 ? LB : 1;
    const SpinScal sa = load_spin(P, 0), sb = load_spin(P, 1);
    const long long *oa = P.off + O_AA, *ob = P.off + O_BB;
    int ra[LA1], ca[LA1], rb[LB1], cb[LB1];
    double da0[LA1 * LA1], da1[LA1 * LA1], db0[LB1 * LB1], db1[LB1 * LB1];
    if constexpr (LA > 0) {
        labels<LA>(bra.ia, bra.ra, ket.ia, sa.nocc, nmo - sa.nocc, ra, ca);
        build_d<LA>(slab + oa[O_X], slab + oa[O_Y], nmo, ra, ca, da0);
        if (sa.m > 0) build_d<LA>(slab + oa[O_X + 1], slab + oa[O_Y + 1], nmo, ra, ca, da1);
        else
#pragma unroll
            for (int k = 0; k < LA * LA; ++k) da1[k] = da0[k];
    }
    if constexpr (LB > 0) {
        labels<LB>(bra.ib, bra.rb, ket.ib, sb.nocc, nmo - sb.nocc, rb, cb);
        build_d<LB>(slab + ob[O_X], slab + ob[O_Y], nmo, rb, cb, db0);
        if (sb.m > 0) build_d<LB>(slab + ob[O_X + 1], slab + ob[O_Y + 1], nmo, rb, cb, db1);
        else
#pragma unroll
            for (int k = 0; k < LB * LB; ++k) db1[k] = db0[k];
    }
    double s_a, h1a, h2aa, s_b, h1b, h2bb;
    same_gen<LA>(slab, oa, sa, nmo, ra, ca, da0, da1, tol, mode, s_a, h1a, h2aa);
    same_gen<LB>(slab, ob, sb, nmo, rb, cb, db0, db1, tol, mode, s_b, h1b, h2bb);
    double h2ab = 0.0;
    for (unsigned ba = 0; ba < (1u << (LA + 1)); ++ba) {
        if (__popc(ba) != sa.m) continue;
        double mixa[LA1 * LA1], cofa[LA1 * LA1];
        double dA = 1.0;
        if constexpr (LA > 0) {
#pragma unroll
            for (int r = 0; r < LA; ++r)
#pragma unroll
                for (int c = 0; c < LA; ++c) mixa[r * LA + c] = bitk(ba >> 1, c) ? da1[r * LA + c] : da0[r * LA + c];
            dA = cofactors<LA>(mixa, cofa);
            if (mode == 0 && !(fabs(dA) > tol)) continue;
        }
        const int ma0 = bitk(ba, 0);
        for (unsigned bb = 0; bb < (1u << (LB + 1)); ++bb) {
            if (__popc(bb) != sb.m) continue;
            double mixb[LB1 * LB1], cofb[LB1 * LB1];
            double dB = 1.0;
            if constexpr (LB > 0) {
#pragma unroll
                for (int r = 0; r < LB; ++r)
#pragma unroll
                    for (int c = 0; c < LB; ++c) mixb[r * LB + c] = bitk(bb >> 1, c) ? db1[r * LB + c] : db0[r * LB + c];
                dB = cofactors<LB>(mixb, cofb);
                if (mode == 0 && !(fabs(dB) > tol)) continue;
            }
            const int mb0 = bitk(bb, 0);
            double contrib = P.meta[M_VAB0 + ma0 * 2 + mb0] * dA * dB;
#pragma unroll
            for (int k = 0; k < LA; ++k) {
                const int mak = bitk(ba, k + 1);
                slab_t vsl = slab + P.off[O_VAB + (ma0 * 2 + mb0) * 2 + mak];
                double v = 0.0;
#pragma unroll
                for (int r = 0; r < LA; ++r) v += vsl[ca[k] * nmo + ra[r]] * cofa[r * LA + k];
                contrib -= v * dB;
            }
#pragma unroll
            for (int k = 0; k < LB; ++k) {
                const int mbk = bitk(bb, k + 1);
                slab_t vsl = slab + P.off[O_VBA + (mb0 * 2 + ma0) * 2 + mbk];
                double v = 0.0;
#pragma unroll
                for (int r = 0; r < LB; ++r) v += vsl[cb[k] * nmo + rb[r]] * cofb[r * LB + k];
                contrib -= v * dA;
            }
#pragma unroll
            for (int i = 0; i < LA; ++i)
#pragma unroll
                for (int j = 0; j < LA; ++j) {
                    const double cij = cofa[i * LA + j];
                    const int ma1 = bitk(ba, j + 1);
#pragma unroll
                    for (int k = 0; k < LB; ++k) {
                        const int mbk = bitk(bb, k + 1);
                        slab_t II = slab + P.off[O_II + ((ma0 * 2 + ma1) * 2 + mb0) * 2 + mbk];
                        double v = 0.0;
#pragma unroll
                        for (int r = 0; r < LB; ++r) v += r4(II, idx4(nmo, ra[i], ca[j], rb[r], cb[k]), P.fp32) * cofb[r * LB + k];
                        contrib += cij * v;
                    }
                }
            h2ab += contrib;
        }
    }
    const double pref = (double)(bra.pha * bra.phb * ket.pha * ket.phb) * sa.pref * sb.pref;
    S = pref * s_a * s_b;
    H = pref * (h1a * s_b + h1b * s_a + 0.5 * h2aa * s_b + 0.5 * h2bb * s_a + h2ab);
    if (mode == 1) H += enuc * S;
}

// ---------------------------------------------------------------------------
// Overlap and generalised-Fock elements per spin: xw_overlap
// (eval/overlap.rs:42-66, sum over L-bit distributions of det D_mix) and
// xw_f (eval/onebody.rs:40-496, sum over (L+1)-bit distributions of
// det f0f[m_op] - sum_b det(D, column b -> ff[m_op][m_b])).  Returned without
// the phase * tilde_s_prod factor.
// ---------------------------------------------------------------------------
template <int L>
__device__ inline void fs_spin(slab_t slab, const long long *o, const SpinScal &sc, int n, const int *rows,
                               const int *cols, const double *d0, const double *d1, double tol, int mode, double &s_out,
                               double &f_out)
{
    constexpr int L1 = L > 0 ? L : 1;
    if (sc.m == 0) {
        double cof[L1 * L1];
        double dd = 1.0;
        if constexpr (L > 0) dd = cofactors<L>(d0, cof);
        double f = dd * sc.f0f[0];
        if constexpr (L > 0) {
            slab_t ff = slab + o[O_FF];
#pragma unroll
            for (int z = 0; z < L; ++z)
#pragma unroll
                for (int e = 0; e < L; ++e) f -= cof[e * L + z] * ff[cols[z] * n + rows[e]];
        }
        s_out = dd;
        f_out = f;
        return;
    }
    double s = 0.0, f = 0.0;
    if (sc.m <= L) {
        for (unsigned bits = 0; bits < (1u << L); ++bits) {
            if (__popc(bits) != sc.m) continue;
            double mix[L1 * L1];
#pragma unroll
            for (int r = 0; r < L; ++r)
#pragma unroll
                for (int c = 0; c < L; ++c) mix[r * L + c] = bitk(bits, c) ? d1[r * L + c] : d0[r * L + c];
            s += det_n<L>(mix);
        }
    }
    for (unsigned bits = 0; bits < (1u << (L + 1)); ++bits) {
        if (__popc(bits) != sc.m) continue;
        const int mi = bitk(bits, 0);
        double mix[L1 * L1], cof[L1 * L1];
        double dd = 1.0;
        if constexpr (L > 0) {
#pragma unroll
            for (int r = 0; r < L; ++r)
#pragma unroll
                for (int c = 0; c < L; ++c) mix[r * L + c] = bitk(bits >> 1, c) ? d1[r * L + c] : d0[r * L + c];
            dd = cofactors<L>(mix, cof);
            if (mode == 0 && !(fabs(dd) > tol)) continue;
        }
        double contrib = dd * sc.f0f[mi];
#pragma unroll
        for (int b = 0; b < L; ++b) {
            const int mj = bitk(bits, b + 1);
            slab_t fsl = slab + o[O_FF + mi * 2 + mj];
            double corr = 0.0;
#pragma unroll
            for (int r = 0; r < L; ++r) corr += (fsl[cols[b] * n + rows[r]] - mix[r * L + b]) * cof[r * L + b];
            contrib -= dd + corr;
        }
        f += contrib;
    }
    s_out = s;
    f_out = f;
}

// calculate_f_pair_wicks (src/noci/fock.rs:175-216) and the overlap
// calculate_s_pair_wicks (src/noci/overlap.rs:106-141) of one pair.
template <int LA, int LB>
__device__ inline void fs_elem(const PairInfo &P, slab_t slab, const PackedDet &bra, const PackedDet &ket,
                               int nmo, double tol, int mode, double &F, double &S)
{
    constexpr int LA1 = LA > 0 ? LA : 1, LB1 = LB > 0 ? LB : 1;
    const SpinScal sa = load_spin(P, 0), sb = load_spin(P, 1);
    const long long *oa = P.off + O_AA, *ob = P.off + O_BB;
    int ra[LA1], ca[LA1], rb[LB1], cb[LB1];
    double da0[LA1 * LA1], da1[LA1 * LA1], db0[LB1 * LB1], db1[LB1 * LB1];
    if constexpr (LA > 0) {
        labels<LA>(bra.ia, bra.ra, ket.ia, sa.nocc, nmo - sa.nocc, ra, ca);
        build_d<LA>(slab + oa[O_X], slab + oa[O_Y], nmo, ra, ca, da0);
        if (sa.m > 0) build_d<LA>(slab + oa[O_X + 1], slab + oa[O_Y + 1], nmo, ra, ca, da1);
    }
    if constexpr (LB > 0) {
        labels<LB>(bra.ib, bra.rb, ket.ib, sb.nocc, nmo - sb.nocc, rb, cb);
        build_d<LB>(slab + ob[O_X], slab + ob[O_Y], nmo, rb, cb, db0);
        if (sb.m > 0) build_d<LB>(slab + ob[O_X + 1], slab + ob[O_Y + 1], nmo, rb, cb, db1);
    }
    double ssa, fsa, ssb, fsb;
    fs_spin<LA>(slab, oa, sa, nmo, ra, ca, da0, da1, tol, mode, ssa, fsa);
    fs_spin<LB>(slab, ob, sb, nmo, rb, cb, db0, db1, tol, mode, ssb, fsb);
    const double pha = (double)(bra.pha * ket.pha), phb = (double)(bra.phb * ket.phb);
    const double s_a = pha * sa.pref * ssa, s_b = phb * sb.pref * ssb;
    double f = 0.0;
    if (s_b != 0.0) f += pha * sa.pref * fsa * s_b;
    if (s_a != 0.0) f += phb * sb.pref * fsb * s_a;
    F = f;
    S = s_a * s_b;
}


// (L_alpha, L_beta) with each <= 4: dispatch table
#define NW_BIN_CASES(X)                                                                               \
    X(0, 0) X(0, 1) X(0, 2) X(0, 3) X(0, 4) X(1, 0) X(1, 1) X(1, 2) X(1, 3) X(1, 4) X(2, 0) X(2, 1) \
        X(2, 2) X(2, 3) X(2, 4) X(3, 0) X(3, 1) X(3, 2) X(3, 3) X(3, 4) X(4, 0) X(4, 1) X(4, 2) X(4, 3) X(4, 4)

}  // namespace nw
