// Determinant-pair evaluators for contraction ranks above four per spin (up
// to eight: bra and ket excitations of rank up to four each), with the rank
// known only at run time.  The arithmetic is the one of the fixed-rank
// templates in nw_eval_dev.h -- the reference's generic paths, m = 0:
// xw_hamiltonian_overlap_m0_gen_prepared (preparehamiltonianoverlap.rs:
// 20304-20533), m > 0: :20550-20966 with eval/helpers.rs, and the one-body
// xw_f / xw_overlap -- with every determinant, cofactor and minor taken by
// LU factorisation with partial pivoting (the reference's adjugate for
// n >= 5 is det * inverse^T from an LU factorisation, adjt_gen, src/maths/
// wick.rs:2320-2365; explicit minors give the same cofactors and also exist
// for singular matrices).  Per-lane arrays of up to 8 x 8 doubles live in
// private memory; these ranks arise only from triple and quadruple
// excitations.
#pragma once
#include "nw_eval_dev.h"

namespace nw {

constexpr int GEN_LMAX = 8;

// det of the n x n matrix a (row stride n), destroying a
__device__ inline double det_lu_rt(double *a, int n)
{
    double det = 1.0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double mx = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            const double v = fabs(a[i * n + k]);
            if (v > mx) {
                mx = v;
                p = i;
            }
        }
        if (mx == 0.0) return 0.0;
        if (p != k) {
            for (int j = k; j < n; ++j) {
                const double t = a[k * n + j];
                a[k * n + j] = a[p * n + j];
                a[p * n + j] = t;
            }
            det = -det;
        }
        const double piv = a[k * n + k];
        det *= piv;
        for (int i = k + 1; i < n; ++i) {
            const double f = a[i * n + k] / piv;
            for (int j = k + 1; j < n; ++j) a[i * n + j] -= f * a[k * n + j];
        }
    }
    return det;
}

// det of a (n x n) with rows r0, r1 and columns c0, c1 removed (-1: none)
__device__ inline double det_sub_rt(const double *a, int n, int r0, int r1, int c0, int c1)
{
    double m[GEN_LMAX * GEN_LMAX];
    int k = 0, nr = 0;
    for (int i = 0; i < n; ++i) {
        if (i == r0 || i == r1) continue;
        ++nr;
        for (int j = 0; j < n; ++j) {
            if (j == c0 || j == c1) continue;
            m[k++] = a[i * n + j];
        }
    }
    return nr == 0 ? 1.0 : det_lu_rt(m, nr);
}

// cofactor matrix cof[r][c] = (-1)^{r+c} det(a[r|c]); returns det(a)
__device__ inline double cofactors_rt(const double *a, int n, double *cof)
{
    if (n == 0) return 1.0;
    if (n == 1) {
        cof[0] = 1.0;
        return a[0];
    }
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            const double v = det_sub_rt(a, n, r, -1, c, -1);
            cof[r * n + c] = ((r + c) & 1) ? -v : v;
        }
    double d = 0.0;
    for (int c = 0; c < n; ++c) d += a[c] * cof[c];
    return d;
}

__device__ inline void labels_rt(const uint8_t *bi, int brank, const uint8_t *ki, int krank, int nocc, int nvirt,
                                 int *rows, int *cols)
{
    for (int i = 0; i < brank; ++i) {
        rows[i] = (int)bi[4 + i] - nocc;
        cols[i] = bi[i];
    }
    for (int k = 0; k < krank; ++k) {
        rows[brank + k] = nvirt + (int)ki[k];
        cols[brank + k] = ki[4 + k];
    }
}

__device__ inline void build_d_rt(slab_t X, slab_t Y, int n, int L, const int *rows, const int *cols, double *d)
{
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) d[i * L + j] = (i >= j) ? X[rows[i] * n + cols[j]] : Y[rows[i] * n + cols[j]];
}

__device__ inline void mix_rt(const double *d0, const double *d1, int L, unsigned cbits, double *mix)
{
    for (int r = 0; r < L; ++r)
        for (int c = 0; c < L; ++c) mix[r * L + c] = bitk(cbits, c) ? d1[r * L + c] : d0[r * L + c];
}

// One spin's labels and endpoint determinants.
struct SpinRT {
    int L;
    int rows[GEN_LMAX], cols[GEN_LMAX];
    double d0[GEN_LMAX * GEN_LMAX], d1[GEN_LMAX * GEN_LMAX];
};

__device__ inline void prep_rt(SpinRT &s, slab_t slab, const long long *o, const SpinScal &sc, int n,
                               const uint8_t *bi, int brank, const uint8_t *ki, int krank)
{
    s.L = brank + krank;
    labels_rt(bi, brank, ki, krank, sc.nocc, n - sc.nocc, s.rows, s.cols);
    build_d_rt(slab + o[O_X], slab + o[O_Y], n, s.L, s.rows, s.cols, s.d0);
    if (sc.m > 0) build_d_rt(slab + o[O_X + 1], slab + o[O_Y + 1], n, s.L, s.rows, s.cols, s.d1);
    else
        for (int k = 0; k < s.L * s.L; ++k) s.d1[k] = s.d0[k];
}

// m_alpha = m_beta = 0: one spin's same-spin sum (same_m0 of nw_eval_dev.h)
__device__ inline double same_m0_rt(slab_t slab, const long long *o, const SpinScal &sc, int n, const SpinRT &s,
                                    const double *cof, double det)
{
    const int L = s.L;
    double acc = (sc.f0h[0] + 0.5 * sc.v0[0]) * det;
    if (L > 0) {
        slab_t hcol0 = slab + o[O_HCOL0];
        for (int z = 0; z < L; ++z)
            for (int e = 0; e < L; ++e) acc -= cof[e * L + z] * hcol0[s.cols[z] * n + s.rows[e]];
    }
    if (L >= 2) {
        slab_t J = slab + o[O_J + 0];
        for (int eta = 0; eta < L; ++eta)
            for (int xi = eta + 1; xi < L; ++xi)
                for (int z = 0; z < L; ++z)
                    for (int y = z + 1; y < L; ++y) {
                        const double second = det_sub_rt(s.d0, L, eta, xi, z, y);
                        const double direct = r4(J, idx4(n, s.rows[eta], s.cols[z], s.rows[xi], s.cols[y]), sc.r4f);
                        const double exch = r4(J, idx4(n, s.rows[eta], s.cols[y], s.rows[xi], s.cols[z]), sc.r4f);
                        const double term = second * (direct - exch);
                        acc += ((eta + xi + z + y) & 1) ? -term : term;
                    }
    }
    return acc;
}

__device__ inline void hs_m0_rt(const PairInfo &P, slab_t slab, const PackedDet &bra, const PackedDet &ket, int nmo,
                                double enuc, double &H, double &S)
{
    const SpinScal sa = load_spin(P, 0), sb = load_spin(P, 1);
    const long long *oa = P.off + O_AA, *ob = P.off + O_BB;
    SpinRT A, B;
    prep_rt(A, slab, oa, sa, nmo, bra.ia, bra.ra, ket.ia, ket.ra);
    prep_rt(B, slab, ob, sb, nmo, bra.ib, bra.rb, ket.ib, ket.rb);
    double cofa[GEN_LMAX * GEN_LMAX], cofb[GEN_LMAX * GEN_LMAX];
    const double deta = cofactors_rt(A.d0, A.L, cofa), detb = cofactors_rt(B.d0, B.L, cofb);
    const double same_a = same_m0_rt(slab, oa, sa, nmo, A, cofa, deta);
    const double same_b = same_m0_rt(slab, ob, sb, nmo, B, cofb, detb);
    double h2ab = P.meta[M_VAB0] * deta * detb;
    if (A.L > 0 && B.L > 0) {
        slab_t II = slab + P.off[O_II];
        for (int z = 0; z < A.L; ++z)
            for (int e = 0; e < A.L; ++e) {
                double inner = 0.0;
                for (int y = 0; y < B.L; ++y)
                    for (int x = 0; x < B.L; ++x)
                        inner += cofb[x * B.L + y] * r4(II, idx4(nmo, A.rows[e], A.cols[z], B.rows[x], B.cols[y]), P.fp32);
                h2ab += cofa[e * A.L + z] * inner;
            }
    }
    const double pref = (double)(bra.pha * bra.phb * ket.pha * ket.phb) * sa.pref * sb.pref;
    const double s = pref * deta * detb;
    S = s;
    H = enuc * s + pref * (same_a * detb + same_b * deta + h2ab);
}

// m > 0: one spin's same-spin distribution sum (same_gen of nw_eval_dev.h)
__device__ inline void same_gen_rt(slab_t slab, const long long *o, const SpinScal &sc, int n, const SpinRT &s,
                                   double tol, int mode, double &s_out, double &h1_out, double &h2_out)
{
    const int L = s.L;
    double sacc = 0.0, h1 = 0.0, h2 = 0.0;
    const int nb = L + 2;
    double mix[GEN_LMAX * GEN_LMAX], cof[GEN_LMAX * GEN_LMAX];
    for (unsigned bits = 0; bits < (1u << nb); ++bits) {
        if (__popc(bits) != sc.m) continue;
        const unsigned cbits = bits >> 2;
        const int m1 = bitk(bits, 0), m2 = bitk(bits, 1);
        mix_rt(s.d0, s.d1, L, cbits, mix);
        const double dd = cofactors_rt(mix, L, cof);
        if (L > 0 && mode == 0 && !(fabs(dd) > tol)) {
            if (m1 == 0 && m2 == 0) sacc += dd;
            continue;
        }
        if (m1 == 0 && m2 == 0) sacc += dd;
        if (m2 == 0) {
            double contrib = dd * sc.f0h[m1];
            for (int k = 0; k < L; ++k) {
                const int mk = bitk(cbits, k);
                slab_t fsl = slab + o[O_FH + m1 * 2 + mk];
                double corr = 0.0;
                for (int r = 0; r < L; ++r) corr += (fsl[s.cols[k] * n + s.rows[r]] - mix[r * L + k]) * cof[r * L + k];
                contrib -= dd + corr;
            }
            h1 += contrib;
        }
        double v0 = sc.v0[m1 + m2];
        if (mode == 1 && m1 + m2 == 1) v0 *= 0.5;
        double contrib = v0 * dd;
        for (int k = 0; k < L; ++k) {
            const int mk = bitk(cbits, k);
            const int vi = mode == 0 ? (m1 * 2 + m2) * 2 + mk : (m1 * 2 + mk) * 2 + m2;
            slab_t vsl = slab + o[O_V + vi];
            double corr = 0.0;
            for (int r = 0; r < L; ++r) corr += (vsl[s.cols[k] * n + s.rows[r]] - mix[r * L + k]) * cof[r * L + k];
            contrib -= 2.0 * (dd + corr);
        }
        if (L >= 2) {
            const int LM = L - 1;
            double mnr[GEN_LMAX * GEN_LMAX], mcof[GEN_LMAX * GEN_LMAX];
            for (int i = 0; i < L; ++i)
                for (int j = 0; j < L; ++j) {
                    const double ph = ((i + j) & 1) ? -1.0 : 1.0;
                    const int ri = s.rows[i], cj = s.cols[j];
                    const int mj = bitk(cbits, j);
                    int k = 0;
                    for (int r = 0; r < L; ++r) {
                        if (r == i) continue;
                        for (int c = 0; c < L; ++c)
                            if (c != j) mnr[k++] = mix[r * L + c];
                    }
                    const double dm = cofactors_rt(mnr, LM, mcof);
                    if (mode == 0 && !(fabs(dm) > tol)) continue;
                    for (int k2 = 0; k2 < LM; ++k2) {
                        const int kf = k2 < j ? k2 : k2 + 1;
                        const int mk = bitk(cbits, kf);
                        int swap;
                        const int slot = mode == 0 ? jslot(m1, m2, mk, mj, swap) : jslot(m1, mk, m2, mj, swap);
                        slab_t J = slab + o[O_J + slot];
                        double v = 0.0;
                        for (int r = 0; r < LM; ++r) {
                            const int rf = r < i ? r : r + 1;
                            const int rr = s.rows[rf], cc = s.cols[kf];
                            v += r4(J, swap ? idx4(n, rr, cc, ri, cj) : idx4(n, ri, cj, rr, cc), sc.r4f) * mcof[r * LM + k2];
                        }
                        contrib += ph * v;
                    }
                }
        }
        h2 += contrib;
    }
    s_out = sacc;
    h1_out = h1;
    h2_out = h2;
}

__device__ inline void hs_gen_rt(const PairInfo &P, slab_t slab, const PackedDet &bra, const PackedDet &ket, int nmo,
                                 double enuc, double tol, int mode, double &H, double &S)
{
    const SpinScal sa = load_spin(P, 0), sb = load_spin(P, 1);
    const long long *oa = P.off + O_AA, *ob = P.off + O_BB;
    SpinRT A, B;
    prep_rt(A, slab, oa, sa, nmo, bra.ia, bra.ra, ket.ia, ket.ra);
    prep_rt(B, slab, ob, sb, nmo, bra.ib, bra.rb, ket.ib, ket.rb);
    double s_a, h1a, h2aa, s_b, h1b, h2bb;
    same_gen_rt(slab, oa, sa, nmo, A, tol, mode, s_a, h1a, h2aa);
    same_gen_rt(slab, ob, sb, nmo, B, tol, mode, s_b, h1b, h2bb);
    const int LA = A.L, LB = B.L;
    double h2ab = 0.0;
    double mixa[GEN_LMAX * GEN_LMAX], cofa[GEN_LMAX * GEN_LMAX], mixb[GEN_LMAX * GEN_LMAX], cofb[GEN_LMAX * GEN_LMAX];
    for (unsigned ba = 0; ba < (1u << (LA + 1)); ++ba) {
        if (__popc(ba) != sa.m) continue;
        mix_rt(A.d0, A.d1, LA, ba >> 1, mixa);
        const double dA = cofactors_rt(mixa, LA, cofa);
        if (LA > 0 && mode == 0 && !(fabs(dA) > tol)) continue;
        const int ma0 = bitk(ba, 0);
        for (unsigned bb = 0; bb < (1u << (LB + 1)); ++bb) {
            if (__popc(bb) != sb.m) continue;
            mix_rt(B.d0, B.d1, LB, bb >> 1, mixb);
            const double dB = cofactors_rt(mixb, LB, cofb);
            if (LB > 0 && mode == 0 && !(fabs(dB) > tol)) continue;
            const int mb0 = bitk(bb, 0);
            double contrib = P.meta[M_VAB0 + ma0 * 2 + mb0] * dA * dB;
            for (int k = 0; k < LA; ++k) {
                slab_t vsl = slab + P.off[O_VAB + (ma0 * 2 + mb0) * 2 + bitk(ba, k + 1)];
                double v = 0.0;
                for (int r = 0; r < LA; ++r) v += vsl[A.cols[k] * nmo + A.rows[r]] * cofa[r * LA + k];
                contrib -= v * dB;
            }
            for (int k = 0; k < LB; ++k) {
                slab_t vsl = slab + P.off[O_VBA + (mb0 * 2 + ma0) * 2 + bitk(bb, k + 1)];
                double v = 0.0;
                for (int r = 0; r < LB; ++r) v += vsl[B.cols[k] * nmo + B.rows[r]] * cofb[r * LB + k];
                contrib -= v * dA;
            }
            for (int i = 0; i < LA; ++i)
                for (int j = 0; j < LA; ++j) {
                    const double cij = cofa[i * LA + j];
                    const int ma1 = bitk(ba, j + 1);
                    for (int k = 0; k < LB; ++k) {
                        slab_t II = slab + P.off[O_II + ((ma0 * 2 + ma1) * 2 + mb0) * 2 + bitk(bb, k + 1)];
                        double v = 0.0;
                        for (int r = 0; r < LB; ++r)
                            v += r4(II, idx4(nmo, A.rows[i], A.cols[j], B.rows[r], B.cols[k]), P.fp32) * cofb[r * LB + k];
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

// Overlap and generalised-Fock sums of one spin (fs_spin of nw_eval_dev.h)
__device__ inline void fs_spin_rt(slab_t slab, const long long *o, const SpinScal &sc, int n, const SpinRT &s,
                                  double tol, int mode, double &s_out, double &f_out)
{
    const int L = s.L;
    double mix[GEN_LMAX * GEN_LMAX], cof[GEN_LMAX * GEN_LMAX];
    if (sc.m == 0) {
        const double dd = cofactors_rt(s.d0, L, cof);
        double f = dd * sc.f0f[0];
        if (L > 0) {
            slab_t ff = slab + o[O_FF];
            for (int z = 0; z < L; ++z)
                for (int e = 0; e < L; ++e) f -= cof[e * L + z] * ff[s.cols[z] * n + s.rows[e]];
        }
        s_out = dd;
        f_out = f;
        return;
    }
    double sacc = 0.0, f = 0.0;
    if (sc.m <= L)
        for (unsigned bits = 0; bits < (1u << L); ++bits) {
            if (__popc(bits) != sc.m) continue;
            mix_rt(s.d0, s.d1, L, bits, mix);
            sacc += det_lu_rt(mix, L);
        }
    for (unsigned bits = 0; bits < (1u << (L + 1)); ++bits) {
        if (__popc(bits) != sc.m) continue;
        const int mi = bitk(bits, 0);
        mix_rt(s.d0, s.d1, L, bits >> 1, mix);
        const double dd = cofactors_rt(mix, L, cof);
        if (L > 0 && mode == 0 && !(fabs(dd) > tol)) continue;
        double contrib = dd * sc.f0f[mi];
        for (int b = 0; b < L; ++b) {
            slab_t fsl = slab + o[O_FF + mi * 2 + bitk(bits, b + 1)];
            double corr = 0.0;
            for (int r = 0; r < L; ++r) corr += (fsl[s.cols[b] * n + s.rows[r]] - mix[r * L + b]) * cof[r * L + b];
            contrib -= dd + corr;
        }
        f += contrib;
    }
    s_out = sacc;
    f_out = f;
}

__device__ inline void fs_elem_rt(const PairInfo &P, slab_t slab, const PackedDet &bra, const PackedDet &ket, int nmo,
                                  double tol, int mode, double &F, double &S)
{
    const SpinScal sa = load_spin(P, 0), sb = load_spin(P, 1);
    const long long *oa = P.off + O_AA, *ob = P.off + O_BB;
    SpinRT A, B;
    prep_rt(A, slab, oa, sa, nmo, bra.ia, bra.ra, ket.ia, ket.ra);
    prep_rt(B, slab, ob, sb, nmo, bra.ib, bra.rb, ket.ib, ket.rb);
    double ssa, fsa, ssb, fsb;
    fs_spin_rt(slab, oa, sa, nmo, A, tol, mode, ssa, fsa);
    fs_spin_rt(slab, ob, sb, nmo, B, tol, mode, ssb, fsb);
    const double pha = (double)(bra.pha * ket.pha), phb = (double)(bra.phb * ket.phb);
    const double s_a = pha * sa.pref * ssa, s_b = phb * sb.pref * ssb;
    double f = 0.0;
    if (s_b != 0.0) f += pha * sa.pref * fsa * s_b;
    if (s_a != 0.0) f += phb * sb.pref * fsb * s_a;
    F = f;
    S = s_a * s_b;
}

// Per-spin overlap sum of s_spin<L> for any L <= 8
__device__ inline double s_spin_rt(slab_t slab, const long long *o, const SpinScal &sc, int n, const uint8_t *bi,
                                   int brank, const uint8_t *ki, int krank)
{
    const int L = brank + krank;
    if (sc.m > L) return 0.0;
    SpinRT s;
    prep_rt(s, slab, o, sc, n, bi, brank, ki, krank);
    double mix[GEN_LMAX * GEN_LMAX], acc = 0.0;
    for (unsigned bits = 0; bits < (1u << L); ++bits) {
        if (__popc(bits) != sc.m) continue;
        mix_rt(s.d0, s.d1, L, bits, mix);
        acc += det_lu_rt(mix, L);
    }
    return acc;
}

}  // namespace nw
