/*
 * nwick_oracle.c -- CPU oracle for the nonorthogonal-Wick H,S matrix-element path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a scalar FP64 restatement of the
 * reference algorithm, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker.  The product library never links, loads or
 * calls it.
 *
 * Restated reference files (paths relative to the noci-rs source tree):
 *   src/nonorthogonalwicks/build.rs        per-reference-pair intermediates
 *   src/nonorthogonalwicks/layout.rs       slab layout (assign_offsets, write*)
 *   src/nonorthogonalwicks/eval/prepare.rs contraction-determinant labels / D
 *   src/nonorthogonalwicks/eval/helpers.rs distributions, replacements, jslot
 *   src/nonorthogonalwicks/eval/preparehamiltonianoverlap.rs
 *                                          m=0 and m>0 H,S evaluators
 *   src/maths/eri.rs                       AO->MO four-index transform
 *   src/maths/wick.rs                      tiny det / adjugate / minors
 *   src/noci/naive.rs                      generalised Slater-Condon route
 *   src/basis/excitation.rs                excitation phases
 *
 * Third-party arithmetic the reference takes from ndarray-linalg 0.15.0 /
 * OpenBLAS 0.10.13 (SVD, det) is restated here as a one-sided Jacobi SVD and
 * a partial-pivot LU determinant.  Singular vectors are gauge-dependent; only
 * gauge-invariant results (matrix elements, m, phase*prod(sigma)) are compared.
 *
 * Pinning (see tests/test_oracle.py): Wick elements agree with the independent
 * generalised Slater-Condon route (the reference's own *_agree check,
 * tests/noci.rs:303-320) including m=1 and m=2 reference pairs, and with a
 * brute-force second-quantised expansion in the full determinant space for
 * tiny systems; reference-pair S equals det(C_x,occ^T S C_w,occ); the Jacobi
 * SVD reproduces numpy/LAPACK singular values.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------- */
/* Dense helpers (row-major).                                                */
/* ------------------------------------------------------------------------- */

/* C[MxN] = op(A) op(B); opA: 0 -> A[M x K] (lda), 1 -> A^T where A is [K x M]. */
static void mm(int M, int N, int K, const double *A, int lda, int ta,
               const double *B, int ldb, int tb, double *C, int ldc)
{
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < N; ++j) {
            double acc = 0.0;
            for (int k = 0; k < K; ++k) {
                double a = ta ? A[(size_t)k * lda + i] : A[(size_t)i * lda + k];
                double b = tb ? B[(size_t)j * ldb + k] : B[(size_t)k * ldb + j];
                acc += a * b;
            }
            C[(size_t)i * ldc + j] = acc;
        }
}

/* sum_{a,b} g[b,a] h[a,b]  (src/maths/einsum.rs:25 einsum_ba_ab_real) */
static double einsum_ba_ab(int n, const double *g, const double *h)
{
    double acc = 0.0;
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
            acc += g[(size_t)b * n + a] * h[(size_t)a * n + b];
    return acc;
}

/* Determinant by LU with partial pivoting (the reference's det_gen for n>=5). */
static double det_lu(int n, const double *a)
{
    if (n == 0) return 1.0;
    double lu[64 * 64];
    memcpy(lu, a, sizeof(double) * n * n);
    double sign = 1.0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double pa = fabs(lu[k * n + k]);
        for (int r = k + 1; r < n; ++r)
            if (fabs(lu[r * n + k]) > pa) { pa = fabs(lu[r * n + k]); p = r; }
        if (pa == 0.0) return 0.0;
        if (p != k) {
            for (int c = 0; c < n; ++c) { double t = lu[k * n + c]; lu[k * n + c] = lu[p * n + c]; lu[p * n + c] = t; }
            sign = -sign;
        }
        for (int r = k + 1; r < n; ++r) {
            double f = lu[r * n + k] / lu[k * n + k];
            for (int c = k + 1; c < n; ++c) lu[r * n + c] -= f * lu[k * n + c];
        }
    }
    double d = sign;
    for (int i = 0; i < n; ++i) d *= lu[i * n + i];
    return d;
}

static double det2(double a00, double a01, double a10, double a11) { return a00 * a11 - a01 * a10; }
static double det3(double a00, double a01, double a02, double a10, double a11, double a12,
                   double a20, double a21, double a22)
{
    return a00 * det2(a11, a12, a21, a22) - a01 * det2(a10, a12, a20, a22) + a02 * det2(a10, a11, a20, a21);
}

/* det (src/maths/wick.rs:1273): closed forms for n<=4, LU beyond. */
static double det_small(const double *a, int n)
{
    switch (n) {
    case 0: return 1.0;
    case 1: return a[0];
    case 2: return det2(a[0], a[1], a[2], a[3]);
    case 3: return det3(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]);
    default: return det_lu(n, a);
    }
}

/* Minor of a (n x n) with row r and column c removed. */
static void build_minor(double *out, const double *a, int n, int r, int c)
{
    int k = 0;
    for (int i = 0; i < n; ++i) {
        if (i == r) continue;
        for (int j = 0; j < n; ++j) {
            if (j == c) continue;
            out[k++] = a[i * n + j];
        }
    }
}

/* adjugate_transpose (src/maths/wick.rs:1923): cofactor matrix
 * adjt[r,c] = (-1)^{r+c} det(minor(r,c)); returns det(a).  For n<=4 the
 * reference never rejects finite input; n>=5 is reached only by ranks
 * beyond singles/doubles and uses the cofactor expansion here. */
static double adjugate_transpose(double *adjt, const double *a, int n)
{
    if (n == 0) return 1.0;
    if (n == 1) { adjt[0] = 1.0; return a[0]; }
    double mn[64 * 64];
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            build_minor(mn, a, n, r, c);
            double md = det_small(mn, n - 1);
            adjt[r * n + c] = (((r + c) & 1) ? -1.0 : 1.0) * md;
        }
    return det_small(a, n);
}

/* adjugate_transpose_generic (eval/helpers.rs:726): reject |det| <= tol. */
static int adjugate_transpose_generic(double *adjt, const double *a, int n, double tol, double *detv)
{
    double d = adjugate_transpose(adjt, a, n);
    if (!isfinite(d)) return 0;
    if (fabs(d) <= tol) return 0;
    *detv = d;
    return 1;
}

/* ------------------------------------------------------------------------- */
/* One-sided Jacobi SVD of a square n x n matrix: A = U diag(s) V^T,          */
/* s descending (LAPACK order).  Zero singular vectors on the left are        */
/* completed orthonormally.  Returns det(U) det(V).                           */
/* ------------------------------------------------------------------------- */
int ow_svd(int n, const double *A, double *U, double *s, double *V)
{
    double *W = (double *)malloc(sizeof(double) * n * n);
    double *Vt = (double *)malloc(sizeof(double) * n * n);
    memcpy(W, A, sizeof(double) * n * n);
    for (int i = 0; i < n * n; ++i) Vt[i] = 0.0;
    for (int i = 0; i < n; ++i) Vt[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 80; ++sweep) {
        int rotated = 0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double al = 0, be = 0, ga = 0;
                for (int i = 0; i < n; ++i) {
                    double wp = W[i * n + p], wq = W[i * n + q];
                    al += wp * wp; be += wq * wq; ga += wp * wq;
                }
                if (fabs(ga) <= 1e-15 * sqrt(al * be) || ga == 0.0) continue;
                rotated = 1;
                double zeta = (be - al) / (2.0 * ga);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                for (int i = 0; i < n; ++i) {
                    double wp = W[i * n + p], wq = W[i * n + q];
                    W[i * n + p] = c * wp - sn * wq;
                    W[i * n + q] = sn * wp + c * wq;
                    double vp = Vt[i * n + p], vq = Vt[i * n + q];
                    Vt[i * n + p] = c * vp - sn * vq;
                    Vt[i * n + q] = sn * vp + c * vq;
                }
            }
        if (!rotated) break;
    }
    /* singular values and descending order */
    int *ord = (int *)malloc(sizeof(int) * n);
    double *sv = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; ++k) {
        double acc = 0;
        for (int i = 0; i < n; ++i) acc += W[i * n + k] * W[i * n + k];
        sv[k] = sqrt(acc);
        ord[k] = k;
    }
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b)
            if (sv[ord[b]] > sv[ord[a]]) { int t = ord[a]; ord[a] = ord[b]; ord[b] = t; }
    double smax = n ? sv[ord[0]] : 0.0;
    for (int k = 0; k < n; ++k) {
        int src = ord[k];
        s[k] = sv[src];
        int good = sv[src] > 1e-13 * (smax > 0 ? smax : 1.0);
        for (int i = 0; i < n; ++i) U[i * n + k] = good ? W[i * n + src] / sv[src] : 0.0;
        for (int i = 0; i < n; ++i) V[i * n + k] = good ? Vt[i * n + src] : 0.0;
        if (!good) {
            /* Numerically zero singular values: the left and right null
             * spaces are completed canonically (Gram-Schmidt of unit vectors
             * against the preceding columns), so the pairing u_k <-> v_k does
             * not depend on the rotation history.  Exact matrix elements are
             * gauge invariant; the reference's m>0 arithmetic is not for
             * m >= 2, which makes a fixed gauge necessary to compare it. */
            for (int side = 0; side < 2; ++side) {
                double *Q = side ? V : U;
                for (int e = 0; e < n; ++e) {
                    double cand[64];
                    for (int i = 0; i < n; ++i) cand[i] = (i == e) ? 1.0 : 0.0;
                    for (int pass = 0; pass < 2; ++pass)
                        for (int j = 0; j < k; ++j) {
                            double d = 0;
                            for (int i = 0; i < n; ++i) d += Q[i * n + j] * cand[i];
                            for (int i = 0; i < n; ++i) cand[i] -= d * Q[i * n + j];
                        }
                    double nn = 0;
                    for (int i = 0; i < n; ++i) nn += cand[i] * cand[i];
                    if (nn > 1e-2) {
                        nn = sqrt(nn);
                        for (int i = 0; i < n; ++i) Q[i * n + k] = cand[i] / nn;
                        break;
                    }
                }
            }
        }
    }
    free(ord); free(sv); free(W); free(Vt);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* AO -> MO four-index transform (src/maths/eri.rs:258-313):                 */
/* out[p,q,r,s] = sum c1[mu,p] c2[nu,q] c3[lam,r] c4[sig,s] (mu nu|lam sig)  */
/* ------------------------------------------------------------------------- */
/* One quarter of the transform as a cache-blocked product:
 * Y[s, row] = sum_l X[row, l] c[l, s] for X rows x nb, c nb x nm, Y nm x rows.
 * Rows go in blocks of QP_RB through a 4 x 8 register tile (the inner loop
 * of 8 columns vectorises), the block's result is written transposed, so the
 * next quarter again contracts the last index. */
#define QP_RB 64
typedef double v4d __attribute__((vector_size(32)));

/* An external BLAS dgemm (Fortran calling convention) for the quarter passes:
 * the reference computes its transforms with ndarray GEMMs on OpenBLAS
 * (ndarray-linalg 0.15 with openblas-system), so the CPU baseline can time the
 * build with that library instead of the loops below.  NULL: built-in loops. */
typedef void (*ow_dgemm_fn)(const char *ta, const char *tb, const int *m, const int *n, const int *k,
                            const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
                            const double *beta, double *c, const int *ldc);
static ow_dgemm_fn ext_dgemm = NULL;
void ow_set_dgemm(void *fn) { ext_dgemm = (ow_dgemm_fn)fn; }

static void qpass(long long rows, int nb, int nm, const double *X, const double *c, double *Y)
{
    if (ext_dgemm && rows < 0x7fffffffLL) {
        /* column-major: Y^T (rows x nm, ld rows) = X^T^T c^T^T with X^T (nb x rows)
         * and c^T (nm x nb) the row-major operands read column-major */
        const int m = (int)rows, n = nm, k = nb, lda = nb, ldb = nm, ldc = (int)rows;
        const double one = 1.0, zero = 0.0;
        ext_dgemm("T", "T", &m, &n, &k, &one, X, &lda, c, &ldb, &zero, Y, &ldc);
        return;
    }
    const int np = (nm + 7) & ~7;
    double *cp = (double *)calloc((size_t)nb * np, sizeof(double));
    for (int l = 0; l < nb; ++l)
        for (int t = 0; t < nm; ++t) cp[(size_t)l * np + t] = c[(size_t)l * nm + t];
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * QP_RB * np);
#pragma omp for schedule(static)
        for (long long r0 = 0; r0 < rows; r0 += QP_RB) {
            const int rn = (int)(rows - r0 < QP_RB ? rows - r0 : QP_RB);
            int rr = 0;
            for (; rr + 4 <= rn; rr += 4) {
                const double *x0 = X + (size_t)(r0 + rr) * nb;
                for (int s0 = 0; s0 < np; s0 += 8) {
                    v4d a00 = {0, 0, 0, 0}, a01 = a00, a10 = a00, a11 = a00, a20 = a00, a21 = a00, a30 = a00, a31 = a00;
                    for (int l = 0; l < nb; ++l) {
                        const double *bl = cp + (size_t)l * np + s0;
                        v4d b0, b1;
                        memcpy(&b0, bl, 32);
                        memcpy(&b1, bl + 4, 32);
                        const double x0l = x0[l], x1l = x0[nb + l], x2l = x0[2 * nb + l], x3l = x0[3 * nb + l];
                        a00 += x0l * b0; a01 += x0l * b1;
                        a10 += x1l * b0; a11 += x1l * b1;
                        a20 += x2l * b0; a21 += x2l * b1;
                        a30 += x3l * b0; a31 += x3l * b1;
                    }
                    double *o = buf + (size_t)rr * np + s0;
                    memcpy(o, &a00, 32); memcpy(o + 4, &a01, 32);
                    memcpy(o + np, &a10, 32); memcpy(o + np + 4, &a11, 32);
                    memcpy(o + 2 * np, &a20, 32); memcpy(o + 2 * np + 4, &a21, 32);
                    memcpy(o + 3 * np, &a30, 32); memcpy(o + 3 * np + 4, &a31, 32);
                }
            }
            for (; rr < rn; ++rr) {
                const double *x0 = X + (size_t)(r0 + rr) * nb;
                for (int t = 0; t < np; ++t) buf[(size_t)rr * np + t] = 0.0;
                for (int l = 0; l < nb; ++l)
                    for (int t = 0; t < np; ++t) buf[(size_t)rr * np + t] += x0[l] * cp[(size_t)l * np + t];
            }
            for (int t = 0; t < nm; ++t) {
                double *dst = Y + (size_t)t * rows + r0;
                for (int i = 0; i < rn; ++i) dst[i] = buf[(size_t)i * np + t];
            }
        }
        free(buf);
    }
    free(cp);
}

/* ------------------------------------------------------------------------- */
/* AO -> MO four-index transform (src/maths/eri.rs:258-313):                 */
/* out[p,q,r,s] = sum c1[mu,p] c2[nu,q] c3[lam,r] c4[sig,s] (mu nu|lam sig)  */
/* as four quarter passes, each contracting the last index and moving the    */
/* new index to the front: [mu nu lam sig] -> [s mu nu lam] -> [r s mu nu]   */
/* -> [q r s mu] -> [p q r s].                                               */
/* ------------------------------------------------------------------------- */
static void ao2mo(int nb, int nm, const double *eri, const double *c1, const double *c2,
                  const double *c3, const double *c4, double *out)
{
    size_t nb2 = (size_t)nb * nb, nm2 = (size_t)nm * nm;
    double *t1 = (double *)malloc(sizeof(double) * nb2 * nb * nm);
    qpass((long long)(nb2 * nb), nb, nm, eri, c4, t1);
    double *t2 = (double *)malloc(sizeof(double) * nb2 * nm2);
    qpass((long long)(nm * nb2), nb, nm, t1, c3, t2);
    free(t1);
    double *t3 = (double *)malloc(sizeof(double) * nb * nm2 * nm);
    qpass((long long)(nm2 * nb), nb, nm, t2, c2, t3);
    free(t2);
    qpass((long long)(nm2 * nm), nb, nm, t3, c1, out);
    free(t3);
}

/* ------------------------------------------------------------------------- */
/* Per-reference-pair build (src/nonorthogonalwicks/build.rs).                */
/* ------------------------------------------------------------------------- */

/* SAME_SPIN_J_BRANCHES (build.rs:19-30) */
static const int JBR[10][4] = {
    {0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 1, 0}, {0, 0, 1, 1}, {0, 1, 0, 1},
    {0, 1, 1, 0}, {0, 1, 1, 1}, {1, 0, 1, 0}, {1, 0, 1, 1}, {1, 1, 1, 1}};

typedef struct {
    int nb, nm, nocc, m;
    double tsp, phase;
    double *mao[2];      /* nb x nb */
    double *cx[2], *xc[2]; /* nb x nm */
    double *x[2], *y[2];  /* nm x nm */
} spin_core;

static void free_core(spin_core *c)
{
    for (int i = 0; i < 2; ++i) { free(c->mao[i]); free(c->cx[i]); free(c->xc[i]); free(c->x[i]); free(c->y[i]); }
}

/* occupied columns, ascending bit order (naive.rs:20 occ_coeffs) */
static int occ_cols(int nb, int nm, const double *c, u128 occ, double *out)
{
    int k = 0;
    for (int i = 0; i < nm; ++i)
        if ((occ >> i) & 1) {
            for (int mu = 0; mu < nb; ++mu) out[(size_t)mu * 64 + k] = c[(size_t)mu * nm + i];
            ++k;
        }
    return k;
}

/* perform_ortho_and_svd_and_rotate (build.rs:401), construct_m (:447),
 * build_cx_xc (:944), contraction_orbitals (:1008), construct_xy (:526). */
static void spin_core_build(int nb, int nm, const double *S, const double *lc, u128 lo,
                            const double *gc, u128 go, double tol, spin_core *o)
{
    memset(o, 0, sizeof(*o));
    o->nb = nb; o->nm = nm;
    double *lco = (double *)calloc((size_t)nb * 64, sizeof(double));
    double *gco = (double *)calloc((size_t)nb * 64, sizeof(double));
    int nocc = occ_cols(nb, nm, lc, lo, lco);
    occ_cols(nb, nm, gc, go, gco);
    o->nocc = nocc;
    /* S_occ = lco^T S gco */
    double *tmp = (double *)calloc((size_t)nb * 64, sizeof(double));
    mm(nb, nocc, nb, S, nb, 0, gco, 64, 0, tmp, 64);
    double socc[64 * 64], U[64 * 64], V[64 * 64], sv[64];
    mm(nocc, nocc, nb, lco, 64, 1, tmp, 64, 0, socc, nocc);
    ow_svd(nocc, socc, U, sv, V);
    o->phase = det_small(U, nocc) * det_small(V, nocc);
    double tsp = 1.0;
    int m = 0, zeros[64];
    for (int k = 0; k < nocc; ++k) {
        if (fabs(sv[k]) > tol) tsp *= sv[k];
        else zeros[m++] = k;
    }
    o->tsp = tsp; o->m = m;
    /* rotated occupied orbitals */
    double *lt = (double *)calloc((size_t)nb * 64, sizeof(double));
    double *gt = (double *)calloc((size_t)nb * 64, sizeof(double));
    mm(nb, nocc, nocc, lco, 64, 0, U, nocc, 0, lt, 64);
    mm(nb, nocc, nocc, gco, 64, 0, V, nocc, 0, gt, 64);
    /* construct_m */
    size_t nb2 = (size_t)nb * nb;
    o->mao[0] = (double *)calloc(nb2, sizeof(double));
    o->mao[1] = (double *)calloc(nb2, sizeof(double));
    for (int mu = 0; mu < nb; ++mu)
        for (int nu = 0; nu < nb; ++nu) {
            double acc = 0;
            for (int k = 0; k < nocc; ++k)
                if (fabs(sv[k]) > tol) acc += (gt[mu * 64 + k] * (1.0 / sv[k])) * lt[nu * 64 + k];
            o->mao[0][mu * nb + nu] = acc;
        }
    for (int z = 0; z < m; ++z) {
        int k = zeros[z];
        for (int mu = 0; mu < nb; ++mu)
            for (int nu = 0; nu < nb; ++nu) o->mao[0][mu * nb + nu] += gt[mu * 64 + k] * gt[nu * 64 + k];
    }
    for (int z = 0; z < m; ++z) {
        int k = zeros[z];
        for (int mu = 0; mu < nb; ++mu)
            for (int nu = 0; nu < nb; ++nu) {
                double p = gt[mu * 64 + k] * lt[nu * 64 + k];
                o->mao[1][mu * nb + nu] += p;
                o->mao[0][mu * nb + nu] += p;
            }
    }
    /* contraction_orbitals: rowc = [l virt | g occ], colc = [l occ | g virt] */
    int nv = nm - nocc;
    double *rowc = (double *)calloc((size_t)nb * nm, sizeof(double));
    double *colc = (double *)calloc((size_t)nb * nm, sizeof(double));
    for (int mu = 0; mu < nb; ++mu) {
        for (int j = 0; j < nv; ++j) rowc[mu * nm + j] = lc[mu * nm + nocc + j];
        for (int j = 0; j < nocc; ++j) rowc[mu * nm + nv + j] = gc[mu * nm + j];
        for (int j = 0; j < nocc; ++j) colc[mu * nm + j] = lc[mu * nm + j];
        for (int j = nocc; j < nm; ++j) colc[mu * nm + j] = gc[mu * nm + j];
    }
    double *t1 = (double *)calloc(nb2, sizeof(double));
    double *t2 = (double *)calloc(nb2, sizeof(double));
    double *t3 = (double *)calloc((size_t)nb * nm, sizeof(double));
    for (int mi = 0; mi < 2; ++mi) {
        const double *M = o->mao[mi];
        /* construct_xy: xrdm = S M S; yrdm = xrdm - S (mi=0) */
        mm(nb, nb, nb, S, nb, 0, M, nb, 0, t1, nb);
        mm(nb, nb, nb, t1, nb, 0, S, nb, 0, t2, nb); /* xrdm */
        o->x[mi] = (double *)calloc((size_t)nm * nm, sizeof(double));
        o->y[mi] = (double *)calloc((size_t)nm * nm, sizeof(double));
        mm(nb, nm, nb, t2, nb, 0, colc, nm, 0, t3, nm);
        mm(nm, nm, nb, rowc, nm, 1, t3, nm, 0, o->x[mi], nm);
        if (mi == 0)
            for (size_t i = 0; i < nb2; ++i) t2[i] -= S[i];
        mm(nb, nm, nb, t2, nb, 0, colc, nm, 0, t3, nm);
        mm(nm, nm, nb, rowc, nm, 1, t3, nm, 0, o->y[mi], nm);
        /* build_cx_xc */
        double omi = (double)(1 - mi);
        double *ms = t1, *mts = t2;
        mm(nb, nb, nb, M, nb, 0, S, nb, 0, ms, nb);
        mm(nb, nb, nb, M, nb, 1, S, nb, 0, mts, nb);
        double *full = (double *)calloc((size_t)nb * nm, sizeof(double));
        o->cx[mi] = (double *)calloc((size_t)nb * nm, sizeof(double));
        o->xc[mi] = (double *)calloc((size_t)nb * nm, sizeof(double));
        /* cx = [M^T S l_c - (1-i) l_c](virt cols) | [M^T S g_c](occ cols) */
        mm(nb, nm, nb, mts, nb, 0, lc, nm, 0, full, nm);
        for (int mu = 0; mu < nb; ++mu)
            for (int j = 0; j < nv; ++j)
                o->cx[mi][mu * nm + j] = full[mu * nm + nocc + j] - omi * lc[mu * nm + nocc + j];
        mm(nb, nm, nb, mts, nb, 0, gc, nm, 0, full, nm);
        for (int mu = 0; mu < nb; ++mu)
            for (int j = 0; j < nocc; ++j) o->cx[mi][mu * nm + nv + j] = full[mu * nm + j];
        /* xc = [M S l_c](occ cols) | [M S g_c - (1-i) g_c](virt cols) */
        mm(nb, nm, nb, ms, nb, 0, lc, nm, 0, full, nm);
        for (int mu = 0; mu < nb; ++mu)
            for (int j = 0; j < nocc; ++j) o->xc[mi][mu * nm + j] = full[mu * nm + j];
        mm(nb, nm, nb, ms, nb, 0, gc, nm, 0, full, nm);
        for (int mu = 0; mu < nb; ++mu)
            for (int j = nocc; j < nm; ++j)
                o->xc[mi][mu * nm + j] = full[mu * nm + j] - omi * gc[mu * nm + j];
        free(full);
    }
    free(t1); free(t2); free(t3); free(rowc); free(colc);
    free(lco); free(gco); free(tmp); free(lt); free(gt);
}

/* build_j_coulomb / build_k_exchange (build.rs:627-683) */
static void coulomb(int n, const double *eri, const double *M, double *J)
{
    size_t n2 = (size_t)n * n;
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s)
        for (int t = 0; t < n; ++t) {
            double acc = 0;
            const double *e = eri + ((size_t)s * n + t) * n2;
            for (size_t k = 0; k < n2; ++k) acc += e[k] * M[k];
            J[s * n + t] = acc;
        }
}
static void exchange(int n, const double *eri, const double *M, double *K)
{
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        double *k = K + (size_t)s * n;
        for (int t = 0; t < n; ++t) k[t] = 0.0;
        for (int mu = 0; mu < n; ++mu)
            for (int nu = 0; nu < n; ++nu) {
                const double m = M[mu * n + nu];
                const double *e = eri + (((size_t)s * n + mu) * n + nu) * n;
                for (int t = 0; t < n; ++t) k[t] += e[t] * m;
            }
    }
}

/* ------------------------------------------------------------------------- */
/* Slab layout (layout.rs:21-189), one ordered reference pair.               */
/* Offsets are relative to the pair's base.                                  */
/* ------------------------------------------------------------------------- */
#define OFF_NONE (-1LL)
enum {
    O_X = 0, O_Y = 2, O_FH = 4, O_FF = 8, O_V = 12, O_J = 20, O_HCOL0 = 30, O_SAME = 31,
    O_AA = 0, O_BB = 31, O_VAB = 62, O_VBA = 70, O_II = 78, O_TOTAL = 94, O_SIZE = 94
};

/* Returns the pair slab length in doubles; off[] has O_TOTAL+1 entries (last = size). */
long long ow_pair_offsets(int ma, int mb, int nmo, long long *off)
{
    long long nn2 = (long long)nmo * nmo, nn4 = nn2 * nn2, i = 0;
    for (int k = 0; k < O_TOTAL; ++k) off[k] = OFF_NONE;
    for (int sp = 0; sp < 2; ++sp) {
        long long *o = off + (sp ? O_BB : O_AA);
        int msp = sp ? mb : ma;
        for (int mi = 0; mi < 2; ++mi) { o[O_X + mi] = i; i += nn2; }
        for (int mi = 0; mi < 2; ++mi) { o[O_Y + mi] = i; i += nn2; }
        for (int k = 0; k < 4; ++k) { o[O_FH + k] = i; i += nn2; }
        for (int k = 0; k < 4; ++k) { o[O_FF + k] = i; i += nn2; }
        for (int k = 0; k < 8; ++k) { o[O_V + k] = i; i += nn2; }
        for (int s = 0; s < 10; ++s)
            if (JBR[s][0] + JBR[s][1] + JBR[s][2] + JBR[s][3] <= msp) { o[O_J + s] = i; i += nn4; }
    }
    for (int k = 0; k < 8; ++k) { off[O_VAB + k] = i; i += nn2; }
    for (int k = 0; k < 8; ++k) { off[O_VBA + k] = i; i += nn2; }
    if (ma == 0 && mb == 0) {
        off[O_AA + O_HCOL0] = i; i += nn2;
        off[O_BB + O_HCOL0] = i; i += nn2;
    }
    for (int a0 = 0; a0 < 2; ++a0)
        for (int aj = 0; aj < 2; ++aj)
            for (int b0 = 0; b0 < 2; ++b0)
                for (int bj = 0; bj < 2; ++bj)
                    if (a0 + aj <= ma && b0 + bj <= mb) {
                        off[O_II + ((a0 * 2 + aj) * 2 + b0) * 2 + bj] = i; i += nn4;
                    }
    off[O_TOTAL] = i;
    return i;
}

static void write2(double *slab, long long o, const double *a, int n) { memcpy(slab + o, a, sizeof(double) * n * n); }
static void write2t(double *slab, long long o, const double *a, int n)
{
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) slab[o + (long long)c * n + r] = a[r * n + c];
}

/* Pair meta (scalars) layout, META_LEN doubles:
 *  per spin (aa at 0, bb at 11): tsp, phase, m, nocc, f0h[2], f0f[2], v0[3]
 *  ab at 22: vab0[2][2], vba0[2][2] */
enum { M_TSP = 0, M_PH = 1, M_M = 2, M_NOCC = 3, M_F0H = 4, M_F0F = 6, M_V0 = 8, M_SPIN = 11,
       M_AA = 0, M_BB = 11, M_VAB0 = 22, M_VBA0 = 26, META_LEN = 32 };

/* Count m per spin for an ordered pair (count_zero_overlap_pairs, build.rs:350). */
int ow_count_zero(int nb, int nm, const double *S, const double *lc, const uint64_t *lo,
                  const double *gc, const uint64_t *go, double tol, double *sig_out)
{
    u128 l = ((u128)lo[1] << 64) | lo[0], g = ((u128)go[1] << 64) | go[0];
    double *lco = (double *)calloc((size_t)nb * 64, sizeof(double));
    double *gco = (double *)calloc((size_t)nb * 64, sizeof(double));
    int nocc = occ_cols(nb, nm, lc, l, lco);
    occ_cols(nb, nm, gc, g, gco);
    double *tmp = (double *)calloc((size_t)nb * 64, sizeof(double));
    double socc[64 * 64], U[64 * 64], V[64 * 64], sv[64];
    mm(nb, nocc, nb, S, nb, 0, gco, 64, 0, tmp, 64);
    mm(nocc, nocc, nb, lco, 64, 1, tmp, 64, 0, socc, nocc);
    ow_svd(nocc, socc, U, sv, V);
    int m = 0;
    for (int k = 0; k < nocc; ++k) {
        if (sig_out) sig_out[k] = sv[k];
        if (fabs(sv[k]) <= tol) ++m;
    }
    free(lco); free(gco); free(tmp);
    return m;
}

/* Build one ordered reference pair (x = bra = l, w = ket = g) into slab (pair
 * base) + meta.  Mirrors src/noci/wicks.rs:37-68 build_wicks_pair:
 * SameSpinBuild alpha, beta, DiffSpinBuild, write_same_spin x2,
 * write_diff_spin, write_hcol0.  eri: dense chemists' [a,b,c,d]. */
int ow_build_pair(int nb, int nm, const double *S, const double *h, const double *eri,
                  const double *lca, const double *lcb, const uint64_t *loa, const uint64_t *lob,
                  const double *gca, const double *gcb, const uint64_t *goa, const uint64_t *gob,
                  double tol, const long long *off, double *slab, double *meta)
{
    u128 la = ((u128)loa[1] << 64) | loa[0], lb = ((u128)lob[1] << 64) | lob[0];
    u128 ga = ((u128)goa[1] << 64) | goa[0], gb = ((u128)gob[1] << 64) | gob[0];
    spin_core core[2];
    spin_core_build(nb, nm, S, lca, la, gca, ga, tol, &core[0]);
    spin_core_build(nb, nm, S, lcb, lb, gcb, gb, tol, &core[1]);
    size_t nb2 = (size_t)nb * nb, nm2 = (size_t)nm * nm, nm4 = nm2 * nm2;
    double *jk[2][2], *jc[2][2];
    for (int sp = 0; sp < 2; ++sp)
        for (int mi = 0; mi < 2; ++mi) {
            jk[sp][mi] = (double *)calloc(nb2, sizeof(double));
            jc[sp][mi] = (double *)calloc(nb2, sizeof(double));
            double *K = (double *)calloc(nb2, sizeof(double));
            coulomb(nb, eri, core[sp].mao[mi], jc[sp][mi]);
            exchange(nb, eri, core[sp].mao[mi], K);
            for (size_t i = 0; i < nb2; ++i) jk[sp][mi][i] = jc[sp][mi][i] - K[i];
            free(K);
        }
    double *tmp = (double *)calloc((size_t)nb * nm, sizeof(double));
    double *blk2 = (double *)calloc(nm2, sizeof(double));
    for (int k = 0; k < META_LEN; ++k) meta[k] = 0.0;
    for (int sp = 0; sp < 2; ++sp) {
        spin_core *c = &core[sp];
        const long long *o = off + (sp ? O_BB : O_AA);
        double *mt = meta + (sp ? M_BB : M_AA);
        int m = c->m;
        mt[M_TSP] = c->tsp; mt[M_PH] = c->phase; mt[M_M] = m; mt[M_NOCC] = c->nocc;
        write2(slab, o[O_X + 0], c->x[0], nm); write2(slab, o[O_X + 1], c->x[1], nm);
        write2(slab, o[O_Y + 0], c->y[0], nm); write2(slab, o[O_Y + 1], c->y[1], nm);
        /* fh = cx^T h xc; f0h = einsum(M, h)  (build.rs:199-207) */
        for (int mi = 0; mi < 2; ++mi)
            for (int mj = 0; mj < 2; ++mj) {
                mm(nb, nm, nb, h, nb, 0, c->xc[mj], nm, 0, tmp, nm);
                mm(nm, nm, nb, c->cx[mi], nm, 1, tmp, nm, 0, blk2, nm);
                write2t(slab, o[O_FH + mi * 2 + mj], blk2, nm);
                memset(blk2, 0, sizeof(double) * nm2);
                write2t(slab, o[O_FF + mi * 2 + mj], blk2, nm);
            }
        mt[M_F0H + 0] = einsum_ba_ab(nb, c->mao[0], h);
        mt[M_F0H + 1] = einsum_ba_ab(nb, c->mao[1], h);
        /* v0 (build.rs:227-236) */
        mt[M_V0 + 0] = einsum_ba_ab(nb, jk[sp][0], c->mao[0]);
        if (m >= 1) mt[M_V0 + 1] = 2.0 * einsum_ba_ab(nb, jk[sp][0], c->mao[1]);
        if (m >= 2) mt[M_V0 + 2] = einsum_ba_ab(nb, jk[sp][1], c->mao[1]);
        /* v[mi][mj][mk] = cx[mi]^T JK[mk] xc[mj]  (build.rs:247-259) */
        for (int mi = 0; mi < 2; ++mi)
            for (int mj = 0; mj < 2; ++mj)
                for (int mk = 0; mk < 2; ++mk) {
                    mm(nb, nm, nb, jk[sp][mk], nb, 0, c->xc[mj], nm, 0, tmp, nm);
                    mm(nm, nm, nb, c->cx[mi], nm, 1, tmp, nm, 0, blk2, nm);
                    write2t(slab, o[O_V + (mi * 2 + mj) * 2 + mk], blk2, nm);
                }
        /* J slots (build.rs:267-313): direct minus exchange transform,
         * written with write4ijrc (layout.rs:367). */
        double *blk = (double *)malloc(sizeof(double) * nm4);
        double *exch = (double *)malloc(sizeof(double) * nm4);
        for (int s = 0; s < 10; ++s) {
            if (o[O_J + s] == OFF_NONE) continue;
            int i = JBR[s][0], j = JBR[s][1], k = JBR[s][2], l = JBR[s][3];
            ao2mo(nb, nm, eri, c->cx[i], c->xc[j], c->cx[k], c->xc[l], blk);
            ao2mo(nb, nm, eri, c->cx[i], c->cx[k], c->xc[j], c->xc[l], exch);
#pragma omp parallel for schedule(static)
            for (int p = 0; p < nm; ++p)
                for (int q = 0; q < nm; ++q)
                    for (int r = 0; r < nm; ++r)
                        for (int t = 0; t < nm; ++t)
                            blk[((p * (size_t)nm + q) * nm + r) * nm + t] -= exch[((p * (size_t)nm + r) * nm + q) * nm + t];
            double *dst = slab + o[O_J + s];
#pragma omp parallel for schedule(static)
            for (int r = 0; r < nm; ++r)
                for (int cc = 0; cc < nm; ++cc)
                    for (int ii = 0; ii < nm; ++ii)
                        for (int jj = 0; jj < nm; ++jj)
                            dst[((ii * (size_t)nm + jj) * nm + r) * nm + cc] = blk[((r * (size_t)nm + cc) * nm + ii) * nm + jj];
        }
        free(blk); free(exch);
    }
    /* DiffSpinBuild (build.rs:728-926) */
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            meta[M_VAB0 + i * 2 + j] = einsum_ba_ab(nb, jc[0][i], core[1].mao[j]);
            meta[M_VBA0 + j * 2 + i] = einsum_ba_ab(nb, jc[1][j], core[0].mao[i]);
        }
    for (int a0 = 0; a0 < 2; ++a0)
        for (int b0 = 0; b0 < 2; ++b0)
            for (int k = 0; k < 2; ++k) {
                /* vab[ma0][mb0][mak] = cx_a[ma0]^T Jb[mb0] xc_a[mak] */
                mm(nb, nm, nb, jc[1][b0], nb, 0, core[0].xc[k], nm, 0, tmp, nm);
                mm(nm, nm, nb, core[0].cx[a0], nm, 1, tmp, nm, 0, blk2, nm);
                write2t(slab, off[O_VAB + (a0 * 2 + b0) * 2 + k], blk2, nm);
                /* vba[mb0][ma0][mbk] = cx_b[mb0]^T Ja[ma0] xc_b[mbk] */
                mm(nb, nm, nb, jc[0][a0], nb, 0, core[1].xc[k], nm, 0, tmp, nm);
                mm(nm, nm, nb, core[1].cx[b0], nm, 1, tmp, nm, 0, blk2, nm);
                write2t(slab, off[O_VBA + (b0 * 2 + a0) * 2 + k], blk2, nm);
            }
    for (int a0 = 0; a0 < 2; ++a0)
        for (int aj = 0; aj < 2; ++aj)
            for (int b0 = 0; b0 < 2; ++b0)
                for (int bj = 0; bj < 2; ++bj) {
                    long long o = off[O_II + ((a0 * 2 + aj) * 2 + b0) * 2 + bj];
                    if (o == OFF_NONE) continue;
                    ao2mo(nb, nm, eri, core[0].cx[a0], core[0].xc[aj], core[1].cx[b0], core[1].xc[bj], slab + o);
                }
    /* write_hcol0 (layout.rs:283) */
    if (off[O_AA + O_HCOL0] != OFF_NONE) {
        for (size_t i = 0; i < nm2; ++i) {
            slab[off[O_AA + O_HCOL0] + i] = slab[off[O_AA + O_FH] + i] + slab[off[O_AA + O_V] + i] + slab[off[O_VAB] + i];
            slab[off[O_BB + O_HCOL0] + i] = slab[off[O_BB + O_FH] + i] + slab[off[O_BB + O_V] + i] + slab[off[O_VBA] + i];
        }
    }
    for (int sp = 0; sp < 2; ++sp)
        for (int mi = 0; mi < 2; ++mi) { free(jk[sp][mi]); free(jc[sp][mi]); }
    free(tmp); free(blk2);
    free_core(&core[0]); free_core(&core[1]);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Evaluation (eval/prepare.rs, eval/helpers.rs, preparehamiltonianoverlap.rs) */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n, nocc, m;
    double tsp, phase, f0h[2], f0f[2], v0[3];
    const double *x[2], *y[2], *fh[2][2], *ff[2][2], *v[2][2][2], *j[10], *hcol0;
} same_view;
typedef struct {
    int n;
    double vab0[2][2], vba0[2][2];
    const double *vab[2][2][2], *vba[2][2][2], *ii[2][2][2][2];
} diff_view;

static const double *ptr_or_null(const double *slab, long long o) { return o == OFF_NONE ? NULL : slab + o; }

static void make_views(const double *slab, const long long *off, const double *meta, int nm,
                       same_view *sv, diff_view *dv)
{
    for (int sp = 0; sp < 2; ++sp) {
        same_view *s = &sv[sp];
        const long long *o = off + (sp ? O_BB : O_AA);
        const double *mt = meta + (sp ? M_BB : M_AA);
        s->n = nm; s->nocc = (int)mt[M_NOCC]; s->m = (int)mt[M_M];
        s->tsp = mt[M_TSP]; s->phase = mt[M_PH];
        s->f0h[0] = mt[M_F0H]; s->f0h[1] = mt[M_F0H + 1];
        s->f0f[0] = mt[M_F0F]; s->f0f[1] = mt[M_F0F + 1];
        for (int k = 0; k < 3; ++k) s->v0[k] = mt[M_V0 + k];
        for (int mi = 0; mi < 2; ++mi) { s->x[mi] = slab + o[O_X + mi]; s->y[mi] = slab + o[O_Y + mi]; }
        for (int k = 0; k < 4; ++k) s->fh[k / 2][k % 2] = slab + o[O_FH + k];
        for (int k = 0; k < 4; ++k) s->ff[k / 2][k % 2] = slab + o[O_FF + k];
        for (int k = 0; k < 8; ++k) s->v[k / 4][(k / 2) % 2][k % 2] = slab + o[O_V + k];
        for (int k = 0; k < 10; ++k) s->j[k] = ptr_or_null(slab, o[O_J + k]);
        s->hcol0 = ptr_or_null(slab, o[O_HCOL0]);
    }
    dv->n = nm;
    for (int k = 0; k < 4; ++k) { dv->vab0[k / 2][k % 2] = meta[M_VAB0 + k]; dv->vba0[k / 2][k % 2] = meta[M_VBA0 + k]; }
    for (int k = 0; k < 8; ++k) {
        dv->vab[k / 4][(k / 2) % 2][k % 2] = slab + off[O_VAB + k];
        dv->vba[k / 4][(k / 2) % 2][k % 2] = slab + off[O_VBA + k];
    }
    for (int k = 0; k < 16; ++k) dv->ii[k / 8][(k / 4) % 2][(k / 2) % 2][k % 2] = ptr_or_null(slab, off[O_II + k]);
}

static int ctz128(u128 v) { uint64_t lo = (uint64_t)v; return lo ? __builtin_ctzll(lo) : 64 + __builtin_ctzll((uint64_t)(v >> 64)); }
static int pop128(u128 v) { return __builtin_popcountll((uint64_t)v) + __builtin_popcountll((uint64_t)(v >> 64)); }

/* construct_determinant_indices (prepare.rs:555-595) */
static int det_indices(const same_view *w, u128 xh, u128 xp, u128 wh, u128 wp, int *rows, int *cols)
{
    int nocc = w->nocc, nv = w->n - nocc, i = 0;
    while (xh) {
        int hole = ctz128(xh), part = ctz128(xp);
        xh &= xh - 1; xp &= xp - 1;
        rows[i] = part - nocc; cols[i] = hole; ++i;
    }
    while (wh) {
        int hole = ctz128(wh), part = ctz128(wp);
        wh &= wh - 1; wp &= wp - 1;
        rows[i] = nv + hole; cols[i] = part; ++i;
    }
    return i;
}

/* build_d (src/maths/wick.rs:59): X below/on diagonal, Y above */
static void build_d(double *d, int l, const double *x, const double *y, int n, const int *rows, const int *cols)
{
    for (int i = 0; i < l; ++i)
        for (int j = 0; j < l; ++j)
            d[i * l + j] = (i >= j) ? x[rows[i] * n + cols[j]] : y[rows[i] * n + cols[j]];
}

typedef struct {
    int l;
    int rows[16], cols[16];
    double det0[256], det1[256];
} prep;

/* prepare_same (prepare.rs:27) */
static void prepare_same(const same_view *w, u128 xh, u128 xp, u128 wh, u128 wp, prep *p)
{
    p->l = det_indices(w, xh, xp, wh, wp, p->rows, p->cols);
    build_d(p->det0, p->l, w->x[0], w->y[0], w->n, p->rows, p->cols);
    if (w->m > 0) build_d(p->det1, p->l, w->x[1], w->y[1], w->n, p->rows, p->cols);
}

static size_t idx4(int n, int a, int b, int c, int d) { return (((size_t)a * n + b) * n + c) * n + d; }

/* xw_hamiltonian_overlap_m0_gen_prepared (pho.rs:20304-20533).  For ranks
 * <= 4 this is the same arithmetic as the fixed-rank kernels (pho.rs:380). */
static void hs_m0(const same_view *wa, const same_view *wb, const diff_view *ab, const prep *pa,
                  const prep *pb, double exph, double enuc, double *H, double *S)
{
    int la = pa->l, lb = pb->l;
    double cofa[256], cofb[256];
    double deta = 1.0, detb = 1.0;
    int have_a = 1, have_b = 1;
    if (la > 0) { deta = adjugate_transpose(cofa, pa->det0, la); if (!isfinite(deta)) { have_a = 0; deta = 0; } }
    if (lb > 0) { detb = adjugate_transpose(cofb, pb->det0, lb); if (!isfinite(detb)) { have_b = 0; detb = 0; } }
    double same[2] = {0, 0};
    const same_view *ws[2] = {wa, wb};
    const prep *ps[2] = {pa, pb};
    const double *cofs[2] = {cofa, cofb};
    double dets[2] = {deta, detb};
    int haves[2] = {have_a, have_b};
    for (int sp = 0; sp < 2; ++sp) {
        if (!haves[sp]) continue;
        const same_view *w = ws[sp];
        const prep *p = ps[sp];
        int l = p->l, n = w->n;
        double acc = (w->f0h[0] + 0.5 * w->v0[0]) * dets[sp];
        if (l > 0) {
            const double *cof = cofs[sp];
            if (haves[1 - sp]) {
                for (int z = 0; z < l; ++z)
                    for (int e = 0; e < l; ++e) acc -= cof[e * l + z] * w->hcol0[p->cols[z] * n + p->rows[e]];
            } else {
                const double *fh = w->fh[0][0], *vv = w->v[0][0][0];
                for (int z = 0; z < l; ++z)
                    for (int e = 0; e < l; ++e)
                        acc -= cof[e * l + z] * (fh[p->cols[z] * n + p->rows[e]] + vv[p->cols[z] * n + p->rows[e]]);
            }
        }
        if (l >= 2) {
            const double *jsl = w->j[0];
            double minor[256];
            for (int eta = 0; eta < l; ++eta)
                for (int xi = eta + 1; xi < l; ++xi)
                    for (int z = 0; z < l; ++z)
                        for (int y = z + 1; y < l; ++y) {
                            int ii = 0;
                            for (int r = 0; r < l; ++r) {
                                if (r == eta || r == xi) continue;
                                int mc = 0;
                                for (int c = 0; c < l; ++c) {
                                    if (c == z || c == y) continue;
                                    minor[ii * (l - 2) + mc] = p->det0[r * l + c];
                                    ++mc;
                                }
                                ++ii;
                            }
                            double second = det_small(minor, l - 2);
                            double direct = jsl[idx4(n, p->rows[eta], p->cols[z], p->rows[xi], p->cols[y])];
                            double exch = jsl[idx4(n, p->rows[eta], p->cols[y], p->rows[xi], p->cols[z])];
                            double term = second * (direct - exch);
                            if (((eta + xi + z + y) & 1) == 0) acc += term; else acc -= term;
                        }
        }
        same[sp] = acc;
    }
    double h2ab = 0.0;
    if (have_a && have_b) {
        h2ab = ab->vab0[0][0] * deta * detb;
        int n = ab->n;
        if (la > 0 && lb > 0) {
            const double *iisl = ab->ii[0][0][0][0];
            for (int z = 0; z < la; ++z)
                for (int e = 0; e < la; ++e) {
                    double inner = 0.0;
                    for (int y = 0; y < lb; ++y)
                        for (int xi = 0; xi < lb; ++xi)
                            inner += cofb[xi * lb + y] * iisl[idx4(n, pa->rows[e], pa->cols[z], pb->rows[xi], pb->cols[y])];
                    h2ab += cofa[e * la + z] * inner;
                }
        }
    }
    double pref = exph * (wa->phase * wa->tsp) * (wb->phase * wb->tsp);
    double s = pref * deta * detb;
    *S = s;
    *H = enuc * s + pref * (same[0] * detb + same[1] * deta + h2ab);
}

/* jslot (helpers.rs:165) */
static int jslot(int mi, int mj, int mk, int ml, int *swap)
{
    int code = ((mi * 2 + mj) * 2 + mk) * 2 + ml;
    static const int slot[16] = {0, 1, 2, 3, 1, 4, 5, 6, 2, 5, 7, 8, 3, 6, 8, 9};
    static const int sw[16] = {0, 0, 0, 0, 1, 0, 0, 0, 1, 1, 0, 0, 1, 1, 1, 0};
    *swap = sw[code];
    return slot[code];
}

/* Gosper enumeration of l-bit masks with popcount m, increasing (helpers.rs:682) */
static int m_combinations(int l, int m, uint64_t *out)
{
    int cnt = 0;
    if (m > l) return 0;
    if (m == 0) { out[0] = 0; return 1; }
    if (m == l) { out[0] = (1ULL << l) - 1; return 1; }
    uint64_t limit = 1ULL << l, x = (1ULL << m) - 1;
    while (x < limit) {
        out[cnt++] = x;
        uint64_t c = x & (~x + 1), r = x + c;
        x = (((r ^ x) >> 2) / c) | r;
    }
    return cnt;
}

static int bit(uint64_t b, int k) { return (int)((b >> k) & 1); }

static void mix_columns(double *dst, const double *d0, const double *d1, int l, uint64_t cbits)
{
    for (int r = 0; r < l; ++r)
        for (int c = 0; c < l; ++c) dst[r * l + c] = bit(cbits, c) ? d1[r * l + c] : d0[r * l + c];
}

/* column_replacement_det (helpers.rs:386): Laplace expansion down column col */
static double col_repl_det(int n, const double *cof, int col, const double *newc)
{
    double v = 0.0;
    for (int r = 0; r < n; ++r) v += newc[r] * cof[r * n + col];
    return v;
}

/* Same-spin sum over distributions of an L+2 bit assignment (pho.rs:20550-20760).
 * `rd` selects, one bit per reading, which literal readings of the
 * reference's m>0 arithmetic apply (RD_ALL = the reference, 0 = the corrected
 * form, which agrees with the generalised Slater-Condon route and with the
 * brute-force determinant expansion):
 *   RD_V0    v0[1] is stored as the sum of the (0,1) and (1,0) assignments
 *            (build.rs:229-232) while the (m1,m2) enumeration visits both and
 *            adds v0[m1+m2] each time (pho.rs:20637); corrected: v0[1]/2 per visit;
 *   RD_VIDX  the one-column term reads v[m1][m2][mk] (pho.rs:20648-20651); the
 *            stored order is v[left][right][JK] (build.rs:242-259), so the
 *            replaced column's bit belongs on the right: v[m1][mk][m2];
 *   RD_JSLOT the two-column term uses jslot(m1, m2, mk, mj) (pho.rs:20690);
 *            corrected: jslot(m1, mk, m2, mj);
 *   RD_DIST  distributions with |det D_mix| <= tol are dropped from H
 *            (adjugate_transpose_generic, helpers.rs:737-748; pho.rs:20577-20586,
 *            and the different-spin sum), their overlap kept (:20735-20739);
 *   RD_MINOR first minors with |det| <= tol are skipped (minor_adjt,
 *            helpers.rs:636-660);
 *   RD_ENUC  E_nuc S is not added (the `_enuc` argument, pho.rs:20555).
 * Cofactors exist for singular matrices, so the corrected form keeps every
 * distribution and minor.
 * RD_VIDX and RD_JSLOT leave the sum unchanged: exchanging the operator
 * assignment m2 with the replaced column's assignment mk is a bijection of
 * the distributions, and the cofactors of column k do not involve column k.
 * Two planted mistakes (test use only, never part of mode 0) show that the
 * checks of those readings would see a wrong index: RD_PLANT_VT reads the
 * one-column matrix untransposed, RD_PLANT_JCOL exchanges the two column
 * labels of the J entry. */
enum { RD_V0 = 1, RD_VIDX = 2, RD_JSLOT = 4, RD_DIST = 8, RD_MINOR = 16, RD_ENUC = 32, RD_ALL = 63,
       RD_PLANT_VT = 256, RD_PLANT_JCOL = 512 };

static void same_gen(const same_view *w, const prep *p, double tol, int rd, double *s_out, double *h1_out, double *h2_out)
{
    int l = p->l, n = w->n;
    uint64_t combos[4096];
    int nc = m_combinations(l + 2, w->m, combos);
    double s = 0, h1 = 0, h2 = 0;
    double mix[256], cof[256], newc[16];
    for (int ci = 0; ci < nc; ++ci) {
        uint64_t bits = combos[ci], cbits = bits >> 2;
        int m1 = bit(bits, 0), m2 = bit(bits, 1);
        mix_columns(mix, p->det0, p->det1, l, cbits);
        double dd = 1.0;
        int ok = 1;
        if (l > 0) {
            if (rd & RD_DIST) ok = adjugate_transpose_generic(cof, mix, l, tol, &dd);
            else dd = adjugate_transpose(cof, mix, l);
        }
        if (ok) {
            if (m1 == 0 && m2 == 0) s += dd;
            if (m2 == 0) {
                double contrib = dd * w->f0h[m1];
                for (int k = 0; k < l; ++k) {
                    int mk = bit(bits, k + 2);
                    const double *fsl = w->fh[m1][mk];
                    int ck = p->cols[k];
                    double corr = 0;
                    for (int r = 0; r < l; ++r) corr += (fsl[ck * n + p->rows[r]] - mix[r * l + k]) * cof[r * l + k];
                    contrib -= dd + corr;
                }
                h1 += contrib;
            }
            double v0 = w->v0[m1 + m2];
            if (!(rd & RD_V0) && m1 + m2 == 1) v0 *= 0.5;
            double contrib = v0 * dd;
            for (int k = 0; k < l; ++k) {
                int mk = bit(bits, k + 2);
                const double *vsl = (rd & RD_VIDX) ? w->v[m1][m2][mk] : w->v[m1][mk][m2];
                int ck = p->cols[k];
                double corr = 0;
                for (int r = 0; r < l; ++r) {
                    double vv = (rd & RD_PLANT_VT) ? vsl[p->rows[r] * n + ck] : vsl[ck * n + p->rows[r]];
                    corr += (vv - mix[r * l + k]) * cof[r * l + k];
                }
                contrib -= 2.0 * (dd + corr);
            }
            if (l >= 2) {
                double minor[256], mcof[256];
                for (int i = 0; i < l; ++i)
                    for (int j = 0; j < l; ++j) {
                        double ph = ((i + j) & 1) ? -1.0 : 1.0;
                        int ri = p->rows[i], cj = p->cols[j];
                        int mj = bit(bits, j + 2);
                        /* minor_adjt (helpers.rs:636): skip |det minor| <= tol */
                        int lm1 = l - 1;
                        build_minor(minor, mix, l, i, j);
                        double dm;
                        if (!(rd & RD_MINOR)) {
                            dm = adjugate_transpose(mcof, minor, lm1);
                        } else if (l <= 4) {
                            dm = adjugate_transpose(mcof, minor, lm1);
                            if (!(fabs(dm) > tol)) continue;
                        } else {
                            if (!adjugate_transpose_generic(mcof, minor, lm1, tol, &dm)) continue;
                        }
                        for (int k2 = 0; k2 < lm1; ++k2) {
                            int kf = k2 < j ? k2 : k2 + 1;
                            int mk = bit(bits, kf + 2);
                            int swap;
                            int slot = (rd & RD_JSLOT) ? jslot(m1, m2, mk, mj, &swap) : jslot(m1, mk, m2, mj, &swap);
                            const double *jsl = w->j[slot];
                            for (int r = 0; r < lm1; ++r) {
                                int rf = r < i ? r : r + 1;
                                int rr = p->rows[rf], cc = p->cols[kf];
                                if (rd & RD_PLANT_JCOL) { int t = cc; cc = cj; cj = t; }
                                newc[r] = swap ? jsl[idx4(n, rr, cc, ri, cj)] : jsl[idx4(n, ri, cj, rr, cc)];
                                if (rd & RD_PLANT_JCOL) cj = p->cols[j];
                            }
                            contrib += ph * col_repl_det(lm1, mcof, k2, newc);
                        }
                    }
            }
            h2 += contrib;
        } else if (m1 == 0 && m2 == 0) {
            s += det_small(mix, l);
        }
    }
    *s_out = s; *h1_out = h1; *h2_out = h2;
}

/* xw_hamiltonian_overlap_gen_prepared (pho.rs:20550-20966); omits E_nuc as the reference does. */
static void hs_gen(const same_view *wa, const same_view *wb, const diff_view *ab, const prep *pa,
                   const prep *pb, double exph, double tol, int rd, double *H, double *S)
{
    int la = pa->l, lb = pb->l, n = ab->n;
    double sa, h1a, h2aa, sb, h1b, h2bb;
    same_gen(wa, pa, tol, rd, &sa, &h1a, &h2aa);
    same_gen(wb, pb, tol, rd, &sb, &h1b, &h2bb);
    double h2ab = 0.0;
    uint64_t ca[4096], cb[4096];
    int nca = m_combinations(la + 1, wa->m, ca);
    int ncb = m_combinations(lb + 1, wb->m, cb);
    double mixa[256], mixb[256], cofa[256], cofb[256], newc[16];
    for (int ia = 0; ia < nca; ++ia) {
        uint64_t ba = ca[ia];
        mix_columns(mixa, pa->det0, pa->det1, la, ba >> 1);
        double da = 1.0;
        if (la > 0) {
            if (rd & RD_DIST) { if (!adjugate_transpose_generic(cofa, mixa, la, tol, &da)) continue; }
            else da = adjugate_transpose(cofa, mixa, la);
        }
        for (int ib = 0; ib < ncb; ++ib) {
            uint64_t bb = cb[ib];
            mix_columns(mixb, pb->det0, pb->det1, lb, bb >> 1);
            double db = 1.0;
            if (lb > 0) {
                if (rd & RD_DIST) { if (!adjugate_transpose_generic(cofb, mixb, lb, tol, &db)) continue; }
                else db = adjugate_transpose(cofb, mixb, lb);
            }
            int ma0 = bit(ba, 0), mb0 = bit(bb, 0);
            double contrib = ab->vab0[ma0][mb0] * da * db;
            for (int k = 0; k < la; ++k) {
                int mak = bit(ba, k + 1);
                const double *vsl = ab->vab[ma0][mb0][mak];
                int ck = pa->cols[k];
                for (int r = 0; r < la; ++r) newc[r] = vsl[ck * n + pa->rows[r]];
                contrib -= col_repl_det(la, cofa, k, newc) * db;
            }
            for (int k = 0; k < lb; ++k) {
                int mbk = bit(bb, k + 1);
                const double *vsl = ab->vba[mb0][ma0][mbk];
                int ck = pb->cols[k];
                for (int r = 0; r < lb; ++r) newc[r] = vsl[ck * n + pb->rows[r]];
                contrib -= col_repl_det(lb, cofb, k, newc) * da;
            }
            for (int i = 0; i < la; ++i) {
                int ra = pa->rows[i];
                for (int j = 0; j < la; ++j) {
                    int cca = pa->cols[j];
                    double cofij = cofa[i * la + j];
                    int ma1 = bit(ba, j + 1);
                    for (int k = 0; k < lb; ++k) {
                        int mbk = bit(bb, k + 1);
                        const double *iisl = ab->ii[ma0][ma1][mb0][mbk];
                        for (int r = 0; r < lb; ++r) newc[r] = iisl[idx4(n, ra, cca, pb->rows[r], pb->cols[k])];
                        contrib += cofij * col_repl_det(lb, cofb, k, newc);
                    }
                }
            }
            h2ab += contrib;
        }
    }
    double pref = exph * (wa->phase * wa->tsp) * (wb->phase * wb->tsp);
    *S = pref * sa * sb;
    *H = pref * (h1a * sb + h1b * sa + 0.5 * h2aa * sb + 0.5 * h2bb * sa + h2ab);
}

/* Excitation masks: ex[0..7] = alpha holes lo,hi, alpha parts lo,hi, beta holes lo,hi, beta parts lo,hi */
static u128 mk128(const uint64_t *p) { return ((u128)p[1] << 64) | p[0]; }

/* xw_hamiltonian_overlap_prepared (pho.rs:53): one determinant pair on one
 * ordered reference pair (slab at the pair base), with the m>0 readings `rd`
 * (RD_* above; m = 0 pairs do not depend on them). */
int ow_hs_readings(const double *slab, const long long *off, const double *meta, int nm,
                   const uint64_t *xex, const uint64_t *wex, double exph, double enuc, double tol,
                   int rd, double *H, double *S)
{
    same_view sv[2];
    diff_view dv;
    make_views(slab, off, meta, nm, sv, &dv);
    prep pa, pb;
    prepare_same(&sv[0], mk128(xex + 0), mk128(xex + 2), mk128(wex + 0), mk128(wex + 2), &pa);
    prepare_same(&sv[1], mk128(xex + 4), mk128(xex + 6), mk128(wex + 4), mk128(wex + 6), &pb);
    if (sv[0].m == 0 && sv[1].m == 0) hs_m0(&sv[0], &sv[1], &dv, &pa, &pb, exph, enuc, H, S);
    else {
        hs_gen(&sv[0], &sv[1], &dv, &pa, &pb, exph, tol, rd, H, S);
        if (!(rd & RD_ENUC)) *H += enuc * *S;
    }
    return 0;
}

/* mode 0: the reference arithmetic (every reading); mode 1: corrected. */
int ow_hs(const double *slab, const long long *off, const double *meta, int nm,
          const uint64_t *xex, const uint64_t *wex, double exph, double enuc, double tol,
          int mode, double *H, double *S)
{
    return ow_hs_readings(slab, off, meta, nm, xex, wex, exph, enuc, tol, mode == 0 ? RD_ALL : 0, H, S);
}

/* construct_f_scalar (build.rs:564-617) as applied by update_wicks_fock
 * (src/noci/wicks.rs:290-330): generalised-Fock scalar and one-column
 * intermediates of one ordered pair from the spin Fock matrices. */
int ow_update_fock_pair(int nb, int nm, const double *S, const double *fa, const double *fb,
                        const double *lca, const double *lcb, const uint64_t *loa, const uint64_t *lob,
                        const double *gca, const double *gcb, const uint64_t *goa, const uint64_t *gob,
                        double tol, const long long *off, double *slab, double *meta)
{
    u128 la = mk128(loa), lb = mk128(lob), ga = mk128(goa), gb = mk128(gob);
    const double *F[2] = {fa, fb};
    const double *lc[2] = {lca, lcb}, *gc[2] = {gca, gcb};
    u128 lo[2] = {la, lb}, go[2] = {ga, gb};
    double *tmp = (double *)calloc((size_t)nb * nm, sizeof(double));
    double *blk = (double *)calloc((size_t)nm * nm, sizeof(double));
    for (int sp = 0; sp < 2; ++sp) {
        spin_core c;
        spin_core_build(nb, nm, S, lc[sp], lo[sp], gc[sp], go[sp], tol, &c);
        const long long *o = off + (sp ? O_BB : O_AA);
        double *mt = meta + (sp ? M_BB : M_AA);
        mt[M_F0F + 0] = einsum_ba_ab(nb, c.mao[0], F[sp]);
        mt[M_F0F + 1] = einsum_ba_ab(nb, c.mao[1], F[sp]);
        for (int mi = 0; mi < 2; ++mi)
            for (int mj = 0; mj < 2; ++mj) {
                mm(nb, nm, nb, F[sp], nb, 0, c.xc[mj], nm, 0, tmp, nm);
                mm(nm, nm, nb, c.cx[mi], nm, 1, tmp, nm, 0, blk, nm);
                write2t(slab, o[O_FF + mi * 2 + mj], blk, nm);
            }
        free_core(&c);
    }
    free(tmp); free(blk);
    return 0;
}

/* xw_overlap (eval/overlap.rs:42-66): phase * tsp * sum over L-bit
 * distributions with popcount m of det D_mix (no tolerance). */
static double xw_overlap_o(const same_view *w, const prep *p)
{
    int l = p->l;
    if (w->m > l) return 0.0;
    uint64_t combos[4096];
    int nc = m_combinations(l, w->m, combos);
    double mix[256], acc = 0.0;
    for (int ci = 0; ci < nc; ++ci) {
        mix_columns(mix, p->det0, p->det1, l, combos[ci]);
        acc += det_small(mix, l);
    }
    return w->phase * w->tsp * acc;
}

/* xw_f (eval/onebody.rs:40-496): phase * tsp * sum over (L+1)-bit
 * distributions (bit 0 = operator) of det f0f[mi] - sum_b det(D, col b ->
 * ff[mi][m_b]); m > 0 distributions with |det| <= tol are dropped in
 * mode 0 (get_det_adjt_same, helpers.rs:521). */
static double xw_f_o(const same_view *w, const prep *p, double tol, int mode)
{
    int l = p->l, n = w->n;
    uint64_t combos[4096];
    int nc = m_combinations(l + 1, w->m, combos);
    double mix[256], cof[256], acc = 0.0;
    for (int ci = 0; ci < nc; ++ci) {
        uint64_t bits = combos[ci];
        int mi = bit(bits, 0);
        mix_columns(mix, p->det0, p->det1, l, bits >> 1);
        double dd = 1.0;
        if (l > 0) {
            dd = adjugate_transpose(cof, mix, l);
            if (w->m > 0 && mode == 0 && !(fabs(dd) > tol)) continue;
        }
        double contrib = dd * w->f0f[mi];
        for (int b = 0; b < l; ++b) {
            int mj = bit(bits, b + 1);
            const double *fsl = w->ff[mi][mj];
            int cb = p->cols[b];
            double corr = 0.0;
            for (int r = 0; r < l; ++r) corr += (fsl[cb * n + p->rows[r]] - mix[r * l + b]) * cof[r * l + b];
            contrib -= dd + corr;
        }
        acc += contrib;
    }
    return w->phase * w->tsp * acc;
}

/* calculate_f_pair_wicks (src/noci/fock.rs:175-216) and
 * calculate_s_pair_wicks (src/noci/overlap.rs:106-141) for one pair;
 * M = F - e0 S is calculate_m_pair_wicks (src/noci/m.rs:160-197). */
int ow_fs(const double *slab, const long long *off, const double *meta, int nm,
          const uint64_t *xex, const uint64_t *wex, double pha, double phb, double tol, int mode,
          double *F, double *Sout)
{
    same_view sv[2];
    diff_view dv;
    make_views(slab, off, meta, nm, sv, &dv);
    prep pa, pb;
    prepare_same(&sv[0], mk128(xex + 0), mk128(xex + 2), mk128(wex + 0), mk128(wex + 2), &pa);
    prepare_same(&sv[1], mk128(xex + 4), mk128(xex + 6), mk128(wex + 4), mk128(wex + 6), &pb);
    double sa = pha * xw_overlap_o(&sv[0], &pa);
    double sb = phb * xw_overlap_o(&sv[1], &pb);
    double f = 0.0;
    if (sb != 0.0) f += pha * xw_f_o(&sv[0], &pa, tol, mode) * sb;
    if (sa != 0.0) f += phb * xw_f_o(&sv[1], &pb, tol, mode) * sa;
    *F = f;
    *Sout = sa * sb;
    return 0;
}

/* One spin's factor of a determinant pair: {S, F} = ph * (xw_overlap, xw_f)
 * of that spin (the generic path of build_spin_one_body_factor_row,
 * src/noci/factorise/onebody.rs:1548-1568: phase * xw_f_overlap_prepared).
 * xh, xp, wh, wp: hole / particle masks {lo, hi} of the bra (x) and ket (w). */
int ow_fs_spin(const double *slab, const long long *off, const double *meta, int nm, int spin,
               const uint64_t *xh, const uint64_t *xp, const uint64_t *wh, const uint64_t *wp, double ph,
               double tol, int mode, double *F, double *Sout)
{
    same_view sv[2];
    diff_view dv;
    make_views(slab, off, meta, nm, sv, &dv);
    prep p;
    prepare_same(&sv[spin], mk128(xh), mk128(xp), mk128(wh), mk128(wp), &p);
    *Sout = ph * xw_overlap_o(&sv[spin], &p);
    *F = ph * xw_f_o(&sv[spin], &p, tol, mode);
    return 0;
}

/* Evaluate a block of determinant pairs on one ordered reference pair:
 * xex/wex: per pair 8 uint64 masks; exph: per pair excitation phase. */
int ow_hs_many(const double *slab, const long long *off, const double *meta, int nm, long long npair,
               const uint64_t *xex, const uint64_t *wex, const double *exph, double enuc, double tol, int mode,
               double *H, double *S)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (long long q = 0; q < npair; ++q)
        ow_hs(slab, off, meta, nm, xex + 8 * q, wex + 8 * q, exph[q], enuc, tol, mode, H + q, S + q);
    return 0;
}

int ow_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* excitation_phase_bits (src/basis/excitation.rs:79-116) */
double ow_excitation_phase(const uint64_t *occ, const uint64_t *holes, const uint64_t *parts)
{
    u128 o = mk128(occ), h = mk128(holes), p = mk128(parts);
    int odd = 0;
    while (h) {
        int i = 127 - (((uint64_t)(h >> 64)) ? __builtin_clzll((uint64_t)(h >> 64)) : 64 + __builtin_clzll((uint64_t)h));
        odd ^= (i > 0) ? (pop128(o & (((u128)1 << i) - 1)) & 1) : 0;
        o &= ~((u128)1 << i);
        h &= ~((u128)1 << i);
    }
    while (p) {
        int a = ctz128(p);
        odd ^= (a > 0) ? (pop128(o & (((u128)1 << a) - 1)) & 1) : 0;
        p &= p - 1;
        if (p) o |= (u128)1 << a;
    }
    return odd ? -1.0 : 1.0;
}

/* ------------------------------------------------------------------------- */
/* Generalised Slater-Condon route (src/noci/naive.rs, src/noci/hs.rs:365)    */
/* ------------------------------------------------------------------------- */
typedef struct {
    double s, s_red, phase;
    int nz, zeros[64];
    double *w, *p_i, *p_j; /* nb x nb */
} naive_pair;

static void naive_build(int nb, int nm, const double *S, const double *lc, u128 lo, const double *gc, u128 go,
                        double tol, naive_pair *pp)
{
    double *lco = (double *)calloc((size_t)nb * 64, sizeof(double));
    double *gco = (double *)calloc((size_t)nb * 64, sizeof(double));
    int nocc = occ_cols(nb, nm, lc, lo, lco);
    occ_cols(nb, nm, gc, go, gco);
    double *tmp = (double *)calloc((size_t)nb * 64, sizeof(double));
    double sij[64 * 64], U[64 * 64], V[64 * 64], sv[64];
    mm(nb, nocc, nb, S, nb, 0, gco, 64, 0, tmp, 64);
    mm(nocc, nocc, nb, lco, 64, 1, tmp, 64, 0, sij, nocc);
    ow_svd(nocc, sij, U, sv, V);
    pp->phase = det_small(U, nocc) * det_small(V, nocc);
    double prod = 1.0;
    pp->s_red = 1.0; pp->nz = 0;
    for (int k = 0; k < nocc; ++k) {
        prod *= sv[k];
        if (fabs(sv[k]) > tol) pp->s_red *= sv[k];
        else pp->zeros[pp->nz++] = k;
    }
    double *lt = (double *)calloc((size_t)nb * 64, sizeof(double));
    double *gt = (double *)calloc((size_t)nb * 64, sizeof(double));
    mm(nb, nocc, nocc, lco, 64, 0, U, nocc, 0, lt, 64);
    mm(nb, nocc, nocc, gco, 64, 0, V, nocc, 0, gt, 64);
    size_t nb2 = (size_t)nb * nb;
    pp->w = pp->p_i = pp->p_j = NULL;
    if (pp->nz <= 1) {
        pp->w = (double *)calloc(nb2, sizeof(double));
        for (int mu = 0; mu < nb; ++mu)
            for (int nu = 0; nu < nb; ++nu) {
                double acc = 0;
                for (int k = 0; k < nocc; ++k) {
                    double wk = fabs(sv[k]) > tol ? 1.0 / sv[k] : 0.0;
                    acc += (gt[mu * 64 + k] * wk) * lt[nu * 64 + k];
                }
                pp->w[mu * nb + nu] = acc;
            }
    }
    for (int z = 0; z < pp->nz && z < 2; ++z) {
        double *P = (double *)calloc(nb2, sizeof(double));
        int k = pp->zeros[z];
        for (int mu = 0; mu < nb; ++mu)
            for (int nu = 0; nu < nb; ++nu) P[mu * nb + nu] = gt[mu * 64 + k] * lt[nu * 64 + k];
        if (z == 0) pp->p_i = P; else pp->p_j = P;
    }
    if (pp->nz > 2) { free(pp->p_i); free(pp->p_j); pp->p_i = pp->p_j = NULL; }
    pp->s = pp->nz == 0 ? pp->phase * prod : 0.0;
    free(lco); free(gco); free(tmp); free(lt); free(gt);
}

/* einsum_ba_abcd_cd (src/maths/einsum.rs:318): sum g[b,a] t[a,b,c,d] h[c,d] */
static double einsum4(int n, const double *g, const double *t, const double *h)
{
    double acc = 0;
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            double gba = g[b * n + a];
            if (gba == 0.0) continue;
            const double *tb = t + ((size_t)a * n + b) * n * n;
            double dot = 0;
            for (size_t cd = 0; cd < (size_t)n * n; ++cd) dot += tb[cd] * h[cd];
            acc += gba * dot;
        }
    return acc;
}

int ow_hs_naive(int nb, int nm, const double *S, const double *h, const double *eri_coul, const double *eri_asym,
                const double *lca, const double *lcb, const uint64_t *loa, const uint64_t *lob,
                const double *gca, const double *gcb, const uint64_t *goa, const uint64_t *gob,
                double enuc, double tol, double *H, double *Sout)
{
    naive_pair pa, pb;
    naive_build(nb, nm, S, lca, mk128(loa), gca, mk128(goa), tol, &pa);
    naive_build(nb, nm, S, lcb, mk128(lob), gcb, mk128(gob), tol, &pb);
    double s = pa.s * pb.s;
    double hnuc = (pa.nz == 0 && pb.nz == 0) ? enuc * s : 0.0;
    naive_pair *ps[2] = {&pa, &pb};
    double h1[2], h2s[2];
    for (int k = 0; k < 2; ++k) {
        naive_pair *p = ps[k];
        double fac = p->phase * p->s_red;
        const double *rho = p->nz == 0 ? p->w : (p->nz == 1 ? p->p_i : NULL);
        h1[k] = rho ? fac * einsum_ba_ab(nb, rho, h) : 0.0;
        if (p->nz == 0) h2s[k] = 0.5 * fac * einsum4(nb, p->w, eri_asym, p->w);
        else if (p->nz == 1) h2s[k] = fac * einsum4(nb, p->p_i, eri_asym, p->w);
        else if (p->nz == 2) h2s[k] = fac * einsum4(nb, p->p_i, eri_asym, p->p_j);
        else h2s[k] = 0.0;
    }
    double h2ab = 0.0;
    if (pa.nz <= 1 && pb.nz <= 1) {
        const double *ra = pa.nz == 0 ? pa.w : pa.p_i;
        const double *rb = pb.nz == 0 ? pb.w : pb.p_i;
        h2ab = (pa.phase * pa.s_red) * (pb.phase * pb.s_red) * einsum4(nb, ra, eri_coul, rb);
    }
    *H = hnuc + (pb.s * h1[0] + pa.s * h1[1]) + (pb.s * h2s[0] + pa.s * h2s[1] + h2ab);
    *Sout = s;
    free(pa.w); free(pa.p_i); free(pa.p_j); free(pb.w); free(pb.p_i); free(pb.p_j);
    return 0;
}

int ow_meta_len(void) { return META_LEN; }
int ow_offsets_len(void) { return O_SIZE + 1; }
