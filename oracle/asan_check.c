/*
 * asan_check.c -- sanitizer driver for the CPU oracle (test infrastructure only).
 *
 * Built with AddressSanitizer and UndefinedBehaviorSanitizer together with
 * nwick_oracle.c (`make -C oracle asan`), it runs every exported oracle
 * routine on a small seeded system: zero counts, the per-pair build of every
 * ordered reference pair (one pair with a zero-overlap orbital, m_alpha = 1),
 * the Fock refresh, and H,S / F,S / per-spin factors of every pair of
 * reference, single and double determinants in both arithmetic modes and
 * with every reading, the batched evaluator, and the generalised
 * Slater-Condon route.  Besides running clean under the sanitizers it checks
 * that the corrected Wick elements equal the Slater-Condon ones (1e-9) and
 * that every value is finite.  Prints "asan_check ok" and exits 0 on success.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int ow_svd(int n, const double *A, double *U, double *s, double *V);
long long ow_pair_offsets(int ma, int mb, int nmo, long long *off);
int ow_count_zero(int nb, int nm, const double *S, const double *lc, const uint64_t *lo, const double *gc,
                  const uint64_t *go, double tol, double *sig_out);
int ow_build_pair(int nb, int nm, const double *S, const double *h, const double *eri, const double *lca,
                  const double *lcb, const uint64_t *loa, const uint64_t *lob, const double *gca, const double *gcb,
                  const uint64_t *goa, const uint64_t *gob, double tol, const long long *off, double *slab,
                  double *meta);
int ow_hs_readings(const double *slab, const long long *off, const double *meta, int nm, const uint64_t *xex,
                   const uint64_t *wex, double exph, double enuc, double tol, int rd, double *H, double *S);
int ow_hs(const double *slab, const long long *off, const double *meta, int nm, const uint64_t *xex,
          const uint64_t *wex, double exph, double enuc, double tol, int mode, double *H, double *S);
int ow_update_fock_pair(int nb, int nm, const double *S, const double *fa, const double *fb, const double *lca,
                        const double *lcb, const uint64_t *loa, const uint64_t *lob, const double *gca,
                        const double *gcb, const uint64_t *goa, const uint64_t *gob, double tol,
                        const long long *off, double *slab, double *meta);
int ow_fs(const double *slab, const long long *off, const double *meta, int nm, const uint64_t *xex,
          const uint64_t *wex, double pha, double phb, double tol, int mode, double *F, double *Sout);
int ow_fs_spin(const double *slab, const long long *off, const double *meta, int nm, int spin, const uint64_t *xh,
               const uint64_t *xp, const uint64_t *wh, const uint64_t *wp, double ph, double tol, int mode,
               double *F, double *Sout);
int ow_hs_many(const double *slab, const long long *off, const double *meta, int nm, long long npair,
               const uint64_t *xex, const uint64_t *wex, const double *exph, double enuc, double tol, int mode,
               double *H, double *S);
double ow_excitation_phase(const uint64_t *occ, const uint64_t *holes, const uint64_t *parts);
int ow_hs_naive(int nb, int nm, const double *S, const double *h, const double *eri_coul, const double *eri_asym,
                const double *lca, const double *lcb, const uint64_t *loa, const uint64_t *lob, const double *gca,
                const double *gcb, const uint64_t *goa, const uint64_t *gob, double enuc, double tol, double *H,
                double *Sout);
int ow_offsets_len(void);
int ow_meta_len(void);

enum { NB = 8, NA = 3, NBETA = 3, NREF = 3, NAUX = 12 };
static const double TOL = 1e-8, ENUC = 0.7;

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double urand(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) / 9007199254740992.0 - 0.5;
}

/* Orthonormal columns of a random n x n matrix (modified Gram-Schmidt), row-major. */
static void random_orthogonal(int n, double *C)
{
    for (int i = 0; i < n * n; ++i) C[i] = urand();
    for (int j = 0; j < n; ++j) {
        for (int k = 0; k < j; ++k) {
            double d = 0.0;
            for (int i = 0; i < n; ++i) d += C[i * n + j] * C[i * n + k];
            for (int i = 0; i < n; ++i) C[i * n + j] -= d * C[i * n + k];
        }
        double nr = 0.0;
        for (int i = 0; i < n; ++i) nr += C[i * n + j] * C[i * n + j];
        nr = sqrt(nr);
        for (int i = 0; i < n; ++i) C[i * n + j] /= nr;
    }
}

/* One determinant: parent, per spin occupation, holes and particles (bit masks). */
typedef struct { int parent; uint64_t occ[2], holes[2], parts[2]; double ph[2]; } det_t;

static int fail_count = 0;
static void check(int ok, const char *what, int x, int w, int i, int j)
{
    if (!ok) {
        if (fail_count < 10) fprintf(stderr, "check failed: %s pair (%d,%d) dets %d %d\n", what, x, w, i, j);
        ++fail_count;
    }
}

int main(void)
{
    const int nb = NB, nm = NB;
    const size_t n2 = (size_t)nb * nb, n4 = n2 * n2;
    double *S = calloc(n2, sizeof(double)), *h = calloc(n2, sizeof(double));
    double *eri = calloc(n4, sizeof(double)), *asym = calloc(n4, sizeof(double));
    double *chol = calloc((size_t)NAUX * n2, sizeof(double));
    for (int i = 0; i < nb; ++i) S[i * nb + i] = 1.0;
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j <= i; ++j) h[i * nb + j] = h[j * nb + i] = 0.1 * urand() + (i == j ? -2.0 + 2.0 * i / nb : 0.0);
    for (int k = 0; k < NAUX; ++k)
        for (int i = 0; i < nb; ++i)
            for (int j = 0; j <= i; ++j) chol[k * n2 + i * nb + j] = chol[k * n2 + j * nb + i] = 0.1 * urand();
    for (size_t p = 0; p < n2; ++p)
        for (size_t q = 0; q < n2; ++q) {
            double v = 0.0;
            for (int k = 0; k < NAUX; ++k) v += chol[k * n2 + p] * chol[k * n2 + q];
            eri[p * n2 + q] = v;
        }
    for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nb; ++c)
                for (int d = 0; d < nb; ++d)
                    asym[((size_t)(a * nb + b) * nb + c) * nb + d] =
                        eri[((size_t)(a * nb + b) * nb + c) * nb + d] - eri[((size_t)(a * nb + c) * nb + b) * nb + d];
    /* references: random orthogonal coefficients per spin; reference 2 takes
     * reference 0's alpha orbitals with its first occupied orbital exchanged
     * for a virtual one, so pair (0, 2) has m_alpha = 1 */
    double *C[NREF][2];
    for (int r = 0; r < NREF; ++r)
        for (int sp = 0; sp < 2; ++sp) {
            C[r][sp] = calloc(n2, sizeof(double));
            random_orthogonal(nb, C[r][sp]);
        }
    memcpy(C[2][0], C[0][0], n2 * sizeof(double));
    for (int i = 0; i < nb; ++i) {
        double t = C[2][0][i * nb + 0];
        C[2][0][i * nb + 0] = C[2][0][i * nb + NA];
        C[2][0][i * nb + NA] = t;
    }
    const int nocc[2] = {NA, NBETA};
    uint64_t refocc[2][2] = {{(1ull << NA) - 1, 0}, {(1ull << NBETA) - 1, 0}};
    /* determinants: every reference, its singles and its doubles of each spin class (first few) */
    det_t *dets = calloc(4096, sizeof(det_t));
    int nd = 0;
    for (int r = 0; r < NREF; ++r) {
        dets[nd++] = (det_t){r, {refocc[0][0], refocc[1][0]}, {0, 0}, {0, 0}, {1.0, 1.0}};
        for (int sp = 0; sp < 2; ++sp)
            for (int i = 0; i < nocc[sp]; ++i)
                for (int a = nocc[sp]; a < nm; ++a) {
                    det_t d = {r, {refocc[0][0], refocc[1][0]}, {0, 0}, {0, 0}, {1.0, 1.0}};
                    d.holes[sp] = 1ull << i;
                    d.parts[sp] = 1ull << a;
                    dets[nd++] = d;
                }
        for (int sp = 0; sp < 2; ++sp)
            for (int i = 0; i < nocc[sp]; ++i)
                for (int j = i + 1; j < nocc[sp]; ++j)
                    for (int a = nocc[sp]; a < nm && a < nocc[sp] + 3; ++a)
                        for (int b = a + 1; b < nm && b < nocc[sp] + 4; ++b) {
                            det_t d = {r, {refocc[0][0], refocc[1][0]}, {0, 0}, {0, 0}, {1.0, 1.0}};
                            d.holes[sp] = (1ull << i) | (1ull << j);
                            d.parts[sp] = (1ull << a) | (1ull << b);
                            dets[nd++] = d;
                        }
        for (int i = 0; i < NA; ++i)
            for (int a = NA; a < nm && a < NA + 2; ++a) {
                det_t d = {r, {refocc[0][0], refocc[1][0]}, {1ull << i, 1ull << (NBETA - 1)},
                           {1ull << a, 1ull << (nm - 1)}, {1.0, 1.0}};
                dets[nd++] = d;
            }
    }
    for (int k = 0; k < nd; ++k)
        for (int sp = 0; sp < 2; ++sp) {
            uint64_t o[2] = {refocc[sp][0], 0}, hm[2] = {dets[k].holes[sp], 0}, pm[2] = {dets[k].parts[sp], 0};
            dets[k].ph[sp] = ow_excitation_phase(o, hm, pm);
            dets[k].occ[sp] = (refocc[sp][0] & ~dets[k].holes[sp]) | dets[k].parts[sp];
        }
    const int olen = ow_offsets_len(), mlen = ow_meta_len();
    double fa[NB * NB], fb[NB * NB];
    for (int i = 0; i < nb; ++i)
        for (int j = 0; j <= i; ++j) {
            fa[i * nb + j] = fa[j * nb + i] = urand();
            fb[i * nb + j] = fb[j * nb + i] = urand();
        }
    int npairs_m1 = 0;
    long long nelem = 0;
    for (int x = 0; x < NREF; ++x)
        for (int w = 0; w < NREF; ++w) {
            double sig[64];
            const int ma = ow_count_zero(nb, nm, S, C[x][0], refocc[0], C[w][0], refocc[0], TOL, sig);
            const int mb = ow_count_zero(nb, nm, S, C[x][1], refocc[1], C[w][1], refocc[1], TOL, sig);
            npairs_m1 += (ma > 0 || mb > 0);
            long long *off = calloc((size_t)olen, sizeof(long long));
            const long long len = ow_pair_offsets(ma, mb, nm, off);
            double *slab = calloc((size_t)len, sizeof(double)), *meta = calloc((size_t)mlen, sizeof(double));
            ow_build_pair(nb, nm, S, h, eri, C[x][0], C[x][1], refocc[0], refocc[1], C[w][0], C[w][1], refocc[0],
                          refocc[1], TOL, off, slab, meta);
            ow_update_fock_pair(nb, nm, S, fa, fb, C[x][0], C[x][1], refocc[0], refocc[1], C[w][0], C[w][1],
                                refocc[0], refocc[1], TOL, off, slab, meta);
            for (long long k = 0; k < len; ++k) check(isfinite(slab[k]), "finite slab", x, w, -1, -1);
            /* every determinant pair of parents (x, w) */
            long long nq = 0;
            uint64_t *xex = calloc((size_t)nd * nd * 8, sizeof(uint64_t)), *wex = calloc((size_t)nd * nd * 8, sizeof(uint64_t));
            double *exph = calloc((size_t)nd * nd, sizeof(double));
            for (int i = 0; i < nd; ++i) {
                if (dets[i].parent != x) continue;
                for (int j = 0; j < nd; ++j) {
                    if (dets[j].parent != w) continue;
                    const det_t *a = &dets[i], *b = &dets[j];
                    uint64_t xe[8] = {a->holes[0], 0, a->parts[0], 0, a->holes[1], 0, a->parts[1], 0};
                    uint64_t we[8] = {b->holes[0], 0, b->parts[0], 0, b->holes[1], 0, b->parts[1], 0};
                    const double ph = a->ph[0] * b->ph[0] * a->ph[1] * b->ph[1];
                    double H0, S0, H1, S1, Hr, Sr, F, Sf, Fs, Ss, Hn, Sn;
                    ow_hs(slab, off, meta, nm, xe, we, ph, ENUC, TOL, 0, &H0, &S0);
                    ow_hs(slab, off, meta, nm, xe, we, ph, ENUC, TOL, 1, &H1, &S1);
                    for (int rd = 1; rd < 64; rd <<= 1) {
                        ow_hs_readings(slab, off, meta, nm, xe, we, ph, ENUC, TOL, rd, &Hr, &Sr);
                        check(isfinite(Hr) && isfinite(Sr), "finite reading", x, w, i, j);
                    }
                    ow_fs(slab, off, meta, nm, xe, we, a->ph[0] * b->ph[0], a->ph[1] * b->ph[1], TOL, 1, &F, &Sf);
                    ow_fs(slab, off, meta, nm, xe, we, a->ph[0] * b->ph[0], a->ph[1] * b->ph[1], TOL, 0, &F, &Sf);
                    for (int sp = 0; sp < 2; ++sp) {
                        uint64_t xh[2] = {a->holes[sp], 0}, xp[2] = {a->parts[sp], 0};
                        uint64_t wh[2] = {b->holes[sp], 0}, wp[2] = {b->parts[sp], 0};
                        ow_fs_spin(slab, off, meta, nm, sp, xh, xp, wh, wp, a->ph[sp] * b->ph[sp], TOL, 1, &Fs, &Ss);
                        check(isfinite(Fs) && isfinite(Ss), "finite spin factor", x, w, i, j);
                    }
                    uint64_t oa[2] = {a->occ[0], 0}, ob[2] = {a->occ[1], 0}, ga[2] = {b->occ[0], 0},
                             gb[2] = {b->occ[1], 0};
                    ow_hs_naive(nb, nm, S, h, eri, asym, C[x][0], C[x][1], oa, ob, C[w][0], C[w][1], ga, gb, ENUC,
                                TOL, &Hn, &Sn);
                    check(isfinite(H0) && isfinite(S0) && isfinite(F) && isfinite(Sf), "finite element", x, w, i, j);
                    check(fabs(H1 - Hn) < 1e-9 && fabs(S1 - Sn) < 1e-9, "corrected Wick = Slater-Condon", x, w, i, j);
                    check(fabs(Sf - S1) < 1e-12, "F,S overlap = H,S overlap", x, w, i, j);
                    memcpy(xex + 8 * nq, xe, sizeof(xe));
                    memcpy(wex + 8 * nq, we, sizeof(we));
                    exph[nq++] = ph;
                    ++nelem;
                }
            }
            double *Hm = calloc((size_t)nq + 1, sizeof(double)), *Sm = calloc((size_t)nq + 1, sizeof(double));
            ow_hs_many(slab, off, meta, nm, nq, xex, wex, exph, ENUC, TOL, 0, Hm, Sm);
            for (long long q = 0; q < nq; ++q) check(isfinite(Hm[q]) && isfinite(Sm[q]), "finite batched", x, w, -1, -1);
            free(Hm); free(Sm); free(xex); free(wex); free(exph);
            free(slab); free(meta); free(off);
        }
    double U[64], sv[8], V[64], A[64];
    for (int i = 0; i < 64; ++i) A[i] = urand();
    ow_svd(8, A, U, sv, V);
    for (int i = 0; i < 8; ++i) check(isfinite(sv[i]), "finite singular value", -1, -1, -1, -1);
    check(npairs_m1 >= 2, "m > 0 pairs present", -1, -1, -1, -1);
    for (int r = 0; r < NREF; ++r) { free(C[r][0]); free(C[r][1]); }
    free(dets); free(S); free(h); free(eri); free(asym); free(chol);
    if (fail_count) {
        fprintf(stderr, "asan_check: %d checks failed\n", fail_count);
        return 1;
    }
    printf("asan_check ok: %d determinants, %lld elements, %d reference pairs with m > 0\n", nd, nelem, npairs_m1);
    return 0;
}
