// Kernel launchers shared between the build, evaluation and API units.
#pragma once
#include "nw_common.h"

namespace nw {

struct GemmDesc { const double *A; const double *B; double *C; };
struct LinDesc { const double *a, *b, *c; double *out; double ca, cb, cc; };  // out = ca a + cb b + cc c
struct TraceDesc { const double *g; const double *h; double *out; double scale; };

void launch_gemm(hipStream_t st, int ta, int tb, int M, int N, int K, int lda, int ldb, int ldc,
                 double alpha, double beta, const GemmDesc *d_descs, int nbatch);
void launch_svd(hipStream_t st, const double *A, int n, int nprob, double *U, double *V, double *sig,
                double *phase, int *m, double *tsp, double tol);
void launch_codensity(hipStream_t st, const double *lt, const double *gt, const double *sig, int n, int no,
                      double tol, double *M, int nprob);
void launch_contraction_orbitals(hipStream_t st, const double *const *cx, const double *const *cw, int n, int nm,
                                 int no, double *rowc, double *colc, int nprob);
void launch_prefill_cxxc(hipStream_t st, const double *rowc, const double *colc, int n, int nm, int no, double *cx,
                         double *xc, int nprob);
void launch_lincomb(hipStream_t st, const LinDesc *d, int nb, long long n);
void launch_einsum(hipStream_t st, const TraceDesc *d, int nd, int n);
void launch_antisym(hipStream_t st, const double *eri, double *out, int n);

// Rotating quarter transform Y[l', m] = sum_l c[l, l'] X[m, l]
// (X: rows x nl row-major, c: nl x nout row-major, Y: nout x rows row-major).
void launch_qtransform(hipStream_t st, const double *X, const double *c, double *Y, long long rows, int nl, int nout);
// FP32 variants (mixed-precision build); return false when the shape is unsupported.
bool launch_qtransform_f32(hipStream_t st, const float *X, const float *c, float *Y, long long rows, int nl, int nout);
bool launch_qtransform_f32_out64(hipStream_t st, const float *X, const float *c, double *Y, long long rows, int nl,
                                 int nout);
void launch_to_f32(hipStream_t st, const double *in, float *out, long long n);

// Evaluation.
struct EvalTile {
    int i0, j0;      // first bra / ket determinant
    int ni, nj;      // extent (<= 16 each)
    int pair;        // x * nref + w
    int tri;         // 1: keep only j >= i
};
struct EvalReq { int a, b; long long out; };

void launch_pack_dets(hipStream_t st, const void *dets, long long ndet, PackedDet *out, int *err);
void launch_eval_tiles(hipStream_t st, int la, int lb, bool m0, int mode, const EvalTile *tiles, int ntiles,
                       const PackedDet *dets, const PairInfo *pairs, const double *slab, const int *nocc_ab, int nmo,
                       double enuc, double tol, double *H, double *S, long long ld);
void launch_eval_reqs(hipStream_t st, int la, int lb, bool m0, int mode, const EvalReq *reqs, int nreq,
                      const PackedDet *dets, const PairInfo *pairs, int nref, const double *slab, const int *nocc_ab,
                      int nmo, double enuc, double tol, double *H, double *S);

}  // namespace nw
