// Kernel launchers shared between the build, evaluation and API units.
#pragma once
#include "nw_common.h"

namespace nw {

struct GemmDesc { const double *A; const double *B; double *C; };
struct LinDesc { const double *a, *b, *c; double *out; double ca, cb, cc; };  // out = ca a + cb b + cc c
struct TraceDesc { const double *g; const double *h; double *out; double scale; };

void launch_gemm(hipStream_t st, int ta, int tb, int M, int N, int K, int lda, int ldb, int ldc,
                 double alpha, double beta, const GemmDesc *d_descs, int nbatch);
// C = A B^T on FP64 MFMA for any K (A: M x K, B: N x K, both K-contiguous).
void launch_gemm_nt_mfma(hipStream_t st, int M, int N, int K, int lda, int ldb, int ldc, const GemmDesc *d_descs,
                         int nbatch);
// The same products with K split ksplit ways (partials in part: nbatch x
// ksplit x M x N doubles), summed in a fixed order into each C.
// gemm_nt_splitk_count: the split for a launch of that shape on ncu CUs (1:
// none needed).
int gemm_nt_splitk_count(int M, int N, int K, int nbatch, int ncu);
void launch_gemm_nt_splitk(hipStream_t st, int M, int N, int K, int lda, int ldb, int ldc, const GemmDesc *d_descs,
                           int nbatch, int ksplit, double *part);
void launch_svd(hipStream_t st, const double *A, int n, int nprob, double *U, double *V, double *sig,
                double *phase, int *m, double *tsp, double tol);
// Strides (in doubles) between consecutive problems of one launch.
void launch_codensity(hipStream_t st, const double *lt, const double *gt, const double *sig, int n, int no,
                      double tol, double *M, int nprob, long long in_stride, long long sig_stride, long long out_stride);
void launch_contraction_orbitals(hipStream_t st, const double *const *cx, const double *const *cw, int n, int nm,
                                 int no, double *rowc, double *colc, int nprob, long long out_stride);
void launch_prefill_cxxc(hipStream_t st, const double *rowc, const double *colc, int n, int nm, int no, double *cx,
                         double *xc, int nprob, long long in_stride, long long out_stride);
void launch_lincomb(hipStream_t st, const LinDesc *d, int nb, long long n);
void launch_einsum(hipStream_t st, const TraceDesc *d, int nd, int n);
void launch_antisym(hipStream_t st, const double *eri, double *out, int n);

// Rotating quarter transform Y[l', m] = sum_l c[l, l'] X[m, l]
// (X: rows x nl row-major, c: nl x nout row-major, Y: nout x rows row-major).
void launch_qtransform(hipStream_t st, const double *X, const double *c, double *Y, long long rows, int nl, int nout);
// Up to QMAX tensors per launch travel as kernel arguments (indexed by a grid
// dimension); the *_many launchers split longer lists.
constexpr int QMAX = 32;
struct QPtrs {
    const void *X[QMAX];
    const void *c[QMAX];
    void *Y[QMAX];
};
void launch_qtransform_many(hipStream_t st, const double *const *X, const double *const *c, double *const *Y, int count,
                            long long rows, int nl, int nout);
// FP32 variants (mixed-precision build; every pass, the last included, writes
// FP32); return false when the shape is unsupported.
bool launch_qtransform_f32_many(hipStream_t st, const float *const *X, const float *const *c, float *const *Y,
                                int count, long long rows, int nl, int nout);
// The same with BF16 MFMA on three-way split operands (FP32-level accuracy).
bool launch_qtransform_bf3_many(hipStream_t st, const float *const *X, const float *const *c, float *const *Y,
                                int count, long long rows, int nl, int nout);
// Three-launch bf16x3 four-index transform (see nw_build.hip): the half
// transform of every (p,q) slice, then two leading-index passes.
struct QPtrs2 {
    const void *X[QMAX];
    const void *c3[QMAX];
    const void *c4[QMAX];
    void *Y[QMAX];
};
bool launch_qhalf_bf3_many(hipStream_t st, const float *const *X, const float *const *c3, const float *const *c4,
                           float *const *Y, int count, int n, int nout);
bool launch_qlead_bf3_many(hipStream_t st, const float *const *X, const float *const *c, float *const *Y, int count,
                           int nl, int nout, long long ncols, int nbatch, long long xstride, long long ystride);
void launch_to_f32(hipStream_t st, const double *in, float *out, long long n);

// Evaluation.  Work is described by blocks: a rectangle of determinant pairs
// (bra rows i0..i0+ni, ket columns j0..j0+nj) sharing one ordered reference
// pair and one contraction-rank bin; tri = 1 keeps only j >= i.  Blocks of one
// bin are launched together; tstart is the prefix sum of their 16x16 tile
// counts, and each workgroup binary-searches its block.
struct EvalBlock {
    long long i0, j0;
    int ni, nj;
    int pair, tri;
    long long tstart;  // first tile (EVAL_PROJECT: first 16-row strip) of the block in its bin
    long long pslot;   // EVAL_PROJECT: the block's first row slot in the partial-sum buffers
};
struct EvalReq { int a, b; long long out; };

enum { EVAL_HS = 0, EVAL_FS = 1, EVAL_MATVEC = 2, EVAL_PROJECT = 3 };

struct EvalArgs {
    const PackedDet *dets;
    const PairInfo *pairs;
    const double *slab;
    int nref, nmo, mode;
    double enuc, tol, e0;
    double *A, *B;        // H,S or F,S outputs (matrix: ld = ldo; batched: indexed by req.out)
    long long ldo;
    const double *x;      // matvec input
    double *y, *ys;       // matvec outputs (ys may be null)
};

// Candidate projections from the per-(block, row) partial sums of
// EVAL_PROJECT: row r (determinant nrow0 + r) of row run k (runs sorted by
// first row, run_i0[k]) sums slot table[k * ncol + c] + (row - run_i0[k]) of
// every column run c in order (-1: none), h and s parts.
void launch_project_reduce(hipStream_t st, const double *part_h, const double *part_s, const long long *run_i0,
                           int nruns, const long long *table, int ncol, long long nrow0, long long nrows, double *v,
                           double *s);

void launch_pack_dets(hipStream_t st, const void *dets, long long ndet, int nref, PackedDet *out, int *err);

// GPU binning of batched requests (nw_hs_batched / nw_fs_batched): k_req_key
// writes each request's bin key (-1: exact zero or error) and the per-key
// counts; k_req_scatter places the requests of every key at start[key].
// ((L_a * 5 + L_b) << 1) | m0 with L_a, L_b <= 4; REQ_KEY_GEN + m0 for the
// pairs with a contraction rank above four (up to eight) on a spin
constexpr int REQ_KEY_GEN = 50;
constexpr int REQ_NKEY = 52;
// the same bin of a rank-4-or-less key, or the generic bin's (L_a, L_b)
// stand-in (5, 0) that routes launch_eval_reqs / launch_eval_blocks to the
// run-time-rank kernels
__host__ __device__ inline void key_ranks(int key, int &la, int &lb)
{
    if (key >= REQ_KEY_GEN) { la = 5; lb = 0; }
    else { la = (key >> 1) / 5; lb = (key >> 1) % 5; }
}
enum { REQ_ERR_INDEX = 1, REQ_ERR_UNBUILT = 2, REQ_ERR_RANK = 3 };
struct ReqBin {
    const uint64_t *req;  // nreq (a, b) index pairs
    long long nreq, ndet;
    const void *raw;      // nw_det records on the device
    const PackedDet *dp;
    const PairInfo *pairs;
    int nref, hs;         // hs: H,S requests (exact zeros and m0 bins apply)
    int *key;
    unsigned *hist, *cursor;  // REQ_NKEY counts each (zeroed by the caller)
    const unsigned *start;    // REQ_NKEY offsets into out
    int *err;
    double *A, *B;
    EvalReq *out;
};
void launch_req_bin(hipStream_t st, const ReqBin &b, bool scatter);
// Workgroups per k_eval_blocks launch: tiles beyond it are visited
// grid-stride (256 lanes x 2^22 workgroups stays below 2^32 work-items).
constexpr long long EVAL_GRID_MAX = 1ll << 22;
long long eval_grid_cap();
void launch_eval_blocks(hipStream_t st, int kind, int la, int lb, bool m0, const EvalBlock *blocks, int nblocks,
                        long long ntiles, const EvalArgs &args);
void launch_eval_reqs(hipStream_t st, int kind, int la, int lb, bool m0, const EvalReq *reqs, int nreq,
                      const EvalArgs &args);
// Per-spin overlap factor of each request (nw_s_same_batched); err takes the
// REQ_ERR_* codes (atomicMax).
void launch_s_same(hipStream_t st, const uint64_t *req, long long nreq, long long ndet, int spin,
                   const EvalArgs &args, double *out, int *err);

// Multi-system fused evaluation (nw_batch_*).  Per-system addressing and one
// work item per tile; see k_eval_fused.
struct BatchSys {
    const PackedDet *dets;
    const PairInfo *pairs;
    const double *slab;
    double *H, *S;
    long long ldo;
    int nref, nmo, mode, pad;
    double enuc, tol;
};
struct BatchTile {
    long long i0, j0;
    int sys, pair;
    short ni, nj;
    unsigned char code, shape, tri, pad;
};

// Register class of a contraction-rank code ((la*5+lb) << 1 | m0): 0 = light
// (<= 64 VGPRs per lane in the per-bin kernels), 1 = medium (<= 128), 2 = heavy.
__host__ __device__ constexpr int fused_class(int code)
{
    const int m0 = code & 1, lab = code >> 1, la = lab / 5, lb = lab % 5;
    const int lo = la < lb ? la : lb, hi = la < lb ? lb : la;
    if (m0) {
        if (hi <= 2 && lo <= 1) return 0;
        if ((lo == 0 && hi <= 4) || (lo == 1 && hi == 3) || (lo == 2 && hi == 2)) return 1;
        return 2;
    }
    if (hi <= 1) return 0;
    if (hi == 2 && lo <= 2) return 1;
    return 2;
}

constexpr int FUSE_ALL = -5;
// sel: a code (>= 0), a register class (-1 - class) or FUSE_ALL.
// reset = zero the work counter first (false when the caller zeroed all
// counters of a plan with one launch_reset_counters).
void launch_eval_fused(hipStream_t st, int sel, bool dyn, const BatchTile *tiles, long long ntiles,
                       const BatchSys *sys, unsigned long long *counter, int grid, bool reset = true);
void launch_reset_counters(hipStream_t st, unsigned long long *counters, int n);

}  // namespace nw
