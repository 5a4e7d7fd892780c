/*
 * nwick.h -- C ABI of the MI355X nonorthogonal-Wick matrix-element engine.
 *
 * Every entry point replaces one piece of noci-rs's Rust surface for the
 * nonorthogonal-Wick path (paths relative to the noci-rs source tree).  Types
 * are plain C: 128-bit occupation strings are split into {lo, hi} uint64_t
 * pairs; matrices are row-major double arrays.
 *
 * Buffers: every pointer argument may be host memory or device memory of the
 * context's GPU; the library detects which and copies only host buffers.
 * The library works on its own HIP stream: device inputs must be complete
 * when a call is made (synchronise the stream that produced them), and every
 * output is complete when the call returns.
 * Calls on one context are serialised internally; distinct contexts are
 * independent.  All functions return 0 (NW_OK) or an nw_status code;
 * nw_last_error() describes the last failure on a context.
 */
#ifndef NWICK_H
#define NWICK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nw_ctx nw_ctx;

enum nw_status {
    NW_OK = 0,
    NW_E_ARG = 1,       /* null pointer / bad value */
    NW_E_SHAPE = 2,     /* dimensions inconsistent or beyond supported limits */
    NW_E_HIP = 3,       /* HIP runtime failure */
    NW_E_MEM = 4,       /* device allocation failed */
    NW_E_RANK = 5,      /* excitation rank per spin beyond 4, or a pair's contraction rank
                           (bra plus ket excitations) per spin beyond 8 (beyond 4 for
                           factor tables and multi-system batches) */
    NW_E_STATE = 6,     /* call order (e.g. evaluation before nw_build) */
    NW_E_NONFINITE = 7, /* non-finite input */
    NW_E_ORDER = 8      /* references not occupied-first */
};

/* Evaluation modes for reference pairs with zero-overlap orbital pairs (m > 0).
 * NW_MODE_REFERENCE reproduces the arithmetic of
 * src/nonorthogonalwicks/eval/preparehamiltonianoverlap.rs:20550-20966
 * literally; NW_MODE_CORRECTED is the exact extended-Wick result (agrees with
 * the generalised Slater-Condon rules of src/noci/naive.rs).  Both give
 * identical results when m_alpha = m_beta = 0. */
enum nw_mode { NW_MODE_REFERENCE = 0, NW_MODE_CORRECTED = 1 };

/* Arithmetic of the per-reference-pair build.  NW_PREC_FP64: FP64 throughout
 * (the reference's arithmetic).  NW_PREC_MIXED: the four-index transforms
 * producing the same-spin J and different-spin II tensors run with FP32
 * operands, FP32 intermediates and FP32 MFMA accumulation; the SVD, zero
 * classification, co-densities, one-body and one-column intermediates and
 * every evaluation stay FP64.  NW_PREC_BF16X3: as NW_PREC_MIXED, with the
 * transforms on BF16 MFMA: every FP32 operand split into three BF16 parts,
 * six cross products accumulated in FP32 (FP32-level error).  With integral
 * factors installed (nw_set_eri_factors) the same two settings choose how
 * the rank-4 block products of pairs above the sigma floor run: FP32 MFMA
 * (NW_PREC_MIXED) or BF16 MFMA on split FP32 factors (NW_PREC_BF16X3). */
enum nw_precision { NW_PREC_FP64 = 0, NW_PREC_MIXED = 1, NW_PREC_BF16X3 = 2 };

/* One determinant: the fields of DetState (src/scalar.rs:32-61) the Wick path
 * reads -- parent reference, excitation phases, occupations and the
 * excitation relative to the parent (Excitation, src/lib.rs:79-119). */
typedef struct nw_det {
    uint32_t parent;
    uint32_t flags; /* reserved, 0 */
    double pha, phb;
    uint64_t oa[2], ob[2];
    uint64_t holes_a[2], parts_a[2];
    uint64_t holes_b[2], parts_b[2];
} nw_det;

/* Per ordered reference pair summary (PairMeta, src/nonorthogonalwicks/types.rs:146). */
typedef struct nw_pair_meta {
    int32_t ma, mb;           /* zero-overlap orbital pairs per spin */
    double tilde_s_a, phase_a; /* SameSpinMeta tilde_s_prod / phase */
    double tilde_s_b, phase_b;
    double sigma_min;         /* smallest singular value above tol, both spins (no reference field) */
    int32_t fp32, pad;        /* 1 when this pair's four-index transforms ran in FP32 (NW_PREC_MIXED) */
} nw_pair_meta;

/* Create a context on `device` holding the AO problem (AoData, src/lib.rs:52-76):
 * S, h: n_bas x n_bas; eri: dense chemists' (ab|cd) stored [a,b,c,d]
 * (src/read.rs:31); enuc: nuclear repulsion. */
int nw_create(nw_ctx **out, int device, int n_bas, const double *S, const double *h,
              const double *eri, double enuc);

/* Cholesky-type factors of the context's two-electron integrals,
 * (ab|cd) = sum_K L[K][a][b] L[K][c][d] with every L[K] symmetric
 * (L: naux x n_bas x n_bas, host or device memory), the other form of the
 * integrals that the AO problem can come in.  With factors installed,
 * nw_build leaves the rank-4 blocks (same-spin J slots, different-spin II
 * blocks) of every reference pair unbuilt, and each
 * evaluation call (nw_hs_matrix, nw_hs_batched, nw_project_hs,
 * nw_hs_matrix_streamed, nw_batch_add) first builds the (row, row) blocks
 * T[p][:][r][:] that its determinants read, as products over K of
 * MO-transformed factors on MFMA, in the pair's build precision.  The slab
 * keeps the reference layout (layout.rs:21-189); nw_pair_slab and
 * nw_save_wicks complete every block first (dense transforms).  The factors
 * are checked against the dense eri of nw_create on a sample of entries and
 * for the symmetry of each L[K]: the largest deviation is returned in
 * max_dev (may be NULL), and one above check_tol is NW_E_ARG.  They take
 * effect at the next nw_build; naux = 0 removes them.  No counterpart in the
 * reference, which transforms the dense integrals (eri_ao2mo_into,
 * src/maths/eri.rs:258-313, called from build.rs:261-313 and 877-917). */
int nw_set_eri_factors(nw_ctx *ctx, const double *L, int naux, double check_tol, double *max_dev);

/* Fraction of the (row, row) blocks of ordered pair (x, w)'s rank-4 tensors
 * that are built: frac3 = {alpha J slots, beta J slots, II blocks}; 1 for a
 * pair built without factors.  Diagnostics; no reference counterpart. */
int nw_rank4_coverage(nw_ctx *ctx, int x, int w, double *frac3);

/* The on-demand rank-4 work of the last evaluation call that built blocks:
 * out10 = {device ms in total, of the factor products, of the rectangle
 * products, of the single blocks, of the exchange products; algorithmic
 * flops of the rectangle products, single blocks, exchange products and
 * factor products; single blocks built}.  Zeros when nothing was built.
 * Diagnostics (measurement); no reference counterpart. */
int nw_rank4_stats(nw_ctx *ctx, double *out10);

/* Generalised Slater-Condon route for batched determinant pairs: replaces
 * calculate_hs_pair_naive (src/noci/hs.rs:372-402: build_s_pair,
 * one_electron, two_electron_same, two_electron_diff of
 * src/noci/naive.rs:73-409) on the GPU, from the reference determinants of
 * nw_set_refs and the integrals of nw_create (no nw_build needed).  req:
 * nreq pairs (bra, ket) of indices into dets; h, s: nreq values each.  The
 * determinants' occupations select their parent's occupied columns (ascending,
 * occ_coeffs naive.rs:20-32); no excitation phase is applied, as in the
 * reference.  O(n_bas^4) per pair (a cross-check route).  Host arrays only;
 * NW_E_ARG for occupation counts that differ from the references'. */
int nw_hs_naive_batched(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq,
                        double *h, double *s);

/* wicks.compare: H, S of every pair by Wick's theorem as nw_hs_matrix
 * (written to H, S), each upper-triangle pair also by the generalised
 * Slater-Condon route (nw_hs_naive_batched); *total = sum over those pairs of
 * |dH| + |dS|, *max_elem = the largest single |dH| or |dS|
 * (compare_hs_pair_wicks_naive, src/noci/hs.rs:141-156, summed as in
 * src/noci/matrix.rs:224-249).  Host arrays; up to 65536 determinants. */
int nw_hs_compare(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, double *H, double *S, int mode, double *total,
                  double *max_elem);

/* The same call's block products by kernel launch class: out8 = {device ms
 * of the FP64 pairs' large-tile products (k_r4_gemm, rectangles, single
 * blocks and exchange products), of their narrow-tile products (ket-particle
 * pairs), of the FP32 pairs' large-tile products (k_r4_bf3 under
 * NW_PREC_BF16X3, else k_r4_gemm on FP32 MFMA), of their narrow-tile
 * products; then the algorithmic flops of the same four}.  Diagnostics
 * (measurement); no reference counterpart. */
int nw_rank4_kernel_stats(nw_ctx *ctx, double *out8);

/* Install the reference determinants (the `refs` of build_wicks_shared,
 * src/noci/wicks.rs:133).  ca, cb: nref x n_bas x nmo row-major MO
 * coefficients; occ: nref x 4 uint64 {oa.lo, oa.hi, ob.lo, ob.hi}.  Occupied
 * columns must come first (src/scf/occupation.rs:75-100). */
int nw_set_refs(nw_ctx *ctx, int nref, int nmo, const double *ca, const double *cb,
                const uint64_t *occ);

/* Build the per-ordered-reference-pair intermediates on the GPU: replaces
 * build_wicks_shared (src/noci/wicks.rs:133-279) with plan_wicks_pairs
 * (:90), assign_offsets (src/nonorthogonalwicks/layout.rs:21),
 * SameSpinBuild::new (build.rs:86) x2, DiffSpinBuild::new (build.rs:728),
 * write_same_spin/write_diff_spin/write_hcol0 (layout.rs:199-300).
 * pair_mask: nref*nref bytes (x*nref + w), nonzero = build; NULL = all pairs.
 * meta_out: nref*nref entries or NULL. */
int nw_build(nw_ctx *ctx, double tol, const uint8_t *pair_mask, nw_pair_meta *meta_out);

/* Batched H,S for arbitrary determinant pairs: replaces
 * calculate_hs_pairs_wicks_batched (src/noci/hs.rs:72-130) and
 * xw_hamiltonian_overlap_prepared_batched (preparehamiltonianoverlap.rs:121).
 * req: nreq x 2 uint64 {bra index, ket index} into dets (host or device
 * memory).  h_out, s_out: nreq (host or device).  The requests are binned by
 * contraction rank on the GPU; an index past ndet is NW_E_ARG, an unbuilt
 * reference pair NW_E_STATE, a contraction rank beyond 8 NW_E_RANK. */
int nw_hs_batched(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, const uint64_t *req,
                  uint64_t nreq, double *h_out, double *s_out, int mode);

/* Full H and S over a determinant list: replaces calculate_matrix_elements
 * (src/noci/matrix.rs:39-81) + scatter_matrix_elements (:93-110) as used by
 * build_noci_hs (:214).  Evaluates pairs i <= j and mirrors them; H, S are
 * ndet x ndet row-major. */
int nw_hs_matrix(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, double *H, double *S, int mode);

/* nw_build + nw_hs_matrix for slabs larger than device memory: replaces
 * build_wicks_shared with a disk-backed slab (src/noci/wicks.rs:213-278,
 * src/nonorthogonalwicks/storage.rs:189-255) followed by
 * calculate_matrix_elements.  The ordered reference pairs the upper triangle
 * of dets touches are sized from their zero-overlap counts, dealt in (x, w)
 * order into groups whose slab blocks fit in slab_bytes, and each group is
 * built, its elements evaluated into H, S, and its slab reused by the next.
 * slab_bytes = 0: the context's current slab allocation plus free device
 * memory, less room for the build's temporaries.  A slab allocation that
 * fits the budget is kept and reused (as by nw_build); a larger one is freed.
 * A pair whose block alone exceeds the budget is NW_E_MEM.  H, S as in
 * nw_hs_matrix; meta_out (nref*nref entries or NULL) receives the meta of
 * every pair built, as from nw_build; ngroups (may be NULL) the number of
 * groups.  The context then holds the last group's build (as nw_build with
 * that mask).  nw_last_timings: total, builds, evaluations. */
int nw_hs_matrix_streamed(nw_ctx *ctx, double tol, const nw_det *dets, uint64_t ndet, double *H,
                          double *S, int mode, uint64_t slab_bytes, nw_pair_meta *meta_out,
                          int *ngroups);

/* Rebuild the generalised-Fock intermediates of every built pair from the
 * spin Fock matrices fa, fb (n_bas x n_bas): replaces update_wicks_fock
 * (src/noci/wicks.rs:290-330) / construct_f_scalar (build.rs:564-617). */
int nw_update_fock(nw_ctx *ctx, const double *fa, const double *fb);

/* Full generalised-Fock and overlap matrices over a determinant list (pairs
 * i <= j, mirrored): F_ab = calculate_f_pair_wicks (src/noci/fock.rs:175),
 * S_ab = calculate_s_pair_wicks (src/noci/overlap.rs:106).  M = F - E0 S is
 * calculate_m_pair_wicks (src/noci/m.rs:160). */
int nw_fs_matrix(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, double *F, double *S, int mode);

/* Batched F, S for arbitrary determinant pairs (req as in nw_hs_batched). */
int nw_fs_batched(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, const uint64_t *req,
                  uint64_t nreq, double *f_out, double *s_out, int mode);

/* Per-spin overlap factors of arbitrary determinant pairs (req as in
 * nw_hs_batched): s_out[q] = the spin-`spin` (0 alpha, 1 beta) overlap of
 * bra req[2q] and ket req[2q+1] with both excitation phases applied,
 * ph_bra ph_ket phase tilde_s sum det D_mix -- xw_overlap_same_f64_batched
 * (src/nonorthogonalwicks/eval/overlap.rs:136-352) over xw_overlap (:42-66)
 * and prepare_same (eval/prepare.rs:27).  The full overlap of the pair is the
 * product of the two spins' factors. */
int nw_s_same_batched(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq, int spin,
                      double *s_out);

/* Shifted generalised-Fock elements of arbitrary determinant pairs:
 * m_out[q] = F - e0 S of (req[2q], req[2q+1]), calculate_m_pair_wicks
 * (src/noci/m.rs:160-197). */
int nw_m_batched(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq, double e0,
                 double *m_out, int mode);

/* Matrix-free y = (F - e0 S) x over all pairs a <= b with the mirror
 * contribution (apply_candidate_m, src/snoci/operators.rs:375-488); if ys is
 * not NULL also ys = S x (apply_candidate_s).  x, y, ys: ndet doubles. */
int nw_matvec_m(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, double e0, const double *x,
                double *y, double *ys, int mode);

/* Projection of candidates on a current NOCI state.  dets = the ncur
 * current-space determinants followed by the candidates; for every candidate
 * a (index ncur + k): v[k] = sum_i H_ai coef_i and s[k] = sum_i S_ai coef_i
 * over the current space i < ncur.  Replaces build_candidate_current_h with
 * build_candidate_v (src/snoci/operators.rs:151-180, 991-999) and the
 * candidate-current overlap of build_snoci_overlaps (:106-149). */
int nw_project_hs(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, uint64_t ncur, const double *coef, double *v,
                  double *s, int mode);

/* NOCI one-particle densities of a state sum_i coef_i |det_i> and its spin
 * Fock matrices, the refresh that precedes nw_update_fock in every SNOCI
 * iteration: noci_density (src/noci/naive.rs:231-289, pair densities
 * :73-229) with zero threshold tol, then fock (src/scf/kernels.rs:46).
 * fa, fb (n_bas x n_bas) required; da, db optional (NULL). */
int nw_noci_fock(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, const double *coef, double tol, double *fa,
                 double *fb, double *da, double *db);

/* Select the build arithmetic (nw_precision) for subsequent nw_build calls.
 * No reference counterpart: the reference computes in FP64 only. */
int nw_set_precision(nw_ctx *ctx, int precision);

/* Under NW_PREC_MIXED, reference pairs whose occupied-overlap matrices have a
 * nonzero singular value below sigma_floor (per spin, values <= tol excluded)
 * keep FP64 transforms: the co-densities carry 1/sigma, which amplifies FP32
 * rounding in the matrix elements roughly as 1/sigma^4.  Default 0.1.  No
 * reference counterpart. */
int nw_set_mixed_floor(nw_ctx *ctx, double sigma_floor);

/* Copy one ordered pair's slab block (layout of layout.rs:21-189) and its
 * offsets (95 int64, relative to the block, -1 = absent, last = length) and
 * scalar meta (32 doubles) out of device memory; out may be NULL to query
 * the length through offsets[94].  For tests and disk interop.  A pair built
 * with reduced-precision transforms (nw_pair_meta.fp32) keeps its rank-4
 * blocks as half-length FP32 blocks on the device; the copy (and its offsets)
 * is in the reference layout, those blocks widened to FP64, as nw_save_wicks
 * writes them. */
int nw_pair_slab(nw_ctx *ctx, int x, int w, double *out, int64_t *offsets, double *meta);

/* Wick disk cache in the reference's format (src/nonorthogonalwicks/storage.rs:
 * WICKS_DISK_CACHE_VERSION 2): `dir/wicks.bin` = the slab (f64, pairs
 * x * nref + w back to back, create_wicks_mmap) and `dir/wicks.meta` = the
 * bincode serialisation of WicksDiskMeta<f64> (offsets and PairMeta).
 * nw_save_wicks needs a build of every ordered pair.  nw_load_wicks replaces
 * nw_build (load_wicks_mmap, :227-255): it checks the version, nref, nmo,
 * occupations and that every pair's layout is assign_offsets'; tol is the
 * zero threshold for later evaluations.  A loaded slab has no co-densities,
 * so nw_update_fock needs a real nw_build. */
int nw_save_wicks(nw_ctx *ctx, const char *dir);
int nw_load_wicks(nw_ctx *ctx, const char *dir, double tol);

/* nw_hs_batched for callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * No counterpart in the reference.  Same requests, outputs and errors as
 * nw_hs_batched, but dets, req, h_out and s_out must be device memory of the
 * context's GPU (NW_E_ARG otherwise), ndet <= 2^18 and nreq <= 2^20
 * (NW_E_SHAPE otherwise), and nothing is allocated: the temporaries live in
 * one fixed device array of about 26 MiB per GPU, which that GPU holds from
 * the first call on it until the process ends.  Calls on one GPU share the
 * array and are serialised within the process (contexts included); the call
 * synchronises with the device (it reads the per-bin counts), so it is not
 * itself graph-capturable. */
int nw_hs_batched_noalloc(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq,
                          double *h_out, double *s_out, int mode);

/* Device time in milliseconds of the kernels in the last nw_build /
 * nw_hs_matrix / nw_hs_batched call, split as {total, transform, evaluate}. */
int nw_last_timings(nw_ctx *ctx, double *ms3);

const char *nw_last_error(nw_ctx *ctx);
/* Frees the context and every factor object created on it (nw_factor_create);
 * such a factor must not be used or passed to nw_factor_destroy afterwards. */
void nw_destroy(nw_ctx *ctx);

/* ---------------------------------------------------------------------------
 * Spin-factorised one-body and overlap operators (src/noci/factorise).
 * A determinant list is split per parent reference P into its unique alpha
 * and beta excitation strings (SpinFactorisation::new, mod.rs:75-93;
 * assign_aids/assign_bids :185-278; build_parent_spin_spaces :315-410).  For
 * every ordered parent pair (Q target, P source) and spin the per-spin
 * overlap and generalised-Fock factors between Q's and P's strings are
 * evaluated from the built Wick intermediates, bra = the parent whose first
 * determinant comes first (ordered_parent_pair, mod.rs:163-183), and
 * y = (F + lambda S) x is applied as two table contractions per pair.
 * Components are numbered per parent by (rank, holes, particles, phase).
 * Every determinant's per-spin rank plus any other parent's must be <= 4.
 * ------------------------------------------------------------------------- */
typedef struct nw_factor nw_factor;

/* Table storage: NW_FACTOR_DEVICE keeps every ordered pair's tables in
 * device memory (SNOCIStorage::RAM in OneBodyFactorisation::new,
 * onebody.rs:152-210); NW_FACTOR_TRANSIENT regenerates one pair's tables
 * per use (SNOCIStorage::None, the Transient blocks). */
enum nw_factor_storage { NW_FACTOR_DEVICE = 0, NW_FACTOR_TRANSIENT = 1 };

/* Factorise dets and build the tables from the context's current build and
 * Fock: SpinFactorisation::new + OneBodyFactorisation::new
 * (build_one_body_factor_tables, onebody.rs:938-1017;
 * build_spin_one_body_factor_row :1300-1569) and the overlap factors of
 * OverlapFactors::build_overlap_factors (overlap.rs:280-397), which are the S
 * halves of the same tables.  mode: nw_mode for m > 0 pairs. */
int nw_factor_create(nw_ctx *ctx, const nw_det *dets, uint64_t ndet, int mode, int storage, nw_factor **out);
/* Rebuild the tables after nw_update_fock (a new OneBodyFactorisation per
 * SNOCI iteration, src/snoci/step.rs:292-305).  Applying stale tables fails
 * with NW_E_STATE. */
int nw_factor_refresh(nw_factor *f);
/* y = (F + lambda S) x over every determinant pair (apply_one_body,
 * onebody.rs:290-345, with apply_one_body_a_first/_b_first :491-648 chosen by
 * select_one_body_contraction :1583); ys = S x if not NULL (the overlap
 * action of apply_overlap_sparse, overlap.rs:434-496, on a dense vector).
 * lambda = -E0 gives the PT2 operator of apply_factorised_shifted_omega_m
 * (src/snoci/operators.rs:762-778) before the projection correction. */
int nw_factor_apply(nw_factor *f, double lambda, const double *x, double *y, double *ys);
/* nw_factor_apply on two vectors at once, interleaved as x2[2i + v] (the
 * real and imaginary parts of a complex vector, v = 0, 1), likewise y2 and
 * ys2: each table row is read once for both.  This is the real-table action
 * on a complex vector that apply_factorised_shifted_omega_m
 * (src/snoci/operators.rs:759-776) takes with T = Complex64 for imaginary
 * shifts; the i eps S x part is formed by the caller from ys2. */
int nw_factor_apply_pair(nw_factor *f, double lambda, const double *x2, double *y2, double *ys2);
/* Diagonals of F + lambda S and of S (one_body_diagonals, onebody.rs:358-405);
 * either output may be NULL. */
int nw_factor_diagonals(nw_factor *f, double lambda, double *mdiag, double *sdiag);
/* Component counts of one parent (parent_component_counts, mod.rs:132-146;
 * parent < 0 skips them) and the bytes of the resident tables
 * (OneBodyFactorisation::storage_bytes, onebody.rs:231-281, counted over
 * every ordered pair). */
int nw_factor_info(const nw_factor *f, int parent, uint64_t *ncomp_a, uint64_t *ncomp_b, uint64_t *table_bytes);
/* Tables that nw_factor_apply reads: NW_PREC_FP64 (default) or NW_PREC_MIXED,
 * FP32 tables and FP32 materialised rows with FP64 sums (rounded copies of
 * the resident FP64 tables, made on first use and kept current by
 * nw_factor_refresh; regenerated tables are written in FP32 directly).
 * Diagonals always come from FP64 tables; nw_factor_info and
 * nw_factor_table report the selected tables.  No counterpart in the
 * reference, which applies the tables in FP64: this is the low-precision
 * operator of an iteratively refined PT2 solve, whose residuals use FP64. */
int nw_factor_set_precision(nw_factor *f, int precision);
/* Table bytes one nw_factor_apply reads with the selected tables: every
 * table row it sums or gathers from, once, and each table that a pass reads
 * for both directions of a parent pair (Tab(p, q) = Tab(q, p)^T) once.  No
 * counterpart in the reference; the algorithmic traffic of the apply. */
int nw_factor_apply_bytes(nw_factor *f, uint64_t *bytes);
/* Source groups whose stage-1 sums run as dense products on FP64 MFMA (the
 * single x single block of each parent; environment NW_FAC_BLOCK=0 turns them
 * off, NW_FAC_BLOCK_MIN_G sets the smallest block, default 16 groups), and the
 * number of such blocks.  Diagnostics; no counterpart in the reference. */
int nw_factor_block_groups(const nw_factor *f, uint64_t *groups, uint64_t *blocks);
/* Per-determinant component ids (SpinFactorisation::aid/bid, mod.rs:106-130). */
int nw_factor_ids(const nw_factor *f, uint32_t *aids, uint32_t *bids);
/* Copy one table: ncomp(source) x ncomp(target) interleaved {S, F} doubles,
 * element [s][t] = the spin factor between target component t of parent
 * `target` and source component s of parent `source`.  For tests. */
int nw_factor_table(nw_factor *f, int target, int source, int spin, double *out);
/* Device time of the last nw_factor_apply {total, table regeneration} in ms. */
int nw_factor_last_ms(const nw_factor *f, double *ms2);
void nw_factor_destroy(nw_factor *f);

/* ---------------------------------------------------------------------------
 * Multi-system batches.  No single reference counterpart: noci-rs evaluates
 * one system per build_noci_hs call (src/noci/matrix.rs:214, pair list and
 * Rayon loop of calculate_matrix_elements :39-81).  A batch takes several
 * built contexts (each keeps its own slab; none may be rebuilt or destroyed
 * while the batch exists) with one determinant list each, and evaluates the
 * upper-triangle H, S of every system in one interleaved plan of persistent
 * fused launches, captured into a HIP graph and replayed.
 * ------------------------------------------------------------------------- */
typedef struct nw_batch nw_batch;

enum nw_batch_flags { NW_BATCH_TUNE = 1, NW_BATCH_GRAPH = 2 };
/* Kernel grouping of the plan: one launch per contraction-rank code, per
 * register class of codes, or a single launch for every code. */
enum nw_batch_grouping { NW_GROUP_CODE = 0, NW_GROUP_CLASS = 1, NW_GROUP_ALL = 2 };

/* One tuned bucket: systems with n_occ = nocc (max over spins) and
 * contraction-rank code ((l_alpha * 5 + l_beta) << 1 | m0).  shape = log2 of
 * the tile width TJ (tiles are 256/TJ bra x TJ ket determinants). */
typedef struct nw_tune_entry {
    int32_t nocc, code, shape, pad;
    double ms_best, ms_default; /* sampled bucket time with the chosen / 16x16 tiles */
} nw_tune_entry;

int nw_batch_create(nw_batch **out, int device);
/* Add one system: a built context and its determinants (as nw_hs_matrix). */
int nw_batch_add(nw_batch *b, nw_ctx *ctx, const nw_det *dets, uint64_t ndet, int mode, int *index);
/* Plan the batch.  NW_BATCH_TUNE times tile shapes per bucket and then the
 * schedule (grouping, static/dynamic distribution, persistent grid) on the
 * batch itself; NW_BATCH_GRAPH captures the plan into a HIP graph. */
int nw_batch_plan(nw_batch *b, int flags);
/* Evaluate every system (graph replay when captured); blocks until done. */
int nw_batch_run(nw_batch *b);
/* Copy system `index`'s ndet x ndet H and S (host or device buffers). */
int nw_batch_result(nw_batch *b, int index, double *H, double *S);
/* The schedule in force and the tuner's per-bucket report. */
int nw_batch_schedule(nw_batch *b, int *grouping, int *dyn, int *grid_per_cu, nw_tune_entry *entries, int cap,
                      int *count);
/* Impose a schedule (e.g. one cached from an earlier tuning run). */
int nw_batch_set_schedule(nw_batch *b, int grouping, int dyn, int grid_per_cu, const nw_tune_entry *entries, int n);
/* Device time of the last nw_batch_run and the number of pairs it evaluated. */
int nw_batch_last_ms(nw_batch *b, double *ms, long long *npairs);
const char *nw_batch_last_error(nw_batch *b);
void nw_batch_destroy(nw_batch *b);

/* ---------------------------------------------------------------------------
 * Learned candidate screener.  No reference counterpart: noci-rs ranks SNOCI
 * and PT2 candidates by exact evaluation only.  A perceptron n_in-128-128-1
 * (ReLU) with BF16 weight copies, FP32 activations and accumulation, FP32
 * master weights trained with Adam on the weighted mean squared error
 * L = sum_r w_r (y_r - t_r)^2 / (2 n).  Parameters are one FP32 vector laid out
 * W1 [128 x n_in], b1 [128], W2 [128 x 128], b2 [128], W3 [128], b3 [1].
 * ------------------------------------------------------------------------- */
typedef struct nw_mlp nw_mlp;

int nw_mlp_create(nw_mlp **out, int device, int n_in, uint64_t seed); /* n_in <= 32; He-uniform init */
int nw_mlp_num_params(nw_mlp *net, long long *n);
int nw_mlp_set_params(nw_mlp *net, const float *params, long long n); /* also resets Adam state */
int nw_mlp_get_params(nw_mlp *net, float *params, long long n);
/* y[r] = network(X[r, :]) for n rows of X (n x n_in, row-major). */
int nw_mlp_forward(nw_mlp *net, const float *X, long long n, float *y);
/* Loss and gradient over all n rows (w may be NULL for unit weights). */
int nw_mlp_grad(nw_mlp *net, const float *X, const float *t, const float *w, long long n, float *loss, float *grad);
/* One full-batch Adam step (lr; beta1 0.9, beta2 0.999, eps 1e-8). */
int nw_mlp_step(nw_mlp *net, const float *X, const float *t, const float *w, long long n, float lr, float *loss);
/* Mini-batch Adam over shuffled rows for `epochs` epochs; epoch_loss[e] =
 * mean batch loss of epoch e (may be NULL). */
int nw_mlp_train(nw_mlp *net, const float *X, const float *t, const float *w, long long n, int epochs, int batch,
                 float lr, uint64_t seed, float *epoch_loss);
const char *nw_mlp_last_error(nw_mlp *net);
void nw_mlp_destroy(nw_mlp *net);

#ifdef __cplusplus
}
#endif
#endif
