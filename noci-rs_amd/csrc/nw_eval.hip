// Per-determinant-pair H,S evaluation kernels (FP64).
//
// One lane evaluates one determinant pair.  Work arrives grouped so that a
// workgroup shares one ordered reference pair and one contraction-rank bin
// (L_alpha, L_beta), so pair scalars are wave-uniform and the rank-specialised
// code has no divergence.  Arithmetic follows
// src/nonorthogonalwicks/eval/preparehamiltonianoverlap.rs (m = 0: the fixed
// kernels :380-19544 / generic m0 path :20304-20533; m > 0: :20550-20966) and
// eval/helpers.rs.
#include <algorithm>
#include <cstdlib>

#include "nw_eval_gen.h"

namespace nw {

// ---------------------------------------------------------------------------
// Determinant packing (ExcitationSpin::cache, src/lib.rs:128-148).
// ---------------------------------------------------------------------------
struct RawDet {
    uint32_t parent, flags;
    double pha, phb;
    uint64_t oa[2], ob[2], ha[2], pa[2], hb[2], pb[2];
};

__device__ inline int cache_spin(const uint64_t *holes, const uint64_t *parts, uint8_t *idx)
{
    int nh = __popcll(holes[0]) + __popcll(holes[1]);
    int np = __popcll(parts[0]) + __popcll(parts[1]);
    for (int i = 0; i < 8; ++i) idx[i] = 0;
    uint64_t h0 = holes[0], h1 = holes[1], p0 = parts[0], p1 = parts[1];
    for (int i = 0; i < 4; ++i) {
        if (h0) { idx[i] = (uint8_t)__ffsll((long long)h0) - 1; h0 &= h0 - 1; }
        else if (h1) { idx[i] = (uint8_t)(63 + __ffsll((long long)h1)); h1 &= h1 - 1; }
        if (p0) { idx[4 + i] = (uint8_t)__ffsll((long long)p0) - 1; p0 &= p0 - 1; }
        else if (p1) { idx[4 + i] = (uint8_t)(63 + __ffsll((long long)p1)); p1 &= p1 - 1; }
    }
    if (nh != np) return -1;
    return nh;
}

// err bit 0: an excitation rank per spin above 4 or unequal holes and
// particles; bit 1: a parent index at or past nref (checked on the full
// 32-bit value before it is narrowed to the packed 16-bit field).
__global__ void k_pack_dets(const RawDet *__restrict__ in, long long n, int nref, PackedDet *__restrict__ out,
                            int *err)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    RawDet d = in[i];
    PackedDet p;
    int ra = cache_spin(d.ha, d.pa, p.ia);
    int rb = cache_spin(d.hb, d.pb, p.ib);
    if (ra < 0 || rb < 0 || ra > 4 || rb > 4) atomicOr(err, 1);
    if (d.parent >= (uint32_t)nref) {
        atomicOr(err, 2);
        d.parent = 0;
    }
    p.ra = (uint8_t)(ra < 0 ? 0 : ra);
    p.rb = (uint8_t)(rb < 0 ? 0 : rb);
    p.parent = (uint16_t)d.parent;
    p.pha = d.pha < 0 ? -1 : 1;
    p.phb = d.phb < 0 ? -1 : 1;
    p.pad[0] = p.pad[1] = 0;
    out[i] = p;
}

void launch_pack_dets(hipStream_t st, const void *dets, long long ndet, int nref, PackedDet *out, int *err)
{
    if (ndet <= 0) return;
    hipLaunchKernelGGL(k_pack_dets, dim3((unsigned)((ndet + 255) / 256)), dim3(256), 0, st,
                       (const RawDet *)dets, ndet, nref, out, err);
}

// ---------------------------------------------------------------------------
// Kernels.  Block mode: one 256-lane workgroup = one 16 x 16 tile of a block.
// ---------------------------------------------------------------------------
__device__ inline int find_block(const EvalBlock *__restrict__ blocks, int nb, long long t)
{
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blocks[mid].tstart <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// One 16 x 16 tile t of the bin's blocks.  The tile loop of k_eval_blocks is
// workgroup-uniform, so the barrier of the matvec branch is reached by every
// lane of the workgroup on every tile.
template <int KIND, int LA, int LB, bool M0, bool GEN>
__device__ inline void eval_tile(const EvalBlock *__restrict__ blocks, int nblocks, const EvalArgs &a, long long t)
{
    const EvalBlock B = blocks[find_block(blocks, nblocks, t)];
    const int ntj = (B.nj + 15) >> 4;
    const long long loc = t - B.tstart;
    const int ti = (int)(loc / ntj), tj = (int)(loc % ntj);
    const int li = threadIdx.x >> 4, lj = threadIdx.x & 15;
    const long long i = B.i0 + ti * 16 + li, j = B.j0 + tj * 16 + lj;
    bool valid = (ti * 16 + li < B.ni) && (tj * 16 + lj < B.nj) && !(B.tri && j < i);
    if constexpr (KIND == EVAL_HS || KIND == EVAL_FS) {
        if (!valid) return;
    }
    const PairInfo &P = a.pairs[B.pair];
    const slab_t slab = (slab_t)(a.slab + P.base);
    double v0 = 0.0, v1 = 0.0;
    if (valid) {
        const PackedDet bra = a.dets[i], ket = a.dets[j];
        if constexpr (GEN) {
            // contraction rank above four on a spin: the rank is the lane's own
            if constexpr (KIND == EVAL_HS) {
                if constexpr (M0) hs_m0_rt(P, slab, bra, ket, a.nmo, a.enuc, v0, v1);
                else hs_gen_rt(P, slab, bra, ket, a.nmo, a.enuc, a.tol, a.mode, v0, v1);
            } else {
                fs_elem_rt(P, slab, bra, ket, a.nmo, a.tol, a.mode, v0, v1);
            }
        } else if constexpr (KIND == EVAL_HS) {
            if constexpr (M0) hs_m0<LA, LB>(P, slab, bra, ket, a.nmo, a.enuc, v0, v1);
            else hs_gen<LA, LB>(P, slab, bra, ket, a.nmo, a.enuc, a.tol, a.mode, v0, v1);
        } else {
            fs_elem<LA, LB>(P, slab, bra, ket, a.nmo, a.tol, a.mode, v0, v1);
        }
    }
    static_assert(KIND != EVAL_PROJECT, "projections run by strips (eval_strip_project)");
    if constexpr (KIND != EVAL_MATVEC) {
        a.A[i * a.ldo + j] = v0;
        a.B[i * a.ldo + j] = v1;
        if (i != j) {
            a.A[j * a.ldo + i] = v0;
            a.B[j * a.ldo + i] = v1;
        }
    } else {
        // apply_candidate_m (src/snoci/operators.rs:430-476): y_a += M_ab x_b,
        // y_b += M_ab x_a (b != a), M = F - e0 S; optionally the same for S.
        __shared__ double colm[16][17], cols_[16][17];
        const double m = valid ? v0 - a.e0 * v1 : 0.0;
        const double s = valid ? v1 : 0.0;
        const double xj = valid ? a.x[j] : 0.0, xi = valid ? a.x[i] : 0.0;
        double rm = m * xj, rs = s * xj;
#pragma unroll
        for (int o = 8; o >= 1; o >>= 1) {
            rm += __shfl_xor(rm, o, 16);
            rs += __shfl_xor(rs, o, 16);
        }
        if (lj == 0 && ti * 16 + li < B.ni) {
            unsafeAtomicAdd(a.y + B.i0 + ti * 16 + li, rm);
            if (a.ys) unsafeAtomicAdd(a.ys + B.i0 + ti * 16 + li, rs);
        }
        const bool off = valid && i != j;
        colm[li][lj] = off ? m * xi : 0.0;
        cols_[li][lj] = off ? s * xi : 0.0;
        __syncthreads();
        if (threadIdx.x < 16 && tj * 16 + (int)threadIdx.x < B.nj) {
            double cm = 0.0, cs = 0.0;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                cm += colm[r][threadIdx.x];
                cs += cols_[r][threadIdx.x];
            }
            unsafeAtomicAdd(a.y + B.j0 + tj * 16 + threadIdx.x, cm);
            if (a.ys) unsafeAtomicAdd(a.ys + B.j0 + tj * 16 + threadIdx.x, cs);
        }
        __syncthreads();  // colm / cols_ are rewritten by the next tile
    }
}

// build_candidate_current_h + build_candidate_v (src/snoci/operators.rs:151,
// 991) over one 16-row strip s of a block: the workgroup walks the strip's
// column tiles in order, each lane summing H_ai c_i and S_ai c_i over its
// columns, the 16 lanes of a row then reduce by shuffles; the row's sums go
// to its slot of the block (no atomics: k_project_reduce adds the blocks of a
// row in a fixed order, so the projections do not depend on scheduling).
template <int LA, int LB, bool M0, bool GEN>
__device__ inline void eval_strip_project(const EvalBlock *__restrict__ blocks, int nblocks, const EvalArgs &a,
                                          long long s)
{
    const EvalBlock B = blocks[find_block(blocks, nblocks, s)];
    const int ntj = (B.nj + 15) >> 4;
    const int ti = (int)(s - B.tstart);
    const int li = threadIdx.x >> 4, lj = threadIdx.x & 15;
    const long long i = B.i0 + ti * 16 + li;
    const bool row_ok = ti * 16 + li < B.ni;
    const PairInfo &P = a.pairs[B.pair];
    const slab_t slab = (slab_t)(a.slab + P.base);
    double rh = 0.0, rs = 0.0;
    for (int tj = 0; tj < ntj; ++tj) {
        const long long j = B.j0 + tj * 16 + lj;
        if (!row_ok || tj * 16 + lj >= B.nj) continue;
        const PackedDet bra = a.dets[i], ket = a.dets[j];
        double v0, v1;
        if constexpr (GEN) {
            if constexpr (M0) hs_m0_rt(P, slab, bra, ket, a.nmo, a.enuc, v0, v1);
            else hs_gen_rt(P, slab, bra, ket, a.nmo, a.enuc, a.tol, a.mode, v0, v1);
        } else {
            if constexpr (M0) hs_m0<LA, LB>(P, slab, bra, ket, a.nmo, a.enuc, v0, v1);
            else hs_gen<LA, LB>(P, slab, bra, ket, a.nmo, a.enuc, a.tol, a.mode, v0, v1);
        }
        const double xj = a.x[j];
        rh += v0 * xj;
        rs += v1 * xj;
    }
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) {
        rh += __shfl_xor(rh, o, 16);
        rs += __shfl_xor(rs, o, 16);
    }
    if (lj == 0 && row_ok) {
        a.y[B.pslot + ti * 16 + li] = rh;
        a.ys[B.pslot + ti * 16 + li] = rs;
    }
}

__global__ void k_project_reduce(const double *__restrict__ part_h, const double *__restrict__ part_s,
                                 const long long *__restrict__ run_i0, int nruns, const long long *__restrict__ table,
                                 int ncol, long long nrow0, long long nrows, double *__restrict__ v,
                                 double *__restrict__ s)
{
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const long long i = nrow0 + r;
    int lo = 0, hi = nruns - 1;  // the last run starting at or before row i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (run_i0[mid] <= i) lo = mid; else hi = mid - 1;
    }
    const long long off = i - run_i0[lo];
    double h = 0.0, o = 0.0;
    for (int c = 0; c < ncol; ++c) {
        const long long slot = table[(size_t)lo * ncol + c];
        if (slot < 0) continue;
        h += part_h[slot + off];
        o += part_s[slot + off];
    }
    v[r] = h;
    s[r] = o;
}

void launch_project_reduce(hipStream_t st, const double *part_h, const double *part_s, const long long *run_i0,
                           int nruns, const long long *table, int ncol, long long nrow0, long long nrows, double *v,
                           double *s)
{
    if (nrows <= 0) return;
    hipLaunchKernelGGL(k_project_reduce, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, st, part_h, part_s,
                       run_i0, nruns, table, ncol, nrow0, nrows, v, s);
}

// XCD-aware tile order: workgroups are dealt round-robin over the 8 XCDs, so
// workgroup b runs on XCD b % 8.  The tile list is cut into 8 contiguous
// chunks and XCD x takes chunk x, in order: neighbouring tiles (one bra strip
// against consecutive ket strips, which gather from the same rank-4 slices
// J[a, i, :, :] of the strip's bra labels) then meet in one XCD's L2 instead
// of being spread over all eight.  Within a chunk the tiles are visited
// grid-stride, so a bin of any size fits one launch whose grid stays within
// the runtime's limit of 2^32 - 1 work-items.
constexpr int EVAL_XCD = 8;

// Pairs without zero-overlap orbitals (M0) are held to two waves per SIMD
// (256 VGPRs with the accumulation registers, a few bytes of spill): the
// rank-4 bins then ran 22% faster than at one wave with AGPR spill space; the
// m > 0 kernels keep the compiler's choice (at two waves they spill hundreds
// of bytes per lane and ran 24% slower).
template <int KIND, int LA, int LB, bool M0, bool GEN = false>
__global__ __launch_bounds__(256, M0 ? 2 : 1) void k_eval_blocks(const EvalBlock *__restrict__ blocks, int nblocks, long long ntiles,
                                                     EvalArgs a)
{
    const long long chunk = (ntiles + EVAL_XCD - 1) / EVAL_XCD;
    const int xcd = blockIdx.x % EVAL_XCD;
    const long long lo = xcd * chunk, hi = min(ntiles, lo + chunk);
    const long long step = gridDim.x / EVAL_XCD;
    for (long long t = lo + blockIdx.x / EVAL_XCD; t < hi; t += step) {
        if constexpr (KIND == EVAL_PROJECT) eval_strip_project<LA, LB, M0, GEN>(blocks, nblocks, a, t);
        else eval_tile<KIND, LA, LB, M0, GEN>(blocks, nblocks, a, t);
    }
}

template <int KIND, int LA, int LB, bool M0, bool GEN = false>
__global__ __launch_bounds__(256) void k_eval_reqs(const EvalReq *__restrict__ reqs, int nreq, EvalArgs a)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nreq) return;
    const EvalReq r = reqs[q];
    const PackedDet bra = a.dets[r.a], ket = a.dets[r.b];
    const PairInfo &P = a.pairs[(int)bra.parent * a.nref + ket.parent];
    const slab_t slab = (slab_t)(a.slab + P.base);
    double v0, v1;
    if constexpr (GEN) {
        if constexpr (KIND == EVAL_HS) {
            if constexpr (M0) hs_m0_rt(P, slab, bra, ket, a.nmo, a.enuc, v0, v1);
            else hs_gen_rt(P, slab, bra, ket, a.nmo, a.enuc, a.tol, a.mode, v0, v1);
        } else {
            fs_elem_rt(P, slab, bra, ket, a.nmo, a.tol, a.mode, v0, v1);
        }
    } else if constexpr (KIND == EVAL_HS) {
        if constexpr (M0) hs_m0<LA, LB>(P, slab, bra, ket, a.nmo, a.enuc, v0, v1);
        else hs_gen<LA, LB>(P, slab, bra, ket, a.nmo, a.enuc, a.tol, a.mode, v0, v1);
    } else {
        fs_elem<LA, LB>(P, slab, bra, ket, a.nmo, a.tol, a.mode, v0, v1);
    }
    a.A[r.out] = v0;
    a.B[r.out] = v1;
}



long long eval_grid_cap()
{
    // NW_EVAL_GRID_CAP lowers the cap (tests of the grid-stride tile loop)
    const char *e = std::getenv("NW_EVAL_GRID_CAP");
    const long long v = e ? std::atoll(e) : 0;
    return v > 0 ? std::min(v, EVAL_GRID_MAX) : EVAL_GRID_MAX;
}

void launch_eval_blocks(hipStream_t st, int kind, int la, int lb, bool m0, const EvalBlock *blocks, int nblocks,
                        long long ntiles, const EvalArgs &args)
{
    if (ntiles <= 0) return;
    if (la > 4 || lb > 4) {
        // generic bin: every pair with a contraction rank above four on a spin
        const long long chunk = (ntiles + EVAL_XCD - 1) / EVAL_XCD;
        const long long per_xcd = std::max(1ll, std::min(chunk, eval_grid_cap() / EVAL_XCD));
        dim3 grid((unsigned)(per_xcd * EVAL_XCD)), block(256);
        if (kind == EVAL_HS) {
            if (m0) hipLaunchKernelGGL((k_eval_blocks<EVAL_HS, 0, 0, true, true>), grid, block, 0, st, blocks, nblocks, ntiles, args);
            else hipLaunchKernelGGL((k_eval_blocks<EVAL_HS, 0, 0, false, true>), grid, block, 0, st, blocks, nblocks, ntiles, args);
        } else if (kind == EVAL_FS) {
            hipLaunchKernelGGL((k_eval_blocks<EVAL_FS, 0, 0, false, true>), grid, block, 0, st, blocks, nblocks, ntiles, args);
        } else if (kind == EVAL_PROJECT) {
            if (m0) hipLaunchKernelGGL((k_eval_blocks<EVAL_PROJECT, 0, 0, true, true>), grid, block, 0, st, blocks, nblocks, ntiles, args);
            else hipLaunchKernelGGL((k_eval_blocks<EVAL_PROJECT, 0, 0, false, true>), grid, block, 0, st, blocks, nblocks, ntiles, args);
        } else {
            hipLaunchKernelGGL((k_eval_blocks<EVAL_MATVEC, 0, 0, false, true>), grid, block, 0, st, blocks, nblocks, ntiles, args);
        }
        return;
    }
    // a multiple of the XCD count: every XCD gets the same number of workgroups
    const long long chunk = (ntiles + EVAL_XCD - 1) / EVAL_XCD;
    const long long per_xcd = std::max(1ll, std::min(chunk, eval_grid_cap() / EVAL_XCD));
    dim3 grid((unsigned)(per_xcd * EVAL_XCD)), block(256);
#define NW_T(A, B)                                                                                             \
    if (la == A && lb == B) {                                                                                  \
        if (kind == EVAL_HS) {                                                                                 \
            if (m0) hipLaunchKernelGGL((k_eval_blocks<EVAL_HS, A, B, true>), grid, block, 0, st, blocks, nblocks, ntiles, args); \
            else hipLaunchKernelGGL((k_eval_blocks<EVAL_HS, A, B, false>), grid, block, 0, st, blocks, nblocks, ntiles, args); \
        } else if (kind == EVAL_FS) {                                                                          \
            hipLaunchKernelGGL((k_eval_blocks<EVAL_FS, A, B, false>), grid, block, 0, st, blocks, nblocks, ntiles, args);  \
        } else if (kind == EVAL_PROJECT) {                                                                     \
            if (m0) hipLaunchKernelGGL((k_eval_blocks<EVAL_PROJECT, A, B, true>), grid, block, 0, st, blocks, nblocks, ntiles, args); \
            else hipLaunchKernelGGL((k_eval_blocks<EVAL_PROJECT, A, B, false>), grid, block, 0, st, blocks, nblocks, ntiles, args); \
        } else {                                                                                               \
            hipLaunchKernelGGL((k_eval_blocks<EVAL_MATVEC, A, B, false>), grid, block, 0, st, blocks, nblocks, ntiles, args); \
        }                                                                                                      \
        return;                                                                                                \
    }
    NW_BIN_CASES(NW_T)
#undef NW_T
}

// ---------------------------------------------------------------------------
// Request binning for the batched entry points on the GPU
// (calculate_hs_pairs_wicks_batched, src/noci/hs.rs:86-109): same-parent
// pairs that differ in more than four spin orbitals are exactly zero (written
// here); every other request gets the key ((L_a * 5 + L_b) << 1) | m0 of its
// contraction-rank bin.  Counting and placement go through LDS first, so a
// workgroup issues one global atomic per key.  The order of requests inside a
// bin is not fixed; each request's value does not depend on it.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_req_key(ReqBin b)
{
    __shared__ unsigned s_hist[REQ_NKEY];
    for (int k = threadIdx.x; k < REQ_NKEY; k += blockDim.x) s_hist[k] = 0;
    __syncthreads();
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < b.nreq) {
        const uint64_t ia = b.req[2 * q], ib = b.req[2 * q + 1];
        int key = -1;
        if (ia >= (uint64_t)b.ndet || ib >= (uint64_t)b.ndet) {
            atomicMax(b.err, REQ_ERR_INDEX);
        } else {
            const PackedDet da = b.dp[ia], db = b.dp[ib];
            bool zero = false;
            if (b.hs && da.parent == db.parent) {
                const RawDet *ra = reinterpret_cast<const RawDet *>(b.raw) + ia;
                const RawDet *rb = reinterpret_cast<const RawDet *>(b.raw) + ib;
                const int diff = __popcll(ra->oa[0] ^ rb->oa[0]) + __popcll(ra->oa[1] ^ rb->oa[1]) +
                                 __popcll(ra->ob[0] ^ rb->ob[0]) + __popcll(ra->ob[1] ^ rb->ob[1]);
                zero = diff > 4;
            }
            if (zero) {
                b.A[q] = 0.0;
                b.B[q] = 0.0;
            } else {
                const PairInfo &pi = b.pairs[(int)da.parent * b.nref + (int)db.parent];
                const int la = da.ra + db.ra, lb = da.rb + db.rb;
                const int m0 = (b.hs && pi.ma == 0 && pi.mb == 0) ? 1 : 0;
                if (!pi.built) atomicMax(b.err, REQ_ERR_UNBUILT);
                else if (la > 8 || lb > 8) atomicMax(b.err, REQ_ERR_RANK);
                else key = (la > 4 || lb > 4) ? REQ_KEY_GEN + m0 : ((la * 5 + lb) << 1) | m0;
            }
        }
        b.key[q] = key;
        if (key >= 0) atomicAdd(&s_hist[key], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < REQ_NKEY; k += blockDim.x)
        if (s_hist[k]) atomicAdd(&b.hist[k], s_hist[k]);
}

__global__ __launch_bounds__(256) void k_req_scatter(ReqBin b)
{
    __shared__ unsigned s_cnt[REQ_NKEY], s_base[REQ_NKEY];
    for (int k = threadIdx.x; k < REQ_NKEY; k += blockDim.x) s_cnt[k] = 0;
    __syncthreads();
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    unsigned local = 0;
    if (q < b.nreq) {
        key = b.key[q];
        if (key >= 0) local = atomicAdd(&s_cnt[key], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < REQ_NKEY; k += blockDim.x)
        s_base[k] = s_cnt[k] ? atomicAdd(&b.cursor[k], s_cnt[k]) : 0u;
    __syncthreads();
    if (key >= 0)
        b.out[(size_t)b.start[key] + s_base[key] + local] = EvalReq{(int)b.req[2 * q], (int)b.req[2 * q + 1], q};
}

void launch_req_bin(hipStream_t st, const ReqBin &b, bool scatter)
{
    if (b.nreq <= 0) return;
    const dim3 grid((unsigned)((b.nreq + 255) / 256)), block(256);
    if (scatter) hipLaunchKernelGGL(k_req_scatter, grid, block, 0, st, b);
    else hipLaunchKernelGGL(k_req_key, grid, block, 0, st, b);
}

// ---------------------------------------------------------------------------
// Per-spin overlap factors of requested determinant pairs
// (xw_overlap_same_f64_batched, eval/overlap.rs:136-352, over xw_overlap
// :42-66): ph_bra ph_ket phase tilde_s sum over the m-bit column
// distributions of det D_mix, zero when m > L.
// ---------------------------------------------------------------------------
template <int L>
__device__ inline double s_spin(slab_t slab, const long long *o, const SpinScal &sc, int n, const uint8_t *bi,
                                int brank, const uint8_t *ki)
{
    constexpr int L1 = L > 0 ? L : 1;
    if (sc.m > L) return 0.0;
    int rows[L1], cols[L1];
    double d0[L1 * L1], d1[L1 * L1];
    if constexpr (L > 0) {
        labels<L>(bi, brank, ki, sc.nocc, n - sc.nocc, rows, cols);
        build_d<L>(slab + o[O_X], slab + o[O_Y], n, rows, cols, d0);
        if (sc.m > 0) build_d<L>(slab + o[O_X + 1], slab + o[O_Y + 1], n, rows, cols, d1);
    }
    if (sc.m == 0) return det_n<L>(d0);
    double acc = 0.0;
    for (unsigned bits = 0; bits < (1u << L); ++bits) {
        if (__popc(bits) != sc.m) continue;
        double mix[L1 * L1];
#pragma unroll
        for (int r = 0; r < L; ++r)
#pragma unroll
            for (int c = 0; c < L; ++c) mix[r * L + c] = bitk(bits, c) ? d1[r * L + c] : d0[r * L + c];
        acc += det_n<L>(mix);
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_s_same(const uint64_t *__restrict__ req, long long nreq, long long ndet,
                                                int spin, EvalArgs a, double *__restrict__ out, int *err)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nreq) return;
    const uint64_t ia = req[2 * q], ib = req[2 * q + 1];
    if (ia >= (uint64_t)ndet || ib >= (uint64_t)ndet) {
        atomicMax(err, REQ_ERR_INDEX);
        return;
    }
    const PackedDet bra = a.dets[ia], ket = a.dets[ib];
    const PairInfo &P = a.pairs[(int)bra.parent * a.nref + ket.parent];
    if (!P.built) {
        atomicMax(err, REQ_ERR_UNBUILT);
        return;
    }
    const int rb = spin ? bra.rb : bra.ra, rk = spin ? ket.rb : ket.ra;
    const uint8_t *bi = spin ? bra.ib : bra.ia, *ki = spin ? ket.ib : ket.ia;
    const SpinScal sc = load_spin(P, spin);
    const long long *o = P.off + (spin ? O_BB : O_AA);
    const slab_t slab = (slab_t)(a.slab + P.base);
    double v = 0.0;
    switch (rb + rk) {
    case 0: v = s_spin<0>(slab, o, sc, a.nmo, bi, rb, ki); break;
    case 1: v = s_spin<1>(slab, o, sc, a.nmo, bi, rb, ki); break;
    case 2: v = s_spin<2>(slab, o, sc, a.nmo, bi, rb, ki); break;
    case 3: v = s_spin<3>(slab, o, sc, a.nmo, bi, rb, ki); break;
    case 4: v = s_spin<4>(slab, o, sc, a.nmo, bi, rb, ki); break;
    default: v = s_spin_rt(slab, o, sc, a.nmo, bi, rb, ki, rk); break;
    }
    const double ph = spin ? (double)(bra.phb * ket.phb) : (double)(bra.pha * ket.pha);
    out[q] = ph * sc.pref * v;
}

void launch_s_same(hipStream_t st, const uint64_t *req, long long nreq, long long ndet, int spin,
                   const EvalArgs &args, double *out, int *err)
{
    if (nreq <= 0) return;
    hipLaunchKernelGGL(k_s_same, dim3((unsigned)((nreq + 255) / 256)), dim3(256), 0, st, req, nreq, ndet, spin, args,
                       out, err);
}

void launch_eval_reqs(hipStream_t st, int kind, int la, int lb, bool m0, const EvalReq *reqs, int nreq,
                      const EvalArgs &args)
{
    if (nreq <= 0) return;
    dim3 grid((nreq + 255) / 256), block(256);
    if (la > 4 || lb > 4) {
        if (kind == EVAL_HS) {
            if (m0) hipLaunchKernelGGL((k_eval_reqs<EVAL_HS, 0, 0, true, true>), grid, block, 0, st, reqs, nreq, args);
            else hipLaunchKernelGGL((k_eval_reqs<EVAL_HS, 0, 0, false, true>), grid, block, 0, st, reqs, nreq, args);
        } else {
            hipLaunchKernelGGL((k_eval_reqs<EVAL_FS, 0, 0, false, true>), grid, block, 0, st, reqs, nreq, args);
        }
        return;
    }
#define NW_R(A, B)                                                                                             \
    if (la == A && lb == B) {                                                                                  \
        if (kind == EVAL_HS) {                                                                                 \
            if (m0) hipLaunchKernelGGL((k_eval_reqs<EVAL_HS, A, B, true>), grid, block, 0, st, reqs, nreq, args);  \
            else hipLaunchKernelGGL((k_eval_reqs<EVAL_HS, A, B, false>), grid, block, 0, st, reqs, nreq, args);   \
        } else {                                                                                               \
            hipLaunchKernelGGL((k_eval_reqs<EVAL_FS, A, B, false>), grid, block, 0, st, reqs, nreq, args);        \
        }                                                                                                      \
        return;                                                                                                \
    }
    NW_BIN_CASES(NW_R)
#undef NW_R
}

}  // namespace nw
