// Per-reference-pair build kernels (FP64): occupied-overlap SVD, co-densities,
// batched small GEMMs and the four-index AO->MO transforms.
//
// The algorithm follows src/nonorthogonalwicks/build.rs; the decomposition
// into GPU launches is this engine's own: every step runs batched over all
// requested ordered reference pairs and both spins at once.
#include <mutex>
#include <type_traits>
#include <unordered_map>

#include "nw_kernels.h"

namespace nw {

// ---------------------------------------------------------------------------
// Batched FP64 GEMM, C = alpha op(A) op(B) + beta C, row-major operands.
// op(A) is M x K: TA=0 reads A[i*lda+k], TA=1 reads A[k*lda+i].
// op(B) is K x N: TB=0 reads B[k*ldb+j], TB=1 reads B[j*ldb+k].
// 64x64 tile per 256-thread workgroup, 16-deep K steps staged in LDS, each
// thread accumulating a 4x4 strided sub-tile.
// ---------------------------------------------------------------------------
template <int TA, int TB>
__device__ inline void gemm_tile(const GemmDesc g, int M, int N, int K, int lda, int ldb, int ldc, double alpha,
                                 double beta)
{
    constexpr int BM = 64, BN = 64, BK = 16;
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN + 1];
    const int tid = threadIdx.x;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int tx = tid & 15, ty = tid >> 4;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;

    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int e4 = 0; e4 < 4; ++e4) {
            int e = tid + e4 * 256;
            int i, k;
            if (TA == 0) { i = e / BK; k = e % BK; } else { k = e / BM; i = e % BM; }
            int gi = m0 + i, gk = k0 + k;
            double v = 0.0;
            if (gi < M && gk < K) v = TA == 0 ? g.A[(size_t)gi * lda + gk] : g.A[(size_t)gk * lda + gi];
            As[k][i] = v;
            int j;
            if (TB == 0) { k = e / BN; j = e % BN; } else { j = e / BK; k = e % BK; }
            int gj = n0 + j;
            gk = k0 + k;
            double w = 0.0;
            if (gj < N && gk < K) w = TB == 0 ? g.B[(size_t)gk * ldb + gj] : g.B[(size_t)gj * ldb + gk];
            Bs[k][j] = w;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < BK; ++k) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[k][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int gi = m0 + ty + 16 * i;
        if (gi >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int gj = n0 + tx + 16 * j;
            if (gj >= N) continue;
            double *c = g.C + (size_t)gi * ldc + gj;
            double v = alpha * acc[i][j];
            if (beta != 0.0) v += beta * *c;
            *c = v;
        }
    }
}

template <int TA, int TB>
__global__ __launch_bounds__(256) void k_gemm(const GemmDesc *__restrict__ descs, int M, int N, int K,
                                               int lda, int ldb, int ldc, double alpha, double beta)
{
    gemm_tile<TA, TB>(descs[blockIdx.z], M, N, K, lda, ldb, ldc, alpha, beta);
}

// Rotating quarter transform of the four-index AO->MO transformation
// (eri_ao2mo_into, src/maths/eri.rs:258-313): Y[l', m] = sum_l c[l, l'] X[m, l].
// The contracted index is the fastest of X and the new index becomes the
// slowest of Y, so four passes return the original index order.
__global__ __launch_bounds__(256) void k_qtransform(const QPtrs q, int rows, int nl, int nout)
{
    GemmDesc g;
    g.A = (const double *)q.c[blockIdx.z];
    g.B = (const double *)q.X[blockIdx.z];
    g.C = (double *)q.Y[blockIdx.z];
    gemm_tile<1, 1>(g, nout, rows, nl, nout, nl, rows, 1.0, 0.0);
}

// MFMA form of the rotating quarter transform (v_mfma_f64_16x16x4f64).
// One workgroup = NW waves; wave w owns output rows l' in [16w, 16w+16) and
// keeps its 16 x nl slice of c in registers as MFMA A fragments (lane ->
// c[4kk + lane/16][16w + lane%16]).  The workgroup walks 64-column tiles of
// X (64 rows m of X, contiguous in HBM) staged in LDS; each wave issues
// KS x 4 MFMAs per tile and stores its 16 x 64 block of Y with 128-byte
// coalesced row segments.  Persistent grid: tiles strided by gridDim.x.
typedef double nw_d4 __attribute__((ext_vector_type(4)));
// Row of the 16x16 FP32 MFMA result held in accumulator register r of lane
// (lane/16 = lk): 4*lk + r on CDNA3/4.
#define NW_F32_ROW(lk, r) (4 * (lk) + (r))

template <int KS>
__global__ __launch_bounds__(512) void k_qtransform_mfma(const QPtrs q, long long rows, int nl, int nout,
                                                          long long ntiles)
{
    const double *__restrict__ X = (const double *)q.X[blockIdx.y];
    const double *__restrict__ c = (const double *)q.c[blockIdx.y];
    double *__restrict__ Y = (double *)q.Y[blockIdx.y];
    extern __shared__ double xs[];
    constexpr int NLP = KS * 4;
    constexpr int LDX = NLP + 1;
    const int nwave = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int lp = wave * 16 + lr;
    double afr[KS];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
        const int l = kk * 4 + lk;
        afr[kk] = (l < nl && lp < nout) ? c[(size_t)l * nout + lp] : 0.0;
    }
    const bool active = wave * 16 < nout;
    for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const long long m0 = t * 64;
        const int mcount = (int)((rows - m0) < 64 ? (rows - m0) : 64);
        __syncthreads();
        for (int r = wave; r < 64; r += nwave) {
            const double *src = X + (size_t)(m0 + r) * nl;
#pragma unroll
            for (int k0 = 0; k0 < NLP; k0 += 64) {
                const int k = k0 + lane;
                if (k < NLP) xs[r * LDX + k] = (r < mcount && k < nl) ? src[k] : 0.0;
            }
        }
        __syncthreads();
        if (!active) continue;
        nw_d4 acc[4];
#pragma unroll
        for (int ns = 0; ns < 4; ++ns) acc[ns] = nw_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const double a = afr[kk];
#pragma unroll
            for (int ns = 0; ns < 4; ++ns) {
                const double b = xs[(ns * 16 + lr) * LDX + kk * 4 + lk];
                acc[ns] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[ns], 0, 0, 0);
            }
        }
#pragma unroll
        for (int ns = 0; ns < 4; ++ns) {
            const int mm = ns * 16 + lr;
            if (mm >= mcount) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wave * 16 + 4 * r + lk;  // D row = 4*reg + lane/16 (probed on gfx950)
                if (row < nout) Y[(size_t)row * rows + m0 + mm] = acc[ns][r];
            }
        }
    }
}

// Pipelined form for even nl: the next 64-row tile of X is fetched into
// registers (16-byte loads, all in flight at once) while the current tile is
// multiplied out of LDS, so HBM latency hides behind the MFMA work.  Rows of
// X are nl long; the LDS tile is padded to NL = 4 KS columns with zeros so
// the K = 4 MFMA steps need no guards.
template <int KS, bool PAD2, int TRD, int NW>
__global__ __launch_bounds__(64 * NW) void k_qtransform_mfma_pipe(const QPtrs q, long long rows, int nout,
                                                                  long long ntiles)
{
    constexpr int nl = KS * 4 - (PAD2 ? 2 : 0);  // row length of X (even)
    const double *__restrict__ X = (const double *)q.X[blockIdx.y];
    const double *__restrict__ c = (const double *)q.c[blockIdx.y];
    double *__restrict__ Y = (double *)q.Y[blockIdx.y];
    constexpr int NL = KS * 4;
    constexpr int LDX = NL + 1;
    constexpr int NT = 64 * NW;
    constexpr int PER = (TRD * NL / 2 + NT - 1) / NT;  // double2 per thread (upper bound)
    extern __shared__ double xs[];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int lp = wave * 16 + lr;
    constexpr int tile2 = TRD * nl / 2;
    constexpr int npad = NL - nl;
    if constexpr (npad > 0)
        for (int e = tid; e < TRD * npad; e += NT) xs[(e / npad) * LDX + nl + e % npad] = 0.0;
    double afr[KS];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
        const int l = kk * 4 + lk;
        afr[kk] = (lp < nout && l < nl) ? c[(size_t)l * nout + lp] : 0.0;
    }
    const bool active = wave * 16 < nout;
    double2 buf[PER];
    auto gload = [&](long long t) {
        const double2 *src = reinterpret_cast<const double2 *>(X + (size_t)t * TRD * nl);
        const long long rem = rows - t * TRD;
        const int lim2 = (int)((rem < TRD ? rem : TRD) * nl / 2);
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e2 = tid + j * NT;
            buf[j] = (e2 < lim2) ? src[e2] : double2{0.0, 0.0};
        }
    };
    long long t = blockIdx.x;
    if (t < ntiles) gload(t);
    for (; t < ntiles; t += gridDim.x) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int e2 = tid + j * NT;
            if (e2 < tile2) {
                const int e = 2 * e2, r = e / nl, k = e % nl;
                xs[r * LDX + k] = buf[j].x;
                xs[r * LDX + k + 1] = buf[j].y;
            }
        }
        __syncthreads();
        const long long tn = t + gridDim.x;
        if (tn < ntiles) gload(tn);
        if (!active) continue;
        const long long m0 = t * TRD;
        const int mcount = (int)((rows - m0) < TRD ? (rows - m0) : TRD);
        nw_d4 acc[TRD / 16];
#pragma unroll
        for (int ns = 0; ns < TRD / 16; ++ns) acc[ns] = nw_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const double a = afr[kk];
#pragma unroll
            for (int ns = 0; ns < TRD / 16; ++ns) {
                const double b = xs[(ns * 16 + lr) * LDX + kk * 4 + lk];
                acc[ns] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[ns], 0, 0, 0);
            }
        }
#pragma unroll
        for (int ns = 0; ns < TRD / 16; ++ns) {
            const int mm = ns * 16 + lr;
            if (mm >= mcount) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wave * 16 + 4 * r + lk;
                if (row < nout) Y[(size_t)row * rows + m0 + mm] = acc[ns][r];
            }
        }
    }
}

// FP32 form used by the mixed-precision build: FP32 operands and
// intermediates, FP32 MFMA accumulation (v_mfma_f32_16x16x4f32), FP32 output
// (the final pass writes the pair's FP32 rank-4 slab block).
// 128-row tiles of X, staged in LDS, next tile prefetched into registers.
typedef float nw_f4 __attribute__((ext_vector_type(4)));

// Rows of X per LDS tile of the pipelined FP64 / FP32 transforms.
#ifndef NW_TR_F64
#define NW_TR_F64 32
#endif
#ifndef NW_TR_F64_SMALL
#define NW_TR_F64_SMALL 64
#endif
#ifndef NW_TR_F32
#define NW_TR_F32 64
#endif

template <int KS, typename OutT, int VEC, int TR, int NW>
__global__ __launch_bounds__(64 * NW) void k_qtransform_f32_pipe(const QPtrs q, long long rows, int nout,
                                                                 long long ntiles)
{
    constexpr int nl = KS * 4 - (VEC == 4 ? 0 : 2);  // row length of X: float4 rows when nl % 4 == 0
    const float *__restrict__ X = (const float *)q.X[blockIdx.y];
    const float *__restrict__ c = (const float *)q.c[blockIdx.y];
    OutT *__restrict__ Y = (OutT *)q.Y[blockIdx.y];
    using VT = typename std::conditional<VEC == 4, float4, float2>::type;
    constexpr int NL = KS * 4;
    constexpr int LDX = NL + 1;
    constexpr int NT = 64 * NW;
    constexpr int PER = (TR * NL / VEC + NT - 1) / NT;     // vectors per thread (upper bound)
    extern __shared__ float xsf[];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int lp = wave * 16 + lr;
    constexpr int tilev = TR * nl / VEC;
    constexpr int npad = NL - nl;
    if constexpr (npad > 0)
        for (int e = tid; e < TR * npad; e += NT) xsf[(e / npad) * LDX + nl + e % npad] = 0.f;
    float afr[KS];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
        const int l = kk * 4 + lk;
        afr[kk] = (lp < nout && l < nl) ? c[(size_t)l * nout + lp] : 0.0f;
    }
    const bool active = wave * 16 < nout;
    VT buf[PER];
    auto gload = [&](long long t) {
        const VT *src = reinterpret_cast<const VT *>(X + (size_t)t * TR * nl);
        const long long rem = rows - t * TR;
        const int limv = (int)((rem < TR ? rem : TR) * nl / VEC);
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int ev = tid + j * NT;
            buf[j] = (ev < limv) ? src[ev] : VT{};
        }
    };
    long long t = blockIdx.x;
    if (t < ntiles) gload(t);
    for (; t < ntiles; t += gridDim.x) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int ev = tid + j * NT;
            if (ev < tilev) {
                const int e = VEC * ev, r = e / nl, k = e % nl;
                float *d = xsf + r * LDX + k;
                d[0] = buf[j].x;
                d[1] = buf[j].y;
                if constexpr (VEC == 4) {
                    d[2] = buf[j].z;
                    d[3] = buf[j].w;
                }
            }
        }
        __syncthreads();
        const long long tn = t + gridDim.x;
        if (tn < ntiles) gload(tn);
        if (!active) continue;
        const long long m0 = t * TR;
        const int mcount = (int)((rows - m0) < TR ? (rows - m0) : TR);
        nw_f4 acc[TR / 16];
#pragma unroll
        for (int ns = 0; ns < TR / 16; ++ns) acc[ns] = nw_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const float a = afr[kk];
#pragma unroll
            for (int ns = 0; ns < TR / 16; ++ns) {
                const float b = xsf[(ns * 16 + lr) * LDX + kk * 4 + lk];
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[ns], 0, 0, 0);
            }
        }
#pragma unroll
        for (int ns = 0; ns < TR / 16; ++ns) {
            const int mm = ns * 16 + lr;
            if (mm >= mcount) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wave * 16 + NW_F32_ROW(lk, r);
                if (row < nout) Y[(size_t)row * rows + m0 + mm] = (OutT)acc[ns][r];
            }
        }
    }
}

__global__ void k_to_f32(const double *__restrict__ in, float *__restrict__ out, long long n)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        out[e] = (float)in[e];
}

void launch_to_f32(hipStream_t st, const double *in, float *out, long long n)
{
    if (n <= 0) return;
    long long g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)g), dim3(256), 0, st, in, out, n);
}

// Workgroups of one kernel instance that are resident on the whole device at
// once (occupancy x CU count).  A persistent grid larger than this leaves a
// second round of workgroups that starts only as the first drains.
template <typename F>
static long long resident_grid(F kern, int block, size_t lds)
{
    static std::mutex mu;
    static std::unordered_map<const void *, long long> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find((const void *)kern);
    if (it != cache.end()) return it->second;
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1) ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, block, lds) != hipSuccess || per < 1) per = 1;
    (void)hipGetLastError();
    const long long g = (long long)per * ncu;
    cache[(const void *)kern] = g;
    return g;
}

// Number of tensors of one batched launch and the persistent grid per tensor.
static unsigned grid_x(long long ntiles, int count, long long target)
{
    long long g = target / count;
    if (g < 1) g = 1;
    if (g > ntiles) g = ntiles;
    return (unsigned)g;
}

// Waves per workgroup: one per 16 output rows.  Narrow transforms (nout <= 64
// with rows of at most 64) get 1- to 4-wave workgroups, so no wave idles.
inline int waves_for(int nout) { return nout <= 64 ? (nout + 15) / 16 : 7; }

template <int K, typename OutT, int TR>
static void (*f32_kernel(bool v4, int nout, int &nw))(const QPtrs, long long, int, long long)
{
    nw = 7;
    if constexpr (K <= 16) {
        nw = waves_for(nout);
        if (nw == 1) return v4 ? k_qtransform_f32_pipe<K, OutT, 4, TR, 1> : k_qtransform_f32_pipe<K, OutT, 2, TR, 1>;
        if (nw == 2) return v4 ? k_qtransform_f32_pipe<K, OutT, 4, TR, 2> : k_qtransform_f32_pipe<K, OutT, 2, TR, 2>;
        if (nw == 3) return v4 ? k_qtransform_f32_pipe<K, OutT, 4, TR, 3> : k_qtransform_f32_pipe<K, OutT, 2, TR, 3>;
        if (nw == 4) return v4 ? k_qtransform_f32_pipe<K, OutT, 4, TR, 4> : k_qtransform_f32_pipe<K, OutT, 2, TR, 4>;
    }
    return v4 ? k_qtransform_f32_pipe<K, OutT, 4, TR, 7> : k_qtransform_f32_pipe<K, OutT, 2, TR, 7>;
}

template <int K, int TRD>
static void (*f64_kernel(bool pad2, int nout, int &nw))(const QPtrs, long long, int, long long)
{
    nw = 7;
    if constexpr (K <= 16) {
        nw = waves_for(nout);
        if (nw == 1) return pad2 ? k_qtransform_mfma_pipe<K, true, TRD, 1> : k_qtransform_mfma_pipe<K, false, TRD, 1>;
        if (nw == 2) return pad2 ? k_qtransform_mfma_pipe<K, true, TRD, 2> : k_qtransform_mfma_pipe<K, false, TRD, 2>;
        if (nw == 3) return pad2 ? k_qtransform_mfma_pipe<K, true, TRD, 3> : k_qtransform_mfma_pipe<K, false, TRD, 3>;
        if (nw == 4) return pad2 ? k_qtransform_mfma_pipe<K, true, TRD, 4> : k_qtransform_mfma_pipe<K, false, TRD, 4>;
    }
    return pad2 ? k_qtransform_mfma_pipe<K, true, TRD, 7> : k_qtransform_mfma_pipe<K, false, TRD, 7>;
}

template <typename OutT>
static bool launch_qtransform_f32_t(hipStream_t st, const float *const *X, const float *const *c, OutT *const *Y,
                                    int count, long long rows, int nl, int nout)
{
    const int ks = (nl + 3) / 4;
    if (nl % 2 != 0 || ks > 31 || (nout + 15) / 16 > 7) return false;
    const bool v4 = nl % 4 == 0;
    constexpr int TR = NW_TR_F32;
    const long long ntiles = (rows + TR - 1) / TR;
    const size_t lds = (size_t)TR * (ks * 4 + 1) * sizeof(float);
    for (int b0 = 0; b0 < count; b0 += QMAX) {
        const int nb = count - b0 < QMAX ? count - b0 : QMAX;
        QPtrs q{};
        for (int k = 0; k < nb; ++k) { q.X[k] = X[b0 + k]; q.c[k] = c[b0 + k]; q.Y[k] = Y[b0 + k]; }
#define NW_QF(K)                                                                                               \
    case K: {                                                                                                  \
        int nw = 7;                                                                                            \
        void (*kern)(const QPtrs, long long, int, long long) = f32_kernel<K, OutT, TR>(v4, nout, nw);          \
        const dim3 g(grid_x(ntiles, nb, resident_grid(kern, 64 * nw, lds)), (unsigned)nb);                    \
        hipLaunchKernelGGL(kern, g, dim3(64 * nw), lds, st, q, rows, nout, ntiles);                           \
        break;                                                                                                 \
    }
        switch (ks) {
            NW_QF(1) NW_QF(2) NW_QF(3) NW_QF(4) NW_QF(5) NW_QF(6) NW_QF(7) NW_QF(8) NW_QF(9) NW_QF(10)
            NW_QF(11) NW_QF(12) NW_QF(13) NW_QF(14) NW_QF(15) NW_QF(16) NW_QF(17) NW_QF(18) NW_QF(19)
            NW_QF(20) NW_QF(21) NW_QF(22) NW_QF(23) NW_QF(24) NW_QF(25) NW_QF(26) NW_QF(27) NW_QF(28)
            NW_QF(29) NW_QF(30) NW_QF(31)
        default: return false;
        }
#undef NW_QF
    }
    return true;
}

bool launch_qtransform_f32_many(hipStream_t st, const float *const *X, const float *const *c, float *const *Y,
                                int count, long long rows, int nl, int nout)
{
    return launch_qtransform_f32_t<float>(st, X, c, Y, count, rows, nl, nout);
}


// ---------------------------------------------------------------------------
// The same quarter transform with BF16 MFMA on split operands ("bf16x3"):
// every FP32 value v is written as h + m + l with h = bf16(v),
// m = bf16(v - h), l = bf16(v - h - m) (round to nearest even), which keeps
// 24 significant bits, and a product as the six terms whose weight is at
// least 2^-16 of h_a h_b:
//   ab ~ h_a h_b + h_a m_b + m_a h_b + h_a l_b + l_a h_b + m_a m_b,
// each an exact BF16 x BF16 product accumulated in FP32 by
// v_mfma_f32_16x16x32_bf16 (K = 32 per instruction; a lane holds 8
// consecutive K values of one row of A and one column of B, the same K
// assignment for both).  The dropped terms (m l, l m, l l) are below
// 2^-24 relative, so the result has FP32-level error at 6/16 of the cost of
// FP32 MFMA per product.  X tiles of 64 rows are split once when staged in
// LDS (three bf16 planes, K padded to a multiple of 32 with zeros); the
// coefficient fragments are split once per wave.
typedef __bf16 nw_bf8 __attribute__((ext_vector_type(8)));

__device__ inline void nw_split3(float v, __bf16 &h, __bf16 &m, __bf16 &l)
{
    h = (__bf16)v;
    const float r = v - (float)h;
    m = (__bf16)r;
    l = (__bf16)(r - (float)m);
}

template <int KS, typename OutT, int VEC>
__global__ __launch_bounds__(448) void k_qtransform_bf3(const QPtrs q, long long rows, int nl, int nout, long long ntiles)
{
    constexpr int NT = 448;
    constexpr int TR = 64;            // rows of X per tile
    constexpr int KP = KS * 32;       // padded K
    constexpr int LDB = KP + 8;       // bf16 per LDS row (16 B pad)
    constexpr int PLANE = TR * LDB;
    const float *__restrict__ X = (const float *)q.X[blockIdx.y];
    const float *__restrict__ c = (const float *)q.c[blockIdx.y];
    OutT *__restrict__ Y = (OutT *)q.Y[blockIdx.y];
    using VT = typename std::conditional<VEC == 4, float4, float2>::type;
    extern __shared__ __bf16 xb[];  // planes h, m, l
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int lp = wave * 16 + lr;
    // zero the K padding of every plane once; tiles only write k < nl
    for (int e = tid; e < 3 * TR * (KP - nl); e += NT) {
        const int pl = e / (TR * (KP - nl)), rem = e % (TR * (KP - nl));
        xb[pl * PLANE + (rem / (KP - nl)) * LDB + nl + rem % (KP - nl)] = (__bf16)0.0f;
    }
    // coefficient fragments: row lp of A = c[:, lp], K = 32 ks + 8 lk + j
    nw_bf8 ah[KS], am[KS], al[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int l = ks * 32 + lk * 8 + j;
            const float v = (lp < nout && l < nl) ? c[(size_t)l * nout + lp] : 0.0f;
            __bf16 h, m, lo;
            nw_split3(v, h, m, lo);
            ah[ks][j] = h;
            am[ks][j] = m;
            al[ks][j] = lo;
        }
    // this thread's vectors of a tile: element e = VEC (tid + j NT) -> row r, column k
    const int tilev = TR * nl / VEC;
    constexpr int PER = (TR * KP / VEC + NT - 1) / NT;  // upper bound (nl <= KP)
    const int step = VEC * NT, dr = step / nl, dk = step % nl;
    const int r0 = (VEC * tid) / nl, k0 = (VEC * tid) % nl;
    VT buf[PER];
    auto gload = [&](long long t) {
        const VT *src = reinterpret_cast<const VT *>(X + (size_t)t * TR * nl);
        const long long rem = rows - t * TR;
        const int limv = (int)((rem < TR ? rem : TR) * nl / VEC);
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int ev = tid + j * NT;
            buf[j] = (ev < limv) ? src[ev] : VT{};
        }
    };
    const bool active = wave * 16 < nout;
    long long t = blockIdx.x;
    if (t < ntiles) gload(t);
    for (; t < ntiles; t += gridDim.x) {
        __syncthreads();
        {
            int r = r0, k = k0;
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int ev = tid + j * NT;
                if (ev < tilev) {
                    __bf16 h[VEC], m[VEC], lo[VEC];
                    nw_split3(buf[j].x, h[0], m[0], lo[0]);
                    nw_split3(buf[j].y, h[1], m[1], lo[1]);
                    if constexpr (VEC == 4) {
                        nw_split3(buf[j].z, h[2], m[2], lo[2]);
                        nw_split3(buf[j].w, h[3], m[3], lo[3]);
                    }
                    __bf16 *d = xb + r * LDB + k;
#pragma unroll
                    for (int v = 0; v < VEC; ++v) {
                        d[v] = h[v];
                        d[PLANE + v] = m[v];
                        d[2 * PLANE + v] = lo[v];
                    }
                }
                r += dr;
                k += dk;
                if (k >= nl) {
                    k -= nl;
                    ++r;
                }
            }
        }
        __syncthreads();
        const long long tn = t + gridDim.x;
        if (tn < ntiles) gload(tn);
        if (!active) continue;
        const long long m0 = t * TR;
        const int mcount = (int)((rows - m0) < TR ? (rows - m0) : TR);
        nw_f4 acc[TR / 16];
#pragma unroll
        for (int ns = 0; ns < TR / 16; ++ns) acc[ns] = nw_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int ns = 0; ns < TR / 16; ++ns) {
                const __bf16 *bp = xb + (ns * 16 + lr) * LDB + ks * 32 + lk * 8;
                const nw_bf8 bh = *reinterpret_cast<const nw_bf8 *>(bp);
                const nw_bf8 bm = *reinterpret_cast<const nw_bf8 *>(bp + PLANE);
                const nw_bf8 bl = *reinterpret_cast<const nw_bf8 *>(bp + 2 * PLANE);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[ks], bh, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bl, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[ks], bm, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[ks], bh, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bm, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bh, acc[ns], 0, 0, 0);
            }
        }
#pragma unroll
        for (int ns = 0; ns < TR / 16; ++ns) {
            const int mm = ns * 16 + lr;
            if (mm >= mcount) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wave * 16 + NW_F32_ROW(lk, r);
                if (row < nout) Y[(size_t)row * rows + m0 + mm] = (OutT)acc[ns][r];
            }
        }
    }
}

template <typename OutT>
static bool launch_qtransform_bf3_t(hipStream_t st, const float *const *X, const float *const *c, OutT *const *Y,
                                    int count, long long rows, int nl, int nout)
{
    const int ks = (nl + 31) / 32;
    if (nl % 2 != 0 || ks > 8 || (nout + 15) / 16 > 7) return false;
    const bool v4 = nl % 4 == 0;
    const long long ntiles = (rows + 63) / 64;
    const size_t lds = (size_t)3 * 64 * (ks * 32 + 8) * sizeof(__bf16);
    const dim3 b(448);
    for (int b0 = 0; b0 < count; b0 += QMAX) {
        const int nb = count - b0 < QMAX ? count - b0 : QMAX;
        QPtrs q{};
        for (int k = 0; k < nb; ++k) { q.X[k] = X[b0 + k]; q.c[k] = c[b0 + k]; q.Y[k] = Y[b0 + k]; }
#define NW_QB(K)                                                                                      \
    case K: {                                                                                         \
        void (*kern)(const QPtrs, long long, int, int, long long) =                                   \
            v4 ? k_qtransform_bf3<K, OutT, 4> : k_qtransform_bf3<K, OutT, 2>;                         \
        const dim3 g(grid_x(ntiles, nb, resident_grid(kern, 448, lds)), (unsigned)nb);                \
        hipLaunchKernelGGL(kern, g, b, lds, st, q, rows, nl, nout, ntiles);                           \
        break;                                                                                        \
    }
        switch (ks) {
            NW_QB(1) NW_QB(2) NW_QB(3) NW_QB(4) NW_QB(5) NW_QB(6) NW_QB(7) NW_QB(8)
        default: return false;
        }
#undef NW_QB
    }
    return true;
}

bool launch_qtransform_bf3_many(hipStream_t st, const float *const *X, const float *const *c, float *const *Y,
                                int count, long long rows, int nl, int nout)
{
    return launch_qtransform_bf3_t<float>(st, X, c, Y, count, rows, nl, nout);
}


// ---------------------------------------------------------------------------
// Three-launch form of the bf16x3 four-index transform, which moves three
// tensors' worth of HBM traffic instead of four:
//   half:  Y2[p q][o3 o4] = sum_{r,s} c3[r][o3] c4[s][o4] X[p q][r s]
//   lead:  Z[o1][q o3 o4]  = sum_p c1[p][o1] Y2[p][q o3 o4]
//   lead:  Y[o1][o2 o3 o4] = sum_q c2[q][o2] Z[o1][q][o3 o4]   (batch over o1)
// which is the same Y[o1 o2 o3 o4] as the four quarter passes.  The half
// transform keeps a whole n x n slice and its first product in LDS; the
// leading passes stream 64-column tiles of a row-major matrix.

// Half transform of one (p,q) slice per step.  NS = ceil(n/16) tiles of 16
// slice rows; seven waves own 16 output rows each in both products.  The
// first product T[o4][r] = sum_s c4[s][o4] X[r][s] goes back to the same LDS
// planes as rows o4 with K = r (zero for r >= n), the B operand of the
// second product Y2[o3][o4] = sum_r c3[r][o3] T[o4][r].
template <int NS, int VEC>
__global__ __launch_bounds__(448) void k_qhalf_bf3(const QPtrs2 q, int n, int nout, long long nslice)
{
    constexpr int NT = 448;
    constexpr int TR = NS * 16;        // slice rows (and T rows) held in LDS
    constexpr int KS = (NS + 1) / 2;   // K steps of 32: KP = 32 KS >= n
    constexpr int KP = KS * 32;
    constexpr int LDB = KP + 8;
    constexpr int PLANE = TR * LDB;
    const float *__restrict__ X = (const float *)q.X[blockIdx.y];
    const float *__restrict__ c3 = (const float *)q.c3[blockIdx.y];
    const float *__restrict__ c4 = (const float *)q.c4[blockIdx.y];
    float *__restrict__ Y = (float *)q.Y[blockIdx.y];
    using VT = typename std::conditional<VEC == 4, float4, float2>::type;
    extern __shared__ __bf16 xb[];  // planes h, m, l
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int lp = wave * 16 + lr;
    // entries that neither the slice nor T ever writes (K columns >= TR, and
    // the K padding before the first T) must read as zero
    for (int e = tid; e < 3 * PLANE; e += NT) xb[e] = (__bf16)0.0f;
    nw_bf8 ah[KS], am[KS], al[KS], bh3[KS], bm3[KS], bl3[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int l = ks * 32 + lk * 8 + j;
            const bool ok = lp < nout && l < n;
            __bf16 h, m, lo;
            nw_split3(ok ? c4[(size_t)l * nout + lp] : 0.0f, h, m, lo);
            ah[ks][j] = h;
            am[ks][j] = m;
            al[ks][j] = lo;
            nw_split3(ok ? c3[(size_t)l * nout + lp] : 0.0f, h, m, lo);
            bh3[ks][j] = h;
            bm3[ks][j] = m;
            bl3[ks][j] = lo;
        }
    const int tilev = n * n / VEC;
    constexpr int PER = (TR * KP / VEC + NT - 1) / NT;  // n <= TR and n <= KP
    const int step = VEC * NT, dr = step / n, dk = step % n;
    const int r0 = (VEC * tid) / n, k0 = (VEC * tid) % n;
    VT buf[PER];
    auto gload = [&](long long t) {
        const VT *src = reinterpret_cast<const VT *>(X + (size_t)t * n * n);
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int ev = tid + j * NT;
            buf[j] = src[ev < tilev ? ev : tilev - 1];  // clamped: no branch per load; staging skips ev >= tilev
        }
    };
    const bool active = wave * 16 < nout;
    long long t = blockIdx.x;
    if (t < nslice) gload(t);
    for (; t < nslice; t += gridDim.x) {
        __syncthreads();  // the previous slice's second product has read the planes
        {
            int r = r0, k = k0;
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int ev = tid + j * NT;
                if (ev < tilev) {
                    __bf16 h[VEC], m[VEC], lo[VEC];
                    nw_split3(buf[j].x, h[0], m[0], lo[0]);
                    nw_split3(buf[j].y, h[1], m[1], lo[1]);
                    if constexpr (VEC == 4) {
                        nw_split3(buf[j].z, h[2], m[2], lo[2]);
                        nw_split3(buf[j].w, h[3], m[3], lo[3]);
                    }
                    __bf16 *d = xb + r * LDB + k;
#pragma unroll
                    for (int v = 0; v < VEC; ++v) {
                        d[v] = h[v];
                        d[PLANE + v] = m[v];
                        d[2 * PLANE + v] = lo[v];
                    }
                }
                r += dr;
                k += dk;
                if (k >= n) {
                    k -= n;
                    ++r;
                }
            }
        }
        __syncthreads();
        const long long tn = t + gridDim.x;
        if (tn < nslice) gload(tn);
        nw_f4 acc[NS];
#pragma unroll
        for (int ns = 0; ns < NS; ++ns) acc[ns] = nw_f4{0.f, 0.f, 0.f, 0.f};
        if (active) {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
                for (int ns = 0; ns < NS; ++ns) {
                    const __bf16 *bp = xb + (ns * 16 + lr) * LDB + ks * 32 + lk * 8;
                    const nw_bf8 bh = *reinterpret_cast<const nw_bf8 *>(bp);
                    const nw_bf8 bm = *reinterpret_cast<const nw_bf8 *>(bp + PLANE);
                    const nw_bf8 bl = *reinterpret_cast<const nw_bf8 *>(bp + 2 * PLANE);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[ks], bh, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bl, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[ks], bm, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[ks], bh, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bm, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bh, acc[ns], 0, 0, 0);
                }
            }
        }
        __syncthreads();  // every wave has read the slice
        // T rows o4 < TR, K columns r < min(TR, KP); r >= n written as zero
#pragma unroll
        for (int ns = 0; ns < NS; ++ns) {
            const int r = ns * 16 + lr;
            if (r >= KP) continue;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int o4 = wave * 16 + NW_F32_ROW(lk, i);
                if (o4 >= TR) continue;
                __bf16 h, m, lo;
                nw_split3(r < n ? acc[ns][i] : 0.0f, h, m, lo);
                __bf16 *d = xb + o4 * LDB + r;
                d[0] = h;
                d[PLANE] = m;
                d[2 * PLANE] = lo;
            }
        }
        __syncthreads();
        if (active) {
#pragma unroll
            for (int ns = 0; ns < NS; ++ns) acc[ns] = nw_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
                for (int ns = 0; ns < NS; ++ns) {
                    const __bf16 *bp = xb + (ns * 16 + lr) * LDB + ks * 32 + lk * 8;
                    const nw_bf8 bh = *reinterpret_cast<const nw_bf8 *>(bp);
                    const nw_bf8 bm = *reinterpret_cast<const nw_bf8 *>(bp + PLANE);
                    const nw_bf8 bl = *reinterpret_cast<const nw_bf8 *>(bp + 2 * PLANE);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl3[ks], bh, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh3[ks], bl, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bm3[ks], bm, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bm3[ks], bh, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh3[ks], bm, acc[ns], 0, 0, 0);
                    acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh3[ks], bh, acc[ns], 0, 0, 0);
                }
            }
            float *yo = Y + (size_t)t * nout * nout;
#pragma unroll
            for (int ns = 0; ns < NS; ++ns) {
                const int o4 = ns * 16 + lr;
                if (o4 >= nout) continue;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int o3 = wave * 16 + NW_F32_ROW(lk, i);
                    if (o3 < nout) yo[(size_t)o3 * nout + o4] = acc[ns][i];
                }
            }
        }
    }
}

// Leading-index quarter transform of a batch of row-major matrices:
// Y[b][o][j] = sum_p c[p][o] X[b][p][j] (p < nl, o < nout, j < ncols).  A
// step is one 64-column tile; wave w stages the load groups w, w+7, ... of
// eight consecutive p for the column of its lane (each load instruction
// reads 64 consecutive floats of one row), so the planes hold the tile as
// [column][p], K-contiguous for the MFMA B operand.
template <int KS, typename OutT>
__global__ __launch_bounds__(448) void k_qlead_bf3(const QPtrs q, int nl, int nout, long long ncols, int nbatch,
                                                   long long xstride, long long ystride)
{
    constexpr int NT = 448, TC = 64;
    constexpr int KP = KS * 32;
    constexpr int LDB = KP + 8;
    constexpr int PLANE = TC * LDB;
    constexpr int NG = KP / 8;              // load groups of eight p
    constexpr int PER = (NG + 6) / 7;       // groups per wave
    const float *__restrict__ X = (const float *)q.X[blockIdx.y];
    const float *__restrict__ c = (const float *)q.c[blockIdx.y];
    OutT *__restrict__ Y = (OutT *)q.Y[blockIdx.y];
    extern __shared__ __bf16 xb[];
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int lp = wave * 16 + lr;
    nw_bf8 ah[KS], am[KS], al[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int l = ks * 32 + lk * 8 + j;
            const float v = (lp < nout && l < nl) ? c[(size_t)l * nout + lp] : 0.0f;
            __bf16 h, m, lo;
            nw_split3(v, h, m, lo);
            ah[ks][j] = h;
            am[ks][j] = m;
            al[ks][j] = lo;
        }
    const long long tpb = (ncols + TC - 1) / TC, ntiles = tpb * nbatch;
    float buf[PER][8];
    // loads from clamped (always valid) addresses, zeroed when staged: no
    // branch or wait per load; the group test is uniform over the wave
    auto gload = [&](long long t) {
        const long long b = t / tpb, j = (t - b * tpb) * TC + lane;
        const float *src = X + b * xstride + (j < ncols ? j : ncols - 1);
#pragma unroll
        for (int g = 0; g < PER; ++g) {
            const int grp = wave + 7 * g;
            if (grp < NG) {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int p = grp * 8 + i;
                    buf[g][i] = src[(size_t)(p < nl ? p : nl - 1) * ncols];
                }
            }
        }
    };
    const bool active = wave * 16 < nout;
    long long t = blockIdx.x;
    if (t < ntiles) gload(t);
    for (; t < ntiles; t += gridDim.x) {
        __syncthreads();
        const bool jok = (t % tpb) * TC + lane < ncols;
#pragma unroll
        for (int g = 0; g < PER; ++g) {
            const int grp = wave + 7 * g;
            if (grp >= NG) continue;
            nw_bf8 h8, m8, l8;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                __bf16 h, m, lo;
                nw_split3((jok && grp * 8 + i < nl) ? buf[g][i] : 0.0f, h, m, lo);
                h8[i] = h;
                m8[i] = m;
                l8[i] = lo;
            }
            __bf16 *d = xb + lane * LDB + grp * 8;
            *reinterpret_cast<nw_bf8 *>(d) = h8;
            *reinterpret_cast<nw_bf8 *>(d + PLANE) = m8;
            *reinterpret_cast<nw_bf8 *>(d + 2 * PLANE) = l8;
        }
        __syncthreads();
        const long long tn = t + gridDim.x;
        if (tn < ntiles) gload(tn);
        if (!active) continue;
        const long long b = t / tpb, j0 = (t - b * tpb) * TC;
        nw_f4 acc[TC / 16];
#pragma unroll
        for (int ns = 0; ns < TC / 16; ++ns) acc[ns] = nw_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int ns = 0; ns < TC / 16; ++ns) {
                const __bf16 *bp = xb + (ns * 16 + lr) * LDB + ks * 32 + lk * 8;
                const nw_bf8 bh = *reinterpret_cast<const nw_bf8 *>(bp);
                const nw_bf8 bm = *reinterpret_cast<const nw_bf8 *>(bp + PLANE);
                const nw_bf8 bl = *reinterpret_cast<const nw_bf8 *>(bp + 2 * PLANE);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[ks], bh, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bl, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[ks], bm, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am[ks], bh, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bm, acc[ns], 0, 0, 0);
                acc[ns] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[ks], bh, acc[ns], 0, 0, 0);
            }
        }
        OutT *yb = Y + b * ystride;
#pragma unroll
        for (int ns = 0; ns < TC / 16; ++ns) {
            const long long j = j0 + ns * 16 + lr;
            if (j >= ncols) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wave * 16 + NW_F32_ROW(lk, r);
                if (row < nout) yb[(size_t)row * ncols + j] = (OutT)acc[ns][r];
            }
        }
    }
}

bool launch_qhalf_bf3_many(hipStream_t st, const float *const *X, const float *const *c3, const float *const *c4,
                           float *const *Y, int count, int n, int nout)
{
    if (n < 2 || n % 2 != 0 || n > 112 || nout < 1 || nout > n) return false;
    const int ns = (n + 15) / 16, ks = (ns + 1) / 2;
    const bool v4 = n % 4 == 0;
    const size_t lds = (size_t)3 * (16 * ns) * (32 * ks + 8) * sizeof(__bf16);
    const long long nslice = (long long)n * n;
    const dim3 b(448);
    for (int b0 = 0; b0 < count; b0 += QMAX) {
        const int nb = count - b0 < QMAX ? count - b0 : QMAX;
        QPtrs2 q{};
        for (int k = 0; k < nb; ++k) {
            q.X[k] = X[b0 + k];
            q.c3[k] = c3[b0 + k];
            q.c4[k] = c4[b0 + k];
            q.Y[k] = Y[b0 + k];
        }
#define NW_QH(NS)                                                                                     \
    case NS: {                                                                                        \
        void (*kern)(const QPtrs2, int, int, long long) = v4 ? k_qhalf_bf3<NS, 4> : k_qhalf_bf3<NS, 2>; \
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
        const dim3 g(grid_x(nslice, nb, resident_grid(kern, 448, lds)), (unsigned)nb);                \
        hipLaunchKernelGGL(kern, g, b, lds, st, q, n, nout, nslice);                                  \
        break;                                                                                        \
    }
        switch (ns) {
            NW_QH(1) NW_QH(2) NW_QH(3) NW_QH(4) NW_QH(5) NW_QH(6) NW_QH(7)
        default: return false;
        }
#undef NW_QH
    }
    return true;
}

template <typename OutT>
static bool launch_qlead_bf3_t(hipStream_t st, const float *const *X, const float *const *c, OutT *const *Y, int count,
                               int nl, int nout, long long ncols, int nbatch, long long xstride, long long ystride)
{
    if (nl < 1 || nl > 128 || nout < 1 || nout > 112 || ncols < 1 || nbatch < 1) return false;
    const int ks = (nl + 31) / 32;
    const size_t lds = (size_t)3 * 64 * (ks * 32 + 8) * sizeof(__bf16);
    const long long ntiles = (ncols + 63) / 64 * nbatch;
    const dim3 b(448);
    for (int b0 = 0; b0 < count; b0 += QMAX) {
        const int nb = count - b0 < QMAX ? count - b0 : QMAX;
        QPtrs q{};
        for (int k = 0; k < nb; ++k) { q.X[k] = X[b0 + k]; q.c[k] = c[b0 + k]; q.Y[k] = Y[b0 + k]; }
#define NW_QL(K)                                                                                      \
    case K: {                                                                                         \
        void (*kern)(const QPtrs, int, int, long long, int, long long, long long) = k_qlead_bf3<K, OutT>; \
        const dim3 g(grid_x(ntiles, nb, resident_grid(kern, 448, lds)), (unsigned)nb);                \
        hipLaunchKernelGGL(kern, g, b, lds, st, q, nl, nout, ncols, nbatch, xstride, ystride);        \
        break;                                                                                        \
    }
        switch (ks) {
            NW_QL(1) NW_QL(2) NW_QL(3) NW_QL(4)
        default: return false;
        }
#undef NW_QL
    }
    return true;
}

bool launch_qlead_bf3_many(hipStream_t st, const float *const *X, const float *const *c, float *const *Y, int count,
                           int nl, int nout, long long ncols, int nbatch, long long xstride, long long ystride)
{
    return launch_qlead_bf3_t<float>(st, X, c, Y, count, nl, nout, ncols, nbatch, xstride, ystride);
}


void launch_qtransform_many(hipStream_t st, const double *const *X, const double *const *c, double *const *Y, int count,
                            long long rows, int nl, int nout)
{
    const int ks = (nl + 3) / 4;
    const int nw = (nout + 15) / 16;
    for (int b0 = 0; b0 < count; b0 += QMAX) {
        const int nb = count - b0 < QMAX ? count - b0 : QMAX;
        QPtrs q{};
        for (int k = 0; k < nb; ++k) { q.X[k] = X[b0 + k]; q.c[k] = c[b0 + k]; q.Y[k] = Y[b0 + k]; }
        bool done = false;
        if (nl % 2 == 0 && ks <= 31 && nw <= 7) {
            done = true;
            // small n (K <= 64): tiles of NW_TR_F64_SMALL rows, so a tile's
            // staging and barriers are spread over more products
#define NW_QP(K)                                                                                        \
    case K: {                                                                                           \
        constexpr int TRK = K <= 16 ? NW_TR_F64_SMALL : NW_TR_F64;                                      \
        const size_t lds = (size_t)TRK * (ks * 4 + 1) * sizeof(double);                                 \
        const long long ntiles = (rows + TRK - 1) / TRK;                                                \
        int nwv = 7;                                                                                    \
        void (*kern)(const QPtrs, long long, int, long long) = f64_kernel<K, TRK>(nl % 4 != 0, nout, nwv); \
        const dim3 g(grid_x(ntiles, nb, resident_grid(kern, 64 * nwv, lds)), (unsigned)nb);            \
        hipLaunchKernelGGL(kern, g, dim3(64 * nwv), lds, st, q, rows, nout, ntiles);                   \
        break;                                                                                          \
    }
            switch (ks) {
                NW_QP(1) NW_QP(2) NW_QP(3) NW_QP(4) NW_QP(5) NW_QP(6) NW_QP(7) NW_QP(8) NW_QP(9) NW_QP(10)
                NW_QP(11) NW_QP(12) NW_QP(13) NW_QP(14) NW_QP(15) NW_QP(16) NW_QP(17) NW_QP(18) NW_QP(19)
                NW_QP(20) NW_QP(21) NW_QP(22) NW_QP(23) NW_QP(24) NW_QP(25) NW_QP(26) NW_QP(27) NW_QP(28)
                NW_QP(29) NW_QP(30) NW_QP(31)
            default: done = false; break;
            }
#undef NW_QP
        }
        if (!done && ks <= 31 && nw <= 8) {  // LDS tile <= 64 KB
            const size_t lds = (size_t)64 * (ks * 4 + 1) * sizeof(double);
            const long long ntiles64 = (rows + 63) / 64;
            dim3 g(grid_x(ntiles64, nb, 1024), (unsigned)nb), b(64 * nw);
            done = true;
#define NW_QT(K)                                                                                        \
    case K:                                                                                             \
        hipLaunchKernelGGL((k_qtransform_mfma<K>), g, b, lds, st, q, rows, nl, nout, ntiles64);        \
        break;
            switch (ks) {
                NW_QT(1) NW_QT(2) NW_QT(3) NW_QT(4) NW_QT(5) NW_QT(6) NW_QT(7) NW_QT(8) NW_QT(9) NW_QT(10)
                NW_QT(11) NW_QT(12) NW_QT(13) NW_QT(14) NW_QT(15) NW_QT(16) NW_QT(17) NW_QT(18) NW_QT(19)
                NW_QT(20) NW_QT(21) NW_QT(22) NW_QT(23) NW_QT(24) NW_QT(25) NW_QT(26) NW_QT(27) NW_QT(28)
                NW_QT(29) NW_QT(30) NW_QT(31) NW_QT(32)
            default: done = false; break;
            }
#undef NW_QT
        }
        if (!done) {
            dim3 grid((unsigned)((rows + 63) / 64), (nout + 63) / 64, (unsigned)nb);
            hipLaunchKernelGGL(k_qtransform, grid, dim3(256), 0, st, q, (int)rows, nl, nout);
        }
    }
}

void launch_qtransform(hipStream_t st, const double *X, const double *c, double *Y, long long rows, int nl, int nout)
{
    launch_qtransform_many(st, &X, &c, &Y, 1, rows, nl, nout);
}

// C = A B^T on FP64 MFMA (A: M x K, B: N x K, both K-contiguous), the shape
// of the J - K and J contractions of every co-density with the n^2 x n^2
// integral matrices.  A workgroup owns a 48 x 64 tile of C (4 waves, each 16
// columns x 3 row blocks of 16), K staged through LDS 32 at a time.  With
// ksplit > 1 workgroup z = (product, split s) sums K in [s kc, (s + 1) kc)
// only and writes its partial tile to part[(product ksplit + s)][M][N]
// (k_splitk_reduce adds the splits in order: the result does not depend on
// scheduling): at n_bas = 100 the products are 144 x 10^4 with K = 10^4, 471
// tiles of 313 chunks each, too few workgroups for 256 CUs.
__global__ __launch_bounds__(256) void k_gemm_nt_mfma(const GemmDesc *__restrict__ descs, int M, int N, int K, int lda,
                                                      int ldb, int ldc, int ksplit, int kc, double *__restrict__ part)
{
    constexpr int MT = 48, NTL = 64, BK = 32, LD = BK + 2;
    constexpr int PA = MT * BK / 2 / 256, PB = NTL * BK / 2 / 256;  // double2 per thread per chunk
    __shared__ double As[MT][LD];
    __shared__ double Bs[NTL][LD];
    // Workgroups go to the 8 XCDs round-robin by linear id; tile T is given
    // to linear id L so that each XCD runs a contiguous range of T, in which
    // the row blocks sharing a B tile (x fastest) meet in that XCD's L2.
    const unsigned total = gridDim.x * gridDim.y * gridDim.z;
    const unsigned L = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned xcd = L % 8, q8 = total / 8, r8 = total % 8;
    const unsigned T = xcd * q8 + min(xcd, r8) + L / 8;
    const int bx = (int)(T % gridDim.x), by = (int)(T / gridDim.x % gridDim.y), bz = (int)(T / (gridDim.x * gridDim.y));
    const int zb = bz / ksplit, sk = bz % ksplit;
    const GemmDesc g = descs[zb];
    const int kbeg = sk * kc, kend = min(K, kbeg + kc);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
    // row blocks of M vary fastest over the grid: the workgroups sharing a B tile run together
    const int m0 = bx * MT, n0 = by * NTL;
    const bool vec = (K % 2 == 0) && (lda % 2 == 0) && (ldb % 2 == 0);
    double2 ra[PA], rb[PB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int j = 0; j < PA; ++j) {
            const int e = tid + j * 256, r = e / (BK / 2), k = 2 * (e % (BK / 2)), gi = m0 + r, gk = k0 + k;
            double2 v{0.0, 0.0};
            if (gi < M) {
                const double *src = g.A + (size_t)gi * lda + gk;
                if (vec && gk + 1 < kend) v = *reinterpret_cast<const double2 *>(src);
                else {
                    if (gk < kend) v.x = src[0];
                    if (gk + 1 < kend) v.y = src[1];
                }
            }
            ra[j] = v;
        }
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            const int e = tid + j * 256, r = e / (BK / 2), k = 2 * (e % (BK / 2)), gj = n0 + r, gk = k0 + k;
            double2 v{0.0, 0.0};
            if (gj < N) {
                const double *src = g.B + (size_t)gj * ldb + gk;
                if (vec && gk + 1 < kend) v = *reinterpret_cast<const double2 *>(src);
                else {
                    if (gk < kend) v.x = src[0];
                    if (gk + 1 < kend) v.y = src[1];
                }
            }
            rb[j] = v;
        }
    };
    nw_d4 acc[3];
#pragma unroll
    for (int ms = 0; ms < 3; ++ms) acc[ms] = nw_d4{0.0, 0.0, 0.0, 0.0};
    fetch(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PA; ++j) {
            const int e = tid + j * 256, r = e / (BK / 2), k = 2 * (e % (BK / 2));
            As[r][k] = ra[j].x;
            As[r][k + 1] = ra[j].y;
        }
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            const int e = tid + j * 256, r = e / (BK / 2), k = 2 * (e % (BK / 2));
            Bs[r][k] = rb[j].x;
            Bs[r][k + 1] = rb[j].y;
        }
        __syncthreads();
        if (k0 + BK < kend) fetch(k0 + BK);  // next chunk in flight during the products
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            const double b = Bs[wave * 16 + lr][kk * 4 + lk];
#pragma unroll
            for (int ms = 0; ms < 3; ++ms)
                acc[ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[ms * 16 + lr][kk * 4 + lk], b, acc[ms], 0, 0, 0);
        }
    }
    const int col = n0 + wave * 16 + lr;
    if (col >= N) return;
    double *dst = part ? part + (size_t)bz * M * N : g.C;
    const size_t ld = part ? (size_t)N : (size_t)ldc;
#pragma unroll
    for (int ms = 0; ms < 3; ++ms)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = m0 + ms * 16 + 4 * r + lk;
            if (row < M) dst[(size_t)row * ld + col] = acc[ms][r];
        }
}

// C[z] = sum over s = 0 .. ksplit-1, in order, of part[(z ksplit + s)].
__global__ void k_splitk_reduce(const GemmDesc *__restrict__ descs, const double *__restrict__ part, int M, int N,
                                int ldc, int ksplit)
{
    const int z = blockIdx.y;
    double *C = descs[z].C;
    const size_t mn = (size_t)M * N;
    const double *p = part + (size_t)z * ksplit * mn;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn; e += (size_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int sk = 0; sk < ksplit; ++sk) v += p[sk * mn + e];
        C[(e / N) * ldc + e % N] = v;
    }
}

void launch_gemm_nt_mfma(hipStream_t st, int M, int N, int K, int lda, int ldb, int ldc, const GemmDesc *d_descs,
                         int nbatch)
{
    if (nbatch <= 0 || M <= 0 || N <= 0) return;
    for (int z0 = 0; z0 < nbatch; z0 += 65535) {
        const int chunk = nbatch - z0 > 65535 ? 65535 : nbatch - z0;
        dim3 grid((M + 47) / 48, (N + 63) / 64, chunk);
        hipLaunchKernelGGL(k_gemm_nt_mfma, grid, dim3(256), 0, st, d_descs + z0, M, N, K, lda, ldb, ldc, 1, K,
                           (double *)nullptr);
    }
}

int gemm_nt_splitk_count(int M, int N, int K, int nbatch, int ncu)
{
    const long long tiles = (long long)((M + 47) / 48) * ((N + 63) / 64) * nbatch;
    int ks = 1;
    while (ks < 8 && tiles * ks < 8ll * ncu && K / (2 * ks) >= 512) ks *= 2;
    return (long long)nbatch * ks <= 65535 ? ks : 1;
}

void launch_gemm_nt_splitk(hipStream_t st, int M, int N, int K, int lda, int ldb, int ldc, const GemmDesc *d_descs,
                           int nbatch, int ksplit, double *part)
{
    if (nbatch <= 0 || M <= 0 || N <= 0) return;
    const int kc = ((K + ksplit - 1) / ksplit + 31) / 32 * 32;
    dim3 grid((M + 47) / 48, (N + 63) / 64, (unsigned)(nbatch * ksplit));
    hipLaunchKernelGGL(k_gemm_nt_mfma, grid, dim3(256), 0, st, d_descs, M, N, K, lda, ldb, ldc, ksplit, kc, part);
    const size_t mn = (size_t)M * N;
    dim3 rgrid((unsigned)std::min<size_t>(2048, (mn + 255) / 256), (unsigned)nbatch);
    hipLaunchKernelGGL(k_splitk_reduce, rgrid, dim3(256), 0, st, d_descs, part, M, N, ldc, ksplit);
}

void launch_gemm(hipStream_t st, int ta, int tb, int M, int N, int K, int lda, int ldb, int ldc,
                 double alpha, double beta, const GemmDesc *d_descs, int nbatch)
{
    if (nbatch <= 0 || M <= 0 || N <= 0) return;
    if (ta == 0 && tb == 1 && alpha == 1.0 && beta == 0.0 && K >= 256 && (long long)N * K >= (1ll << 20) &&
        (N + 63) / 64 <= 65535) {
        // the large NT products (co-densities x integral matrices) on MFMA
        launch_gemm_nt_mfma(st, M, N, K, lda, ldb, ldc, d_descs, nbatch);
        return;
    }
    int nz = nbatch;
    while (nz > 0) {
        int chunk = nz > 65535 ? 65535 : nz;
        dim3 grid((N + 63) / 64, (M + 63) / 64, chunk);
        const GemmDesc *d = d_descs + (nbatch - nz);
        if (ta == 0 && tb == 0) hipLaunchKernelGGL((k_gemm<0, 0>), grid, dim3(256), 0, st, d, M, N, K, lda, ldb, ldc, alpha, beta);
        else if (ta == 0 && tb == 1) hipLaunchKernelGGL((k_gemm<0, 1>), grid, dim3(256), 0, st, d, M, N, K, lda, ldb, ldc, alpha, beta);
        else if (ta == 1 && tb == 0) hipLaunchKernelGGL((k_gemm<1, 0>), grid, dim3(256), 0, st, d, M, N, K, lda, ldb, ldc, alpha, beta);
        else hipLaunchKernelGGL((k_gemm<1, 1>), grid, dim3(256), 0, st, d, M, N, K, lda, ldb, ldc, alpha, beta);
        nz -= chunk;
    }
}

// ---------------------------------------------------------------------------
// One-sided Jacobi SVD of one n x n occupied-overlap matrix per 64-lane
// workgroup (n <= 64), FP64 throughout.  Produces singular values in
// descending order, U (left vectors; zero-singular-value columns completed
// orthonormally), V, phase = det(U) det(V), m = #{sigma <= tol} and the
// product of the nonzero singular values (build.rs:123-142, 401-424).
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return __shfl(v, 0, 64);
}

__global__ __launch_bounds__(64) void k_svd(const double *__restrict__ A, int n, double *__restrict__ U,
                                             double *__restrict__ V, double *__restrict__ sig,
                                             double *__restrict__ phase, int *__restrict__ mout,
                                             double *__restrict__ tsp, double tol)
{
    __shared__ double W[MAX_NOCC * MAX_NOCC];   // column-major: W[c*64 + r]
    __shared__ double Vs[MAX_NOCC * MAX_NOCC];
    __shared__ double Us[MAX_NOCC * MAX_NOCC];
    __shared__ double Vo[MAX_NOCC * MAX_NOCC];  // V columns in descending-sigma order
    __shared__ double sv[MAX_NOCC];
    __shared__ int ord[MAX_NOCC];
    __shared__ int piv_s;
    const int prob = blockIdx.x;
    const int r = threadIdx.x;
    const double *a = A + (size_t)prob * n * n;
    for (int c = 0; c < n; ++c) {
        W[c * 64 + r] = (r < n) ? a[r * n + c] : 0.0;
        Vs[c * 64 + r] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < 80; ++sweep) {
        int rotated = 0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double wp = W[p * 64 + r], wq = W[q * 64 + r];
                double al = wave_sum(wp * wp), be = wave_sum(wq * wq), ga = wave_sum(wp * wq);
                if (ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
                rotated = 1;
                double zeta = (be - al) / (2.0 * ga);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                W[p * 64 + r] = c * wp - s * wq;
                W[q * 64 + r] = s * wp + c * wq;
                double vp = Vs[p * 64 + r], vq = Vs[q * 64 + r];
                Vs[p * 64 + r] = c * vp - s * vq;
                Vs[q * 64 + r] = s * vp + c * vq;
                __syncthreads();
            }
        if (!rotated) break;
    }
    __syncthreads();
    if (r < n) {
        double acc = 0.0;
        for (int i = 0; i < n; ++i) acc += W[r * 64 + i] * W[r * 64 + i];
        sv[r] = sqrt(acc);
    }
    __syncthreads();
    if (r == 0) {
        for (int k = 0; k < n; ++k) ord[k] = k;
        for (int x = 0; x < n; ++x)
            for (int y = x + 1; y < n; ++y)
                if (sv[ord[y]] > sv[ord[x]]) { int t = ord[x]; ord[x] = ord[y]; ord[y] = t; }
    }
    __syncthreads();
    const double smax = n > 0 ? sv[ord[0]] : 0.0;
    // U columns are the normalised W columns; numerically zero singular values
    // get canonical left and right null vectors (Gram-Schmidt of unit vectors
    // against the preceding columns) so the u_k <-> v_k pairing is independent
    // of the rotation history.
    for (int k = 0; k < n; ++k) {
        int src = ord[k];
        double s = sv[src];
        bool good = s > 1e-13 * (smax > 0 ? smax : 1.0);
        if (good) {
            Us[k * 64 + r] = (r < n) ? W[src * 64 + r] / s : 0.0;
            Vo[k * 64 + r] = (r < n) ? Vs[src * 64 + r] : 0.0;
        } else {
            for (int side = 0; side < 2; ++side) {
                double *Q = side ? Vo : Us;
                for (int e = 0; e < n; ++e) {
                    double cand = (r == e) ? 1.0 : 0.0;
                    for (int pass = 0; pass < 2; ++pass)
                        for (int j = 0; j < k; ++j) {
                            double d = wave_sum(Q[j * 64 + r] * cand);
                            cand -= d * Q[j * 64 + r];
                        }
                    double nn = wave_sum(cand * cand);
                    if (nn > 1e-2) {
                        Q[k * 64 + r] = cand / sqrt(nn);
                        break;
                    }
                }
            }
        }
        __syncthreads();
    }
    // outputs: row-major U, V (n x n), sigma
    double *u = U + (size_t)prob * n * n, *v = V + (size_t)prob * n * n;
    if (r < n) {
        for (int k = 0; k < n; ++k) {
            u[r * n + k] = Us[k * 64 + r];
            v[r * n + k] = Vo[k * 64 + r];
        }
        sig[(size_t)prob * n + r] = sv[ord[r]];
    }
    __syncthreads();
    // det(U) and det(V) by partial-pivot LU; lane r owns row r (W reused for U, Vs permuted into W? use Us/W)
    double dets[2];
    for (int which = 0; which < 2; ++which) {
        // copy into W as row-major-by-lane: W[c*64 + r] = M[r][c]
        for (int c = 0; c < n; ++c) W[c * 64 + r] = (r < n) ? (which == 0 ? Us[c * 64 + r] : Vo[c * 64 + r]) : 0.0;
        __syncthreads();
        double sign = 1.0, d = 1.0;
        for (int k = 0; k < n; ++k) {
            if (r == 0) {
                int p = k;
                double pa = fabs(W[k * 64 + k]);
                for (int i = k + 1; i < n; ++i)
                    if (fabs(W[k * 64 + i]) > pa) { pa = fabs(W[k * 64 + i]); p = i; }
                piv_s = p;
            }
            __syncthreads();
            int p = piv_s;
            if (p != k) {
                sign = -sign;
                if (r < n) {
                    // swap rows k and p across columns (lane r handles column r)
                    double t = W[r * 64 + k];
                    W[r * 64 + k] = W[r * 64 + p];
                    W[r * 64 + p] = t;
                }
            }
            __syncthreads();
            double pivot = W[k * 64 + k];
            d *= pivot;
            if (pivot != 0.0 && r > k && r < n) {
                double f = W[k * 64 + r] / pivot;
                for (int c = k + 1; c < n; ++c) W[c * 64 + r] -= f * W[c * 64 + k];
            }
            __syncthreads();
        }
        dets[which] = sign * d;
        __syncthreads();
    }
    if (r == 0) {
        phase[prob] = dets[0] * dets[1];
        int m = 0;
        double prod = 1.0;
        for (int k = 0; k < n; ++k) {
            double s = sv[ord[k]];
            if (fabs(s) > tol) prod *= s; else ++m;
        }
        mout[prob] = m;
        tsp[prob] = prod;
    }
}

void launch_svd(hipStream_t st, const double *A, int n, int nprob, double *U, double *V, double *sig,
                double *phase, int *m, double *tsp, double tol)
{
    if (nprob <= 0) return;
    hipLaunchKernelGGL(k_svd, dim3(nprob), dim3(64), 0, st, A, n, U, V, sig, phase, m, tsp, tol);
}

// ---------------------------------------------------------------------------
// Co-densities M0, M1 (construct_m, build.rs:447-503), one problem per z.
// lt, gt: n x no rotated occupied orbitals (row-major, ld = no).
// ---------------------------------------------------------------------------
__global__ void k_codensity(const double *__restrict__ lt, const double *__restrict__ gt,
                            const double *__restrict__ sig, int n, int no, double tol,
                            double *__restrict__ M /* per prob: [2][n][n] */, long long in_stride,
                            long long sig_stride, long long out_stride)
{
    const int prob = blockIdx.z;
    const int mu = blockIdx.y, nu = blockIdx.x * blockDim.x + threadIdx.x;
    if (nu >= n) return;
    const double *L = lt + (size_t)prob * in_stride, *G = gt + (size_t)prob * in_stride;
    const double *s = sig + (size_t)prob * sig_stride;
    double m0 = 0.0, m1 = 0.0, zz = 0.0;
    for (int k = 0; k < no; ++k) {
        double g = G[mu * no + k], l = L[nu * no + k];
        if (fabs(s[k]) > tol) m0 += (g * (1.0 / s[k])) * l;
    }
    for (int k = 0; k < no; ++k) {
        if (fabs(s[k]) > tol) continue;
        double g = G[mu * no + k];
        zz += g * G[nu * no + k];
        m1 += g * L[nu * no + k];
    }
    double *out = M + (size_t)prob * out_stride;
    out[mu * n + nu] = m0 + zz + m1;
    out[n * n + mu * n + nu] = m1;
}

void launch_codensity(hipStream_t st, const double *lt, const double *gt, const double *sig, int n, int no,
                      double tol, double *M, int nprob, long long in_stride, long long sig_stride, long long out_stride)
{
    if (nprob <= 0) return;
    dim3 grid((n + 63) / 64, n, nprob);
    hipLaunchKernelGGL(k_codensity, grid, dim3(64), 0, st, lt, gt, sig, n, no, tol, M, in_stride, sig_stride,
                       out_stride);
}

// rowc = [Cx[:, no:] | Cw[:, :no]], colc = [Cx[:, :no] | Cw[:, no:]]
// (contraction_orbitals, build.rs:1008-1031) and the masked copies used for
// the (1 - m_i) subtractions of build_cx_xc (build.rs:944-991).
__global__ void k_contraction_orbitals(const double *const *__restrict__ cx_ptrs, const double *const *__restrict__ cw_ptrs,
                                       int n, int nm, int no, double *__restrict__ rowc, double *__restrict__ colc,
                                       long long out_stride)
{
    const int prob = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nm) return;
    const int mu = e / nm, j = e % nm;
    const double *Cx = cx_ptrs[prob], *Cw = cw_ptrs[prob];
    const int nv = nm - no;
    rowc[(size_t)prob * out_stride + e] = (j < nv) ? Cx[mu * nm + no + j] : Cw[mu * nm + (j - nv)];
    colc[(size_t)prob * out_stride + e] = (j < no) ? Cx[mu * nm + j] : Cw[mu * nm + j];
}

void launch_contraction_orbitals(hipStream_t st, const double *const *cx, const double *const *cw, int n, int nm,
                                 int no, double *rowc, double *colc, int nprob, long long out_stride)
{
    if (nprob <= 0) return;
    dim3 grid((n * nm + 255) / 256, nprob);
    hipLaunchKernelGGL(k_contraction_orbitals, grid, dim3(256), 0, st, cx, cw, n, nm, no, rowc, colc, out_stride);
}

// cx[mi] prefill = -(1-mi) rowc on columns < nv, 0 elsewhere;
// xc[mi] prefill = -(1-mi) colc on columns >= no, 0 elsewhere.
__global__ void k_prefill_cxxc(const double *__restrict__ rowc, const double *__restrict__ colc, int n, int nm, int no,
                               double *__restrict__ cx, double *__restrict__ xc, long long in_stride,
                               long long out_stride)
{
    const int prob = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * nm) return;
    const int j = e % nm;
    const int nv = nm - no;
    const size_t nn = (size_t)n * nm;
    double rc = rowc[prob * in_stride + e], cc = colc[prob * in_stride + e];
    double *cxp = cx + prob * out_stride, *xcp = xc + prob * out_stride;
    cxp[e] = (j < nv) ? -rc : 0.0;
    cxp[nn + e] = 0.0;
    xcp[e] = (j >= no) ? -cc : 0.0;
    xcp[nn + e] = 0.0;
}

void launch_prefill_cxxc(hipStream_t st, const double *rowc, const double *colc, int n, int nm, int no, double *cx,
                         double *xc, int nprob, long long in_stride, long long out_stride)
{
    if (nprob <= 0) return;
    dim3 grid((n * nm + 255) / 256, nprob);
    hipLaunchKernelGGL(k_prefill_cxxc, grid, dim3(256), 0, st, rowc, colc, n, nm, no, cx, xc, in_stride, out_stride);
}

// out = ca*a + cb*b + cc*c elementwise (null operands skipped), batch over y.
__global__ void k_lincomb(const LinDesc *__restrict__ d, long long n)
{
    const LinDesc s = d[blockIdx.y];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        double v = 0.0;
        if (s.a) v += s.ca * s.a[e];
        if (s.b) v += s.cb * s.b[e];
        if (s.c) v += s.cc * s.c[e];
        s.out[e] = v;
    }
}

void launch_lincomb(hipStream_t st, const LinDesc *d, int nb, long long n)
{
    if (nb <= 0 || n <= 0) return;
    long long gx = (n + 255) / 256;
    if (gx > 1024) gx = 1024;
    int y = 0;
    while (y < nb) {
        int chunk = nb - y > 65535 ? 65535 : nb - y;
        hipLaunchKernelGGL(k_lincomb, dim3((unsigned)gx, chunk), dim3(256), 0, st, d + y, n);
        y += chunk;
    }
}

// sum_{a,b} g[b,a] h[a,b] * scale  (einsum_ba_ab, src/maths/einsum.rs:25)
__global__ __launch_bounds__(256) void k_einsum_ba_ab(const TraceDesc *__restrict__ d, int n)
{
    const TraceDesc t = d[blockIdx.x];
    double acc = 0.0;
    for (int e = threadIdx.x; e < n * n; e += 256) {
        int a = e / n, b = e % n;
        acc += t.g[b * n + a] * t.h[a * n + b];
    }
    __shared__ double red[4];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) *t.out = t.scale * (red[0] + red[1] + red[2] + red[3]);
}

void launch_einsum(hipStream_t st, const TraceDesc *d, int nd, int n)
{
    if (nd <= 0) return;
    hipLaunchKernelGGL(k_einsum_ba_ab, dim3(nd), dim3(256), 0, st, d, n);
}

// eri_asym[a,b,c,d] = (ab|cd) - (ac|bd)  (src/read.rs:17-20)
__global__ void k_antisym(const double *__restrict__ eri, double *__restrict__ out, int n)
{
    const long long total = (long long)n * n * n * n;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        long long d = e % n, t = e / n;
        long long c = t % n; t /= n;
        long long b = t % n, a = t / n;
        out[e] = eri[e] - eri[((a * n + c) * n + b) * n + d];
    }
}

void launch_antisym(hipStream_t st, const double *eri, double *out, int n)
{
    hipLaunchKernelGGL(k_antisym, dim3(2048), dim3(256), 0, st, eri, out, n);
}

}  // namespace nw
