// Generalised Slater-Condon route on the GPU: calculate_hs_pair_naive
// (src/noci/hs.rs:372-402) with build_s_pair, one_electron,
// two_electron_same and two_electron_diff (src/noci/naive.rs:73-409), for
// batched determinant pairs, and the wicks.compare cross-check
// (compare_hs_pair_wicks_naive, src/noci/hs.rs:141-156, reported as in
// src/noci/matrix.rs:224-249).
//
// Per pair and spin: the occupied-orbital overlap S_ij = C_l^T S C_g is a
// gather from O_xw = C_x^T S C_w of the two parents (the determinant's
// occupied columns, ascending, are columns of its parent's C), its SVD
// (k_svd: U, V, sigma descending, phase = det U det V, zero count m, product
// of the nonzero sigma), the rotated orbitals C~_l = C_l U, C~_g = C_g V and
// the AO co-densities W = sum_k C~_g,k w_k C~_l,k^T (w_k = 1/sigma_k, 0 for
// zeros), P_i, P_j = C~_g,z C~_l,z^T of the first two zero singular values.
// The two-electron terms sum_{abcd} g[b,a] T[a,b,c,d] h[c,d]
// (einsum_ba_abcd_cd, src/maths/einsum.rs:318) are dot products of g^T with
// T vec(h): the T vec(h) of every pair in a chunk are one A B^T product of
// the stacked vec(h) with the n^2 x n^2 integral matrix (FP64 MFMA for
// n_bas >= 16), T = (ab||cd) for the same-spin terms and (ab|cd) for the
// opposite-spin one.  O(n_bas^4) per pair: a cross-check route, as in the
// reference.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nw_internal.h"

using namespace nw;
using namespace nwi;

namespace {

// one determinant's occupied orbitals per spin, ascending (occ_coeffs, naive.rs:20-32)
struct NaiveOcc {
    uint8_t o[2][MAX_NOCC];
};

// A[(q * 2 + sp)] (no x no, row-major) = O[(x, w, sp)][occ_l[a]][occ_g[b]] for request q.
__global__ void k_naive_sij(const double *__restrict__ O, const NaiveOcc *__restrict__ occ,
                            const uint64_t *__restrict__ req, const uint32_t *__restrict__ par, int nref, int nmo,
                            int no, int sp, double *__restrict__ A)
{
    const long long q = blockIdx.x;
    const uint64_t l = req[2 * q], g = req[2 * q + 1];
    const double *o = O + (((size_t)par[l] * nref + par[g]) * 2 + sp) * (size_t)nmo * nmo;
    const uint8_t *ol = occ[l].o[sp], *og = occ[g].o[sp];
    for (int e = threadIdx.x; e < no * no; e += blockDim.x) {
        const int a = e / no, b = e % no;
        A[(size_t)q * no * no + e] = o[(size_t)ol[a] * nmo + og[b]];
    }
}

// The rotated occupied orbitals of request q, spin sp: Lt = C_l U, Gt = C_g V
// (n x no row-major), then the co-densities W (zero count <= 1), P_i (>= 1
// zero), P_j (>= 2 zeros) as n x n matrices, rows mu of the ket side.
__global__ void k_naive_rho(const double *__restrict__ C, const NaiveOcc *__restrict__ occ,
                            const uint64_t *__restrict__ req, const uint32_t *__restrict__ par, int n, int nmo, int no,
                            int sp, const double *__restrict__ U, const double *__restrict__ V,
                            const double *__restrict__ sig, const int *__restrict__ mz, double tol,
                            double *__restrict__ Lt, double *__restrict__ Gt, double *__restrict__ W,
                            double *__restrict__ P1, double *__restrict__ P2)
{
    const long long q = blockIdx.x;
    const uint64_t l = req[2 * q], g = req[2 * q + 1];
    const double *cl = C + (size_t)par[l] * n * nmo, *cg = C + (size_t)par[g] * n * nmo;
    const uint8_t *ol = occ[l].o[sp], *og = occ[g].o[sp];
    const double *u = U + (size_t)q * no * no, *v = V + (size_t)q * no * no, *s = sig + (size_t)q * no;
    double *lt = Lt + (size_t)q * n * no, *gt = Gt + (size_t)q * n * no;
    for (int e = threadIdx.x; e < n * no; e += blockDim.x) {
        const int mu = e / no, k = e % no;
        double al = 0.0, ag = 0.0;
        for (int a = 0; a < no; ++a) {
            al += cl[(size_t)mu * nmo + ol[a]] * u[a * no + k];
            ag += cg[(size_t)mu * nmo + og[a]] * v[a * no + k];
        }
        lt[e] = al;
        gt[e] = ag;
    }
    __syncthreads();
    const int m = mz[q];
    // zero singular values sit at the end of the descending list
    const int z0 = no - m, z1 = no - m + 1;
    double *w = W + (size_t)q * n * n, *p1 = P1 + (size_t)q * n * n, *p2 = P2 + (size_t)q * n * n;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int mu = e / n, nu = e % n;
        double acc = 0.0;
        if (m <= 1)
            for (int k = 0; k < no; ++k)
                if (fabs(s[k]) > tol) acc += (gt[mu * no + k] * (1.0 / s[k])) * lt[nu * no + k];
        w[e] = acc;
        p1[e] = m >= 1 ? gt[mu * no + z0] * lt[nu * no + z0] : 0.0;
        p2[e] = m >= 2 ? gt[mu * no + z1] * lt[nu * no + z1] : 0.0;
    }
}

// Rows of the A B^T products (the vec(h) of each contraction): per request
// q, row 2q + sp of HA (same-spin, spin sp: W for zero counts 0 and 1, P_j
// for 2) and row q of HC (opposite-spin: the beta co-density, W or P_i).
__global__ void k_naive_hrows(const double *__restrict__ W, const double *__restrict__ P1,
                              const double *__restrict__ P2, const int *__restrict__ mz, int sp, long long nn,
                              double *__restrict__ HA, double *__restrict__ HC)
{
    const long long q = blockIdx.x;
    const int m = mz[q];
    const double *src = m == 2 ? P2 : W;
    double *dst = HA + ((size_t)q * 2 + sp) * nn;
    for (long long e = threadIdx.x; e < nn; e += blockDim.x) dst[e] = src[(size_t)q * nn + e];
    if (sp == 1) {
        const double *cs = m == 1 ? P1 : W;
        for (long long e = threadIdx.x; e < nn; e += blockDim.x) HC[(size_t)q * nn + e] = cs[(size_t)q * nn + e];
    }
}

__device__ inline double block_sum(double v, double *red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
    return t;
}

struct NaiveSpin {
    const double *W, *P1, *phase, *tsp;
    const int *m;
};

// H and S of request q (calculate_hs_pair_naive, src/noci/hs.rs:372-402).
// YA rows 2q + sp: (ab||cd) vec(h_sp); YC row q: (ab|cd) vec(rho_beta).
__global__ __launch_bounds__(256) void k_naive_final(NaiveSpin sa, NaiveSpin sb, const double *__restrict__ hao,
                                                     const double *__restrict__ YA, const double *__restrict__ YC,
                                                     int n, double enuc, double *__restrict__ H, double *__restrict__ S)
{
    __shared__ double red[4];
    const long long q = blockIdx.x;
    const size_t nn = (size_t)n * n;
    const NaiveSpin sp[2] = {sa, sb};
    double s_[2], fac[2], h1[2], h2[2];
    int m[2];
    for (int k = 0; k < 2; ++k) {
        m[k] = sp[k].m[q];
        fac[k] = sp[k].phase[q] * sp[k].tsp[q];  // phase * s_red
        s_[k] = m[k] == 0 ? fac[k] : 0.0;        // phase * product of all sigma
        // g: the bra-side co-density of each term (W, or P_i with zeros)
        const double *g = (m[k] == 0 ? sp[k].W : sp[k].P1) + (size_t)q * nn;
        const double *ya = YA + ((size_t)q * 2 + k) * nn;
        double a1 = 0.0, a2 = 0.0;
        for (size_t e = threadIdx.x; e < nn; e += blockDim.x) {
            const size_t a = e / n, b = e % n;
            const double gba = g[b * n + a];
            a1 += gba * hao[e];  // einsum_ba_ab(rho, h)
            a2 += gba * ya[e];   // einsum_ba_abcd_cd(g, (ab||cd), h)
        }
        a1 = block_sum(a1, red);
        a2 = block_sum(a2, red);
        h1[k] = m[k] <= 1 ? fac[k] * a1 : 0.0;
        h2[k] = m[k] == 0 ? 0.5 * fac[k] * a2 : (m[k] <= 2 ? fac[k] * a2 : 0.0);
    }
    double d = 0.0;
    {
        const double *ga = (m[0] == 0 ? sa.W : sa.P1) + (size_t)q * nn;
        const double *yc = YC + (size_t)q * nn;
        for (size_t e = threadIdx.x; e < nn; e += blockDim.x) {
            const size_t a = e / n, b = e % n;
            d += ga[b * n + a] * yc[e];
        }
        d = block_sum(d, red);
    }
    if (threadIdx.x == 0) {
        const double s = s_[0] * s_[1];
        const double hnuc = (m[0] == 0 && m[1] == 0) ? enuc * s : 0.0;
        const double h2ab = (m[0] <= 1 && m[1] <= 1) ? fac[0] * fac[1] * d : 0.0;
        H[q] = hnuc + (s_[1] * h1[0] + s_[0] * h1[1]) + (s_[1] * h2[0] + s_[0] * h2[1] + h2ab);
        S[q] = s;
    }
}

int naive_batched(nw_ctx *c, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq, double *Hout,
                  double *Sout)
{
    if (c->nref <= 0 || !c->d_c[0]) return fail(c, NW_E_STATE, "nw_set_refs first");
    if (!dets || !req || !Hout || !Sout) return fail(c, NW_E_ARG, "null arguments");
    if (nreq == 0) return NW_OK;
    if (is_device_ptr(dets) || is_device_ptr(req) || is_device_ptr(Hout) || is_device_ptr(Sout))
        return fail(c, NW_E_ARG, "host arrays expected");
    const int n = c->n, nm = c->nmo, nref = c->nref;
    const size_t nn = (size_t)n * n;
    // occupied orbitals per determinant and spin, ascending; counts must be the references'
    std::vector<NaiveOcc> occ(ndet);
    std::vector<uint32_t> par(ndet);
    for (uint64_t i = 0; i < ndet; ++i) {
        const nw_det &d = dets[i];
        if (d.parent >= (uint32_t)nref) return fail(c, NW_E_ARG, "determinant parent out of range");
        par[i] = d.parent;
        for (int sp = 0; sp < 2; ++sp) {
            const uint64_t *w = sp ? d.ob : d.oa;
            int k = 0;
            for (int b = 0; b < 128; ++b)
                if ((w[b >> 6] >> (b & 63)) & 1ull) {
                    if (b >= nm || k >= MAX_NOCC) return fail(c, NW_E_ARG, "occupation outside the MO space");
                    occ[i].o[sp][k++] = (uint8_t)b;
                }
            if (k != c->nocc[sp]) return fail(c, NW_E_ARG, "occupation count differs from the references'");
        }
    }
    for (uint64_t q = 0; q < nreq; ++q)
        if (req[2 * q] >= ndet || req[2 * q + 1] >= ndet) return fail(c, NW_E_ARG, "request index out of range");
    const int no[2] = {c->nocc[0], c->nocc[1]};
    // O_xw = C_x^T S C_w for every ordered parent pair and spin (nref^2 x 2 x nmo x nmo)
    double *SC = c->scratchA.get<double>((size_t)2 * nref * n * nm + (size_t)nref * nref * 2 * nm * nm);
    double *O = SC + (size_t)2 * nref * n * nm;
    Uploader up(c, 16u << 20);
    {
        std::vector<GemmDesc> g1, g2;
        for (int sp = 0; sp < 2; ++sp)
            for (int p = 0; p < nref; ++p)
                g1.push_back({c->d_S, c->d_c[sp] + (size_t)p * n * nm, SC + ((size_t)sp * nref + p) * n * nm});
        launch_gemm(c->st, 0, 0, n, nm, n, n, nm, nm, 1.0, 0.0, up.put(g1), (int)g1.size());
        for (int x = 0; x < nref; ++x)
            for (int w = 0; w < nref; ++w)
                for (int sp = 0; sp < 2; ++sp)
                    g2.push_back({c->d_c[sp] + (size_t)x * n * nm, SC + ((size_t)sp * nref + w) * n * nm,
                                  O + (((size_t)x * nref + w) * 2 + sp) * nm * nm});
        launch_gemm(c->st, 1, 0, nm, nm, n, nm, nm, nm, 1.0, 0.0, up.put(g2), (int)g2.size());
    }
    const NaiveOcc *docc = (const NaiveOcc *)to_device(c, c->dets_raw, occ.data(), ndet * sizeof(NaiveOcc));
    const uint32_t *dpar = (const uint32_t *)to_device(c, c->dets_packed, par.data(), ndet * sizeof(uint32_t));
    // requests per chunk: per spin the SVD problem and factors (3 k^2 + k + 3), the rotated orbitals
    // (2 n k) and the co-densities (3 n^2); the product rows and results (6 n^2 + 2); within ~512 MiB
    const size_t kmax = (size_t)std::max(1, std::max(no[0], no[1]));
    const size_t per_q = 2 * (3 * kmax * kmax + kmax + 3 + 2 * n * kmax + 3 * nn) + 6 * nn + 2;
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(nreq, (512ull << 20) / (per_q * 8)));
    double *buf = c->scratchB.get<double>(chunk * per_q + 64 * 32);  // + the 32-double alignment of 28 pieces
    std::vector<double> hh(nreq), ss(nreq);
    for (uint64_t q0 = 0; q0 < nreq; q0 += chunk) {
        const int R = (int)std::min<uint64_t>(chunk, nreq - q0);
        const uint64_t *dreq = (const uint64_t *)to_device(c, c->req_in, req + 2 * q0, (size_t)R * 16);
        double *p = buf;
        auto take = [&](size_t cnt) {
            double *r = p;
            p += (cnt + 31) & ~size_t(31);
            return r;
        };
        NaiveSpin ns[2];
        double *HA = take((size_t)R * 2 * nn), *HC = take((size_t)R * nn);
        double *YA = take((size_t)R * 2 * nn), *YC = take((size_t)R * nn);
        double *dH = take(R), *dS = take(R);
        for (int sp = 0; sp < 2; ++sp) {
            const int k = std::max(1, no[sp]);
            double *A = take((size_t)R * k * k), *U = take((size_t)R * k * k), *V = take((size_t)R * k * k);
            double *sig = take((size_t)R * k), *ph = take(R), *tsp = take(R);
            int *mz = (int *)take(R);
            double *Lt = take((size_t)R * n * k), *Gt = take((size_t)R * n * k);
            double *W = take((size_t)R * nn), *P1 = take((size_t)R * nn), *P2 = take((size_t)R * nn);
            if (no[sp] > 0)
                hipLaunchKernelGGL(k_naive_sij, dim3(R), dim3(64), 0, c->st, O, docc, dreq, dpar, nref, nm, no[sp],
                                   sp, A);
            launch_svd(c->st, A, no[sp], R, U, V, sig, ph, mz, tsp, c->tol);
            hipLaunchKernelGGL(k_naive_rho, dim3(R), dim3(256), 0, c->st, c->d_c[sp], docc, dreq, dpar, n, nm,
                               no[sp], sp, U, V, sig, mz, c->tol, Lt, Gt, W, P1, P2);
            hipLaunchKernelGGL(k_naive_hrows, dim3(R), dim3(256), 0, c->st, W, P1, P2, mz, sp, (long long)nn, HA, HC);
            ns[sp] = NaiveSpin{W, P1, ph, tsp, mz};
        }
        // T vec(h) for every row: (ab||cd) for the same-spin rows, (ab|cd) for the opposite-spin ones
        const std::vector<GemmDesc> ga = {{HA, c->d_easym, YA}}, gc = {{HC, c->d_eri, YC}};
        launch_gemm(c->st, 0, 1, 2 * R, (int)nn, (int)nn, (int)nn, (int)nn, (int)nn, 1.0, 0.0, up.put(ga), 1);
        launch_gemm(c->st, 0, 1, R, (int)nn, (int)nn, (int)nn, (int)nn, (int)nn, 1.0, 0.0, up.put(gc), 1);
        hipLaunchKernelGGL(k_naive_final, dim3(R), dim3(256), 0, c->st, ns[0], ns[1], c->d_h, YA, YC, n, c->enuc,
                           dH, dS);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(hh.data() + q0, dH, (size_t)R * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(ss.data() + q0, dS, (size_t)R * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
    }
    std::memcpy(Hout, hh.data(), nreq * 8);
    std::memcpy(Sout, ss.data(), nreq * 8);
    return NW_OK;
}

}  // namespace

extern "C" {

int nw_hs_naive_batched(nw_ctx *c, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq, double *H,
                        double *S)
{
    return guarded(c, [&]() -> int { return naive_batched(c, dets, ndet, req, nreq, H, S); });
}

int nw_hs_compare(nw_ctx *c, const nw_det *dets, uint64_t ndet, double *H, double *S, int mode, double *total,
                  double *max_elem)
{
    if (!c) return NW_E_ARG;
    if (!dets || !H || !S || !total || !max_elem || ndet == 0) return fail(c, NW_E_ARG, "null arguments");
    if (ndet > (1ull << 16)) return fail(c, NW_E_SHAPE, "cross-check limited to 65536 determinants");
    // Wick's H, S (calculate_hs_pair_wicks for every pair), then the naive
    // route for the same upper-triangle pairs
    int rc = nw_hs_matrix(c, dets, ndet, H, S, mode);
    if (rc != NW_OK) return rc;
    std::vector<uint64_t> req;
    req.reserve(ndet * (ndet + 1));
    for (uint64_t i = 0; i < ndet; ++i)
        for (uint64_t j = i; j < ndet; ++j) {
            req.push_back(i);
            req.push_back(j);
        }
    const uint64_t nreq = req.size() / 2;
    std::vector<double> hn(nreq), sn(nreq);
    rc = nw_hs_naive_batched(c, dets, ndet, req.data(), nreq, hn.data(), sn.data());
    if (rc != NW_OK) return rc;
    double td = 0.0, md = 0.0;
    for (uint64_t k = 0; k < nreq; ++k) {
        const uint64_t i = req[2 * k], j = req[2 * k + 1];
        const double dh = std::fabs(hn[k] - H[i * ndet + j]), ds = std::fabs(sn[k] - S[i * ndet + j]);
        td += dh + ds;
        md = std::max(md, std::max(dh, ds));
    }
    *total = td;
    *max_elem = md;
    return NW_OK;
}

}  // extern "C"
