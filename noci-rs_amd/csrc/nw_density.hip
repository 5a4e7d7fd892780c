// NOCI one-particle densities and the spin Fock matrices built from them on
// the GPU: the per-SNOCI-iteration refresh that precedes nw_update_fock.
//
// noci_density (src/noci/naive.rs:231-289) sums, over ordered determinant
// pairs (i, j) and per spin, c_i c_j ph_i ph_j s_other rho_sigma with the
// Loewdin-paired transition density of naive.rs:73-229:
//   SVD  C_i,occ^T S C_j,occ = U diag(s) V^T,  phase = det U det V,
//   m = 0: rho = phase s~ sum_k (C_j V)_k (C_i U)_k^T / s_k,
//   m = 1: rho = phase s~ (C_j V)_z (C_i U)_z^T for the zero column z,
//   m > 1: rho = 0,   s_sigma = phase prod s (0 when m > 0).
// Every determinant's occupied orbitals are columns of its parent's MO
// matrix, so with O_xw = C_x^T S C_w the occupied overlap is a gather and
//   rho = C_j[:, occ_j] Z C_i[:, occ_i]^T,   Z = V diag(w) U^T  (nocc x nocc).
// Z is scattered into an (nref nmo)^2 matrix T indexed by (parent j, orbital)
// x (parent i, orbital), and D = CC T CC^T with CC = [C_0 | C_1 | ...].
// fock (src/scf/kernels.rs:46): F_s = h + J[Da + Db] - K[D_s], evaluated as
// F_s = h + (ab||cd) D_s + (ab|cd) D_other and mirrored from its lower triangle.
#include <algorithm>
#include <cmath>
#include <vector>

#include "nw_internal.h"

using namespace nw;
using namespace nwi;

namespace {

// Occupied orbital indices (ascending) of each determinant and spin.
__global__ void k_occ_lists(const uint64_t *__restrict__ occbits /* ndet x 4 */, long long ndet, int no_a, int no_b,
                            int nmo, uint8_t *__restrict__ occ /* ndet x 2 x MAX_NOCC */, int *__restrict__ err)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ndet) return;
    for (int sp = 0; sp < 2; ++sp) {
        uint64_t lo = occbits[i * 4 + 2 * sp], hi = occbits[i * 4 + 2 * sp + 1];
        const int want = sp ? no_b : no_a;
        int k = 0;
        while (lo && k < MAX_NOCC) {
            occ[(i * 2 + sp) * MAX_NOCC + k++] = (uint8_t)(__ffsll((long long)lo) - 1);
            lo &= lo - 1;
        }
        while (hi && k < MAX_NOCC) {
            occ[(i * 2 + sp) * MAX_NOCC + k++] = (uint8_t)(63 + __ffsll((long long)hi));
            hi &= hi - 1;
        }
        if (k != want || lo || hi || (k > 0 && occ[(i * 2 + sp) * MAX_NOCC + k - 1] >= nmo)) atomicOr(err, 1);
    }
}

// Socc of pairs (i, j) = (p0 + q) / nd ... one problem per pair of this chunk.
__global__ void k_pair_socc(const double *__restrict__ O /* nref x nref x nmo x nmo */, const uint32_t *__restrict__ par,
                            const uint8_t *__restrict__ occ, int sp, int no, int nmo, int nref, long long ndet,
                            long long p0, long long np, double *__restrict__ A /* np x no x no */)
{
    const long long q = blockIdx.x;
    if (q >= np) return;
    const long long pr = p0 + q, i = pr / ndet, j = pr % ndet;
    const double *Oxw = O + ((size_t)par[i] * nref + par[j]) * nmo * nmo;
    const uint8_t *oi = occ + (i * 2 + sp) * MAX_NOCC, *oj = occ + (j * 2 + sp) * MAX_NOCC;
    for (int e = threadIdx.x; e < no * no; e += blockDim.x) {
        const int a = e / no, b = e % no;
        A[(size_t)q * no * no + e] = Oxw[(size_t)oi[a] * nmo + oj[b]];
    }
}

// Z = V diag(w) U^T of both spins of each pair, scattered (atomically) into
// T[(par_j, occ_j[a]), (par_i, occ_i[b])].
__global__ void k_pair_z(const double *__restrict__ U0, const double *__restrict__ V0, const double *__restrict__ s0,
                         const double *__restrict__ ph0, const int *__restrict__ m0, const double *__restrict__ t0,
                         const double *__restrict__ U1, const double *__restrict__ V1, const double *__restrict__ s1,
                         const double *__restrict__ ph1, const int *__restrict__ m1, const double *__restrict__ t1,
                         const uint32_t *__restrict__ par, const uint8_t *__restrict__ occ,
                         const double *__restrict__ coef, const double *__restrict__ dph, int no_a, int no_b, int nmo,
                         int nref, long long ndet, long long p0, long long np, double tol, double *__restrict__ Ta,
                         double *__restrict__ Tb)
{
    const long long q = blockIdx.x;
    if (q >= np) return;
    const long long pr = p0 + q, i = pr / ndet, j = pr % ndet;
    const double c = coef[i] * coef[j] * dph[i] * dph[j];
    const double sa = m0[q] == 0 ? ph0[q] * t0[q] : 0.0;
    const double sb = m1[q] == 0 ? ph1[q] * t1[q] : 0.0;
    const long long ldt = (long long)nref * nmo;
    for (int sp = 0; sp < 2; ++sp) {
        const int no = sp ? no_b : no_a;
        const int m = sp ? m1[q] : m0[q];
        const double f = c * (sp ? sa : sb) * (sp ? ph1[q] * t1[q] : ph0[q] * t0[q]);
        if (m > 1 || f == 0.0) continue;
        const double *U = (sp ? U1 : U0) + (size_t)q * no * no, *V = (sp ? V1 : V0) + (size_t)q * no * no;
        const double *sv = (sp ? s1 : s0) + (size_t)q * no;
        double *T = sp ? Tb : Ta;
        const uint8_t *oi = occ + (i * 2 + sp) * MAX_NOCC, *oj = occ + (j * 2 + sp) * MAX_NOCC;
        const long long r0 = (long long)par[j] * nmo, c0 = (long long)par[i] * nmo;
        for (int e = threadIdx.x; e < no * no; e += blockDim.x) {
            const int a = e / no, b = e % no;  // Z[a][b], a: ket orbital, b: bra orbital
            double z = 0.0;
            for (int k = 0; k < no; ++k) {
                const bool zero = fabs(sv[k]) <= tol;
                const double w = m == 0 ? 1.0 / sv[k] : (zero ? 1.0 : 0.0);
                z += V[a * no + k] * w * U[b * no + k];
            }
            unsafeAtomicAdd(T + (r0 + oj[a]) * ldt + c0 + oi[b], f * z);
        }
    }
}

// y = A x for an n2 x n2 row-major A (one wave per row).
__global__ void k_gemv_rows(const double *__restrict__ A, const double *__restrict__ x, long long n2,
                            double *__restrict__ y)
{
    const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n2) return;
    const double *a = A + (size_t)row * n2;
    double acc = 0.0;
    for (long long k = lane; k < n2; k += 64) acc += a[k] * x[k];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) y[row] = acc;
}

// F = h + J1 + J2 on the lower triangle, mirrored to the upper one.
__global__ void k_fock_mirror(const double *__restrict__ h, const double *__restrict__ g1, const double *__restrict__ g2,
                              int n, double *__restrict__ F)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * n) return;
    const int p = e / n, q = e % n;
    const int lo = p >= q ? p * n + q : q * n + p;  // (max, min)
    F[e] = h[lo] + g1[lo] + g2[lo];
}

}  // namespace

extern "C" int nw_noci_fock(nw_ctx *c, const nw_det *dets, uint64_t ndet, const double *coef, double tol, double *fa,
                            double *fb, double *da, double *db)
{
    return guarded(c, [&]() -> int {
        if (c->nref == 0) return fail(c, NW_E_STATE, "nw_set_refs first");
        if (!dets || !coef || !fa || !fb || ndet == 0) return fail(c, NW_E_ARG, "null arguments");
        if (ndet > (1ull << 20)) return fail(c, NW_E_SHAPE, "too many determinants");
        const int n = c->n, nm = c->nmo, nref = c->nref;
        const int no[2] = {c->nocc[0], c->nocc[1]};
        if (no[0] < 1 || no[1] < 1) return fail(c, NW_E_SHAPE, "both spins need at least one electron");
        const size_t nn = (size_t)n * n, nnm = (size_t)n * nm, ld = (size_t)nref * nm;
        Uploader up(c, 64u << 20);
        // determinant records -> parents, phases, occupied lists
        std::vector<nw_det> hd(ndet);
        HIPCHK(hipMemcpyAsync(hd.data(), dets, ndet * sizeof(nw_det), hipMemcpyDefault, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        std::vector<uint32_t> par(ndet);
        std::vector<double> dph(ndet);
        std::vector<uint64_t> bits(ndet * 4);
        for (uint64_t i = 0; i < ndet; ++i) {
            if ((int)hd[i].parent >= nref) throw NwError(NW_E_ARG, "determinant parent out of range");
            par[i] = hd[i].parent;
            dph[i] = hd[i].pha * hd[i].phb;
            bits[4 * i] = hd[i].oa[0]; bits[4 * i + 1] = hd[i].oa[1];
            bits[4 * i + 2] = hd[i].ob[0]; bits[4 * i + 3] = hd[i].ob[1];
        }
        const uint32_t *dpar = up.put(par);
        const double *ddph = up.put(dph);
        const uint64_t *dbits = up.put(bits);
        const double *dcoef = (const double *)to_device(c, c->outH, coef, ndet * 8);
        uint8_t *docc = dalloc<uint8_t>(ndet * 2 * MAX_NOCC);
        std::vector<void *> owned = {docc};
        auto release = [&]() {
            for (void *p : owned) (void)hipFree(p);
        };
        try {
            int *dflag = c->flag.get<int>(1);
            HIPCHK(hipMemsetAsync(dflag, 0, sizeof(int), c->st));
            hipLaunchKernelGGL(k_occ_lists, dim3((unsigned)((ndet + 255) / 256)), dim3(256), 0, c->st, dbits,
                               (long long)ndet, no[0], no[1], nm, docc, dflag);
            int flag = 0;
            HIPCHK(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            if (flag) throw NwError(NW_E_ARG, "determinant occupations must have the references' electron counts within nmo");
            // O_xw = C_x^T S C_w per spin (nref^2 x nmo x nmo), and CC = [C_0 | C_1 | ...] (n x nref nmo)
            double *O = dalloc<double>((size_t)2 * nref * nref * nm * nm), *SC = dalloc<double>((size_t)nref * nnm);
            double *CC = dalloc<double>((size_t)2 * n * ld), *T = dalloc<double>((size_t)2 * ld * ld);
            double *Y = dalloc<double>(ld * n), *D = dalloc<double>(2 * nn), *G = dalloc<double>(4 * nn);
            owned.insert(owned.end(), {O, SC, CC, T, Y, D, G});
            for (int sp = 0; sp < 2; ++sp) {
                const double *C = c->d_c[sp];
                std::vector<GemmDesc> gsc, go, gcc;
                for (int w = 0; w < nref; ++w) gsc.push_back({c->d_S, C + (size_t)w * nnm, SC + (size_t)w * nnm});
                for (int x = 0; x < nref; ++x)
                    for (int w = 0; w < nref; ++w)
                        go.push_back({C + (size_t)x * nnm, SC + (size_t)w * nnm,
                                      O + (((size_t)sp * nref + x) * nref + w) * nm * nm});
                launch_gemm(c->st, 0, 0, n, nm, n, n, nm, nm, 1.0, 0.0, up.put(gsc), (int)gsc.size());
                launch_gemm(c->st, 1, 0, nm, nm, n, nm, nm, nm, 1.0, 0.0, up.put(go), (int)go.size());
                // CC rows mu: [C_0[mu, :] | C_1[mu, :] | ...]
                for (int r = 0; r < nref; ++r)
                    HIPCHK(hipMemcpy2DAsync(CC + (size_t)sp * n * ld + (size_t)r * nm, ld * 8, C + (size_t)r * nnm,
                                            nm * 8, nm * 8, n, hipMemcpyDeviceToDevice, c->st));
            }
            HIPCHK(hipMemsetAsync(T, 0, 2 * ld * ld * 8, c->st));
            // pairs in chunks: Socc gather, SVD per spin, Z scatter
            const long long npairs = (long long)ndet * (long long)ndet;
            const int nomax = std::max(1, std::max(no[0], no[1]));
            const long long chunk = std::min<long long>(npairs, (1LL << 23) / (nomax * nomax));
            double *A = dalloc<double>((size_t)chunk * nomax * nomax * 2);
            double *U = dalloc<double>((size_t)chunk * nomax * nomax * 2), *V = dalloc<double>((size_t)chunk * nomax * nomax * 2);
            double *sg = dalloc<double>((size_t)chunk * nomax * 2), *ph = dalloc<double>((size_t)chunk * 2);
            double *tp = dalloc<double>((size_t)chunk * 2);
            int *mm = dalloc<int>((size_t)chunk * 2);
            owned.insert(owned.end(), {A, U, V, sg, ph, tp, (double *)mm});
            for (long long p0 = 0; p0 < npairs; p0 += chunk) {
                const long long np = std::min(chunk, npairs - p0);
                for (int sp = 0; sp < 2; ++sp) {
                    const size_t off = (size_t)sp * chunk * nomax * nomax;
                    hipLaunchKernelGGL(k_pair_socc, dim3((unsigned)np), dim3(64), 0, c->st,
                                       O + (size_t)sp * nref * nref * nm * nm, dpar, docc, sp, no[sp], nm, nref,
                                       (long long)ndet, p0, np, A + off);
                    launch_svd(c->st, A + off, no[sp], (int)np, U + off, V + off, sg + (size_t)sp * chunk * nomax,
                               ph + (size_t)sp * chunk, mm + (size_t)sp * chunk, tp + (size_t)sp * chunk, tol);
                }
                const size_t o1 = (size_t)chunk * nomax * nomax;
                hipLaunchKernelGGL(k_pair_z, dim3((unsigned)np), dim3(64), 0, c->st, U, V, sg, ph, mm, tp, U + o1,
                                   V + o1, sg + (size_t)chunk * nomax, ph + chunk, mm + chunk, tp + chunk, dpar, docc,
                                   dcoef, ddph, no[0], no[1], nm, nref, (long long)ndet, p0, np, tol, T, T + ld * ld);
            }
            // D_s = CC T_s CC^T, then the Fock matrices
            for (int sp = 0; sp < 2; ++sp) {
                const double *CCs = CC + (size_t)sp * n * ld;
                std::vector<GemmDesc> g1 = {{T + (size_t)sp * ld * ld, CCs, Y}};
                launch_gemm(c->st, 0, 1, (int)ld, n, (int)ld, (int)ld, (int)ld, n, 1.0, 0.0, up.put(g1), 1);
                std::vector<GemmDesc> g2 = {{CCs, Y, D + (size_t)sp * nn}};
                launch_gemm(c->st, 0, 0, n, n, (int)ld, (int)ld, n, n, 1.0, 0.0, up.put(g2), 1);
            }
            // G = [(ab||cd) Da, (ab||cd) Db, (ab|cd) Da, (ab|cd) Db]
            const long long n2 = (long long)nn;
            for (int k = 0; k < 4; ++k)
                hipLaunchKernelGGL(k_gemv_rows, dim3((unsigned)((n2 + 3) / 4)), dim3(256), 0, c->st,
                                   k < 2 ? c->d_easym : c->d_eri, D + (size_t)(k & 1) * nn, n2, G + (size_t)k * nn);
            const bool fadev = is_device_ptr(fa), fbdev = is_device_ptr(fb);
            double *F = c->fockbuf.get<double>(2 * nn);
            double *Fa = fadev ? fa : F, *Fb = fbdev ? fb : F + nn;
            const unsigned gb = (unsigned)((nn + 255) / 256);
            hipLaunchKernelGGL(k_fock_mirror, dim3(gb), dim3(256), 0, c->st, c->d_h, G, G + 3 * nn, n, Fa);
            hipLaunchKernelGGL(k_fock_mirror, dim3(gb), dim3(256), 0, c->st, c->d_h, G + nn, G + 2 * nn, n, Fb);
            HIPCHK(hipGetLastError());
            if (!fadev) HIPCHK(hipMemcpyAsync(fa, Fa, nn * 8, hipMemcpyDeviceToHost, c->st));
            if (!fbdev) HIPCHK(hipMemcpyAsync(fb, Fb, nn * 8, hipMemcpyDeviceToHost, c->st));
            if (da) HIPCHK(hipMemcpyAsync(da, D, nn * 8, hipMemcpyDefault, c->st));
            if (db) HIPCHK(hipMemcpyAsync(db, D + nn, nn * 8, hipMemcpyDefault, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
        } catch (...) {
            (void)hipStreamSynchronize(c->st);
            release();
            throw;
        }
        release();
        return NW_OK;
    });
}
