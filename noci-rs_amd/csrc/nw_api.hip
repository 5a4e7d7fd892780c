// C ABI of the engine (include/nwick.h): context, per-reference-pair build
// orchestration and the evaluation entry points.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "nw_internal.h"

using namespace nw;
using namespace nwi;

namespace {



void gemm(nw_ctx *c, Uploader &up, int ta, int tb, int M, int N, int K, int lda, int ldb, int ldc, double alpha,
          double beta, const std::vector<GemmDesc> &descs)
{
    if (descs.empty()) return;
    GemmDesc *d = up.put(descs);
    launch_gemm(c->st, ta, tb, M, N, K, lda, ldb, ldc, alpha, beta, d, (int)descs.size());
}

// NW_TRANSFORM_PASSES=3 runs the bf16x3 transform as three launches (half
// transform + two leading passes; even n_bas <= 112, nmo <= n_bas).  It
// moves 3/4 of the four-pass traffic but measured slower at n_bas = 100
// (900 vs 753 us per tensor: the half transform is MFMA/LDS-bound at one
// workgroup per CU, the leading passes read 256-byte row pieces), so the
// four quarter passes are the default.
bool three_launch_transform(int n, int nm)
{
    const char *e = std::getenv("NW_TRANSFORM_PASSES");
    if (!e || std::strcmp(e, "3") != 0) return false;
    return n % 2 == 0 && n <= 112 && nm <= n;
}


}  // namespace


extern "C" {

int nw_create(nw_ctx **out, int device, int n_bas, const double *S, const double *h, const double *eri, double enuc)
{
    if (!out || !S || !h || !eri || n_bas <= 0) return NW_E_ARG;
    if (n_bas > MAX_NMO) return NW_E_SHAPE;
    nw_ctx *c = new nw_ctx();
    c->device = device;
    c->n = n_bas;
    c->enuc = enuc;
    int rc = guarded(c, [&]() -> int {
        HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
        const size_t n2 = (size_t)n_bas * n_bas, n4 = n2 * n2;
        c->d_S = dalloc<double>(n2);
        c->d_h = dalloc<double>(n2);
        c->d_eri = dalloc<double>(n4);
        c->d_easym = dalloc<double>(n4);
        HIPCHK(hipMemcpyAsync(c->d_S, S, n2 * 8, hipMemcpyDefault, c->st));
        HIPCHK(hipMemcpyAsync(c->d_h, h, n2 * 8, hipMemcpyDefault, c->st));
        HIPCHK(hipMemcpyAsync(c->d_eri, eri, n4 * 8, hipMemcpyDefault, c->st));
        launch_antisym(c->st, c->d_eri, c->d_easym, n_bas);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->st));
        return NW_OK;
    });
    if (rc != NW_OK) {
        std::fprintf(stderr, "nw_create: %s\n", c->err.c_str());
        delete c;
        return rc;
    }
    *out = c;
    return NW_OK;
}

int nw_set_refs(nw_ctx *c, int nref, int nmo, const double *ca, const double *cb, const uint64_t *occ)
{
    return guarded(c, [&]() -> int {
        if (!ca || !cb || !occ || nref <= 0 || nmo <= 0) return fail(c, NW_E_ARG, "null reference data");
        if (nmo > MAX_NMO || nmo > c->n) return fail(c, NW_E_SHAPE, "nmo must be <= n_bas and <= 128");
        if (nref > 65535) return fail(c, NW_E_SHAPE, "too many references");
        int no[2] = {-1, -1};
        for (int r = 0; r < nref; ++r)
            for (int sp = 0; sp < 2; ++sp) {
                uint64_t lo = occ[r * 4 + sp * 2], hi = occ[r * 4 + sp * 2 + 1];
                int cnt = __builtin_popcountll(lo) + __builtin_popcountll(hi);
                // occupied-first: bits 0..cnt-1 set exactly
                bool first = true;
                for (int b = 0; b < 128; ++b) {
                    bool set = b < 64 ? (lo >> b) & 1 : (hi >> (b - 64)) & 1;
                    if (set != (b < cnt)) first = false;
                }
                if (!first) return fail(c, NW_E_ORDER, "reference occupations must be occupied-first");
                if (no[sp] < 0) no[sp] = cnt;
                else if (no[sp] != cnt) return fail(c, NW_E_SHAPE, "all references need the same electron count per spin");
            }
        if (no[0] > MAX_NOCC || no[1] > MAX_NOCC || no[0] >= nmo || no[1] >= nmo)
            return fail(c, NW_E_SHAPE, "nocc must be < nmo and <= 64");
        const size_t len = (size_t)nref * c->n * nmo;
        for (int sp = 0; sp < 2; ++sp) {
            if (c->d_c[sp]) HIPCHK(hipFree(c->d_c[sp]));
            c->d_c[sp] = dalloc<double>(len);
            HIPCHK(hipMemcpyAsync(c->d_c[sp], sp ? cb : ca, len * 8, hipMemcpyDefault, c->st));
        }
        HIPCHK(hipStreamSynchronize(c->st));
        c->nref = nref;
        c->nmo = nmo;
        c->nocc[0] = no[0];
        c->nocc[1] = no[1];
        c->built = false;
        return NW_OK;
    });
}

int nw_build(nw_ctx *c, double tol, const uint8_t *pair_mask, nw_pair_meta *meta_out)
{
    return guarded(c, [&]() -> int { return build_impl(c, tol, pair_mask, meta_out, nullptr); });
}

}  // extern "C"

namespace nwi {

// Body of nw_build.  With `sizes` set, only the occupied-overlap SVDs run:
// sizes receives, per masked pair in (x, w) order, the slab length (in
// doubles) its block would take, and no slab is written.
int build_impl(nw_ctx *c, double tol, const uint8_t *pair_mask, nw_pair_meta *meta_out, std::vector<long long> *sizes)
{
    {
        if (c->nref == 0) return fail(c, NW_E_STATE, "nw_set_refs first");
        const int n = c->n, nm = c->nmo, nref = c->nref;
        const size_t nn = (size_t)n * n, nnm = (size_t)n * nm, nm2 = (size_t)nm * nm;
        std::vector<int> px, pw;
        for (int x = 0; x < nref; ++x)
            for (int w = 0; w < nref; ++w)
                if (!pair_mask || pair_mask[x * nref + w]) { px.push_back(x); pw.push_back(w); }
        const int P = (int)px.size();
        c->built = false;
        c->build_gen++;
        c->fock_gen++;
        c->tol = tol;
        HIPCHK(hipEventRecord(c->ev[0], c->st));
        Uploader up(c, 64u << 20);
        // ---- scratch per (pair, spin) -------------------------------------
        int nomax = std::max(c->nocc[0], c->nocc[1]);
        // The build's temporaries live in one grow-only arena of the context:
        // in steady state a build allocates and frees nothing (hipFree waits
        // for the whole device), so builds of different contexts overlap.
        const size_t Pz = (size_t)P, nz = (size_t)n, noz = (size_t)nomax;
        enum { T_TOCC, T_SOCC, T_U, T_V, T_SIG, T_PHASE, T_TSP, T_M, T_LT, T_GT, T_ROWC, T_COLC, T_MS, T_MTS, T_Q, T_R,
               T_Z, T_JK, T_JC, T_HXC, T_JKXC, T_META, T_COUNT };
        const size_t req[T_COUNT] = {2 * Pz * nz * noz, 2 * Pz * noz * noz, 2 * Pz * noz * noz, 2 * Pz * noz * noz,
                                     2 * Pz * noz, 2 * Pz, 2 * Pz, Pz, 2 * Pz * nz * noz, 2 * Pz * nz * noz,
                                     2 * Pz * nnm, 2 * Pz * nnm, 4 * Pz * nn, 4 * Pz * nn, 4 * Pz * nnm, 4 * Pz * nnm,
                                     2 * Pz * nm2, 4 * Pz * nn, 4 * Pz * nn, 4 * Pz * nnm, 8 * Pz * nnm, Pz * M_LEN};
        size_t need = 0;
        for (size_t r : req) need += (r + 31) & ~size_t(31);
        double *arena = c->buildscratch.get<double>(need);
        double *tp[T_COUNT];
        for (size_t k = 0, cur = 0; k < T_COUNT; ++k) {
            tp[k] = arena + cur;
            cur += (req[k] + 31) & ~size_t(31);
        }
        double *Tocc = tp[T_TOCC], *Socc = tp[T_SOCC];
        double *U = tp[T_U], *V = tp[T_V];
        double *sig = tp[T_SIG], *phase = tp[T_PHASE], *tsp = tp[T_TSP];
        int *mdev = (int *)tp[T_M];
        double *Lt = tp[T_LT], *Gt = tp[T_GT];
        // co-densities and the cx/xc contractions stay resident for nw_update_fock
        double *M = c->keepM.get<double>((size_t)4 * P * nn);  // [(p*2+sp)*2+mi][n][n]
        double *rowc = tp[T_ROWC], *colc = tp[T_COLC];
        double *cx = c->keepCX.get<double>((size_t)4 * P * nnm), *xc = c->keepXC.get<double>((size_t)4 * P * nnm);
        double *MS = tp[T_MS], *MtS = tp[T_MTS];
        double *Q = tp[T_Q], *R = tp[T_R];
        double *Z = tp[T_Z];
        double *JK = tp[T_JK], *Jc = tp[T_JC];
        double *hxc = tp[T_HXC];
        double *JKxc = tp[T_JKXC];
        double *dmeta = tp[T_META];
        HIPCHK(hipMemsetAsync(dmeta, 0, (size_t)P * M_LEN * 8, c->st));
        auto Cref = [&](int sp, int r) { return c->d_c[sp] + (size_t)r * nnm; };
        auto prob = [&](int p, int sp) { return p * 2 + sp; };
        // spin-local index for no-dependent arrays: (sp*P + p)
        for (int sp = 0; sp < 2; ++sp) {
            const int no = c->nocc[sp];
            std::vector<GemmDesc> g1, g2;
            for (int p = 0; p < P; ++p) {
                size_t q = (size_t)sp * P + p;
                g1.push_back({c->d_S, Cref(sp, pw[p]), Tocc + q * n * nomax});
                g2.push_back({Cref(sp, px[p]), Tocc + q * n * nomax, Socc + q * nomax * nomax});
            }
            gemm(c, up, 0, 0, n, no, n, n, nm, no, 1.0, 0.0, g1);   // S Cw,occ
            gemm(c, up, 1, 0, no, no, n, nm, no, no, 1.0, 0.0, g2);  // Cx,occ^T (S Cw,occ)
        }
        // The SVD kernel reads problem q at A + q*no*no; repack into U as staging.
        // With equal electron counts per spin both spins' problems go in one
        // launch (one wave per problem: twice the problems run side by side).
        if (c->nocc[0] == c->nocc[1]) {
            const int no = c->nocc[0];
            std::vector<LinDesc> cp;
            double *stage = R;  // 4P*n*nm >= 2P*no*no
            for (int q = 0; q < 2 * P; ++q)
                cp.push_back({Socc + (size_t)q * nomax * nomax, nullptr, nullptr, stage + (size_t)q * no * no, 1.0, 0.0,
                              0.0});
            launch_lincomb(c->st, up.put(cp), 2 * P, (long long)no * no);
            // no == nomax: problem sp * P + p lands where the per-spin launches put it
            launch_svd(c->st, stage, no, 2 * P, U, V, sig, phase, mdev, tsp, tol);
            std::vector<int> mh(2 * P);
            HIPCHK(hipMemcpyAsync(mh.data(), mdev, 2 * P * sizeof(int), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            c->pinfo.assign((size_t)nref * nref, PairInfo{});
            for (int p = 0; p < P; ++p) {
                PairInfo &pi = c->pinfo[(size_t)px[p] * nref + pw[p]];
                pi.ma = mh[p];
                pi.mb = mh[P + p];
            }
        }
        for (int sp = 0; sp < 2 && c->nocc[0] != c->nocc[1]; ++sp) {
            const int no = c->nocc[sp];
            std::vector<LinDesc> cp;
            double *stage = R;  // reuse R as staging (size 4P*n*nm >= P*no*no)
            for (int p = 0; p < P; ++p) {
                size_t q = (size_t)sp * P + p;
                cp.push_back({Socc + q * nomax * nomax, nullptr, nullptr, stage + (size_t)p * no * no, 1.0, 0.0, 0.0});
            }
            launch_lincomb(c->st, up.put(cp), P, (long long)no * no);
            double *Usp = U + (size_t)sp * P * nomax * nomax, *Vsp = V + (size_t)sp * P * nomax * nomax;
            launch_svd(c->st, stage, no, P, Usp, Vsp, sig + (size_t)sp * P * nomax, phase + (size_t)sp * P,
                       mdev + 0, tsp + (size_t)sp * P, tol);
            // copy m out per spin (mdev has P ints)
            std::vector<int> mh(P);
            HIPCHK(hipMemcpyAsync(mh.data(), mdev, P * sizeof(int), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            if (sp == 0) c->pinfo.assign((size_t)nref * nref, PairInfo{});
            for (int p = 0; p < P; ++p) {
                PairInfo &pi = c->pinfo[(size_t)px[p] * nref + pw[p]];
                if (sp == 0) pi.ma = mh[p]; else pi.mb = mh[p];
            }
        }
        std::vector<double> ph_h(2 * P), tsp_h(2 * P), sig_h((size_t)2 * P * nomax);
        HIPCHK(hipMemcpyAsync(ph_h.data(), phase, 2 * P * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(tsp_h.data(), tsp, 2 * P * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(sig_h.data(), sig, sig_h.size() * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        // smallest singular value above the zero threshold per pair (both spins):
        // the co-densities carry 1/sigma, so FP32 rounding in the four-index
        // transforms is amplified roughly as 1/sigma^4 in the matrix elements
        std::vector<double> smin(P, 1.0);
        for (int sp = 0; sp < 2; ++sp)
            for (int p = 0; p < P; ++p)
                for (int k = 0; k < c->nocc[sp]; ++k) {
                    const double sv = std::fabs(sig_h[(size_t)sp * P * nomax + (size_t)p * c->nocc[sp] + k]);
                    if (sv > tol) smin[p] = std::min(smin[p], sv);
                }
        // ---- slab layout ------------------------------------------------------
        // A pair whose smallest nonzero singular value clears the floor runs its
        // four-index transforms in reduced precision and keeps its rank-4 blocks
        // as FP32 (half-length blocks).
        const bool mixed = c->precision != NW_PREC_FP64 && n % 2 == 0;
        long long total = 0;
        for (int p = 0; p < P; ++p) {
            PairInfo &pi = c->pinfo[(size_t)px[p] * nref + pw[p]];
            pi.smin = smin[p];
            pi.fp32 = mixed && smin[p] >= c->mixed_floor ? 1 : 0;
            long long sz = pair_offsets(pi.ma, pi.mb, nm, pi.off, pi.fp32 != 0);
            pi.base = total;
            pi.built = 1;
            pi.bidx = p;
            total += sz;
            if (sizes) sizes->push_back(sz);
        }
        if (sizes) return NW_OK;
        if ((size_t)total > c->slab_cap) {
            if (c->d_slab) HIPCHK(hipFree(c->d_slab));
            c->d_slab = nullptr;
            c->slab_cap = 0;
            c->d_slab = dalloc<double>((size_t)total);
            c->slab_cap = (size_t)total;
        }
        double *slab = c->d_slab;
        auto SL = [&](int p, long long rel) { return slab + c->pinfo[(size_t)px[p] * nref + pw[p]].base + rel; };
        auto OFF = [&](int p) { return c->pinfo[(size_t)px[p] * nref + pw[p]].off; };
        // ---- rotated occupied orbitals and co-densities ----------------------
        for (int sp = 0; sp < 2; ++sp) {
            const int no = c->nocc[sp];
            std::vector<GemmDesc> gl, gg;
            for (int p = 0; p < P; ++p) {
                size_t q = (size_t)sp * P + p;
                gl.push_back({Cref(sp, px[p]), U + (size_t)sp * P * nomax * nomax + (size_t)p * no * no, Lt + q * n * nomax});
                gg.push_back({Cref(sp, pw[p]), V + (size_t)sp * P * nomax * nomax + (size_t)p * no * no, Gt + q * n * nomax});
            }
            gemm(c, up, 0, 0, n, no, no, nm, no, no, 1.0, 0.0, gl);
            gemm(c, up, 0, 0, n, no, no, nm, no, no, 1.0, 0.0, gg);
            // codensity of every pair for this spin, into M[(p*2+sp)*2 + mi]
            launch_codensity(c->st, Lt + (size_t)sp * P * n * nomax, Gt + (size_t)sp * P * n * nomax,
                             sig + (size_t)sp * P * nomax, n, no, tol, M + (size_t)prob(0, sp) * 2 * nn, P,
                             (long long)n * nomax, no, (long long)(prob(1, sp) - prob(0, sp)) * 2 * nn);
            std::vector<const double *> cxp, cwp;
            for (int p = 0; p < P; ++p) { cxp.push_back(Cref(sp, px[p])); cwp.push_back(Cref(sp, pw[p])); }
            // contraction orbitals of every pair into rowc/colc at prob(p, sp)
            const double *const *dcx = up.put(cxp);
            const double *const *dcw = up.put(cwp);
            const long long pstride = (long long)(prob(1, sp) - prob(0, sp));
            launch_contraction_orbitals(c->st, dcx, dcw, n, nm, no, rowc + (size_t)prob(0, sp) * nnm,
                                        colc + (size_t)prob(0, sp) * nnm, P, pstride * (long long)nnm);
            launch_prefill_cxxc(c->st, rowc + (size_t)prob(0, sp) * nnm, colc + (size_t)prob(0, sp) * nnm, n, nm, no,
                                cx + (size_t)prob(0, sp) * 2 * nnm, xc + (size_t)prob(0, sp) * 2 * nnm, P,
                                pstride * (long long)nnm, pstride * 2 * (long long)nnm);
        }
        // ---- M S, M^T S, cx, xc (build_cx_xc, build.rs:944-991) ---------------
        {
            std::vector<GemmDesc> gms, gmts, gcx, gxc, gq, gr, gsc, gz;
            for (int k = 0; k < 4 * P; ++k) {
                gms.push_back({M + (size_t)k * nn, c->d_S, MS + (size_t)k * nn});
                gmts.push_back({M + (size_t)k * nn, c->d_S, MtS + (size_t)k * nn});
                const int pr = k / 2;
                gcx.push_back({MtS + (size_t)k * nn, rowc + (size_t)pr * nnm, cx + (size_t)k * nnm});
                gxc.push_back({MS + (size_t)k * nn, colc + (size_t)pr * nnm, xc + (size_t)k * nnm});
                gq.push_back({MS + (size_t)k * nn, colc + (size_t)pr * nnm, Q + (size_t)k * nnm});
                gr.push_back({c->d_S, Q + (size_t)k * nnm, R + (size_t)k * nnm});
            }
            gemm(c, up, 0, 0, n, n, n, n, n, n, 1.0, 0.0, gms);
            gemm(c, up, 1, 0, n, n, n, n, n, n, 1.0, 0.0, gmts);
            gemm(c, up, 0, 0, n, nm, n, n, nm, nm, 1.0, 1.0, gcx);
            gemm(c, up, 0, 0, n, nm, n, n, nm, nm, 1.0, 1.0, gxc);
            gemm(c, up, 0, 0, n, nm, n, n, nm, nm, 1.0, 0.0, gq);
            gemm(c, up, 0, 0, n, nm, n, n, nm, nm, 1.0, 0.0, gr);
            // X[mi] = rowc^T (S M S colc) into the slab; Z = rowc^T S colc
            std::vector<GemmDesc> gx;
            for (int p = 0; p < P; ++p)
                for (int sp = 0; sp < 2; ++sp) {
                    const long long *o = OFF(p) + (sp ? O_BB : O_AA);
                    const int pr = prob(p, sp);
                    for (int mi = 0; mi < 2; ++mi)
                        gx.push_back({rowc + (size_t)pr * nnm, R + (size_t)(pr * 2 + mi) * nnm, SL(p, o[O_X + mi])});
                    gsc.push_back({c->d_S, colc + (size_t)pr * nnm, Q + (size_t)pr * nnm});  // reuse Q rows as S colc
                }
            gemm(c, up, 1, 0, nm, nm, n, nm, nm, nm, 1.0, 0.0, gx);
            // Q is free now (its contents were consumed by R); compute S colc into Q[pr]
            gemm(c, up, 0, 0, n, nm, n, n, nm, nm, 1.0, 0.0, gsc);
            for (int p = 0; p < P; ++p)
                for (int sp = 0; sp < 2; ++sp) {
                    const int pr = prob(p, sp);
                    gz.push_back({rowc + (size_t)pr * nnm, Q + (size_t)pr * nnm, Z + (size_t)pr * nm2});
                }
            gemm(c, up, 1, 0, nm, nm, n, nm, nm, nm, 1.0, 0.0, gz);
            std::vector<LinDesc> ly;
            for (int p = 0; p < P; ++p)
                for (int sp = 0; sp < 2; ++sp) {
                    const long long *o = OFF(p) + (sp ? O_BB : O_AA);
                    const int pr = prob(p, sp);
                    ly.push_back({SL(p, o[O_X]), Z + (size_t)pr * nm2, nullptr, SL(p, o[O_Y]), 1.0, -1.0, 0.0});
                    ly.push_back({SL(p, o[O_X + 1]), nullptr, nullptr, SL(p, o[O_Y + 1]), 1.0, 0.0, 0.0});
                    for (int k = 0; k < 4; ++k)  // ff intermediates start at zero (build.rs:216-220)
                        ly.push_back({nullptr, nullptr, nullptr, SL(p, o[O_FF + k]), 0.0, 0.0, 0.0});
                }
            launch_lincomb(c->st, up.put(ly), (int)ly.size(), (long long)nm2);
        }
        // ---- J - K and J for every co-density (build.rs:179-191, 812-820) ----
        // one launch for both products, K split when they have too few
        // output tiles to fill the device
        {
            const std::vector<GemmDesc> g = {{M, c->d_easym, JK}, {M, c->d_eri, Jc}};
            const int rows = 4 * P, nni = (int)nn;
            int ncu = 0;
            HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
            const int ks = nni >= 256 ? gemm_nt_splitk_count(rows, nni, nni, 2, std::max(1, ncu)) : 1;
            if (ks > 1) {
                double *part = c->splitk.get<double>((size_t)2 * ks * rows * nn);
                launch_gemm_nt_splitk(c->st, rows, nni, nni, nni, nni, nni, up.put(g), 2, ks, part);
            } else {
                gemm(c, up, 0, 1, rows, nni, nni, nni, nni, nni, 1.0, 0.0, g);
            }
        }
        // ---- scalar intermediates --------------------------------------------
        {
            std::vector<TraceDesc> tr;
            for (int p = 0; p < P; ++p) {
                const PairInfo &pi = c->pinfo[(size_t)px[p] * nref + pw[p]];
                double *mt = dmeta + (size_t)p * M_LEN;
                for (int sp = 0; sp < 2; ++sp) {
                    double *ms = mt + (sp ? M_BB : M_AA);
                    const int m = sp ? pi.mb : pi.ma;
                    const int pr = prob(p, sp);
                    const double *M0 = M + (size_t)(pr * 2) * nn, *M1 = M + (size_t)(pr * 2 + 1) * nn;
                    const double *JK0 = JK + (size_t)(pr * 2) * nn, *JK1 = JK + (size_t)(pr * 2 + 1) * nn;
                    tr.push_back({M0, c->d_h, ms + M_F0H, 1.0});
                    tr.push_back({M1, c->d_h, ms + M_F0H + 1, 1.0});
                    tr.push_back({JK0, M0, ms + M_V0, 1.0});
                    if (m >= 1) tr.push_back({JK0, M1, ms + M_V0 + 1, 2.0});
                    if (m >= 2) tr.push_back({JK1, M1, ms + M_V0 + 2, 1.0});
                }
                const int pa = prob(p, 0), pb = prob(p, 1);
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j) {
                        tr.push_back({Jc + (size_t)(pa * 2 + i) * nn, M + (size_t)(pb * 2 + j) * nn, mt + M_VAB0 + i * 2 + j, 1.0});
                        tr.push_back({Jc + (size_t)(pb * 2 + j) * nn, M + (size_t)(pa * 2 + i) * nn, mt + M_VBA0 + j * 2 + i, 1.0});
                    }
            }
            launch_einsum(c->st, up.put(tr), (int)tr.size(), n);
        }
        // ---- one-column intermediates (transposed, write2t layout.rs:339) ----
        {
            std::vector<GemmDesc> gh, gfh, gjx, gv, gva, gvb;
            for (int k = 0; k < 4 * P; ++k) gh.push_back({c->d_h, xc + (size_t)k * nnm, hxc + (size_t)k * nnm});
            gemm(c, up, 0, 0, n, nm, n, n, nm, nm, 1.0, 0.0, gh);
            for (int p = 0; p < P; ++p)
                for (int sp = 0; sp < 2; ++sp) {
                    const long long *o = OFF(p) + (sp ? O_BB : O_AA);
                    const int pr = prob(p, sp);
                    for (int mi = 0; mi < 2; ++mi)
                        for (int mj = 0; mj < 2; ++mj)
                            gfh.push_back({hxc + (size_t)(pr * 2 + mj) * nnm, cx + (size_t)(pr * 2 + mi) * nnm,
                                           SL(p, o[O_FH + mi * 2 + mj])});
                    for (int mk = 0; mk < 2; ++mk)
                        for (int mj = 0; mj < 2; ++mj)
                            gjx.push_back({JK + (size_t)(pr * 2 + mk) * nn, xc + (size_t)(pr * 2 + mj) * nnm,
                                           JKxc + (size_t)((pr * 2 + mk) * 2 + mj) * nnm});
                    for (int mi = 0; mi < 2; ++mi)
                        for (int mj = 0; mj < 2; ++mj)
                            for (int mk = 0; mk < 2; ++mk)
                                gv.push_back({JKxc + (size_t)((pr * 2 + mk) * 2 + mj) * nnm, cx + (size_t)(pr * 2 + mi) * nnm,
                                              SL(p, o[O_V + (mi * 2 + mj) * 2 + mk])});
                }
            gemm(c, up, 1, 0, nm, nm, n, nm, nm, nm, 1.0, 0.0, gfh);
            gemm(c, up, 0, 0, n, nm, n, n, nm, nm, 1.0, 0.0, gjx);
            gemm(c, up, 1, 0, nm, nm, n, nm, nm, nm, 1.0, 0.0, gv);
            // different-spin: reuse JKxc as Jc xc staging
            std::vector<GemmDesc> gja;
            for (int p = 0; p < P; ++p) {
                const int pa = prob(p, 0), pb = prob(p, 1);
                for (int b0 = 0; b0 < 2; ++b0)
                    for (int k = 0; k < 2; ++k) {
                        // Jc_b[b0] xc_a[k]  -> slot ((pa*2+b0)*2+k)
                        gja.push_back({Jc + (size_t)(pb * 2 + b0) * nn, xc + (size_t)(pa * 2 + k) * nnm,
                                       JKxc + (size_t)((pa * 2 + b0) * 2 + k) * nnm});
                        // Jc_a[b0] xc_b[k]  -> slot ((pb*2+b0)*2+k)
                        gja.push_back({Jc + (size_t)(pa * 2 + b0) * nn, xc + (size_t)(pb * 2 + k) * nnm,
                                       JKxc + (size_t)((pb * 2 + b0) * 2 + k) * nnm});
                    }
            }
            gemm(c, up, 0, 0, n, nm, n, n, nm, nm, 1.0, 0.0, gja);
            for (int p = 0; p < P; ++p) {
                const int pa = prob(p, 0), pb = prob(p, 1);
                const long long *o = OFF(p);
                for (int a0 = 0; a0 < 2; ++a0)
                    for (int b0 = 0; b0 < 2; ++b0)
                        for (int k = 0; k < 2; ++k) {
                            // vab[a0][b0][k] = cx_a[a0]^T Jb[b0] xc_a[k]
                            gva.push_back({JKxc + (size_t)((pa * 2 + b0) * 2 + k) * nnm, cx + (size_t)(pa * 2 + a0) * nnm,
                                           SL(p, o[O_VAB + (a0 * 2 + b0) * 2 + k])});
                            // vba[b0][a0][k] = cx_b[b0]^T Ja[a0] xc_b[k]
                            gvb.push_back({JKxc + (size_t)((pb * 2 + a0) * 2 + k) * nnm, cx + (size_t)(pb * 2 + b0) * nnm,
                                           SL(p, o[O_VBA + (b0 * 2 + a0) * 2 + k])});
                        }
            }
            gemm(c, up, 1, 0, nm, nm, n, nm, nm, nm, 1.0, 0.0, gva);
            gemm(c, up, 1, 0, nm, nm, n, nm, nm, nm, 1.0, 0.0, gvb);
            std::vector<LinDesc> lh;
            for (int p = 0; p < P; ++p) {
                const long long *o = OFF(p);
                if (o[O_AA + O_HCOL0] < 0) continue;
                lh.push_back({SL(p, o[O_AA + O_FH]), SL(p, o[O_AA + O_V]), SL(p, o[O_VAB]), SL(p, o[O_AA + O_HCOL0]), 1, 1, 1});
                lh.push_back({SL(p, o[O_BB + O_FH]), SL(p, o[O_BB + O_V]), SL(p, o[O_VBA]), SL(p, o[O_BB + O_HCOL0]), 1, 1, 1});
            }
            launch_lincomb(c->st, up.put(lh), (int)lh.size(), (long long)nm2);
        }
        HIPCHK(hipGetLastError());
        // ---- four-index transforms (build.rs:261-313, 877-917) ----------------
        // With integral factors installed, every pair leaves its rank-4 blocks
        // to the evaluation calls (chol_ensure).
        HIPCHK(hipEventRecord(c->ev[1], c->st));
        std::vector<uint8_t> lazy(P, 0);
        std::vector<int> keys_all, keys_dense;
        {
            const bool route = chol_route(c);
            for (int p = 0; p < P; ++p) {
                const int key = px[p] * nref + pw[p];
                keys_all.push_back(key);
                if (route) lazy[p] = 1;
                else keys_dense.push_back(key);
            }
        }
        run_rank4_transforms(c, keys_dense);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[2], c->st));
        // ---- meta to host, pair table to device ------------------------------
        std::vector<double> mh((size_t)P * M_LEN);
        HIPCHK(hipMemcpyAsync(mh.data(), dmeta, mh.size() * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (int p = 0; p < P; ++p) {
            PairInfo &pi = c->pinfo[(size_t)px[p] * nref + pw[p]];
            for (int k = 0; k < M_LEN; ++k) pi.meta[k] = mh[(size_t)p * M_LEN + k];
            for (int sp = 0; sp < 2; ++sp) {
                double *ms = pi.meta + (sp ? M_BB : M_AA);
                ms[M_TSP] = tsp_h[(size_t)sp * P + p];
                ms[M_PH] = ph_h[(size_t)sp * P + p];
                ms[M_M] = sp ? pi.mb : pi.ma;
                ms[M_NOCC] = c->nocc[sp];
            }
            if (meta_out) {
                nw_pair_meta &o = meta_out[(size_t)px[p] * nref + pw[p]];
                o.ma = pi.ma; o.mb = pi.mb;
                o.tilde_s_a = pi.meta[M_AA + M_TSP]; o.phase_a = pi.meta[M_AA + M_PH];
                o.tilde_s_b = pi.meta[M_BB + M_TSP]; o.phase_b = pi.meta[M_BB + M_PH];
                o.sigma_min = pi.smin; o.fp32 = pi.fp32; o.pad = 0;
            }
        }
        if (c->pinfo.size() > c->pinfo_cap) {
            if (c->d_pinfo) HIPCHK(hipFree(c->d_pinfo));
            c->d_pinfo = nullptr;
            c->pinfo_cap = 0;
            c->d_pinfo = (PairInfo *)dalloc<char>(sizeof(PairInfo) * c->pinfo.size());
            c->pinfo_cap = c->pinfo.size();
        }
        HIPCHK(hipMemcpyAsync(c->d_pinfo, c->pinfo.data(), sizeof(PairInfo) * c->pinfo.size(), hipMemcpyHostToDevice, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->t_ms[0] = elapsed(c->ev[0], c->ev[2]);
        c->t_ms[1] = elapsed(c->ev[1], c->ev[2]);
        c->t_ms[2] = 0.0;
        c->built = true;
        c->have_codensity = true;
        chol_on_build(c, keys_all, lazy);
        return NW_OK;
    }
}

// The four-index transforms of every rank-4 tensor of the given built pairs
// (build.rs:261-313, 877-917), from the cx/xc contractions the last build
// kept: one job per tensor, the J slots per spin and the II blocks.
void run_rank4_transforms(nw_ctx *c, const std::vector<int> &keys)
{
    if (keys.empty()) return;
    const int n = c->n, nm = c->nmo;
    const size_t nnm = (size_t)n * nm;
    int P = 0;
    for (const PairInfo &q : c->pinfo)
        if (q.built) P = std::max(P, q.bidx + 1);
    double *slab = c->d_slab;
    double *cx = (double *)c->keepCX.p, *xc = (double *)c->keepXC.p;
    auto prob = [&](int p, int sp) { return p * 2 + sp; };
    const bool bf3 = c->precision == NW_PREC_BF16X3;
    // One job per rank-4 tensor: the J slots per spin and the II blocks.
    // J slab block [a,b,c,d] = blk[c,d,a,b] with blk = (cx_i xc_j || cx_k xc_l);
    // the pair symmetry of (ab||cd) lets one transform produce it directly.
    struct TJob { bool f32; bool asym; int c1, c2, c3, c4; double *out; };  // cN: matrix index into cx/xc
    std::vector<TJob> jobs;
    auto CX = [&](int pr, int mi) { return 2 * (pr * 2 + mi); };       // even: cx, odd: xc
    auto XC = [&](int pr, int mi) { return 2 * (pr * 2 + mi) + 1; };
    for (int key : keys) {
        PairInfo &pi = c->pinfo[key];
        const int p = pi.bidx;
        const bool f32 = pi.fp32 != 0;
        for (int sp = 0; sp < 2; ++sp) {
            const long long *o = pi.off + (sp ? O_BB : O_AA);
            const int pr = prob(p, sp);
            for (int s = 0; s < 10; ++s) {
                if (o[O_J + s] < 0) continue;
                const int i = jbranch(s, 0), j = jbranch(s, 1), k = jbranch(s, 2), l = jbranch(s, 3);
                jobs.push_back({f32, true, CX(pr, k), XC(pr, l), CX(pr, i), XC(pr, j), slab + pi.base + o[O_J + s]});
            }
        }
        const int pa = prob(p, 0), pb = prob(p, 1);
        for (int a0 = 0; a0 < 2; ++a0)
            for (int aj = 0; aj < 2; ++aj)
                for (int b0 = 0; b0 < 2; ++b0)
                    for (int bj = 0; bj < 2; ++bj) {
                        long long oo = pi.off[O_II + ((a0 * 2 + aj) * 2 + b0) * 2 + bj];
                        if (oo < 0) continue;
                        jobs.push_back({f32, false, CX(pa, a0), XC(pa, aj), CX(pb, b0), XC(pb, bj), slab + pi.base + oo});
                    }
    }
    float *cx32 = nullptr, *xc32 = nullptr;
    bool any32 = false;
    for (auto &jb : jobs) any32 |= jb.f32;
    if (any32) {
        // FP32 operands, intermediates and MFMA accumulation; the rank-4 slab
        // blocks they produce hold FP32 values (PairInfo::fp32).
        const size_t n4 = (size_t)n * n * n * n;
        if (!c->d_eri32) {
            c->d_eri32 = (float *)dalloc<char>(n4 * sizeof(float));
            c->d_easym32 = (float *)dalloc<char>(n4 * sizeof(float));
            launch_to_f32(c->st, c->d_eri, c->d_eri32, (long long)n4);
            launch_to_f32(c->st, c->d_easym, c->d_easym32, (long long)n4);
        }
        cx32 = c->cx32.get<float>((size_t)4 * P * nnm);
        xc32 = c->xc32.get<float>((size_t)4 * P * nnm);
        launch_to_f32(c->st, cx, cx32, (long long)4 * P * nnm);
        launch_to_f32(c->st, xc, xc32, (long long)4 * P * nnm);
    }
    // Tensors in flight per launch group: each needs two scratch slots of
    // max(n^4, nmo n^3) elements; the group is sized to 1 GiB per buffer.
    const size_t big = std::max((size_t)n * n * n * n, (size_t)nm * n * n * n);
    for (int f32 = 0; f32 < 2; ++f32) {
        std::vector<const TJob *> todo;
        for (auto &jb : jobs)
            if (jb.f32 == (f32 == 1)) todo.push_back(&jb);
        if (todo.empty()) continue;
        const size_t slot = big * (f32 ? sizeof(float) : sizeof(double));
        const int g = (int)std::max<size_t>(1, std::min<size_t>(QMAX, (size_t(1) << 30) / slot));
        char *A = c->scratchA.get<char>(slot * std::min<size_t>(g, todo.size()));
        char *B = c->scratchB.get<char>(slot * std::min<size_t>(g, todo.size()));
        for (size_t j0 = 0; j0 < todo.size(); j0 += g) {
            const int nb = (int)std::min<size_t>(g, todo.size() - j0);
            if (f32) {
                std::vector<const float *> X(nb), C1(nb), C2(nb), C3(nb), C4(nb);
                std::vector<float *> TA(nb), TB(nb), OUT(nb);
                auto cm = [&](int idx) { return (idx & 1 ? xc32 : cx32) + (size_t)(idx >> 1) * nnm; };
                for (int k = 0; k < nb; ++k) {
                    const TJob &jb = *todo[j0 + k];
                    X[k] = jb.asym ? c->d_easym32 : c->d_eri32;
                    C1[k] = cm(jb.c1); C2[k] = cm(jb.c2); C3[k] = cm(jb.c3); C4[k] = cm(jb.c4);
                    TA[k] = (float *)(A + k * slot);
                    TB[k] = (float *)(B + k * slot);
                    OUT[k] = (float *)jb.out;  // the pair's half-length FP32 rank-4 block
                }
                std::vector<const float *> TAc(TA.begin(), TA.end()), TBc(TB.begin(), TB.end());
                auto pass = bf3 ? launch_qtransform_bf3_many : launch_qtransform_f32_many;
                bool ok;
                if (bf3 && three_launch_transform(n, nm)) {
                    // half transform of the (p,q) slices, then p and q as leading indices
                    const long long nm2 = (long long)nm * nm;
                    ok = launch_qhalf_bf3_many(c->st, X.data(), C3.data(), C4.data(), TA.data(), nb, n, nm) &&
                         launch_qlead_bf3_many(c->st, TAc.data(), C1.data(), TB.data(), nb, n, nm, n * nm2, 1, 0, 0) &&
                         launch_qlead_bf3_many(c->st, TBc.data(), C2.data(), OUT.data(), nb, n, nm, nm2, nm,
                                               n * nm2, nm * nm2);
                } else {
                    ok = pass(c->st, X.data(), C4.data(), TA.data(), nb, (long long)n * n * n, n, nm) &&
                         pass(c->st, TAc.data(), C3.data(), TB.data(), nb, (long long)nm * n * n, n, nm) &&
                         pass(c->st, TBc.data(), C2.data(), TA.data(), nb, (long long)nm * nm * n, n, nm) &&
                         pass(c->st, TAc.data(), C1.data(), OUT.data(), nb, (long long)nm * nm * nm, n, nm);
                }
                if (!ok) throw NwError(NW_E_SHAPE, "mixed-precision transform needs an even n_bas and nmo <= 112");
            } else {
                std::vector<const double *> X(nb), C1(nb), C2(nb), C3(nb), C4(nb);
                std::vector<double *> TA(nb), TB(nb), OUT(nb);
                auto cm = [&](int idx) { return (idx & 1 ? xc : cx) + (size_t)(idx >> 1) * nnm; };
                for (int k = 0; k < nb; ++k) {
                    const TJob &jb = *todo[j0 + k];
                    X[k] = jb.asym ? c->d_easym : c->d_eri;
                    C1[k] = cm(jb.c1); C2[k] = cm(jb.c2); C3[k] = cm(jb.c3); C4[k] = cm(jb.c4);
                    TA[k] = (double *)(A + k * slot);
                    TB[k] = (double *)(B + k * slot);
                    OUT[k] = jb.out;
                }
                std::vector<const double *> TAc(TA.begin(), TA.end()), TBc(TB.begin(), TB.end());
                launch_qtransform_many(c->st, X.data(), C4.data(), TA.data(), nb, (long long)n * n * n, n, nm);
                launch_qtransform_many(c->st, TAc.data(), C3.data(), TB.data(), nb, (long long)nm * n * n, n, nm);
                launch_qtransform_many(c->st, TBc.data(), C2.data(), TA.data(), nb, (long long)nm * nm * n, n, nm);
                launch_qtransform_many(c->st, TAc.data(), C1.data(), OUT.data(), nb, (long long)nm * nm * nm, n, nm);
            }
        }
    }
}

// assign_offsets (src/nonorthogonalwicks/layout.rs:21-189) for one pair.
// With r4_f32 the rank-4 blocks (J slots, II) hold nmo^4 FP32 values in
// half-length blocks; every other block and the block order are the
// reference's.
long long pair_offsets(int ma, int mb, int nmo, long long *off, bool r4_f32)
{
    const long long nn2 = (long long)nmo * nmo, nn4 = r4_f32 ? (nn2 * nn2 + 1) / 2 : nn2 * nn2;
    long long i = 0;
    for (int k = 0; k < O_TOTAL; ++k) off[k] = -1;
    for (int sp = 0; sp < 2; ++sp) {
        long long *o = off + (sp ? O_BB : O_AA);
        const int msp = sp ? mb : ma;
        for (int mi = 0; mi < 2; ++mi) { o[O_X + mi] = i; i += nn2; }
        for (int mi = 0; mi < 2; ++mi) { o[O_Y + mi] = i; i += nn2; }
        for (int k = 0; k < 4; ++k) { o[O_FH + k] = i; i += nn2; }
        for (int k = 0; k < 4; ++k) { o[O_FF + k] = i; i += nn2; }
        for (int k = 0; k < 8; ++k) { o[O_V + k] = i; i += nn2; }
        for (int s = 0; s < 10; ++s)
            if (jbranch(s, 0) + jbranch(s, 1) + jbranch(s, 2) + jbranch(s, 3) <= msp) { o[O_J + s] = i; i += nn4; }
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
                        off[O_II + ((a0 * 2 + aj) * 2 + b0) * 2 + bj] = i;
                        i += nn4;
                    }
    off[O_TOTAL] = i;
    return i;
}

// One pair's block in the reference layout (FP64 throughout, the offsets of
// pair_offsets(ma, mb, nmo)) from a host copy `blk` of its device block, whose
// rank-4 blocks are half-length FP32 blocks when the pair was built in
// reduced precision.
void export_pair(const PairInfo &pi, int nmo, const double *blk, double *out)
{
    long long ref[O_LEN];
    pair_offsets(pi.ma, pi.mb, nmo, ref);
    const size_t nn2 = (size_t)nmo * nmo, nn4 = nn2 * nn2;
    auto rank4 = [](int k) {
        return (k >= O_AA + O_J && k < O_AA + O_J + 10) || (k >= O_BB + O_J && k < O_BB + O_J + 10) ||
               (k >= O_II && k < O_II + 16);
    };
    for (int k = 0; k < O_TOTAL; ++k) {
        if (pi.off[k] < 0) continue;
        if (rank4(k) && pi.fp32) {
            float f;
            const char *src = (const char *)(blk + pi.off[k]);
            for (size_t i = 0; i < nn4; ++i) {
                std::memcpy(&f, src + i * sizeof(float), sizeof(float));
                out[ref[k] + i] = (double)f;
            }
        } else {
            std::memcpy(out + ref[k], blk + pi.off[k], (rank4(k) ? nn4 : nn2) * sizeof(double));
        }
    }
}


// Pack determinants on the device and return the host copy of the packed records.
// The evaluation kernels cover contraction ranks (bra + ket excitations per
// spin) up to four, the fixed-rank kernels of the reference.
const char *const RANK_MSG = "contraction rank per spin (bra plus ket excitations) must be <= 8";

std::vector<PackedDet> pack(nw_ctx *c, const nw_det *dets, uint64_t ndet, PackedDet *&dpacked, const void **draw_out)
{
    const void *draw = to_device(c, c->dets_raw, dets, ndet * sizeof(nw_det));
    if (draw_out) *draw_out = draw;
    dpacked = c->dets_packed.get<PackedDet>(ndet);
    int *dflag = c->flag.get<int>(1);
    HIPCHK(hipMemsetAsync(dflag, 0, sizeof(int), c->st));
    launch_pack_dets(c->st, draw, (long long)ndet, c->nref, dpacked, dflag);
    std::vector<PackedDet> h(ndet);
    int flag = 0;
    HIPCHK(hipMemcpyAsync(h.data(), dpacked, ndet * sizeof(PackedDet), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(&flag, dflag, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (flag & 2) throw NwError(NW_E_ARG, "determinant parent out of range");
    if (flag & 1) throw NwError(NW_E_RANK, "excitation rank per spin must be <= 4 with equal holes and particles");
    return h;
}

// Bin key of a block: ((L_a * 5 + L_b) << 1) | m0 for contraction ranks up
// to four per spin (the fixed-rank kernels), REQ_KEY_GEN + m0 above (the
// run-time-rank kernels).
int bin_key(int la, int lb, bool m0)
{
    return (la > 4 || lb > 4) ? REQ_KEY_GEN + (m0 ? 1 : 0) : ((la * 5 + lb) << 1) | (m0 ? 1 : 0);
}

// Upper-triangle pairs (i <= j) of a determinant list grouped into blocks of
// equal (parent, rank_a, rank_b) runs: calculate_matrix_elements' pair list
// (src/noci/matrix.rs:55-59), partitioned by contraction-rank bin.
std::map<int, std::vector<EvalBlock>> triangle_blocks(nw_ctx *c, const std::vector<PackedDet> &h, bool need_hcol,
                                                      bool built_only)
{
    std::vector<Run> runs;
    for (long long i = 0; i < (long long)h.size(); ++i) {
        const PackedDet &d = h[i];
        if (!runs.empty() && runs.back().parent == d.parent && runs.back().ra == d.ra && runs.back().rb == d.rb)
            runs.back().len++;
        else
            runs.push_back({i, 1, d.parent, d.ra, d.rb});
    }
    std::map<int, std::vector<EvalBlock>> bins;  // key = ((la*5+lb) << 1) | m0
    for (size_t r1 = 0; r1 < runs.size(); ++r1)
        for (size_t r2 = r1; r2 < runs.size(); ++r2) {
            const Run &A = runs[r1], &B = runs[r2];
            const int pair = A.parent * c->nref + B.parent;
            const PairInfo &pi = c->pinfo[pair];
            const int la = A.ra + B.ra, lb = A.rb + B.rb;
            if (la > 8 || lb > 8) throw NwError(NW_E_RANK, RANK_MSG);
            if (!pi.built) {
                if (built_only) continue;
                throw NwError(NW_E_STATE, "reference pair (" + std::to_string(A.parent) + "," + std::to_string(B.parent) +
                                              ") was not built");
            }
            const bool m0 = need_hcol && pi.ma == 0 && pi.mb == 0;
            auto &v = bins[bin_key(la, lb, m0)];
            v.push_back({A.start, B.start, (int)A.len, (int)B.len, pair, r1 == r2 ? 1 : 0, 0});
        }
    for (auto &kv : bins) {
        long long t = 0;
        for (auto &b : kv.second) {
            b.tstart = t;
            t += (long long)((b.ni + 15) / 16) * ((b.nj + 15) / 16);
        }
    }
    return bins;
}

// Rows [nrow0, n) against columns [0, ncol) of one determinant list: every
// pair (row, col) once, grouped by (parent, rank_a, rank_b) runs as in
// triangle_blocks (no triangle restriction).
std::map<int, std::vector<EvalBlock>> rect_blocks(nw_ctx *c, const std::vector<PackedDet> &h, long long nrow0,
                                                  long long ncol, bool need_hcol)
{
    auto runs_of = [&](long long a, long long b) {
        std::vector<Run> runs;
        for (long long i = a; i < b; ++i) {
            const PackedDet &d = h[i];
            if (!runs.empty() && runs.back().parent == d.parent && runs.back().ra == d.ra && runs.back().rb == d.rb)
                runs.back().len++;
            else
                runs.push_back({i, 1, d.parent, d.ra, d.rb});
        }
        return runs;
    };
    const std::vector<Run> rows = runs_of(nrow0, (long long)h.size()), cols = runs_of(0, ncol);
    std::map<int, std::vector<EvalBlock>> bins;
    for (const Run &A : rows)
        for (const Run &B : cols) {
            const int pair = A.parent * c->nref + B.parent;
            const PairInfo &pi = c->pinfo[pair];
            if (!pi.built)
                throw NwError(NW_E_STATE, "reference pair (" + std::to_string(A.parent) + "," + std::to_string(B.parent) +
                                              ") was not built");
            const int la = A.ra + B.ra, lb = A.rb + B.rb;
            if (la > 8 || lb > 8) throw NwError(NW_E_RANK, RANK_MSG);
            const bool m0 = need_hcol && pi.ma == 0 && pi.mb == 0;
            bins[bin_key(la, lb, m0)].push_back({A.start, B.start, (int)A.len, (int)B.len, pair, 0, 0});
        }
    for (auto &kv : bins) {
        long long t = 0;
        for (auto &b : kv.second) {
            b.tstart = t;
            t += (long long)((b.ni + 15) / 16) * ((b.nj + 15) / 16);
        }
    }
    return bins;
}

long long tiles_of(const std::vector<EvalBlock> &v)
{
    const EvalBlock &b = v.back();
    return b.tstart + (long long)((b.ni + 15) / 16) * ((b.nj + 15) / 16);
}

EvalArgs base_args(nw_ctx *c, const PackedDet *dp, int mode)
{
    EvalArgs a{};
    a.dets = dp;
    a.pairs = c->d_pinfo;
    a.slab = c->d_slab;
    a.nref = c->nref;
    a.nmo = c->nmo;
    a.mode = mode;
    a.enuc = c->enuc;
    a.tol = c->tol;
    return a;
}

// Shared body of nw_hs_matrix / nw_fs_matrix.
int eval_matrix(nw_ctx *c, int kind, const nw_det *dets, uint64_t ndet, double *A, double *B, int mode)
{
    if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
    if (!dets || !A || !B || ndet == 0) return fail(c, NW_E_ARG, "null arguments");
    if (ndet > (1ull << 31)) return fail(c, NW_E_SHAPE, "too many determinants");
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    PackedDet *dp = nullptr;
    std::vector<PackedDet> h = pack(c, dets, ndet, dp);
    if (kind == EVAL_HS) {
        const std::vector<uint8_t> reads = triangle_pairs(c, h);
        chol_ensure(c, h, false, &reads);
    }
    auto bins = triangle_blocks(c, h, kind == EVAL_HS);
    const bool adev = is_device_ptr(A), bdev = is_device_ptr(B);
    double *dA = adev ? A : c->outH.get<double>(ndet * ndet);
    double *dB = bdev ? B : c->outS.get<double>(ndet * ndet);
    EvalArgs args = base_args(c, dp, mode);
    args.A = dA;
    args.B = dB;
    args.ldo = (long long)ndet;
    Uploader up(c, 64u << 20);
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    for (auto &kv : bins) {
        const int key = kv.first;
        int la, lb;
        key_ranks(key, la, lb);
        EvalBlock *db = up.put(kv.second);
        launch_eval_blocks(c->st, kind, la, lb, key & 1, db, (int)kv.second.size(), tiles_of(kv.second), args);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[2], c->st));
    if (!adev) HIPCHK(hipMemcpyAsync(A, dA, ndet * ndet * 8, hipMemcpyDeviceToHost, c->st));
    if (!bdev) HIPCHK(hipMemcpyAsync(B, dB, ndet * ndet * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->t_ms[0] = elapsed(c->ev[0], c->ev[2]);
    c->t_ms[1] = kind == EVAL_HS ? chol_last_ms(c) : 0.0;
    c->t_ms[2] = elapsed(c->ev[1], c->ev[2]);
    return NW_OK;
}

// Shared body of nw_hs_batched / nw_fs_batched.
int eval_batched(nw_ctx *c, int kind, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq, double *A,
                 double *B, int mode)
{
    if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
    if (!dets || !req || !A || !B) return fail(c, NW_E_ARG, "null arguments");
    if (nreq == 0) return NW_OK;
    if (nreq > (1ull << 31)) return fail(c, NW_E_SHAPE, "too many requests");
    HIPCHK(hipEventRecord(c->ev[0], c->st));
    PackedDet *dp = nullptr;
    const void *draw = nullptr;
    {
        std::vector<PackedDet> h = pack(c, dets, ndet, dp, &draw);
        if (kind == EVAL_HS) chol_ensure(c, h);
    }
    const uint64_t *dreq = (const uint64_t *)to_device(c, c->req_in, req, nreq * 16);
    const bool adev = is_device_ptr(A), bdev = is_device_ptr(B);
    double *dA = adev ? A : c->outH.get<double>(nreq);
    double *dB = bdev ? B : c->outS.get<double>(nreq);
    // calculate_hs_pairs_wicks_batched (src/noci/hs.rs:86-109): exact zeros and
    // contraction-rank bins, on the GPU (k_req_key / k_req_scatter)
    unsigned *bins = c->req_bins.get<unsigned>(2 * REQ_NKEY + 1);
    ReqBin rb{};
    rb.req = dreq;
    rb.nreq = (long long)nreq;
    rb.ndet = (long long)ndet;
    rb.raw = draw;
    rb.dp = dp;
    rb.pairs = c->d_pinfo;
    rb.nref = c->nref;
    rb.hs = kind == EVAL_HS ? 1 : 0;
    rb.key = c->req_key.get<int>(nreq);
    rb.hist = bins;
    rb.cursor = bins + REQ_NKEY;
    rb.err = (int *)(bins + 2 * REQ_NKEY);
    rb.A = dA;
    rb.B = dB;
    rb.out = c->req_sorted.get<EvalReq>(nreq);
    HIPCHK(hipMemsetAsync(bins, 0, (2 * REQ_NKEY + 1) * sizeof(unsigned), c->st));
    launch_req_bin(c->st, rb, false);
    HIPCHK(hipGetLastError());
    unsigned hb[REQ_NKEY + 1];
    HIPCHK(hipMemcpyAsync(hb, bins, REQ_NKEY * sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(hb + REQ_NKEY, rb.err, sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (hb[REQ_NKEY] == REQ_ERR_INDEX) throw NwError(NW_E_ARG, "request index out of range");
    if (hb[REQ_NKEY] == REQ_ERR_UNBUILT) throw NwError(NW_E_STATE, "reference pair not built");
    if (hb[REQ_NKEY] == REQ_ERR_RANK) throw NwError(NW_E_RANK, RANK_MSG);
    std::vector<unsigned> start(REQ_NKEY);
    unsigned tot = 0;
    for (int k = 0; k < REQ_NKEY; ++k) {
        start[k] = tot;
        tot += hb[k];
    }
    Uploader up(c, 4096);
    rb.start = up.put(start);
    launch_req_bin(c->st, rb, true);
    EvalArgs args = base_args(c, dp, mode);
    args.A = dA;
    args.B = dB;
    HIPCHK(hipEventRecord(c->ev[1], c->st));
    for (int k = 0; k < REQ_NKEY; ++k) {
        if (!hb[k]) continue;
        int la, lb;
        key_ranks(k, la, lb);
        launch_eval_reqs(c->st, kind, la, lb, k & 1, rb.out + start[k], (int)hb[k], args);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[2], c->st));
    if (!adev) HIPCHK(hipMemcpyAsync(A, dA, nreq * 8, hipMemcpyDeviceToHost, c->st));
    if (!bdev) HIPCHK(hipMemcpyAsync(B, dB, nreq * 8, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    c->t_ms[0] = elapsed(c->ev[0], c->ev[2]);
    c->t_ms[1] = kind == EVAL_HS ? chol_last_ms(c) : 0.0;
    c->t_ms[2] = elapsed(c->ev[1], c->ev[2]);
    return NW_OK;
}

}  // namespace

extern "C" {

int nw_hs_matrix(nw_ctx *c, const nw_det *dets, uint64_t ndet, double *H, double *S, int mode)
{
    return guarded(c, [&]() -> int { return eval_matrix(c, EVAL_HS, dets, ndet, H, S, mode); });
}

int nw_hs_matrix_streamed(nw_ctx *c, double tol, const nw_det *dets, uint64_t ndet, double *H, double *S, int mode,
                          uint64_t slab_bytes, nw_pair_meta *meta_out, int *ngroups)
{
    return guarded(c, [&]() -> int {
        if (c->nref == 0) return fail(c, NW_E_STATE, "nw_set_refs first");
        if (!dets || !H || !S || ndet == 0) return fail(c, NW_E_ARG, "null arguments");
        if (ndet > (1ull << 31)) return fail(c, NW_E_SHAPE, "too many determinants");
        const int nref = c->nref;
        const auto t0 = std::chrono::steady_clock::now();
        PackedDet *dp = nullptr;
        std::vector<PackedDet> h = pack(c, dets, ndet, dp);
        // the ordered reference pairs of the upper triangle's run pairs (triangle_blocks)
        std::vector<int> runp;
        for (size_t i = 0; i < h.size(); ++i)
            if (i == 0 || h[i].parent != h[i - 1].parent || h[i].ra != h[i - 1].ra || h[i].rb != h[i - 1].rb)
                runp.push_back(h[i].parent);
        std::vector<uint8_t> need((size_t)nref * nref, 0);
        for (size_t r1 = 0; r1 < runp.size(); ++r1)
            for (size_t r2 = r1; r2 < runp.size(); ++r2) need[(size_t)runp[r1] * nref + runp[r2]] = 1;
        // the slab length of every needed pair, from its zero-overlap counts
        std::vector<long long> sizes;
        int rc = build_impl(c, tol, need.data(), nullptr, &sizes);
        if (rc != NW_OK) return rc;
        std::vector<int> pid;
        for (int p = 0; p < nref * nref; ++p)
            if (need[p]) pid.push_back(p);
        // outputs first, then the budget.  The slab allocation of an earlier build
        // is kept when it fits the budget: the groups reuse it, and so do later
        // calls (freeing and re-allocating hundreds of GB costs seconds, while the
        // driver clears the freed memory).
        const bool adev = is_device_ptr(H), bdev = is_device_ptr(S);
        double *dA = adev ? H : c->outH.get<double>(ndet * ndet);
        double *dB = bdev ? S : c->outS.get<double>(ndet * ndet);
        uint64_t budget = slab_bytes;
        if (budget == 0) {
            size_t fr = 0, tot = 0;
            HIPCHK(hipMemGetInfo(&fr, &tot));
            // room for the build's temporaries: two transform scratch buffers of up to
            // 1 GiB each, the per-pair arenas and the FP32 integral copies
            const uint64_t n4 = (uint64_t)c->n * c->n * c->n * c->n;
            const uint64_t reserve = (6ull << 30) + n4 * 8, cap = (uint64_t)c->slab_cap * 8;
            const double avail = (double)cap + (double)fr - (double)reserve;
            // keep a slab that already holds most of what is available: growing it
            // by a few percent would cost a free and a re-allocation
            budget = (double)cap >= 0.85 * avail ? cap : (uint64_t)std::max(0.0, 0.9 * avail);
        }
        if (c->d_slab && (uint64_t)c->slab_cap * 8 > budget) {
            HIPCHK(hipFree(c->d_slab));
            c->d_slab = nullptr;
            c->slab_cap = 0;
        }
        // pairs in (x, w) order, greedily into groups whose blocks fit the budget
        std::vector<std::vector<int>> groups;
        uint64_t cur = 0;
        for (size_t k = 0; k < pid.size(); ++k) {
            const uint64_t b = (uint64_t)sizes[k] * 8;
            if (b > budget)
                return fail(c, NW_E_MEM, "the slab block of reference pair (" + std::to_string(pid[k] / nref) + "," +
                                             std::to_string(pid[k] % nref) + ") takes " + std::to_string(b) +
                                             " bytes, more than the slab budget of " + std::to_string(budget));
            if (groups.empty() || cur + b > budget) {
                groups.emplace_back();
                cur = 0;
            }
            groups.back().push_back(pid[k]);
            cur += b;
        }
        double build_ms = 0.0, eval_ms = 0.0;
        std::vector<uint8_t> mask((size_t)nref * nref);
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            const auto &g = groups[gi];
            std::fill(mask.begin(), mask.end(), 0);
            for (int p : g) mask[p] = 1;
            rc = build_impl(c, tol, mask.data(), meta_out, nullptr);
            if (rc != NW_OK) return rc;
            build_ms += c->t_ms[0];
            chol_ensure(c, h, gi > 0);  // on-demand rank-4 blocks of the group (integral factors)
            build_ms += chol_last_ms(c);
            // every element of the group's pairs; each pair is in exactly one group
            auto bins = triangle_blocks(c, h, true, true);
            EvalArgs args = base_args(c, dp, mode);
            args.A = dA;
            args.B = dB;
            args.ldo = (long long)ndet;
            Uploader up(c, 64u << 20);
            HIPCHK(hipEventRecord(c->ev[1], c->st));
            for (auto &kv : bins) {
                int la, lb;
                key_ranks(kv.first, la, lb);
                EvalBlock *db = up.put(kv.second);
                launch_eval_blocks(c->st, EVAL_HS, la, lb, kv.first & 1, db, (int)kv.second.size(), tiles_of(kv.second),
                                   args);
            }
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(c->ev[2], c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            eval_ms += elapsed(c->ev[1], c->ev[2]);
        }
        if (!adev) HIPCHK(hipMemcpyAsync(H, dA, ndet * ndet * 8, hipMemcpyDeviceToHost, c->st));
        if (!bdev) HIPCHK(hipMemcpyAsync(S, dB, ndet * ndet * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (ngroups) *ngroups = (int)groups.size();
        c->t_ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c->t_ms[1] = build_ms;
        c->t_ms[2] = eval_ms;
        return NW_OK;
    });
}

int nw_fs_matrix(nw_ctx *c, const nw_det *dets, uint64_t ndet, double *F, double *S, int mode)
{
    return guarded(c, [&]() -> int { return eval_matrix(c, EVAL_FS, dets, ndet, F, S, mode); });
}

int nw_hs_batched(nw_ctx *c, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq, double *h_out,
                  double *s_out, int mode)
{
    return guarded(c, [&]() -> int { return eval_batched(c, EVAL_HS, dets, ndet, req, nreq, h_out, s_out, mode); });
}

int nw_fs_batched(nw_ctx *c, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq, double *f_out,
                  double *s_out, int mode)
{
    return guarded(c, [&]() -> int { return eval_batched(c, EVAL_FS, dets, ndet, req, nreq, f_out, s_out, mode); });
}

int nw_s_same_batched(nw_ctx *c, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq, int spin,
                      double *s_out)
{
    return guarded(c, [&]() -> int {
        if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
        if (!dets || !req || !s_out || (spin != 0 && spin != 1)) return fail(c, NW_E_ARG, "bad arguments");
        if (nreq == 0) return NW_OK;
        if (nreq > (1ull << 31)) return fail(c, NW_E_SHAPE, "too many requests");
        HIPCHK(hipEventRecord(c->ev[0], c->st));
        PackedDet *dp = nullptr;
        pack(c, dets, ndet, dp);
        const uint64_t *dreq = (const uint64_t *)to_device(c, c->req_in, req, nreq * 16);
        const bool odev = is_device_ptr(s_out);
        double *dout = odev ? s_out : c->outS.get<double>(nreq);
        int *derr = c->flag.get<int>(1);
        HIPCHK(hipMemsetAsync(derr, 0, sizeof(int), c->st));
        EvalArgs args = base_args(c, dp, 0);
        HIPCHK(hipEventRecord(c->ev[1], c->st));
        launch_s_same(c->st, dreq, (long long)nreq, (long long)ndet, spin, args, dout, derr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[2], c->st));
        int err = 0;
        HIPCHK(hipMemcpyAsync(&err, derr, sizeof(int), hipMemcpyDeviceToHost, c->st));
        if (!odev) HIPCHK(hipMemcpyAsync(s_out, dout, nreq * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (err == REQ_ERR_INDEX) return fail(c, NW_E_ARG, "request index out of range");
        if (err == REQ_ERR_UNBUILT) return fail(c, NW_E_STATE, "reference pair not built");
        if (err == REQ_ERR_RANK) return fail(c, NW_E_RANK, RANK_MSG);
        c->t_ms[0] = elapsed(c->ev[0], c->ev[2]);
        c->t_ms[1] = 0.0;
        c->t_ms[2] = elapsed(c->ev[1], c->ev[2]);
        return NW_OK;
    });
}

int nw_m_batched(nw_ctx *c, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq, double e0,
                 double *m_out, int mode)
{
    return guarded(c, [&]() -> int {
        if (!m_out) return fail(c, NW_E_ARG, "null output");
        if (nreq == 0) return NW_OK;
        // F and S of every request into device temporaries, then M = F - e0 S
        double *f = c->scratchA.get<double>(2 * nreq), *s = f + nreq;
        int rc = eval_batched(c, EVAL_FS, dets, ndet, req, nreq, f, s, mode);
        if (rc != NW_OK) return rc;
        const bool mdev = is_device_ptr(m_out);
        double *dm = mdev ? m_out : c->outH.get<double>(nreq);
        std::vector<LinDesc> lc = {{f, s, nullptr, dm, 1.0, -e0, 0.0}};
        Uploader up(c, 4096);
        launch_lincomb(c->st, up.put(lc), 1, (long long)nreq);
        HIPCHK(hipGetLastError());
        if (!mdev) HIPCHK(hipMemcpyAsync(m_out, dm, nreq * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        return NW_OK;
    });
}

int nw_update_fock(nw_ctx *c, const double *fa, const double *fb)
{
    return guarded(c, [&]() -> int {
        if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
        if (!c->have_codensity)
            return fail(c, NW_E_STATE, "the Fock refresh needs the co-densities of an nw_build (not a loaded cache)");
        if (!fa || !fb) return fail(c, NW_E_ARG, "null Fock matrix");
        const int n = c->n, nm = c->nmo, nref = c->nref;
        const size_t nn = (size_t)n * n, nnm = (size_t)n * nm;
        c->fock_gen++;
        double *F = c->fockbuf.get<double>(2 * nn);
        HIPCHK(hipMemcpyAsync(F, fa, nn * 8, hipMemcpyDefault, c->st));
        HIPCHK(hipMemcpyAsync(F + nn, fb, nn * 8, hipMemcpyDefault, c->st));
        std::vector<int> built;
        for (int k = 0; k < nref * nref; ++k)
            if (c->pinfo[k].built) built.push_back(k);
        const int P = (int)built.size();
        int maxb = 0;
        for (int k : built) maxb = std::max(maxb, c->pinfo[k].bidx + 1);
        double *M = (double *)c->keepM.p, *cx = (double *)c->keepCX.p, *xc = (double *)c->keepXC.p;
        Uploader up(c, 64u << 20);
        double *tmp = c->scratchA.get<double>((size_t)8 * maxb * nnm);
        double *dmeta = c->scratchB.get<double>((size_t)P * 4);
        std::vector<GemmDesc> gfx, gff;
        std::vector<TraceDesc> tr;
        for (int q = 0; q < P; ++q) {
            const PairInfo &pi = c->pinfo[built[q]];
            const int p = pi.bidx;
            for (int sp = 0; sp < 2; ++sp) {
                const int pr = p * 2 + sp;
                const double *Fs = F + sp * nn;
                const long long *o = pi.off + (sp ? O_BB : O_AA);
                for (int mi = 0; mi < 2; ++mi)
                    tr.push_back({M + (size_t)(pr * 2 + mi) * nn, Fs, dmeta + (size_t)q * 4 + sp * 2 + mi, 1.0});
                for (int mj = 0; mj < 2; ++mj)
                    gfx.push_back({Fs, xc + (size_t)(pr * 2 + mj) * nnm, tmp + (size_t)(pr * 2 + mj) * nnm});
                for (int mi = 0; mi < 2; ++mi)
                    for (int mj = 0; mj < 2; ++mj)
                        gff.push_back({tmp + (size_t)(pr * 2 + mj) * nnm, cx + (size_t)(pr * 2 + mi) * nnm,
                                       c->d_slab + pi.base + o[O_FF + mi * 2 + mj]});
            }
        }
        // f0f[mi] = sum M[b,a] F[a,b]; ff[mi][mj] = cx[mi]^T F xc[mj] (construct_f_scalar, build.rs:564-617)
        launch_einsum(c->st, up.put(tr), (int)tr.size(), n);
        gemm(c, up, 0, 0, n, nm, n, n, nm, nm, 1.0, 0.0, gfx);
        gemm(c, up, 1, 0, nm, nm, n, nm, nm, nm, 1.0, 0.0, gff);
        std::vector<double> mh((size_t)P * 4);
        HIPCHK(hipMemcpyAsync(mh.data(), dmeta, mh.size() * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        for (int q = 0; q < P; ++q) {
            PairInfo &pi = c->pinfo[built[q]];
            for (int sp = 0; sp < 2; ++sp)
                for (int mi = 0; mi < 2; ++mi) pi.meta[(sp ? M_BB : M_AA) + M_F0F + mi] = mh[(size_t)q * 4 + sp * 2 + mi];
        }
        HIPCHK(hipMemcpyAsync(c->d_pinfo, c->pinfo.data(), sizeof(PairInfo) * c->pinfo.size(), hipMemcpyHostToDevice,
                              c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        return NW_OK;
    });
}

int nw_matvec_m(nw_ctx *c, const nw_det *dets, uint64_t ndet, double e0, const double *x, double *y, double *ys,
                int mode)
{
    return guarded(c, [&]() -> int {
        if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
        if (!dets || !x || !y || ndet == 0) return fail(c, NW_E_ARG, "null arguments");
        HIPCHK(hipEventRecord(c->ev[0], c->st));
        PackedDet *dp = nullptr;
        std::vector<PackedDet> h = pack(c, dets, ndet, dp);
        auto bins = triangle_blocks(c, h, false);
        const double *dx = (const double *)to_device(c, c->outH, x, ndet * 8);
        const bool ydev = is_device_ptr(y), ysdev = ys && is_device_ptr(ys);
        double *dy = ydev ? y : c->outS.get<double>(2 * ndet);
        double *dys = ys ? (ysdev ? ys : c->outS.get<double>(2 * ndet) + ndet) : nullptr;
        std::vector<LinDesc> lz = {{nullptr, nullptr, nullptr, dy, 0, 0, 0}};
        if (dys) lz.push_back({nullptr, nullptr, nullptr, dys, 0, 0, 0});
        Uploader up(c, 64u << 20);
        launch_lincomb(c->st, up.put(lz), (int)lz.size(), (long long)ndet);
        EvalArgs args = base_args(c, dp, mode);
        args.e0 = e0;
        args.x = dx;
        args.y = dy;
        args.ys = dys;
        HIPCHK(hipEventRecord(c->ev[1], c->st));
        for (auto &kv : bins) {
            int la, lb;
            key_ranks(kv.first, la, lb);
            EvalBlock *db = up.put(kv.second);
            launch_eval_blocks(c->st, EVAL_MATVEC, la, lb, false, db, (int)kv.second.size(),
                               tiles_of(kv.second), args);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[2], c->st));
        if (!ydev) HIPCHK(hipMemcpyAsync(y, dy, ndet * 8, hipMemcpyDeviceToHost, c->st));
        if (ys && !ysdev) HIPCHK(hipMemcpyAsync(ys, dys, ndet * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->t_ms[0] = elapsed(c->ev[0], c->ev[2]);
        c->t_ms[1] = 0.0;
        c->t_ms[2] = elapsed(c->ev[1], c->ev[2]);
        return NW_OK;
    });
}

int nw_project_hs(nw_ctx *c, const nw_det *dets, uint64_t ndet, uint64_t ncur, const double *coef, double *v,
                  double *s, int mode)
{
    return guarded(c, [&]() -> int {
        if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
        if (!dets || !coef || !v || !s || ncur == 0 || ncur >= ndet) return fail(c, NW_E_ARG, "bad arguments");
        if (ndet > (1ull << 31)) return fail(c, NW_E_SHAPE, "too many determinants");
        const uint64_t nc = ndet - ncur;
        HIPCHK(hipEventRecord(c->ev[0], c->st));
        PackedDet *dp = nullptr;
        std::vector<PackedDet> h = pack(c, dets, ndet, dp);
        chol_ensure(c, h);
        auto bins = rect_blocks(c, h, (long long)ncur, (long long)ncur, true);
        // Each block (candidate row run x current column run) sums its rows
        // over its columns into its own slots; a row's projections are then its
        // blocks' sums in column-run order (k_project_reduce), a fixed order.
        std::vector<long long> ri0, cj0;  // row and column runs by first determinant
        long long nslot = 0;
        for (auto &kv : bins) {
            long long t = 0;
            for (EvalBlock &b : kv.second) {
                b.tstart = t;  // strips of 16 rows
                t += (b.ni + 15) / 16;
                b.pslot = nslot;
                nslot += b.ni;
                ri0.push_back(b.i0);
                cj0.push_back(b.j0);
            }
        }
        for (auto *r : {&ri0, &cj0}) {
            std::sort(r->begin(), r->end());
            r->erase(std::unique(r->begin(), r->end()), r->end());
        }
        const int nrun = (int)ri0.size(), ncrun = (int)cj0.size();
        std::vector<long long> table((size_t)nrun * ncrun, -1);
        for (auto &kv : bins)
            for (const EvalBlock &b : kv.second) {
                const size_t k = std::lower_bound(ri0.begin(), ri0.end(), b.i0) - ri0.begin();
                const size_t col = std::lower_bound(cj0.begin(), cj0.end(), b.j0) - cj0.begin();
                table[k * ncrun + col] = b.pslot;
            }
        const double *dx = (const double *)to_device(c, c->outH, coef, ncur * 8);
        const bool vdev = is_device_ptr(v), sdev = is_device_ptr(s);
        double *dv = vdev ? v : c->fockbuf.get<double>(2 * nc);
        double *ds = sdev ? s : (vdev ? c->fockbuf.get<double>(nc) : dv + nc);
        double *part = c->scratchA.get<double>(2 * (size_t)std::max<long long>(1, nslot));
        Uploader up(c, (64u << 20) + (ri0.size() + table.size()) * sizeof(long long));
        EvalArgs args = base_args(c, dp, mode);
        args.x = dx;
        args.y = part;
        args.ys = part + nslot;
        HIPCHK(hipEventRecord(c->ev[1], c->st));
        for (auto &kv : bins) {
            const int key = kv.first;
            int la, lb;
            key_ranks(key, la, lb);
            EvalBlock *db = up.put(kv.second);
            const EvalBlock &lastb = kv.second.back();
            launch_eval_blocks(c->st, EVAL_PROJECT, la, lb, key & 1, db, (int)kv.second.size(),
                               lastb.tstart + (lastb.ni + 15) / 16, args);
        }
        launch_project_reduce(c->st, part, part + nslot, up.put(ri0), nrun, up.put(table), ncrun, (long long)ncur,
                              (long long)nc, dv, ds);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[2], c->st));
        if (!vdev) HIPCHK(hipMemcpyAsync(v, dv, nc * 8, hipMemcpyDeviceToHost, c->st));
        if (!sdev) HIPCHK(hipMemcpyAsync(s, ds, nc * 8, hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->t_ms[0] = elapsed(c->ev[0], c->ev[2]);
        c->t_ms[1] = 0.0;
        c->t_ms[2] = elapsed(c->ev[1], c->ev[2]);
        return NW_OK;
    });
}

int nw_set_mixed_floor(nw_ctx *c, double sigma_floor)
{
    return guarded(c, [&]() -> int {
        if (!(sigma_floor >= 0.0) || !std::isfinite(sigma_floor)) return fail(c, NW_E_ARG, "bad sigma floor");
        c->mixed_floor = sigma_floor;
        c->built = false;
        return NW_OK;
    });
}

int nw_set_precision(nw_ctx *c, int precision)
{
    return guarded(c, [&]() -> int {
        if (precision != NW_PREC_FP64 && precision != NW_PREC_MIXED && precision != NW_PREC_BF16X3)
            return fail(c, NW_E_ARG, "unknown precision");
        c->precision = precision;
        c->built = false;
        return NW_OK;
    });
}

int nw_pair_slab(nw_ctx *c, int x, int w, double *out, int64_t *offsets, double *meta)
{
    return guarded(c, [&]() -> int {
        if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
        if (x < 0 || w < 0 || x >= c->nref || w >= c->nref) return fail(c, NW_E_ARG, "pair out of range");
        const PairInfo &pi = c->pinfo[(size_t)x * c->nref + w];
        if (!pi.built) return fail(c, NW_E_STATE, "pair not built");
        if (out) chol_complete(c);  // every rank-4 block of on-demand pairs
        // the reference layout, whatever precision the pair was built in
        long long ref[O_LEN];
        pair_offsets(pi.ma, pi.mb, c->nmo, ref);
        if (offsets)
            for (int k = 0; k < O_LEN; ++k) offsets[k] = ref[k];
        if (meta)
            for (int k = 0; k < M_LEN; ++k) meta[k] = pi.meta[k];
        if (out) {
            std::vector<double> blk((size_t)pi.off[O_TOTAL]);
            HIPCHK(hipMemcpyAsync(blk.data(), c->d_slab + pi.base, blk.size() * 8, hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            export_pair(pi, c->nmo, blk.data(), out);
        }
        return NW_OK;
    });
}

int nw_last_timings(nw_ctx *c, double *ms3)
{
    if (!c || !ms3) return NW_E_ARG;
    for (int k = 0; k < 3; ++k) ms3[k] = c->t_ms[k];
    return NW_OK;
}

const char *nw_last_error(nw_ctx *c) { return c ? c->err.c_str() : "null context"; }

void nw_destroy(nw_ctx *c)
{
    if (!c) return;
    destroy_factors(c);
    delete c;
}

}  // extern "C"
