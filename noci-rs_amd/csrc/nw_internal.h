// Internal host-side definitions shared by the C ABI translation units:
// error handling, device buffers, the context object and the determinant
// packing / pair-block planning used by every evaluation entry point.
#pragma once
#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nwick.h"
#include "nw_kernels.h"

namespace nwi {
using namespace nw;


struct NwError : std::runtime_error {
    int code;
    NwError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw NwError(NW_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));                \
    } while (0)

template <class T>
T *dalloc(size_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw NwError(NW_E_MEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
    }
    return (T *)p;
}

struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DBuf() { if (p) (void)hipFree(p); }
    template <class T>
    T *get(size_t n)
    {
        size_t need = std::max<size_t>(n * sizeof(T), 8);
        if (need > bytes) {
            if (p) HIPCHK(hipFree(p));
            p = nullptr;
            bytes = 0;
            HIPCHK(hipMalloc(&p, need));
            bytes = need;
        }
        return (T *)p;
    }
};

// Pinned host staging for descriptor uploads: a bump allocator over chunks of
// hipHostMalloc memory, reset at the start of every API call (each call ends
// with its stream drained, so no copy from an earlier call is pending).
// Copies from pinned memory are asynchronous DMA; from pageable memory the
// runtime stages them and the host can wait on work queued before.
struct PinnedArena {
    std::vector<std::pair<char *, size_t>> chunks;
    size_t ci = 0, used = 0;
    char *get(size_t bytes)
    {
        bytes = (bytes + 255) & ~size_t(255);
        while (ci < chunks.size() && used + bytes > chunks[ci].second) {
            ++ci;
            used = 0;
        }
        if (ci == chunks.size()) {
            const size_t sz = std::max<size_t>(bytes, size_t(32) << 20);
            void *p = nullptr;
            HIPCHK(hipHostMalloc(&p, sz, hipHostMallocDefault));
            chunks.push_back({(char *)p, sz});
            used = 0;
        }
        char *p = chunks[ci].first + used;
        used += bytes;
        return p;
    }
    void reset()
    {
        ci = 0;
        used = 0;
    }
    ~PinnedArena()
    {
        for (auto &ch : chunks) (void)hipHostFree(ch.first);
    }
};

inline bool is_device_ptr(const void *p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

}  // namespace nwi


namespace nwi {
// Destroys the factor objects still alive on a context (nw_factor.hip).
void destroy_factors(nw_ctx *c);
// Integral factors and on-demand rank-4 blocks of a context (nw_chol.hip).
struct CholState;
void destroy_chol(nw_ctx *c);
}  // namespace nwi

struct nw_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev[4] = {};
    std::mutex mu;
    std::string err;
    int n = 0;
    double enuc = 0.0;
    double *d_S = nullptr, *d_h = nullptr, *d_eri = nullptr, *d_easym = nullptr;
    float *d_eri32 = nullptr, *d_easym32 = nullptr;  // FP32 copies for the mixed build
    int precision = NW_PREC_FP64;
    double mixed_floor = 0.1;  // pairs with a nonzero singular value below this build in FP64
    int nref = 0, nmo = 0, nocc[2] = {0, 0};
    double *d_c[2] = {nullptr, nullptr};  // nref x n x nmo per spin
    // build state
    bool built = false;
    unsigned long long build_gen = 0;  // incremented by every nw_build / nw_set_refs
    unsigned long long fock_gen = 0;   // incremented by every nw_build / nw_update_fock / nw_load_wicks
    bool have_codensity = false;       // co-densities and cx/xc of the last nw_build are kept (nw_update_fock)
    double tol = 1e-8;
    double *d_slab = nullptr;
    size_t slab_cap = 0;
    std::set<nw_factor *> factors;  // live factor objects on this context (under mu)
    std::vector<nw::PairInfo> pinfo;
    nw::PairInfo *d_pinfo = nullptr;
    size_t pinfo_cap = 0;
    double t_ms[3] = {0, 0, 0};
    // scratch
    nwi::DBuf buildscratch, arena_desc, scratchA, scratchB, cx32, xc32, keepM, keepCX, keepXC, fockbuf, dets_raw, dets_packed, tiles_buf, outH, outS, flag;
    nwi::DBuf req_in, req_key, req_bins, req_sorted;  // batched requests binned on the GPU
    nwi::DBuf splitk;                                 // partial products of the split-K J - K contractions
    std::deque<std::vector<char>> keep;
    nwi::PinnedArena pinned;  // staging of Uploader copies
    nwi::CholState *chol = nullptr;  // integral factors (nw_set_eri_factors) and block coverage

    ~nw_ctx()
    {
        (void)hipSetDevice(device);
        if (st) (void)hipStreamSynchronize(st);
        nwi::destroy_chol(this);
        for (double *p : {d_S, d_h, d_eri, d_easym, d_c[0], d_c[1], d_slab})
            if (p) (void)hipFree(p);
        for (float *p : {d_eri32, d_easym32})
            if (p) (void)hipFree(p);
        if (d_pinfo) (void)hipFree(d_pinfo);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};


namespace nwi {


// Upload host descriptor arrays into a per-call device arena (append only).
struct Uploader {
    nw_ctx *c;
    char *base;
    size_t cap, used = 0;
    Uploader(nw_ctx *ctx, size_t bytes) : c(ctx)
    {
        base = c->arena_desc.get<char>(bytes);
        cap = bytes;
    }
    template <class T>
    T *put(const std::vector<T> &v)
    {
        size_t bytes = v.size() * sizeof(T);
        size_t off = (used + 255) & ~size_t(255);
        if (off + bytes > cap) throw NwError(NW_E_MEM, "descriptor arena exhausted");
        if (bytes) {
            char *h = c->pinned.get(bytes);
            std::memcpy(h, v.data(), bytes);
            HIPCHK(hipMemcpyAsync(base + off, h, bytes, hipMemcpyHostToDevice, c->st));
        }
        used = off + bytes;
        return (T *)(base + off);
    }
};

inline int fail(nw_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

template <class F>
int guarded(nw_ctx *c, F &&f)
{
    if (!c) return NW_E_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    try {
        hipError_t e = hipSetDevice(c->device);
        if (e != hipSuccess) return fail(c, NW_E_HIP, hipGetErrorString(e));
        c->keep.clear();
        if (c->st) HIPCHK(hipStreamSynchronize(c->st));  // no copy from the staging is pending
        c->pinned.reset();
        int rc = f();
        c->keep.clear();
        return rc;
    } catch (const NwError &e) {
        (void)hipStreamSynchronize(c->st);
        c->keep.clear();
        return fail(c, e.code, e.what());
    } catch (const std::exception &e) {
        c->keep.clear();
        return fail(c, NW_E_ARG, e.what());
    }
}

// Copy count*size bytes to a device buffer unless the pointer is already device memory.
inline const void *to_device(nw_ctx *c, DBuf &buf, const void *p, size_t bytes)
{
    if (is_device_ptr(p)) return p;
    char *d = buf.get<char>(bytes);
    HIPCHK(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, c->st));
    return d;
}

inline float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.f;
    }
    return ms;
}


struct Run { long long start, len; int parent, ra, rb; };
long long pair_offsets(int ma, int mb, int nmo, long long *off, bool r4_f32 = false);
int build_impl(nw_ctx *c, double tol, const uint8_t *pair_mask, nw_pair_meta *meta_out, std::vector<long long> *sizes);
void export_pair(const PairInfo &pi, int nmo, const double *blk, double *out);
std::vector<PackedDet> pack(nw_ctx *c, const nw_det *dets, uint64_t ndet, PackedDet *&dpacked,
                            const void **draw = nullptr);
int bin_key(int la, int lb, bool m0);
std::map<int, std::vector<EvalBlock>> triangle_blocks(nw_ctx *c, const std::vector<PackedDet> &h, bool need_hcol,
                                                      bool built_only = false);
long long tiles_of(const std::vector<EvalBlock> &v);
std::map<int, std::vector<EvalBlock>> rect_blocks(nw_ctx *c, const std::vector<PackedDet> &h, long long nrow0,
                                                  long long ncol, bool need_hcol);
EvalArgs base_args(nw_ctx *c, const PackedDet *dp, int mode);

// The four-index transforms of every rank-4 tensor of the given built pairs
// (keys x*nref + w) from the co-densities of the context's last build.
void run_rank4_transforms(nw_ctx *c, const std::vector<int> &keys);

// On-demand rank-4 blocks (nw_chol.hip).  chol_route: pairs of the build
// about to run may leave their rank-4 blocks to chol_ensure.  chol_on_build:
// record which pairs of the new build did (lazy[p] for build index p).
// chol_ensure: build every block the packed determinants h read on any
// pair (superset: every determinant of a pair's parents as bra and ket), or
// on the pairs marked in reads (x * nref + w); the stage times and flops it
// records add to the previous call's when accumulate is set (the groups of
// one streamed call).  triangle_pairs: the ordered reference pairs an
// upper-triangle evaluation of h reads.
// chol_missing: true if chol_ensure would have work for h (no device
// allocation).  chol_complete: every rank-4 block of every built pair.
bool chol_route(nw_ctx *c);
void chol_on_build(nw_ctx *c, const std::vector<int> &keys, const std::vector<uint8_t> &lazy);
void chol_ensure(nw_ctx *c, const std::vector<PackedDet> &h, bool accumulate = false,
                 const std::vector<uint8_t> *reads = nullptr);
std::vector<uint8_t> triangle_pairs(const nw_ctx *c, const std::vector<PackedDet> &h);
bool chol_missing(nw_ctx *c, const std::vector<PackedDet> &h);
void chol_complete(nw_ctx *c);
void chol_forget(nw_ctx *c);  // a loaded slab: every pair complete
double chol_last_ms(nw_ctx *c);  // device time of the last chol_ensure's blocks
bool chol_any_lazy(nw_ctx *c);
// Packed records of host determinants (ranks and ascending holes / particles
// per spin; ranks above 4 are clipped), for host-side planning.
std::vector<PackedDet> host_pack(const nw_det *dets, uint64_t ndet);

}  // namespace nwi
