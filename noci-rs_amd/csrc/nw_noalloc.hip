// nw_hs_batched_noalloc: batched H,S requests with no device allocation.
//
// The same kernels as nw_hs_batched (determinant packing, request binning,
// per-bin evaluation), with every temporary -- packed determinants, request
// keys, the binned request list, per-bin counts, cursors, offsets and error
// flags -- in one file-scope device array of fixed size.  Each GPU holds its
// own copy of the array from the first call on it until the process ends.
// The array's host-side name is not a device address: its device address is
// taken once per device with hipGetSymbolAddress and only that address is
// passed to kernels and copies.
#include <mutex>

#include "nw_internal.h"

using namespace nw;
using namespace nwi;

namespace {

constexpr size_t NA_MAX_DET = size_t(1) << 18;
constexpr size_t NA_MAX_REQ = size_t(1) << 20;
constexpr size_t round256(size_t b) { return (b + 255) & ~size_t(255); }
// scratch layout (byte offsets)
constexpr size_t NA_PACKED = 0;
constexpr size_t NA_KEY = NA_PACKED + round256(NA_MAX_DET * sizeof(PackedDet));
constexpr size_t NA_SORTED = NA_KEY + round256(NA_MAX_REQ * sizeof(int));
constexpr size_t NA_BINS = NA_SORTED + round256(NA_MAX_REQ * sizeof(EvalReq));
// bins: hist[REQ_NKEY], cursor[REQ_NKEY], err, pack flag, then start[REQ_NKEY]
constexpr int NA_NCLEAR = 2 * REQ_NKEY + 2;
constexpr size_t NA_START = NA_BINS + round256(NA_NCLEAR * sizeof(unsigned));
constexpr size_t NA_BYTES = NA_START + round256(REQ_NKEY * sizeof(unsigned));
static_assert(sizeof(PackedDet) == 24 && sizeof(EvalReq) == 16, "scratch layout");

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu[NA_MAX_DEVICES];      // serialises the calls on one device within the process
unsigned char *na_ptr[NA_MAX_DEVICES];  // device address of `scratch` per device (under na_mu)

}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

__global__ void k_na_clear(unsigned *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// The device address of this GPU's scratch array (the current device is the
// context's); called with na_mu[device] held.
unsigned char *na_scratch(int device)
{
    if (!na_ptr[device]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        na_ptr[device] = (unsigned char *)p;
    }
    return na_ptr[device];
}

}  // namespace

extern "C" int nw_hs_batched_noalloc(nw_ctx *c, const nw_det *dets, uint64_t ndet, const uint64_t *req, uint64_t nreq,
                                     double *h_out, double *s_out, int mode)
{
    return guarded(c, [&]() -> int {
        if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
        if (!dets || !req || !h_out || !s_out) return fail(c, NW_E_ARG, "null arguments");
        if (c->device < 0 || c->device >= NA_MAX_DEVICES) return fail(c, NW_E_ARG, "device index out of range");
        if (ndet > NA_MAX_DET || nreq > NA_MAX_REQ)
            return fail(c, NW_E_SHAPE, "nw_hs_batched_noalloc takes at most 2^18 determinants and 2^20 requests");
        if (nreq == 0) return NW_OK;
        if (!is_device_ptr(dets) || !is_device_ptr(req) || !is_device_ptr(h_out) || !is_device_ptr(s_out))
            return fail(c, NW_E_ARG, "nw_hs_batched_noalloc needs device arrays for dets, req, h_out and s_out");
        if (chol_any_lazy(c)) {
            // pairs built from integral factors: the blocks these determinants read
            // must exist already (no device allocation here to build them)
            std::vector<nw_det> hd(ndet);
            HIPCHK(hipMemcpyAsync(hd.data(), dets, ndet * sizeof(nw_det), hipMemcpyDeviceToHost, c->st));
            HIPCHK(hipStreamSynchronize(c->st));
            if (chol_missing(c, host_pack(hd.data(), ndet)))
                return fail(c, NW_E_STATE, "rank-4 blocks these determinants read are not built yet (pairs built "
                                           "from integral factors): evaluate them once with nw_hs_batched first");
        }
        std::lock_guard<std::mutex> lk(na_mu[c->device]);
        unsigned char *base = na_scratch(c->device);
        PackedDet *dp = (PackedDet *)(base + NA_PACKED);
        unsigned *bins = (unsigned *)(base + NA_BINS);
        unsigned *dstart = (unsigned *)(base + NA_START);
        int *perr = (int *)(bins + 2 * REQ_NKEY), *pflag = perr + 1;
        HIPCHK(hipEventRecord(c->ev[0], c->st));
        // every count, cursor and flag starts at zero; the rest is overwritten before it is read
        hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(256), 0, c->st, bins, NA_NCLEAR);
        launch_pack_dets(c->st, dets, (long long)ndet, c->nref, dp, pflag);
        // calculate_hs_pairs_wicks_batched (src/noci/hs.rs:86-109): exact zeros
        // and contraction-rank bins on the GPU, as nw_hs_batched
        ReqBin rb{};
        rb.req = req;
        rb.nreq = (long long)nreq;
        rb.ndet = (long long)ndet;
        rb.raw = dets;
        rb.dp = dp;
        rb.pairs = c->d_pinfo;
        rb.nref = c->nref;
        rb.hs = 1;
        rb.key = (int *)(base + NA_KEY);
        rb.hist = bins;
        rb.cursor = bins + REQ_NKEY;
        rb.err = perr;
        rb.A = h_out;
        rb.B = s_out;
        rb.out = (EvalReq *)(base + NA_SORTED);
        launch_req_bin(c->st, rb, false);
        HIPCHK(hipGetLastError());
        unsigned hb[REQ_NKEY + 2];
        HIPCHK(hipMemcpyAsync(hb, bins, REQ_NKEY * sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipMemcpyAsync(hb + REQ_NKEY, perr, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        if (hb[REQ_NKEY + 1] & 2) throw NwError(NW_E_ARG, "determinant parent out of range");
        if (hb[REQ_NKEY + 1] & 1)
            throw NwError(NW_E_RANK, "excitation rank per spin must be <= 4 with equal holes and particles");
        if (hb[REQ_NKEY] == REQ_ERR_INDEX) throw NwError(NW_E_ARG, "request index out of range");
        if (hb[REQ_NKEY] == REQ_ERR_UNBUILT) throw NwError(NW_E_STATE, "reference pair not built");
        if (hb[REQ_NKEY] == REQ_ERR_RANK)
            throw NwError(NW_E_RANK, "contraction rank per spin (bra plus ket excitations) must be <= 8");
        unsigned start[REQ_NKEY];
        unsigned tot = 0;
        for (int k = 0; k < REQ_NKEY; ++k) {
            start[k] = tot;
            tot += hb[k];
        }
        HIPCHK(hipMemcpyAsync(dstart, start, sizeof(start), hipMemcpyHostToDevice, c->st));
        rb.start = dstart;
        launch_req_bin(c->st, rb, true);
        EvalArgs args = base_args(c, dp, mode);
        args.A = h_out;
        args.B = s_out;
        HIPCHK(hipEventRecord(c->ev[1], c->st));
        for (int k = 0; k < REQ_NKEY; ++k) {
            if (!hb[k]) continue;
            int la, lb;
            key_ranks(k, la, lb);
            launch_eval_reqs(c->st, EVAL_HS, la, lb, k & 1, rb.out + start[k], (int)hb[k], args);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[2], c->st));
        HIPCHK(hipStreamSynchronize(c->st));
        c->t_ms[0] = elapsed(c->ev[0], c->ev[2]);
        c->t_ms[1] = 0.0;
        c->t_ms[2] = elapsed(c->ev[1], c->ev[2]);
        return NW_OK;
    });
}
