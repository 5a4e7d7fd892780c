// Multi-system batched H,S evaluation (nw_batch_* in include/nwick.h).
//
// Several built contexts (one per molecular system, each with its own slab)
// and their determinant lists are planned into one tile list that interleaves
// every system's upper-triangle determinant pairs.  The plan runs as
// persistent fused kernel launches (k_eval_fused) on up to four streams,
// captured once into a HIP graph and replayed.  A tuner times tile shapes per
// occupation bucket (system n_occ x contraction-rank code) and then the
// schedule (kernel grouping, static or dynamic work distribution, persistent
// grid size) on the actual batch.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <numeric>

#include "nw_internal.h"

using namespace nw;
using namespace nwi;

namespace {

constexpr int NSTREAM = 4;
constexpr int SHAPES[] = {4, 2, 3, 5, 6};  // log2(TJ); 4 = 16 x 16 is the default
constexpr int DEFAULT_SHAPE = 4;

// Algorithmic bytes of one m = 0 element of contraction ranks (la, lb):
// X/Y gathers, J and II entries, H/S out and the two determinant records.
double elem_bytes(int code)
{
    const int lab = code >> 1, la = lab / 5, lb = lab % 5;
    auto c2 = [](int l) { return l * (l - 1) / 2; };
    return 8.0 * (2 * (la * la + lb * lb) + 2 * c2(la) * c2(la) + 2 * c2(lb) * c2(lb) + la * la * lb * lb) + 60.0;
}

struct BSys {
    nw_ctx *ctx = nullptr;
    unsigned long long gen = 0;  // the context's build generation at nw_batch_add
    uint64_t ndet = 0;
    int mode = 0, nocc = 0;
    PackedDet *dets = nullptr;
    double *H = nullptr, *S = nullptr;
    std::map<int, std::vector<EvalBlock>> bins;  // code -> blocks
};

struct Launch {
    int sel;
    long long off, n;  // tile range in the device tile array
    double cost;
    int stream;
};

}  // namespace

struct nw_batch {
    int device = 0, ncu = 256;
    hipStream_t st[NSTREAM] = {};
    hipEvent_t ev[2] = {}, fork = nullptr, join[NSTREAM] = {};
    std::mutex mu;
    std::string err;
    std::vector<BSys> sys;
    BatchSys *d_sys = nullptr;
    // schedule
    std::map<std::pair<int, int>, int> shape;  // (nocc, code) -> log2(TJ)
    int grouping = 1;                          // NW_GROUP_CODE / CLASS / ALL
    int dyn = 1, grid_per_cu = 4;
    std::vector<nw_tune_entry> report;
    // plan
    bool planned = false;
    std::vector<Launch> launches;
    DBuf tiles, counters;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double last_ms = 0.0;
    long long npairs = 0;

    ~nw_batch()
    {
        (void)hipSetDevice(device);
        for (auto &s : st)
            if (s) (void)hipStreamSynchronize(s);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &s : sys) {
            for (void *p : {(void *)s.dets, (void *)s.H, (void *)s.S})
                if (p) (void)hipFree(p);
        }
        if (d_sys) (void)hipFree(d_sys);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        for (auto &e : join)
            if (e) (void)hipEventDestroy(e);
        for (auto &s : st)
            if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

int bfail(nw_batch *b, int code, const std::string &m)
{
    if (b) b->err = m;
    return code;
}

// The plan addresses each context's slab and pair table directly, so a
// context rebuilt after nw_batch_add would leave it pointing at freed memory.
void check_generations(nw_batch *b)
{
    for (size_t k = 0; k < b->sys.size(); ++k)
        if (b->sys[k].ctx->build_gen != b->sys[k].gen || !b->sys[k].ctx->built)
            throw NwError(NW_E_STATE, "system " + std::to_string(k) + ": context rebuilt after nw_batch_add");
}

template <class F>
int bguard(nw_batch *b, F &&f)
{
    if (!b) return NW_E_ARG;
    std::lock_guard<std::mutex> lk(b->mu);
    try {
        hipError_t e = hipSetDevice(b->device);
        if (e != hipSuccess) return bfail(b, NW_E_HIP, hipGetErrorString(e));
        return f();
    } catch (const NwError &e) {
        for (auto &s : b->st) (void)hipStreamSynchronize(s);
        return bfail(b, e.code, e.what());
    } catch (const std::exception &e) {
        return bfail(b, NW_E_ARG, e.what());
    }
}

int shape_of(const nw_batch *b, int nocc, int code)
{
    auto it = b->shape.find({nocc, code});
    return it == b->shape.end() ? DEFAULT_SHAPE : it->second;
}

// Cut the blocks of one (system, code) into tiles of TI x TJ = 256 pairs.
void tile_blocks(int sysi, int code, int sh, const std::vector<EvalBlock> &blocks, std::vector<BatchTile> &out,
                 std::vector<double> *cost)
{
    const int TJ = 1 << sh, TI = 256 >> sh;
    const double eb = elem_bytes(code);
    for (const EvalBlock &B : blocks) {
        for (int ti = 0; ti * TI < B.ni; ++ti)
            for (int tj = 0; tj * TJ < B.nj; ++tj) {
                const long long r0 = B.i0 + (long long)ti * TI, c0 = B.j0 + (long long)tj * TJ;
                if (B.tri && c0 + TJ - 1 < r0) continue;  // entirely below the diagonal
                BatchTile t{};
                t.i0 = r0;
                t.j0 = c0;
                t.sys = sysi;
                t.pair = B.pair;
                t.ni = (short)std::min(TI, B.ni - ti * TI);
                t.nj = (short)std::min(TJ, B.nj - tj * TJ);
                t.code = (unsigned char)code;
                t.shape = (unsigned char)sh;
                t.tri = (unsigned char)B.tri;
                out.push_back(t);
                if (cost) cost->push_back(eb * t.ni * t.nj);
            }
    }
}

int sel_of(int grouping, int code)
{
    if (grouping == NW_GROUP_CODE) return code;
    if (grouping == NW_GROUP_CLASS) return -1 - fused_class(code);
    return FUSE_ALL;
}

// Build the launch list and upload the tiles for the current schedule.
void make_plan(nw_batch *b)
{
    std::map<int, std::vector<std::pair<double, BatchTile>>> groups;  // sel -> (cost, tile)
    b->npairs = 0;
    for (size_t s = 0; s < b->sys.size(); ++s) {
        const BSys &S = b->sys[s];
        for (auto &kv : S.bins) {
            std::vector<BatchTile> t;
            std::vector<double> c;
            tile_blocks((int)s, kv.first, shape_of(b, S.nocc, kv.first), kv.second, t, &c);
            auto &g = groups[sel_of(b->grouping, kv.first)];
            for (size_t k = 0; k < t.size(); ++k) g.push_back({c[k], t[k]});
        }
        b->npairs += (long long)S.ndet * (long long)(S.ndet + 1) / 2;
    }
    std::vector<BatchTile> all;
    b->launches.clear();
    for (auto &kv : groups) {
        auto &g = kv.second;
        // costliest tiles first: static striding then hands every workgroup
        // a similar mix, and the dynamic counter ends on the cheap tail
        std::stable_sort(g.begin(), g.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
        Launch L{kv.first, (long long)all.size(), (long long)g.size(), 0.0, 0};
        for (auto &p : g) {
            all.push_back(p.second);
            L.cost += p.first;
        }
        b->launches.push_back(L);
    }
    // greedy balance of launches over the streams, largest first
    std::vector<size_t> order(b->launches.size());
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return b->launches[x].cost > b->launches[y].cost; });
    double load[NSTREAM] = {0, 0, 0, 0};
    for (size_t k : order) {
        int s = (int)(std::min_element(load, load + NSTREAM) - load);
        b->launches[k].stream = s;
        load[s] += b->launches[k].cost;
    }
    BatchTile *dt = b->tiles.get<BatchTile>(std::max<size_t>(all.size(), 1));
    if (!all.empty())
        HIPCHK(hipMemcpyAsync(dt, all.data(), all.size() * sizeof(BatchTile), hipMemcpyHostToDevice, b->st[0]));
    b->counters.get<unsigned long long>(std::max<size_t>(b->launches.size(), 1));
    HIPCHK(hipStreamSynchronize(b->st[0]));
    if (b->exec) {
        HIPCHK(hipGraphExecDestroy(b->exec));
        b->exec = nullptr;
    }
    if (b->graph) {
        HIPCHK(hipGraphDestroy(b->graph));
        b->graph = nullptr;
    }
    b->planned = true;
}

// Enqueue the plan: fork the side streams off stream 0, one persistent launch
// per group, join back.  Safe under stream capture (no memset, no host sync).
void enqueue(nw_batch *b)
{
    BatchTile *dt = (BatchTile *)b->tiles.p;
    unsigned long long *dc = (unsigned long long *)b->counters.p;
    const int grid = b->grid_per_cu * b->ncu;
    launch_reset_counters(b->st[0], dc, (int)b->launches.size());  // one zero-fill for every launch's counter
    HIPCHK(hipEventRecord(b->fork, b->st[0]));
    for (int s = 1; s < NSTREAM; ++s) HIPCHK(hipStreamWaitEvent(b->st[s], b->fork, 0));
    for (size_t k = 0; k < b->launches.size(); ++k) {
        const Launch &L = b->launches[k];
        launch_eval_fused(b->st[L.stream], L.sel, b->dyn != 0, dt + L.off, L.n, b->d_sys, dc + k, grid, false);
    }
    HIPCHK(hipGetLastError());
    for (int s = 1; s < NSTREAM; ++s) {
        HIPCHK(hipEventRecord(b->join[s], b->st[s]));
        HIPCHK(hipStreamWaitEvent(b->st[0], b->join[s], 0));
    }
}

double timed(nw_batch *b, int reps)
{
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
        HIPCHK(hipEventRecord(b->ev[0], b->st[0]));
        enqueue(b);
        HIPCHK(hipEventRecord(b->ev[1], b->st[0]));
        HIPCHK(hipEventSynchronize(b->ev[1]));
        best = std::min(best, (double)elapsed(b->ev[0], b->ev[1]));
    }
    return best;
}

// Time one (nocc, code) bucket alone with a given tile shape on a sample of
// its blocks (the same sample for every shape).
double time_bucket(nw_batch *b, const std::vector<std::pair<int, const std::vector<EvalBlock> *>> &parts, int code,
                   int sh)
{
    std::vector<BatchTile> t;
    for (auto &p : parts) tile_blocks(p.first, code, sh, *p.second, t, nullptr);
    if (t.empty()) return 0.0;
    BatchTile *dt = b->tiles.get<BatchTile>(t.size());
    HIPCHK(hipMemcpyAsync(dt, t.data(), t.size() * sizeof(BatchTile), hipMemcpyHostToDevice, b->st[0]));
    unsigned long long *dc = b->counters.get<unsigned long long>(1);
    const int grid = 4 * b->ncu;
    double best = 1e300;
    for (int r = 0; r < 4; ++r) {
        HIPCHK(hipEventRecord(b->ev[0], b->st[0]));
        launch_eval_fused(b->st[0], code, true, dt, (long long)t.size(), b->d_sys, dc, grid);
        HIPCHK(hipEventRecord(b->ev[1], b->st[0]));
        HIPCHK(hipEventSynchronize(b->ev[1]));
        if (r > 0) best = std::min(best, (double)elapsed(b->ev[0], b->ev[1]));
    }
    return best;
}

void tune(nw_batch *b)
{
    b->report.clear();
    b->shape.clear();
    // Phase 1: tile shape per (n_occ, code) bucket.
    std::map<std::pair<int, int>, std::vector<std::pair<int, const std::vector<EvalBlock> *>>> buckets;
    for (size_t s = 0; s < b->sys.size(); ++s)
        for (auto &kv : b->sys[s].bins) buckets[{b->sys[s].nocc, kv.first}].push_back({(int)s, &kv.second});
    std::vector<std::vector<EvalBlock>> samples;
    samples.reserve(4096);
    for (auto &kv : buckets) {
        // sample: every k-th block so that about 2^21 pairs are timed
        long long pairs = 0;
        for (auto &p : kv.second)
            for (auto &B : *p.second) pairs += (long long)B.ni * B.nj;
        const long long target = 1LL << 21;
        const long long stride = std::max<long long>(1, pairs / target);
        std::vector<std::pair<int, const std::vector<EvalBlock> *>> parts;
        for (auto &p : kv.second) {
            samples.emplace_back();
            long long k = 0;
            for (auto &B : *p.second)
                if (k++ % stride == 0) samples.back().push_back(B);
            parts.push_back({p.first, &samples.back()});
        }
        nw_tune_entry e{};
        e.nocc = kv.first.first;
        e.code = kv.first.second;
        e.ms_default = time_bucket(b, parts, e.code, DEFAULT_SHAPE);
        e.shape = DEFAULT_SHAPE;
        e.ms_best = e.ms_default;
        for (int sh : SHAPES) {
            if (sh == DEFAULT_SHAPE) continue;
            double t = time_bucket(b, parts, e.code, sh);
            if (t < e.ms_best * 0.98) {
                e.ms_best = t;
                e.shape = sh;
            }
        }
        b->shape[kv.first] = e.shape;
        b->report.push_back(e);
    }
    // Phase 2: schedule on the whole batch.
    double best = 1e300;
    int bg = 1, bd = 1, bgrid = 4;
    for (int g : {NW_GROUP_CODE, NW_GROUP_CLASS, NW_GROUP_ALL})
        for (int d : {1, 0}) {
            if (g == NW_GROUP_CODE && d == 0) continue;
            for (int gp : {2, 4, 8}) {
                b->grouping = g;
                b->dyn = d;
                b->grid_per_cu = gp;
                make_plan(b);
                timed(b, 1);
                double t = timed(b, 2);
                if (t < best) {
                    best = t;
                    bg = g;
                    bd = d;
                    bgrid = gp;
                }
            }
        }
    b->grouping = bg;
    b->dyn = bd;
    b->grid_per_cu = bgrid;
}

}  // namespace

extern "C" {

int nw_batch_create(nw_batch **out, int device)
{
    if (!out) return NW_E_ARG;
    *out = nullptr;
    nw_batch *b = new nw_batch();
    b->device = device;
    int rc = bguard(b, [&]() -> int {
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        b->ncu = prop.multiProcessorCount;
        for (auto &s : b->st) HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        for (auto &e : b->ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        for (auto &e : b->join) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return NW_OK;
    });
    if (rc != NW_OK) {
        delete b;
        return rc;
    }
    *out = b;
    return NW_OK;
}

int nw_batch_add(nw_batch *b, nw_ctx *ctx, const nw_det *dets, uint64_t ndet, int mode, int *index)
{
    return bguard(b, [&]() -> int {
        if (!ctx || !dets || ndet == 0) return bfail(b, NW_E_ARG, "null arguments");
        if (mode != NW_MODE_REFERENCE && mode != NW_MODE_CORRECTED) return bfail(b, NW_E_ARG, "bad mode");
        if (ctx->device != b->device) return bfail(b, NW_E_ARG, "context on another device");
        if (ndet > (1ull << 20)) return bfail(b, NW_E_SHAPE, "too many determinants for one system");
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (!ctx->built) return bfail(b, NW_E_STATE, "context not built");
        BSys S;
        S.ctx = ctx;
        S.gen = ctx->build_gen;
        S.ndet = ndet;
        S.mode = mode;
        S.nocc = std::max(ctx->nocc[0], ctx->nocc[1]);
        PackedDet *dp = nullptr;
        std::vector<PackedDet> h = pack(ctx, dets, ndet, dp);
        chol_ensure(ctx, h);  // on-demand rank-4 blocks (integral factors)
        S.bins = triangle_blocks(ctx, h, true);
        for (auto &kv : S.bins)
            if (kv.first >= REQ_KEY_GEN)
                throw NwError(NW_E_RANK, "multi-system batches take contraction ranks up to 4 per spin");
        S.dets = dalloc<PackedDet>(ndet);
        HIPCHK(hipMemcpyAsync(S.dets, dp, ndet * sizeof(PackedDet), hipMemcpyDeviceToDevice, ctx->st));
        HIPCHK(hipStreamSynchronize(ctx->st));
        ctx->keep.clear();
        ctx->pinned.reset();
        S.H = dalloc<double>(ndet * ndet);
        S.S = dalloc<double>(ndet * ndet);
        b->sys.push_back(std::move(S));
        std::vector<BatchSys> hs;
        for (auto &x : b->sys) {
            BatchSys d{};
            d.dets = x.dets;
            d.pairs = x.ctx->d_pinfo;
            d.slab = x.ctx->d_slab;
            d.H = x.H;
            d.S = x.S;
            d.ldo = (long long)x.ndet;
            d.nref = x.ctx->nref;
            d.nmo = x.ctx->nmo;
            d.mode = x.mode;
            d.enuc = x.ctx->enuc;
            d.tol = x.ctx->tol;
            hs.push_back(d);
        }
        if (b->d_sys) HIPCHK(hipFree(b->d_sys));
        b->d_sys = nullptr;
        b->d_sys = dalloc<BatchSys>(hs.size());
        HIPCHK(hipMemcpy(b->d_sys, hs.data(), hs.size() * sizeof(BatchSys), hipMemcpyHostToDevice));
        b->planned = false;
        if (index) *index = (int)b->sys.size() - 1;
        return NW_OK;
    });
}

int nw_batch_plan(nw_batch *b, int flags)
{
    return bguard(b, [&]() -> int {
        if (b->sys.empty()) return bfail(b, NW_E_STATE, "no systems added");
        check_generations(b);
        if (flags & NW_BATCH_TUNE) tune(b);
        make_plan(b);
        if (flags & NW_BATCH_GRAPH) {
            HIPCHK(hipStreamBeginCapture(b->st[0], hipStreamCaptureModeThreadLocal));
            try {
                enqueue(b);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(b->st[0], &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIPCHK(hipStreamEndCapture(b->st[0], &b->graph));
            HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        return NW_OK;
    });
}

int nw_batch_run(nw_batch *b)
{
    return bguard(b, [&]() -> int {
        if (!b->planned) return bfail(b, NW_E_STATE, "nw_batch_plan first");
        check_generations(b);
        HIPCHK(hipEventRecord(b->ev[0], b->st[0]));
        if (b->exec) HIPCHK(hipGraphLaunch(b->exec, b->st[0]));
        else enqueue(b);
        HIPCHK(hipEventRecord(b->ev[1], b->st[0]));
        HIPCHK(hipEventSynchronize(b->ev[1]));
        b->last_ms = elapsed(b->ev[0], b->ev[1]);
        return NW_OK;
    });
}

int nw_batch_result(nw_batch *b, int index, double *H, double *S)
{
    return bguard(b, [&]() -> int {
        if (index < 0 || index >= (int)b->sys.size()) return bfail(b, NW_E_ARG, "system index out of range");
        const BSys &X = b->sys[index];
        const size_t bytes = X.ndet * X.ndet * sizeof(double);
        if (H) HIPCHK(hipMemcpyAsync(H, X.H, bytes, hipMemcpyDefault, b->st[0]));
        if (S) HIPCHK(hipMemcpyAsync(S, X.S, bytes, hipMemcpyDefault, b->st[0]));
        HIPCHK(hipStreamSynchronize(b->st[0]));
        return NW_OK;
    });
}

int nw_batch_schedule(nw_batch *b, int *grouping, int *dyn, int *grid_per_cu, nw_tune_entry *entries, int cap,
                      int *count)
{
    return bguard(b, [&]() -> int {
        if (grouping) *grouping = b->grouping;
        if (dyn) *dyn = b->dyn;
        if (grid_per_cu) *grid_per_cu = b->grid_per_cu;
        if (count) *count = (int)b->report.size();
        for (int k = 0; k < cap && k < (int)b->report.size(); ++k) entries[k] = b->report[k];
        return NW_OK;
    });
}

int nw_batch_set_schedule(nw_batch *b, int grouping, int dyn, int grid_per_cu, const nw_tune_entry *entries, int n)
{
    return bguard(b, [&]() -> int {
        if (grouping < NW_GROUP_CODE || grouping > NW_GROUP_ALL) return bfail(b, NW_E_ARG, "bad grouping");
        if (grid_per_cu < 1 || grid_per_cu > 64) return bfail(b, NW_E_ARG, "bad grid_per_cu");
        if (grouping == NW_GROUP_CODE && !dyn) return bfail(b, NW_E_ARG, "per-code grouping is dynamic only");
        b->grouping = grouping;
        b->dyn = dyn ? 1 : 0;
        b->grid_per_cu = grid_per_cu;
        b->shape.clear();
        for (int k = 0; k < n; ++k) {
            const int sh = entries[k].shape;
            if (sh < 0 || sh > 8) return bfail(b, NW_E_ARG, "bad tile shape");
            b->shape[{entries[k].nocc, entries[k].code}] = sh;
        }
        b->planned = false;
        return NW_OK;
    });
}

int nw_batch_last_ms(nw_batch *b, double *ms, long long *npairs)
{
    return bguard(b, [&]() -> int {
        if (ms) *ms = b->last_ms;
        if (npairs) *npairs = b->npairs;
        return NW_OK;
    });
}

const char *nw_batch_last_error(nw_batch *b)
{
    return b ? b->err.c_str() : "null batch";
}

void nw_batch_destroy(nw_batch *b)
{
    delete b;
}

}  // extern "C"
