// Wick disk cache in the reference's format: `wicks.bin` (the raw slab, f64,
// pairs x * nref + w back to back) and `wicks.meta` (bincode 1 serialisation
// of WicksDiskMeta<f64>, version 2; src/nonorthogonalwicks/storage.rs:19,
// 163-255, types.rs:20-160).  bincode 1's default encoding: integers and
// usize as fixed-width little-endian u32/u64, f64 as 8 bytes, fixed arrays
// and structs as their fields in order, Vec as a u64 length then elements.
// Offsets in the file are absolute slab indices; tensors that are not
// allocated carry usize::MAX (J slots, hcol0, II) and the NOCC-only RDM
// blocks stay 0 (layout.rs:21-189 without the nocc feature).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "nw_internal.h"

using namespace nw;
using namespace nwi;

namespace {

constexpr uint32_t WICKS_DISK_CACHE_VERSION = 2;
constexpr uint64_t NONE = ~0ull;

struct Writer {
    std::vector<char> b;
    void u32(uint32_t v) { put(&v, 4); }
    void u64(uint64_t v) { put(&v, 8); }
    void f64(double v) { put(&v, 8); }
    void put(const void *p, size_t n) { b.insert(b.end(), (const char *)p, (const char *)p + n); }
};

struct Reader {
    const std::vector<char> &b;
    size_t i = 0;
    explicit Reader(const std::vector<char> &buf) : b(buf) {}
    void get(void *p, size_t n)
    {
        if (i + n > b.size()) throw NwError(NW_E_ARG, "wicks.meta is truncated");
        std::memcpy(p, b.data() + i, n);
        i += n;
    }
    uint32_t u32() { uint32_t v; get(&v, 4); return v; }
    uint64_t u64() { uint64_t v; get(&v, 8); return v; }
    double f64() { double v; get(&v, 8); return v; }
};

// SameSpinOffset field order: x[2], y[2], xrdm[2], yrdm[2], fh[2][2], ff[2][2], v[2][2][2], hcol0, j[10].
void put_same_off(Writer &w, const long long *o, long long base)
{
    auto abs = [&](long long rel) -> uint64_t { return rel < 0 ? NONE : (uint64_t)(base + rel); };
    for (int k = 0; k < 2; ++k) w.u64(abs(o[O_X + k]));
    for (int k = 0; k < 2; ++k) w.u64(abs(o[O_Y + k]));
    for (int k = 0; k < 4; ++k) w.u64(0);  // xrdm, yrdm (NOCC only)
    for (int k = 0; k < 4; ++k) w.u64(abs(o[O_FH + k]));
    for (int k = 0; k < 4; ++k) w.u64(abs(o[O_FF + k]));
    for (int k = 0; k < 8; ++k) w.u64(abs(o[O_V + k]));
    w.u64(abs(o[O_HCOL0]));
    for (int k = 0; k < 10; ++k) w.u64(abs(o[O_J + k]));
}

void get_same_off(Reader &r, long long *o, long long base)
{
    auto rel = [&](uint64_t a) -> long long { return a == NONE ? -1 : (long long)a - base; };
    for (int k = 0; k < 2; ++k) o[O_X + k] = rel(r.u64());
    for (int k = 0; k < 2; ++k) o[O_Y + k] = rel(r.u64());
    for (int k = 0; k < 4; ++k) (void)r.u64();
    for (int k = 0; k < 4; ++k) o[O_FH + k] = rel(r.u64());
    for (int k = 0; k < 4; ++k) o[O_FF + k] = rel(r.u64());
    for (int k = 0; k < 8; ++k) o[O_V + k] = rel(r.u64());
    o[O_HCOL0] = rel(r.u64());
    for (int k = 0; k < 10; ++k) o[O_J + k] = rel(r.u64());
}

// SameSpinMeta field order: tilde_s_prod, phase, m, nmo, nocc, f0f[2], f0h[2], v0[3].
void put_same_meta(Writer &w, const double *m, int nmo)
{
    w.f64(m[M_TSP]);
    w.f64(m[M_PH]);
    w.u64((uint64_t)m[M_M]);
    w.u64((uint64_t)nmo);
    w.u64((uint64_t)m[M_NOCC]);
    for (int k = 0; k < 2; ++k) w.f64(m[M_F0F + k]);
    for (int k = 0; k < 2; ++k) w.f64(m[M_F0H + k]);
    for (int k = 0; k < 3; ++k) w.f64(m[M_V0 + k]);
}

void get_same_meta(Reader &r, double *m, uint64_t &nmo)
{
    m[M_TSP] = r.f64();
    m[M_PH] = r.f64();
    m[M_M] = (double)r.u64();
    nmo = r.u64();
    m[M_NOCC] = (double)r.u64();
    for (int k = 0; k < 2; ++k) m[M_F0F + k] = r.f64();
    for (int k = 0; k < 2; ++k) m[M_F0H + k] = r.f64();
    for (int k = 0; k < 3; ++k) m[M_V0 + k] = r.f64();
}

}  // namespace

extern "C" {

int nw_save_wicks(nw_ctx *c, const char *dir)
{
    return guarded(c, [&]() -> int {
        if (!dir) return fail(c, NW_E_ARG, "null directory");
        if (!c->built) return fail(c, NW_E_STATE, "nw_build first");
        const int nref = c->nref, nm = c->nmo;
        for (auto &pi : c->pinfo)
            if (!pi.built) return fail(c, NW_E_STATE, "the disk cache holds every ordered pair: build without a mask");
        chol_complete(c);  // every rank-4 block of on-demand pairs
        // slab in pair order x * nref + w, in the reference layout; offsets absolute
        std::vector<long long> base(nref * nref), refoff((size_t)nref * nref * O_LEN);
        long long total = 0;
        for (int p = 0; p < nref * nref; ++p) {
            const PairInfo &pi = c->pinfo[p];
            base[p] = total;
            total += pair_offsets(pi.ma, pi.mb, nm, &refoff[(size_t)p * O_LEN]);
        }
        Writer w;
        w.u32(WICKS_DISK_CACHE_VERSION);
        w.u64((uint64_t)nref);
        w.u64((uint64_t)total);
        w.u64((uint64_t)nref * nref);
        for (int p = 0; p < nref * nref; ++p) {
            const long long *off = &refoff[(size_t)p * O_LEN];
            put_same_off(w, off + O_AA, base[p]);
            put_same_off(w, off + O_BB, base[p]);
            for (int k = 0; k < 8; ++k) w.u64((uint64_t)(base[p] + off[O_VAB + k]));
            for (int k = 0; k < 8; ++k) w.u64((uint64_t)(base[p] + off[O_VBA + k]));
            for (int k = 0; k < 16; ++k) w.u64(off[O_II + k] < 0 ? NONE : (uint64_t)(base[p] + off[O_II + k]));
        }
        w.u64((uint64_t)nref * nref);
        for (int p = 0; p < nref * nref; ++p) {
            const PairInfo &pi = c->pinfo[p];
            put_same_meta(w, pi.meta + M_AA, nm);
            put_same_meta(w, pi.meta + M_BB, nm);
            w.u64((uint64_t)nm);  // DiffSpinMeta: nmo, vab0[2][2], vba0[2][2]
            for (int k = 0; k < 4; ++k) w.f64(pi.meta[M_VAB0 + k]);
            for (int k = 0; k < 4; ++k) w.f64(pi.meta[M_VBA0 + k]);
        }
        const std::string d(dir);
        {
            std::ofstream f(d + "/wicks.meta", std::ios::binary | std::ios::trunc);
            if (!f) return fail(c, NW_E_ARG, "cannot write " + d + "/wicks.meta");
            f.write(w.b.data(), (std::streamsize)w.b.size());
            if (!f) return fail(c, NW_E_ARG, "write failed: " + d + "/wicks.meta");
        }
        std::ofstream f(d + "/wicks.bin", std::ios::binary | std::ios::trunc);
        if (!f) return fail(c, NW_E_ARG, "cannot write " + d + "/wicks.bin");
        std::vector<double> buf, out;
        for (int p = 0; p < nref * nref; ++p) {
            const PairInfo &pi = c->pinfo[p];
            buf.resize((size_t)pi.off[O_TOTAL]);
            HIPCHK(hipMemcpy(buf.data(), c->d_slab + pi.base, buf.size() * 8, hipMemcpyDeviceToHost));
            const size_t len = (size_t)refoff[(size_t)p * O_LEN + O_TOTAL];
            out.resize(len);
            export_pair(pi, nm, buf.data(), out.data());
            f.write((const char *)out.data(), (std::streamsize)(len * 8));
            if (!f) return fail(c, NW_E_ARG, "write failed: " + d + "/wicks.bin");
        }
        return NW_OK;
    });
}

int nw_load_wicks(nw_ctx *c, const char *dir, double tol)
{
    return guarded(c, [&]() -> int {
        if (!dir) return fail(c, NW_E_ARG, "null directory");
        if (c->nref == 0) return fail(c, NW_E_STATE, "nw_set_refs first");
        const int nref = c->nref, nm = c->nmo;
        const std::string d(dir);
        std::ifstream fm(d + "/wicks.meta", std::ios::binary);
        if (!fm) return fail(c, NW_E_ARG, "cannot read " + d + "/wicks.meta");
        std::vector<char> mb((std::istreambuf_iterator<char>(fm)), std::istreambuf_iterator<char>());
        Reader r(mb);
        const uint32_t version = r.u32();
        if (version != WICKS_DISK_CACHE_VERSION)
            return fail(c, NW_E_ARG, "Wick disk-cache version " + std::to_string(version) +
                                         " is incompatible with expected version 2");
        const uint64_t fnref = r.u64(), slab_len = r.u64();
        if (fnref != (uint64_t)nref) return fail(c, NW_E_SHAPE, "cache holds a different number of references");
        if (r.u64() != fnref * fnref) return fail(c, NW_E_ARG, "cache offset table has the wrong length");
        std::vector<PairInfo> pinfo((size_t)nref * nref, PairInfo{});
        for (int p = 0; p < nref * nref; ++p) {
            PairInfo &pi = pinfo[p];
            // the block starts at the alpha X^(0) offset
            Reader peek = r;
            const long long base = (long long)peek.u64();
            get_same_off(r, pi.off + O_AA, base);
            get_same_off(r, pi.off + O_BB, base);
            for (int k = 0; k < 8; ++k) pi.off[O_VAB + k] = (long long)r.u64() - base;
            for (int k = 0; k < 8; ++k) pi.off[O_VBA + k] = (long long)r.u64() - base;
            for (int k = 0; k < 16; ++k) {
                const uint64_t a = r.u64();
                pi.off[O_II + k] = a == NONE ? -1 : (long long)a - base;
            }
            pi.base = base;
        }
        if (r.u64() != fnref * fnref) return fail(c, NW_E_ARG, "cache meta table has the wrong length");
        for (int p = 0; p < nref * nref; ++p) {
            PairInfo &pi = pinfo[p];
            uint64_t nmo_a, nmo_b;
            get_same_meta(r, pi.meta + M_AA, nmo_a);
            get_same_meta(r, pi.meta + M_BB, nmo_b);
            const uint64_t nmo_ab = r.u64();
            for (int k = 0; k < 4; ++k) pi.meta[M_VAB0 + k] = r.f64();
            for (int k = 0; k < 4; ++k) pi.meta[M_VBA0 + k] = r.f64();
            if (nmo_a != (uint64_t)nm || nmo_b != (uint64_t)nm || nmo_ab != (uint64_t)nm)
                return fail(c, NW_E_SHAPE, "cache nmo differs from the references'");
            if ((int)pi.meta[M_AA + M_NOCC] != c->nocc[0] || (int)pi.meta[M_BB + M_NOCC] != c->nocc[1])
                return fail(c, NW_E_SHAPE, "cache occupation differs from the references'");
            pi.ma = (int)pi.meta[M_AA + M_M];
            pi.mb = (int)pi.meta[M_BB + M_M];
            // the stored layout must be the one assign_offsets gives for (ma, mb)
            long long want[O_LEN];
            const long long len = pair_offsets(pi.ma, pi.mb, nm, want);
            for (int k = 0; k < O_TOTAL; ++k)
                if (want[k] != pi.off[k]) return fail(c, NW_E_ARG, "cache slab layout differs from assign_offsets");
            pi.off[O_TOTAL] = len;
            if (pi.base < 0 || (uint64_t)(pi.base + len) > slab_len) return fail(c, NW_E_ARG, "cache offsets out of range");
            pi.built = 1;
            pi.bidx = p;
            pi.smin = 0.0;
            pi.fp32 = 0;
        }
        std::ifstream fb(d + "/wicks.bin", std::ios::binary | std::ios::ate);
        if (!fb) return fail(c, NW_E_ARG, "cannot read " + d + "/wicks.bin");
        if ((uint64_t)fb.tellg() != slab_len * 8) return fail(c, NW_E_ARG, "wicks.bin size differs from slab_len");
        fb.seekg(0);
        c->built = false;
        c->build_gen++;
        c->fock_gen++;
        chol_forget(c);  // a loaded slab is complete
        if (slab_len > c->slab_cap) {
            if (c->d_slab) HIPCHK(hipFree(c->d_slab));
            c->d_slab = nullptr;
            c->slab_cap = 0;
            c->d_slab = dalloc<double>(slab_len);
            c->slab_cap = slab_len;
        }
        std::vector<double> buf(std::min<uint64_t>(slab_len, 1ull << 24));
        for (uint64_t o = 0; o < slab_len; o += buf.size()) {
            const uint64_t n = std::min<uint64_t>(buf.size(), slab_len - o);
            fb.read((char *)buf.data(), (std::streamsize)(n * 8));
            if (!fb) return fail(c, NW_E_ARG, "read failed: " + d + "/wicks.bin");
            HIPCHK(hipMemcpy(c->d_slab + o, buf.data(), n * 8, hipMemcpyHostToDevice));
        }
        c->pinfo = pinfo;
        if (c->d_pinfo) HIPCHK(hipFree(c->d_pinfo));
        c->d_pinfo = nullptr;
        c->pinfo_cap = 0;
        c->d_pinfo = (PairInfo *)dalloc<char>(sizeof(PairInfo) * pinfo.size());
        c->pinfo_cap = pinfo.size();
        HIPCHK(hipMemcpy(c->d_pinfo, pinfo.data(), sizeof(PairInfo) * pinfo.size(), hipMemcpyHostToDevice));
        c->tol = tol;
        c->have_codensity = false;  // nw_update_fock needs the co-densities of a build
        c->built = true;
        return NW_OK;
    });
}

}  // extern "C"
