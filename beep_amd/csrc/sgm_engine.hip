// sgm_engine.hip -- device residency and the launch sequence of one
// enumeration (the host driver).
//
// Replaces count_subgraphs (subgraph_matching.cuh:1037-1279): the reference
// copies __constant__ tables, sizes a grid by occupancy, allocates and memsets
// per-block global scratch, launches the persistent kernel, adds the device
// counter into a managed counter and prints it.  Here the prepared graph is
// uploaded once (beep_prep_upload) and stays in HBM; beep_count launches
//   on one stream, per hub round:  arena clear -> hub_hash -> hub_encode -> hub_dfs
//   on side streams, per lane class: lds_batch_kernel<P>
// and reads back the counters.  Every buffer of the sequence is carved from
// one scratch region sized before anything is enqueued.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "engine.hpp"

namespace beep {

static_assert(sizeof(RootRec) == kRootRecBytes, "per-root record size");

hipError_t launch_lds_batches(int lg, int cus, hipStream_t st, const KGraph &g, const KPlan &pl,
                              const RootRec *recs, const uint32_t *bf, const uint32_t *bc,
                              uint32_t b0, uint32_t b1, unsigned long long *cursor, const KOut &out,
                              int *grid_out, bool dense = false);
hipError_t launch_hub_hash(int grid, hipStream_t st, const KGraph &g, const HubRec *hubs,
                           const uint32_t *hor, uint64_t g0, uint64_t g1, uint64_t *kv,
                           const unsigned long long *dl);
hipError_t launch_hub_encode(int grid, hipStream_t st, const KGraph &g, const HubRec *hubs,
                             const uint32_t *hor, const uint64_t *hrange, const uint64_t *rows, uint64_t nrows,
                             const uint64_t *kv, uint64_t *E, int upper, const KOut &out);
hipError_t launch_hub_dfs(int grid, hipStream_t st, const KGraph &g, const KPlan &pl, const HubRec *hubs,
                          const uint32_t *hoi, uint64_t i0, uint64_t i1, unsigned long long *cursor,
                          const uint64_t *E, uint64_t *stk, uint32_t wmax, const KOut &out,
                          const HubQueue &hq, int mode);
int hub_dfs_occupancy(int nlev, uint32_t wmax, bool pair);

template <class T>
static T *dev_alloc(size_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    return (T *)p;
}

template <class T>
static T *dev_upload(const std::vector<T> &v)
{
    T *p = dev_alloc<T>(v.size());
    if (!v.empty()) HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

// ------------------------------------------------------------ residency

DevGraph::~DevGraph()
{
    if (device >= 0) (void)hipSetDevice(device);
    if (ocol) (void)hipFree(ocol);
    if (vinfo) (void)hipFree(vinfo);
    if (htab) (void)hipFree(htab);
}

Variant::~Variant() { delete dev; }

void DevRoots::release()
{
    for (void *p : {(void *)recs, (void *)batch_first, (void *)batch_count, (void *)hubs, (void *)hub_of_row,
                    (void *)hub_of_item, (void *)hub_range, (void *)enc_rows})
        if (p) (void)hipFree(p);
    recs = nullptr;
    batch_first = batch_count = hub_of_row = hub_of_item = nullptr;
    hubs = nullptr;
    hub_range = enc_rows = nullptr;
}

void Lane::create(int dev, int ns)
{
    if (main) return;
    device = dev;
    nside = std::max(1, std::min(ns, kMaxSide));
    HIP_CHECK(hipStreamCreateWithFlags(&main, hipStreamNonBlocking));
    for (int i = 0; i < nside; i++) HIP_CHECK(hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking));
    for (auto &e : ev) HIP_CHECK(hipEventCreate(&e));
    for (int i = 0; i < nside; i++) HIP_CHECK(hipEventCreate(&side_done[i]));
    for (auto &e : cls_done) HIP_CHECK(hipEventCreate(&e));
    for (auto &e : cls_start) HIP_CHECK(hipEventCreate(&e));
}

Lane::~Lane()
{
    if (device >= 0) (void)hipSetDevice(device);
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : side_done)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : cls_start)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : cls_done)
        if (e) (void)hipEventDestroy(e);
    for (auto &s : side)
        if (s) (void)hipStreamDestroy(s);
    if (main) (void)hipStreamDestroy(main);
}

DevPrep::~DevPrep()
{
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = device;
    if (device >= 0) (void)hipSetDevice(device);
    roots.release();
    if (region) (void)hipFree(region);
    if (device >= 0) (void)hipSetDevice(cur);
}

Prep::~Prep() { delete dev; }

int device_cus(int device)
{
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, device));
    return prop.multiProcessorCount;
}

uint64_t wall_clock_khz(int device)
{
    int khz = 0;
    HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
    return khz > 0 ? (uint64_t)khz : 100000ull;
}

static void upload_variant(Variant &V, int device)
{
    std::lock_guard<std::mutex> lk(V.dev_mu);
    if (V.dev) {
        if (V.dev->device != device) throw Error(BEEP_EINVAL, "the prepared graph is resident on another device");
        return;
    }
    if (V.m >= (1ull << kRangeShift) || V.max_degree >= (1u << (64 - kRangeShift)))
        throw Error(BEEP_ERANGE, "graph too large for the staged row ranges (2^40 entries, degree 2^24)");
    auto *dg = new DevGraph();
    try {
        dg->device = device;
        dg->ocol = dev_upload(V.ocol);
        std::vector<VInfo> vi(V.n);
        parallel_for(V.n, [&](size_t b, size_t e2) {
            for (size_t v = b; v < e2; v++) {
                vi[v].hbeg = V.rowptr[v] + (V.asc ? V.nlow[v] : 0u);
                vi[v].hlen = V.hlen[v];
                vi[v].deg = (uint32_t)(V.rowptr[v + 1] - V.rowptr[v]);
            }
        });
        dg->vinfo = dev_upload(vi);
        {
            // lookup tables: vertex v's larger-key range [hbeg, hbeg + hlen)
            // in the 2 hlen slots at 2 hbeg, linear probing (KGraph::htab)
            std::vector<uint32_t> ht(2 * V.m, 0u);
            parallel_for(V.n, [&](size_t b, size_t e2) {
                for (size_t v = b; v < e2; v++) {
                    const uint64_t hb = vi[v].hbeg, hl = vi[v].hlen;
                    const uint32_t size = (uint32_t)(2 * hl);
                    uint32_t *t = ht.data() + 2 * hb;
                    for (uint64_t x = 0; x < hl; x++) {
                        const uint32_t w = V.ocol[hb + x];
                        uint32_t s = lookup_slot(w, size);
                        while (t[s]) s = s + 1 == size ? 0 : s + 1;
                        t[s] = w + 1;
                    }
                }
            }, 1024);
            dg->htab = dev_upload(ht);
        }
        progress("variant uploaded (%llu entries)", (unsigned long long)V.m);
    } catch (...) {
        delete dg;
        throw;
    }
    V.dev = dg;
}

// What hub_encode streams for every hub row j (vertex v = A[j]): v's
// larger-key range.  A row below the first kept row (rlo) sets bits only in
// kept columns, i.e. for entries not before A[rlo] in the row order (every
// row is sorted by the same key), so its range is cut to that suffix; with
// no kept row at all it is empty.  Plans that keep only the bit in the row of
// the smaller index (upper) need nothing from such a row: it is not scanned.
static std::vector<uint64_t> hub_scan_ranges(const Variant &V, const std::vector<HubRec> &hubs, bool upper)
{
    uint64_t rows = 0;
    for (const auto &h : hubs) rows += h.d;
    std::vector<uint64_t> out(rows);
    auto before = [&](uint32_t a, uint32_t b) {  // a precedes b in a row
        const uint32_t da = V.deg(a), db = V.deg(b);
        if (da != db) return V.asc ? da < db : da > db;
        return a < b;
    };
    parallel_for(hubs.size(), [&](size_t b, size_t e) {
        for (size_t hi = b; hi < e; hi++) {
            const HubRec &h = hubs[hi];
            const uint32_t *A = V.ocol.data() + h.rbeg;
            for (uint32_t j = 0; j < h.d; j++) {
                const uint32_t v = A[j];
                uint64_t beg = V.rowptr[v] + (V.asc ? V.nlow[v] : 0u);
                uint64_t len = V.hlen[v];
                if (j < h.rlo) {
                    if (h.rlo >= h.d || upper) {
                        len = 0;
                    } else {
                        const uint32_t p = A[h.rlo];
                        const uint32_t *rg = V.ocol.data() + beg;
                        const uint64_t cut = std::partition_point(rg, rg + len, [&](uint32_t x) { return before(x, p); }) - rg;
                        beg += cut;
                        len -= cut;
                    }
                }
                out[h.joff + j] = beg | (len << kRangeShift);
            }
        }
    }, 1);
    return out;
}

void upload(Prep &pp, int device, int p_skip_split, uint64_t arena_budget)
{
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw Error(BEEP_EDEVICE, "no such HIP device");
    HIP_CHECK(hipSetDevice(device));
    upload_variant(*pp.var, device);
    if (pp.dev && pp.dev->device == device && pp.dev->roots.skip_split == p_skip_split &&
        (arena_budget == 0 || pp.dev->roots.arena_budget == arena_budget))
        return;
    delete pp.dev;
    pp.dev = nullptr;
    const Variant &V = *pp.var;
    // built aside and published at the end, so a failure leaves no half-built state
    std::unique_ptr<DevPrep> dp(new DevPrep());
    DevPrep *d = dp.get();
    d->device = device;
    d->cus = device_cus(device);
    DevRoots &R = d->roots;
    R.skip_split = p_skip_split;

    // LDS batches: per-root records with batch-relative offsets.
    {
        std::vector<RootRec> recs;
        std::vector<uint32_t> bf, bc;
        for (const auto &b : pp.batches) {
            bf.push_back((uint32_t)recs.size());
            bc.push_back(b.count);
            uint32_t doff = 0, eoff = 0, hoff = 0;
            for (uint32_t i = 0; i < b.count; i++) {
                const uint32_t wi = b.first + i;
                const uint32_t u = pp.wu[wi];
                RootRec r{};
                r.rbeg = V.rowptr[u] + pp.woff[wi];
                r.u = u;
                r.d = pp.wlen[wi];
                r.split = pp.reduced ? 0u : V.nlow[u];
                r.doff = doff;
                r.eoff = eoff;
                r.hoff = hoff;
                doff += r.d;
                eoff += r.d * ((r.d + 63) / 64);
                hoff += hash_slots_for(r.d);
                recs.push_back(r);
            }
        }
        R.recs = dev_upload(recs);
        R.batch_first = dev_upload(bf);
        R.batch_count = dev_upload(bc);
        R.nbatches = (uint32_t)bf.size();
    }

    // Hubs: records, row -> hub map and rounds that fit the arena.
    {
        std::vector<HubRec> hubs;
        std::vector<uint32_t> hor, hoi;
        uint64_t joff = 0, ioff = 0;
        for (uint32_t h = 0; h < pp.n_hub; h++) {
            const uint32_t u = pp.wu[h];
            HubRec r{};
            r.rbeg = V.rowptr[u] + pp.woff[h];
            r.u = u;
            r.d = pp.wlen[h];
            r.split = pp.reduced ? 0u : V.nlow[u];
            r.joff = joff;
            r.ioff = ioff;
            // x_1 < split items are skipped (sgm_kernelsLD.cuh:60): they are not listed.
            r.jbeg = (p_skip_split && r.split < r.d) ? r.split : (p_skip_split ? r.d : 0u);
            r.rlo = pp.rows_above_split ? std::min(r.split, r.d) : 0u;
            // every pair the bitmap keeps has an endpoint in a kept row, and the
            // scan ranges only hold entries not before A[rlo]: the hash needs the
            // kept rows only
            r.hslots = hash_slots_for(r.d - r.rlo);
            joff += r.d;
            ioff += r.d - r.jbeg;
            for (uint32_t j = 0; j < r.d; j++) hor.push_back(h);
            for (uint32_t j = r.jbeg; j < r.d; j++) hoi.push_back(h);
            const uint64_t W = (r.d + 63) / 64;
            R.wmax = std::max<uint32_t>(R.wmax, (uint32_t)W);
            // donation: at most one slot of W words and ceil(d/16)+1 records per item
            R.don_words += (uint64_t)(r.d - r.jbeg) * W;
            R.don_records += (uint64_t)(r.d - r.jbeg) * ((r.d + 15) / 16 + 1);
            hubs.push_back(r);
        }
        if (arena_budget == 0) {
            size_t freeb = 0, totalb = 0;
            HIP_CHECK(hipMemGetInfo(&freeb, &totalb));
            arena_budget = std::min<uint64_t>((uint64_t)(freeb * 0.35), 48ull << 30);
        }
        R.arena_budget = arena_budget;
        uint64_t ew = 0, hsl = 0;
        uint32_t h0 = 0;
        for (uint32_t h = 0; h < (uint32_t)hubs.size(); h++) {
            // plans that never read a bitmap row need neither bitmap nor hash
            uint64_t w = pp.uses_bitmap ? (uint64_t)(hubs[h].d - hubs[h].rlo) * ((hubs[h].d + 63) / 64) : 0;
            uint64_t s = pp.uses_bitmap ? hubs[h].hslots : 0;
            if (h > h0 && (ew + w) * 8 + (hsl + s) * 8 > arena_budget) {
                R.rounds.push_back({h0, h, hubs[h0].joff, hubs[h].joff, hubs[h0].ioff, hubs[h].ioff, ew, hsl});
                h0 = h;
                ew = hsl = 0;
            }
            hubs[h].eoff = ew;
            hubs[h].hoff = hsl;
            ew += w;
            hsl += s;
            if ((ew * 8 + hsl * 8) > arena_budget)
                throw Error(BEEP_ENOMEM, "hub root bitmap exceeds the HBM arena budget");
        }
        if (!hubs.empty())
            R.rounds.push_back({h0, (uint32_t)hubs.size(), hubs[h0].joff, joff, hubs[h0].ioff, ioff, ew, hsl});
        for (const auto &r : R.rounds) {
            R.arena_words = std::max(R.arena_words, r.ewords);
            R.arena_slots = std::max(R.arena_slots, r.hslots);
        }
        if (pp.uses_bitmap) {
            const std::vector<uint64_t> ranges = hub_scan_ranges(V, hubs, pp.upper_rows != 0);
            // the bitmap build streams only rows with a non-empty range (a
            // third of the diamond's hub rows on chunglu1m have none)
            std::vector<uint64_t> rows;
            for (auto &r : R.rounds) {
                r.er0 = rows.size();
                for (uint64_t g = r.g0; g < r.g1; g++)
                    if (ranges[g] >> kRangeShift) rows.push_back(g);
                r.er1 = rows.size();
            }
            if (!rows.empty()) R.enc_rows = dev_upload(rows);
            R.hub_range = dev_upload(ranges);
        }
        R.hubs = dev_upload(hubs);
        R.hub_of_row = dev_upload(hor);
        R.hub_of_item = dev_upload(hoi);
        R.nhubs = (uint32_t)hubs.size();
    }
    pp.dev = dp.release();
}

// ------------------------------------------------------------ scratch

static uint64_t align256(uint64_t b) { return (b + 255) & ~255ull; }

uint64_t ScratchNeeds::bytes() const
{
    return align256(kCtrTotal * 8) + align256(arena_words * 8) + align256(hash_slots * 8) +
           align256(stk_words * 8) + align256(queue_cap * sizeof(HubChunk)) + align256(slot_cap * 8) +
           align256(dbuf_words * 4) + align256(prof_words * 8);
}

void carve(Scratch &s, char *base, const ScratchNeeds &n)
{
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        char *p = base + off;
        off += align256(bytes);
        return p;
    };
    s.ctr = (unsigned long long *)take(kCtrTotal * 8);
    s.arena = n.arena_words ? (uint64_t *)take(n.arena_words * 8) : nullptr;
    s.kv = n.hash_slots ? (uint64_t *)take(n.hash_slots * 8) : nullptr;
    s.stk = n.stk_words ? (uint64_t *)take(n.stk_words * 8) : nullptr;
    s.queue = n.queue_cap ? (HubChunk *)take(n.queue_cap * sizeof(HubChunk)) : nullptr;
    s.slots = n.slot_cap ? (uint64_t *)take(n.slot_cap * 8) : nullptr;
    s.dbuf = n.dbuf_words ? (uint32_t *)take(n.dbuf_words * 4) : nullptr;
    s.prof = n.prof_words ? (unsigned long long *)take(n.prof_words * 8) : nullptr;
    s.cap = n;
}

// Count-only plans whose last level does not intersect with x_{L-2}'s row sum
// that level in closed form (KPlan::tail).
static bool tail_summed(const Plan &p, int emit)
{
    return emit == BEEP_EMIT_NONE && p.L >= 4 && !((p.qmask[p.L - 1] >> (p.L - 2)) & 1u);
}

// Count-only plans whose last three levels pair up (KPlan::pair).
static bool pair_summed(const Plan &p, int emit)
{
    if (emit != BEEP_EMIT_NONE || p.L < 5) return false;
    const int a = p.L - 3, b = p.L - 2, c = p.L - 1;
    const uint32_t below_a = (1u << a) - 1u;
    if (p.qmask[c] != (uint32_t)(p.qmask[a] | (1u << b))) return false;
    if (!((p.qmask[b] >> a) & 1u) || (p.qmask[b] & ~(1u << a) & ~below_a)) return false;
    if (p.smask[c] != (uint32_t)(p.smask[a] | (1u << a))) return false;
    return (p.smask[b] & ~below_a) == 0;
}

// Hub items donate level-2 subtrees; not when everything below x_1 is one
// summed level (L = 4 with a summed tail): an item then costs one pass over
// its d/64 words whatever |C_2| is.  Items of 5-vertex plans with paired
// levels donate ranges of the paired pass's words instead.
static bool hub_donates(const Plan &p, const beep_opts &o, const DevRoots &R)
{
    return o.donate && R.nhubs && p.L >= 4 && !(p.L == 4 && tail_summed(p, o.emit));
}

static int hub_grid_of(const DevRoots &R, int nlev, bool pair, int cus)
{
    return hub_dfs_occupancy(nlev, std::max<uint32_t>(R.wmax, 1), pair) * cus;
}

ScratchNeeds scratch_needs(const Prep &pp, const Plan &p, const beep_opts &o, int cus)
{
    const DevRoots &R = pp.dev->roots;
    ScratchNeeds n;
    if (R.nhubs && pp.uses_bitmap) {
        n.arena_words = R.arena_words;
        n.hash_slots = R.arena_slots;
    }
    const int nlev = p.L > 3 ? p.L - 3 : 1;
    if (R.nhubs) {
        // per wave: nlev stack rows + the row of x_1 (used when they do not fit LDS)
        n.stk_words = (uint64_t)hub_grid_of(R, nlev, pair_summed(p, o.emit), cus) * (kBlock / 64) *
                      hub_wave_words(nlev, std::max<uint32_t>(R.wmax, 1));
    }
    if (hub_donates(p, o, R)) {
        n.queue_cap = std::max<uint64_t>(1, std::min<uint64_t>(1ull << 22, R.don_records));
        n.slot_cap = std::max<uint64_t>(1, std::min<uint64_t>(1ull << 26, R.don_words));
    }
    if (o.emit == BEEP_EMIT_BUFFER && o.match_cap) n.dbuf_words = o.match_cap * (uint64_t)p.L;
    if (o.profile) n.prof_words = (uint64_t)kProfPhases * kProfSlots;
    return n;
}

static KPlan make_kplan(const Plan &p, int emit, int build_mode)
{
    KPlan k{};
    k.k = p.k;
    k.L = p.L;
    k.skip_split = p.skip_split;
    k.emit = emit;
    k.bin_ratio = build_mode == BEEP_BUILD_SCAN ? ~0u : build_mode == BEEP_BUILD_SEARCH ? 0u : kLookupRatio;
    k.tail = tail_summed(p, emit) ? 1 : 0;
    k.pair = pair_summed(p, emit) ? 1 : 0;
    k.upper = p.upper_rows;
    for (int l = 0; l < kMaxLevels; l++) {
        k.qmask[l] = p.qmask[l];
        k.smask[l] = p.smask[l];
    }
    return k;
}

LaunchCfg make_launch(const Prep &pp, const Plan &p, const beep_opts &o, const Scratch &s, int cus,
                      double lds_deadline_ms)
{
    const Variant &V = *pp.var;
    if (p.min_qdeg != V.min_qdeg || p.max_qdeg != pp.max_qdeg || p.asc != V.asc ||
        p.above_split != pp.reduced || p.rows_above_split != pp.rows_above_split ||
        p.uses_bitmap != pp.uses_bitmap || p.upper_rows != pp.upper_rows ||
        p.skip_split != pp.dev->roots.skip_split)
        throw Error(BEEP_EINVAL, "plan does not match the prepared graph");
    const DevRoots &R = pp.dev->roots;
    LaunchCfg c{};
    const bool buf = o.emit == BEEP_EMIT_BUFFER && o.match_cap && s.dbuf;
    c.out = KOut{s.ctr + kCtrAcc, buf ? s.dbuf : nullptr, buf ? o.match_cap : 0, s.ctr + kCtrDeadline,
                 s.ctr + kCtrStop, s.prof};
    c.kp = make_kplan(p, o.emit != BEEP_EMIT_NONE ? 1 : 0, o.build_mode);
    c.kg = KGraph{V.dev->ocol, V.dev->vinfo, V.dev->htab, V.asc};
    c.nlev = p.L > 3 ? p.L - 3 : 1;
    c.hub_grid = hub_grid_of(R, c.nlev, c.kp.pair != 0, cus);
    // donation ring (hub items with |C_2| > donate_min)
    c.donate = hub_donates(p, o, R) && s.queue && s.slots;
    if (c.donate) {
        c.hq.q = s.queue;
        c.hq.cap = s.cap.queue_cap;
        c.hq.slots = s.slots;
        c.hq.slot_cap = s.cap.slot_cap;
        c.hq.tail = s.ctr + kCtrQTail;
        c.hq.head = s.ctr + kCtrHead;
        c.hq.slot_tail = s.ctr + kCtrSlot;
        c.hq.donated = s.ctr + kCtrDonated;
        c.hq.split_min = o.donate_min ? o.donate_min : 64;
        c.hq.chunk = c.hq.split_min;
    }
    c.lmin = 0;
    if (o.lanes > 0)
        while ((1 << c.lmin) < o.lanes && c.lmin < 6) c.lmin++;
    auto ticks = [&](double ms) -> uint64_t {
        if (!(ms > 0)) return 0;
        const uint64_t t = (uint64_t)(ms * (double)wall_clock_khz(pp.dev->device));
        return t ? t : 1;
    };
    c.deadline_ticks = ticks(o.deadline_ms);
    c.deadline_ticks_lds = lds_deadline_ms < 0 ? c.deadline_ticks : ticks(lds_deadline_ms);
    return c;
}

// ------------------------------------------------------------ launch sequence

// Hub rounds on one stream: clear, hash, bitmap rows, DFS + donation drain.
void enqueue_hubs(const Prep &pp, const LaunchCfg &c, const Scratch &s, hipStream_t st)
{
    if (c.kp.L < 3) return;  // k = 3 with an unmaterialised last level: no level to count
    const DevRoots &R = pp.dev->roots;
    const uint32_t wmax = std::max<uint32_t>(R.wmax, 1);
    const int cus = pp.dev->cus;
    for (size_t ri = 0; ri < R.rounds.size(); ri++) {
        const HubRound &r = R.rounds[ri];
        const uint64_t rows = r.g1 - r.g0;
        if (pp.uses_bitmap) {
            HIP_CHECK(launch_fill_zero(s.arena, r.ewords * 8, st, c.out.dl));
            HIP_CHECK(launch_fill_zero(s.kv, r.hslots * 8, st, c.out.dl));
            int g1 = (int)std::min<uint64_t>((rows + kBlock - 1) / kBlock, (uint64_t)cus * 16);
            HIP_CHECK(launch_hub_hash(std::max(g1, 1), st, c.kg, R.hubs, R.hub_of_row, r.g0, r.g1, s.kv, c.out.dl));
            const uint64_t nenc = r.er1 - r.er0;
            int g2 = (int)std::min<uint64_t>((nenc + 255) / 256, (uint64_t)cus * 16);
            if (nenc)
                HIP_CHECK(launch_hub_encode(std::max(g2, 1), st, c.kg, R.hubs, R.hub_of_row, R.hub_range,
                                            R.enc_rows + r.er0, nenc, s.kv, s.arena, c.kp.upper, c.out));
        }
        unsigned long long *cur = s.ctr + kCtrHub + (uint64_t)(ri % 64) * kCurWords;
        if (ri >= 64) HIP_CHECK(launch_fill_zero(cur, kCurWords * 8, st));
        const uint64_t items = r.i1 - r.i0;
        if (items) {
            if (c.donate && ri > 0) {  // queue tail, slot tail and the drain cursor start over
                HIP_CHECK(launch_fill_zero(s.ctr + kCtrQTail, 2 * 8, st));
                HIP_CHECK(launch_fill_zero(s.ctr + kCtrHead, kCurWords * 8, st));
            }
            int g3 = (int)std::min<uint64_t>((items + 3) / 4, (uint64_t)c.hub_grid);
            HIP_CHECK(launch_hub_dfs(std::max(g3, 1), st, c.kg, c.kp, R.hubs, R.hub_of_item, r.i0, r.i1, cur,
                                     s.arena, s.stk, wmax, c.out, c.hq, 0));
            if (c.donate)
                HIP_CHECK(launch_hub_dfs(c.hub_grid, st, c.kg, c.kp, R.hubs, R.hub_of_item, r.i0, r.i1, cur,
                                         s.arena, s.stk, wmax, c.out, c.hq, 1));
        }
    }
}

// LDS lane classes, one launch each, round-robin over the given streams.
int enqueue_lds_classes(const Prep &pp, const LaunchCfg &c, const Scratch &s, hipStream_t *streams, int nst,
                        hipEvent_t *cls_done, uint8_t *cls_lg)
{
    if (c.kp.L < 3) return 0;
    const DevRoots &R = pp.dev->roots;
    uint32_t b = 0;
    int cls = 0;
    while (b < R.nbatches) {
        int lg = std::max<int>(pp.batch_lanes_log2[b], c.lmin);
        uint32_t e = b + 1;
        while (e < R.nbatches && std::max<int>(pp.batch_lanes_log2[e], c.lmin) == lg) e++;
        if (cls >= 32) throw Error(BEEP_ERANGE, "too many lane classes");
        unsigned long long *cur = s.ctr + kCtrLds + (uint64_t)cls * kCurWords;
        HIP_CHECK(launch_lds_batches(lg, pp.dev->cus, streams[cls % nst], c.kg, c.kp, R.recs, R.batch_first,
                                     R.batch_count, b, e, cur, c.out, nullptr));
        if (cls_done) {
            HIP_CHECK(hipEventRecord(cls_done[cls], streams[cls % nst]));
            cls_lg[cls] = (uint8_t)lg;
        }
        b = e;
        cls++;
    }
    return cls;
}

void class_times(const Lane &lane, double *t_by_lg)
{
    for (int i = 0; i < 8; i++) t_by_lg[i] = 0;
    for (int i = 0; i < lane.ncls; i++) {
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, lane.cls_start[i], lane.cls_done[i]));
        if (lane.cls_lg[i] < 8) t_by_lg[lane.cls_lg[i]] += t;
    }
}

double lds_phase_ms(const Lane &lane)
{
    float tmax = 0;
    for (int i = 0; i < lane.ncls; i++) {
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, lane.ev[0], lane.cls_done[i]));
        tmax = std::max(tmax, t);
    }
    return tmax;
}

// Lane classes of a query: [batch begin, end) and lane class, with the bitmap
// build units as the estimate of their work.
struct LdsClass {
    uint32_t b0, b1;
    int lg;
    uint64_t work;
};

static std::vector<LdsClass> lds_classes(const Prep &pp, const LaunchCfg &c)
{
    std::vector<LdsClass> out;
    const DevRoots &R = pp.dev->roots;
    uint32_t b = 0;
    while (b < R.nbatches) {
        const int lg = std::max<int>(pp.batch_lanes_log2[b], c.lmin);
        uint32_t e = b + 1;
        uint64_t work = 0;
        while (e < R.nbatches && std::max<int>(pp.batch_lanes_log2[e], c.lmin) == lg) e++;
        for (int g = 0; g < 8; g++)
            if (std::max(g, c.lmin) == lg) work += pp.cls_units[g] + 8 * pp.cls_roots[g];
        out.push_back({b, e, lg, work});
        b = e;
    }
    return out;
}

void enqueue_query(const Prep &pp, const LaunchCfg &c, const Scratch &s, Lane &lane, unsigned long long *host_ctr)
{
    hipStream_t st = lane.main;
    HIP_CHECK(launch_begin_sequence(s.ctr, kCtrTotal, kCtrDeadline, c.deadline_ticks, c.deadline_ticks_lds, st));
    if (s.prof) HIP_CHECK(launch_fill_zero(s.prof, s.cap.prof_words * 8, st));
    HIP_CHECK(hipEventRecord(lane.ev[0], st));
    // With hub roots, the heaviest LDS lane class starts on the main stream at
    // once (a wait on another stream's event costs tens of microseconds); the
    // hub rounds, the longest chain of dependent kernels, are queued next, on
    // side[0]; the other classes follow widest first (their batches take
    // longest, so they start early and the narrow batches fill in around
    // them), one on each further side stream.  Classes left over queue on main
    // behind the first one, which holds the CUs' LDS while it runs: they start
    // as it drains.  Without hubs, the widest class takes the main stream and
    // the others side[0], side[1], ...
    std::vector<LdsClass> cls = c.kp.L >= 3 ? lds_classes(pp, c) : std::vector<LdsClass>();
    if (cls.size() > (size_t)Lane::kMaxClasses) throw Error(BEEP_ERANGE, "too many lane classes");
    std::vector<int> order(cls.size());
    for (size_t i = 0; i < cls.size(); i++) order[i] = (int)i;
    const DevRoots &R = pp.dev->roots;
    const bool hubs = c.kp.L >= 3 && !R.rounds.empty();
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cls[a].work > cls[b].work; });
    if (!hubs)  // widest first, on main; the narrow classes fill in around it
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cls[a].lg > cls[b].lg; });
    else if (order.size() > 2)
        std::stable_sort(order.begin() + 1, order.end(), [&](int a, int b) { return cls[a].lg > cls[b].lg; });
    // stream k: 0 = main, 1 + i = side[i]
    const int nstreams = 1 + lane.nside;
    std::vector<bool> waited(nstreams, false);
    waited[0] = true;
    auto side_stream = [&](int k) -> hipStream_t {
        if (!waited[k]) {
            HIP_CHECK(hipStreamWaitEvent(lane.side[k - 1], lane.ev[0], 0));
            waited[k] = true;
        }
        return lane.side[k - 1];
    };
    lane.ncls = 0;
    hipStream_t hs = nullptr;
    for (size_t n = 0; n < order.size(); n++) {
        const LdsClass &lc = cls[order[n]];
        // class n >= 1 on side stream first_side + n - 1 while there is one
        const int first_side = hubs ? 2 : 1;
        const int k = n > 0 && first_side + (int)n - 1 < nstreams ? first_side + (int)n - 1 : 0;
        hipStream_t sk = k == 0 ? st : side_stream(k);
        const int i = lane.ncls++;
        unsigned long long *cur = s.ctr + kCtrLds + (uint64_t)order[n] * kCurWords;
        // without hub kernels beside them, the classes take the registers
        // for 7 workgroups per CU
        HIP_CHECK(hipEventRecord(lane.cls_start[i], sk));
        HIP_CHECK(launch_lds_batches(lc.lg, pp.dev->cus, sk, c.kg, c.kp, R.recs, R.batch_first, R.batch_count,
                                     lc.b0, lc.b1, cur, c.out, nullptr, !hubs));
        HIP_CHECK(hipEventRecord(lane.cls_done[i], sk));
        lane.cls_lg[i] = (uint8_t)lc.lg;
        if (n == 0) {
            hs = side_stream(1);
            if (hubs) enqueue_hubs(pp, c, s, hs);
            HIP_CHECK(hipEventRecord(lane.ev[1], hs));
        }
    }
    if (!hs) {
        hs = side_stream(1);
        if (hubs) enqueue_hubs(pp, c, s, hs);
        HIP_CHECK(hipEventRecord(lane.ev[1], hs));
    }
    for (int k = 1; k < nstreams; k++)
        if (waited[k]) {
            HIP_CHECK(hipEventRecord(lane.side_done[k - 1], lane.side[k - 1]));
            HIP_CHECK(hipStreamWaitEvent(st, lane.side_done[k - 1], 0));
        }
    HIP_CHECK(hipEventRecord(lane.ev[2], st));
    HIP_CHECK(hipMemcpyAsync(host_ctr, s.ctr, sizeof(unsigned long long) * kCtrWords, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipEventRecord(lane.ev[3], st));
}

void fill_result(const beep_opts &o, const unsigned long long *acc, beep_result &res)
{
    uint64_t count = 0, hsum = 0, hxor = 0;
    for (int t = 0; t < kStripes; t++) {
        const unsigned long long *a = acc + kCtrAcc + t * kLine;
        count += a[0];
        hsum += a[1];
        hxor ^= a[2];
    }
    res.count = count;
    res.hash_sum = hsum;
    res.hash_xor = hxor;
    res.n_emitted = (o.emit != BEEP_EMIT_NONE) ? count : 0;
    res.n_written =
        (o.emit == BEEP_EMIT_BUFFER && o.match_cap) ? std::min<uint64_t>(acc[kCtrAcc + 3], o.match_cap) : 0;
    res.donations = acc[kCtrDonated];
    res.stopped = acc[kCtrStop] ? 1u : 0u;
}

void fill_profile(const unsigned long long *prof, beep_result &res)
{
    auto balance = [](const std::vector<double> &b) {
        double sum = 0, mx = 0;
        size_t n = 0;
        for (double x : b)
            if (x > 0) {
                sum += x;
                mx = std::max(mx, x);
                n++;
            }
        return n && mx > 0 ? sum / n / mx : 0.0;
    };
    std::vector<double> lds(kProfSlots), hub(kProfSlots);
    double own = 0, don = 0, lds_sum = 0, build = 0;
    for (int i = 0; i < kProfSlots; i++) {
        lds_sum += (double)prof[i];
        build += (double)prof[3 * kProfSlots + i];
        lds[i] = (double)prof[i];
        const double a = (double)prof[kProfSlots + i], b = (double)prof[2 * kProfSlots + i];
        hub[i] = a + b;
        own += a;
        don += b;
    }
    res.balance_lds = balance(lds);
    res.balance_hub = balance(hub);
    std::vector<double> cu(kProfSlots);
    uint64_t busy = 0;
    for (int i = 0; i < kProfSlots; i++) {
        cu[i] = (double)prof[kProfCuPhase * kProfSlots + i];
        busy += prof[kProfCuPhase * kProfSlots + i] ? 1 : 0;
    }
    res.balance_cu = balance(cu);
    res.cus_busy = busy;
    res.donated_frac = (own + don) > 0 ? don / (own + don) : 0.0;
    res.lds_build_frac = lds_sum > 0 ? build / lds_sum : 0.0;
}

// Sizes (grow-only) and carves the owned scratch region of a single query.
static void own_scratch(DevPrep &d, const ScratchNeeds &n)
{
    const uint64_t b = n.bytes();
    if (b > d.region_bytes) {
        if (d.region) HIP_CHECK(hipFree(d.region));
        d.region = nullptr;
        d.region_bytes = 0;
        HIP_CHECK(hipMalloc((void **)&d.region, b));
        d.region_bytes = b;
    }
    carve(d.scratch, d.region, n);
}

void count_on_device(Prep &pp, const Plan &p, const beep_opts &o, beep_result &res)
{
    upload(pp, o.device, p.skip_split);
    DevPrep &d = *pp.dev;
    HIP_CHECK(hipSetDevice(d.device));
    own_scratch(d, scratch_needs(pp, p, o, d.cus));
    d.lane.create(d.device, Lane::kMaxSide);
    const LaunchCfg c = make_launch(pp, p, o, d.scratch, d.cus);
    unsigned long long acc[kCtrWords];
    enqueue_query(pp, c, d.scratch, d.lane, acc);
    HIP_CHECK(hipStreamSynchronize(d.lane.main));
    float t_all = 0, t_hub = 0;
    HIP_CHECK(hipEventElapsedTime(&t_all, d.lane.ev[0], d.lane.ev[3]));
    HIP_CHECK(hipEventElapsedTime(&t_hub, d.lane.ev[0], d.lane.ev[1]));
    fill_result(o, acc, res);
    res.t_enum_ms = t_all;
    res.t_hub_ms = t_hub;
    res.t_lds_ms = lds_phase_ms(d.lane);
    class_times(d.lane, res.t_lds_class_ms);
    if (c.out.buf && o.match_buf && res.n_written)
        HIP_CHECK(hipMemcpy(o.match_buf, c.out.buf, res.n_written * (uint64_t)p.L * sizeof(uint32_t),
                            hipMemcpyDeviceToHost));
    if (d.scratch.prof) {
        std::vector<unsigned long long> prof((size_t)kProfPhases * kProfSlots);
        HIP_CHECK(hipMemcpy(prof.data(), d.scratch.prof, prof.size() * 8, hipMemcpyDeviceToHost));
        fill_profile(prof.data(), res);
    }
}

void upload_prep(Prep &pp, int device, int skip_split) { upload(pp, device, skip_split); }

// ------------------------------------------------------------------ batch
//
// Several prepared queries in one hipGraph: per query a counter clear and its
// deadline, one persistent lds_multi_kernel over every query's LDS batches,
// heaviest first, and then every query's hub rounds on a branch of its own.

struct Batch {
    int device = -1;
    std::vector<Prep *> preps;
    std::vector<const DevPrep *> devs;  // device state the graph was built from
    std::vector<const Plan *> plans;
    std::vector<beep_opts> opts;
    std::vector<LaunchCfg> cfgs;
    std::vector<char *> regions;        // per query scratch, owned by the batch
    std::vector<Scratch> scr;
    MultiQuery *dqs = nullptr;
    MultiWork *dwork = nullptr;
    uint32_t nwork = 0;
    uint32_t nlev = 1;
    // lane classes: work records [cls_w0, cls_w1) of width cls_lg, one launch
    // and one striped cursor (kCurWords words) each
    std::vector<int> cls_lg;
    std::vector<uint32_t> cls_w0, cls_w1;
    unsigned long long *cursor = nullptr;
    hipStream_t s0 = nullptr;
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> done;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~Batch()
    {
        if (device >= 0) (void)hipSetDevice(device);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void *p : {(void *)dqs, (void *)dwork, (void *)cursor})
            if (p) (void)hipFree(p);
        for (char *p : regions)
            if (p) (void)hipFree(p);
        for (auto e : done)
            if (e) (void)hipEventDestroy(e);
        for (auto e : {fork, t0, t1})
            if (e) (void)hipEventDestroy(e);
        for (auto s : branch)
            if (s) (void)hipStreamDestroy(s);
        if (s0) (void)hipStreamDestroy(s0);
    }
};

hipError_t launch_lds_multi(int cus, hipStream_t st, const MultiQuery *qs, const MultiWork *work, uint32_t nwork,
                            unsigned long long *cursor, uint32_t nlev, int lg, int *grid_out);

Batch *batch_create(Prep *const *preps, const Plan *const *plans, const beep_opts *opts, int n)
{
    if (n <= 0) throw Error(BEEP_EINVAL, "empty batch");
    auto *b = new Batch();
    try {
        b->device = opts[0].device;
        std::vector<MultiQuery> qs(n);
        std::vector<std::pair<uint64_t, MultiWork>> work;
        for (int i = 0; i < n; i++) {
            if (opts[i].emit == BEEP_EMIT_BUFFER) throw Error(BEEP_EINVAL, "batches do not return tuple buffers");
            if (opts[i].device != b->device) throw Error(BEEP_EINVAL, "a batch runs on one device");
            Prep &pp = *preps[i];
            upload(pp, opts[i].device, plans[i]->skip_split);
            DevPrep &d = *pp.dev;
            const ScratchNeeds need = scratch_needs(pp, *plans[i], opts[i], d.cus);
            char *region = nullptr;
            HIP_CHECK(hipMalloc((void **)&region, need.bytes()));
            b->regions.push_back(region);
            b->scr.emplace_back();
            carve(b->scr.back(), region, need);
            b->preps.push_back(preps[i]);
            b->devs.push_back(pp.dev);
            b->plans.push_back(plans[i]);
            b->opts.push_back(opts[i]);
            b->cfgs.push_back(make_launch(pp, *plans[i], opts[i], b->scr.back(), d.cus));
            const LaunchCfg &c = b->cfgs.back();
            qs[i] = MultiQuery{c.kg, c.kp, d.roots.recs, d.roots.batch_first, d.roots.batch_count, c.out};
            b->nlev = std::max<uint32_t>(b->nlev, (uint32_t)c.nlev);
            for (uint32_t bi = 0; c.kp.L >= 3 && bi < (uint32_t)pp.batches.size(); bi++) {
                const LdsBatch &lb = pp.batches[bi];
                uint64_t cost = 0;
                for (uint32_t r = 0; r < lb.count; r++) {
                    uint64_t dd = pp.wlen[lb.first + r];
                    cost += dd * dd;
                }
                uint32_t lg = std::max<uint32_t>(pp.batch_lanes_log2[bi], (uint32_t)c.lmin);
                work.push_back({cost, MultiWork{(uint32_t)i, bi, lg, 0}});
            }
        }
        // by lane class, heaviest first within a class
        std::stable_sort(work.begin(), work.end(), [](const auto &a, const auto &c) {
            if (a.second.lanes_log2 != c.second.lanes_log2) return a.second.lanes_log2 > c.second.lanes_log2;
            return a.first > c.first;
        });
        std::vector<MultiWork> w;
        for (auto &x : work) {
            const int lg = (int)x.second.lanes_log2;
            if (b->cls_lg.empty() || b->cls_lg.back() != lg) {
                b->cls_lg.push_back(lg);
                b->cls_w0.push_back((uint32_t)w.size());
                b->cls_w1.push_back((uint32_t)w.size());
            }
            w.push_back(x.second);
            b->cls_w1.back() = (uint32_t)w.size();
        }
        HIP_CHECK(hipSetDevice(b->device));
        b->dqs = dev_upload(qs);
        b->dwork = dev_upload(w);
        b->nwork = (uint32_t)w.size();
        b->cursor = dev_alloc<unsigned long long>(kCurWords * std::max<size_t>(1, b->cls_lg.size()));
        HIP_CHECK(hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking));
        b->branch.resize(n + b->cls_lg.size() + 1);
        b->done.resize(n + b->cls_lg.size() + 1);
        for (auto &s : b->branch) HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        for (auto &e : b->done) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        HIP_CHECK(hipEventCreate(&b->t0));
        HIP_CHECK(hipEventCreate(&b->t1));
    } catch (...) {
        delete b;
        throw;
    }
    return b;
}

static void batch_capture(Batch &b)
{
    const int n = (int)b.preps.size();
    HIP_CHECK(hipStreamBeginCapture(b.s0, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < n; i++) {
        const Scratch &s = b.scr[i];
        HIP_CHECK(launch_begin_sequence(s.ctr, kCtrTotal, kCtrDeadline, b.cfgs[i].deadline_ticks,
                                        b.cfgs[i].deadline_ticks_lds, b.s0));
        if (s.prof) HIP_CHECK(launch_fill_zero(s.prof, s.cap.prof_words * 8, b.s0));
    }
    const int ncls = (int)b.cls_lg.size();
    HIP_CHECK(launch_fill_zero(b.cursor, (uint64_t)kCurWords * 8 * std::max(ncls, 1), b.s0));
    // Every query's hub rounds on a branch of its own, then one persistent
    // launch per lane class (each with that class's LDS layout, so the narrow
    // classes keep their occupancy), widest first, each on a branch: they all
    // share the GPU, as the classes and hub rounds of one query do.
    const DevPrep &d0 = *b.preps[0]->dev;
    HIP_CHECK(hipEventRecord(b.fork, b.s0));
    for (int i = 0; i < n; i++) {
        HIP_CHECK(hipStreamWaitEvent(b.branch[i], b.fork, 0));
        enqueue_hubs(*b.preps[i], b.cfgs[i], b.scr[i], b.branch[i]);
        HIP_CHECK(hipEventRecord(b.done[i], b.branch[i]));
    }
    for (int k = 0; k < ncls; k++) {
        hipStream_t sk = b.branch[n + k];
        HIP_CHECK(hipStreamWaitEvent(sk, b.fork, 0));
        HIP_CHECK(launch_lds_multi(d0.cus, sk, b.dqs, b.dwork + b.cls_w0[k], b.cls_w1[k] - b.cls_w0[k],
                                   b.cursor + (uint64_t)k * kCurWords, b.nlev, b.cls_lg[k], nullptr));
        HIP_CHECK(hipEventRecord(b.done[n + k], sk));
    }
    for (int i = 0; i < n + ncls; i++) HIP_CHECK(hipStreamWaitEvent(b.s0, b.done[i], 0));
    HIP_CHECK(hipStreamEndCapture(b.s0, &b.graph));
    HIP_CHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
}

void batch_run(Batch &b, beep_result *res, double *t_ms)
{
    HIP_CHECK(hipSetDevice(b.device));
    for (size_t i = 0; i < b.preps.size(); i++)
        if (b.preps[i]->dev != b.devs[i])
            throw Error(BEEP_EINVAL, "a prepared query was uploaded again after the batch was created");
    if (!b.exec) batch_capture(b);
    HIP_CHECK(hipEventRecord(b.t0, b.s0));
    HIP_CHECK(hipGraphLaunch(b.exec, b.s0));
    HIP_CHECK(hipEventRecord(b.t1, b.s0));
    const int n = (int)b.preps.size();
    std::vector<unsigned long long> acc((size_t)n * kCtrWords);
    for (int i = 0; i < n; i++)
        HIP_CHECK(hipMemcpyAsync(acc.data() + (size_t)i * kCtrWords, b.scr[i].ctr,
                                 sizeof(unsigned long long) * kCtrWords, hipMemcpyDeviceToHost, b.s0));
    HIP_CHECK(hipStreamSynchronize(b.s0));
    float t = 0;
    HIP_CHECK(hipEventElapsedTime(&t, b.t0, b.t1));
    for (int i = 0; i < n; i++) {
        if (res) {
            memset(&res[i], 0, sizeof(beep_result));
            fill_result(b.opts[i], acc.data() + (size_t)i * kCtrWords, res[i]);
            res[i].t_enum_ms = t;
        }
    }
    if (t_ms) *t_ms = t;
}

void batch_free(Batch *b) { delete b; }

}  // namespace beep
