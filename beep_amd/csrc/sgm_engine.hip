// sgm_engine.hip -- device residency and the launch sequence of one
// enumeration (the host driver).
//
// Replaces count_subgraphs (subgraph_matching.cuh:1037-1279): the reference
// copies __constant__ tables, sizes a grid by occupancy, allocates and memsets
// per-block global scratch, launches the persistent kernel, adds the device
// counter into a managed counter and prints it.  Here the prepared graph is
// uploaded once (beep_prep_upload) and stays in HBM; beep_count launches, on
// one stream,
//   per hub round:   arena clear -> hub_hash -> hub_encode -> hub_dfs
//   per lane class:  lds_batch_kernel<P>
// and reads back four counters.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "internal.hpp"
#include "sgm_kernels.hpp"

namespace beep {

hipError_t launch_lds_batches(int lg, int cus, hipStream_t st, const KGraph &g, const KPlan &pl,
                              const RootRec *recs, const uint32_t *bf, const uint32_t *bc,
                              uint32_t b0, uint32_t b1, unsigned int *cursor, uint32_t ewords,
                              const KOut &out, int *grid_out);
hipError_t launch_hub_hash(int grid, hipStream_t st, const KGraph &g, const HubRec *hubs,
                           const uint32_t *hor, uint64_t g0, uint64_t g1, uint32_t *keys, uint32_t *vals);
hipError_t launch_hub_encode(int grid, hipStream_t st, const KGraph &g, const HubRec *hubs,
                             const uint32_t *hor, uint64_t g0, uint64_t g1, const uint32_t *keys,
                             const uint32_t *vals, uint64_t *E);
hipError_t launch_hub_dfs(int grid, hipStream_t st, const KGraph &g, const KPlan &pl, const HubRec *hubs,
                          const uint32_t *hoi, uint64_t i0, uint64_t i1, unsigned long long *cursor,
                          const uint64_t *E, uint64_t *stk, uint32_t wmax, const KOut &out,
                          const HubQueue &hq, int mode);
int hub_dfs_occupancy(int nlev, uint32_t wmax);

#define HIP_CHECK(x)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess)                                                                   \
            throw Error(BEEP_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_));         \
    } while (0)

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

constexpr int kCtrAcc = 0, kCtrHub = 4, kCtrLds = 68, kCtrQTail = 84, kCtrQHead = 85, kCtrSlot = 86,
              kCtrDonated = 87, kCtrWords = 88;

struct HubRound {
    uint32_t h0, h1;      // hub index range
    uint64_t g0, g1;      // global row range (bitmap build)
    uint64_t i0, i1;      // global item range (DFS)
    uint64_t ewords;      // bitmap words used
    uint64_t hslots;      // hash slots used
};

struct DevPrep {
    int device = -1;
    int cus = 0;
    uint32_t *ocol = nullptr;
    VInfo *vinfo = nullptr;
    // LDS tier
    RootRec *recs = nullptr;
    uint32_t *batch_first = nullptr, *batch_count = nullptr;
    uint32_t nbatches = 0;
    // hub tier
    HubRec *hubs = nullptr;
    uint32_t *hub_of_row = nullptr, *hub_of_item = nullptr;
    int skip_split = -1;  // plan property the hub items were built for
    uint32_t nhubs = 0;
    uint32_t wmax = 0;
    std::vector<HubRound> rounds;
    uint64_t arena_words = 0, arena_slots = 0;
    uint64_t *arena = nullptr;
    uint32_t *keys = nullptr, *vals = nullptr;
    uint64_t *stk = nullptr;
    size_t stk_words = 0;
    // donation queue of the hub DFS
    HubChunk *queue = nullptr;
    uint64_t *slots = nullptr;
    uint64_t queue_cap = 0, slot_cap = 0;
    // counters: acc[4] | hub cursors[64] | lds cursors[16] | queue tail, head,
    // slot tail, donated
    unsigned long long *ctr = nullptr;
    hipStream_t stream = nullptr;
    static constexpr int kSide = 3;      // streams of the LDS lane classes
    hipStream_t side[kSide] = {};
    hipEvent_t ev[4] = {};
    hipEvent_t side_done[kSide] = {};

    ~DevPrep()
    {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) cur = device;
        if (device >= 0) (void)hipSetDevice(device);
        for (void *p : {(void *)ocol, (void *)vinfo, (void *)recs, (void *)batch_first,
                        (void *)batch_count, (void *)hubs, (void *)hub_of_row, (void *)hub_of_item, (void *)arena,
                        (void *)keys, (void *)vals, (void *)stk, (void *)ctr, (void *)queue, (void *)slots})
            if (p) (void)hipFree(p);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : side_done)
            if (e) (void)hipEventDestroy(e);
        for (auto &s : side)
            if (s) (void)hipStreamDestroy(s);
        if (stream) (void)hipStreamDestroy(stream);
        if (device >= 0) (void)hipSetDevice(cur);
    }
};

Prep::~Prep() { delete dev; }

static void upload(Prep &pp, int device, int p_skip_split)
{
    if (pp.dev && pp.dev->device == device && pp.dev->skip_split == p_skip_split) return;
    delete pp.dev;
    pp.dev = nullptr;
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw Error(BEEP_EDEVICE, "no such HIP device");
    HIP_CHECK(hipSetDevice(device));
    auto *d = new DevPrep();
    pp.dev = d;
    d->device = device;
    d->skip_split = p_skip_split;
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, device));
    d->cus = prop.multiProcessorCount;
    HIP_CHECK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    for (auto &s : d->side) HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (auto &e : d->ev) HIP_CHECK(hipEventCreate(&e));
    for (auto &e : d->side_done) HIP_CHECK(hipEventCreate(&e));
    d->ocol = dev_upload(pp.ocol);
    {
        std::vector<VInfo> vi(pp.n);
        parallel_for(pp.n, [&](size_t b, size_t e2) {
            for (size_t v = b; v < e2; v++) {
                vi[v].hbeg = pp.rowptr[v] + (pp.asc ? pp.nlow[v] : 0u);
                vi[v].hlen = pp.hlen[v];
                vi[v].deg = (uint32_t)(pp.rowptr[v + 1] - pp.rowptr[v]);
            }
        });
        d->vinfo = dev_upload(vi);
    }

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
                r.rbeg = pp.rowptr[u] + pp.woff[wi];
                r.u = u;
                r.d = pp.wlen[wi];
                r.split = pp.reduced ? 0u : pp.nlow[u];
                r.doff = doff;
                r.eoff = eoff;
                r.hoff = hoff;
                doff += r.d;
                eoff += r.d * ((r.d + 63) / 64);
                hoff += hash_slots_for(r.d);
                recs.push_back(r);
            }
        }
        d->recs = dev_upload(recs);
        d->batch_first = dev_upload(bf);
        d->batch_count = dev_upload(bc);
        d->nbatches = (uint32_t)bf.size();
    }

    // Hubs: records, row -> hub map and rounds that fit the arena.
    {
        std::vector<HubRec> hubs;
        std::vector<uint32_t> hor, hoi;
        uint64_t joff = 0, ioff = 0;
        for (uint32_t h = 0; h < pp.n_hub; h++) {
            const uint32_t u = pp.wu[h];
            HubRec r{};
            r.rbeg = pp.rowptr[u] + pp.woff[h];
            r.u = u;
            r.d = pp.wlen[h];
            r.split = pp.reduced ? 0u : pp.nlow[u];
            r.hslots = hash_slots_for(r.d);
            r.joff = joff;
            r.ioff = ioff;
            // x_1 < split items are skipped (sgm_kernelsLD.cuh:60): they are not listed.
            r.jbeg = (p_skip_split && r.split < r.d) ? r.split : (p_skip_split ? r.d : 0u);
            r.rlo = pp.rows_above_split ? std::min(r.split, r.d) : 0u;
            joff += r.d;
            ioff += r.d - r.jbeg;
            for (uint32_t j = 0; j < r.d; j++) hor.push_back(h);
            for (uint32_t j = r.jbeg; j < r.d; j++) hoi.push_back(h);
            d->wmax = std::max<uint32_t>(d->wmax, (r.d + 63) / 64);
            hubs.push_back(r);
        }
        size_t freeb = 0, totalb = 0;
        HIP_CHECK(hipMemGetInfo(&freeb, &totalb));
        const uint64_t budget_bytes = std::min<uint64_t>((uint64_t)(freeb * 0.35), 48ull << 30);
        uint64_t ew = 0, hsl = 0;
        uint32_t h0 = 0;
        for (uint32_t h = 0; h < (uint32_t)hubs.size(); h++) {
            uint64_t w = (uint64_t)(hubs[h].d - hubs[h].rlo) * ((hubs[h].d + 63) / 64);
            uint64_t s = hubs[h].hslots;
            if (h > h0 && (ew + w) * 8 + (hsl + s) * 8 > budget_bytes) {
                d->rounds.push_back({h0, h, hubs[h0].joff, hubs[h].joff, hubs[h0].ioff, hubs[h].ioff, ew, hsl});
                h0 = h;
                ew = hsl = 0;
            }
            hubs[h].eoff = ew;
            hubs[h].hoff = hsl;
            ew += w;
            hsl += s;
            if ((ew * 8 + hsl * 8) > budget_bytes)
                throw Error(BEEP_ENOMEM, "hub root bitmap exceeds the HBM arena budget");
        }
        if (!hubs.empty())
            d->rounds.push_back({h0, (uint32_t)hubs.size(), hubs[h0].joff, joff, hubs[h0].ioff, ioff, ew, hsl});
        for (const auto &r : d->rounds) {
            d->arena_words = std::max(d->arena_words, r.ewords);
            d->arena_slots = std::max(d->arena_slots, r.hslots);
        }
        d->hubs = dev_upload(hubs);
        d->hub_of_row = dev_upload(hor);
        d->hub_of_item = dev_upload(hoi);
        d->nhubs = (uint32_t)hubs.size();
        if (d->nhubs) {
            d->arena = dev_alloc<uint64_t>(d->arena_words);
            d->keys = dev_alloc<uint32_t>(d->arena_slots);
            d->vals = dev_alloc<uint32_t>(d->arena_slots);
        }
    }
    d->ctr = dev_alloc<unsigned long long>(kCtrWords);
}

static KPlan make_kplan(const Plan &p, int emit)
{
    KPlan k{};
    k.k = p.k;
    k.L = p.L;
    k.skip_split = p.skip_split;
    k.emit = emit;
    for (int l = 0; l < kMaxLevels; l++) {
        k.qmask[l] = p.qmask[l];
        k.smask[l] = p.smask[l];
    }
    return k;
}

void count_on_device(Prep &pp, const Plan &p, const beep_opts &o, beep_result &res)
{
    upload(pp, o.device, p.skip_split);
    DevPrep &d = *pp.dev;
    HIP_CHECK(hipSetDevice(d.device));
    if (p.min_qdeg != pp.min_qdeg || p.max_qdeg != pp.max_qdeg || p.asc != pp.asc)
        throw Error(BEEP_EINVAL, "plan does not match the prepared graph");
    const int width = p.L;
    uint32_t *dbuf = nullptr;
    if (o.emit == BEEP_EMIT_BUFFER && o.match_cap) dbuf = dev_alloc<uint32_t>(o.match_cap * (uint64_t)width);
    KOut out{d.ctr + kCtrAcc, dbuf, (o.emit == BEEP_EMIT_BUFFER) ? o.match_cap : 0};
    KPlan kp = make_kplan(p, o.emit != BEEP_EMIT_NONE ? 1 : 0);
    KGraph kg{d.ocol, d.vinfo};
    hipStream_t st = d.stream;
    const int nlev = p.L > 3 ? p.L - 3 : 1;
    // hub DFS stack slots
    int hub_grid = hub_dfs_occupancy(nlev, std::max<uint32_t>(d.wmax, 1)) * d.cus;
    if (d.nhubs) {
        // per wave: nlev stack rows + the row of x_1 (used when they do not fit LDS)
        size_t need = (size_t)hub_grid * (kBlock / 64) * (nlev + 1) * std::max<uint32_t>(d.wmax, 1);
        if (need > d.stk_words) {
            if (d.stk) HIP_CHECK(hipFree(d.stk));
            d.stk = dev_alloc<uint64_t>(need);
            d.stk_words = need;
        }
    }
    // donation queue (hub items with |C_2| > donate_min), allocated on first use
    const bool donate = o.donate && d.nhubs && p.L >= 4;
    if (donate && !d.queue) {
        d.queue_cap = 1ull << 22;
        d.slot_cap = 1ull << 26;
        d.queue = dev_alloc<HubChunk>(d.queue_cap);
        d.slots = dev_alloc<uint64_t>(d.slot_cap);
    }
    HubQueue hq{};
    if (donate) {
        hq.q = d.queue;
        hq.cap = d.queue_cap;
        hq.slots = d.slots;
        hq.slot_cap = d.slot_cap;
        hq.tail = d.ctr + kCtrQTail;
        hq.head = d.ctr + kCtrQHead;
        hq.slot_tail = d.ctr + kCtrSlot;
        hq.donated = d.ctr + kCtrDonated;
        hq.split_min = o.donate_min ? o.donate_min : 64;
        hq.chunk = hq.split_min;
    }
    HIP_CHECK(hipMemsetAsync(d.ctr, 0, sizeof(unsigned long long) * kCtrWords, st));
    HIP_CHECK(hipEventRecord(d.ev[0], st));
    // The hub rounds run on the main stream; the LDS lane classes are
    // independent of them and of each other and run on the side streams.
    for (auto &s : d.side) HIP_CHECK(hipStreamWaitEvent(s, d.ev[0], 0));
    // LDS phase: consecutive batches with the same group width share a launch
    int lmin = 0;
    if (o.lanes > 0) {
        while ((1 << lmin) < o.lanes && lmin < 6) lmin++;
    }
    uint32_t b = 0;
    int cls = 0;
    while (b < d.nbatches) {
        int lg = std::max<int>(pp.batch_lanes_log2[b], lmin);
        uint32_t e = b + 1;
        while (e < d.nbatches && std::max<int>(pp.batch_lanes_log2[e], lmin) == lg) e++;
        uint32_t ewords = std::min<uint32_t>(LdsBudget::kMaxEWords, LdsBudget::kMaxSumDeg << std::min(lg, 3));
        if (cls >= 32) throw Error(BEEP_ERANGE, "too many lane classes");
        unsigned int *cur = (unsigned int *)(d.ctr + kCtrLds) + cls;
        HIP_CHECK(launch_lds_batches(lg, d.cus, d.side[cls % DevPrep::kSide], kg, kp, d.recs, d.batch_first,
                                     d.batch_count, b, e, cur, ewords, out, nullptr));
        b = e;
        cls++;
    }
    for (int i = 0; i < DevPrep::kSide; i++) HIP_CHECK(hipEventRecord(d.side_done[i], d.side[i]));
    // hub phase
    for (size_t ri = 0; ri < d.rounds.size(); ri++) {
        const HubRound &r = d.rounds[ri];
        HIP_CHECK(hipMemsetAsync(d.arena, 0, r.ewords * 8, st));
        HIP_CHECK(hipMemsetAsync(d.keys, 0, r.hslots * 4, st));
        const uint64_t rows = r.g1 - r.g0;
        int g1 = (int)std::min<uint64_t>((rows + kBlock - 1) / kBlock, (uint64_t)d.cus * 16);
        HIP_CHECK(launch_hub_hash(std::max(g1, 1), st, kg, d.hubs, d.hub_of_row, r.g0, r.g1, d.keys, d.vals));
        int g2 = (int)std::min<uint64_t>((rows + 255) / 256, (uint64_t)d.cus * 16);
        HIP_CHECK(launch_hub_encode(std::max(g2, 1), st, kg, d.hubs, d.hub_of_row, r.g0, r.g1, d.keys,
                                    d.vals, d.arena));
        unsigned long long *cur = d.ctr + kCtrHub + (ri % 64);
        if (ri >= 64) HIP_CHECK(hipMemsetAsync(cur, 0, 8, st));
        const uint64_t items = r.i1 - r.i0;
        if (items) {
            if (donate && ri > 0) HIP_CHECK(hipMemsetAsync(d.ctr + kCtrQTail, 0, 3 * 8, st));
            int g3 = (int)std::min<uint64_t>((items + 3) / 4, (uint64_t)hub_grid);
            HIP_CHECK(launch_hub_dfs(std::max(g3, 1), st, kg, kp, d.hubs, d.hub_of_item, r.i0, r.i1, cur,
                                     d.arena, d.stk, std::max<uint32_t>(d.wmax, 1), out, hq, 0));
            if (donate)
                HIP_CHECK(launch_hub_dfs(hub_grid, st, kg, kp, d.hubs, d.hub_of_item, r.i0, r.i1, cur, d.arena,
                                         d.stk, std::max<uint32_t>(d.wmax, 1), out, hq, 1));
        }
    }
    HIP_CHECK(hipEventRecord(d.ev[1], st));
    for (int i = 0; i < DevPrep::kSide; i++) HIP_CHECK(hipStreamWaitEvent(st, d.side_done[i], 0));
    HIP_CHECK(hipEventRecord(d.ev[2], st));
    unsigned long long acc[kCtrWords];
    HIP_CHECK(hipMemcpyAsync(acc, d.ctr, sizeof(acc), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipEventRecord(d.ev[3], st));
    HIP_CHECK(hipStreamSynchronize(st));
    float t_all = 0, t_hub = 0, t_lds = 0;
    HIP_CHECK(hipEventElapsedTime(&t_all, d.ev[0], d.ev[3]));
    HIP_CHECK(hipEventElapsedTime(&t_hub, d.ev[0], d.ev[1]));
    for (int i = 0; i < DevPrep::kSide; i++) {
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, d.ev[0], d.side_done[i]));
        t_lds = std::max(t_lds, t);
    }
    res.count = acc[0];
    res.hash_sum = acc[1];
    res.hash_xor = acc[2];
    res.n_emitted = (o.emit != BEEP_EMIT_NONE) ? acc[0] : 0;
    res.n_written = dbuf ? std::min<uint64_t>(acc[3], o.match_cap) : 0;
    res.t_enum_ms = t_all;
    res.t_hub_ms = t_hub;
    res.t_lds_ms = t_lds;
    res.donations = acc[kCtrDonated];
    if (dbuf) {
        if (o.match_buf && res.n_written)
            HIP_CHECK(hipMemcpy(o.match_buf, dbuf, res.n_written * width * sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIP_CHECK(hipFree(dbuf));
    }
}

void upload_prep(Prep &pp, int device, int skip_split) { upload(pp, device, skip_split); }

}  // namespace beep
