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
hipError_t launch_fill_zero(void *p, uint64_t bytes, hipStream_t st);
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
            // plans that never read a bitmap row need neither bitmap nor hash
            uint64_t w = pp.uses_bitmap ? (uint64_t)(hubs[h].d - hubs[h].rlo) * ((hubs[h].d + 63) / 64) : 0;
            uint64_t s = pp.uses_bitmap ? hubs[h].hslots : 0;
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

// Everything one enumeration launch sequence needs; built (and buffers
// allocated) before anything is enqueued, so the sequence can be captured.
struct LaunchCfg {
    KPlan kp;
    KGraph kg;
    KOut out;
    HubQueue hq;
    int hub_grid = 0;
    int nlev = 1;
    int lmin = 0;
    bool donate = false;
};

static LaunchCfg prepare_launch(Prep &pp, const Plan &p, const beep_opts &o, uint32_t *dbuf)
{
    upload(pp, o.device, p.skip_split);
    DevPrep &d = *pp.dev;
    HIP_CHECK(hipSetDevice(d.device));
    if (p.min_qdeg != pp.min_qdeg || p.max_qdeg != pp.max_qdeg || p.asc != pp.asc ||
        p.above_split != pp.reduced || p.rows_above_split != pp.rows_above_split)
        throw Error(BEEP_EINVAL, "plan does not match the prepared graph");
    LaunchCfg c;
    c.out = KOut{d.ctr + kCtrAcc, dbuf, (o.emit == BEEP_EMIT_BUFFER) ? o.match_cap : 0};
    c.kp = make_kplan(p, o.emit != BEEP_EMIT_NONE ? 1 : 0);
    c.kg = KGraph{d.ocol, d.vinfo};
    c.nlev = p.L > 3 ? p.L - 3 : 1;
    c.hub_grid = hub_dfs_occupancy(c.nlev, std::max<uint32_t>(d.wmax, 1)) * d.cus;
    if (d.nhubs) {
        // per wave: nlev stack rows + the row of x_1 (used when they do not fit LDS)
        size_t need = (size_t)c.hub_grid * (kBlock / 64) * (c.nlev + 1) * std::max<uint32_t>(d.wmax, 1);
        if (need > d.stk_words) {
            if (d.stk) HIP_CHECK(hipFree(d.stk));
            d.stk = dev_alloc<uint64_t>(need);
            d.stk_words = need;
        }
    }
    // donation queue (hub items with |C_2| > donate_min), allocated on first use
    c.donate = o.donate && d.nhubs && p.L >= 4;
    if (c.donate && !d.queue) {
        d.queue_cap = 1ull << 22;
        d.slot_cap = 1ull << 26;
        d.queue = dev_alloc<HubChunk>(d.queue_cap);
        d.slots = dev_alloc<uint64_t>(d.slot_cap);
    }
    if (c.donate) {
        c.hq.q = d.queue;
        c.hq.cap = d.queue_cap;
        c.hq.slots = d.slots;
        c.hq.slot_cap = d.slot_cap;
        c.hq.tail = d.ctr + kCtrQTail;
        c.hq.head = d.ctr + kCtrQHead;
        c.hq.slot_tail = d.ctr + kCtrSlot;
        c.hq.donated = d.ctr + kCtrDonated;
        c.hq.split_min = o.donate_min ? o.donate_min : 64;
        c.hq.chunk = c.hq.split_min;
    }
    c.lmin = 0;
    if (o.lanes > 0)
        while ((1 << c.lmin) < o.lanes && c.lmin < 6) c.lmin++;
    return c;
}

// Hub rounds on one stream: clear, hash, bitmap rows, DFS + donation drain.
static void enqueue_hubs(Prep &pp, const LaunchCfg &c, hipStream_t st)
{
    DevPrep &d = *pp.dev;
    for (size_t ri = 0; ri < d.rounds.size(); ri++) {
        const HubRound &r = d.rounds[ri];
        const uint64_t rows = r.g1 - r.g0;
        if (pp.uses_bitmap) {
            HIP_CHECK(launch_fill_zero(d.arena, r.ewords * 8, st));
            HIP_CHECK(launch_fill_zero(d.keys, r.hslots * 4, st));
            int g1 = (int)std::min<uint64_t>((rows + kBlock - 1) / kBlock, (uint64_t)d.cus * 16);
            HIP_CHECK(launch_hub_hash(std::max(g1, 1), st, c.kg, d.hubs, d.hub_of_row, r.g0, r.g1, d.keys, d.vals));
            int g2 = (int)std::min<uint64_t>((rows + 255) / 256, (uint64_t)d.cus * 16);
            HIP_CHECK(launch_hub_encode(std::max(g2, 1), st, c.kg, d.hubs, d.hub_of_row, r.g0, r.g1, d.keys,
                                        d.vals, d.arena));
        }
        unsigned long long *cur = d.ctr + kCtrHub + (ri % 64);
        if (ri >= 64) HIP_CHECK(launch_fill_zero(cur, 8, st));
        const uint64_t items = r.i1 - r.i0;
        if (items) {
            if (c.donate && ri > 0) HIP_CHECK(launch_fill_zero(d.ctr + kCtrQTail, 3 * 8, st));
            int g3 = (int)std::min<uint64_t>((items + 3) / 4, (uint64_t)c.hub_grid);
            HIP_CHECK(launch_hub_dfs(std::max(g3, 1), st, c.kg, c.kp, d.hubs, d.hub_of_item, r.i0, r.i1, cur,
                                     d.arena, d.stk, std::max<uint32_t>(d.wmax, 1), c.out, c.hq, 0));
            if (c.donate)
                HIP_CHECK(launch_hub_dfs(c.hub_grid, st, c.kg, c.kp, d.hubs, d.hub_of_item, r.i0, r.i1, cur,
                                         d.arena, d.stk, std::max<uint32_t>(d.wmax, 1), c.out, c.hq, 1));
        }
    }
}

// LDS lane classes, one launch each, round-robin over the given streams.
static void enqueue_lds_classes(Prep &pp, const LaunchCfg &c, hipStream_t *streams, int nst)
{
    DevPrep &d = *pp.dev;
    uint32_t b = 0;
    int cls = 0;
    while (b < d.nbatches) {
        int lg = std::max<int>(pp.batch_lanes_log2[b], c.lmin);
        uint32_t e = b + 1;
        while (e < d.nbatches && std::max<int>(pp.batch_lanes_log2[e], c.lmin) == lg) e++;
        uint32_t ewords = std::min<uint32_t>(LdsBudget::kMaxEWords, LdsBudget::kMaxSumDeg << std::min(lg, 3));
        if (cls >= 32) throw Error(BEEP_ERANGE, "too many lane classes");
        unsigned int *cur = (unsigned int *)(d.ctr + kCtrLds) + cls;
        HIP_CHECK(launch_lds_batches(lg, d.cus, streams[cls % nst], c.kg, c.kp, d.recs, d.batch_first,
                                     d.batch_count, b, e, cur, ewords, c.out, nullptr));
        b = e;
        cls++;
    }
}

static void fill_result(const DevPrep &d, const beep_opts &o, const unsigned long long *acc, beep_result &res)
{
    res.count = acc[kCtrAcc + 0];
    res.hash_sum = acc[kCtrAcc + 1];
    res.hash_xor = acc[kCtrAcc + 2];
    res.n_emitted = (o.emit != BEEP_EMIT_NONE) ? acc[0] : 0;
    res.n_written = (o.emit == BEEP_EMIT_BUFFER && o.match_cap) ? std::min<uint64_t>(acc[3], o.match_cap) : 0;
    res.donations = acc[kCtrDonated];
    (void)d;
}

void count_on_device(Prep &pp, const Plan &p, const beep_opts &o, beep_result &res)
{
    const int width = p.L;
    uint32_t *dbuf = nullptr;
    upload(pp, o.device, p.skip_split);
    HIP_CHECK(hipSetDevice(pp.dev->device));
    if (o.emit == BEEP_EMIT_BUFFER && o.match_cap) dbuf = dev_alloc<uint32_t>(o.match_cap * (uint64_t)width);
    LaunchCfg c = prepare_launch(pp, p, o, dbuf);
    DevPrep &d = *pp.dev;
    hipStream_t st = d.stream;
    HIP_CHECK(launch_fill_zero(d.ctr, sizeof(unsigned long long) * kCtrWords, st));
    HIP_CHECK(hipEventRecord(d.ev[0], st));
    // The LDS lane classes run on the side streams, concurrently with the hub
    // rounds on the main stream.
    for (auto &s : d.side) HIP_CHECK(hipStreamWaitEvent(s, d.ev[0], 0));
    enqueue_lds_classes(pp, c, d.side, DevPrep::kSide);
    for (int i = 0; i < DevPrep::kSide; i++) HIP_CHECK(hipEventRecord(d.side_done[i], d.side[i]));
    enqueue_hubs(pp, c, st);
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
    fill_result(d, o, acc, res);
    res.t_enum_ms = t_all;
    res.t_hub_ms = t_hub;
    res.t_lds_ms = t_lds;
    if (dbuf) {
        if (o.match_buf && res.n_written)
            HIP_CHECK(hipMemcpy(o.match_buf, dbuf, res.n_written * width * sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIP_CHECK(hipFree(dbuf));
    }
}

void upload_prep(Prep &pp, int device, int skip_split) { upload(pp, device, skip_split); }

// ------------------------------------------------------------------ batch
//
// Several prepared queries in one hipGraph: per query a counter clear and its
// hub rounds on a branch of their own, and one persistent lds_multi_kernel
// over every query's LDS batches, heaviest first.

struct Batch {
    int device = -1;
    std::vector<Prep *> preps;
    std::vector<const Plan *> plans;
    std::vector<beep_opts> opts;
    std::vector<LaunchCfg> cfgs;
    MultiQuery *dqs = nullptr;
    MultiWork *dwork = nullptr;
    uint32_t nwork = 0;
    uint32_t nlev = 1;
    unsigned int *cursor = nullptr;
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
                            unsigned int *cursor, uint32_t nlev, int *grid_out);

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
            b->preps.push_back(preps[i]);
            b->plans.push_back(plans[i]);
            b->opts.push_back(opts[i]);
            b->cfgs.push_back(prepare_launch(*preps[i], *plans[i], opts[i], nullptr));
            const LaunchCfg &c = b->cfgs.back();
            Prep &pp = *preps[i];
            DevPrep &d = *pp.dev;
            qs[i] = MultiQuery{c.kg, c.kp, d.recs, d.batch_first, d.batch_count, c.out};
            b->nlev = std::max<uint32_t>(b->nlev, (uint32_t)c.nlev);
            for (uint32_t bi = 0; bi < (uint32_t)pp.batches.size(); bi++) {
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
        std::stable_sort(work.begin(), work.end(), [](const auto &a, const auto &c) { return a.first > c.first; });
        std::vector<MultiWork> w;
        for (auto &x : work) w.push_back(x.second);
        HIP_CHECK(hipSetDevice(b->device));
        b->dqs = dev_upload(qs);
        b->dwork = dev_upload(w);
        b->nwork = (uint32_t)w.size();
        b->cursor = dev_alloc<unsigned int>(1);
        HIP_CHECK(hipStreamCreateWithFlags(&b->s0, hipStreamNonBlocking));
        b->branch.resize(n + 1);
        b->done.resize(n + 1);
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
    for (int i = 0; i < n; i++)
        HIP_CHECK(launch_fill_zero(b.preps[i]->dev->ctr, sizeof(unsigned long long) * kCtrWords, b.s0));
    HIP_CHECK(launch_fill_zero(b.cursor, 4, b.s0));
    HIP_CHECK(hipEventRecord(b.fork, b.s0));
    for (int i = 0; i < n; i++) {
        HIP_CHECK(hipStreamWaitEvent(b.branch[i], b.fork, 0));
        enqueue_hubs(*b.preps[i], b.cfgs[i], b.branch[i]);
        HIP_CHECK(hipEventRecord(b.done[i], b.branch[i]));
    }
    HIP_CHECK(hipStreamWaitEvent(b.branch[n], b.fork, 0));
    const DevPrep &d0 = *b.preps[0]->dev;
    if (b.nwork)
        HIP_CHECK(launch_lds_multi(d0.cus, b.branch[n], b.dqs, b.dwork, b.nwork, b.cursor, b.nlev, nullptr));
    HIP_CHECK(hipEventRecord(b.done[n], b.branch[n]));
    for (int i = 0; i <= n; i++) HIP_CHECK(hipStreamWaitEvent(b.s0, b.done[i], 0));
    HIP_CHECK(hipStreamEndCapture(b.s0, &b.graph));
    HIP_CHECK(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
}

void batch_run(Batch &b, beep_result *res, double *t_ms)
{
    HIP_CHECK(hipSetDevice(b.device));
    if (!b.exec) batch_capture(b);
    HIP_CHECK(hipEventRecord(b.t0, b.s0));
    HIP_CHECK(hipGraphLaunch(b.exec, b.s0));
    HIP_CHECK(hipEventRecord(b.t1, b.s0));
    const int n = (int)b.preps.size();
    std::vector<unsigned long long> acc((size_t)n * kCtrWords);
    for (int i = 0; i < n; i++)
        HIP_CHECK(hipMemcpyAsync(acc.data() + (size_t)i * kCtrWords, b.preps[i]->dev->ctr,
                                 sizeof(unsigned long long) * kCtrWords, hipMemcpyDeviceToHost, b.s0));
    HIP_CHECK(hipStreamSynchronize(b.s0));
    float t = 0;
    HIP_CHECK(hipEventElapsedTime(&t, b.t0, b.t1));
    for (int i = 0; i < n; i++) {
        if (res) {
            memset(&res[i], 0, sizeof(beep_result));
            fill_result(*b.preps[i]->dev, b.opts[i], acc.data() + (size_t)i * kCtrWords, res[i]);
            res[i].t_enum_ms = t;
        }
    }
    if (t_ms) *t_ms = t;
}

void batch_free(Batch *b) { delete b; }

}  // namespace beep
