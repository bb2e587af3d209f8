// server.hip -- the serving loop: one resident data graph, queries admitted
// continuously as lanes and scratch memory free up.
//
// The reference runs one query per process (it reloads and re-prepares the
// graph every time, main.cu:70-341).  A server keeps, per data graph,
//   * the peeled + oriented variants, keyed by (min_qdeg, asc), in HBM;
//   * the prepared root queues, keyed by every plan property they depend on;
// and one pool of HBM from which each running query's scratch (hub bitmap
// arena, hash tables, DFS stacks, donation ring, tuple buffer) is carved.
// Submitted queries wait in arrival order; the worker admits the oldest as
// soon as a lane (a main and a side stream) is free and its predicted
// footprint fits the free part of the pool, so queries that arrive while
// others run are started right away and share the GPU with them.  A query
// can carry a deadline after which its kernels stop (the result is then
// marked stopped).  With a budget, each new plan first runs as a pilot on a
// stratified 1/stride sample of its root queue (hub roots sampled 8x finer;
// deadlines 1.5 x budget / the stride of each phase):
// the larger of the pilot's hub and LDS phase times, each scaled by its
// stride, over the average number of lanes busy while it ran, predicts the
// full run alone, and plans predicted above 1.5 x the budget are refused
// without running.  Predictions are cached per plan.  A pilot runs alone
// (beside persistent kernels that hold the CUs its time says little), and a
// query with a prediction starts only when the predictions of the queries
// running beside it and its own fit the budget together (or when nothing
// else runs), so sharing the GPU does not push admitted queries past their
// deadline.

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>
#include <tuple>

#include "engine.hpp"

namespace beep {

using Clock = std::chrono::steady_clock;

// A plan is refused when its predicted device time exceeds kPilotSlack x the
// budget (the pilot's deadlines scale the same way).  Sampling error makes
// predictions noisy; refusing a query that would finish costs a completion,
// admitting one that would not costs at most its deadline.
constexpr double kPilotSlack = 1.5;

static double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit allocator over one device region (offsets, 256-byte granules).
class PoolAlloc {
  public:
    void reset(uint64_t bytes)
    {
        free_.clear();
        free_[0] = bytes;
        used_ = peak_ = 0;
    }
    static constexpr uint64_t npos = ~0ull;
    uint64_t alloc(uint64_t bytes)
    {
        bytes = (bytes + 255) & ~255ull;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const uint64_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return off;
        }
        return npos;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        bytes = (bytes + 255) & ~255ull;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }

  private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t used_ = 0, peak_ = 0;
};

struct Server {
    struct Job {
        uint64_t ticket = 0;
        Plan plan;
        beep_opts opts{};
        Clock::time_point t_submit, t_admit;
        std::shared_ptr<Prep> pp;
        std::shared_ptr<Prep> pilot;  // sampled root queue while a pilot runs first
        bool piloting = false;
        uint64_t busy_at_pilot = 1;  // lanes running when the pilot started (itself included)
        double predicted_ms = -1.0;
        double charged_ms = 0;       // its share of running_pred while it runs
        bool counted = false;        // counted in running_full / running_pred
        std::string sig;
        ScratchNeeds need;
        uint64_t off = 0, bytes = 0;
    };
    struct Slot {
        Lane lane;
        unsigned long long *host_ctr = nullptr;
        bool busy = false;
        bool busy_pilot = false;  // its launch sequence is a pilot
        Job job;
        Scratch s;
        LaunchCfg c;
    };
    using RootKey = std::tuple<int, int, int, int, int, int, int, int>;

    const Graph *g = nullptr;
    int device = 0, cus = 0;
    uint64_t pool_bytes = 0, arena_cap = 0;
    double budget_ms = 0;
    uint32_t stride = 32;
    char *pool = nullptr;
    PoolAlloc alloc;
    std::vector<std::unique_ptr<Slot>> slots;

    std::mutex mu;  // incoming, done, stopping, counters
    std::condition_variable cv_in, cv_out;
    std::deque<Job> incoming;
    std::deque<beep_server_done> done;
    bool stopping = false;
    uint64_t next_ticket = 1;
    beep_server_stats st{};

    std::mutex cache_mu;  // variants, rootq, pilotq
    std::map<std::pair<int, int>, std::shared_ptr<Variant>> variants;
    std::map<RootKey, std::shared_ptr<Prep>> rootq, pilotq;

    std::deque<Job> pending;  // worker only
    std::map<std::string, double> predicted;  // worker only: plan -> predicted device ms
    // worker only: queries (not pilots) running, the sum of their predictions,
    // and pilots running
    int running_full = 0;
    double running_pred = 0;
    int running_pilots = 0;
    bool fits_time(const Job &j) const
    {
        return budget_ms <= 0 || j.predicted_ms < 0 || running_full == 0 ||
               running_pred + j.predicted_ms <= budget_ms;
    }
    std::thread worker;

    ~Server();
    std::shared_ptr<Prep> get_prep(const Plan &p, uint32_t sample_stride = 1);
    void run();
    void launch(Slot &sl);
    beep_opts pilot_opts(const beep_opts &o) const;
    void finish(Slot &sl, int status, const std::string &err);
    void fail(Job &j, int status);
};

std::shared_ptr<Prep> Server::get_prep(const Plan &p, uint32_t sample_stride)
{
    std::lock_guard<std::mutex> lk(cache_mu);
    const RootKey key{p.min_qdeg, p.asc, p.max_qdeg, p.above_split, p.rows_above_split, p.uses_bitmap,
                      p.upper_rows, p.skip_split};
    auto &cache = sample_stride > 1 ? pilotq : rootq;
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    auto &v = variants[{p.min_qdeg, p.asc}];
    if (!v) v = make_variant(*g, p.min_qdeg, p.asc);
    std::shared_ptr<Prep> pp(prepare_on(v, p, sample_stride).release());
    upload(*pp, device, p.skip_split, arena_cap);
    cache[key] = pp;
    return pp;
}

// Everything of a plan the enumeration depends on.
// Prediction key: the plan, and whether the query emits tuples (count-only
// runs may sum the last level instead of walking it).
static std::string plan_signature(const Plan &p, bool emits)
{
    std::string s;
    auto put = [&](const void *b, size_t n) { s.append((const char *)b, n); };
    const int head[] = {p.k, p.mode, p.lunmat, p.asc, p.max_qdeg, p.min_qdeg, p.L, emits ? 1 : 0};
    put(head, sizeof(head));
    put(p.seq, sizeof(p.seq));
    put(p.qedge_ptr, sizeof(p.qedge_ptr));
    put(p.qedge, sizeof(p.qedge));
    put(p.sym_ptr, sizeof(p.sym_ptr));
    put(p.sym, sizeof(p.sym));
    return s;
}

static ScratchNeeds max_needs(const ScratchNeeds &a, const ScratchNeeds &b)
{
    ScratchNeeds n;
    n.arena_words = std::max(a.arena_words, b.arena_words);
    n.hash_slots = std::max(a.hash_slots, b.hash_slots);
    n.stk_words = std::max(a.stk_words, b.stk_words);
    n.queue_cap = std::max(a.queue_cap, b.queue_cap);
    n.slot_cap = std::max(a.slot_cap, b.slot_cap);
    n.dbuf_words = std::max(a.dbuf_words, b.dbuf_words);
    n.prof_words = std::max(a.prof_words, b.prof_words);
    return n;
}

beep_opts Server::pilot_opts(const beep_opts &o) const
{
    beep_opts po = o;
    // an emitting query walks every tuple: its pilot does too (hashes only)
    po.emit = o.emit != BEEP_EMIT_NONE ? BEEP_EMIT_HASH : BEEP_EMIT_NONE;
    po.match_cap = 0;
    po.match_buf = nullptr;
    // hub phase: long enough for the finer-sampled hubs to reach kPilotSlack x
    // budget; the LDS phase has its own deadline (kPilotSlack x budget / stride)
    po.deadline_ms = kPilotSlack * budget_ms / hub_sample_stride(stride);
    return po;
}

// Enqueues the slot's current stage: the pilot, or the query itself.
void Server::launch(Slot &sl)
{
    Job &j = sl.job;
    if (j.piloting) {
        const beep_opts po = pilot_opts(j.opts);
        sl.c = make_launch(*j.pilot, j.plan, po, sl.s, cus, kPilotSlack * budget_ms / stride);
        enqueue_query(*j.pilot, sl.c, sl.s, sl.lane, sl.host_ctr);
        running_pilots++;
    } else {
        sl.c = make_launch(*j.pp, j.plan, j.opts, sl.s, cus);
        enqueue_query(*j.pp, sl.c, sl.s, sl.lane, sl.host_ctr);
        running_full++;
        j.charged_ms = j.predicted_ms > 0 ? j.predicted_ms : 0;
        running_pred += j.charged_ms;
        j.counted = true;
    }
}

void Server::fail(Job &j, int status)
{
    beep_server_done d{};
    d.ticket = j.ticket;
    d.status = status;
    d.footprint = j.bytes;
    d.predicted_ms = j.predicted_ms;
    const auto now = Clock::now();
    d.latency_ms = ms_between(j.t_submit, now);
    std::lock_guard<std::mutex> lk(mu);
    if (status == BEEP_ERANGE) st.refused++;
    else st.failed++;
    st.pending--;
    done.push_back(d);
    cv_out.notify_all();
}

void Server::finish(Slot &sl, int status, const std::string &err)
{
    Job &j = sl.job;
    beep_server_done d{};
    d.ticket = j.ticket;
    d.footprint = j.bytes;
    d.predicted_ms = j.predicted_ms;
    if (status == 0) {
        try {
            fill_result(j.opts, sl.host_ctr, d.result);
            float t_all = 0, t_hub = 0;
            HIP_CHECK(hipEventElapsedTime(&t_all, sl.lane.ev[0], sl.lane.ev[3]));
            HIP_CHECK(hipEventElapsedTime(&t_hub, sl.lane.ev[0], sl.lane.ev[1]));
            d.result.t_enum_ms = t_all;
            d.result.t_hub_ms = t_hub;
            d.result.t_lds_ms = lds_phase_ms(sl.lane);
            if (sl.c.out.buf && j.opts.match_buf && d.result.n_written) {
                HIP_CHECK(hipMemcpyAsync(j.opts.match_buf, sl.c.out.buf,
                                         d.result.n_written * (uint64_t)j.plan.L * sizeof(uint32_t),
                                         hipMemcpyDeviceToHost, sl.lane.main));
                HIP_CHECK(hipStreamSynchronize(sl.lane.main));
            }
            if (d.result.stopped) status = BEEP_ETIMEDOUT;
        } catch (const Error &e) {
            status = e.code;
            set_last_error(e.what());
        }
    } else {
        set_last_error(err);
    }
    d.status = status;
    const auto now = Clock::now();
    d.queue_ms = ms_between(j.t_submit, j.t_admit);
    d.service_ms = ms_between(j.t_admit, now);
    d.latency_ms = ms_between(j.t_submit, now);
    alloc.release(j.off, j.bytes);
    sl.busy = false;
    if (j.piloting && sl.busy_pilot) {  // a pilot ended with an error
        running_pilots--;
        sl.busy_pilot = false;
    }
    if (j.counted) {  // a query (not a pilot) ends: it no longer shares the GPU
        running_full--;
        running_pred -= j.charged_ms;
        if (running_full == 0 || running_pred < 0) running_pred = 0;
        j.counted = false;
        j.charged_ms = 0;
    }
    j.pp.reset();
    j.pilot.reset();
    std::lock_guard<std::mutex> lk(mu);
    st.running--;
    st.pool_used = alloc.used();
    if (status == 0) st.completed++;
    else if (status == BEEP_ETIMEDOUT) st.stopped++;
    else if (status == BEEP_ERANGE) st.refused++;
    else st.failed++;
    done.push_back(d);
    cv_out.notify_all();
}

void Server::run()
{
    (void)hipSetDevice(device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            if (incoming.empty() && pending.empty() && st.running == 0) {
                if (stopping) break;
                cv_in.wait(lk, [&] { return stopping || !incoming.empty(); });
            }
            while (!incoming.empty()) {
                pending.push_back(std::move(incoming.front()));
                incoming.pop_front();
            }
        }
        bool progress = false;
        // completions
        for (auto &sp : slots) {
            Slot &sl = *sp;
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.lane.ev[3]);
            if (q == hipErrorNotReady) continue;
            progress = true;
            if (q == hipSuccess && sl.job.piloting) {
                // pilot done: predict, then refuse or run the query on the same lane
                running_pilots--;
                sl.busy_pilot = false;
                float t_hub = 0, t_lds = 0;
                bool ok = hipEventElapsedTime(&t_hub, sl.lane.ev[0], sl.lane.ev[1]) == hipSuccess;
                for (int i = 0; ok && i < sl.lane.ncls; i++) {
                    float t = 0;
                    ok = hipEventElapsedTime(&t, sl.lane.ev[0], sl.lane.cls_done[i]) == hipSuccess;
                    t_lds = std::max(t_lds, t);
                }
                // hub and LDS phases run side by side; each scales by its own stride
                const double t = std::max((double)t_hub * hub_sample_stride(stride) / stride, (double)t_lds);
                const bool stopped = sl.host_ctr[kCtrStop] != 0;
                // the pilot shared the GPU with the other running queries: its
                // time over the average number of busy lanes estimates it alone
                uint64_t busy_now = 1;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    busy_now = std::max<uint64_t>(st.running, 1);
                }
                const double share = 0.5 * (double)(sl.job.busy_at_pilot + busy_now);
                const double pred = (stopped || !ok) ? INFINITY : (double)t * stride / share;
                predicted[sl.job.sig] = pred;
                sl.job.predicted_ms = pred;
                sl.job.piloting = false;
                if (pred > kPilotSlack * budget_ms) {
                    finish(sl, BEEP_ERANGE, "predicted device time above the budget");
                    continue;
                }
                if (!fits_time(sl.job)) {
                    // it would not finish in time beside the running queries:
                    // give the lane and memory back and wait at the queue's head
                    alloc.release(sl.job.off, sl.job.bytes);
                    sl.busy = false;
                    pending.push_front(std::move(sl.job));
                    std::lock_guard<std::mutex> lk(mu);
                    st.running--;
                    st.pending++;
                    st.pool_used = alloc.used();
                    continue;
                }
                try {
                    launch(sl);
                } catch (const Error &e) {
                    (void)hipStreamSynchronize(sl.lane.main);
                    finish(sl, e.code, e.what());
                }
                continue;
            }
            finish(sl, q == hipSuccess ? 0 : BEEP_EDEVICE, std::string("hipEventQuery: ") + hipGetErrorString(q));
        }
        // admissions, oldest first
        while (!pending.empty()) {
            Job &j = pending.front();
            if (!j.pp) {
                try {
                    j.pp = get_prep(j.plan);
                    j.need = scratch_needs(*j.pp, j.plan, j.opts, cus);
                    if (budget_ms > 0) {
                        j.sig = plan_signature(j.plan, j.opts.emit != BEEP_EMIT_NONE);
                        auto it = predicted.find(j.sig);
                        if (it != predicted.end()) {
                            j.predicted_ms = it->second;
                            if (j.predicted_ms > kPilotSlack * budget_ms)
                                throw Error(BEEP_ERANGE, "predicted device time above the budget");
                        } else {
                            j.pilot = get_prep(j.plan, stride);
                            j.need = max_needs(j.need, scratch_needs(*j.pilot, j.plan, pilot_opts(j.opts), cus));
                            j.piloting = true;
                        }
                    }
                    j.bytes = j.need.bytes();
                    if (j.bytes > pool_bytes)
                        throw Error(BEEP_ENOMEM, "the query's predicted scratch exceeds the server pool");
                } catch (const Error &e) {
                    set_last_error(e.what());
                    fail(j, e.code);
                    pending.pop_front();
                    progress = true;
                    continue;
                }
            }
            // Pilots run alone: beside running queries (persistent kernels that
            // hold the CUs) their time says little about the plan.  Queries
            // wait while a pilot runs and until the running ones leave room.
            if (j.piloting ? (running_full > 0 || running_pilots > 0) : (running_pilots > 0 || !fits_time(j)))
                break;
            Slot *free_slot = nullptr;
            for (auto &sp : slots)
                if (!sp->busy) {
                    free_slot = sp.get();
                    break;
                }
            if (!free_slot) break;
            const uint64_t off = alloc.alloc(j.bytes);
            if (off == PoolAlloc::npos) break;
            Slot &sl = *free_slot;
            sl.job = std::move(j);
            pending.pop_front();
            sl.job.off = off;
            sl.busy = true;
            sl.job.t_admit = Clock::now();
            {
                std::lock_guard<std::mutex> lk(mu);
                st.pending--;
                st.running++;
                st.peak_running = std::max(st.peak_running, st.running);
                st.pool_used = alloc.used();
                st.pool_peak = alloc.peak();
            }
            try {
                carve(sl.s, pool + off, sl.job.need);
                if (sl.job.piloting) {
                    std::lock_guard<std::mutex> lk(mu);
                    st.pilots++;
                    sl.job.busy_at_pilot = std::max<uint64_t>(st.running, 1);
                }
                launch(sl);
                sl.busy_pilot = sl.job.piloting;
            } catch (const Error &e) {
                (void)hipStreamSynchronize(sl.lane.main);
                finish(sl, e.code, e.what());
            }
            progress = true;
        }
        if (!progress) {
            std::unique_lock<std::mutex> lk(mu);
            cv_in.wait_for(lk, std::chrono::microseconds(50), [&] { return stopping || !incoming.empty(); });
        }
    }
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu);
        stopping = true;
    }
    cv_in.notify_all();
    if (worker.joinable()) worker.join();
    (void)hipSetDevice(device);
    for (auto &sp : slots)
        if (sp->host_ctr) (void)hipHostFree(sp->host_ctr);
    slots.clear();
    rootq.clear();
    variants.clear();
    if (pool) (void)hipFree(pool);
}

Server *server_create(const Graph *g, const beep_server_opts &o)
{
    if (o.lanes < 1 || o.lanes > 16) throw Error(BEEP_EINVAL, "lanes must be in 1..16");
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (o.device < 0 || o.device >= ndev) throw Error(BEEP_EDEVICE, "no such HIP device");
    HIP_CHECK(hipSetDevice(o.device));
    std::unique_ptr<Server> s(new Server());
    s->g = g;
    s->device = o.device;
    s->cus = device_cus(o.device);
    size_t freeb = 0, totalb = 0;
    HIP_CHECK(hipMemGetInfo(&freeb, &totalb));
    s->pool_bytes = o.pool_bytes ? o.pool_bytes : (uint64_t)(freeb / 4);
    s->pool_bytes &= ~255ull;
    if (s->pool_bytes < (1ull << 20)) throw Error(BEEP_EINVAL, "pool too small");
    s->arena_cap = o.arena_cap ? o.arena_cap : s->pool_bytes / (2ull * o.lanes);
    if (!(o.budget_ms >= 0)) throw Error(BEEP_EINVAL, "bad budget");
    s->budget_ms = o.budget_ms;
    s->stride = o.pilot_stride ? o.pilot_stride : 32;
    HIP_CHECK(hipMalloc((void **)&s->pool, s->pool_bytes));
    s->alloc.reset(s->pool_bytes);
    s->st.pool_bytes = s->pool_bytes;
    for (int i = 0; i < o.lanes; i++) {
        auto sl = std::make_unique<Server::Slot>();
        sl->lane.create(o.device, 1);
        HIP_CHECK(hipHostMalloc((void **)&sl->host_ctr, sizeof(unsigned long long) * kCtrWords));
        s->slots.push_back(std::move(sl));
    }
    Server *raw = s.get();
    s->worker = std::thread([raw] { raw->run(); });
    return s.release();
}

void server_prepare(Server &s, const Plan &p)
{
    (void)s.get_prep(p);
    if (s.budget_ms > 0) (void)s.get_prep(p, s.stride);
}

uint64_t server_submit(Server &s, const Plan &p, const beep_opts &o)
{
    Server::Job j;
    j.plan = p;
    j.opts = o;
    j.opts.device = s.device;
    j.t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.stopping) throw Error(BEEP_EINVAL, "server is shutting down");
    j.ticket = s.next_ticket++;
    const uint64_t t = j.ticket;
    s.incoming.push_back(std::move(j));
    s.st.submitted++;
    s.st.pending++;
    s.cv_in.notify_one();
    return t;
}

bool server_poll(Server &s, beep_server_done &out, int timeout_ms)
{
    std::unique_lock<std::mutex> lk(s.mu);
    auto ready = [&] { return !s.done.empty(); };
    if (timeout_ms < 0) s.cv_out.wait(lk, ready);
    else if (!s.cv_out.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return false;
    out = s.done.front();
    s.done.pop_front();
    return true;
}

void server_stats(Server &s, beep_server_stats &out)
{
    {
        std::lock_guard<std::mutex> lk(s.mu);
        out = s.st;
    }
    std::lock_guard<std::mutex> lk(s.cache_mu);
    out.variants = s.variants.size();
    out.root_queues = s.rootq.size() + s.pilotq.size();
}

void server_free(Server *s) { delete s; }

}  // namespace beep
