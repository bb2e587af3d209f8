// capi.cpp -- the extern "C" boundary declared in include/beep.h.
//
// Every entry point converts exceptions into an errno-style return code and a
// per-thread message (the reference exits the process on any error,
// utils.cuh:27-40, Config.h:86-87).

#include <sched.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "internal.hpp"

struct beep_graph { std::unique_ptr<beep::Graph> g; };
struct beep_plan { std::unique_ptr<beep::Plan> p; };
struct beep_prep { std::unique_ptr<beep::Prep> pp; };

namespace beep {
struct Batch;
Batch *batch_create(Prep *const *preps, const Plan *const *plans, const beep_opts *opts, int n);
void batch_run(Batch &b, beep_result *res, double *t_ms);
void batch_free(Batch *b);
}  // namespace beep

struct beep_batch { beep::Batch *b; };

namespace beep {
struct Server;
Server *server_create(const Graph *g, const beep_server_opts &o);
void server_prepare(Server &s, const Plan &p);
uint64_t server_submit(Server &s, const Plan &p, const beep_opts &o);
bool server_poll(Server &s, beep_server_done &out, int timeout_ms);
void server_stats(Server &s, beep_server_stats &out);
void server_free(Server *s);
}  // namespace beep

struct beep_server { beep::Server *s; };

namespace beep {

void count_on_device(Prep &pp, const Plan &p, const beep_opts &o, beep_result &res);
void count_noalloc(Prep &pp, const Plan &p, const beep_opts &o, void *stream, beep_result &res);
void upload_prep(Prep &pp, int device, int skip_split);

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

void progress(const char *fmt, ...)
{
    static const bool on = getenv("BEEP_VERBOSE") != nullptr;
    if (!on) return;
    static const auto t0 = std::chrono::steady_clock::now();
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    fprintf(stderr, "[beep %8.2fs] %s\n", s, buf);
    fflush(stderr);
}

unsigned host_threads()
{
    static unsigned cached = 0;
    if (cached) return cached;
    unsigned t = 0;
    if (const char *e = getenv("BEEP_THREADS")) t = (unsigned)atoi(e);
    if (!t) {
        if (const char *e = getenv("OMP_NUM_THREADS")) t = (unsigned)atoi(e);
    }
    if (!t) {
        cpu_set_t cs;
        CPU_ZERO(&cs);
        if (sched_getaffinity(0, sizeof(cs), &cs) == 0) t = (unsigned)CPU_COUNT(&cs);
    }
    if (!t) t = std::thread::hardware_concurrency();
    if (t > 64) t = 64;
    cached = t ? t : 1;
    return cached;
}

}  // namespace beep

using beep::Error;

template <class F>
static int guarded(F &&f)
{
    try {
        f();
        return BEEP_OK;
    } catch (const Error &e) {
        beep::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        beep::set_last_error("out of host memory");
        return BEEP_ENOMEM;
    } catch (const std::exception &e) {
        beep::set_last_error(e.what());
        return BEEP_EINVAL;
    }
}

// Checks the fields of beep_opts every entry point shares.
static void check_opts(const beep_opts &o)
{
    if (o.emit < BEEP_EMIT_NONE || o.emit > BEEP_EMIT_BUFFER) throw Error(BEEP_EINVAL, "bad emit mode");
    if (o.emit == BEEP_EMIT_BUFFER && o.match_cap && !o.match_buf)
        throw Error(BEEP_EINVAL, "match_buf is required with a match_cap");
    if (!(o.deadline_ms >= 0)) throw Error(BEEP_EINVAL, "bad deadline");
    if (o.build_mode < BEEP_BUILD_AUTO || o.build_mode > BEEP_BUILD_SEARCH) throw Error(BEEP_EINVAL, "bad build mode");
    if (o.lanes < 0 || o.lanes > 64) throw Error(BEEP_EINVAL, "lanes must be 0 (automatic) or 1..64");
}

extern "C" {

const char *beep_last_error(void) { return beep::g_last_error.c_str(); }
const char *beep_version(void) { return "beep-mi355x 0.1"; }

int beep_graph_load(const char *path, int sort_edges, beep_graph **out)
{
    return guarded([&] {
        if (!path || !out) throw Error(BEEP_EINVAL, "null argument");
        std::vector<uint32_t> s, d;
        beep::read_edges(path, s, d);
        if (sort_edges) {
            std::vector<uint64_t> k(s.size());
            for (size_t i = 0; i < s.size(); i++) k[i] = ((uint64_t)s[i] << 32) | d[i];
            std::sort(k.begin(), k.end());
            for (size_t i = 0; i < s.size(); i++) {
                s[i] = (uint32_t)(k[i] >> 32);
                d[i] = (uint32_t)k[i];
            }
        }
        auto g = beep::csr_from_edges(s.data(), d.data(), s.size());
        *out = new beep_graph{std::move(g)};
    });
}

int beep_graph_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t n, beep_graph **out)
{
    return guarded([&] {
        if ((!src || !dst) && n) throw Error(BEEP_EINVAL, "null edge arrays");
        if (!out) throw Error(BEEP_EINVAL, "null out");
        *out = new beep_graph{beep::csr_from_edges(src, dst, n)};
    });
}

int beep_graph_rmat(uint32_t scale, uint64_t samples, double a, double b, double c, uint64_t seed,
                    uint32_t n_keep, beep_graph **out)
{
    return guarded([&] {
        if (!out) throw Error(BEEP_EINVAL, "null out");
        *out = new beep_graph{beep::gen_rmat(scale, samples, a, b, c, seed, n_keep)};
    });
}

int beep_graph_chung_lu(uint32_t n, uint64_t samples, double gamma, double i0, uint64_t seed,
                        beep_graph **out)
{
    return guarded([&] {
        if (!out) throw Error(BEEP_EINVAL, "null out");
        *out = new beep_graph{beep::gen_chung_lu(n, samples, gamma, i0, seed)};
    });
}

int beep_graph_describe(const beep_graph *g, beep_graph_desc *o)
{
    return guarded([&] {
        if (!g || !o) throw Error(BEEP_EINVAL, "null argument");
        const auto &G = *g->g;
        o->n = G.n;
        o->m = G.m;
        o->max_degree = G.max_degree;
        o->cutoff = G.cutoff;
        o->mean = G.mean;
        o->sd = G.sd;
        o->rowptr = G.rowptr.data();
        o->col = G.col.data();
    });
}

int beep_graph_write_bel(const beep_graph *g, const char *path)
{
    return guarded([&] {
        if (!g || !path) throw Error(BEEP_EINVAL, "null argument");
        beep::write_bel(*g->g, path);
    });
}

void beep_graph_free(beep_graph *g) { delete g; }

int beep_convert(int kind, const char *src_path, const char *dst_path)
{
    return guarded([&] {
        if (!src_path || !dst_path) throw Error(BEEP_EINVAL, "null path");
        beep::convert(kind, src_path, dst_path);
    });
}

int beep_plan_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t n, int mode, beep_plan **out)
{
    return guarded([&] {
        if (!src || !dst || !out) throw Error(BEEP_EINVAL, "null argument");
        *out = new beep_plan{beep::plan_from_edges(src, dst, n, mode)};
    });
}

int beep_plan_from_mtx(const char *path, int mode, beep_plan **out)
{
    return guarded([&] {
        if (!path || !out) throw Error(BEEP_EINVAL, "null argument");
        std::vector<uint32_t> s, d;
        beep::read_edges(path, s, d);
        *out = new beep_plan{beep::plan_from_edges(s.data(), d.data(), s.size(), mode)};
    });
}

int beep_plan_describe(const beep_plan *p, beep_plan_desc *o)
{
    return guarded([&] {
        if (!p || !o) throw Error(BEEP_EINVAL, "null argument");
        const auto &P = *p->p;
        o->k = P.k;
        o->mode = P.mode;
        o->lunmat = P.lunmat;
        o->asc = P.asc;
        o->max_qdeg = P.max_qdeg;
        o->min_qdeg = P.min_qdeg;
        memcpy(o->seq, P.seq, sizeof(P.seq));
        memcpy(o->qdeg, P.qdeg, sizeof(P.qdeg));
        memcpy(o->qedge_ptr, P.qedge_ptr, sizeof(P.qedge_ptr));
        memcpy(o->qedge, P.qedge, sizeof(P.qedge));
        memcpy(o->sym_ptr, P.sym_ptr, sizeof(P.sym_ptr));
        memcpy(o->sym, P.sym, sizeof(P.sym));
        o->upper_rows = P.upper_rows;
    });
}

void beep_plan_free(beep_plan *p) { delete p; }

int beep_prepare(const beep_graph *g, const beep_plan *p, beep_prep **out)
{
    return guarded([&] {
        if (!g || !p || !out) throw Error(BEEP_EINVAL, "null argument");
        *out = new beep_prep{beep::prepare(*g->g, *p->p)};
    });
}

int beep_prep_describe(const beep_prep *pp, beep_prep_desc *o)
{
    return guarded([&] {
        if (!pp || !o) throw Error(BEEP_EINVAL, "null argument");
        const auto &P = *pp->pp;
        const auto &V = *P.var;
        o->n = V.n;
        o->m = V.m;
        o->n_roots = P.roots.size();
        o->q0_size = P.q0_size;
        o->boundary = P.boundary;
        o->cutoff = V.cutoff;
        o->max_degree = V.max_degree;
        o->max_root_degree = P.max_root_degree;
        o->peel_rounds = V.peel_rounds;
        o->asc = V.asc;
        o->alg_bytes = P.alg_bytes;
        o->n_hub_roots = P.n_hub;
        o->n_lds_batches = P.batches.size();
        o->rowptr = V.rowptr.data();
        o->col = V.col.data();
        o->ocol = V.ocol.data();
        o->nlow = V.nlow.data();
        o->roots = P.roots.data();
        o->p1 = (uint64_t)V.max_degree < (uint64_t)V.cutoff * 100u ? 1 : 0;
        o->p2 = (!V.asc || P.boundary == 0 || P.q0_head_larger == 0) ? 1 : 0;
        o->q0_head = P.q0_head;
        o->q0_head_larger = P.q0_head_larger;
        for (int c = 0; c < 8; c++) {
            o->lds_class_roots[c] = P.cls_roots[c];
            o->lds_class_units[c] = P.cls_units[c];
            o->lds_class_bytes[c] = P.cls_bytes[c];
        }
    });
}

int beep_prep_subset(const beep_prep *pp, const beep_plan *p, const uint64_t *ranks, uint64_t n,
                     beep_prep **out)
{
    return guarded([&] {
        if (!pp || !p || !out || (n && !ranks)) throw Error(BEEP_EINVAL, "null argument");
        *out = new beep_prep{beep::prepare_subset(*pp->pp, *p->p, ranks, n)};
    });
}

int beep_prep_upload(beep_prep *pp, int device)
{
    return guarded([&] {
        if (!pp) throw Error(BEEP_EINVAL, "null argument");
        beep::upload_prep(*pp->pp, device, pp->pp->skip_split);
    });
}

void beep_prep_free(beep_prep *pp) { delete pp; }

void beep_opts_default(beep_opts *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->device = 0;
    o->emit = BEEP_EMIT_NONE;
    o->lanes = 0;
    o->donate = 1;
    o->donate_min = 64;
    o->timing = 1;
}

int beep_count(beep_prep *pp, const beep_plan *p, const beep_opts *o, beep_result *out)
{
    return guarded([&] {
        if (!pp || !p || !out) throw Error(BEEP_EINVAL, "null argument");
        beep_opts def;
        beep_opts_default(&def);
        const beep_opts &opt = o ? *o : def;
        check_opts(opt);
        memset(out, 0, sizeof(*out));
        beep::count_on_device(*pp->pp, *p->p, opt, *out);
        if (out->stopped) throw Error(BEEP_ETIMEDOUT, "deadline passed: the counts are partial");
    });
}

int beep_count_noalloc(beep_prep *pp, const beep_plan *p, const beep_opts *o, void *stream, beep_result *out)
{
    return guarded([&] {
        if (!pp || !p || !out) throw Error(BEEP_EINVAL, "null argument");
        beep_opts def;
        beep_opts_default(&def);
        const beep_opts &opt = o ? *o : def;
        check_opts(opt);
        memset(out, 0, sizeof(*out));
        beep::count_noalloc(*pp->pp, *p->p, opt, stream, *out);
        if (out->stopped) throw Error(BEEP_ETIMEDOUT, "deadline passed: the counts are partial");
    });
}

int beep_batch_create(beep_prep *const *preps, const beep_plan *const *plans, const beep_opts *opts, int n,
                      beep_batch **out)
{
    return guarded([&] {
        if (!preps || !plans || !out || n <= 0) throw Error(BEEP_EINVAL, "null argument");
        std::vector<beep::Prep *> pv(n);
        std::vector<const beep::Plan *> lv(n);
        std::vector<beep_opts> ov(n);
        for (int i = 0; i < n; i++) {
            if (!preps[i] || !plans[i]) throw Error(BEEP_EINVAL, "null query");
            pv[i] = preps[i]->pp.get();
            lv[i] = plans[i]->p.get();
            if (opts) ov[i] = opts[i];
            else beep_opts_default(&ov[i]);
            check_opts(ov[i]);
        }
        *out = new beep_batch{beep::batch_create(pv.data(), lv.data(), ov.data(), n)};
    });
}

int beep_batch_run(beep_batch *b, beep_result *results, double *t_ms)
{
    return guarded([&] {
        if (!b) throw Error(BEEP_EINVAL, "null argument");
        beep::batch_run(*b->b, results, t_ms);
    });
}

void beep_batch_free(beep_batch *b)
{
    if (!b) return;
    beep::batch_free(b->b);
    delete b;
}

void beep_server_opts_default(beep_server_opts *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->device = 0;
    o->lanes = 4;
}

int beep_server_create(const beep_graph *g, const beep_server_opts *o, beep_server **out)
{
    return guarded([&] {
        if (!g || !out) throw Error(BEEP_EINVAL, "null argument");
        beep_server_opts def;
        beep_server_opts_default(&def);
        *out = new beep_server{beep::server_create(g->g.get(), o ? *o : def)};
    });
}

int beep_server_prepare(beep_server *s, const beep_plan *p)
{
    return guarded([&] {
        if (!s || !p) throw Error(BEEP_EINVAL, "null argument");
        beep::server_prepare(*s->s, *p->p);
    });
}

int beep_server_submit(beep_server *s, const beep_plan *p, const beep_opts *o, uint64_t *ticket)
{
    return guarded([&] {
        if (!s || !p || !ticket) throw Error(BEEP_EINVAL, "null argument");
        beep_opts def;
        beep_opts_default(&def);
        const beep_opts &opt = o ? *o : def;
        check_opts(opt);
        *ticket = beep::server_submit(*s->s, *p->p, opt);
    });
}

int beep_server_poll(beep_server *s, beep_server_done *out, int timeout_ms)
{
    return guarded([&] {
        if (!s || !out) throw Error(BEEP_EINVAL, "null argument");
        if (!beep::server_poll(*s->s, *out, timeout_ms)) throw Error(BEEP_ETIMEDOUT, "no query finished in time");
    });
}

int beep_server_stats_get(beep_server *s, beep_server_stats *out)
{
    return guarded([&] {
        if (!s || !out) throw Error(BEEP_EINVAL, "null argument");
        beep::server_stats(*s->s, *out);
    });
}

void beep_server_free(beep_server *s)
{
    if (!s) return;
    beep::server_free(s->s);
    delete s;
}

uint64_t beep_tuple_hash(const uint32_t *t, int len)
{
    auto mix = [](uint64_t z) {
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    };
    uint64_t h = 0x6a09e667f3bcc909ull ^ (uint64_t)len;
    for (int i = 0; i < len; i++) h = mix(h + 0x9e3779b97f4a7c15ull * (uint64_t)(i + 1) + t[i]);
    return h;
}

}  // extern "C"
