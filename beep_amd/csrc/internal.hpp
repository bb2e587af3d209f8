// internal.hpp -- host-side data structures shared by the libbeep sources.
#pragma once

#include <atomic>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/beep.h"

namespace beep {

constexpr int kMaxK = BEEP_MAXK;

// Error carrying an errno-style code; the C ABI converts it into a return value
// and the thread's last-error message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);
// One line on stderr per finished preparation step when BEEP_VERBOSE is set.
void progress(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// ---------------------------------------------------------------- threading
unsigned host_threads();

// Dynamic parallel loop over [0, n) in chunks of `grain`.
template <class F>
void parallel_for(size_t n, F &&f, size_t grain = 1 << 14)
{
    unsigned T = host_threads();
    if (n <= grain || T <= 1) {
        if (n) f(size_t(0), n);
        return;
    }
    size_t chunks = (n + grain - 1) / grain;
    if (T > chunks) T = (unsigned)chunks;
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (;;) {
            size_t b = next.fetch_add(grain);
            if (b >= n) break;
            f(b, b + grain < n ? b + grain : n);
        }
    };
    std::vector<std::thread> th;
    th.reserve(T - 1);
    for (unsigned t = 1; t < T; ++t) th.emplace_back(worker);
    worker();
    for (auto &x : th) x.join();
}

// ---------------------------------------------------------------- graph
struct Graph {
    uint32_t n = 0;
    uint64_t m = 0;
    std::vector<uint64_t> rowptr;  // n + 1
    std::vector<uint32_t> col;     // m
    uint32_t max_degree = 0;
    uint32_t cutoff = 0;
    double mean = 0, sd = 0;
    bool rows_sorted = true;       // every row ascending by id
    void finalize();               // degree stats + cutoff
};

// Edge list -> CSR with CSRCOO::from_edgelist rules (CSRCOO.cuh:37-80).
std::unique_ptr<Graph> csr_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t ne);
// Reads .bel/.tsv/.mtx edges with the reference's byte semantics.
void read_edges(const std::string &path, std::vector<uint32_t> &src, std::vector<uint32_t> &dst);
// cutoff = ceil(mean + 6 sd) (utils.cuh:186-261).
uint32_t compute_cutoff(const Graph &g, double *mean, double *sd);
void write_bel(const Graph &g, const std::string &path);
// Format converters (BEEP_CONV_*), convert.cpp.
void convert(int kind, const std::string &src, const std::string &dst);

std::unique_ptr<Graph> gen_rmat(uint32_t scale, uint64_t samples, double a, double b, double c,
                                uint64_t seed, uint32_t n_keep);
std::unique_ptr<Graph> gen_chung_lu(uint32_t n, uint64_t samples, double gamma, double i0,
                                    uint64_t seed);
// Sorted unique undirected edge keys (u<v packed as u<<32|v) -> symmetric CSR.
std::unique_ptr<Graph> csr_from_undirected_keys(std::vector<uint64_t> &keys, uint32_t n);

// ---------------------------------------------------------------- plan
struct Plan {
    int k = 0, mode = BEEP_MODE_SGM, lunmat = 0, asc = 0, max_qdeg = 0, min_qdeg = 0;
    uint32_t seq[kMaxK] = {}, qdeg[kMaxK] = {};
    uint32_t qedge_ptr[kMaxK + 1] = {}, qedge[kMaxK * (kMaxK - 1) / 2] = {};
    uint32_t sym_ptr[kMaxK + 1] = {}, sym[kMaxK * (kMaxK - 1) / 2] = {};
    // derived, used by the kernels
    int L = 0;            // last materialised level + 1 = k - lunmat
    int skip_split = 0;   // SYMNODE_PTR[2] == 1 (sgm_kernelsLD.cuh:60)
    // Every candidate of every level is bound above x_1 (through the chain of
    // symmetry constraints) and x_1 >= split: the enumeration only touches row
    // indices >= split, so a root can be reduced to its larger-key neighbours.
    int above_split = 0;
    // Only bitmap rows >= split are ever read (rows of QEDGE positions).
    int rows_above_split = 0;
    // Some level intersects with a bitmap row (QEDGE(l)[1:] non-empty).
    int uses_bitmap = 0;
    // Every row x_q a level intersects with is used above x_q only: of each
    // adjacent pair, bitmaps need the bit in the row of the smaller index.
    int upper_rows = 0;
    uint16_t qmask[kMaxK] = {};  // positions QEDGE(l)[1:]
    uint16_t smask[kMaxK] = {};  // positions s in SYM(l), s > 0
};

std::unique_ptr<Plan> plan_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t ne,
                                      int mode);

// ---------------------------------------------------------------- prep
struct DevPrep;  // device mirror (sgm_engine.hip)

// One LDS-resident batch of consecutive roots.
struct LdsBatch {
    uint32_t first;   // index into roots
    uint32_t count;   // number of roots
};

struct DevGraph;  // device copy of a variant (sgm_engine.hip)

// Peeled and degree-oriented graph.  It depends on the query only through
// (min_qdeg, asc), so every query with the same pair can share one variant
// (and its copy in HBM).
struct Variant {
    uint32_t n = 0;
    uint64_t m = 0;
    int asc = 0, min_qdeg = 0, peel_rounds = 0;
    uint32_t cutoff = 0, max_degree = 0;
    std::vector<uint64_t> rowptr;  // peeled
    std::vector<uint32_t> col;     // peeled, id ascending
    std::vector<uint32_t> ocol;    // oriented
    std::vector<uint32_t> nlow;    // #smaller keys
    std::vector<uint32_t> hlen;    // length of the bitmap-build scan range of each row
    uint32_t deg(uint32_t v) const { return (uint32_t)(rowptr[v + 1] - rowptr[v]); }

    std::mutex dev_mu;             // guards dev
    DevGraph *dev = nullptr;
    ~Variant();
};

std::shared_ptr<Variant> make_variant(const Graph &g, int min_qdeg, int asc);

// One query's root queue and work split over a variant.
struct Prep {
    std::shared_ptr<Variant> var;
    int max_qdeg = 0;
    uint32_t max_root_degree = 0;
    std::vector<uint32_t> roots;   // R
    uint64_t q0_size = 0, boundary = 0;
    uint32_t q0_head = ~0u;        // Q0[0]
    uint32_t q0_head_larger = 0;   // its neighbours with a larger key
    uint64_t alg_bytes = 0;

    // Work list: the roots that carry work, with the part of the oriented row
    // the enumeration uses (offset, length), ordered by length descending.
    // work[0, n_hub) keep their bitmap in HBM ("hubs"); the rest are packed
    // into LDS batches.
    int reduced = 0;                  // rows start at the split (above_split plans)
    int skip_split = 0;               // plan's x_1 < split skip
    int rows_above_split = 0;         // hub bitmaps keep rows >= split only
    int uses_bitmap = 1;              // hub bitmaps are built at all
    int upper_rows = 0;               // bitmaps hold no mirrored bits
    std::vector<uint32_t> wu, woff, wlen;
    uint32_t n_hub = 0;
    std::vector<LdsBatch> batches;
    std::vector<uint8_t> batch_lanes_log2;  // per batch: log2 of DFS-group lanes
    // Per lane class lg (one lds_batch_kernel launch each): roots, bitmap-build
    // work units (range entries streamed, or root neighbours looked up) and
    // the bytes the launch must read at least once (root records and rows,
    // the row vertices' records and scan ranges, batch arrays).
    uint64_t cls_roots[8] = {}, cls_units[8] = {}, cls_bytes[8] = {};

    DevPrep *dev = nullptr;
    ~Prep();
};

// Fresh variant + root queue.
std::unique_ptr<Prep> prepare(const Graph &g, const Plan &p);
// Root queue over an existing variant (which must match the plan's
// min_qdeg and asc).  sample_stride > 1 keeps a stratified sample of the work
// list (ordered by row length): every sample_stride-th LDS root and every
// hub_sample_stride()-th hub root.
std::unique_ptr<Prep> prepare_on(const std::shared_ptr<Variant> &v, const Plan &p, uint32_t sample_stride = 1);
// A prepared query over the roots of src at the given ranks of src.roots
// (strictly ascending), sharing src's variant.
std::unique_ptr<Prep> prepare_subset(const Prep &src, const Plan &p, const uint64_t *ranks, uint64_t n);
void build_work(Prep &pp, const Plan &p, uint32_t sample_stride);
// Stride at which a sampled root queue keeps hub roots (finer: hubs are few and heavy).
inline uint32_t hub_sample_stride(uint32_t sample_stride) { return sample_stride >= 8 ? sample_stride / 8 : 1; }

// LDS budget of the batch kernel (words / slots); shared with the kernels.
// A batch of lane class lg holds roots of at most 64 << lg neighbours and a
// degree sum of at most sumdeg(lg), so its bitmaps take at most
// sumdeg(lg) << lg words.  The budget is sized so that the narrow classes,
// which carry most roots, fit six to seven workgroups on a CU (the batches
// wait on dependent memory accesses; more of them in flight hide it); every
// budget grows with lg, so the widest class's layout serves every batch.
struct LdsBudget {
    static constexpr uint32_t kMaxRoots = 128;     // roots per batch
    static constexpr uint32_t kMaxSumDeg = 512;    // largest sum of degrees per batch
    static constexpr uint32_t kMaxEWords = 4096;   // bitmap words (u64)
    static constexpr uint32_t kMaxHash = 2048;     // largest hash slots per batch
    static constexpr int kMaxClass = 6;
    static constexpr uint32_t sumdeg(int lg) { return lg >= 3 ? kMaxSumDeg : kMaxSumDeg / 2; }
    static constexpr uint32_t hash(int lg) { return 4 * sumdeg(lg); }
    static constexpr uint32_t ewords(int lg)
    {
        return (sumdeg(lg) << (lg < 3 ? lg : 3)) < kMaxEWords ? (sumdeg(lg) << (lg < 3 ? lg : 3)) : kMaxEWords;
    }
    // a root of d neighbours fits a batch of its lane class
    static bool fits(uint32_t d, uint32_t hash_slots)
    {
        const uint32_t w = (d + 63) / 64;
        int lg = 0;
        while ((1u << lg) < w) lg++;
        return (uint64_t)d * w <= ewords(lg) && d <= sumdeg(lg) && hash_slots <= hash(lg);
    }
};

uint32_t hash_slots_for(uint32_t d);
// Automatic LDS bitmap build: a row vertex whose larger-key range is longer
// than kLookupRatio x the root's degree looks the root's neighbours up in the
// range's lookup table instead of streaming the range.
constexpr uint32_t kLookupRatio = 4;
void build_tiers(Prep &pp);
void class_figures(Prep &pp);
// Bytes of the kernels' per-root record of an LDS batch (RootRec).
constexpr uint64_t kRootRecBytes = 32;

}  // namespace beep
