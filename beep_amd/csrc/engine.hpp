// engine.hpp -- device-side state of prepared queries and the launch
// sequence, shared by the single-query engine, batches and the server.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "internal.hpp"
#include "sgm_kernels.hpp"

namespace beep {

#define HIP_CHECK(x)                                                                            \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess)                                                                   \
            throw Error(BEEP_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_));         \
    } while (0)

// Counter words of one launch sequence (u64), in 128-byte lines of kLine
// words.  Read back after the run (kCtrWords): the kStripes accumulator
// stripes (kCtrAcc; see KOut::acc), then one line with the donation tail,
// slot tail and donated count, the deadlines of the hub and the LDS kernels
// and the stop flag.  Cleared but not read back: striped work cursors
// (kCurWords each) of the hub rounds (64), the donation drain and the LDS
// lane classes (32).
constexpr int kCtrAcc = 0;
constexpr int kCtrQTail = kStripes * kLine, kCtrSlot = kCtrQTail + 1, kCtrDonated = kCtrQTail + 2,
              kCtrDeadline = kCtrQTail + 3, kCtrStop = kCtrQTail + 5;
constexpr int kCtrWords = (kStripes + 1) * kLine;
constexpr int kCurWords = kStripes * kLine;
constexpr int kCtrHub = kCtrWords, kCtrHead = kCtrHub + 64 * kCurWords, kCtrLds = kCtrHead + kCurWords,
              kCtrTotal = kCtrLds + 32 * kCurWords;

// HBM copy of a variant: oriented rows and per-vertex records.
struct DevGraph {
    int device = -1;
    uint32_t *ocol = nullptr;
    VInfo *vinfo = nullptr;
    uint32_t *htab = nullptr;  // lookup tables of the larger-key ranges (2 m slots)
    ~DevGraph();
};

struct HubRound {
    uint32_t h0, h1;      // hub index range
    uint64_t g0, g1;      // global row range (bitmap build)
    uint64_t i0, i1;      // global item range (DFS)
    uint64_t ewords;      // bitmap words used
    uint64_t hslots;      // hash slots used
    uint64_t er0 = 0, er1 = 0;  // its rows with a non-empty scan range: enc_rows[er0, er1)
};

// Read-only device records of one prepared root queue.
struct DevRoots {
    RootRec *recs = nullptr;
    uint32_t *batch_first = nullptr, *batch_count = nullptr;
    uint32_t nbatches = 0;
    HubRec *hubs = nullptr;
    uint32_t *hub_of_row = nullptr, *hub_of_item = nullptr;
    // per hub row: the part of the row vertex's larger-key range the bitmap
    // build streams (start | length << kRangeShift)
    uint64_t *hub_range = nullptr;
    uint64_t *enc_rows = nullptr;  // per round, the hub rows whose scan range is not empty
    int skip_split = -1;  // plan property the hub items were built for
    uint32_t nhubs = 0;
    uint32_t wmax = 0;
    std::vector<HubRound> rounds;
    uint64_t arena_budget = 0;            // bytes the rounds were cut for
    uint64_t arena_words = 0, arena_slots = 0;
    uint64_t don_words = 0, don_records = 0;  // upper bounds of the donation ring
    void release();
};

// Scratch sizes of one launch sequence.
struct ScratchNeeds {
    uint64_t arena_words = 0, hash_slots = 0, stk_words = 0;
    uint64_t queue_cap = 0, slot_cap = 0, dbuf_words = 0, prof_words = 0;
    uint64_t bytes() const;
};

// Scratch of one launch sequence, carved out of one device region.
struct Scratch {
    uint64_t *arena = nullptr;
    uint64_t *kv = nullptr;  // hub hash slots: key (vertex + 1) | row index << 32
    uint64_t *stk = nullptr;
    HubChunk *queue = nullptr;
    uint64_t *slots = nullptr;
    uint32_t *dbuf = nullptr;
    unsigned long long *ctr = nullptr;
    unsigned long long *prof = nullptr;
    ScratchNeeds cap;
};
void carve(Scratch &s, char *base, const ScratchNeeds &n);

// Streams and events of one launch sequence: the heaviest LDS lane class on
// `main` right after the counters are cleared, the hub rounds on side[0], the
// other lane classes on the side streams (side[1..] when there are more).
struct Lane {
    static constexpr int kMaxSide = 3;  // with main: the 4 hardware queues a process gets by default
    int device = -1;
    int nside = 0;
    hipStream_t main = nullptr;
    hipStream_t side[kMaxSide] = {};
    hipEvent_t ev[4] = {};  // start, hubs done, all done, counters copied
    hipEvent_t side_done[kMaxSide] = {};
    // start and end of each LDS lane class's launch (recorded on its stream
    // right before and after it) and its lane class
    static constexpr int kMaxClasses = 32;
    hipEvent_t cls_start[kMaxClasses] = {};
    hipEvent_t cls_done[kMaxClasses] = {};
    uint8_t cls_lg[kMaxClasses] = {};
    int ncls = 0;
    void create(int device, int nside);
    ~Lane();
};

// Single-query device state of a Prep: its records, an owned scratch
// region and a lane.
struct DevPrep {
    int device = -1;
    int cus = 0;
    DevRoots roots;
    char *region = nullptr;
    uint64_t region_bytes = 0;
    Scratch scratch;
    Lane lane;
    ~DevPrep();
};

// Everything one enumeration launch sequence needs.
struct LaunchCfg {
    KPlan kp{};
    KGraph kg{};
    KOut out{};
    HubQueue hq{};  // q == nullptr: donation off
    int hub_grid = 0;
    int nlev = 1;
    int lmin = 0;
    bool donate = false;
    uint64_t deadline_ticks = 0;      // hub kernels
    uint64_t deadline_ticks_lds = 0;  // LDS kernels
};

// Zero fill kernel (no memset nodes, so sequences can be captured) and the
// first kernel of a launch sequence, which clears n counter words and writes
// the deadlines at words dl, dl + 1.
// dl: skip the clear once that deadline has passed (hub arenas only)
hipError_t launch_fill_zero(void *p, uint64_t bytes, hipStream_t st, const unsigned long long *dl = nullptr);
hipError_t launch_begin_sequence(unsigned long long *ctr, uint64_t n, uint64_t dl, uint64_t ticks_hub,
                                 uint64_t ticks_lds, hipStream_t st);

int device_cus(int device);
uint64_t wall_clock_khz(int device);

// Uploads the variant (once per variant) and the root records of pp, with hub
// rounds cut to arena_budget bytes (0: the default budget).
void upload(Prep &pp, int device, int skip_split, uint64_t arena_budget = 0);
ScratchNeeds scratch_needs(const Prep &pp, const Plan &p, const beep_opts &o, int cus);
// lds_deadline_ms < 0: the LDS kernels share o.deadline_ms
LaunchCfg make_launch(const Prep &pp, const Plan &p, const beep_opts &o, const Scratch &s, int cus,
                      double lds_deadline_ms = -1.0);
// clear counters + deadline, LDS classes on the side streams and hub rounds on
// main, join, counters -> host_ctr; records lane.ev[0..3].
void enqueue_query(const Prep &pp, const LaunchCfg &c, const Scratch &s, Lane &lane,
                   unsigned long long *host_ctr);
void enqueue_hubs(const Prep &pp, const LaunchCfg &c, const Scratch &s, hipStream_t st);
// LDS lane classes, one launch each, round-robin over nst streams; with
// cls_done, the end of class i's launch is recorded in cls_done[i] and its
// lane class in cls_lg[i].  Returns the number of classes.
int enqueue_lds_classes(const Prep &pp, const LaunchCfg &c, const Scratch &s, hipStream_t *streams, int nst,
                        hipEvent_t *cls_done = nullptr, uint8_t *cls_lg = nullptr);
// Device time of each lane class's launch (ms, by lane class) from a lane's
// events, and of the whole LDS phase (ev[0] to the last class's end).
void class_times(const Lane &lane, double *t_by_lg);
double lds_phase_ms(const Lane &lane);
void fill_result(const beep_opts &o, const unsigned long long *acc, beep_result &res);
// Balance figures from the busy-time accumulators (kProfPhases x kProfSlots).
void fill_profile(const unsigned long long *prof, beep_result &res);

}  // namespace beep
