// sgm_kernels.hpp -- kernel-side data layout shared by the HIP kernels and the
// host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace beep {

constexpr int kBlock = 256;     // threads per workgroup (4 wave64s)
constexpr int kMaxLevels = 10;  // query positions (DEPTH)

// Device-wide atomics on one address serialise across the whole GPU, so
// result accumulators and work cursors are spread over kStripes counters,
// one per 128-byte line of kLine u64 words.
constexpr int kLine = 16, kStripes = 16;

// Query plan as the kernels consume it (by value).  Level l in [2, L) is
// the candidate set C_l:
//   C_l = {i in [0,d) \ {x_1..x_{l-1}} : E[x_q][i] = 1 for q in qmask[l],
//          i > x_s for s in smask[l]}
// which is count_tri for l = 2 and compute_intersection for l >= 3
// (device_utils.cuh:179-209, 426-456).
struct KPlan {
    int32_t k;
    int32_t L;           // k - LUNMAT
    int32_t skip_split;  // skip x_1 < split(u) (sgm_kernelsLD.cuh:60)
    int32_t emit;        // BEEP_EMIT_*
    // LDS bitmap build: a row vertex whose larger-key range is longer than
    // bin_ratio x the root's degree looks the root's larger-key neighbours up
    // in the range's lookup table instead of streaming the range (0: always,
    // ~0u: never)
    uint32_t bin_ratio;
    // Count-only plans whose last level L-1 does not intersect with the row
    // of x_{L-2}: C_{L-1} depends on x_{L-2} only through the exclusion of
    // x_{L-2} and, when L-2 is in smask[L-1], the bound x_{L-1} > x_{L-2}.
    // The kernels then sum |C_{L-1}| over all x_{L-2} in C_{L-2} in one pass
    // (prefix popcounts) instead of walking x_{L-2}.
    int32_t tail;
    // Count-only plans whose last three levels a = L-3, b = L-2, c = L-1 have
    //   qmask[c] = qmask[a] + {b},  smask[c] = smask[a] + {a},
    //   a in qmask[b], and qmask[b] - {a}, smask[b] below a:
    // x_c ranges over the same set S = C_a as x_a, restricted to the row of
    // x_b and above x_a, and x_b does not depend on x_a otherwise.  For fixed
    // x_1..x_{a-1}, the levels a..c then add up to
    //   sum over x_b in B of C(|S n E[x_b]|, 2),
    // B = the candidates of b without the a constraints (hub + 4-cycle, W5).
    int32_t pair;
    // Every row a level intersects with is used above its own index only
    // (through the symmetry bounds): the bitmap builds set, of each adjacent
    // pair, only the bit in the row of the smaller index.
    int32_t upper;
    uint32_t qmask[kMaxLevels];
    uint32_t smask[kMaxLevels];
};

// Per-vertex record read when the vertex appears in a root's row: where its
// bitmap-build scan range starts in ocol and how long it is (one 16-byte
// gather instead of three scattered ones).
struct VInfo {
    uint64_t hbeg;  // rowptr[v] (+ nlow[v] for ascending rows)
    uint32_t hlen;  // neighbours with a larger key in the orientation's order
    uint32_t deg;
};

struct KGraph {
    const uint32_t *ocol;    // degree-oriented rows
    const VInfo *vinfo;      // n records
    // Lookup table of every vertex's larger-key range: the range [hbeg, hbeg +
    // hlen) hashed into the 2 hlen slots at 2 hbeg (key w + 1, 0 empty,
    // linear probing from lookup_slot(w, 2 hlen)).
    const uint32_t *htab;
    int32_t asc;             // rows ascend by (degree, id); else degree descends, ids ascend
};

// First slot of w in a lookup table of `size` slots (host and device).
__host__ __device__ __forceinline__ uint32_t lookup_slot(uint32_t w, uint32_t size)
{
    uint32_t h = w * 0x9E3779B1u;
    h ^= h >> 15;
    return (uint32_t)(((uint64_t)h * size) >> 32);
}

// Staged range of a row vertex: start (40 bits) and length (24 bits).
constexpr int kRangeShift = 40;

// Per-root record of an LDS batch, offsets relative to the batch.
struct RootRec {
    uint64_t rbeg;   // rowptr[u]
    uint32_t u;      // root vertex
    uint32_t d;      // degree
    uint32_t split;  // nlow[u]
    uint32_t doff;   // offset of the root's row in the batch's concatenated rows
    uint32_t eoff;   // offset of the root's bitmap (u64 words)
    uint32_t hoff;   // offset of the root's hash table (slots)
};

// Per-hub record (bitmap and hash table in HBM).
struct HubRec {
    uint64_t rbeg;   // rowptr[u]
    uint64_t eoff;   // bitmap offset in the arena (u64 words)
    uint64_t hoff;   // hash offset in the arena (slots)
    uint32_t u, d, split, hslots;
    uint64_t joff;   // first global row index of this hub (bitmap build)
    uint64_t ioff;   // first global DFS item of this hub
    uint32_t jbeg;   // first x_1 of the hub's items
    uint32_t rlo;    // first bitmap row kept (rows below are never read)
};

// Donated level-2 subtree of a hub item: x_2 in C_2 n [lo, hi), with C_2
// published at slots[slot, slot + W).
struct HubChunk {
    uint32_t hub;
    uint32_t x1;
    uint64_t slot;
    uint32_t lo, hi;
};

// Every field defaults to null / 0, i.e. donation off, unless set on purpose.
struct HubQueue {
    HubChunk *q = nullptr;                    // null: donation off
    unsigned long long *tail = nullptr;       // records reserved
    unsigned long long *head = nullptr;       // striped cursor of the draining pass
    uint64_t cap = 0;
    uint64_t *slots = nullptr;                // published C_2 bitmaps
    unsigned long long *slot_tail = nullptr;  // words reserved
    uint64_t slot_cap = 0;
    unsigned long long *donated = nullptr;    // records published over the whole call
    uint32_t split_min = 0;                   // donate when |C_2| > split_min
    uint32_t chunk = 0;                       // candidates per record
};

// Words (u64) of one hub-DFS wave's buffer: nlev candidate-stack rows and the
// bitmap row of x_1.
__host__ __device__ inline uint64_t hub_wave_words(int nlev, uint32_t wmax)
{
    return (uint64_t)(nlev + 1) * wmax;
}

struct KOut {
    // kStripes lines: stripe s holds [0] count [1] hash_sum [2] hash_xor of
    // the waves that flush to it; word 3 of stripe 0 counts tuples written
    unsigned long long *acc = nullptr;
    uint32_t *buf = nullptr;  // device tuple buffer (cap * L words) or null
    uint64_t cap = 0;
    // Deadlines in wall_clock64() ticks, written on the device by the launch
    // sequence's first kernel (all ones: none): dl[0] for the hub kernels,
    // dl[1] for the LDS kernels.  Kernels stop claiming work and leave their
    // DFS once theirs has passed, and set *stop.
    const unsigned long long *dl = nullptr;
    unsigned long long *stop = nullptr;
    // Optional busy-time accumulators (wall_clock64 ticks), kProfPhases x
    // kProfSlots: phase 0 LDS workgroups, 1 hub waves on their own items, 2 hub
    // waves on donated subtrees, 3 the LDS workgroups' staging and bitmap
    // build; slot = workgroup (LDS) or wave (hub) index.  Phase 4 sums
    // phases 0-2 by compute unit (slot = XCC, shader engine, array, CU).
    unsigned long long *prof = nullptr;
};

constexpr int kProfSlots = 4096, kProfPhases = 5, kProfCuPhase = 4;

// One prepared query of a multi-query launch.
struct MultiQuery {
    KGraph g;
    KPlan pl;
    const RootRec *recs;
    const uint32_t *batch_first;
    const uint32_t *batch_count;
    KOut out;
};

// One LDS batch of one query.
struct MultiWork {
    uint32_t query;
    uint32_t batch;
    uint32_t lanes_log2;
    uint32_t pad;
};

}  // namespace beep
