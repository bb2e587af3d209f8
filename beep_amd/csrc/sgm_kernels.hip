// sgm_kernels.hip -- CDNA4 (gfx950) kernels of the enumeration path.
//
// The reference runs one persistent CUDA kernel in which a 320-thread block
// claims a root, builds the root's d x ceil(d/32) neighbourhood bitmap in
// global memory by binary search, and 16-lane partitions walk the DFS over
// 32-bit words (sgm_kernelsLD.cuh:7-273, device_utils.cuh:60-456).
//
// Here the same enumeration is organised for wave64:
//  * bitmap words are 64 bits; a DFS group is P lanes of one wave (P = 1..64,
//    lane i of a group owns words i, i+P, ...), the next candidate is found
//    with one wave ballot over the group's words and one ds_bpermute;
//  * roots whose bitmap fits are packed into LDS batches: a workgroup stages
//    the rows, an LDS hash (vertex -> row index) and the bitmaps, then its
//    groups claim (root, x_1) items from an LDS counter;
//  * the bitmap is built by streaming, for every row vertex v, only the part
//    of v's oriented row that holds neighbours with a larger (degree, id) key
//    and probing the hash -- each adjacent pair inside N(u) is found once and
//    both bits are set -- instead of d^2/2 binary searches; when that part is
//    long next to the root's degree, the root's neighbours on v's larger-key
//    side are looked up in a hash table of it instead (one load each);
//  * roots too large for LDS ("hubs") keep bitmap + hash in an HBM arena built
//    by grid-wide kernels, and every wave of the grid takes (hub, x_1) items,
//    so one hub is shared by the whole GPU.
//
// Candidate sets follow count_tri (device_utils.cuh:179-209) for level 2 and
// compute_intersection (:426-456) for deeper levels; counting stops at level
// L-1 = k-1-LUNMAT as backtrack does (:294-326).

#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"
#include "sgm_kernels.hpp"

namespace beep {

// ------------------------------------------------------------------ helpers

// bits [0, d) of word w
__device__ __forceinline__ uint64_t valid_mask(uint32_t w, uint32_t d)
{
    const uint32_t lo = w << 6;
    if (lo + 64 <= d) return ~0ull;
    if (lo >= d) return 0ull;
    return (1ull << (d - lo)) - 1ull;
}

// bits > x of word w (explicit branches: shifts by 64 are never issued)
__device__ __forceinline__ uint64_t gt_mask(uint32_t x, uint32_t w)
{
    const uint32_t xw = x >> 6;
    if (xw > w) return 0ull;
    if (xw < w) return ~0ull;
    const uint32_t b = x & 63u;
    return b == 63u ? 0ull : (~0ull << (b + 1u));
}

__device__ __forceinline__ uint64_t bit_at(uint32_t x, uint32_t w)
{
    return (x >> 6) == w ? (1ull << (x & 63u)) : 0ull;
}

// bits < x of word w
__device__ __forceinline__ uint64_t lt_mask(uint32_t x, uint32_t w)
{
    const uint32_t xw = x >> 6;
    if (xw > w) return ~0ull;
    if (xw < w) return 0ull;
    return (1ull << (x & 63u)) - 1ull;
}

// Pairs (x, y) with x in S, y in F and y > x, inside one word: C(|S|, 2) when
// S = F, else a loop over the sparser operand.
__device__ __forceinline__ uint32_t pairs_in_word(uint64_t S, uint64_t F)
{
    if (S == F) {
        const uint32_t k = (uint32_t)__popcll(S);
        return k * (k - 1) / 2;
    }
    uint32_t n = 0;
    if (__popcll(S) <= __popcll(F)) {
        while (S) {
            const uint32_t b = (uint32_t)__builtin_ctzll(S);
            S &= S - 1;
            n += (uint32_t)__popcll((F >> b) >> 1);
        }
    } else {
        while (F) {
            const uint32_t b = (uint32_t)__builtin_ctzll(F);
            F &= F - 1;
            n += (uint32_t)__popcll(S & ((1ull << b) - 1ull));
        }
    }
    return n;
}

__device__ __forceinline__ uint32_t hash_slot(uint32_t v, uint32_t H)
{
    uint32_t h = v * 0x9E3779B1u;
    h ^= h >> 15;
    return h & (H - 1u);
}

// An adjacent pair (j, i) of an LDS neighbourhood bitmap (W words per row):
// both bits, or with `upper` only the one in the row of the smaller index
// (the symmetry bounds never read a row below its own index then).
__device__ __forceinline__ void set_pair(uint64_t *E, uint32_t W, uint32_t j, uint32_t i, int upper)
{
    if (!upper || i > j) atomicOr((unsigned long long *)&E[j * W + (i >> 6)], 1ull << (i & 63));
    if (!upper || j > i) atomicOr((unsigned long long *)&E[i * W + (j >> 6)], 1ull << (j & 63));
}

// smallest power of two >= max(4, 4d): at most a quarter of the slots are
// taken, so a probe for an absent key rarely goes past its first slot
__host__ __device__ __forceinline__ uint32_t hash_slots_dev(uint32_t d)
{
    if (d <= 1) return 4;
    return 1u << (32 - __builtin_clz(4 * d - 1));
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t hash_step(uint64_t h, int pos, uint32_t v)
{
    return mix64(h + 0x9e3779b97f4a7c15ull * (uint64_t)(pos + 1) + v);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint64_t wave_xor_u64(uint64_t v)
{
    for (int o = 32; o > 0; o >>= 1) v ^= __shfl_xor(v, o, 64);
    return v;
}

// Wave inclusive scan.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    return v;
}

// Deadline check: wall_clock64() is a constant-rate device clock.
__device__ __forceinline__ bool past(uint64_t dl) { return (uint64_t)wall_clock64() > dl; }

__device__ __forceinline__ void mark_stop(const KOut &out) { atomicOr(out.stop, 1ull); }

// DFS steps between two deadline checks.
constexpr uint32_t kClockEvery = 1024;

// Paired last levels of the hub DFS: bit-sliced counters of kPairSlices bits
// count up to kPairMembers members of S (more than any hub row holds below the
// root window's 100 x cutoff on graphs of up to 2^26 vertices).  The
// instantiation with them is held to 3 waves per SIMD (168 VGPRs, a few spilled).
constexpr int kPairSlices = 20;
constexpr uint32_t kPairMembers = (1u << kPairSlices) - 1u;
// A 5-vertex pair plan's hub item of more than kPairSplitWork member-rounds
// (|C_2| x rounds of 64 words of B) is donated in parts of about
// kPairChunkWork member-rounds, whole rounds each.
constexpr uint64_t kPairSplitWork = 1u << 13;
constexpr uint32_t kPairChunkWork = 1u << 12;

// Compute unit the wave runs on, as a slot below 2048: the XCC (HW_REG_XCC_ID
// bits 2:0) and, within it, shader engine, shader array and CU (HW_REG_HW_ID
// bits 15:8).  The reference reads %smid for its per-SM times (utils.cuh:66-99).
__device__ __forceinline__ uint32_t cu_slot()
{
    const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | (0 << 6) | 4);    // HW_ID, 32 bits
    const uint32_t xcc = __builtin_amdgcn_s_getreg((15 << 11) | (0 << 6) | 20);  // XCC_ID, 16 bits
    return ((xcc & 7u) << 8) | ((hw >> 8) & 0xffu);
}

// Busy time since t0 into phase `phase`, slot `slot`; phases 0-2 (work, not
// the build share) also into the compute unit's slot of phase kProfCuPhase.
__device__ __forceinline__ void prof_add(const KOut &out, int phase, uint32_t slot, uint64_t t0)
{
    if (out.prof) {
        const unsigned long long dt = (unsigned long long)((uint64_t)wall_clock64() - t0);
        atomicAdd(&out.prof[phase * kProfSlots + (slot % kProfSlots)], dt);
        if (phase < 3) atomicAdd(&out.prof[kProfCuPhase * kProfSlots + cu_slot()], dt);
    }
}

// Flushes per-lane accumulators: one atomic per wave and counter, into the
// wave's stripe.
__device__ __forceinline__ void flush_acc(const KOut &out, uint64_t cnt, uint64_t hs, uint64_t hx)
{
    cnt = wave_sum_u64(cnt);
    hs = wave_sum_u64(hs);
    hx = wave_xor_u64(hx);
    if ((threadIdx.x & 63) == 0) {
        const uint32_t wave = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
        unsigned long long *a = out.acc + (uint64_t)(wave % kStripes) * kLine;
        if (cnt) atomicAdd(&a[0], (unsigned long long)cnt);
        if (hs) atomicAdd(&a[1], (unsigned long long)hs);
        if (hx) atomicXor(&a[2], (unsigned long long)hx);
    }
}

// Claims up to `step` indices of [0, n) from a striped cursor (kStripes
// counters kLine words apart).  Part t of [0, n) holds the indices t,
// t + kStripes, t + 2 kStripes, ..., so every part walks the work in its
// order (heaviest first where the work is sorted so).  A claimant starts at
// part `home` and moves on once a part is exhausted; on entering a part, a
// plain load usually tells an exhausted one without an atomic.  k (start at
// 0) holds the parts left behind and whether the current part has been
// claimed from.  On success the claimed indices are first + q kStripes for
// q < count; returns false when every part is exhausted.
constexpr uint32_t kPartEntered = 1u << 16;
__device__ __forceinline__ bool claim_striped(unsigned long long *cur, uint64_t n, uint32_t home, uint32_t &k,
                                              uint64_t step, uint64_t &first, uint64_t &count)
{
    while ((k & 0xffffu) < (uint32_t)kStripes) {
        const uint32_t t = (home + (k & 0xffffu)) % kStripes;
        const uint64_t len = t < n ? (n - t + kStripes - 1) / kStripes : 0;
        unsigned long long *p = cur + (uint64_t)t * kLine;
        if (!(k & kPartEntered) && __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len) {
            k = (k & 0xffffu) + 1;
            continue;
        }
        k |= kPartEntered;
        const uint64_t o = atomicAdd(p, (unsigned long long)step);
        if (o < len) {
            first = t + o * kStripes;
            count = o + step < len ? step : len - o;
            return true;
        }
        k = (k & 0xffffu) + 1;
    }
    return false;
}

// Emits the tuples (u, A[x_1], ..., A[x_{l-1}], A[i]) for every bit i of word
// w of the last level: hash accumulation and optional buffer append.
template <class XGet>
__device__ __forceinline__ void emit_word(const KOut &out, const KPlan &pl, int l, uint32_t u,
                                          const uint32_t *A, uint64_t t, uint32_t w, XGet xget,
                                          uint64_t &hs, uint64_t &hx)
{
    const int len = l + 1;
    uint64_t h = 0x6a09e667f3bcc909ull ^ (uint64_t)len;
    h = hash_step(h, 0, u);
    for (int q = 1; q < l; q++) h = hash_step(h, q, A[xget(q)]);
    uint64_t base = 0;
    if (out.buf) base = atomicAdd(&out.acc[3], (unsigned long long)__popcll(t));
    uint64_t idx = 0;
    while (t) {
        const uint32_t b = (uint32_t)__builtin_ctzll(t);
        t &= t - 1;
        const uint32_t vi = A[(w << 6) + b];
        const uint64_t hv = hash_step(h, l, vi);
        hs += hv;
        hx ^= hv;
        if (out.buf && base + idx < out.cap) {
            uint32_t *dst = out.buf + (base + idx) * (uint64_t)len;
            dst[0] = u;
            for (int q = 1; q < l; q++) dst[q] = A[xget(q)];
            dst[l] = vi;
        }
        idx++;
    }
}

__device__ __forceinline__ bool use_search(uint32_t hlen, uint32_t d, uint32_t ratio)
{
    return (uint64_t)hlen > (uint64_t)d * ratio;
}

// ------------------------------------------------------------ LDS batches

struct LdsLayout {
    uint32_t oE, oHBeg, oStk, oA, oPre, oHKey, oRoot, oHVal, oX, oJRoot, oMisc, bytes;
    uint32_t ewords, nlev;
};

__host__ __device__ inline uint32_t align_up(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }

// Layout of lane class lg's batches (the widest class serves every batch);
// nlev + 3 >= L positions of x are kept per thread.
__host__ inline LdsLayout make_lds_layout(int lg, uint32_t nlev)
{
    const uint32_t sd = LdsBudget::sumdeg(lg), hs = LdsBudget::hash(lg), ew = LdsBudget::ewords(lg);
    LdsLayout L{};
    uint32_t o = 0;
    L.ewords = ew;
    L.nlev = nlev;
    L.oE = o;       o += ew * 8;
    L.oHBeg = o;    o += sd * 8;
    L.oStk = o;     o += nlev * kBlock * 8;
    L.oRoot = o;    o += LdsBudget::kMaxRoots * (uint32_t)sizeof(RootRec);
    L.oA = o;       o += sd * 4;
    L.oPre = o;     o += align_up((sd + 1) * 4, 8);
    L.oHKey = o;    o += hs * 4;
    L.oMisc = o;    o += 32 * 4;
    L.oHVal = o;    o += hs * 2;
    L.oX = o;       o += (nlev + 3) * kBlock * 2;
    L.oJRoot = o;   o += sd;
    L.bytes = align_up(o, 16);
    return L;
}

// Misc slots
enum { M_BATCH = 0, M_NEXT = 1, M_QM = 4, M_SM = 14 };

// Block-wide exclusive scan of a[0..n) in place (n <= 4 * kBlock); a[n] = total.
__device__ void block_excl_scan(uint32_t *a, uint32_t n, uint32_t *tmp4)
{
    const int tid = threadIdx.x;
    const int per = (int)((n + kBlock - 1) / kBlock);
    const uint32_t b = (uint32_t)(tid * per);
    uint32_t loc = 0;
    for (int i = 0; i < per; i++)
        if (b + i < n) loc += a[b + i];
    uint32_t inc = wave_incl_scan(loc);
    const int wv = tid >> 6;
    if ((tid & 63) == 63) tmp4[wv] = inc;
    __syncthreads();
    uint32_t wofs = 0;
    for (int i = 0; i < wv; i++) wofs += tmp4[i];
    uint32_t run = wofs + inc - loc;
    for (int i = 0; i < per; i++)
        if (b + i < n) {
            uint32_t x = a[b + i];
            a[b + i] = run;
            run += x;
        }
    if (tid == kBlock - 1) a[n] = run;
    __syncthreads();
}

// LDS carve-up of one workgroup.
struct LdsView {
    uint64_t *sE, *sHBeg, *sStk;
    RootRec *sRoot;
    uint32_t *sA, *sPre, *sHKey, *sMisc;
    uint16_t *sHVal, *sX;
    uint8_t *sJRoot;
};

__device__ __forceinline__ LdsView lds_view(unsigned char *smem, const LdsLayout &lay)
{
    LdsView v;
    v.sE = (uint64_t *)(smem + lay.oE);
    v.sHBeg = (uint64_t *)(smem + lay.oHBeg);
    v.sStk = (uint64_t *)(smem + lay.oStk);
    v.sRoot = (RootRec *)(smem + lay.oRoot);
    v.sA = (uint32_t *)(smem + lay.oA);
    v.sPre = (uint32_t *)(smem + lay.oPre);
    v.sHKey = (uint32_t *)(smem + lay.oHKey);
    v.sMisc = (uint32_t *)(smem + lay.oMisc);
    v.sHVal = (uint16_t *)(smem + lay.oHVal);
    v.sX = (uint16_t *)(smem + lay.oX);
    v.sJRoot = (uint8_t *)(smem + lay.oJRoot);
    return v;
}

// Processes one LDS batch: stage rows, hash and bitmaps, then the DFS over
// its (root, x_1) items.  PC > 0 fixes the group width at compile time; PC == 0
// takes it from Prt.  The plan's level masks are in sMisc[M_QM..], [M_SM..].
template <int PC>
__device__ void process_batch(const KGraph &g, const KPlan &pl, const LdsView &v,
                              const RootRec *__restrict__ recs, uint32_t first, uint32_t nr,
                              const KOut &out, uint64_t dl, int Prt, uint64_t &cnt, uint64_t &hs,
                              uint64_t &hx)
{
    const int P = PC > 0 ? PC : Prt;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int lx = lane & (P - 1);
    const int pbase = lane & ~(P - 1);
    const int L = pl.L;
    uint64_t *sE = v.sE;
    uint64_t *sHBeg = v.sHBeg;
    uint64_t *sStk = v.sStk;
    RootRec *sRoot = v.sRoot;
    uint32_t *sA = v.sA, *sPre = v.sPre, *sHKey = v.sHKey, *sMisc = v.sMisc;
    uint16_t *sHVal = v.sHVal, *sX = v.sX;
    uint8_t *sJRoot = v.sJRoot;
    auto xget = [&](int q) -> uint32_t { return sX[q * kBlock + tid]; };
    // bitmaps are built only when some level intersects with a bitmap row
    bool needE = false;
    for (int l = 2; l < L; l++) needE |= sMisc[M_QM + l] != 0u;
    const uint64_t t_build = wall_clock64();

    if (tid < (int)nr) sRoot[tid] = recs[first + tid];
    __syncthreads();
    const RootRec last = sRoot[nr - 1];
    const uint32_t sumd = last.doff + last.d;
    const uint32_t ew = last.eoff + last.d * ((last.d + 63) >> 6);
    const uint32_t hsl = last.hoff + hash_slots_dev(last.d);
    if (needE) {
        for (uint32_t i = tid; i < ew; i += kBlock) sE[i] = 0;
        for (uint32_t i = tid; i < hsl; i += kBlock) sHKey[i] = 0;
    }
    if (tid == 0) sMisc[M_NEXT] = 0;
    __syncthreads();

    // rows, hash tables and "higher" ranges of every row vertex.  A thread
    // holds at most kRowsPer rows; their row entries, then their vertex
    // records, are loaded together.
    constexpr int kRowsPer = (int)(LdsBudget::kMaxSumDeg / kBlock);
    static_assert(LdsBudget::kMaxSumDeg % kBlock == 0, "rows per thread");
    uint32_t rvx[kRowsPer], rlo[kRowsPer];
#pragma unroll
    for (int q = 0; q < kRowsPer; q++) {
        const uint32_t j = (uint32_t)tid + (uint32_t)q * kBlock;
        rvx[q] = 0;
        rlo[q] = 0;
        if (j < sumd) {
            uint32_t lo = 0, hi = nr;
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (sRoot[mid].doff <= j) lo = mid; else hi = mid;
            }
            rlo[q] = lo;
            const RootRec rr = sRoot[lo];
            rvx[q] = g.ocol[rr.rbeg + (j - rr.doff)];
        }
    }
    VInfo rvi[kRowsPer];
    if (needE) {
#pragma unroll
        for (int q = 0; q < kRowsPer; q++)
            if ((uint32_t)tid + (uint32_t)q * kBlock < sumd) rvi[q] = g.vinfo[rvx[q]];
    }
#pragma unroll
    for (int q = 0; q < kRowsPer; q++) {
        const uint32_t j = (uint32_t)tid + (uint32_t)q * kBlock;
        if (j >= sumd) break;
        const RootRec rr = sRoot[rlo[q]];
        const uint32_t jl = j - rr.doff;
        const uint32_t vx = rvx[q];
        sA[j] = vx;
        sJRoot[j] = (uint8_t)rlo[q];
        if (!needE) continue;
        const uint32_t H = hash_slots_dev(rr.d);
        uint32_t s = hash_slot(vx, H);
        for (;;) {
            if (atomicCAS(&sHKey[rr.hoff + s], 0u, vx + 1u) == 0u) {
                sHVal[rr.hoff + s] = (uint16_t)jl;
                break;
            }
            s = (s + 1) & (H - 1);
        }
        const VInfo vi = rvi[q];
        sHBeg[j] = vi.hbeg | ((uint64_t)vi.hlen << kRangeShift);
        // work units: the range's entries, or the root's neighbours on the
        // larger-key side of the row (later in ascending rows, earlier in
        // descending ones) to look up in the range's table
        sPre[j] = use_search(vi.hlen, rr.d, pl.bin_ratio) ? (g.asc ? rr.d - 1 - jl : jl) : vi.hlen;
    }
    __syncthreads();
    if (needE) {
        block_excl_scan(sPre, sumd, sMisc + 28);
        const uint32_t total = sPre[sumd];
        // bitmap: each adjacent pair (v, w), key(v) < key(w), is found from v.
        // Every thread takes kChunk consecutive work units: one search for the
        // row of the first, then the rows advance with the units.  The units'
        // entries are loaded in a first pass, back to back, and probed in a
        // second, so a thread waits once for its kChunk loads.
        constexpr uint32_t kChunk = 8;
        for (uint32_t e0 = tid * kChunk; e0 < total; e0 += kBlock * kChunk) {
            uint32_t lo = 0, hi = sumd;  // largest j with sPre[j] <= e0
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (sPre[mid] <= e0) lo = mid; else hi = mid;
            }
            const uint32_t jfirst = lo;
            const uint32_t e1 = e0 + kChunk < total ? e0 + kChunk : total;
            // per-row context, refreshed when the units cross into the next row
            uint32_t j, jn;  // row, first unit of the next row (sPre[sumd] = total)
            RootRec rr;
            uint32_t j0, jl, W, H, hl;
            uint64_t hb;
            bool srch;
            auto load_row = [&]() {
                rr = sRoot[sJRoot[j]];
                j0 = sPre[j];
                jl = j - rr.doff;
                W = (rr.d + 63) >> 6;
                H = hash_slots_dev(rr.d);
                const uint64_t hv = sHBeg[j];
                hb = hv & ((1ull << kRangeShift) - 1ull);
                hl = (uint32_t)(hv >> kRangeShift);
                srch = use_search(hl, rr.d, pl.bin_ratio);
            };
            auto restart = [&]() {
                j = jfirst;
                jn = sPre[j + 1];
                load_row();
            };
            auto seek = [&](uint32_t e) {
                if (e >= jn) {
                    do {
                        j++;
                        jn = sPre[j + 1];
                    } while (e >= jn);
                    load_row();
                }
            };
            uint32_t ent[kChunk];
            restart();
#pragma unroll
            for (uint32_t k = 0; k < kChunk; k++) {
                const uint32_t e = e0 + k;
                ent[k] = 0;
                if (e < e1) {
                    seek(e);
                    if (!srch) ent[k] = g.ocol[hb + (e - j0)];
                }
            }
            restart();
#pragma unroll
            for (uint32_t k = 0; k < kChunk; k++) {
                const uint32_t e = e0 + k;
                if (e >= e1) break;
                seek(e);
                const uint32_t off = e - j0;
                if (srch) {
                    // neighbour i of the root on the larger-key side of row j:
                    // is A[i] in A[j]'s larger-key range?  (its lookup table)
                    const uint32_t i = g.asc ? jl + 1 + off : off;
                    const uint32_t a = sA[rr.doff + i] + 1u, size = 2 * hl;
                    const uint32_t *t = g.htab + 2 * hb;
                    uint32_t s = lookup_slot(a - 1u, size);
                    for (;;) {
                        const uint32_t x = t[s];
                        if (x == a) {
                            set_pair(sE + rr.eoff, W, jl, i, pl.upper);
                            break;
                        }
                        if (x == 0u) break;
                        s = s + 1 == size ? 0 : s + 1;
                    }
                    continue;
                }
                const uint32_t w = ent[k];
                uint32_t s = hash_slot(w, H);
                for (;;) {
                    const uint32_t key = sHKey[rr.hoff + s];
                    if (key == 0u) break;
                    if (key == w + 1u) {  // a row holds each vertex once
                        set_pair(sE + rr.eoff, W, jl, sHVal[rr.hoff + s], pl.upper);
                        break;
                    }
                    s = (s + 1) & (H - 1);
                }
            }
        }
        __syncthreads();
    }

    if (tid == 0) prof_add(out, 3, blockIdx.x, t_build);  // staging + bitmap build

    // DFS over (root, x_1) items
    uint32_t spins = 0;
    for (;;) {
        uint32_t item = 0;
        if (lx == 0) {
            if (past(dl)) {
                item = 0xffffffffu;
                mark_stop(out);
            } else {
                item = atomicAdd(&sMisc[M_NEXT], 1u);
            }
        }
        item = __shfl(item, 0, P);
        if (item >= sumd) break;
        const RootRec rr = sRoot[sJRoot[item]];
        const uint32_t jl = item - rr.doff;
        if (pl.skip_split && jl < rr.split) continue;
        if (L < 3) continue;
        const uint32_t d = rr.d, W = (d + 63) >> 6;
        const uint64_t *E = sE + rr.eoff;
        const uint32_t *A = sA + rr.doff;
        const uint32_t w = (uint32_t)lx;
        const bool own = w < W;
        sX[1 * kBlock + tid] = (uint16_t)jl;

        auto level_word = [&](int l) -> uint64_t {  // masks first; no row reads once empty
            uint64_t t = valid_mask(w, d);
            for (int s = 1; s < l; s++) t &= ~bit_at(xget(s), w);
            uint32_t sm = sMisc[M_SM + l];
            while (sm) {
                const int s = __builtin_ctz(sm);
                sm &= sm - 1;
                t &= gt_mask(xget(s), w);
            }
            uint32_t qm = sMisc[M_QM + l];
            while (qm && t) {
                const int q = __builtin_ctz(qm);
                qm &= qm - 1;
                t &= E[xget(q) * W + w];
            }
            return t;
        };

        // Summed last level (pl.tail): F = the part of C_{L-1} fixed by
        // x_1..x_{L-3}; for x_{L-2} = x, |C_{L-1}| = |F n (x, inf)| when
        // L-2 is in smask[L-1], else |F \ {x}|.  Summed over x in S = C_{L-2}:
        //   sum_w [ |S_w| (|F| - |F_{<=w}|) + pairs_w(S, F) ]  or  |S||F| - |S n F|
        // with one group prefix sum of |F_w|.
        auto tail_sum = [&](int l) -> uint64_t {
            const int lv = l - 1;
            const uint64_t S = own ? sStk[(l - 3) * kBlock + tid] : 0ull;
            uint64_t F = 0;
            if (own) {
                F = valid_mask(w, d);
                uint32_t qm = sMisc[M_QM + l];
                while (qm) {
                    const int q = __builtin_ctz(qm);
                    qm &= qm - 1;
                    F &= E[xget(q) * W + w];
                }
                for (int s = 1; s < lv; s++) F &= ~bit_at(xget(s), w);
                uint32_t sm = sMisc[M_SM + l] & ~(1u << lv);
                while (sm) {
                    const int s = __builtin_ctz(sm);
                    sm &= sm - 1;
                    F &= gt_mask(xget(s), w);
                }
            }
            const uint32_t sc = (uint32_t)__popcll(S);
            uint32_t fin = (uint32_t)__popcll(F);
            for (int o = 1; o < P; o <<= 1) {
                const uint32_t y = __shfl_up(fin, o, P);
                if (lx >= o) fin += y;
            }
            const uint32_t ftot = __shfl(fin, P - 1, P);
            if ((sMisc[M_SM + l] >> lv) & 1u) return (uint64_t)sc * (ftot - fin) + pairs_in_word(S, F);
            return (uint64_t)sc * ftot - (uint64_t)__popcll(S & F);
        };

        // Paired last levels (pl.pair), l = L-2 = b: S = C_{l-1} = C_a with
        // x_1..x_{a-1} fixed; adds C(|S n E[x_b]|, 2) over x_b in B, the
        // candidates of level b without its constraints on x_a.  Only an x_b
        // adjacent to some member of S can add anything, so B is first cut to
        // the union of their rows.
        auto pair_sum = [&](int l) -> uint64_t {
            const uint64_t S = own ? sStk[(l - 3) * kBlock + tid] : 0ull;
            uint64_t U = 0;
            for (uint64_t rem = S;;) {
                const uint64_t bal = __ballot(rem != 0ull);
                const uint64_t mine = (P == 64) ? bal : ((bal >> pbase) & ((1ull << P) - 1ull));
                if (mine == 0ull) break;
                const int f = __builtin_ctzll(mine);
                const uint64_t vf = __shfl(rem, f, P);
                const uint32_t xs = ((uint32_t)f << 6) + (uint32_t)__builtin_ctzll(vf);
                if (lx == f) rem &= rem - 1ull;
                if (own) U |= E[xs * W + w];
            }
            uint64_t B = 0;
            if (own) {
                B = valid_mask(w, d) & U;
                uint32_t qm = sMisc[M_QM + l] & ~(1u << (l - 1));
                while (qm) {
                    const int q = __builtin_ctz(qm);
                    qm &= qm - 1;
                    B &= E[xget(q) * W + w];
                }
                for (int s = 1; s < l - 1; s++) B &= ~bit_at(xget(s), w);
                uint32_t sm = sMisc[M_SM + l];
                while (sm) {
                    const int s = __builtin_ctz(sm);
                    sm &= sm - 1;
                    B &= gt_mask(xget(s), w);
                }
            }
            uint64_t sum = 0;
            for (;;) {
                const uint64_t bal = __ballot(B != 0ull);
                const uint64_t mine = (P == 64) ? bal : ((bal >> pbase) & ((1ull << P) - 1ull));
                if (mine == 0ull) break;
                const int f = __builtin_ctzll(mine);
                const uint64_t vf = __shfl(B, f, P);
                const uint32_t xb = ((uint32_t)f << 6) + (uint32_t)__builtin_ctzll(vf);
                if (lx == f) B &= B - 1ull;  // lane f's lowest bit is x_b
                uint32_t pc = own ? (uint32_t)__popcll(S & E[xb * W + w]) : 0u;
                for (int o = 1; o < P; o <<= 1) pc += __shfl_xor(pc, o, P);
                if (lx == 0) sum += (uint64_t)pc * (pc - 1u) / 2u;
            }
            return sum;
        };

        const uint64_t t2 = own ? level_word(2) : 0ull;
        if (L == 3) {
            cnt += (uint64_t)__popcll(t2);
            if (pl.emit && t2) emit_word(out, pl, 2, rr.u, A, t2, w, xget, hs, hx);
            continue;
        }
        sStk[tid] = t2;
        int l = 3;
        uint32_t pos = 0;
        for (;;) {
            if (++spins == kClockEvery) {
                spins = 0;
                if (past(dl)) break;  // the next claim records the stop
            }
            if (pl.tail && l == L - 1) {  // every x_{L-2} of C_{L-2} at once
                cnt += tail_sum(l);
                if (l == 3) break;
                --l;
                pos = (uint32_t)xget(l - 1) + 1u;
                continue;
            }
            if (pl.pair && l == L - 2) {  // levels L-3..L-1 at once
                cnt += pair_sum(l);
                if (l == 3) break;
                --l;
                pos = (uint32_t)xget(l - 1) + 1u;
                continue;
            }
            uint64_t vv = own ? sStk[(l - 3) * kBlock + tid] : 0ull;
            const uint32_t pw = pos >> 6;
            if (w < pw) vv = 0ull;
            else if (w == pw) vv &= (~0ull << (pos & 63u));
            const uint64_t bal = __ballot(vv != 0ull);
            const uint64_t mine = (P == 64) ? bal : ((bal >> pbase) & ((1ull << P) - 1ull));
            if (mine == 0ull) {
                if (l == 3) break;
                --l;
                pos = (uint32_t)xget(l - 1) + 1u;
                continue;
            }
            const int f = __builtin_ctzll(mine);
            const uint64_t vf = __shfl(vv, f, P);
            const uint32_t xi = ((uint32_t)f << 6) + (uint32_t)__builtin_ctzll(vf);
            sX[(l - 1) * kBlock + tid] = (uint16_t)xi;
            const uint64_t t = own ? level_word(l) : 0ull;
            if (l == L - 1) {
                cnt += (uint64_t)__popcll(t);
                if (pl.emit && t) emit_word(out, pl, l, rr.u, A, t, w, xget, hs, hx);
                pos = xi + 1u;
            } else {
                sStk[(l - 2) * kBlock + tid] = t;
                ++l;
                pos = 0;
            }
        }
    }
    __syncthreads();
}

// MinWaves > 1: the compiler keeps the registers to that many waves per SIMD
// (7 workgroups per CU instead of 5 for the narrow classes, at the cost of a
// few bytes of spill); used when no hub kernels of the query share the CUs.
template <int P, int MinWaves>
__global__ __launch_bounds__(kBlock, MinWaves) void lds_batch_kernel(KGraph g, KPlan pl,
                                                           const RootRec *__restrict__ recs,
                                                           const uint32_t *__restrict__ batch_first,
                                                           const uint32_t *__restrict__ batch_count,
                                                           uint32_t b0, uint32_t b1,
                                                           unsigned long long *cursor, LdsLayout lay,
                                                           KOut out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const LdsView v = lds_view(smem, lay);
    const int tid = threadIdx.x;
    if (tid < kMaxLevels) {
        v.sMisc[M_QM + tid] = pl.qmask[tid];
        v.sMisc[M_SM + tid] = pl.smask[tid];
    }
    uint64_t cnt = 0, hs = 0, hx = 0;
    const uint64_t dl = out.dl[1];
    const uint64_t t_start = wall_clock64();
    uint32_t parts = 0;  // thread 0's claim state
    for (;;) {
        if (tid == 0) {
            uint64_t bb = 0, be = 0;
            if (past(dl)) {
                v.sMisc[M_BATCH] = b1;
                mark_stop(out);
            } else if (claim_striped(cursor, b1 - b0, blockIdx.x % kStripes, parts, 1, bb, be)) {
                v.sMisc[M_BATCH] = b0 + (uint32_t)bb;
            } else {
                v.sMisc[M_BATCH] = b1;
            }
        }
        __syncthreads();
        const uint32_t b = v.sMisc[M_BATCH];
        if (b >= b1) break;
        process_batch<P>(g, pl, v, recs, batch_first[b], batch_count[b], out, dl, P, cnt, hs, hx);
    }
    flush_acc(out, cnt, hs, hx);
    if (tid == 0) prof_add(out, 0, blockIdx.x, t_start);
}

// One persistent launch over the LDS batches of several prepared queries.
// Work records are ordered heaviest first; a workgroup takes the next record,
// loads that query's plan and graph and runs the batch with the record's
// group width, then flushes its counts to that query's counters.
__global__ __launch_bounds__(kBlock) void lds_multi_kernel(const MultiQuery *__restrict__ qs,
                                                           const MultiWork *__restrict__ work,
                                                           uint32_t nwork, unsigned long long *cursor,
                                                           LdsLayout lay)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const LdsView v = lds_view(smem, lay);
    const int tid = threadIdx.x;
    uint32_t parts = 0;  // thread 0's claim state
    for (;;) {
        if (tid == 0) {
            uint64_t bb = 0, be = 0;
            v.sMisc[M_BATCH] = claim_striped(cursor, nwork, blockIdx.x % kStripes, parts, 1, bb, be)
                                   ? (uint32_t)bb : nwork;
        }
        __syncthreads();
        const uint32_t wi = v.sMisc[M_BATCH];
        if (wi >= nwork) break;
        const MultiWork w = work[wi];
        const MultiQuery &q = qs[w.query];
        const KPlan pl = q.pl;
        if (tid < kMaxLevels) {
            v.sMisc[M_QM + tid] = pl.qmask[tid];
            v.sMisc[M_SM + tid] = pl.smask[tid];
        }
        __syncthreads();
        uint64_t cnt = 0, hs = 0, hx = 0;
        const uint64_t dl = q.out.dl[1];
        const uint64_t t_start = wall_clock64();
        process_batch<0>(q.g, pl, v, q.recs, q.batch_first[w.batch], q.batch_count[w.batch], q.out, dl,
                         1 << w.lanes_log2, cnt, hs, hx);
        flush_acc(q.out, cnt, hs, hx);
        if (tid == 0) prof_add(q.out, 0, blockIdx.x, t_start);
    }
}

// ------------------------------------------------------------------- hubs

// Hash of every hub row: vertex -> row index (linear probing, 0 = empty).
__global__ __launch_bounds__(kBlock) void hub_hash_kernel(KGraph g, const HubRec *__restrict__ hubs,
                                                          const uint32_t *__restrict__ hub_of_row,
                                                          uint64_t g0, uint64_t g1, uint64_t *kv,
                                                          const unsigned long long *dl)
{
    if (past(*dl)) return;  // the encode and the DFS that follow stop at once
    for (uint64_t r = g0 + blockIdx.x * (uint64_t)kBlock + threadIdx.x; r < g1;
         r += (uint64_t)gridDim.x * kBlock) {
        const HubRec h = hubs[hub_of_row[r]];
        const uint32_t jl = (uint32_t)(r - h.joff);
        if (jl < h.rlo) continue;  // rows below rlo are never found by a scan range
        const uint32_t v = g.ocol[h.rbeg + jl];
        uint32_t s = hash_slot(v, h.hslots);
        for (;;) {
            // slot = key (low word) | row index (high word): a hit costs one load
            uint32_t *slot = (uint32_t *)&kv[h.hoff + s];
            if (atomicCAS(&slot[0], 0u, v + 1u) == 0u) {
                slot[1] = jl;
                break;
            }
            s = (s + 1) & (h.hslots - 1);
        }
    }
}

// Hub bitmaps: waves take chunks of 64 rows (of the round's rows whose scan
// range is not empty, listed on the host), scan the lengths of their scan
// ranges (the larger-key range of the row vertex, cut on the host to the part
// that can hold kept columns) and stream the concatenation lane-parallel.  A
// chunk takes every nchunks-th row, so the long ranges of one hub's rows
// spread over many waves.
// Each lane loads its kChunk entries back to back before probing the hash.
__global__ __launch_bounds__(kBlock) void hub_encode_kernel(KGraph g, const HubRec *__restrict__ hubs,
                                                            const uint32_t *__restrict__ hub_of_row,
                                                            const uint64_t *__restrict__ hub_range,
                                                            const uint64_t *__restrict__ rows, uint64_t nrows,
                                                            const uint64_t *__restrict__ kv,
                                                            uint64_t *E, int upper, KOut out)
{
    __shared__ uint64_t sBeg[kBlock];
    __shared__ uint32_t sPre[kBlock];
    __shared__ uint32_t sHub[kBlock];
    __shared__ uint32_t sJl[kBlock];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t nchunks = (nrows + 63) / 64;
    const uint64_t wstride = (uint64_t)gridDim.x * (kBlock / 64);
    const uint64_t dl = *out.dl;
    for (uint64_t c = blockIdx.x * (uint64_t)(kBlock / 64) + wv; c < nchunks; c += wstride) {
        if (past(dl)) {  // the bitmap stays incomplete; the DFS that follows stops at once
            if (lane == 0) mark_stop(out);
            break;
        }
        const uint64_t x = (uint64_t)lane * nchunks + c;  // rows strided over the chunks
        uint32_t len = 0;
        if (x < nrows) {
            const uint64_t r = rows[x];
            const uint32_t hi = hub_of_row[r];
            const uint64_t rg = hub_range[r];
            sBeg[tid] = rg & ((1ull << kRangeShift) - 1ull);
            len = (uint32_t)(rg >> kRangeShift);
            sHub[tid] = hi;
            sJl[tid] = (uint32_t)(r - hubs[hi].joff);
        }
        const uint32_t inc = wave_incl_scan(len);
        sPre[tid] = inc;
        const uint32_t total = __shfl(inc, 63, 64);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int wb = wv * 64;
        // every lane takes kChunk consecutive elements: one search for the row
        // of the first, then the rows advance with the elements
        constexpr uint32_t kChunk = 8;
        for (uint32_t e0 = lane * kChunk; e0 < total; e0 += 64 * kChunk) {
            int kfirst = 0;
            {
                int hi = 63;  // smallest k with pre[k] > e0
                while (kfirst < hi) {
                    const int mid = (kfirst + hi) >> 1;
                    if (sPre[wb + mid] > e0) hi = mid; else kfirst = mid + 1;
                }
            }
            int k = 0;
            uint32_t kbeg = 0, kend = 0;  // elements of row k: [kbeg, kend)
            HubRec h;
            uint32_t jl = 0;
            uint64_t W = 0, beg = 0;
            auto load_row = [&]() {
                h = hubs[sHub[wb + k]];
                jl = sJl[wb + k];
                W = (h.d + 63) >> 6;
                beg = sBeg[wb + k];
            };
            auto restart = [&]() {
                k = kfirst;
                kend = sPre[wb + k];
                kbeg = k ? sPre[wb + k - 1] : 0u;
                load_row();
            };
            auto seek = [&](uint32_t e) {
                if (e >= kend) {
                    do {
                        k++;
                        kbeg = kend;
                        kend = sPre[wb + k];
                    } while (e >= kend);
                    load_row();
                }
            };
            const uint32_t e1 = e0 + kChunk < total ? e0 + kChunk : total;
            uint32_t ent[kChunk];
            restart();
#pragma unroll
            for (uint32_t q = 0; q < kChunk; q++) {
                const uint32_t e = e0 + q;
                ent[q] = 0;
                if (e < e1) {
                    seek(e);
                    ent[q] = g.ocol[beg + (e - kbeg)];
                }
            }
            restart();
#pragma unroll
            for (uint32_t q = 0; q < kChunk; q++) {
                const uint32_t e = e0 + q;
                if (e >= e1) break;
                seek(e);
                const uint32_t wv2 = ent[q];
                uint32_t s = hash_slot(wv2, h.hslots);
                uint64_t slot = kv[h.hoff + s];
                for (;;) {
                    const uint32_t key = (uint32_t)slot;
                    if (key == 0u) break;
                    if (key == wv2 + 1u) {
                        const uint32_t i = (uint32_t)(slot >> 32);
                        // upper plans keep the bit in the row of the smaller index only
                        const bool fwd = !upper || i > jl, mir = !upper || jl > i;
                        if (fwd && jl >= h.rlo)
                            atomicOr((unsigned long long *)&E[h.eoff + (uint64_t)(jl - h.rlo) * W + (i >> 6)],
                                     1ull << (i & 63));
                        if (mir && i >= h.rlo)
                            atomicOr((unsigned long long *)&E[h.eoff + (uint64_t)(i - h.rlo) * W + (jl >> 6)],
                                     1ull << (jl & 63));
                        break;  // a hub row holds each vertex once
                    }
                    s = (s + 1) & (h.hslots - 1);
                    slot = kv[h.hoff + s];
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Copies a W-word row (HBM) into a wave's buffer, lane i taking words i,
// i + 64, ...; four loads per lane are in flight at a time.
__device__ __forceinline__ void copy_row(uint64_t *dst, const uint64_t *__restrict__ src, uint32_t W, int lane)
{
    for (uint32_t w0 = 0; w0 < W; w0 += 4 * 64) {
        uint64_t v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t w = w0 + k * 64 + lane;
            v[k] = w < W ? src[w] : 0ull;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t w = w0 + k * 64 + lane;
            if (w < W) dst[w] = v[k];
        }
    }
}

// k-th (0-based) set bit of w
__device__ __forceinline__ uint32_t select_bit(uint64_t w, uint32_t k)
{
    for (uint32_t i = 0; i < k; i++) w &= w - 1;
    return (uint32_t)__builtin_ctzll(w);
}

// Hub DFS: one wave per (hub, x_1) item; lane i owns words i, i+64, ...
// The candidate stack (nlev x wmax words per wave) and, when the plan reads
// it, the bitmap row of x_1 live in LDS (LDSBUF) or, for very wide hubs, in a
// per-wave HBM slot.  Words below the largest symmetry lower bound of a level
// are neither computed nor scanned.
//
// mode 0 walks the items.  An item whose level-2 candidate set holds more
// than hq.split_min candidates is donated instead of walked: C_2 is copied to
// a donation slot and cut into chunks of hq.chunk consecutive candidates,
// each a (hub, x_1, slot, [lo, hi)) record appended to the queue with one
// atomic.  mode 1 (the next launch on the stream) drains the queue: every
// wave of the grid takes chunks and walks the x_2 in [lo, hi) of its chunk.
// PAIR compiles in the paired last levels; a 5-vertex pair plan's item is one
// paired pass, so its records hold ranges [lo, hi) of the pass's positions
// of B (whole rounds of 64 words) instead of x_2 ranges.
template <bool LDSBUF, bool PAIR>
__global__ __launch_bounds__(kBlock, PAIR ? 3 : 1) void hub_dfs_kernel(KGraph g, KPlan pl,
                                                         const HubRec *__restrict__ hubs,
                                                         const uint32_t *__restrict__ hub_of_item,
                                                         uint64_t i0, uint64_t i1,
                                                         unsigned long long *cursor,
                                                         const uint64_t *__restrict__ E,
                                                         uint64_t *stk, uint32_t wmax, KOut out,
                                                         HubQueue hq, int mode)
{
    extern __shared__ __align__(16) uint64_t hsm[];
    __shared__ uint32_t sX[kBlock / 64][kMaxLevels];
    __shared__ uint32_t sWlo[kBlock / 64][kMaxLevels];
    __shared__ uint32_t sQM[kMaxLevels], sSM[kMaxLevels];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid < kMaxLevels) {
        sQM[tid] = pl.qmask[tid];
        sSM[tid] = pl.smask[tid];
    }
    __syncthreads();
    const int L = pl.L;
    const int nlev = L > 3 ? L - 3 : 1;
    // x_1's bitmap row is cached when some level intersects with it
    bool row1 = false;
    for (int l = 2; l < L; l++) row1 |= (sQM[l] & 2u) != 0;
    uint64_t *my;
    const uint64_t wave_words = hub_wave_words(nlev, wmax);
    if (LDSBUF) my = hsm + (uint64_t)wv * wave_words;
    else my = stk + ((uint64_t)blockIdx.x * (kBlock / 64) + wv) * wave_words;
    uint64_t *e1 = my + (uint64_t)nlev * wmax;
    uint64_t cnt = 0, hs = 0, hx = 0;
    const uint64_t dl = *out.dl;
    const uint64_t t_start = wall_clock64();
    const uint32_t wslot = blockIdx.x * (kBlock / 64) + wv;
    uint32_t spins = 0;
    auto xget = [&](int q) -> uint32_t { return sX[wv][q]; };

    // per-item context
    const uint64_t *Eh = nullptr;
    const uint32_t *A = nullptr;
    uint32_t d = 0, W = 0, R = 0, u = 0, rlo = 0;

    // first word of level l that can hold a candidate
    auto word_lo = [&](int l) -> uint32_t {
        uint32_t lo = 0;
        uint32_t sm = sSM[l];
        while (sm) {
            const int s = __builtin_ctz(sm);
            sm &= sm - 1;
            const uint32_t v = xget(s) >> 6;
            lo = v > lo ? v : lo;
        }
        return lo;
    };
    // The masks go first, then the rows (x_1's from LDS first), and a lane
    // whose word is already empty loads no further row word.
    auto level_word = [&](int l, uint32_t w) -> uint64_t {
        uint64_t t = valid_mask(w, d);
        for (int s = 1; s < l; s++) t &= ~bit_at(xget(s), w);
        uint32_t sm = sSM[l];
        while (sm) {
            const int s = __builtin_ctz(sm);
            sm &= sm - 1;
            t &= gt_mask(xget(s), w);
        }
        uint32_t qm = sQM[l];
        if (row1 && (qm & 2u)) {
            t &= e1[w];
            qm &= ~2u;
        }
        while (qm && t) {
            const int q = __builtin_ctz(qm);
            qm &= qm - 1;
            t &= Eh[(uint64_t)(xget(q) - rlo) * W + w];
        }
        return t;
    };

    // Summed last level (pl.tail), l = L-1: adds |C_l| over every x_{l-1} in
    // S = C_{l-1} n [pos0, hi2).  F = the part of C_l fixed by x_1..x_{l-2};
    // with carry-chained wave prefix sums of |F_w| over the rounds,
    //   sum_x |F n (x, inf)| = |S||F| - sum_w (|S_w| |F_{<=w}| - pairs_w(S, F))
    //   sum_x |F \ {x}|      = |S||F| - |S n F|
    // (words below both S's and F's first word add nothing to either).
    const bool tail_gt = pl.tail && ((sSM[L - 1] >> (L - 2)) & 1u);
    auto tail_count = [&](int l, uint32_t pos0, uint32_t hi2) {
        const int lv = l - 1;
        const uint32_t wl = sWlo[wv][l - 1];
        const uint32_t ws = (pos0 >> 6) > wl ? (pos0 >> 6) : wl;  // first word of S
        uint32_t fl = 0;  // first word where F can be non-zero
        uint32_t fsm = sSM[l] & ~(1u << lv);
        while (fsm) {
            const int s = __builtin_ctz(fsm);
            fsm &= fsm - 1;
            const uint32_t v = xget(s) >> 6;
            fl = v > fl ? v : fl;
        }
        const uint32_t w0 = tail_gt ? ws : (ws < fl ? ws : fl);
        uint64_t acc = 0, sl = 0;
        uint32_t carry = 0;
        for (uint32_t rr = w0 >> 6; rr < R; rr++) {
            const uint32_t w = rr * 64 + lane;
            uint64_t S = 0, F = 0;
            if (w < W && w >= w0) {
                if (w >= ws) S = my[(uint64_t)(l - 3) * wmax + w] & ~lt_mask(pos0, w) & lt_mask(hi2, w);
                if (w >= fl) {
                    F = valid_mask(w, d);
                    uint32_t qm = sQM[l];
                    while (qm && F) {  // lanes whose word is already empty load nothing
                        const int q = __builtin_ctz(qm);
                        qm &= qm - 1;
                        if (q == 1 && row1) F &= e1[w];
                        else F &= Eh[(uint64_t)(xget(q) - rlo) * W + w];
                    }
                    for (int s = 1; s < lv; s++) F &= ~bit_at(xget(s), w);
                    uint32_t sm = sSM[l] & ~(1u << lv);
                    while (sm) {
                        const int s = __builtin_ctz(sm);
                        sm &= sm - 1;
                        F &= gt_mask(xget(s), w);
                    }
                }
            }
            const uint32_t sc = (uint32_t)__popcll(S);
            const uint32_t fin = wave_incl_scan((uint32_t)__popcll(F)) + carry;
            acc += tail_gt ? (uint64_t)sc * fin - pairs_in_word(S, F) : (uint64_t)__popcll(S & F);
            sl += sc;
            carry = __shfl(fin, 63, 64);
        }
        const uint64_t stot = wave_sum_u64(sl);
        if (lane == 0) cnt += stot * (uint64_t)carry;
        cnt -= acc;  // the wave's sum is exact modulo 2^64
    };

    // Paired last levels (pl.pair), l = L-2 = b: S = C_{l-1} in the stack with
    // x_1..x_{l-2} fixed, B = level b without its constraints on x_a = x_{l-1},
    // cut to the positions [plo, phi) (a donated part of an item; words at
    // multiples of 64).  The three levels add sum over x_b in B of
    // C(|S n E[x_b]|, 2).  Returns true when the deadline passed before it was
    // done (a hub of hundreds of thousands of neighbours makes one call long):
    // the item is left with part of its count.
    //
    // Pair plans keep both bits of every adjacent pair (a paired level c
    // intersects with the row of x_b and is bound only above x_a < x_b), so
    // |S n E[p]| = c_p = |{s in S : E[s][p]}|, and the sum is
    //   sum over p in B of C(c_p, 2)
    // over the rows of S alone.  In each round of 64 words a lane keeps the
    // c_p of its word's 64 positions as bit-sliced counters (slice j = bit j
    // of every c_p), adds the members' words with a ripple carry, and at the
    // round's end takes
    //   sum_p C(c_p, 2) = sum_j 2^(j-1) (2^j - 1) |sl_j|
    //                     + sum_{i<j} 2^(i+j) |sl_i & sl_j|.
    // Every member row is read once per round where B can hold candidates.
    // The members come from a list in the next stack row (u32, two rows in
    // flight) when it holds them, else straight from S's words.  An S with more
    // members than the counters hold takes the walk at the end: B restricted
    // to the union of the members' rows, then C(|S n E[x_b]|, 2) for every x_b
    // in it.
    auto pair_count = [&](int l, uint32_t plo, uint32_t phi) -> bool {
        uint32_t steps = 0;
        const uint64_t *S = my + (uint64_t)(l - 3) * wmax;
        uint64_t *Bv = my + (uint64_t)(l - 2) * wmax;
        uint32_t blo = word_lo(l);
        if ((plo >> 6) > blo) blo = plo >> 6;
        const uint32_t bhi = phi >= (W << 6) ? W : (phi >> 6);  // B's words: [blo, bhi)
        const uint32_t slo = sWlo[wv][l - 1];
        if (blo >= bhi) return false;
        // B's mask on word w (x_a's constraints left out)
        auto b_word = [&](uint32_t w) -> uint64_t {
            uint64_t t = valid_mask(w, d);
            uint32_t qm = sQM[l] & ~(1u << (l - 1));
            while (qm && t) {
                const int q = __builtin_ctz(qm);
                qm &= qm - 1;
                if (q == 1 && row1) t &= e1[w];
                else t &= Eh[(uint64_t)(xget(q) - rlo) * W + w];
            }
            for (int s = 1; s < l - 1; s++) t &= ~bit_at(xget(s), w);
            uint32_t sm = sSM[l];
            while (sm) {
                const int s = __builtin_ctz(sm);
                sm &= sm - 1;
                t &= gt_mask(xget(s), w);
            }
            return t;
        };
        {
            uint32_t *list = (uint32_t *)Bv;
            const uint32_t cap = 2u * wmax < kPairMembers ? 2u * wmax : kPairMembers;
            uint32_t ns = 0;
            for (uint32_t rr = slo >> 6; rr < R; rr++) {
                const uint32_t w = rr * 64 + lane;
                uint64_t v = (w < W && w >= slo) ? S[w] : 0ull;
                const uint32_t pc = (uint32_t)__popcll(v);
                const uint32_t incl = wave_incl_scan(pc);
                uint32_t at = ns + incl - pc;
                if (ns + incl <= cap) {
                    while (v) {
                        list[at++] = (w << 6) + (uint32_t)__builtin_ctzll(v);
                        v &= v - 1;
                    }
                }
                ns += __shfl(incl, 63, 64);
            }
            if (ns == 0u) return false;
            if (ns <= kPairMembers) {
                const bool listed = ns <= cap;
                if (listed) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
                const int nb = 32 - __builtin_clz(ns);  // bits of the largest c_p
                for (uint32_t rr = blo >> 6; rr * 64 < bhi; rr++) {
                    if ((++steps & 15u) == 0u && past(dl)) return true;
                    const uint32_t w = rr * 64 + lane;
                    const uint64_t bc = (w < bhi && w >= blo) ? b_word(w) : 0ull;
                    if (__ballot(bc != 0ull) == 0ull) continue;
                    uint64_t sl[kPairSlices];
#pragma unroll
                    for (int j = 0; j < kPairSlices; j++) sl[j] = 0ull;
                    auto add = [&](uint64_t c) {
#pragma unroll
                        for (int j = 0; j < kPairSlices; j++) {
                            if (c) {
                                const uint64_t t = sl[j] & c;
                                sl[j] ^= c;
                                c = t;
                            }
                        }
                    };
                    auto word_of = [&](uint32_t s) -> uint64_t {
                        return bc ? Eh[(uint64_t)(s - rlo) * W + w] & bc : 0ull;
                    };
                    if (listed) {
                        for (uint32_t b0 = 0; b0 < ns; b0 += 64) {
                            const uint32_t mine = b0 + lane < ns ? list[b0 + lane] : 0u;
                            const uint32_t nk = ns - b0 < 64u ? ns - b0 : 64u;
                            uint32_t k = 0;
                            for (; k + 2 <= nk; k += 2) {  // two member words in flight
                                const uint64_t r0 = word_of(__builtin_amdgcn_readlane(mine, k));
                                const uint64_t r1 = word_of(__builtin_amdgcn_readlane(mine, k + 1));
                                add(r0);
                                add(r1);
                            }
                            for (; k < nk; k++) add(word_of(__builtin_amdgcn_readlane(mine, k)));
                        }
                    } else {
                        for (uint32_t sr = slo >> 6; sr < R; sr++) {  // S's words, 64 at a time
                            const uint32_t sw = sr * 64 + lane;
                            uint64_t v = (sw < W && sw >= slo) ? S[sw] : 0ull;
                            for (;;) {
                                const uint64_t bal = __ballot(v != 0ull);
                                if (bal == 0ull) break;
                                const int f = __builtin_ctzll(bal);
                                const uint64_t vf = __shfl(v, f, 64);
                                if (lane == f) v &= v - 1ull;
                                add(word_of(((sr * 64 + (uint32_t)f) << 6) + (uint32_t)__builtin_ctzll(vf)));
                            }
                        }
                    }
                    uint64_t acc = 0;
#pragma unroll
                    for (int j = 1; j < kPairSlices; j++) {
                        if (j < nb) {
                            acc += ((1ull << (j - 1)) * ((1ull << j) - 1ull)) * (uint64_t)__popcll(sl[j]);
#pragma unroll
                            for (int i = 0; i < j; i++) acc += (uint64_t)__popcll(sl[i] & sl[j]) << (i + j);
                        }
                    }
                    cnt += acc;
                }
                return false;
            }
        }
        // only an x_b adjacent to some member of S adds anything: B starts as
        // the union of their rows (words [blo, bhi))
        for (uint32_t rr = blo >> 6; rr < R; rr++) {
            const uint32_t w = rr * 64 + lane;
            if (w < bhi && w >= blo) Bv[w] = 0ull;
        }
        for (uint32_t pos = slo << 6;;) {
            uint32_t xs = 0xffffffffu;
            const uint32_t pw = pos >> 6;
            for (uint32_t rr = pw >> 6; rr < R; rr++) {
                const uint32_t w = rr * 64 + lane;
                uint64_t v = (w < W && w >= pw) ? S[w] : 0ull;
                if (w == pw) v &= (~0ull << (pos & 63u));
                const uint64_t bal = __ballot(v != 0ull);
                if (bal) {
                    const int f = __builtin_ctzll(bal);
                    const uint64_t vf = __shfl(v, f, 64);
                    xs = ((rr * 64 + (uint32_t)f) << 6) + (uint32_t)__builtin_ctzll(vf);
                    break;
                }
            }
            if (xs == 0xffffffffu) break;
            if ((++steps & 63u) == 0u && past(dl)) return true;
            const uint64_t *Es = Eh + (uint64_t)(xs - rlo) * W;  // x_a's row: kept, as the walk reads it
            for (uint32_t rr = blo >> 6; rr < R; rr++) {
                const uint32_t w = rr * 64 + lane;
                if (w < bhi && w >= blo) Bv[w] |= Es[w];
            }
            pos = xs + 1u;
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t rr = 0; rr < R; rr++) {
            const uint32_t w = rr * 64 + lane;
            if (w >= W) break;
            Bv[w] = (w >= blo && w < bhi && Bv[w]) ? Bv[w] & b_word(w) : 0ull;
        }
        __builtin_amdgcn_wave_barrier();
        uint32_t pos = blo << 6;
        for (;;) {
            uint32_t xb = 0xffffffffu;
            const uint32_t pw = pos >> 6;
            for (uint32_t rr = pw >> 6; rr < R; rr++) {
                const uint32_t w = rr * 64 + lane;
                uint64_t v = (w < W && w >= pw) ? Bv[w] : 0ull;
                if (w == pw) v &= (~0ull << (pos & 63u));
                const uint64_t bal = __ballot(v != 0ull);
                if (bal) {
                    const int f = __builtin_ctzll(bal);
                    const uint64_t vf = __shfl(v, f, 64);
                    xb = ((rr * 64 + (uint32_t)f) << 6) + (uint32_t)__builtin_ctzll(vf);
                    break;
                }
            }
            if (xb == 0xffffffffu) break;
            if ((++steps & 63u) == 0u && past(dl)) return true;
            const uint64_t *Eb = Eh + (uint64_t)(xb - rlo) * W;
            uint32_t pc = 0;
            for (uint32_t rr = slo >> 6; rr < R; rr++) {
                const uint32_t w = rr * 64 + lane;
                if (w < W && w >= slo) {
                    const uint64_t sw = S[w];
                    if (sw) pc += (uint32_t)__popcll(sw & Eb[w]);  // x_b's row only where S has members
                }
            }
            pc = (uint32_t)wave_sum_u64(pc);
            if (lane == 0) cnt += (uint64_t)pc * (pc - 1u) / 2u;
            pos = xb + 1u;
        }
        return false;
    };

    // DFS below level 2 with C_2 in my[0, W): x_2 runs over C_2 n [pos0, hi2).
    auto dfs3 = [&](uint32_t pos0, uint32_t hi2) {
        int l = 3;
        uint32_t pos = pos0;
        for (;;) {
            if (++spins == kClockEvery) {
                spins = 0;
                if (past(dl)) break;  // the next claim records the stop
            }
            if (pl.tail && l == L - 1) {  // every x_{L-2} of the level at once
                tail_count(l, pos, l == 3 ? hi2 : 0xffffffffu);
                if (l == 3) break;
                --l;
                pos = xget(l - 1) + 1u;
                continue;
            }
            if (PAIR && l == L - 2) {  // levels L-3..L-1 at once; at l = 3 [pos0, hi2) cuts B
                const bool top = l == 3;
                if (pair_count(l, top ? pos0 : 0u, top ? hi2 : 0xffffffffu)) break;  // the next claim records the stop
                if (l == 3) break;
                --l;
                pos = xget(l - 1) + 1u;
                continue;
            }
            uint32_t xi = 0xffffffffu;
            uint32_t pw = pos >> 6;
            const uint32_t wl = sWlo[wv][l - 1];
            const uint32_t start = pw > wl ? pw : wl;
            for (uint32_t rr = start >> 6; rr < R; rr++) {
                const uint32_t w = rr * 64 + lane;
                uint64_t v = (w < W && w >= start) ? my[(uint64_t)(l - 3) * wmax + w] : 0ull;
                if (w == pw) v &= (~0ull << (pos & 63u));
                const uint64_t bal = __ballot(v != 0ull);
                if (bal) {
                    const int f = __builtin_ctzll(bal);
                    const uint64_t vf = __shfl(v, f, 64);
                    xi = ((rr * 64 + (uint32_t)f) << 6) + (uint32_t)__builtin_ctzll(vf);
                    break;
                }
            }
            if (l == 3 && xi != 0xffffffffu && xi >= hi2) xi = 0xffffffffu;
            if (xi == 0xffffffffu) {
                if (l == 3) break;
                --l;
                pos = xget(l - 1) + 1u;
                continue;
            }
            if (lane == 0) sX[wv][l - 1] = xi;
            __builtin_amdgcn_wave_barrier();
            const uint32_t lo = word_lo(l);
            if (l == L - 1) {
                for (uint32_t rr = lo >> 6; rr < R; rr++) {
                    const uint32_t w = rr * 64 + lane;
                    if (w >= W) break;
                    if (w < lo) continue;
                    const uint64_t t = level_word(l, w);
                    cnt += (uint64_t)__popcll(t);
                    if (pl.emit && t) emit_word(out, pl, l, u, A, t, w, xget, hs, hx);
                }
                pos = xi + 1u;
            } else {
                for (uint32_t rr = lo >> 6; rr < R; rr++) {
                    const uint32_t w = rr * 64 + lane;
                    if (w < W && w >= lo) my[(uint64_t)(l - 2) * wmax + w] = level_word(l, w);
                }
                if (lane == 0) sWlo[wv][l] = lo;
                __builtin_amdgcn_wave_barrier();
                ++l;
                pos = 0;
            }
        }
    };

    auto bind = [&](const HubRec &h, uint32_t x1) {
        d = h.d;
        W = (d + 63) >> 6;
        R = (W + 63) >> 6;
        u = h.u;
        rlo = h.rlo;
        Eh = E + h.eoff;
        A = g.ocol + h.rbeg;
        if (lane == 0) {
            sX[wv][1] = x1;
            sWlo[wv][2] = 0;
        }
        __builtin_amdgcn_wave_barrier();
        if (row1) copy_row(e1, Eh + (uint64_t)(x1 - rlo) * W, W, lane);
    };

    if (mode == 1) {
        const uint64_t ntail = *hq.tail;
        const uint64_t nq = ntail < hq.cap ? ntail : hq.cap;
        uint32_t parts = 0;  // lane 0's claim state
        for (;;) {
            unsigned long long c = ~0ull;
            if (lane == 0) {
                uint64_t cb = 0, ce = 0;
                if (past(dl)) mark_stop(out);
                else if (claim_striped(hq.head, nq, wslot % kStripes, parts, 1, cb, ce)) c = cb;
            }
            c = __shfl(c, 0, 64);
            if (c >= nq) break;
            const HubChunk ch = hq.q[c];
            if (ch.lo >= ch.hi) continue;
            bind(hubs[ch.hub], ch.x1);
            copy_row(my, hq.slots + ch.slot, W, lane);
            dfs3(ch.lo, ch.hi);
        }
        flush_acc(out, cnt, hs, hx);
        if (lane == 0) prof_add(out, 2, wslot, t_start);
        return;
    }

    constexpr uint32_t kClaim = 4;  // items per cursor claim
    uint64_t nxt = 0, left = 0;
    uint32_t parts = 0;  // lane 0's claim state
    for (;;) {
        if (left == 0) {
            unsigned long long cb = ~0ull, cc = 0;
            if (lane == 0) {
                uint64_t b = 0, n = 0;
                if (past(dl)) mark_stop(out);
                else if (claim_striped(cursor, i1 - i0, wslot % kStripes, parts, kClaim, b, n)) {
                    cb = b;
                    cc = n;
                }
            }
            cb = __shfl(cb, 0, 64);
            cc = __shfl(cc, 0, 64);
            if (cb == ~0ull) break;
            nxt = i0 + cb;
            left = cc;
        }
        const uint64_t item = nxt;
        nxt += kStripes;
        left--;
        const uint32_t hidx = hub_of_item[item];
        const HubRec h = hubs[hidx];
        const uint32_t jl = h.jbeg + (uint32_t)(item - h.ioff);
        if (L < 3) continue;
        bind(h, jl);
        const uint32_t lo2 = word_lo(2);
        if (L == 3) {
            for (uint32_t rr = lo2 >> 6; rr < R; rr++) {
                const uint32_t w = rr * 64 + lane;
                if (w >= W) break;
                if (w < lo2) continue;
                const uint64_t t = level_word(2, w);
                cnt += (uint64_t)__popcll(t);
                if (pl.emit && t) emit_word(out, pl, 2, u, A, t, w, xget, hs, hx);
            }
            continue;
        }
        uint32_t c2 = 0;
        for (uint32_t rr = 0; rr < R; rr++) {
            const uint32_t w = rr * 64 + lane;
            if (w < W) {
                const uint64_t t = w >= lo2 ? level_word(2, w) : 0ull;
                my[w] = t;
                c2 += (uint32_t)__popcll(t);
            }
        }
        c2 = (uint32_t)wave_sum_u64(c2);
        if (PAIR && L == 5) {
            // The item is one paired pass over B's words with S = C_2; a long one
            // donates ranges of B's rounds (the sum splits over B's positions;
            // B's bounds depend on x_1 only).
            const uint32_t r0 = word_lo(3) >> 6;
            const uint32_t nr = R > r0 ? R - r0 : 0u;
            uint32_t rpc = c2 ? kPairChunkWork / c2 : nr;
            if (rpc < 1u) rpc = 1u;
            const uint32_t nc = (nr + rpc - 1u) / rpc;
            if (hq.q && nc >= 2u && (uint64_t)c2 * nr > kPairSplitWork) {
                unsigned long long sb = 0;
                if (lane == 0) sb = atomicAdd(hq.slot_tail, (unsigned long long)W);
                sb = __shfl(sb, 0, 64);
                if (sb + W <= hq.slot_cap) {
                    unsigned long long qb = 0;
                    if (lane == 0) qb = atomicAdd(hq.tail, (unsigned long long)nc);
                    qb = __shfl(qb, 0, 64);
                    if (qb + nc <= hq.cap) {
                        if (lane == 0) atomicAdd(hq.donated, (unsigned long long)nc);
                        for (uint32_t rr = 0; rr < R; rr++) {
                            const uint32_t w = rr * 64 + lane;
                            if (w < W) hq.slots[sb + w] = my[w];
                        }
                        for (uint32_t c = lane; c < nc; c += 64) {
                            HubChunk *rec = hq.q + qb + c;
                            rec->hub = hidx;
                            rec->x1 = jl;
                            rec->slot = sb;
                            rec->lo = c ? (r0 + c * rpc) << 12 : 0u;
                            rec->hi = c + 1u == nc ? (W << 6) : (r0 + (c + 1u) * rpc) << 12;
                        }
                        continue;
                    }
                    for (uint64_t c = qb + lane; c < qb + nc && c < hq.cap; c += 64) {
                        hq.q[c].lo = 0;
                        hq.q[c].hi = 0;
                    }
                }
            }
            dfs3(0, 0xffffffffu);
            continue;
        }
        if (hq.q && c2 > hq.split_min) {
            // donate: publish C_2, then cut it into chunks of hq.chunk candidates
            unsigned long long sb = 0;
            if (lane == 0) sb = atomicAdd(hq.slot_tail, (unsigned long long)W);
            sb = __shfl(sb, 0, 64);
            if (sb + W <= hq.slot_cap) {
                const uint32_t nc = (c2 + hq.chunk - 1) / hq.chunk;
                unsigned long long qb = 0;
                if (lane == 0) qb = atomicAdd(hq.tail, (unsigned long long)nc);
                qb = __shfl(qb, 0, 64);
                if (qb + nc <= hq.cap) {
                    if (lane == 0) atomicAdd(hq.donated, (unsigned long long)nc);
                    uint32_t carry = 0;
                    for (uint32_t rr = 0; rr < R; rr++) {
                        const uint32_t w = rr * 64 + lane;
                        const uint64_t t = (w < W) ? my[w] : 0ull;
                        if (w < W) hq.slots[sb + w] = t;
                        const uint32_t pc = (uint32_t)__popcll(t);
                        const uint32_t incl = wave_incl_scan(pc) + carry;
                        const uint32_t excl = incl - pc;
                        for (uint32_t c = (excl + hq.chunk - 1) / hq.chunk; c * hq.chunk < incl; c++) {
                            const uint32_t idx = (w << 6) + select_bit(t, c * hq.chunk - excl);
                            HubChunk *rec = hq.q + qb + c;
                            rec->hub = hidx;
                            rec->x1 = jl;
                            rec->slot = sb;
                            rec->lo = idx;
                            if (c > 0) hq.q[qb + c - 1].hi = idx;
                        }
                        carry = __shfl(incl, 63, 64);
                    }
                    if (lane == 0) hq.q[qb + nc - 1].hi = W << 6;
                    continue;
                }
                // queue full: neutralise the reserved records that exist, walk inline
                for (uint64_t c = qb + lane; c < qb + nc && c < hq.cap; c += 64) {
                    hq.q[c].lo = 0;
                    hq.q[c].hi = 0;
                }
            }
        }
        dfs3(0, 0xffffffffu);
    }
    flush_acc(out, cnt, hs, hx);
    if (lane == 0) prof_add(out, 1, wslot, t_start);
}

// ------------------------------------------------------------ launchers

template <int P, int MinWaves>
static hipError_t launch_lds_p(int grid, size_t lds, hipStream_t st, const KGraph &g, const KPlan &pl,
                               const RootRec *recs, const uint32_t *bf, const uint32_t *bc,
                               uint32_t b0, uint32_t b1, unsigned long long *cursor, const LdsLayout &lay,
                               const KOut &out)
{
    hipLaunchKernelGGL((lds_batch_kernel<P, MinWaves>), dim3(grid), dim3(kBlock), lds, st, g, pl, recs, bf, bc,
                       b0, b1, cursor, lay, out);
    return hipGetLastError();
}

// Workgroups of a kernel per CU with `lds` bytes of dynamic LDS, and its
// dynamic-LDS attribute raised to at least `lds`.  Both are host calls of
// several microseconds on the launch path of every query, so they are made
// once per (kernel, LDS size); the attribute only ever grows.
struct LdsLaunchCache {
    std::mutex mu;
    size_t attr = 0;
    size_t lds[16] = {};
    int occ[16] = {};
    int n = 0;
    int get(const void *fn, size_t bytes)
    {
        std::lock_guard<std::mutex> lk(mu);
        for (int i = 0; i < n; i++)
            if (lds[i] == bytes) return occ[i];
        if (bytes > attr) {
            (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            attr = bytes;
        }
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kBlock, bytes) != hipSuccess || nb < 1) nb = 1;
        if (n < 16) {
            lds[n] = bytes;
            occ[n++] = nb;
        }
        return nb;
    }
};

template <int P, int MinWaves>
static int occupancy_lds_p(size_t lds)
{
    static LdsLaunchCache cache;
    return cache.get((const void *)lds_batch_kernel<P, MinWaves>, lds);
}

// Dispatch on the DFS group width.
hipError_t launch_lds_batches(int lg, int cus, hipStream_t st, const KGraph &g, const KPlan &pl,
                              const RootRec *recs, const uint32_t *bf, const uint32_t *bc,
                              uint32_t b0, uint32_t b1, unsigned long long *cursor, const KOut &out,
                              int *grid_out, bool dense)
{
    const uint32_t nlev = pl.L > 3 ? (uint32_t)(pl.L - 3) : 1u;
    LdsLayout lay = make_lds_layout(lg, nlev);
    size_t lds = lay.bytes;
    int occ = 1;
    switch (lg) {
#define BEEP_CASE(LG)                                                                              \
    case LG:                                                                                       \
        occ = dense ? occupancy_lds_p<(1 << LG), 7>(lds) : occupancy_lds_p<(1 << LG), 1>(lds);    \
        break;
        BEEP_CASE(0) BEEP_CASE(1) BEEP_CASE(2) BEEP_CASE(3) BEEP_CASE(4) BEEP_CASE(5) BEEP_CASE(6)
#undef BEEP_CASE
    default: return hipErrorInvalidValue;
    }
    int grid = occ * cus;
    if ((uint32_t)grid > b1 - b0) grid = (int)(b1 - b0);
    if (grid < 1) grid = 1;
    if (grid_out) *grid_out = grid;
    switch (lg) {
#define BEEP_CASE(LG)                                                                                    \
    case LG:                                                                                             \
        return dense ? launch_lds_p<(1 << LG), 7>(grid, lds, st, g, pl, recs, bf, bc, b0, b1, cursor, lay, out) \
                     : launch_lds_p<(1 << LG), 1>(grid, lds, st, g, pl, recs, bf, bc, b0, b1, cursor, lay, out);
        BEEP_CASE(0) BEEP_CASE(1) BEEP_CASE(2) BEEP_CASE(3) BEEP_CASE(4) BEEP_CASE(5) BEEP_CASE(6)
#undef BEEP_CASE
    }
    return hipErrorInvalidValue;
}

// Zero fill (used instead of memset nodes so launch sequences can be captured).
// With dl set, a workgroup that starts after that deadline clears nothing: a
// hub round's arena is only read by kernels that stop at the same deadline.
__global__ void fill_zero_u32_kernel(uint32_t *p, uint64_t n, const unsigned long long *dl)
{
    if (dl && past(*dl)) return;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

hipError_t launch_fill_zero(void *p, uint64_t bytes, hipStream_t st, const unsigned long long *dl)
{
    const uint64_t n = (bytes + 3) / 4;
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 4 * kBlock - 1) / (4 * kBlock);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_zero_u32_kernel, dim3((uint32_t)blocks), dim3(kBlock), 0, st, (uint32_t *)p, n, dl);
    return hipGetLastError();
}

// First kernel of a launch sequence: clears the n counter words and writes
// the deadlines of the hub and the LDS kernels at words dl and dl + 1:
// now + ticks, or none (all ones).
__global__ void begin_sequence_kernel(unsigned long long *ctr, uint64_t n, uint64_t dl, uint64_t ticks_hub,
                                      uint64_t ticks_lds)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        unsigned long long v = 0;
        if (i == dl || i == dl + 1) {
            const uint64_t t = i == dl ? ticks_hub : ticks_lds;
            v = t ? (unsigned long long)wall_clock64() + t : ~0ull;
        }
        ctr[i] = v;
    }
}

hipError_t launch_begin_sequence(unsigned long long *ctr, uint64_t n, uint64_t dl, uint64_t ticks_hub,
                                 uint64_t ticks_lds, hipStream_t st)
{
    const uint64_t blocks = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(begin_sequence_kernel, dim3((uint32_t)(blocks ? blocks : 1)), dim3(kBlock), 0, st, ctr, n, dl,
                       ticks_hub, ticks_lds);
    return hipGetLastError();
}

// One lane class's batches (lg: the widest of them) of several queries.
hipError_t launch_lds_multi(int cus, hipStream_t st, const MultiQuery *qs, const MultiWork *work,
                            uint32_t nwork, unsigned long long *cursor, uint32_t nlev, int lg, int *grid_out)
{
    LdsLayout lay = make_lds_layout(lg, nlev);
    size_t lds = lay.bytes;
    static LdsLaunchCache cache;
    const int occ = cache.get((const void *)lds_multi_kernel, lds);
    int grid = occ * cus;
    if ((uint32_t)grid > nwork) grid = (int)nwork;
    if (grid < 1) grid = 1;
    if (grid_out) *grid_out = grid;
    hipLaunchKernelGGL(lds_multi_kernel, dim3(grid), dim3(kBlock), lds, st, qs, work, nwork, cursor, lay);
    return hipGetLastError();
}

hipError_t launch_hub_hash(int grid, hipStream_t st, const KGraph &g, const HubRec *hubs,
                           const uint32_t *hor, uint64_t g0, uint64_t g1, uint64_t *kv,
                           const unsigned long long *dl)
{
    hipLaunchKernelGGL(hub_hash_kernel, dim3(grid), dim3(kBlock), 0, st, g, hubs, hor, g0, g1, kv, dl);
    return hipGetLastError();
}

hipError_t launch_hub_encode(int grid, hipStream_t st, const KGraph &g, const HubRec *hubs,
                             const uint32_t *hor, const uint64_t *hrange, const uint64_t *rows, uint64_t nrows,
                             const uint64_t *kv, uint64_t *E, int upper, const KOut &out)
{
    hipLaunchKernelGGL(hub_encode_kernel, dim3(grid), dim3(kBlock), 0, st, g, hubs, hor, hrange, rows, nrows, kv, E,
                       upper, out);
    return hipGetLastError();
}

// Dynamic LDS of the hub DFS: per wave nlev + 1 rows of wmax words.
static size_t hub_lds_bytes(int nlev, uint32_t wmax) { return (size_t)(kBlock / 64) * hub_wave_words(nlev, wmax) * 8; }
constexpr size_t kHubLdsMax = 96 * 1024;

bool hub_dfs_in_lds(int nlev, uint32_t wmax) { return hub_lds_bytes(nlev, wmax) <= kHubLdsMax; }

// one cache per instantiation: [in LDS][pair]
static LdsLaunchCache &hub_dfs_cache(bool in_lds, bool pair)
{
    static LdsLaunchCache c[2][2];
    return c[in_lds][pair];
}

// The instantiation a plan runs: the stack in LDS or HBM, and the paired last
// levels compiled in only for the plans that have them (their counters cost
// registers, hence occupancy, that the other plans keep).
static const void *hub_dfs_fn(int nlev, uint32_t wmax, bool pair)
{
    if (hub_dfs_in_lds(nlev, wmax))
        return pair ? (const void *)hub_dfs_kernel<true, true> : (const void *)hub_dfs_kernel<true, false>;
    return pair ? (const void *)hub_dfs_kernel<false, true> : (const void *)hub_dfs_kernel<false, false>;
}

hipError_t launch_hub_dfs(int grid, hipStream_t st, const KGraph &g, const KPlan &pl, const HubRec *hubs,
                          const uint32_t *hoi, uint64_t i0, uint64_t i1, unsigned long long *cursor,
                          const uint64_t *E, uint64_t *stk, uint32_t wmax, const KOut &out,
                          const HubQueue &hq, int mode)
{
    const int nlev = pl.L > 3 ? pl.L - 3 : 1;
    const bool in_lds = hub_dfs_in_lds(nlev, wmax);
    const size_t lds = in_lds ? hub_lds_bytes(nlev, wmax) : 0;
    const void *fn = hub_dfs_fn(nlev, wmax, pl.pair != 0);
    if (in_lds) (void)hub_dfs_cache(true, pl.pair != 0).get(fn, lds);
    void *args[] = {(void *)&g,  (void *)&pl,     (void *)&hubs, (void *)&hoi, (void *)&i0,
                    (void *)&i1, (void *)&cursor, (void *)&E,    (void *)&stk, (void *)&wmax,
                    (void *)&out, (void *)&hq,    (void *)&mode};
    return hipLaunchKernel(fn, dim3(grid), dim3(kBlock), args, lds, st);
}

int hub_dfs_occupancy(int nlev, uint32_t wmax, bool pair)
{
    const void *fn = hub_dfs_fn(nlev, wmax, pair);
    const bool in_lds = hub_dfs_in_lds(nlev, wmax);
    return hub_dfs_cache(in_lds, pair).get(fn, in_lds ? hub_lds_bytes(nlev, wmax) : 0);
}

}  // namespace beep
