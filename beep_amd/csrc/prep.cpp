// prep.cpp -- data preparation on the host: peeling, degree orientation, the
// root queue, and the split of roots into LDS batches and HBM hub roots.
//
// Semantics follow SG_Match::peel_data (subgraph_matching.cuh:685-764),
// degree_ordering (:1282-1336 with host_utils.cuh:89-134, 230-246) and the
// root selection of count_subgraphs (:1080-1131 with GraphQueue.cuh:85-120,
// device_utils.cuh:60-97).  The reference runs these as a chain of
// synchronous kernels over managed memory; here they are parallel host loops
// whose output is uploaded once and stays resident in HBM.

#include <algorithm>
#include <numeric>

#include "internal.hpp"

namespace beep {

// Hash slots of a root's row (LDS batches and hub arena): the smallest power
// of two >= max(4, 4d), as the kernels' hash_slots_dev.
uint32_t hash_slots_for(uint32_t d)
{
    uint32_t h = 4;
    while (h < 4 * d) h <<= 1;
    return h;
}

namespace {

// Peel: S = {v : 1 <= d(v) < min_qdeg}; every edge at S is removed in both
// directions (the reverse entry is the first occurrence of v in the
// neighbour's row, common_utils.cuh:217-246); repeat while |S| > 0.05 n with
// n fixed (subgraph_matching.cuh:758).
int peel(const Graph &g, int min_qdeg, std::vector<uint64_t> &rp, std::vector<uint32_t> &col)
{
    const uint32_t n = g.n;
    rp = g.rowptr;
    col = g.col;
    std::vector<uint32_t> deg(n);
    parallel_for(n, [&](size_t b, size_t e) {
        for (size_t v = b; v < e; v++) deg[v] = (uint32_t)(rp[v + 1] - rp[v]);
    });
    const uint32_t lim = (uint32_t)std::max(min_qdeg, 0);
    int rounds = 0;
    std::vector<uint8_t> keep;
    for (;;) {
        rounds++;
        std::vector<uint32_t> S;
        for (uint32_t v = 0; v < n; v++)
            if (deg[v] >= 1 && deg[v] < lim) S.push_back(v);
        if (S.empty()) {  // the round removes nothing: the rows stay as they are
            progress("peel: round %d removed the edges of 0 vertices", rounds);
            break;
        }
        const uint64_t m = rp[n];
        keep.assign(m, 1);
        const bool sorted = g.rows_sorted;
        parallel_for(S.size(), [&](size_t b, size_t e) {
            for (size_t s = b; s < e; s++) {
                uint32_t v = S[s];
                for (uint64_t x = rp[v]; x < rp[v + 1]; x++) {
                    __atomic_store_n(&keep[x], (uint8_t)0, __ATOMIC_RELAXED);
                    uint32_t w = col[x];
                    const uint32_t *row = col.data() + rp[w];
                    const uint64_t len = rp[w + 1] - rp[w];
                    uint64_t f = len;
                    if (sorted) {
                        f = (uint64_t)(std::lower_bound(row, row + len, v) - row);
                        if (f < len && row[f] != v) f = len;
                    } else {
                        for (uint64_t y = 0; y < len; y++)
                            if (row[y] == v) { f = y; break; }
                    }
                    if (f < len) __atomic_store_n(&keep[rp[w] + f], (uint8_t)0, __ATOMIC_RELAXED);
                }
            }
        }, 256);
        std::vector<uint64_t> nrp(n + 1, 0);
        parallel_for(n, [&](size_t b, size_t e) {
            for (size_t v = b; v < e; v++) {
                uint64_t c = 0;
                for (uint64_t x = rp[v]; x < rp[v + 1]; x++) c += keep[x];
                nrp[v + 1] = c;
            }
        });
        for (uint32_t v = 0; v < n; v++) nrp[v + 1] += nrp[v];
        std::vector<uint32_t> ncol(nrp[n]);
        parallel_for(n, [&](size_t b, size_t e) {
            for (size_t v = b; v < e; v++) {
                uint64_t o = nrp[v];
                for (uint64_t x = rp[v]; x < rp[v + 1]; x++)
                    if (keep[x]) ncol[o++] = col[x];
                deg[v] = (uint32_t)(nrp[v + 1] - nrp[v]);
            }
        });
        rp.swap(nrp);
        col.swap(ncol);
        progress("peel: round %d removed the edges of %zu vertices", rounds, S.size());
        if (!((double)S.size() > 0.05 * (double)n)) break;
    }
    return rounds;
}

}  // namespace

std::shared_ptr<Variant> make_variant(const Graph &g, int min_qdeg, int asc_flag)
{
    // The reference's bitmap build binary-searches A[i] in the id-ordered row
    // of A[j] (host_utils.cuh:284-316), and its orientation keeps the input's
    // order among neighbours of equal degree (a stable sort on degree only,
    // subgraph_matching.cuh:1305-1331): its counts are defined only for rows
    // sorted by id, which the reference's -s option produces (main.cu:88-91).
    if (!g.rows_sorted)
        throw Error(BEEP_EINVAL, "graph rows are not sorted by destination id; load the graph with "
                                 "edge sorting on (-s 1, beep_graph_load(path, 1, ...))");
    auto vp = std::make_shared<Variant>();
    Variant &V = *vp;
    V.n = g.n;
    V.asc = asc_flag;
    V.min_qdeg = min_qdeg;
    V.cutoff = g.cutoff;
    V.peel_rounds = peel(g, min_qdeg, V.rowptr, V.col);
    const uint32_t n = g.n;
    V.m = V.rowptr[n];
    const auto &rp = V.rowptr;
    std::vector<uint32_t> deg(n);
    parallel_for(n, [&](size_t b, size_t e) {
        for (size_t v = b; v < e; v++) deg[v] = (uint32_t)(rp[v + 1] - rp[v]);
    });
    // Orientation: key(v) = (deg(v), v).  asc rows ascend by key; desc rows
    // descend by degree with ids ascending (the stable sorts of :1305-1331 on an
    // id-ascending input).  nlow(u) = #neighbours with a smaller key.
    V.ocol.resize(V.m);
    V.nlow.assign(n, 0);
    V.hlen.assign(n, 0);
    const bool asc = asc_flag != 0;
    // in slices, so very large graphs report progress between them
    const uint32_t nslices = V.m > (1ull << 30) ? 16 : 1;
    for (uint32_t sl = 0; sl < nslices; sl++) {
    const size_t s0 = (size_t)n * sl / nslices, s1 = (size_t)n * (sl + 1) / nslices;
    parallel_for(s1 - s0, [&](size_t b, size_t e) {
        // rows are sorted on 64-bit keys (degree, or its complement for desc
        // rows, above the id): the order of the comparison on (deg, id)
        // without a degree lookup per comparison
        std::vector<uint64_t> key;
        for (size_t u = s0 + b; u < s0 + e; u++) {
            const uint64_t rb = rp[u], re = rp[u + 1];
            const uint32_t du = deg[u];
            uint32_t lo = 0, hc = 0;
            key.resize(re - rb);
            for (uint64_t x = rb; x < re; x++) {
                const uint32_t w = V.col[x], dw = deg[w];
                if (dw < du || (dw == du && w < (uint32_t)u)) lo++;
                if (dw > du || (dw == du && w < (uint32_t)u)) hc++;
                key[x - rb] = (uint64_t)(asc ? dw : ~dw) << 32 | w;
            }
            std::sort(key.begin(), key.end());
            for (uint64_t x = rb; x < re; x++) V.ocol[x] = (uint32_t)key[x - rb];
            V.nlow[u] = lo;
            // Bitmap-build scan range: every adjacent pair {v, w} is found from
            // exactly one endpoint.  asc rows: the suffix of larger keys.  desc
            // rows (degree descending, ids ascending): the prefix of larger
            // degrees followed by equal degrees with smaller ids.
            V.hlen[u] = asc ? du - lo : hc;
        }
    }, 4096);
    if (nslices > 1) progress("orientation: rows %zu..%zu done", s0, s1);
    }
    uint32_t maxd = 0;
    for (uint32_t v = 0; v < n; v++) maxd = std::max(maxd, deg[v]);
    V.max_degree = maxd;
    progress("variant (min_qdeg %d, %s): %llu entries oriented", min_qdeg, asc ? "asc" : "desc",
             (unsigned long long)V.m);
    return vp;
}

std::unique_ptr<Prep> prepare(const Graph &g, const Plan &p)
{
    return prepare_on(make_variant(g, p.min_qdeg, p.asc), p);
}

std::unique_ptr<Prep> prepare_on(const std::shared_ptr<Variant> &vp, const Plan &p, uint32_t sample_stride)
{
    if (!vp || vp->min_qdeg != p.min_qdeg || vp->asc != p.asc)
        throw Error(BEEP_EINVAL, "variant does not match the plan");
    const Variant &V = *vp;
    auto pp = std::make_unique<Prep>();
    pp->var = vp;
    pp->max_qdeg = p.max_qdeg;
    const uint32_t n = V.n, maxd = V.max_degree;
    const auto &rp = V.rowptr;
    const bool asc = V.asc != 0;
    auto deg = [&](uint32_t v) { return (uint32_t)(rp[v + 1] - rp[v]); };
    // Roots: Q0 = {max_qdeg <= d < 100 cutoff} ordered by degree descending
    // with ids ascending; boundary = #{d > cutoff}; the head of Q0 is skipped
    // when ascending and boundary >= 1 (work_list_head1 starts at 1).
    const uint64_t hi = (uint64_t)V.cutoff * 100u;
    const uint32_t lo = (uint32_t)p.max_qdeg;
    std::vector<uint64_t> bucket((size_t)maxd + 2, 0);
    uint64_t nq = 0;
    for (uint32_t v = 0; v < n; v++)
        if (deg(v) >= lo && deg(v) < hi) { bucket[maxd - deg(v)]++; nq++; }
    uint64_t s = 0;
    for (auto &c : bucket) { uint64_t t = c; c = s; s += t; }
    std::vector<uint32_t> q0(nq);
    for (uint32_t v = 0; v < n; v++)
        if (deg(v) >= lo && deg(v) < hi) q0[bucket[maxd - deg(v)]++] = v;
    uint64_t bnd = 0;
    while (bnd < nq && deg(q0[bnd]) > V.cutoff) bnd++;
    const uint64_t skip = (asc && bnd >= 1) ? 1 : 0;
    pp->roots.assign(q0.begin() + (ptrdiff_t)skip, q0.end());
    pp->q0_size = nq;
    pp->boundary = bnd;
    if (nq) {
        const uint32_t h = q0[0];
        pp->q0_head = h;
        pp->q0_head_larger = deg(h) - V.nlow[h];
    }
    build_work(*pp, p, sample_stride);
    return pp;
}

std::unique_ptr<Prep> prepare_subset(const Prep &src, const Plan &p, const uint64_t *ranks, uint64_t n)
{
    const Variant &V = *src.var;
    if (V.min_qdeg != p.min_qdeg || V.asc != p.asc || src.max_qdeg != p.max_qdeg)
        throw Error(BEEP_EINVAL, "plan does not match the prepared query");
    auto pp = std::make_unique<Prep>();
    pp->var = src.var;
    pp->max_qdeg = src.max_qdeg;
    pp->q0_size = src.q0_size;
    pp->boundary = src.boundary;
    pp->q0_head = src.q0_head;
    pp->q0_head_larger = src.q0_head_larger;
    pp->roots.resize(n);
    for (uint64_t i = 0; i < n; i++) {
        if (ranks[i] >= src.roots.size()) throw Error(BEEP_ERANGE, "root rank out of range");
        if (i && ranks[i] <= ranks[i - 1]) throw Error(BEEP_EINVAL, "root ranks must be strictly ascending");
        pp->roots[i] = src.roots[ranks[i]];
    }
    build_work(*pp, p, 1);
    return pp;
}

// Root-dependent parts of a prepared query: B_alg, the work list (with a
// stratified sample when sample_stride > 1) and its LDS / hub tiers.
void build_work(Prep &pp_, const Plan &p, uint32_t sample_stride)
{
    Prep *pp = &pp_;
    const Variant &V = *pp->var;
    const auto &rp = V.rowptr;
    auto deg = [&](uint32_t v) { return (uint32_t)(rp[v + 1] - rp[v]); };
    pp->max_root_degree = 0;
    for (uint32_t u : pp->roots) pp->max_root_degree = std::max(pp->max_root_degree, deg(u));
    // B_alg = sum over roots of 4(d(u)+2) + sum over neighbours 4(d(v)+2).
    std::atomic<uint64_t> bytes{0};
    parallel_for(pp->roots.size(), [&](size_t b, size_t e) {
        uint64_t acc = 0;
        for (size_t r = b; r < e; r++) {
            uint32_t u = pp->roots[r];
            acc += 4ull * (deg(u) + 2);
            for (uint64_t x = rp[u]; x < rp[u + 1]; x++) acc += 4ull * (deg(V.col[x]) + 2);
        }
        bytes.fetch_add(acc);
    }, 1024);
    pp->alg_bytes = bytes.load();
    // Work list.  With above_split plans a root keeps only its suffix of
    // larger keys (asc rows), which drops the skipped x_1 < split items and
    // every bitmap entry they alone would use; roots left empty carry no work.
    pp->reduced = p.above_split;
    pp->skip_split = p.skip_split;
    pp->rows_above_split = p.rows_above_split;
    pp->uses_bitmap = p.uses_bitmap;
    pp->upper_rows = p.upper_rows;
    {
        auto len_of = [&](uint32_t r) {
            uint32_t u = pp->roots[r];
            return pp->reduced ? deg(u) - V.nlow[u] : deg(u);
        };
        std::vector<uint32_t> idx;
        idx.reserve(pp->roots.size());
        for (uint32_t r = 0; r < (uint32_t)pp->roots.size(); r++)
            if (len_of(r)) idx.push_back(r);
        // roots are in degree-descending order already; only reduced lengths
        // need a (stable) sort
        if (pp->reduced)
            std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return len_of(a) > len_of(b); });
        if (sample_stride > 1) {
            // Hub roots (rows too long for an LDS batch, a prefix of the
            // length-sorted list) are kept every hub_sample_stride(), the rest
            // every sample_stride; the middle entry of every stride, so the
            // heaviest root alone does not stand for the ones after it.
            auto fits_lds = [](uint32_t d) { return LdsBudget::fits(d, hash_slots_for(d)); };
            size_t nh = 0;
            while (nh < idx.size() && !fits_lds(len_of(idx[nh]))) nh++;
            const uint32_t hs = hub_sample_stride(sample_stride);
            size_t o = 0;
            for (size_t i = hs / 2; i < nh; i += hs) idx[o++] = idx[i];
            for (size_t i = nh + sample_stride / 2; i < idx.size(); i += sample_stride) idx[o++] = idx[i];
            idx.resize(o);
        }
        pp->wu.resize(idx.size());
        pp->woff.resize(idx.size());
        pp->wlen.resize(idx.size());
        parallel_for(idx.size(), [&](size_t b, size_t e) {
            for (size_t i = b; i < e; i++) {
                uint32_t u = pp->roots[idx[i]];
                pp->wu[i] = u;
                pp->woff[i] = pp->reduced ? V.nlow[u] : 0u;
                pp->wlen[i] = len_of(idx[i]);
            }
        });
    }
    build_tiers(*pp);
    progress("root queue: %zu roots, %u hubs, %zu LDS batches", pp->roots.size(), pp->n_hub, pp->batches.size());
}

// Roots whose d x ceil(d/64) bitmap fits the batch kernel's LDS budget are
// packed greedily (in queue order) into batches; larger roots keep their
// bitmap in HBM and are processed by the hub kernels.
void build_tiers(Prep &pp)
{
    auto deg = [&](uint32_t i) { return pp.wlen[i]; };
    auto fits_lds = [&](uint32_t d) { return LdsBudget::fits(d, hash_slots_for(d)); };
    uint32_t nh = 0;
    while (nh < pp.wlen.size() && !fits_lds(deg(nh))) nh++;
    pp.n_hub = nh;
    pp.batches.clear();
    pp.batch_lanes_log2.clear();
    // Batches of one lane class should outnumber the workgroups the GPU runs
    // at once, so the per-batch degree sum is capped at ~total/1024 per class.
    const uint32_t nr = (uint32_t)pp.wlen.size();
    auto lg_of = [&](uint32_t d) {
        uint32_t w = (d + 63) / 64, lg = 0;
        while ((1u << lg) < w) lg++;
        return lg;
    };
    uint64_t class_sum[8] = {};
    for (uint32_t j = nh; j < nr; j++) class_sum[lg_of(deg(j))] += deg(j);
    uint32_t i = nh;
    while (i < nr) {
        LdsBatch b{i, 0};
        uint64_t sd = 0, sw = 0, sh = 0;
        const uint32_t lg = lg_of(deg(i));
        uint64_t cap = class_sum[lg] / 1024;
        if (cap < 64) cap = 64;
        if (cap > LdsBudget::sumdeg((int)lg)) cap = LdsBudget::sumdeg((int)lg);
        while (i < nr && b.count < LdsBudget::kMaxRoots) {
            uint32_t d = deg(i);
            uint64_t w = (d + 63) / 64;
            if (b.count && (sd + d > cap || sw + d * w > LdsBudget::ewords((int)lg) ||
                            sh + hash_slots_for(d) > LdsBudget::hash((int)lg) || lg_of(d) != lg))
                break;
            sd += d;
            sw += d * w;
            sh += hash_slots_for(d);
            b.count++;
            i++;
        }
        pp.batches.push_back(b);
        pp.batch_lanes_log2.push_back((uint8_t)lg);
    }
    class_figures(pp);
}

// Work and compulsory bytes of each lane class's launch, with the automatic
// build mode (a row vertex whose scan range is over kLookupRatio x the root's
// degree looks the root's neighbours on its larger-key side up in the range's
// lookup table, one 4-byte slot each).
void class_figures(Prep &pp)
{
    const Variant &V = *pp.var;
    for (int c = 0; c < 8; c++) pp.cls_roots[c] = pp.cls_units[c] = pp.cls_bytes[c] = 0;
    std::vector<uint64_t> bu(pp.batches.size()), bb(pp.batches.size());
    parallel_for(pp.batches.size(), [&](size_t b0, size_t b1) {
    for (size_t bi = b0; bi < b1; bi++) {
        const LdsBatch &b = pp.batches[bi];
        uint64_t units = 0, bytes = 8;  // batch first / count
        for (uint32_t r = b.first; r < b.first + b.count; r++) {
            const uint32_t u = pp.wu[r], d = pp.wlen[r];
            bytes += kRootRecBytes + 4ull * d;
            if (!pp.uses_bitmap) continue;
            const uint64_t rb = V.rowptr[u] + pp.woff[r];
            for (uint32_t j = 0; j < d; j++) {
                const uint32_t v = V.ocol[rb + j];
                const uint64_t hl = V.hlen[v];
                bytes += 16;  // the row vertex's record
                if (hl > (uint64_t)kLookupRatio * d) {
                    // the root's neighbours on the larger-key side of v, one
                    // table slot each
                    const uint64_t c = V.asc ? d - 1 - j : j;
                    units += c;
                    bytes += 4ull * c;
                } else {
                    units += hl;
                    bytes += 4ull * hl;
                }
            }
        }
        bu[bi] = units;
        bb[bi] = bytes;
    }
    }, 256);
    for (size_t bi = 0; bi < pp.batches.size(); bi++) {
        const int lg = pp.batch_lanes_log2[bi];
        pp.cls_roots[lg] += pp.batches[bi].count;
        pp.cls_units[lg] += bu[bi];
        pp.cls_bytes[lg] += bb[bi];
    }
}

}  // namespace beep
