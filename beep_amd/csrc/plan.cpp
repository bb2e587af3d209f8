// plan.cpp -- query planner: matching order, back-edge lists, symmetry-breaking
// chain, unmaterialised last level and the ascending/descending choice.
//
// Follows SG_Match::preprocess_query / detect_symmetry / detect_reuse /
// check_unmat (subgraph_matching.cuh:257-631).  nauty's densenauty call
// (:434) is replaced by an automorphism search of our own over the at most
// 10-vertex position graph: candidate images are filtered by degree and
// already-fixed neighbourhoods, and the orbit partition is the union of the
// cycles of every automorphism found.

#include <algorithm>

#include "internal.hpp"

namespace beep {

namespace {

struct SmallGraph {
    int k = 0;
    uint16_t adj[kMaxK] = {};  // bit j of adj[i]: edge i-j
    int deg(int v) const { return __builtin_popcount(adj[v]); }
};

// Orbits of the automorphisms of g fixing every vertex in `fixed` (bitmask).
// orbit[v] = smallest vertex of v's orbit.
void automorphism_orbits(const SmallGraph &g, uint16_t fixed, int *orbit)
{
    const int k = g.k;
    int parent[kMaxK];
    for (int v = 0; v < k; v++) parent[v] = v;
    auto find = [&](int v) {
        while (parent[v] != v) v = parent[v] = parent[parent[v]];
        return v;
    };
    int img[kMaxK];
    uint16_t used = 0;
    // depth-first assignment img[0], img[1], ...; a partial map is kept only
    // when it preserves adjacency to every already-mapped vertex.
    struct Frame { int next; };
    Frame st[kMaxK + 1];
    int depth = 0;
    st[0].next = 0;
    while (depth >= 0) {
        if (depth == k) {
            for (int v = 0; v < k; v++) {
                int a = find(v), b = find(img[v]);
                if (a != b) parent[std::max(a, b)] = std::min(a, b);
            }
            --depth;
            if (depth >= 0) used &= (uint16_t)~(1u << img[depth]);
            continue;
        }
        int c = st[depth].next;
        bool placed = false;
        for (; c < k; c++) {
            if (used & (1u << c)) continue;
            if ((fixed >> depth) & 1u) { if (c != depth) continue; }
            if (g.deg(c) != g.deg(depth)) continue;
            bool ok = true;
            for (int p = 0; p < depth && ok; p++)
                if (((g.adj[depth] >> p) & 1u) != ((g.adj[c] >> img[p]) & 1u)) ok = false;
            if (!ok) continue;
            placed = true;
            break;
        }
        if (!placed) {
            --depth;
            if (depth >= 0) used &= (uint16_t)~(1u << img[depth]);
            continue;
        }
        st[depth].next = c + 1;
        img[depth] = c;
        used |= (uint16_t)(1u << c);
        ++depth;
        if (depth <= k) st[depth].next = 0;
    }
    for (int v = 0; v < k; v++) orbit[v] = find(v);
}

}  // namespace

std::unique_ptr<Plan> plan_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t ne,
                                      int mode)
{
    if (mode != BEEP_MODE_SGM && mode != BEEP_MODE_SGC) throw Error(BEEP_EINVAL, "mode must be sgm or sgc");
    // Pattern CSR exactly as the data path builds it (main.cu:322).
    std::unique_ptr<Graph> q = csr_from_edges(src, dst, ne);
    const int k = (int)q->n;
    if (k < 3 || k > kMaxK)
        throw Error(BEEP_EINVAL, "query must have between 3 and 10 vertices");
    for (int i = 0; i < k; i++)
        for (uint64_t a = q->rowptr[i]; a < q->rowptr[i + 1]; a++)
            for (uint64_t b = a + 1; b < q->rowptr[i + 1]; b++)
                if (q->col[a] == q->col[b])
                    throw Error(BEEP_EINVAL, "query lists an edge twice");
    auto p = std::make_unique<Plan>();
    p->k = k;
    p->mode = mode;
    uint32_t degree[kMaxK];
    p->max_qdeg = 0;
    p->min_qdeg = 1 << 30;
    for (int i = 0; i < k; i++) {
        degree[i] = (uint32_t)(q->rowptr[i + 1] - q->rowptr[i]);
        p->max_qdeg = std::max<int>(p->max_qdeg, (int)degree[i]);
        p->min_qdeg = std::min<int>(p->min_qdeg, (int)degree[i]);
    }
    // Matching order: repeatedly take the unplaced vertex maximising
    // dM*k + degree (dM = placed neighbours), first index on ties (:297-351).
    int dM[kMaxK] = {};
    bool placed[kMaxK] = {};
    for (int i = 0; i < k; i++) {
        uint32_t best = 0;
        int idx = 0;
        for (int j = 0; j < k; j++) {
            if (placed[j]) continue;
            uint32_t lk = (uint32_t)dM[j] * (uint32_t)k + degree[j];
            if (lk > best) { best = lk; idx = j; }
        }
        p->seq[i] = (uint32_t)idx;
        p->qdeg[i] = degree[idx];
        placed[idx] = true;
        for (uint64_t e = q->rowptr[idx]; e < q->rowptr[idx + 1]; e++)
            if (!placed[q->col[e]]) dM[q->col[e]]++;
        // back edges: earlier positions adjacent to this vertex, ascending
        p->qedge_ptr[i + 1] = p->qedge_ptr[i];
        for (int j = 0; j < i; j++)
            for (uint64_t e = q->rowptr[idx]; e < q->rowptr[idx + 1]; e++)
                if (q->col[e] == p->seq[j]) p->qedge[p->qedge_ptr[i + 1]++] = (uint32_t)j;
    }
    // Symmetry-breaking chain over the position graph (:397-505): take the
    // largest orbit (smallest representative on ties) of the automorphisms that
    // fix the representatives chosen so far; every other member of that orbit
    // must map above the representative.
    SmallGraph pg;
    pg.k = k;
    for (int i = 0; i < k; i++)
        for (uint32_t e = p->qedge_ptr[i]; e < p->qedge_ptr[i + 1]; e++) {
            pg.adj[i] |= (uint16_t)(1u << p->qedge[e]);
            pg.adj[p->qedge[e]] |= (uint16_t)(1u << i);
        }
    uint16_t fixed = 0;
    std::vector<int> symlist[kMaxK];
    for (;;) {
        int orbit[kMaxK];
        automorphism_orbits(pg, fixed, orbit);
        int size[kMaxK] = {};
        for (int v = 0; v < k; v++) size[orbit[v]]++;
        int rep = -1, best = 1;
        for (int v = 0; v < k; v++)
            if (size[v] > best) { best = size[v]; rep = v; }
        if (rep < 0) break;
        for (int v = 0; v < k; v++)
            if (orbit[v] == rep && v != rep) symlist[v].push_back(rep);
        fixed |= (uint16_t)(1u << rep);
    }
    for (int i = 0; i < k; i++) {
        p->sym_ptr[i + 1] = p->sym_ptr[i];
        for (int s : symlist[i]) p->sym[p->sym_ptr[i + 1]++] = (uint32_t)s;
    }
    // LUNMAT (sgc only): the second-to-last position is not a back edge of the
    // last one (:612-631).
    if (mode == BEEP_MODE_SGC) {
        bool found = false;
        for (uint32_t e = p->qedge_ptr[k - 1]; e < p->qedge_ptr[k]; e++)
            if (p->qedge[e] == (uint32_t)(k - 2)) found = true;
        p->lunmat = found ? 0 : 1;
    }
    // Ascending symmetry breaking iff position 1 is bound to position 0
    // (first_sym_level <= 2, :555-562, 1299).  With no symmetry at all the
    // reference reads an uninitialised member; it is taken as descending.
    p->asc = (p->sym_ptr[2] - p->sym_ptr[1]) > 0 ? 1 : 0;
    p->skip_split = (p->sym_ptr[2] == 1) ? 1 : 0;
    p->L = k - p->lunmat;
    for (int l = 0; l < k; l++) {
        uint16_t qm = 0, sm = 0;
        for (uint32_t e = p->qedge_ptr[l] + 1; e < p->qedge_ptr[l + 1]; e++) qm |= (uint16_t)(1u << p->qedge[e]);
        for (uint32_t e = p->sym_ptr[l]; e < p->sym_ptr[l + 1]; e++)
            if (p->sym[e] > 0) sm |= (uint16_t)(1u << p->sym[e]);
        p->qmask[l] = qm;
        p->smask[l] = sm;
    }
    // x_1 >= split when skip_split; level l is bound above x_1 when one of its
    // symmetry constraints refers to a position that is.
    bool bound[kMaxK] = {};
    bound[1] = true;
    bool all = p->skip_split != 0;
    for (int l = 2; l < p->L; l++) {
        for (int s = 1; s < l; s++)
            if (((p->smask[l] >> s) & 1u) && bound[s]) bound[l] = true;
        if (!bound[l]) all = false;
    }
    p->above_split = all ? 1 : 0;
    // Bitmap rows are read only for positions that appear in some QEDGE(l)[1:];
    // when all of them are bound above x_1 >= split, rows below the split are
    // never read.
    bool rows = p->skip_split != 0;
    for (int l = 2; l < p->L; l++)
        for (int q = 1; q < l; q++)
            if (((p->qmask[l] >> q) & 1u) && !bound[q]) rows = false;
    p->rows_above_split = rows ? 1 : 0;
    p->uses_bitmap = 0;
    for (int l = 2; l < p->L; l++)
        if (p->qmask[l]) p->uses_bitmap = 1;
    // Whenever level l intersects with the row of x_q (q in QEDGE(l)[1:]),
    // its candidates are also bound above some x_s, s in SYM(l), with
    // x_s >= x_q (s = q, or x_s > x_q through the chain of symmetry
    // constraints): only the bits of row x_q above x_q are ever used, so of
    // each adjacent pair the bitmaps need the bit in the row of the smaller
    // index only.
    bool gt[kMaxK][kMaxK] = {};  // gt[a][b]: x_a > x_b for every candidate tuple
    for (int l = 1; l < p->L; l++)
        for (int s = 1; s < l; s++)
            if ((p->smask[l] >> s) & 1u) {
                gt[l][s] = true;
                for (int b = 1; b < s; b++)
                    if (gt[s][b]) gt[l][b] = true;
            }
    bool upper = p->uses_bitmap != 0;
    for (int l = 2; l < p->L; l++)
        for (int q = 1; q < l; q++) {
            if (!((p->qmask[l] >> q) & 1u)) continue;
            bool ok = false;
            for (int s = 1; s < l; s++)
                if (((p->smask[l] >> s) & 1u) && (s == q || gt[s][q])) ok = true;
            if (!ok) upper = false;
        }
    p->upper_rows = upper ? 1 : 0;
    return p;
}

}  // namespace beep
