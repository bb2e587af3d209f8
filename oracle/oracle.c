/*
 * oracle.c -- CPU restatement of BEEP's subgraph-enumeration path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The product never links this file.
 * Citations are file:line in the reference tree.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <errno.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void or_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* Graph I/O                                                                 */
/* ------------------------------------------------------------------------ */

static int ends_with(const char *s, const char *suf)
{
    size_t a = strlen(s), b = strlen(suf);
    return a >= b && memcmp(s + a - b, suf, b) == 0;
}

typedef struct {
    uint32_t *src, *dst;
    size_t n, cap;
} edge_vec;

static int ev_push(edge_vec *v, uint32_t s, uint32_t d)
{
    if (v->n == v->cap) {
        size_t nc = v->cap ? v->cap * 2 : 1024;
        uint32_t *a = realloc(v->src, nc * sizeof(uint32_t));
        if (!a) return -1;
        v->src = a;
        uint32_t *b = realloc(v->dst, nc * sizeof(uint32_t));
        if (!b) return -1;
        v->dst = b;
        v->cap = nc;
    }
    v->src[v->n] = s;
    v->dst[v->n] = d;
    v->n++;
    return 0;
}

/* .bel: 24-byte records, dst u64 at +0, src u64 at +8, weight at +16,
 * truncated to uint32 (FIleReader.cuh:52-95).  A trailing partial record is
 * not read (fread of whole items). */
static int64_t read_bel(const char *path, edge_vec *v)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    unsigned char rec[24 * 4096];
    size_t got;
    while ((got = fread(rec, 24, 4096, f)) > 0) {
        for (size_t i = 0; i < got; i++) {
            uint64_t s, d;
            memcpy(&d, rec + i * 24 + 0, 8);
            memcpy(&s, rec + i * 24 + 8, 8);
            if (ev_push(v, (uint32_t)s, (uint32_t)d)) { fclose(f); return -1; }
        }
        if (got < 4096) break;
    }
    fclose(f);
    return (int64_t)v->n;
}

/* .tsv: whitespace separated "dst src weight" triples (FIleReader.cuh:101-127). */
static int64_t read_tsv(const char *path, edge_vec *v)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    for (;;) {
        unsigned long long d, s, w;
        int k = fscanf(f, "%llu %llu %llu", &d, &s, &w);
        if (k != 3) {
            if (feof(f)) break;
            fclose(f);
            return -1; /* the reference exits on a malformed triple */
        }
        if (ev_push(v, (uint32_t)s, (uint32_t)d)) { fclose(f); return -1; }
    }
    fclose(f);
    return (int64_t)v->n;
}

/* Reads one line (without the newline) into *buf.  Returns 0 at EOF with no
 * characters read (std::getline failing), 1 otherwise. */
static int get_line(FILE *f, char **buf, size_t *cap)
{
    ssize_t r = getline(buf, cap, f);
    if (r < 0) {
        if (*cap == 0 || !*buf) { *cap = 16; *buf = malloc(16); }
        (*buf)[0] = 0;
        return 0;
    }
    if (r > 0 && (*buf)[r - 1] == '\n') (*buf)[r - 1] = 0;
    return 1;
}

/* .mtx: header = first line not starting with '%'; "%%" lines set the banner
 * flags (FIleReader.cuh:299-353).  Body lines are 0-based "row col [val]"
 * pairs read 100 per batch; a batch ends at the first line that does not parse,
 * and reading stops at the first batch that yields no edge (main.cu:83-86,
 * 317-320; FIleReader.cuh:130-212).  The "symmetric" and "array" banners are
 * rejected: the reference reader leaves unwritten slots for them
 * (FIleReader.cuh:202-208). */
static int64_t read_mtx(const char *path, edge_vec *v)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char *line = NULL;
    size_t cap = 0;
    int symmetric = 0, array = 0, pattern = 0;
    for (;;) {
        get_line(f, &line, &cap);
        if (line[0] != '%') break;
        if (strlen(line) >= 2 && line[1] == '%') {
            symmetric = strstr(line, "symmetric") != NULL;
            array = strstr(line, "array") != NULL;
            pattern = strstr(line, "pattern") != NULL;
        }
    }
    if (symmetric || array) { free(line); fclose(f); return -2; }
    for (;;) {
        size_t batch = 0;
        for (; batch < 100; batch++) {
            get_line(f, &line, &cap);
            long long r, c;
            double val;
            int ok;
            if (!pattern) ok = sscanf(line, "%lld %lld %lf", &r, &c, &val) >= 2;
            else ok = sscanf(line, "%lld %lld", &r, &c) == 2;
            if (!ok) break;
            if (ev_push(v, (uint32_t)r, (uint32_t)c)) { free(line); fclose(f); return -1; }
        }
        if (batch == 0) break;
    }
    free(line);
    fclose(f);
    return (int64_t)v->n;
}

int64_t or_read_edges(const char *path, uint32_t **src, uint32_t **dst)
{
    edge_vec v = {0};
    int64_t r;
    if (ends_with(path, ".bel")) r = read_bel(path, &v);
    else if (ends_with(path, ".tsv")) r = read_tsv(path, &v);
    else if (ends_with(path, ".mtx")) r = read_mtx(path, &v);
    else r = -3;
    if (r < 0) { free(v.src); free(v.dst); return r; }
    *src = v.src ? v.src : malloc(4);
    *dst = v.dst ? v.dst : malloc(4);
    return r;
}

/* ------------------------------------------------------------------------ */
/* CSR (CSRCOO.cuh:37-80) and cutoff (utils.cuh:186-261)                      */
/* ------------------------------------------------------------------------ */

int or_csr_build(const uint32_t *src, const uint32_t *dst, uint64_t ne, or_csr *g)
{
    memset(g, 0, sizeof(*g));
    uint64_t nrow = 0, nnz = 0;
    /* pass 1: validate grouping, count rows and entries */
    for (uint64_t e = 0; e < ne; e++) {
        if (src[e] == dst[e]) continue;           /* self loops dropped (:51-52) */
        if ((uint64_t)src[e] + 1 < nrow) return -1; /* input not grouped by src (:59) */
        if ((uint64_t)src[e] + 1 > nrow) nrow = (uint64_t)src[e] + 1;
        nnz++;
    }
    g->n = (uint32_t)nrow;
    g->m = nnz;
    g->rowptr = calloc(nrow + 1, sizeof(uint64_t));
    g->col = malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    if (!g->rowptr || !g->col) return -2;
    uint64_t k = 0;
    for (uint64_t e = 0; e < ne; e++) {
        if (src[e] == dst[e]) continue;
        g->rowptr[src[e] + 1]++;
        g->col[k++] = dst[e];
    }
    for (uint64_t i = 0; i < nrow; i++) g->rowptr[i + 1] += g->rowptr[i];
    for (uint64_t e = 0; e < nnz; e++)
        if (g->col[e] >= g->n) return -3; /* destination without a row */
    return 0;
}

void or_csr_free(or_csr *g)
{
    free(g->rowptr);
    free(g->col);
    memset(g, 0, sizeof(*g));
}

/* cutoff = (uint)ceil(mean + 6 sd); totalRow accumulates in 32-bit unsigned
 * arithmetic as T = uint in the reference (utils.cuh:190-241), the variance is
 * a sequential double sum over all rows (:243-255). */
uint32_t or_cutoff(const or_csr *g, double *mean_out, double *sd_out)
{
    uint32_t nr = g->n;
    uint32_t total = 0;
    for (uint32_t i = 0; i < nr; i++) total += (uint32_t)(g->rowptr[i + 1] - g->rowptr[i]);
    double mean = 1.0 * total / nr;
    double sd = 0;
    for (uint64_t i = 0; i < nr; i++) {
        uint32_t rc = (uint32_t)(g->rowptr[i + 1] - g->rowptr[i]);
        double diff = rc - mean;
        sd += diff * diff;
    }
    double var = sd / nr;
    sd = sqrt(var);
    if (mean_out) *mean_out = mean;
    if (sd_out) *sd_out = sd;
    return (uint32_t)ceil(mean + 6.0 * sd);
}

/* ------------------------------------------------------------------------ */
/* Query planner                                                             */
/* ------------------------------------------------------------------------ */

/* Orbits of the automorphisms of the k-vertex graph adj that fix every vertex
 * with fixed[v] != 0, by exhaustive backtracking over permutations.  orbit[v]
 * receives the smallest vertex of v's orbit (nauty's convention). */
static void brute_orbits(int k, const unsigned char adj[OR_MAXK][OR_MAXK], const int *fixed,
                         int *orbit)
{
    int perm[OR_MAXK], used[OR_MAXK] = {0};
    for (int v = 0; v < k; v++) orbit[v] = v;
    /* iterative DFS over partial permutations */
    int depth = 0;
    int choice[OR_MAXK];
    for (int i = 0; i < OR_MAXK; i++) choice[i] = -1;
    while (depth >= 0) {
        if (depth == k) {
            /* perm is an automorphism: merge orbits */
            for (int v = 0; v < k; v++) {
                int a = v, b = perm[v];
                while (orbit[a] != a) a = orbit[a];
                while (orbit[b] != b) b = orbit[b];
                if (a != b) { if (a < b) orbit[b] = a; else orbit[a] = b; }
            }
            depth--;
            if (depth >= 0) used[perm[depth]] = 0;
            continue;
        }
        int c = choice[depth] + 1;
        int placed = 0;
        for (; c < k; c++) {
            if (used[c]) continue;
            if (fixed[depth] && c != depth) continue;
            int ok = 1;
            for (int p = 0; p < depth && ok; p++)
                if (adj[depth][p] != adj[c][perm[p]]) ok = 0;
            if (!ok) continue;
            choice[depth] = c;
            perm[depth] = c;
            used[c] = 1;
            placed = 1;
            break;
        }
        if (placed) {
            depth++;
            if (depth < k) choice[depth] = -1;
        } else {
            choice[depth] = -1;
            depth--;
            if (depth >= 0) used[perm[depth]] = 0;
        }
    }
    for (int v = 0; v < k; v++) {
        int a = v;
        while (orbit[a] != a) a = orbit[a];
        orbit[v] = a;
    }
}

int or_plan_build(const uint32_t *src, const uint32_t *dst, uint64_t ne, int sgc, or_plan *p)
{
    memset(p, 0, sizeof(*p));
    or_csr q;
    if (or_csr_build(src, dst, ne, &q)) return -1;
    int k = (int)q.n;
    if (k < 3 || k > OR_MAXK) { or_csr_free(&q); return -2; }
    /* duplicate entries would duplicate QEDGE lists; they are rejected */
    for (int i = 0; i < k; i++)
        for (uint64_t a = q.rowptr[i]; a < q.rowptr[i + 1]; a++)
            for (uint64_t b = a + 1; b < q.rowptr[i + 1]; b++)
                if (q.col[a] == q.col[b]) { or_csr_free(&q); return -3; }
    p->k = k;
    p->sgc = sgc;
    uint32_t degree[OR_MAXK];
    p->max_qdeg = 0;
    p->min_qdeg = 0x7fffffff;
    for (int i = 0; i < k; i++) {
        degree[i] = (uint32_t)(q.rowptr[i + 1] - q.rowptr[i]);
        if ((int)degree[i] > p->max_qdeg) p->max_qdeg = degree[i];
        if ((int)degree[i] < p->min_qdeg) p->min_qdeg = degree[i];
    }
    /* MLE order: argmax dM*k + deg, strict '>' from ml = 0 (:297-351) */
    int dM[OR_MAXK] = {0};
    p->qedge_ptr[0] = 0;
    for (int i = 0; i < k; i++) {
        uint32_t ml = 0, idx = 0;
        for (int j = 0; j < k; j++) {
            if (dM[j] >= 0) {
                uint32_t lk = (uint32_t)dM[j] * (uint32_t)k + degree[j];
                if (lk > ml) { ml = lk; idx = (uint32_t)j; }
            }
        }
        p->seq[i] = idx;
        p->qdeg[i] = degree[idx];
        dM[idx] = -1;
        for (uint64_t e = q.rowptr[idx]; e < q.rowptr[idx + 1]; e++)
            if (dM[q.col[e]] != -1) dM[q.col[e]]++;
        p->qedge_ptr[i + 1] = p->qedge_ptr[i];
        for (int j = 0; j < i; j++)
            for (uint64_t e = q.rowptr[idx]; e < q.rowptr[idx + 1]; e++)
                if (q.col[e] == p->seq[j]) p->qedge[p->qedge_ptr[i + 1]++] = (uint32_t)j;
    }
    /* position graph and orbit chain (:397-505) */
    unsigned char adj[OR_MAXK][OR_MAXK];
    memset(adj, 0, sizeof(adj));
    for (int i = 0; i < k; i++)
        for (uint32_t e = p->qedge_ptr[i]; e < p->qedge_ptr[i + 1]; e++) {
            adj[i][p->qedge[e]] = 1;
            adj[p->qedge[e]][i] = 1;
        }
    int fixed[OR_MAXK] = {0};
    int symlist[OR_MAXK][OR_MAXK];
    int symlen[OR_MAXK] = {0};
    for (;;) {
        int orbit[OR_MAXK];
        brute_orbits(k, adj, fixed, orbit);
        int cnt[OR_MAXK] = {0};
        for (int i = 0; i < k; i++) cnt[orbit[i]]++;
        int maxlen = 0, maxidx = -1;
        for (int i = 0; i < k; i++)
            if (cnt[i] > maxlen) { maxlen = cnt[i]; maxidx = i; }
        if (maxlen == 1) break;
        for (int i = 0; i < k; i++)
            if (orbit[i] == maxidx && i != maxidx) symlist[i][symlen[i]++] = maxidx;
        fixed[maxidx] = 1;
    }
    p->sym_ptr[0] = 0;
    for (int i = 0; i < k; i++) {
        p->sym_ptr[i + 1] = p->sym_ptr[i];
        for (int j = 0; j < symlen[i]; j++) p->sym[p->sym_ptr[i + 1]++] = (uint32_t)symlist[i][j];
    }
    /* LUNMAT (sgc only): k-2 not in QEDGE(k-1) (:612-631) */
    p->lunmat = 0;
    if (sgc) {
        int found = 0;
        for (uint32_t e = p->qedge_ptr[k - 1]; e < p->qedge_ptr[k]; e++)
            if (p->qedge[e] == (uint32_t)(k - 2)) found = 1;
        p->lunmat = found ? 0 : 1;
    }
    /* first_sym_level <= 2  <=>  SYM(1) non-empty (:555-562, 1299); with no
     * symmetry at all the reference reads an uninitialised value, taken here as
     * descending. */
    p->asc = (p->sym_ptr[2] - p->sym_ptr[1]) > 0;
    or_csr_free(&q);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Data preparation                                                          */
/* ------------------------------------------------------------------------ */

/* key(v) = (deg(v), v) ascending */
static int cmp_key_asc(const void *a, const void *b, void *ctx)
{
    const uint32_t *deg = ctx;
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    if (deg[x] != deg[y]) return deg[x] < deg[y] ? -1 : 1;
    return x < y ? -1 : x > y;
}

/* degree descending, id ascending */
static int cmp_key_desc(const void *a, const void *b, void *ctx)
{
    const uint32_t *deg = ctx;
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    if (deg[x] != deg[y]) return deg[x] > deg[y] ? -1 : 1;
    return x < y ? -1 : x > y;
}

int or_prepare(const or_csr *g, const or_plan *p, uint32_t cutoff, or_prep *out)
{
    memset(out, 0, sizeof(*out));
    uint32_t n = g->n;
    uint64_t m = g->m;
    /* The bitmap build binary-searches id-ordered rows (host_utils.cuh:284-316)
     * and the orientation keeps the input order among equal degrees
     * (subgraph_matching.cuh:1305-1331): rows must be sorted by id (-s). */
    for (uint32_t v = 0; v < n; v++)
        for (uint64_t e = g->rowptr[v] + 1; e < g->rowptr[v + 1]; e++)
            if (g->col[e - 1] > g->col[e]) return -3;
    uint64_t *rp = malloc((n + 1) * sizeof(uint64_t));
    uint32_t *col = malloc((m ? m : 1) * sizeof(uint32_t));
    uint32_t *deg = malloc((n ? n : 1) * sizeof(uint32_t));
    if (!rp || !col || !deg) return -1;
    memcpy(rp, g->rowptr, (n + 1) * sizeof(uint64_t));
    memcpy(col, g->col, m * sizeof(uint32_t));
    for (uint32_t v = 0; v < n; v++) deg[v] = (uint32_t)(rp[v + 1] - rp[v]);

    /* Peel (subgraph_matching.cuh:685-764): S = {v : 1 <= d < min_qdeg};
     * remove both directions of every edge at S (common_utils.cuh:217-246,
     * first occurrence of the reverse entry); repeat while |S| > 0.05 n. */
    unsigned char *keep = malloc(m ? m : 1);
    int rounds = 0;
    for (;;) {
        rounds++;
        uint64_t ns = 0;
        memset(keep, 1, m);
        for (uint32_t v = 0; v < n; v++) {
            if (deg[v] >= 1 && deg[v] < (uint32_t)p->min_qdeg) {
                ns++;
                for (uint64_t e = rp[v]; e < rp[v + 1]; e++) {
                    keep[e] = 0;
                    uint32_t w = col[e];
                    for (uint64_t f = rp[w]; f < rp[w + 1]; f++)
                        if (col[f] == v) { keep[f] = 0; break; }
                }
            }
        }
        uint64_t k = 0;
        uint64_t *nrp = malloc((n + 1) * sizeof(uint64_t));
        nrp[0] = 0;
        for (uint32_t v = 0; v < n; v++) {
            for (uint64_t e = rp[v]; e < rp[v + 1]; e++)
                if (keep[e]) col[k++] = col[e];
            nrp[v + 1] = k;
            deg[v] = (uint32_t)(nrp[v + 1] - nrp[v]);
        }
        free(rp);
        rp = nrp;
        m = k;
        if (!((double)ns > 0.05 * (double)n)) break;
    }
    free(keep);

    /* Orientation (subgraph_matching.cuh:1282-1336; host_utils.cuh:89-134,
     * 230-246): key(v) = (deg(v), v); rows ascending by key (asc) or by degree
     * descending with ids ascending (desc); nlow = #smaller keys. */
    uint32_t *ocol = malloc((m ? m : 1) * sizeof(uint32_t));
    uint32_t *nlow = malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t maxd = 0;
    for (uint32_t u = 0; u < n; u++) {
        uint64_t b = rp[u], e = rp[u + 1];
        uint32_t d = (uint32_t)(e - b);
        if (d > maxd) maxd = d;
        uint32_t lo = 0;
        for (uint64_t x = b; x < e; x++) {
            uint32_t w = col[x];
            if (deg[w] < deg[u] || (deg[w] == deg[u] && w < u)) lo++;
        }
        nlow[u] = lo;
        /* keys are distinct for distinct vertices, so any sort by key gives
         * the reference's stable order; equal entries are duplicates */
        for (uint64_t x = b; x < e; x++) ocol[x] = col[x];
        qsort_r(ocol + b, d, sizeof(uint32_t), p->asc ? cmp_key_asc : cmp_key_desc, deg);
    }

    /* Roots (subgraph_matching.cuh:1085-1131, 1167-1170, 1201;
     * device_utils.cuh:60-97): Q0 = {max_qdeg <= d < 100 cutoff} by degree
     * descending, ties by id ascending; boundary = #{d > cutoff};
     * R = Q0 \ Q0[0] iff asc and boundary >= 1. */
    uint32_t hi = cutoff * 100u;
    uint64_t nq = 0;
    for (uint32_t v = 0; v < n; v++)
        if (deg[v] >= (uint32_t)p->max_qdeg && deg[v] < hi) nq++;
    uint32_t *q0 = malloc((nq ? nq : 1) * sizeof(uint32_t));
    /* counting sort by degree descending keeps id order within a degree */
    {
        uint64_t c = 0;
        for (uint32_t v = 0; v < n; v++)
            if (deg[v] >= (uint32_t)p->max_qdeg && deg[v] < hi) q0[c++] = v;
        /* stable sort by degree descending: use a merge-free approach via
         * bucketing on degree */
        uint64_t *cnt = calloc((size_t)maxd + 2, sizeof(uint64_t));
        for (uint64_t i = 0; i < nq; i++) cnt[maxd - deg[q0[i]]]++;
        uint64_t s = 0;
        for (uint32_t t = 0; t <= maxd; t++) { uint64_t c2 = cnt[t]; cnt[t] = s; s += c2; }
        uint32_t *tmp = malloc((nq ? nq : 1) * sizeof(uint32_t));
        for (uint64_t i = 0; i < nq; i++) tmp[cnt[maxd - deg[q0[i]]]++] = q0[i];
        memcpy(q0, tmp, nq * sizeof(uint32_t));
        free(tmp);
        free(cnt);
    }
    uint64_t bnd = 0;
    while (bnd < nq && deg[q0[bnd]] > cutoff) bnd++;
    uint64_t skip = (p->asc && bnd >= 1) ? 1 : 0;
    out->nroots = nq - skip;
    out->roots = malloc((out->nroots ? out->nroots : 1) * sizeof(uint32_t));
    memcpy(out->roots, q0 + skip, out->nroots * sizeof(uint32_t));
    free(q0);

    out->n = n;
    out->m = m;
    out->rowptr = rp;
    out->col = col;
    out->ocol = ocol;
    out->nlow = nlow;
    out->q0_size = nq;
    out->boundary = bnd;
    out->cutoff = cutoff;
    out->max_deg = maxd;
    out->peel_rounds = rounds;
    out->asc = p->asc;
    free(deg);
    return 0;
}

void or_prep_free(or_prep *pp)
{
    free(pp->rowptr);
    free(pp->col);
    free(pp->ocol);
    free(pp->nlow);
    free(pp->roots);
    memset(pp, 0, sizeof(*pp));
}

/* ------------------------------------------------------------------------ */
/* Enumeration                                                               */
/* ------------------------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t or_tuple_hash(const uint32_t *t, int len)
{
    uint64_t h = 0x6a09e667f3bcc909ULL ^ (uint64_t)len;
    for (int i = 0; i < len; i++) h = mix64(h + 0x9e3779b97f4a7c15ULL * (uint64_t)(i + 1) + t[i]);
    return h;
}

typedef struct {
    const or_prep *g;
    const or_plan *p;
    int L;
    /* per-root state */
    uint32_t u;
    const uint32_t *A;
    uint32_t d, W;
    uint64_t *E;   /* d x W */
    uint64_t *C;   /* OR_MAXK x W candidate sets */
    uint32_t x[OR_MAXK];
    /* output */
    uint64_t count, hsum, hxor, emitted, written;
    uint32_t *buf;
    uint64_t cap;
    int want_hash;
} dfs_ctx;

static inline int bsearch_u32(const uint32_t *a, uint64_t n, uint32_t v)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (a[mid] == v) return 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return 0;
}

/* E[j][i] = [A[i] in N(A[j])] by binary search of A[i], i >= j, in the
 * id-sorted row of A[j]; both (j,i) and (i,j) set (device_utils.cuh:99-127;
 * host_utils.cuh:284-316). */
static void encode(dfs_ctx *c)
{
    const or_prep *g = c->g;
    memset(c->E, 0, (size_t)c->d * c->W * sizeof(uint64_t));
    for (uint32_t j = 0; j < c->d; j++) {
        uint32_t v = c->A[j];
        const uint32_t *B = g->col + g->rowptr[v];
        uint64_t bsz = g->rowptr[v + 1] - g->rowptr[v];
        for (uint32_t i = j; i < c->d; i++) {
            if (bsearch_u32(B, bsz, c->A[i])) {
                c->E[(size_t)j * c->W + i / 64] |= 1ULL << (i % 64);
                c->E[(size_t)i * c->W + j / 64] |= 1ULL << (j % 64);
            }
        }
    }
}

static void emit_level(dfs_ctx *c, const uint64_t *set, int lastpos)
{
    /* tuples (u, A[x1], ..., A[x_{lastpos-1}], A[i]) for i in set */
    uint32_t t[OR_MAXK];
    t[0] = c->u;
    for (int q = 1; q < lastpos; q++) t[q] = c->A[c->x[q]];
    for (uint32_t w = 0; w < c->W; w++) {
        uint64_t b = set[w];
        while (b) {
            uint32_t i = w * 64 + (uint32_t)__builtin_ctzll(b);
            b &= b - 1;
            t[lastpos] = c->A[i];
            c->emitted++;
            if (c->want_hash) {
                uint64_t h = or_tuple_hash(t, lastpos + 1);
                c->hsum += h;
                c->hxor ^= h;
            }
            if (c->buf && c->written < c->cap) {
                memcpy(c->buf + c->written * (uint64_t)(lastpos + 1), t, (lastpos + 1) * sizeof(uint32_t));
                c->written++;
            }
        }
    }
}

static uint64_t popc_set(const uint64_t *s, uint32_t W)
{
    uint64_t r = 0;
    for (uint32_t w = 0; w < W; w++) r += (uint64_t)__builtin_popcountll(s[w]);
    return r;
}

/* C_l for l >= 3 (compute_intersection, device_utils.cuh:426-456): row0 minus
 * x_1..x_{l-1}, AND E[x_q] for q in QEDGE(l)[1:], keep i > x_s for s in SYM(l)
 * with s > 0. */
static void level_set(dfs_ctx *c, int l, uint64_t *out)
{
    const or_plan *p = c->p;
    for (uint32_t w = 0; w < c->W; w++) {
        uint64_t t;
        uint32_t lo = w * 64;
        if (c->d >= lo + 64) t = ~0ULL;
        else t = (c->d > lo) ? ((1ULL << (c->d - lo)) - 1) : 0;
        for (int q = 1; q < l; q++)
            if (c->x[q] / 64 == w) t &= ~(1ULL << (c->x[q] % 64));
        for (uint32_t e = p->qedge_ptr[l] + 1; e < p->qedge_ptr[l + 1]; e++)
            t &= c->E[(size_t)c->x[p->qedge[e]] * c->W + w];
        for (uint32_t e = p->sym_ptr[l]; e < p->sym_ptr[l + 1]; e++) {
            uint32_t s = p->sym[e];
            if (s > 0) {
                uint32_t xs = c->x[s];
                /* keep bits > xs; bits <= xs cleared (get_mask semantics) */
                if (xs / 64 > w) t = 0;
                else if (xs / 64 == w) {
                    uint32_t r = xs % 64;
                    t &= (r == 63) ? 0 : ~((1ULL << (r + 1)) - 1);
                }
            }
        }
        out[w] = t;
    }
}

static void dfs(dfs_ctx *c, int l)
{
    /* choose x_{l-1} from C_{l-1} in ascending order, compute C_l */
    uint64_t *prev = c->C + (size_t)(l - 1) * c->W;
    uint64_t *cur = c->C + (size_t)l * c->W;
    for (uint32_t w = 0; w < c->W; w++) {
        uint64_t b = prev[w];
        while (b) {
            uint32_t i = w * 64 + (uint32_t)__builtin_ctzll(b);
            b &= b - 1;
            c->x[l - 1] = i;
            level_set(c, l, cur);
            if (l == c->L - 1) {
                c->count += popc_set(cur, c->W);
                if (c->want_hash || c->buf) emit_level(c, cur, l);
            } else {
                dfs(c, l + 1);
            }
        }
    }
}

static void enum_root(dfs_ctx *c, uint32_t u)
{
    const or_prep *g = c->g;
    const or_plan *p = c->p;
    c->u = u;
    c->A = g->ocol + g->rowptr[u];
    c->d = (uint32_t)(g->rowptr[u + 1] - g->rowptr[u]);
    c->W = (c->d + 63) / 64;
    if (c->d == 0) return;
    encode(c);
    uint32_t split = g->nlow[u];
    int skip_asc = (p->sym_ptr[2] == 1);
    int q2 = (int)(p->qedge_ptr[3] - p->qedge_ptr[2]) == 2;
    int s2 = 0;
    for (uint32_t e = p->sym_ptr[2]; e < p->sym_ptr[3]; e++)
        if (p->sym[e] > 0) s2 = 1;
    uint64_t *C2 = c->C + 2 * (size_t)c->W;
    for (uint32_t j = 0; j < c->d; j++) {
        if (skip_asc && j < split) continue; /* sgm_kernelsLD.cuh:60 */
        c->x[1] = j;
        if (c->L <= 2) continue; /* nothing is ever counted (A.6, L = 2) */
        /* level 2 (count_tri, device_utils.cuh:179-209) */
        for (uint32_t w = 0; w < c->W; w++) {
            uint32_t lo = w * 64;
            uint64_t row0;
            if (c->d >= lo + 64) row0 = ~0ULL;
            else row0 = (1ULL << (c->d - lo)) - 1;
            if (j / 64 == w) row0 &= ~(1ULL << (j % 64));
            uint64_t t = q2 ? c->E[(size_t)j * c->W + w] : row0;
            if (s2) {
                uint64_t below; /* bits < j in word w */
                if (j / 64 > w) below = ~0ULL;
                else if (j / 64 < w) below = 0;
                else below = (j % 64) ? ((1ULL << (j % 64)) - 1) : 0;
                t &= ~(row0 & below);
            }
            C2[w] = t;
        }
        if (c->L == 3) {
            c->count += popc_set(C2, c->W);
            if (c->want_hash || c->buf) emit_level(c, C2, 2);
        } else {
            dfs(c, 3);
        }
    }
}

/* Roots roots[r] for r in [r0, r1), or roots[ranks[i]] for i in [r0, r1)
 * when ranks is given. */
static int run_roots(const or_prep *pp, const or_plan *p, const uint64_t *ranks, uint64_t r0, uint64_t r1,
                     int nthreads, int want_hash, uint32_t *match_buf, uint64_t match_cap, or_result *res)
{
    memset(res, 0, sizeof(*res));
    int L = p->k - p->lunmat;
    uint32_t maxd = 0;
    for (uint64_t r = r0; r < r1; r++) {
        uint32_t u = pp->roots[ranks ? ranks[r] : r];
        uint32_t d = (uint32_t)(pp->rowptr[u + 1] - pp->rowptr[u]);
        if (d > maxd) maxd = d;
    }
    uint32_t maxW = (maxd + 63) / 64;
    if (match_buf) nthreads = 1;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    uint64_t cnt = 0, hs = 0, hx = 0, em = 0, wr = 0;
    int fail = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : cnt, hs, em, wr) reduction(^ : hx) reduction(| : fail)
    {
        dfs_ctx c;
        memset(&c, 0, sizeof(c));
        c.g = pp;
        c.p = p;
        c.L = L;
        c.want_hash = want_hash;
        c.buf = match_buf;
        c.cap = match_cap;
        c.E = malloc(((size_t)maxd * maxW + 1) * sizeof(uint64_t));
        c.C = malloc(((size_t)(OR_MAXK + 1) * maxW + 1) * sizeof(uint64_t));
        if (!c.E || !c.C) fail = 1;
        else {
#pragma omp for schedule(dynamic, 1)
            for (uint64_t r = r0; r < r1; r++) enum_root(&c, pp->roots[ranks ? ranks[r] : r]);
        }
        free(c.E);
        free(c.C);
        cnt += c.count;
        hs += c.hsum;
        hx ^= c.hxor;
        em += c.emitted;
        wr += c.written;
    }
    if (fail) return -1;
    res->count = cnt;
    res->hash_sum = hs;
    res->hash_xor = hx;
    res->n_emitted = em;
    res->n_written = wr;
    return 0;
}

int or_count(const or_prep *pp, const or_plan *p, int nthreads, int want_hash,
             uint32_t *match_buf, uint64_t match_cap, or_result *res)
{
    return run_roots(pp, p, NULL, 0, pp->nroots, nthreads, want_hash, match_buf, match_cap, res);
}

int or_count_range(const or_prep *pp, const or_plan *p, uint64_t r0, uint64_t r1, int nthreads,
                   or_result *res)
{
    if (r1 > pp->nroots) r1 = pp->nroots;
    if (r0 > r1) r0 = r1;
    return run_roots(pp, p, NULL, r0, r1, nthreads, 0, NULL, 0, res);
}

int or_count_ranks(const or_prep *pp, const or_plan *p, const uint64_t *ranks, uint64_t n, int nthreads,
                   int want_hash, or_result *res)
{
    for (uint64_t i = 0; i < n; i++)
        if (ranks[i] >= pp->nroots) return -1;
    return run_roots(pp, p, ranks, 0, n, nthreads, want_hash, NULL, 0, res);
}

uint64_t or_alg_bytes(const or_prep *pp)
{
    uint64_t b = 0;
    for (uint64_t r = 0; r < pp->nroots; r++) {
        uint32_t u = pp->roots[r];
        uint64_t d = pp->rowptr[u + 1] - pp->rowptr[u];
        b += 4 * (d + 2);
        for (uint64_t e = pp->rowptr[u]; e < pp->rowptr[u + 1]; e++) {
            uint32_t v = pp->col[e];
            b += 4 * ((pp->rowptr[v + 1] - pp->rowptr[v]) + 2);
        }
    }
    return b;
}
