/*
 * oracle.h -- CPU restatement of BEEP's subgraph-enumeration path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (beep_amd/, include/, the
 * CLI) links, loads or calls this code.  It is used by tests/ as the checker,
 * by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg
 * as the timed CPU baseline ("port").
 *
 * Every function restates the reference's behaviour from its source, cited as
 * file:line relative to the reference tree (researchgroup-zx93920/BEEP).
 *
 * Pinning (what checks each part independently of this code):
 *   or_read_edges ....... hand-written byte fixtures (tests/golden/io_*.json)
 *   or_csr_build ........ hand-written edge lists with self-loops / duplicates
 *   or_cutoff ........... closed form mean + 6*sd on tiny degree sequences
 *   or_plan_build ....... hand-derived query-plan tables for the named patterns
 *                         and a brute-force automorphism check in Python
 *   or_prepare .......... small hand graphs with known peel rounds / splits /
 *                         boundary, plus networkx-free brute force in Python
 *   or_count ............ brute-force monomorphism counts / |Aut| on tiny random
 *                         graphs (all 207 central-node patterns on 3-7 vertices),
 *                         closed forms (triangles = tr(A^3)/6, k-star sums,
 *                         clique counts) on mid-size graphs, and invariants.
 *   nauty orbit chain with ptn[n-1]=1 (subgraph_matching.cuh:463-488) and the
 *   atomic-append tie order of the root queue (common_utils.cuh:201-215) are
 *   parity unpinned: this oracle assumes stabiliser orbits and id-ascending ties.
 */
#ifndef BEEP_ORACLE_H
#define BEEP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXK 10 /* DEPTH, subgraph_matching/include/host_utils.cuh:10 */

typedef struct {
    uint32_t n;       /* number of rows = max src + 1 */
    uint64_t m;       /* number of stored (directed) entries */
    uint64_t *rowptr; /* n + 1 */
    uint32_t *col;    /* m */
} or_csr;

typedef struct {
    int k;          /* query vertices (KCCOUNT) */
    int sgc;        /* 1 = -m sgc, 0 = -m sgm */
    int lunmat;     /* LUNMAT */
    int asc;        /* ascending symmetry breaking (first_sym_level <= 2) */
    int max_qdeg, min_qdeg;
    uint32_t seq[OR_MAXK];
    uint32_t qdeg[OR_MAXK];
    uint32_t qedge_ptr[OR_MAXK + 1];
    uint32_t qedge[OR_MAXK * (OR_MAXK - 1) / 2];
    uint32_t sym_ptr[OR_MAXK + 1];
    uint32_t sym[OR_MAXK * (OR_MAXK - 1) / 2];
} or_plan;

typedef struct {
    uint32_t n;
    uint64_t m;         /* peeled entries */
    uint64_t *rowptr;   /* peeled row pointers, n + 1 */
    uint32_t *col;      /* peeled rows, id ascending */
    uint32_t *ocol;     /* peeled rows, degree oriented */
    uint32_t *nlow;     /* #neighbours with smaller key (deg, id) */
    uint32_t *roots;    /* root set R in processing order */
    uint64_t nroots;
    uint64_t q0_size;   /* |Q0| before the head skip */
    uint64_t boundary;  /* #roots in Q0 with degree > cutoff */
    uint32_t cutoff;
    uint32_t max_deg;   /* max peeled degree */
    int peel_rounds;
    int asc;
} or_prep;

typedef struct {
    uint64_t count;
    uint64_t hash_sum;  /* sum over emitted tuples of or_tuple_hash */
    uint64_t hash_xor;  /* xor over emitted tuples of or_tuple_hash */
    uint64_t n_emitted; /* tuples enumerated at the last materialised level */
    uint64_t n_written; /* tuples written to the caller's buffer (<= cap) */
} or_result;

/* ---- graph I/O (include/FIleReader.cuh) ---- */
int64_t or_read_edges(const char *path, uint32_t **src, uint32_t **dst);
void or_free(void *p);

/* ---- CSR + cutoff (include/CSRCOO.cuh, include/utils.cuh) ---- */
int or_csr_build(const uint32_t *src, const uint32_t *dst, uint64_t ne, or_csr *g);
void or_csr_free(or_csr *g);
uint32_t or_cutoff(const or_csr *g, double *mean_out, double *sd_out);

/* ---- query planner (subgraph_matching.cuh:257-631) ---- */
int or_plan_build(const uint32_t *src, const uint32_t *dst, uint64_t ne, int sgc, or_plan *p);

/* ---- data preparation (subgraph_matching.cuh:66-109, 633-764, 1037-1131, 1282-1336) ---- */
int or_prepare(const or_csr *g, const or_plan *p, uint32_t cutoff, or_prep *out);
void or_prep_free(or_prep *pp);

/* ---- enumeration (sgm_kernelsLD.cuh:9-131, device_utils.cuh:60-456) ----
 * nthreads <= 0 uses the OpenMP default.  When match_buf is non-NULL the run is
 * single threaded and tuples are written in root / DFS order (L words each).   */
int or_count(const or_prep *pp, const or_plan *p, int nthreads, int want_hash,
             uint32_t *match_buf, uint64_t match_cap, or_result *res);

/* Count over an explicit subset of the root list (rank range [r0, r1)). */
int or_count_range(const or_prep *pp, const or_plan *p, uint64_t r0, uint64_t r1,
                   int nthreads, or_result *res);

/* Count over the roots at the given ranks of the root list. */
int or_count_ranks(const or_prep *pp, const or_plan *p, const uint64_t *ranks, uint64_t n, int nthreads,
                   int want_hash, or_result *res);

uint64_t or_tuple_hash(const uint32_t *t, int len);

/* B_alg = sum_{u in R} [4(d(u)+2) + sum_{v in N(u)} 4(d(v)+2)], peeled degrees. */
uint64_t or_alg_bytes(const or_prep *pp);

#ifdef __cplusplus
}
#endif
#endif
