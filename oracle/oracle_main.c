/*
 * oracle_main.c -- command-line driver for the CPU restatement (test
 * infrastructure only).  Mirrors the reference CLI surface for the path
 * (src/main.cu:27-344; include/Config.h:336-429): -g data graph, -t query
 * .mtx, -m sgm|sgc, -s sort edges; -n threads selects OpenMP threads.
 * Prints the reference's final line "------------- Counter = N"
 * (subgraph_matching.cuh:1277) and a timing line for enumeration only.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

static int cmp_edge(const void *a, const void *b)
{
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

int main(int argc, char **argv)
{
    const char *gpath = NULL, *tpath = NULL;
    int sgc = 0, sortE = 0, nth = 0, opt;
    while ((opt = getopt(argc, argv, "g:t:m:s:n:")) >= 0) {
        switch (opt) {
        case 'g': gpath = optarg; break;
        case 't': tpath = optarg; break;
        case 'm':
            if (!strcmp(optarg, "sgc")) sgc = 1;
            else if (!strcmp(optarg, "sgm")) sgc = 0;
            else { fprintf(stderr, "unknown -m %s\n", optarg); return 2; }
            break;
        case 's': sortE = 1; break;
        case 'n': nth = atoi(optarg); break;
        default: return 2;
        }
    }
    if (!gpath || !tpath) { fprintf(stderr, "usage: -g graph -t query.mtx -m sgm|sgc\n"); return 2; }
    uint32_t *s, *d;
    int64_t ne = or_read_edges(gpath, &s, &d);
    if (ne < 0) { fprintf(stderr, "cannot read %s\n", gpath); return 1; }
    if (sortE) {
        uint64_t *k = malloc((size_t)ne * 8);
        for (int64_t i = 0; i < ne; i++) k[i] = ((uint64_t)s[i] << 32) | d[i];
        qsort(k, (size_t)ne, 8, cmp_edge);
        for (int64_t i = 0; i < ne; i++) { s[i] = (uint32_t)(k[i] >> 32); d[i] = (uint32_t)k[i]; }
        free(k);
    }
    or_csr g;
    if (or_csr_build(s, d, (uint64_t)ne, &g)) { fprintf(stderr, "bad edge list\n"); return 1; }
    free(s); free(d);
    double mean, sd;
    uint32_t cutoff = or_cutoff(&g, &mean, &sd);
    uint32_t *qs, *qd;
    int64_t qe = or_read_edges(tpath, &qs, &qd);
    if (qe < 0) { fprintf(stderr, "cannot read %s\n", tpath); return 1; }
    or_plan p;
    if (or_plan_build(qs, qd, (uint64_t)qe, sgc, &p)) { fprintf(stderr, "bad query\n"); return 1; }
    or_prep pp;
    if (or_prepare(&g, &p, cutoff, &pp)) { fprintf(stderr, "prepare failed\n"); return 1; }
    double t0 = omp_get_wtime();
    or_result r;
    if (or_count(&pp, &p, nth, 0, NULL, 0, &r)) { fprintf(stderr, "count failed\n"); return 1; }
    double t1 = omp_get_wtime();
    printf("n=%u m=%llu cutoff=%u roots=%llu threads=%d enum_s=%.6f\n", g.n,
           (unsigned long long)g.m, cutoff, (unsigned long long)pp.nroots,
           nth > 0 ? nth : omp_get_max_threads(), t1 - t0);
    printf("------------- Counter = %llu\n", (unsigned long long)r.count);
    or_prep_free(&pp);
    or_csr_free(&g);
    return 0;
}
