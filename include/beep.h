/*
 * beep.h -- C ABI of the MI355X-native subgraph-enumeration path.
 *
 * The reference (researchgroup-zx93920/BEEP) is a monolithic CUDA executable;
 * its path sits behind three nested boundaries, each replaced here:
 *   1. the CLI contract  (src/main.cu:27-344, include/Config.h:336-429)
 *        -> beep_amd/bin/beep, a thin client of this ABI;
 *   2. the in-process C++ boundary graph::SG_Match<uint>(task, by, dev, cutoff,
 *      ndev).run(COOCSRGraph_d&, COOCSRGraph&) (subgraph_matching.cuh:113,178)
 *        -> beep_plan_* + beep_prepare + beep_count below;
 *   3. the kernel ABI sgm_kernel_central_node_function<uint,320,16>(GLOBAL_HANDLE,
 *      queue...) plus __constant__ tables (sgm_kernelsLD.cuh:7-11,
 *      common_utils.cuh:6-16, host_utils.cuh:12-24)
 *        -> internal to libbeep.so (beep_amd/csrc/sgm_kernels.hip).
 *
 * Conventions: opaque handles with caller-owned *_free; every entry point
 * returns 0 on success and a negative errno-style code otherwise (never exits);
 * beep_last_error() describes the last failure of the calling thread.  Handles
 * are not thread-safe except a beep_server, whose submit/poll/stats are.
 * Device memory belongs to the handle that allocated it.
 */
#ifndef BEEP_H
#define BEEP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BEEP_OK 0
#define BEEP_EIO (-5)
#define BEEP_ENOMEM (-12)
#define BEEP_EDEVICE (-19)
#define BEEP_EINVAL (-22)
#define BEEP_ERANGE (-34)
#define BEEP_ENOSYS (-38)
#define BEEP_ETIMEDOUT (-110)

#define BEEP_MODE_SGM 0 /* -m sgm, Config.h:80-81 */
#define BEEP_MODE_SGC 1 /* -m sgc, Config.h:83-84 */

#define BEEP_MAXK 10 /* DEPTH, host_utils.cuh:10 */

/* Emission of the last materialised level (sgm: full matches). */
#define BEEP_EMIT_NONE 0   /* count only (the reference's behaviour)          */
#define BEEP_EMIT_HASH 1   /* + order-independent multiset hash of the tuples */
#define BEEP_EMIT_BUFFER 2 /* + hash + first match_cap tuples into a buffer   */

typedef struct beep_graph beep_graph;
typedef struct beep_plan beep_plan;
typedef struct beep_prep beep_prep;
typedef struct beep_server beep_server;

/* ------------------------------------------------------------------------ */
/* Graph loading.  Replaces EdgeListFile::get_edges + CSRCOO::from_edgelist +  */
/* get_stats (main.cu:70-132; FIleReader.cuh:52-95,101-127,130-212,299-389;    */
/* CSRCOO.cuh:37-80; utils.cuh:186-261).                                       */
/* ------------------------------------------------------------------------ */

/* .bel / .tsv / .mtx chosen by extension; sort_edges mirrors -s (main.cu:88-91). */
int beep_graph_load(const char *path, int sort_edges, beep_graph **out);

/* Edges in CSR-build order (grouped by ascending src); self loops dropped. */
int beep_graph_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t n_edges,
                          beep_graph **out);

/* Deterministic synthetic generators (symmetrised, no self loops or
 * duplicates, rows sorted).  R-MAT draws `samples` edges with quadrant
 * probabilities (a,b,c, d = 1-a-b-c); ids >= n_keep are rejected when
 * n_keep < 2^scale.  Chung-Lu draws `samples` endpoint pairs with weights
 * w_i proportional to (i + i0)^(-1/(gamma-1)). */
int beep_graph_rmat(uint32_t scale, uint64_t samples, double a, double b, double c,
                    uint64_t seed, uint32_t n_keep, beep_graph **out);
int beep_graph_chung_lu(uint32_t n, uint64_t samples, double gamma, double i0, uint64_t seed,
                        beep_graph **out);

typedef struct {
    uint32_t n;          /* rows (max src + 1) */
    uint64_t m;          /* stored directed entries */
    uint32_t max_degree;
    uint32_t cutoff;     /* ceil(mean + 6 sd), utils.cuh:260 */
    double mean, sd;
    const uint64_t *rowptr; /* host views, valid while the handle lives */
    const uint32_t *col;
} beep_graph_desc;

int beep_graph_describe(const beep_graph *g, beep_graph_desc *out);
/* Writes 24-byte {dst, src, 0} records (FIleReader.cuh:613-637). */
int beep_graph_write_bel(const beep_graph *g, const char *path);
void beep_graph_free(beep_graph *g);

/* Format converters, -m mtx-bel | tsv-bel | tsv-mtx | bel-mtx | txt-bel
 * (main.cu:40-68; MtB_Writer, FIleReader.cuh:459-912), byte for byte as the
 * reference writes them.  Replaces MtB_Writer::write_market_bel /
 * write_tsv_bel / write_tsv_market / write_bel_market / write_txt_bel. */
#define BEEP_CONV_MTX_BEL 1
#define BEEP_CONV_TSV_BEL 2
#define BEEP_CONV_TSV_MTX 3
#define BEEP_CONV_BEL_MTX 4
#define BEEP_CONV_TXT_BEL 5
int beep_convert(int kind, const char *src_path, const char *dst_path);

/* ------------------------------------------------------------------------ */
/* Query plan.  Replaces SG_Match::preprocess_query / detect_symmetry /       */
/* detect_reuse / check_unmat (subgraph_matching.cuh:257-631), with its own   */
/* automorphism-orbit search in place of nauty's densenauty.                  */
/* ------------------------------------------------------------------------ */

int beep_plan_from_mtx(const char *path, int mode, beep_plan **out);
int beep_plan_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t n_edges, int mode,
                         beep_plan **out);

typedef struct {
    int k, mode, lunmat, asc, max_qdeg, min_qdeg;
    uint32_t seq[BEEP_MAXK];
    uint32_t qdeg[BEEP_MAXK];
    uint32_t qedge_ptr[BEEP_MAXK + 1];
    uint32_t qedge[BEEP_MAXK * (BEEP_MAXK - 1) / 2];
    uint32_t sym_ptr[BEEP_MAXK + 1];
    uint32_t sym[BEEP_MAXK * (BEEP_MAXK - 1) / 2];
    /* Every bitmap row a level intersects with is used above its own position
     * only (through the symmetry bounds), so the bitmaps keep, of each adjacent
     * pair, only the bit in the row of the smaller position (cliques, for
     * instance). */
    int upper_rows;
} beep_plan_desc;

int beep_plan_describe(const beep_plan *p, beep_plan_desc *out);
void beep_plan_free(beep_plan *p);

/* ------------------------------------------------------------------------ */
/* Data preparation.  Replaces initialize / peel_data / degree_ordering and  */
/* the root queue of count_subgraphs (subgraph_matching.cuh:633-764,          */
/* 1080-1131, 1282-1336; GraphQueue.cuh:85-120).  The graph handle is not    */
/* mutated (the reference swaps arrays inside COOCSRGraph_d, :746-750).       */
/* ------------------------------------------------------------------------ */

int beep_prepare(const beep_graph *g, const beep_plan *p, beep_prep **out);

typedef struct {
    uint32_t n;
    uint64_t m;              /* peeled entries */
    uint64_t n_roots;        /* |R| */
    uint64_t q0_size;        /* |Q0| */
    uint64_t boundary;       /* #{d > cutoff} in Q0 */
    uint32_t cutoff, max_degree, max_root_degree;
    int peel_rounds, asc;
    uint64_t alg_bytes;      /* B_alg */
    uint64_t n_hub_roots;    /* roots whose bitmap lives in HBM */
    uint64_t n_lds_batches;  /* LDS-resident root batches */
    const uint64_t *rowptr;  /* host views of the prepared graph */
    const uint32_t *col;     /* peeled, id ascending */
    const uint32_t *ocol;    /* peeled, degree oriented */
    const uint32_t *nlow;    /* #neighbours with smaller (degree, id) key */
    const uint32_t *roots;   /* R in processing order */
    /* Preconditions under which the counts are the true counts of the
     * pattern (count_subgraphs' root window and head skip,
     * subgraph_matching.cuh:1052, 1081-1124, 1201; device_utils.cuh:66-94):
     *   p1: max peeled degree < 100 * cutoff, i.e. no vertex is left out of
     *       the root window by its upper bound;
     *   p2: the head skip drops nothing -- the plan is not ascending, or
     *       boundary = 0, or Q0[0] has no neighbour with a larger key. */
    int p1, p2;
    uint32_t q0_head;        /* Q0[0] (UINT32_MAX when Q0 is empty) */
    uint32_t q0_head_larger; /* neighbours of Q0[0] with a larger (degree, id) key */
    /* Per lane class lg (LDS batches of roots with at most 64 << lg row
     * entries; one lds_batch_kernel launch each): roots, bitmap-build work
     * units (scan-range entries streamed, or root neighbours looked up) and
     * the bytes the launch must read at least once (root records and rows,
     * row-vertex records and scan ranges, batch arrays). */
    uint64_t lds_class_roots[8], lds_class_units[8], lds_class_bytes[8];
} beep_prep_desc;

int beep_prep_describe(const beep_prep *pp, beep_prep_desc *out);
/* The same query over a subset of pp's roots: ranks index beep_prep_desc.roots
 * (R in processing order) and must be strictly ascending.  The result shares
 * pp's peeled, oriented graph (and its HBM copy).  The reference has no such
 * call: its launcher takes every root of R (subgraph_matching.cuh:1080-1131);
 * its multi-GPU split hands each device a part of R the same way
 * (sgm_kernelsLD.cuh:60, subgraph_matching.cuh:1167-1170). */
int beep_prep_subset(const beep_prep *pp, const beep_plan *p, const uint64_t *ranks, uint64_t n,
                     beep_prep **out);
/* Copies the prepared graph to HBM on `device` (idempotent). */
int beep_prep_upload(beep_prep *pp, int device);
void beep_prep_free(beep_prep *pp);

/* ------------------------------------------------------------------------ */
/* Enumeration.  Replaces count_subgraphs' launch of                          */
/* sgm_kernel_central_node_function + final_counter + the Counter print       */
/* (subgraph_matching.cuh:1138-1277; sgm_kernelsLD.cuh:7-273).                */
/* ------------------------------------------------------------------------ */

typedef struct {
    int device;             /* HIP device ordinal */
    int emit;               /* BEEP_EMIT_* */
    uint64_t match_cap;     /* BEEP_EMIT_BUFFER: tuples to keep */
    uint32_t *match_buf;    /* host buffer of match_cap * (k - lunmat) u32, or NULL */
    int lanes;              /* 0 = automatic, else lanes per DFS group (1..64) */
    int donate;             /* 0 = off; 1 = donate level-2 subtrees of hub roots
                               (count-only 5-vertex plans with paired last levels,
                               such as W5: ranges of the paired pass instead) */
    uint32_t donate_min;    /* donation threshold: min |C_2| of a donated subtree
                               (not used by the paired-pass ranges) */
    int timing;             /* 1 = record HIP events around the enumeration */
    double deadline_ms;     /* > 0: stop the enumeration this long after it starts on
                               the device; the result is then partial and marked stopped */
    int profile;            /* 1 = busy-time balance of workgroups / waves (beep_count) */
    int build_mode;         /* BEEP_BUILD_*: how LDS bitmaps find adjacent row pairs */
} beep_opts;

#define BEEP_BUILD_AUTO 0   /* stream ranges, look up the root's neighbours when a
                               range is over 32x the root's degree */
#define BEEP_BUILD_SCAN 1   /* always stream the row vertices' larger-key ranges */
#define BEEP_BUILD_SEARCH 2 /* always binary-search the root's neighbours */

typedef struct {
    uint64_t count;         /* "------------- Counter = N" */
    uint64_t hash_sum, hash_xor;
    uint64_t n_emitted, n_written;
    double t_enum_ms;       /* first enumeration launch -> counter readback */
    double t_hub_ms, t_lds_ms; /* device time of the hub and LDS phases */
    uint64_t donations;     /* subtrees donated through the ring */
    uint64_t stopped;       /* 1: the deadline passed; counts are partial */
    /* with opts.profile (the reference's TIMER build, sgm_kernelsLD_timer.cuh:7-64,
     * 339-349, as busy-time balance): mean / max busy time over the LDS
     * workgroups and over the hub-DFS waves (1 = perfectly even), and the
     * share of hub-DFS busy time spent on donated subtrees, and the share of
     * LDS workgroup time spent staging rows and building bitmaps */
    double balance_lds, balance_hub, donated_frac, lds_build_frac;
    /* device time of each lane class's lds_batch_kernel launch (beep_count; 0: none) */
    double t_lds_class_ms[8];
    /* with opts.profile: mean / max busy time over the compute units that ran
     * LDS workgroups or hub waves (the reference's per-SM utilisation,
     * sgm_kernelsLD_timer.cuh:339-349), and how many compute units did */
    double balance_cu;
    uint64_t cus_busy;
} beep_result;

void beep_opts_default(beep_opts *o);
/* Returns BEEP_ETIMEDOUT (with the partial result filled in) when the deadline passed. */
int beep_count(beep_prep *pp, const beep_plan *p, const beep_opts *o, beep_result *out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code running under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc: beep_count with every scratch buffer of the launch
 * sequence (counters, work cursors, hub bitmap arena and hash, DFS stacks,
 * donation ring, tuple buffer) carved from one fixed device array of
 * BEEP_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on
 * until the process ends.  It calls no allocator.  pp must already be
 * resident on o->device (beep_prep_upload).  The work runs on `stream` (a
 * hipStream_t; NULL: the null stream), LDS batches before hub rounds, and the
 * call waits for it; t_enum_ms is host wall time, t_hub_ms / t_lds_ms are 0.
 * A query whose counters, hub arena, stacks and tuple buffer exceed the array
 * fails with BEEP_ERANGE; the donation ring gets what is left (a hub item
 * whose subtree finds the ring full is walked where it is).  opts.profile is
 * refused (BEEP_EINVAL).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU race and must be serialised by the caller. */
#define BEEP_NOALLOC_SCRATCH_BYTES (64ull << 20)
int beep_count_noalloc(beep_prep *pp, const beep_plan *p, const beep_opts *o, void *stream,
                       beep_result *out);

/* ------------------------------------------------------------------------ */
/* Batches.  No counterpart in the reference (one query per process there):  */
/* several prepared queries over resident graphs run as one captured         */
/* hipGraph -- one persistent kernel over every query's LDS batches, then    */
/* per query its hub rounds on a branch of their own.  The graph is captured */
/* on the first run and replayed afterwards.                                 */
/* ------------------------------------------------------------------------ */

typedef struct beep_batch beep_batch;

/* opts: n entries or NULL (defaults); every query must use the same device;
 * BEEP_EMIT_BUFFER is not supported in a batch. */
int beep_batch_create(beep_prep *const *preps, const beep_plan *const *plans, const beep_opts *opts,
                      int n, beep_batch **out);
/* results: n entries (t_enum_ms is the time of the whole graph launch). */
int beep_batch_run(beep_batch *b, beep_result *results, double *t_ms);
void beep_batch_free(beep_batch *b);

/* ------------------------------------------------------------------------ */
/* Serving.  No counterpart in the reference (it reloads and re-prepares the */
/* graph for every query, main.cu:70-341).  A server keeps one data graph    */
/* resident: peeled/oriented variants and prepared root queues are cached in */
/* HBM and shared by every query that needs them.  Submitted queries wait in */
/* arrival order; the oldest is admitted as soon as a lane (two streams) is  */
/* free and its predicted scratch footprint -- hub bitmap arena, hash        */
/* tables, DFS stacks, donation ring, tuple buffer -- fits the free part of  */
/* the server's HBM pool.  Admitted queries run concurrently.  submit, poll  */
/* and stats may be called from any thread.                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    int device;
    int lanes;              /* queries running at once, 1..16 */
    uint64_t pool_bytes;    /* scratch pool; 0 = a quarter of the free HBM */
    uint64_t arena_cap;     /* hub-arena bytes per query; 0 = pool / (2 * lanes) */
    double budget_ms;       /* > 0: predict each new plan's device time with a pilot run
                               on a stratified sample of its root queue (1/pilot_stride of
                               the LDS roots, 8x more of the hub roots; it runs alone) and
                               refuse, with BEEP_ERANGE, plans predicted above 1.5 budget_ms */
    uint32_t pilot_stride;  /* sample stride of the pilot (0 = 32) */
} beep_server_opts;

typedef struct {
    uint64_t ticket;
    int status;             /* 0, BEEP_ETIMEDOUT (deadline passed, partial counts),
                               BEEP_ERANGE (predicted above the budget, not run) or an error */
    beep_result result;     /* t_enum_ms: device time of the query's launch sequence */
    double queue_ms;        /* submit -> admission */
    double service_ms;      /* admission -> completion */
    double latency_ms;      /* submit -> completion */
    uint64_t footprint;     /* predicted scratch bytes the query was admitted with */
    double predicted_ms;    /* pilot prediction of the device time (< 0: none) */
} beep_server_done;

typedef struct {
    uint64_t submitted, completed, stopped, failed;
    uint64_t pending, running, peak_running;
    uint64_t pool_bytes, pool_used, pool_peak;
    uint64_t variants, root_queues;
    uint64_t refused, pilots;   /* refused by the prediction; pilot runs */
} beep_server_stats;

void beep_server_opts_default(beep_server_opts *o);
/* g must outlive the server. */
int beep_server_create(const beep_graph *g, const beep_server_opts *o, beep_server **out);
/* Builds (if needed) and uploads the variant and root queue p needs. */
int beep_server_prepare(beep_server *s, const beep_plan *p);
/* o: emit, match_cap, match_buf (kept valid by the caller until the ticket is
 * polled), lanes, donate, donate_min, deadline_ms; o->device is ignored. */
int beep_server_submit(beep_server *s, const beep_plan *p, const beep_opts *o, uint64_t *ticket);
/* Takes one finished query; BEEP_ETIMEDOUT when none finished within
 * timeout_ms (< 0: wait without limit). */
int beep_server_poll(beep_server *s, beep_server_done *out, int timeout_ms);
int beep_server_stats_get(beep_server *s, beep_server_stats *out);
/* Waits for every submitted query, then releases the server's memory. */
void beep_server_free(beep_server *s);

/* Tuple hash used by BEEP_EMIT_HASH: sum and xor over emitted tuples. */
uint64_t beep_tuple_hash(const uint32_t *t, int len);

const char *beep_last_error(void);
const char *beep_version(void);

#ifdef __cplusplus
}
#endif
#endif
