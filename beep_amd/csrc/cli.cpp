// cli.cpp -- the `beep` command, a thin client of the C ABI with the
// reference's flag surface (include/Config.h:309-429; src/main.cu:27-344).
//
//   beep -g <data.bel|.tsv|.mtx> -t <query.mtx> -m sgm|sgc [-d dev] [-s 1]
//
// Output ends with the reference's "------------- Counter = N" line
// (subgraph_matching.cuh:1277).  -m mtx-bel | tsv-bel | tsv-mtx | bel-mtx |
// txt-bel convert -g into -r; -x prints the statistics block.  Flags the path
// does not use (-w -o -a -v -k -p -e -q -c) are accepted and ignored; -c is
// overwritten by the computed cutoff in the reference as well (main.cu:132).
// -n > 1 is refused: this build drives one GPU.

#include <unistd.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <map>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/beep.h"

static void usage()
{
    fprintf(stderr,
            "\nUsage:  beep [options]\n"
            "\nOptions:\n"
            "    -g <Src graph FileName>       Name of file with input graph (.bel, .tsv, .mtx)\n"
            "    -t <Pattern graph filename>   Query graph (.mtx, 0-based, both directions listed)\n"
            "    -d <Device Id>                HIP device id (default = 0)\n"
            "    -m <MainTask>                 sgm (enumerate) or sgc (count); or a conversion\n"
            "                                  mtx-bel, tsv-bel, tsv-mtx, bel-mtx, txt-bel (-g -> -r)\n"
            "    -r <Dst graph FileName>       Output of a conversion\n"
            "    -s <sort>                     Sort read edges by src then dst\n"
            "    -n <num devices>              Only 1 is supported\n"
            "    -h                            Help\n"
            "  Environment: BEEP_EMIT=none|hash (default none), BEEP_LANES=<1..64>\n\n");
}

// -x: the reference's statistics block (main.cu:135-170): MatrixStats and
// PrintMtarixStruct (utils.cuh:110-183, 264-306) and the bitmap-encoding
// element counts, with its 32-bit unsigned arithmetic and float cell
// counters.  Its trailing single-row intersection prints nothing and is not
// repeated; where it divides by zero (fewer than 10 rows) or indexes past its
// 10 x 10 grid (the last rows when n is not a multiple of 10) this build uses
// one row per cell and clamps the cell index instead.
static void print_stats(const beep_graph_desc &gd)
{
    typedef unsigned int T;
    const T n = gd.n, nnz = (T)gd.m;
    std::map<T, T> dict;
    for (T i = 0; i < n; i++) dict[(T)(gd.rowptr[i + 1] - gd.rowptr[i])]++;
    T minRow = 0, maxRow = 1, totalRow = 0, median = 0, countMin = 0, countMax = 0, countMedian = 0;
    for (const auto &kv : dict) {
        if (kv.first == 0) continue;
        if (minRow == 0) {
            minRow = kv.first;
            countMin = kv.second;
        }
        if (countMedian < kv.second) {
            median = kv.first;
            countMedian = kv.second;
        }
        totalRow += kv.second * kv.first;
        maxRow = kv.first;
        countMax = kv.second;
    }
    const double mean = 1.0 * totalRow / n;
    double sd = 0;
    for (T i = 0; i < n; i++) {
        const double diff = (T)(gd.rowptr[i + 1] - gd.rowptr[i]) - mean;
        sd += diff * diff;
    }
    sd = std::sqrt(sd / n);
    printf("%u,%u,%u,%u,%.1f,%.1f,%u,%u,%u,%u,%u\n", nnz, n, n, minRow, mean, sd, median, maxRow, countMin,
           countMedian, countMax);

    const int res = 10;
    float st[res][res] = {};
    T unit = n / res;
    if (unit == 0) unit = 1;
    for (T i = 0; i < n; i++) {
        T ri = i / unit;
        if (ri >= (T)res) ri = res - 1;
        for (uint64_t j = gd.rowptr[i]; j < gd.rowptr[i + 1]; j++) {
            T ci = gd.col[j] / unit;
            if (ci >= (T)res) ci = res - 1;
            st[ri][ci]++;
        }
    }
    for (int i = 0; i < res; i++) {
        for (int j = 0; j < res; j++) {
            const float perc = st[i][j] / nnz;
            if (perc >= 0.001) printf("%d,", (int)(1000 * st[i][j] / nnz));
            else if (perc > 0) printf("+,");
            else printf("-,");
        }
        printf("\n");
    }

    printf("Now # of bytes we need to make this matrix binary encoded !!\n");
    const T dv = 32;
    uint64_t sum = 0, sumc = 0;
    for (T i = 0; i < n; i++) {
        const T deg = (T)(gd.rowptr[i + 1] - gd.rowptr[i]);
        sum += (uint64_t)(T)(deg * (deg + dv - 1) / dv);
        const uint64_t nelem8 = deg / dv;
        const uint64_t rem = deg - nelem8 * dv;
        sumc += dv * nelem8 * (1 + nelem8) / 2;
        sumc += rem * (1 + nelem8);
    }
    printf("n = %u, m = %u, elements = %llu\n", n, nnz, (unsigned long long)sum);
    printf("n = %u, m = %u, elements = %llu\n", n, nnz, (unsigned long long)sumc);
}

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv)
{
    const char *gpath = nullptr, *tpath = nullptr, *rpath = nullptr;
    int conv = 0, stats = 0;
    int dev = 0, ndev = 1, sort_edges = 0;
    int mode = -1;
    const char *task = "kc";
    int opt;
    while ((opt = getopt(argc, argv, ":g:r:d:m:x:o:a:k:h:v:s:p:e:q:w:t:c:n:")) >= 0) {
        switch (opt) {
        case 'g': gpath = optarg; break;
        case 't': tpath = optarg; break;
        case 'd': dev = atoi(optarg); break;
        case 'm':
            task = optarg;
            if (!strcmp(optarg, "sgm")) mode = BEEP_MODE_SGM;
            else if (!strcmp(optarg, "sgc")) mode = BEEP_MODE_SGC;
            else if (!strcmp(optarg, "mtx-bel")) conv = BEEP_CONV_MTX_BEL;
            else if (!strcmp(optarg, "tsv-bel")) conv = BEEP_CONV_TSV_BEL;
            else if (!strcmp(optarg, "tsv-mtx")) conv = BEEP_CONV_TSV_MTX;
            else if (!strcmp(optarg, "bel-mtx")) conv = BEEP_CONV_BEL_MTX;
            else if (!strcmp(optarg, "txt-bel")) conv = BEEP_CONV_TXT_BEL;
            else if (!strcmp(optarg, "tc") || !strcmp(optarg, "kcore") || !strcmp(optarg, "ktruss") ||
                     !strcmp(optarg, "kc") || !strcmp(optarg, "lkc") || !strcmp(optarg, "mc") ||
                     !strcmp(optarg, "cd"))
                mode = -1;
            else {
                // the reference exits with status 0 here (Config.h:86-87)
                fprintf(stderr, "Unrecognized -mt option (Main TASK): %s\n", optarg);
                return 0;
            }
            break;
        case 's': sort_edges = 1; break;
        case 'n': ndev = atoi(optarg); break;
        case 'h': usage(); return 0;
        case ':':
            if (optopt == 'h') { usage(); return 0; }
            fprintf(stderr, "\nUnrecognized option!\n");
            usage();
            return 0;
        case 'r': rpath = optarg; break;
        case 'x': stats = 1; break;
        case 'o': case 'a': case 'v': case 'k': case 'p': case 'e': case 'q':
        case 'w': case 'c': break;
        default:  // the reference prints usage and exits with status 0 (Config.h:421-424)
            fprintf(stderr, "\nUnrecognized option!\n");
            usage();
            return 0;
        }
    }
    if (ndev != 1) {
        fprintf(stderr, "beep: -n %d requested; this build drives exactly one GPU\n", ndev);
        return 2;
    }
    printf("Welcome ---------------------\n");
    printf("    Graph: %s\n", gpath ? gpath : "");
    printf("    Pattern Graph: %s\n", tpath ? tpath : "");
    printf("    Device Id = %d\n", dev);
    printf("    Main Task = %s\n", task);
    if (conv) {  // format conversion -g -> -r, nothing else (main.cu:40-68)
        printf("    DST Graph: %s\n", rpath ? rpath : "");
        if (!gpath || !rpath) { usage(); return 2; }
        if (beep_convert(conv, gpath, rpath)) {
            fprintf(stderr, "beep: %s\n", beep_last_error());
            return 1;
        }
        return 0;
    }
    if (!gpath) { usage(); return 2; }

    double t0 = now_s();
    beep_graph *g = nullptr;
    if (beep_graph_load(gpath, sort_edges, &g)) {
        fprintf(stderr, "beep: %s\n", beep_last_error());
        return 1;
    }
    beep_graph_desc gd;
    beep_graph_describe(g, &gd);
    printf("Read graph time: %f s\n", now_s() - t0);
    printf("%llu,%u,%u,%.1f,%.1f,%u\n", (unsigned long long)gd.m, gd.n, gd.n, gd.mean, gd.sd, gd.max_degree);
    if (stats) print_stats(gd);
    if (mode < 0) {  // tasks other than sgm/sgc only load the graph in the reference
        beep_graph_free(g);
        return 0;
    }
    if (!tpath) { usage(); beep_graph_free(g); return 2; }
    beep_plan *p = nullptr;
    if (beep_plan_from_mtx(tpath, mode, &p)) {
        fprintf(stderr, "beep: %s\n", beep_last_error());
        beep_graph_free(g);
        return 1;
    }
    beep_plan_desc pd;
    beep_plan_describe(p, &pd);
    printf("Node Sequence:\n");
    for (int i = 0; i < pd.k; i++) printf("i: %d;\tnode: %u (Degree: %u)\n", i, pd.seq[i], pd.qdeg[i]);
    printf("Query edges:\n");
    for (int i = 0; i < pd.k; i++) {
        printf("i: %d\t", i);
        for (uint32_t j = pd.qedge_ptr[i]; j < pd.qedge_ptr[i + 1]; j++) printf("%u,", pd.qedge[j]);
        printf("\n");
    }
    printf("Number of levels (other than last) to unmaterialize: %d\n", pd.lunmat);
    printf("Symmetrical nodes:\n");
    for (int i = 0; i < pd.k; i++) {
        printf("i: %d\t", i);
        for (uint32_t j = pd.sym_ptr[i]; j < pd.sym_ptr[i + 1]; j++) printf("%u,", pd.sym[j]);
        printf("\n");
    }
    double tp = now_s();
    beep_prep *pp = nullptr;
    if (beep_prepare(g, p, &pp) || beep_prep_upload(pp, dev)) {
        fprintf(stderr, "beep: %s\n", beep_last_error());
        return 1;
    }
    printf("%s SB\n", pd.asc ? "Ascending" : "Descending");
    printf("Preprocessing TIME: %f ms\n", (now_s() - tp) * 1000.0);
    beep_opts o;
    beep_opts_default(&o);
    o.device = dev;
    if (const char *e = getenv("BEEP_EMIT")) o.emit = !strcmp(e, "hash") ? BEEP_EMIT_HASH : BEEP_EMIT_NONE;
    if (const char *e = getenv("BEEP_LANES")) o.lanes = atoi(e);
    beep_result r;
    double tc = now_s();
    if (beep_count(pp, p, &o, &r)) {
        fprintf(stderr, "beep: %s\n", beep_last_error());
        return 1;
    }
    printf("device %d time: %f s\n", dev, r.t_enum_ms / 1000.0);
    printf("final count time %f s\n", now_s() - tc);
    if (o.emit) printf("match hash: %016llx %016llx\n", (unsigned long long)r.hash_sum, (unsigned long long)r.hash_xor);
    printf("------------- Counter = %llu\n", (unsigned long long)r.count);
    beep_prep_free(pp);
    beep_plan_free(p);
    beep_graph_free(g);
    return 0;
}
