// graph_io.cpp -- edge-list readers, CSR construction, cutoff statistic and the
// deterministic synthetic generators.
//
// Reader semantics follow include/FIleReader.cuh of the reference; the CSR
// build follows CSRCOO::from_edgelist (include/CSRCOO.cuh:37-80); the cutoff
// follows get_stats (include/utils.cuh:186-261).

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>

#include "internal.hpp"

namespace beep {

static bool ends_with(const std::string &s, const char *suf)
{
    size_t b = strlen(suf);
    return s.size() >= b && s.compare(s.size() - b, b, suf) == 0;
}

// .bel: 24-byte records {u64 dst, u64 src, u64 weight}; values truncated to
// uint32 (FIleReader.cuh:52-95).  The file is memory-mapped and decoded in
// parallel; a trailing partial record is ignored as fread of whole items does.
static void read_bel(const std::string &path, std::vector<uint32_t> &src,
                     std::vector<uint32_t> &dst)
{
    int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) throw Error(BEEP_EIO, "unable to open \"" + path + "\"");
    struct stat st;
    if (fstat(fd, &st)) { close(fd); throw Error(BEEP_EIO, "cannot stat " + path); }
    uint64_t nrec = (uint64_t)st.st_size / 24;
    src.resize(nrec);
    dst.resize(nrec);
    if (nrec == 0) { close(fd); return; }
    void *map = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) throw Error(BEEP_EIO, "cannot map " + path);
    madvise(map, (size_t)st.st_size, MADV_SEQUENTIAL);
    const unsigned char *base = (const unsigned char *)map;
    parallel_for(nrec, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) {
            uint64_t d, s;
            memcpy(&d, base + i * 24 + 0, 8);
            memcpy(&s, base + i * 24 + 8, 8);
            src[i] = (uint32_t)s;
            dst[i] = (uint32_t)d;
        }
    }, 1 << 16);
    munmap(map, (size_t)st.st_size);
}

// .tsv: "dst src weight" triples separated by white space
// (FIleReader.cuh:101-127); a malformed triple is an error (the reference exits).
static void read_tsv(const std::string &path, std::vector<uint32_t> &src,
                     std::vector<uint32_t> &dst)
{
    FILE *f = fopen(path.c_str(), "r");
    if (!f) throw Error(BEEP_EIO, "unable to open \"" + path + "\"");
    for (;;) {
        unsigned long long d, s, w;
        int k = fscanf(f, "%llu %llu %llu", &d, &s, &w);
        if (k != 3) {
            bool eof = feof(f);
            fclose(f);
            if (eof) break;
            throw Error(BEEP_EIO, "malformed triple in " + path);
        }
        src.push_back((uint32_t)s);
        dst.push_back((uint32_t)d);
    }
}

// .mtx (FIleReader.cuh:130-212, 299-353; main.cu:83-86, 317-320): banner lines
// start with '%' ("%%" lines set the format flags), the first other line is the
// size header; body lines are 0-based "row col [value]" read in batches of 100
// -- a batch ends at the first line that does not parse and reading stops at
// the first batch that yields no edge.  "symmetric" and "array" banners are
// rejected: the reference reader leaves unwritten slots for them
// (FIleReader.cuh:202-208).
static void read_mtx(const std::string &path, std::vector<uint32_t> &src,
                     std::vector<uint32_t> &dst)
{
    std::ifstream in(path);
    if (!in) throw Error(BEEP_EIO, "unable to open \"" + path + "\"");
    std::string line;
    bool symmetric = false, array = false, pattern = false;
    for (;;) {
        if (!std::getline(in, line)) line.clear();
        if (line.empty() || line[0] != '%') break;
        if (line.size() >= 2 && line[1] == '%') {
            symmetric = line.find("symmetric") != std::string::npos;
            array = line.find("array") != std::string::npos;
            pattern = line.find("pattern") != std::string::npos;
        }
    }
    if (symmetric || array)
        throw Error(BEEP_EINVAL, path + ": symmetric/array MatrixMarket banners are not supported");
    for (;;) {
        int got = 0;
        for (; got < 100; got++) {
            if (!std::getline(in, line)) line.clear();
            long long r, c;
            double v;
            bool ok = pattern ? sscanf(line.c_str(), "%lld %lld", &r, &c) == 2
                              : sscanf(line.c_str(), "%lld %lld %lf", &r, &c, &v) >= 2;
            if (!ok) break;
            src.push_back((uint32_t)r);
            dst.push_back((uint32_t)c);
        }
        if (got == 0) break;
    }
}

void read_edges(const std::string &path, std::vector<uint32_t> &src, std::vector<uint32_t> &dst)
{
    src.clear();
    dst.clear();
    if (ends_with(path, ".bel")) read_bel(path, src, dst);
    else if (ends_with(path, ".tsv")) read_tsv(path, src, dst);
    else if (ends_with(path, ".mtx")) read_mtx(path, src, dst);
    else throw Error(BEEP_EINVAL, "no reader for file " + path);
}

uint32_t compute_cutoff(const Graph &g, double *mean_out, double *sd_out)
{
    // totalRow is accumulated as a 32-bit unsigned value (T = uint) and the
    // squared deviations as a sequential double sum in row order.
    uint32_t nr = g.n;
    if (nr == 0) {
        if (mean_out) *mean_out = 0;
        if (sd_out) *sd_out = 0;
        return 0;
    }
    uint32_t total = (uint32_t)g.m;
    double mean = 1.0 * total / nr;
    double acc = 0;
    for (uint32_t i = 0; i < nr; i++) {
        uint32_t rc = (uint32_t)(g.rowptr[i + 1] - g.rowptr[i]);
        double diff = rc - mean;
        acc += diff * diff;
    }
    double sd = std::sqrt(acc / nr);
    if (mean_out) *mean_out = mean;
    if (sd_out) *sd_out = sd;
    return (uint32_t)std::ceil(mean + 6.0 * sd);
}

void Graph::finalize()
{
    max_degree = 0;
    for (uint32_t v = 0; v < n; v++)
        max_degree = std::max<uint32_t>(max_degree, (uint32_t)(rowptr[v + 1] - rowptr[v]));
    cutoff = compute_cutoff(*this, &mean, &sd);
}

std::unique_ptr<Graph> csr_from_edges(const uint32_t *src, const uint32_t *dst, uint64_t ne)
{
    auto g = std::make_unique<Graph>();
    uint64_t nrow = 0, nnz = 0;
    for (uint64_t e = 0; e < ne; e++) {
        if (src[e] == dst[e]) continue;  // self loops dropped (CSRCOO.cuh:51-52)
        uint64_t r = (uint64_t)src[e] + 1;
        if (r < nrow) throw Error(BEEP_EINVAL, "edge list is not grouped by ascending source");
        nrow = r;
        nnz++;
    }
    if (nrow > 0xffffffffull) throw Error(BEEP_ERANGE, "too many vertices");
    g->n = (uint32_t)nrow;
    g->m = nnz;
    g->rowptr.assign(nrow + 1, 0);
    g->col.resize(nnz);
    uint64_t k = 0;
    for (uint64_t e = 0; e < ne; e++) {
        if (src[e] == dst[e]) continue;
        g->rowptr[src[e] + 1]++;
        g->col[k++] = dst[e];
    }
    for (uint64_t i = 0; i < nrow; i++) g->rowptr[i + 1] += g->rowptr[i];
    bool sorted = true;
    for (uint32_t v = 0; v < g->n && sorted; v++)
        for (uint64_t e = g->rowptr[v] + 1; e < g->rowptr[v + 1]; e++)
            if (g->col[e - 1] > g->col[e]) { sorted = false; break; }
    g->rows_sorted = sorted;
    for (uint64_t e = 0; e < nnz; e++)
        if (g->col[e] >= g->n)
            throw Error(BEEP_EINVAL, "edge destination has no row (graph not symmetric)");
    g->finalize();
    return g;
}

void write_bel(const Graph &g, const std::string &path)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) throw Error(BEEP_EIO, "cannot create " + path);
    const size_t shard = 1 << 16;
    std::vector<uint64_t> buf(3 * shard);
    size_t fill = 0;
    for (uint32_t u = 0; u < g.n; u++) {
        for (uint64_t e = g.rowptr[u]; e < g.rowptr[u + 1]; e++) {
            buf[3 * fill + 0] = g.col[e];
            buf[3 * fill + 1] = u;
            buf[3 * fill + 2] = 0;
            if (++fill == shard) {
                if (fwrite(buf.data(), 8, 3 * fill, f) != 3 * fill) { fclose(f); throw Error(BEEP_EIO, "short write"); }
                fill = 0;
            }
        }
    }
    if (fill && fwrite(buf.data(), 8, 3 * fill, f) != 3 * fill) { fclose(f); throw Error(BEEP_EIO, "short write"); }
    if (fclose(f)) throw Error(BEEP_EIO, "close failed");
}

// ------------------------------------------------------------- generators
//
// Every sample i draws from its own xoshiro256** stream seeded by splitmix64
// from (seed, i), so the result does not depend on the thread count.

static inline uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    Xoshiro(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed * 0xd1342543de82ef95ull + stream * 0x9e3779b97f4a7c15ull + 1;
        for (auto &v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// Symmetric CSR from endpoint pairs: self loops dropped, duplicates removed,
// rows ascending; n = max id + 1 (what from_edgelist gives on the written file).
static std::unique_ptr<Graph> csr_from_pairs(const std::vector<uint32_t> &a,
                                             const std::vector<uint32_t> &b)
{
    const size_t S = a.size();
    uint32_t nmax = 0;
    {
        std::vector<uint32_t> part(host_threads() * 2 + 1, 0);
        std::atomic<uint32_t> mx{0};
        parallel_for(S, [&](size_t lo, size_t hi) {
            uint32_t m = 0;
            for (size_t i = lo; i < hi; i++)
                if (a[i] != b[i]) m = std::max(m, std::max(a[i], b[i]) + 1);
            uint32_t cur = mx.load();
            while (m > cur && !mx.compare_exchange_weak(cur, m)) {}
        }, 1 << 18);
        nmax = mx.load();
    }
    auto g = std::make_unique<Graph>();
    g->n = nmax;
    std::vector<std::atomic<uint64_t>> cnt(nmax + 1);
    for (auto &c : cnt) c.store(0, std::memory_order_relaxed);
    parallel_for(S, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++)
            if (a[i] != b[i]) {
                cnt[a[i]].fetch_add(1, std::memory_order_relaxed);
                cnt[b[i]].fetch_add(1, std::memory_order_relaxed);
            }
    }, 1 << 18);
    progress("csr: %zu pairs over %u vertices counted", S, nmax);
    std::vector<uint64_t> off(nmax + 1, 0);
    for (uint32_t v = 0; v < nmax; v++) off[v + 1] = off[v] + cnt[v].load(std::memory_order_relaxed);
    std::vector<uint32_t> tmp(off[nmax]);
    for (uint32_t v = 0; v < nmax; v++) cnt[v].store(off[v], std::memory_order_relaxed);
    parallel_for(S, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++)
            if (a[i] != b[i]) {
                tmp[cnt[a[i]].fetch_add(1, std::memory_order_relaxed)] = b[i];
                tmp[cnt[b[i]].fetch_add(1, std::memory_order_relaxed)] = a[i];
            }
    }, 1 << 18);
    progress("csr: %llu entries scattered", (unsigned long long)off[nmax]);
    std::vector<uint64_t> len(nmax, 0);
    const uint32_t nslices = off[nmax] > (1ull << 30) ? 16 : 1;
    for (uint32_t sl = 0; sl < nslices; sl++) {
        const size_t s0 = (size_t)nmax * sl / nslices, s1 = (size_t)nmax * (sl + 1) / nslices;
        parallel_for(s1 - s0, [&](size_t lo, size_t hi) {
            for (size_t v = s0 + lo; v < s0 + hi; v++) {
                auto bgn = tmp.begin() + (ptrdiff_t)off[v], end = tmp.begin() + (ptrdiff_t)off[v + 1];
                std::sort(bgn, end);
                len[v] = (uint64_t)(std::unique(bgn, end) - bgn);
            }
        }, 1 << 12);
        if (nslices > 1) progress("csr: rows %zu..%zu sorted", s0, s1);
    }
    progress("csr: rows sorted and deduplicated");
    g->rowptr.assign((size_t)nmax + 1, 0);
    for (uint32_t v = 0; v < nmax; v++) g->rowptr[v + 1] = g->rowptr[v] + len[v];
    g->m = g->rowptr[nmax];
    g->col.resize(g->m);
    parallel_for(nmax, [&](size_t lo, size_t hi) {
        for (size_t v = lo; v < hi; v++)
            std::copy(tmp.begin() + (ptrdiff_t)off[v], tmp.begin() + (ptrdiff_t)(off[v] + len[v]),
                      g->col.begin() + (ptrdiff_t)g->rowptr[v]);
    }, 1 << 12);
    g->rows_sorted = true;
    g->finalize();
    return g;
}

std::unique_ptr<Graph> csr_from_undirected_keys(std::vector<uint64_t> &keys, uint32_t)
{
    std::vector<uint32_t> a(keys.size()), b(keys.size());
    for (size_t i = 0; i < keys.size(); i++) {
        a[i] = (uint32_t)(keys[i] >> 32);
        b[i] = (uint32_t)keys[i];
    }
    return csr_from_pairs(a, b);
}

std::unique_ptr<Graph> gen_rmat(uint32_t scale, uint64_t samples, double pa, double pb, double pc,
                                uint64_t seed, uint32_t n_keep)
{
    if (scale == 0 || scale > 31) throw Error(BEEP_EINVAL, "R-MAT scale must be in [1, 31]");
    if (pa < 0 || pb < 0 || pc < 0 || pa + pb + pc > 1) throw Error(BEEP_EINVAL, "bad R-MAT probabilities");
    uint64_t full = 1ull << scale;
    uint64_t keep = (n_keep == 0 || n_keep > full) ? full : n_keep;
    std::vector<uint32_t> a(samples), b(samples);
    const double ab = pa + pb, abc = pa + pb + pc;
    parallel_for(samples, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) {
            Xoshiro r(seed, i);
            uint64_t s = 0, d = 0;
            for (uint32_t l = 0; l < scale; l++) {
                double u = r.uniform();
                uint64_t sb = 0, db = 0;
                if (u < pa) {}
                else if (u < ab) db = 1;
                else if (u < abc) sb = 1;
                else { sb = 1; db = 1; }
                s = (s << 1) | sb;
                d = (d << 1) | db;
            }
            if (s >= keep || d >= keep) { s = d = 0; }  // rejected: becomes a dropped loop
            a[i] = (uint32_t)s;
            b[i] = (uint32_t)d;
        }
    }, 1 << 16);
    progress("rmat: %llu samples drawn", (unsigned long long)samples);
    return csr_from_pairs(a, b);
}

std::unique_ptr<Graph> gen_chung_lu(uint32_t n, uint64_t samples, double gamma, double i0,
                                    uint64_t seed)
{
    if (n < 2 || gamma <= 1.0 || i0 <= 0) throw Error(BEEP_EINVAL, "bad Chung-Lu parameters");
    const double alpha = 1.0 / (gamma - 1.0);
    std::vector<double> cdf(n);
    double acc = 0;
    for (uint32_t i = 0; i < n; i++) {
        acc += std::pow((double)i + i0, -alpha);
        cdf[i] = acc;
    }
    for (auto &x : cdf) x /= acc;
    cdf[n - 1] = 1.0;
    std::vector<uint32_t> a(samples), b(samples);
    parallel_for(samples, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) {
            Xoshiro r(seed, i);
            double u1 = r.uniform(), u2 = r.uniform();
            a[i] = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u1) - cdf.begin());
            b[i] = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u2) - cdf.begin());
            if (a[i] >= n) a[i] = n - 1;
            if (b[i] >= n) b[i] = n - 1;
        }
    }, 1 << 16);
    return csr_from_pairs(a, b);
}

}  // namespace beep
