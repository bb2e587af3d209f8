// convert.cpp -- the reference's format converters (-m mtx-bel | tsv-bel |
// tsv-mtx | bel-mtx | txt-bel, main.cu:40-68; MtB_Writer in
// FIleReader.cuh:459-912), with their byte semantics:
//
//   .bel  24-byte records {u64 dst, u64 src, u64 0}
//   mtx-bel  header flags from the %-lines (each line of length >= 2 resets
//            symmetric / skew / array / pattern), "rows cols nnz", nnz entry
//            lines taken as they are (no 1-based shift), both directions when
//            symmetric, ids cut to 32 bits, sorted by (src, dst), duplicates
//            kept; reading stops at the first malformed entry line and what
//            was read is still written (FIleReader.cuh:463-640)
//   tsv-bel  "f1 f2 f3" lines until the first that does not hold three
//            numbers; record {f1, f2, 0}, 64-bit ids, input order (:645-703)
//   tsv-mtx  "f1 f2 w": edge (f2, f1) with 32-bit ids, weight = low 32 bits
//            as int; header n = first id of the last edge (:706-754)
//   bel-mtx  records -> lines "src dst w" (w = low 32 bits of the record's
//            third word as int), header as tsv-mtx (:757-810)
//   txt-bel  "a b" lines, %-lines and malformed lines skipped, a == b
//            dropped, both directions, 32-bit ids, sorted and deduplicated
//            (:813-910)
//
// Where the reference has undefined behaviour -- an empty edge list in the
// *-mtx converters (fileEdges[m - 1]), an invalid mtx size line, empty text
// lines in txt-bel (its fscanf loop never advances), lines longer than its
// 100-byte buffer -- this build returns an error (the first two) or reads the
// line whole (the last two).

#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace beep {

namespace {

struct File {
    FILE *f = nullptr;
    File(const std::string &p, const char *mode)
    {
        f = fopen(p.c_str(), mode);
        if (!f) throw Error(BEEP_EIO, "cannot open " + p);
    }
    ~File()
    {
        if (f) fclose(f);
    }
    void close()
    {
        if (f && fclose(f)) {
            f = nullptr;
            throw Error(BEEP_EIO, "close failed");
        }
        f = nullptr;
    }
};

using Pair32 = std::pair<uint32_t, uint32_t>;

void put_bel(File &w, uint64_t a, uint64_t b)
{
    const uint64_t rec[3] = {a, b, 0};
    if (fwrite(rec, 8, 3, w.f) != 3) throw Error(BEEP_EIO, "short write");
}

// Whole line without its newline; false at end of file.
bool get_line(FILE *f, std::string &line)
{
    line.clear();
    int c;
    bool any = false;
    while ((c = fgetc(f)) != EOF) {
        any = true;
        if (c == '\n') return true;
        line.push_back((char)c);
    }
    return any;
}

void write_market(const std::string &dst, const std::vector<Pair32> &e, const std::vector<int> &w)
{
    if (e.empty()) throw Error(BEEP_EINVAL, "no edges to write");
    File out(dst, "w");
    const unsigned long long n = e.back().first;
    fprintf(out.f, "%%%%MatrixMarket matrix coordinate general\n%llu %llu %llu\n", n, n,
            (unsigned long long)e.size());
    for (size_t i = 0; i < e.size(); i++) fprintf(out.f, "%u %u %d\n", e[i].first, e[i].second, w[i]);
    out.close();
}

void mtx_bel(const std::string &src, const std::string &dst)
{
    std::ifstream in(src);
    if (!in) throw Error(BEEP_EIO, "cannot open " + src);
    bool symmetric = false, array = false, pattern = false;
    std::string line;
    for (;;) {
        if (!std::getline(in, line)) line.clear();
        if (line.empty() || line[0] != '%') break;
        if (line.size() >= 2) {
            symmetric = line.find("symmetric") != std::string::npos;
            array = line.find("array") != std::string::npos;
            pattern = line.find("pattern") != std::string::npos;
        }
    }
    long long nx = 0, ny = 0, nnz = 0;
    const int got = sscanf(line.c_str(), "%lld %lld %lld", &nx, &ny, &nnz);
    if (array && got == 2) nnz = nx * ny;
    else if (!array && got == 3) {
        if (nx != ny) throw Error(BEEP_EINVAL, "MARKET graph is not square");
    } else {
        throw Error(BEEP_EINVAL, "invalid MARKET size line");
    }
    std::vector<Pair32> e;
    for (long long i = 0; i < nnz; i++) {
        if (!std::getline(in, line)) line.clear();
        long long r = 0, c = 0;
        double v = 0;
        int k = pattern ? sscanf(line.c_str(), "%lld %lld", &r, &c) : sscanf(line.c_str(), "%lld %lld %lf", &r, &c, &v);
        if (array && k == 1) {
            c = i / nx;
            r = i - nx * c;
        } else if (pattern ? (array || k != 2) : (array || k < 2)) {
            break;  // malformed entry: stop, keep what was read
        }
        e.emplace_back((uint32_t)r, (uint32_t)c);
        if (symmetric && c != r) e.emplace_back((uint32_t)c, (uint32_t)r);
    }
    std::sort(e.begin(), e.end());
    File out(dst, "wb");
    for (const auto &p : e) put_bel(out, p.second, p.first);
    out.close();
}

void tsv_bel(const std::string &src, const std::string &dst)
{
    File in(src, "r");
    File out(dst, "wb");
    unsigned long long f1, f2, f3;
    while (fscanf(in.f, "%llu %llu %llu", &f1, &f2, &f3) == 3) put_bel(out, f1, f2);
    out.close();
}

void tsv_mtx(const std::string &src, const std::string &dst)
{
    std::vector<Pair32> e;
    std::vector<int> w;
    {
        File in(src, "r");
        unsigned long long f1, f2, f3;
        while (fscanf(in.f, "%llu %llu %llu", &f1, &f2, &f3) == 3) {
            e.emplace_back((uint32_t)f2, (uint32_t)f1);
            w.push_back((int)(uint32_t)f3);
        }
    }
    write_market(dst, e, w);
}

void bel_mtx(const std::string &src, const std::string &dst)
{
    std::vector<Pair32> e;
    std::vector<int> w;
    {
        File in(src, "rb");
        uint64_t rec[3];
        while (fread(rec, 24, 1, in.f) == 1) {
            e.emplace_back((uint32_t)rec[1], (uint32_t)rec[0]);
            w.push_back((int)(uint32_t)rec[2]);
        }
    }
    write_market(dst, e, w);
}

void txt_bel(const std::string &src, const std::string &dst)
{
    std::vector<Pair32> e;
    {
        File in(src, "r");
        std::string line;
        while (get_line(in.f, line)) {
            if (line.empty() || line[0] == '%') continue;
            unsigned long long a, b;
            if (sscanf(line.c_str(), "%llu %llu", &a, &b) != 2 || a == b) continue;
            e.emplace_back((uint32_t)a, (uint32_t)b);
            e.emplace_back((uint32_t)b, (uint32_t)a);
        }
    }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    File out(dst, "wb");
    for (const auto &p : e) put_bel(out, p.second, p.first);
    out.close();
}

}  // namespace

void convert(int kind, const std::string &src, const std::string &dst)
{
    switch (kind) {
    case BEEP_CONV_MTX_BEL: return mtx_bel(src, dst);
    case BEEP_CONV_TSV_BEL: return tsv_bel(src, dst);
    case BEEP_CONV_TSV_MTX: return tsv_mtx(src, dst);
    case BEEP_CONV_BEL_MTX: return bel_mtx(src, dst);
    case BEEP_CONV_TXT_BEL: return txt_bel(src, dst);
    default: throw Error(BEEP_EINVAL, "unknown conversion");
    }
}

}  // namespace beep
