// pgd_reader.cpp -- host parser with the reference reader's rules.
//
// read_graph (graphlet_core.cpp:259-283) dispatches on the file extension:
// ".mtx" goes to read_mtx (459-522), everything else to read_edge_list
// (291-457).  This file turns either format into a COO list in line order
// plus the reader's vertex count; the CSR itself is assembled on the device
// (pgd_graph_from_coo).  Rules kept from the reference:
//   * .mtx: the first non-comment line is the header and is never an edge;
//     ids are decremented; n = 1 + the largest endpoint of a non-loop line;
//     no duplicate removal; if the header's first two fields contain a 0 the
//     reader reads no edges at all.
//   * edge lists: the delimiter is detected from the first non-comment line
//     (',' then tab then three spaces then one space, else by extension); ids
//     are decremented only if no line has a 0 id; ids are renumbered in order
//     of first appearance; n = number of distinct endpoints; duplicates are
//     removed later keeping the first occurrence.
//   * fields are parsed as the reference does: `>>` for the first line and
//     the .mtx lines, atoi on delimiter-separated tokens otherwise, and
//     values left unchanged by a failed parse carry over to the next line.
#include <cctype>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pgd_hip.h"

namespace {

struct Text {
    std::vector<char> buf;
    size_t pos = 0;
    bool load(const char* path) {
        FILE* f = std::fopen(path, "rb");
        if (!f) return false;
        std::fseek(f, 0, SEEK_END);
        long sz = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        buf.resize((size_t)(sz > 0 ? sz : 0));
        size_t got = sz > 0 ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
        buf.resize(got);
        std::fclose(f);
        return true;
    }
    // std::getline: a final segment without '\n' is still a line
    bool line(const char*& s, size_t& n) {
        if (pos >= buf.size()) return false;
        const char* b = buf.data() + pos;
        const void* nl = std::memchr(b, '\n', buf.size() - pos);
        n = nl ? (size_t)((const char*)nl - b) : buf.size() - pos;
        s = b;
        pos += n + (nl ? 1 : 0);
        return true;
    }
};

inline bool comment(const char* s, size_t n) { return n > 0 && (s[0] == '#' || s[0] == '%'); }

// `stream >> int` as libstdc++ performs it (see file comment)
struct StreamInts {
    const char* p;
    const char* e;
    bool failed = false;
    StreamInts(const char* s, size_t n) : p(s), e(s + n) {}
    void get(int& out) {
        if (failed) return;
        while (p < e && std::isspace((unsigned char)*p)) p++;
        if (p >= e) { failed = true; return; }
        bool neg = false;
        if (*p == '+' || *p == '-') { neg = *p == '-'; p++; }
        long long acc = 0;
        int digits = 0;
        bool over = false;
        for (; p < e && *p >= '0' && *p <= '9'; p++, digits++) {
            if (!over) {
                acc = acc * 10 + (*p - '0');
                if (acc > (long long)INT_MAX + 1) over = true;
            }
        }
        if (digits == 0) { out = 0; failed = true; return; }
        long long v = neg ? -acc : acc;
        if (over || v > INT_MAX || v < INT_MIN) { out = neg ? INT_MIN : INT_MAX; failed = true; return; }
        out = (int)v;
    }
};

// atoi on a token (glibc: (int)strtol)
inline int atoi_token(const char* b, size_t n) {
    char tmp[64];
    if (n >= sizeof(tmp)) n = sizeof(tmp) - 1;
    std::memcpy(tmp, b, n);
    tmp[n] = 0;
    return (int)std::strtol(tmp, nullptr, 10);
}

long find_from(const char* s, size_t n, const std::string& d, size_t from) {
    if (d.empty()) return from <= n ? (long)from : -1;
    for (size_t i = from; i + d.size() <= n; i++)
        if (std::memcmp(s + i, d.data(), d.size()) == 0) return (long)i;
    return -1;
}

// graphlet_core.cpp:229-238
inline void token(int& v, const char* s, size_t n, const std::string& d, size_t& prev) {
    long pos = find_from(s, n, d, prev);
    if (pos >= 0) {
        if ((size_t)pos > prev) v = atoi_token(s + prev, (size_t)pos - prev);
        prev = (size_t)pos + 1;
    } else if (prev < n) {
        v = atoi_token(s + prev, n - prev);
    }
}

std::string extension(const std::string& path) {
    size_t d = path.rfind('.');
    return d == std::string::npos ? std::string() : path.substr(d + 1);
}

pgd_status read_mtx(Text& t, std::vector<int32_t>& a, std::vector<int32_t>& b, int64_t& n, int64_t& loops) {
    const char* s = nullptr;
    size_t len = 0;
    bool have;
    while ((have = t.line(s, len)) && comment(s, len)) {}
    if (!have) { s = ""; len = 0; }
    int v = -1, u = -1;
    bool fix = true;
    if (len > 0) {
        StreamInts in(s, len);
        in.get(v);
        in.get(u);
        if (v == 0 || u == 0) fix = false;
    }
    int64_t maxid = -1;
    loops = 0;
    while (fix && t.line(s, len)) {
        if (len == 0 || comment(s, len)) continue;
        StreamInts in(s, len);
        in.get(v);
        in.get(u);
        v--;
        u--;
        if (v == u) { loops++; continue; }
        if (v < 0 || u < 0) return PGD_ERR_INVALID;  // the reference never terminates on these
        a.push_back(v);
        b.push_back(u);
        if (v > maxid) maxid = v;
        if (u > maxid) maxid = u;
    }
    n = maxid + 1;
    return PGD_OK;
}

pgd_status read_edge_list(Text& t, const std::string& path, std::vector<int32_t>& a, std::vector<int32_t>& b,
                          int64_t& n, int64_t& loops) {
    const char* s = nullptr;
    size_t len = 0;
    bool have;
    std::string delim;
    while ((have = t.line(s, len)) && comment(s, len)) {
        static const char tag[] = "MatrixMarket matrix coordinate pattern symmetric";
        if (len >= sizeof(tag) - 1 && memmem(s, len, tag, sizeof(tag) - 1)) delim = " ";
    }
    if (!have) { s = ""; len = 0; }
    if (delim.empty()) {  // detect_delim, graphlet_core.cpp:196-226
        if (std::memchr(s, ',', len)) delim = ",";
        else if (std::memchr(s, '\t', len)) delim = "\t";
        else if (memmem(s, len, "   ", 3)) delim = "   ";
        else if (std::memchr(s, ' ', len)) delim = " ";
    }
    if (delim.empty()) {
        std::string ext = extension(path);
        delim = ext == "csv" ? "," : ext == "tab" ? "\t" : " ";
    }
    int v = -1, u = -1;
    bool fix = true;
    if (len > 0) {
        StreamInts in(s, len);
        in.get(v);
        in.get(u);
        if (v == 0 || u == 0) fix = false;
    }
    int maxv = 0;
    while (t.line(s, len)) {  // first pass starts after the first line
        if (len == 0 || comment(s, len)) continue;
        size_t prev = 0;
        token(v, s, len, delim, prev);
        token(u, s, len, delim, prev);
        if (v == 0 || u == 0) fix = false;
        if (v > maxv) maxv = v;
        if (u > maxv) maxv = u;
    }
    t.pos = 0;
    std::vector<int32_t> lookup((size_t)maxv + 1, -1);
    int32_t next_id = 0;
    loops = 0;
    while (t.line(s, len)) {
        if (len == 0 || comment(s, len)) continue;
        size_t prev = 0;
        token(v, s, len, delim, prev);
        token(u, s, len, delim, prev);
        if (fix) { v--; u--; }
        if (v == u) { loops++; continue; }
        if (v < 0 || u < 0 || v > maxv || u > maxv) return PGD_ERR_INVALID;  // out of the reference's table
        if (lookup[(size_t)v] < 0) lookup[(size_t)v] = next_id++;
        int32_t x = lookup[(size_t)v];
        if (lookup[(size_t)u] < 0) lookup[(size_t)u] = next_id++;
        int32_t y = lookup[(size_t)u];
        a.push_back(x);
        b.push_back(y);
    }
    n = next_id;
    return PGD_OK;
}

}  // namespace

extern "C" pgd_status pgd_read_graph(const char* path, int64_t* n, int64_t* nlines, int32_t** src, int32_t** dst,
                                     int32_t* dedup, int64_t* self_loops) {
    if (!path || !n || !nlines || !src || !dst) return PGD_ERR_INVALID;
    Text t;
    if (!t.load(path)) return PGD_ERR_INVALID;
    std::vector<int32_t> a, b;
    int64_t nv = 0, loops = 0;
    std::string p(path);
    bool mtx = extension(p) == "mtx";
    pgd_status st = mtx ? read_mtx(t, a, b, nv, loops) : read_edge_list(t, p, a, b, nv, loops);
    if (st != PGD_OK) return st;
    size_t k = a.size();
    *src = (int32_t*)std::malloc((k ? k : 1) * sizeof(int32_t));
    *dst = (int32_t*)std::malloc((k ? k : 1) * sizeof(int32_t));
    if (!*src || !*dst) { std::free(*src); std::free(*dst); return PGD_ERR_OOM; }
    if (k) {
        std::memcpy(*src, a.data(), k * sizeof(int32_t));
        std::memcpy(*dst, b.data(), k * sizeof(int32_t));
    }
    *n = nv;
    *nlines = (int64_t)k;
    if (dedup) *dedup = mtx ? 0 : 1;
    if (self_loops) *self_loops = loops;
    return PGD_OK;
}

extern "C" void pgd_free(void* p) { std::free(p); }
