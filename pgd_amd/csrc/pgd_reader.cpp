// pgd_reader.cpp -- the graph reader with the reference reader's rules, its
// text parsed by several host threads.
//
// read_graph (graphlet_core.cpp:259-283) dispatches on the file extension:
// ".mtx" goes to read_mtx (459-522), everything else to read_edge_list
// (291-457).  Rules kept from the reference:
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
//     the .mtx lines, atoi on delimiter-separated tokens otherwise, and a
//     field that a line's parse leaves unset keeps the previous line's value.
//   * a line is skipped when empty or when it starts with '#' or '%'; the
//     leading comment block ends at the first line that is not a comment
//     (an empty line ends it too).
//
// Parallel scheme: the text is cut into chunks at line starts; each thread
// parses whole lines into (value, was-set) pairs, which depend on the line's
// text alone.  The edge-list reader's first pass (graphlet_core.cpp:350-371:
// the largest id, and whether any id is 0) sees carried values unchanged, so
// it resolves them by a two-level scan in line order (the last set value of
// each chunk, propagated across chunks).  Its second pass (383-446) and the
// .mtx pass decrement, and renumber, the line variables in place, so a
// carried field depends on the lines before it: when every line sets both of
// its fields the lines are independent and stay parallel (renumbered later,
// on the device or the host), otherwise the file is replayed line by line.
#include "pgd_reader.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

// a read-only view of a whole file (mapped; an empty file maps nothing)
struct FileView {
    const char* p = nullptr;
    size_t n = 0;
    void* map = nullptr;
    bool open(const char* path) {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        if (fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
            ::close(fd);
            return false;
        }
        n = (size_t)st.st_size;
        if (n) {
            map = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
            if (map == MAP_FAILED) {
                map = nullptr;
                ::close(fd);
                return false;
            }
            (void)madvise(map, n, MADV_SEQUENTIAL);
            p = (const char*)map;
        }
        ::close(fd);
        return true;
    }
    ~FileView() {
        if (map) munmap(map, n);
    }
    // the line starting at `pos` (std::getline: a last line without '\n'
    // still counts); returns the start of the next line
    size_t line(size_t pos, const char*& s, size_t& len) const {
        s = p + pos;
        const void* nl = std::memchr(s, '\n', n - pos);
        len = nl ? (size_t)((const char*)nl - s) : n - pos;
        return pos + len + (nl ? 1 : 0);
    }
};

inline bool comment(const char* s, size_t n) { return n > 0 && (s[0] == '#' || s[0] == '%'); }

// `stream >> int` as libstdc++ performs it; get() returns whether it wrote
// `out` (at the end of the text, or after a failed extraction, it does not)
struct StreamInts {
    const char* p;
    const char* e;
    bool failed = false;
    StreamInts(const char* s, size_t n) : p(s), e(s + n) {}
    bool get(int32_t& out) {
        if (failed) return false;
        while (p < e && std::isspace((unsigned char)*p)) p++;
        if (p >= e) {
            failed = true;
            return false;
        }
        bool neg = false;
        if (*p == '+' || *p == '-') {
            neg = *p == '-';
            p++;
        }
        long long acc = 0;
        int digits = 0;
        bool over = false;
        for (; p < e && *p >= '0' && *p <= '9'; p++, digits++) {
            if (!over) {
                acc = acc * 10 + (*p - '0');
                if (acc > (long long)INT_MAX + 1) over = true;
            }
        }
        if (digits == 0) {
            out = 0;
            failed = true;
            return true;
        }
        long long v = neg ? -acc : acc;
        if (over || v > INT_MAX || v < INT_MIN) {
            out = neg ? INT_MIN : INT_MAX;
            failed = true;
            return true;
        }
        out = (int32_t)v;
        return true;
    }
};

// atoi on a token (glibc: (int)strtol)
inline int32_t atoi_token(const char* b, size_t n) {
    char tmp[64];
    if (n >= sizeof(tmp)) n = sizeof(tmp) - 1;
    std::memcpy(tmp, b, n);
    tmp[n] = 0;
    return (int32_t)std::strtol(tmp, nullptr, 10);
}

inline long find_from(const char* s, size_t n, const std::string& d, size_t from) {
    if (d.size() == 1) {
        if (from >= n) return -1;
        const void* q = std::memchr(s + from, d[0], n - from);
        return q ? (long)((const char*)q - s) : -1;
    }
    for (size_t i = from; i + d.size() <= n; i++)
        if (std::memcmp(s + i, d.data(), d.size()) == 0) return (long)i;
    return -1;
}

// get_token (graphlet_core.cpp:229-238); returns whether it wrote v
inline bool token(int32_t& v, const char* s, size_t n, const std::string& d, size_t& prev) {
    const long pos = find_from(s, n, d, prev);
    if (pos >= 0) {
        bool set = false;
        if ((size_t)pos > prev) {
            v = atoi_token(s + prev, (size_t)pos - prev);
            set = true;
        }
        prev = (size_t)pos + 1;
        return set;
    }
    if (prev < n) {
        v = atoi_token(s + prev, n - prev);
        return true;
    }
    return false;
}

// one line's two fields: bit 0 set when v was written, bit 1 when u was
inline uint8_t parse_tokens(const char* s, size_t n, const std::string& d, int32_t& v, int32_t& u) {
    size_t prev = 0;
    uint8_t f = 0;
    if (token(v, s, n, d, prev)) f |= 1;
    if (token(u, s, n, d, prev)) f |= 2;
    return f;
}
inline uint8_t parse_stream(const char* s, size_t n, int32_t& v, int32_t& u) {
    StreamInts in(s, n);
    uint8_t f = 0;
    if (in.get(v)) f |= 1;
    if (in.get(u)) f |= 2;
    return f;
}

std::string extension(const std::string& path) {
    size_t d = path.rfind('.');
    return d == std::string::npos ? std::string() : path.substr(d + 1);
}

// the parsed fields of every data line of [begin, end of file), in order
struct Parsed {
    std::vector<int32_t> v, u;
    std::vector<uint8_t> set;
};

struct Chunk {
    size_t b = 0, e = 0;  // lines starting in [b, e)
    std::vector<int32_t> v, u;
    std::vector<uint8_t> set;
};

template <class F>
void run_threads(int threads, size_t items, F&& fn) {
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i = next++; i < items; i = next++) fn(i);
    };
    const int t = (int)std::min<size_t>((size_t)threads, std::max<size_t>(items, 1));
    std::vector<std::thread> pool;
    for (int k = 1; k < t; k++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

int thread_count(int threads) {
    if (threads > 0) return threads;
    const unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

// chunk size: PGD_READER_CHUNK bytes if set (for tests), else about four
// chunks per thread of at least 1 MiB
size_t chunk_bytes(size_t n, int threads) {
    if (const char* s = std::getenv("PGD_READER_CHUNK")) {
        const long long c = std::atoll(s);
        if (c > 0) return (size_t)c;
    }
    return std::max<size_t>((size_t)1 << 20, n / ((size_t)threads * 4) + 1);
}

Parsed parse_lines(const FileView& f, size_t begin, bool stream, const std::string& d, int threads) {
    std::vector<Chunk> ch;
    const size_t cb = chunk_bytes(f.n - begin, threads);
    for (size_t b = begin; b < f.n;) {
        size_t e = std::min(f.n, b + cb);
        if (e < f.n) {  // move the cut to the start of the next line
            const void* nl = std::memchr(f.p + e - 1, '\n', f.n - (e - 1));
            e = nl ? (size_t)((const char*)nl - f.p) + 1 : f.n;
        }
        Chunk c;
        c.b = b;
        c.e = e;
        ch.push_back(std::move(c));
        b = e;
    }
    run_threads(threads, ch.size(), [&](size_t i) {
        Chunk& c = ch[i];
        for (size_t pos = c.b; pos < c.e;) {
            const char* s;
            size_t len;
            pos = f.line(pos, s, len);
            if (len == 0 || comment(s, len)) continue;
            int32_t v = 0, u = 0;
            const uint8_t fl = stream ? parse_stream(s, len, v, u) : parse_tokens(s, len, d, v, u);
            c.v.push_back(v);
            c.u.push_back(u);
            c.set.push_back(fl);
        }
    });
    size_t total = 0;
    for (auto& c : ch) total += c.v.size();
    Parsed out;
    out.v.resize(total);
    out.u.resize(total);
    out.set.resize(total);
    std::vector<size_t> off(ch.size() + 1, 0);
    for (size_t i = 0; i < ch.size(); i++) off[i + 1] = off[i] + ch[i].v.size();
    run_threads(threads, ch.size(), [&](size_t i) {
        std::copy(ch[i].v.begin(), ch[i].v.end(), out.v.begin() + off[i]);
        std::copy(ch[i].u.begin(), ch[i].u.end(), out.u.begin() + off[i]);
        std::copy(ch[i].set.begin(), ch[i].set.end(), out.set.begin() + off[i]);
        std::vector<int32_t>().swap(ch[i].v);
        std::vector<int32_t>().swap(ch[i].u);
        std::vector<uint8_t>().swap(ch[i].set);
    });
    return out;
}

// Carried values in line order over entries [lo, hi) of x: an entry whose
// `bit` is clear takes the value before it (`seed` for the first).  With
// write, x is resolved in place; every resolved value is passed to agg(v)
// (called from several threads: agg must reduce per slot `k`).  Returns the
// last resolved value (seed when the range is empty).
template <class Agg>
int32_t resolve(std::vector<int32_t>& x, const std::vector<uint8_t>& set, uint8_t bit, size_t lo, size_t hi,
                int32_t seed, bool write, int threads, Agg&& agg) {
    if (hi <= lo) return seed;
    const size_t parts = (size_t)threads * 4, step = (hi - lo + parts - 1) / parts;
    const size_t np = (hi - lo + step - 1) / step;
    std::vector<int32_t> last(np);
    std::vector<uint8_t> has(np, 0);
    run_threads(threads, np, [&](size_t k) {
        const size_t b = lo + k * step, e = std::min(hi, b + step);
        for (size_t i = e; i > b; i--)
            if (set[i - 1] & bit) {
                last[k] = x[i - 1];
                has[k] = 1;
                break;
            }
    });
    std::vector<int32_t> in(np);
    int32_t cur = seed;
    for (size_t k = 0; k < np; k++) {
        in[k] = cur;
        if (has[k]) cur = last[k];
    }
    run_threads(threads, np, [&](size_t k) {
        const size_t b = lo + k * step, e = std::min(hi, b + step);
        int32_t c = in[k];
        for (size_t i = b; i < e; i++) {
            if (set[i] & bit) c = x[i];
            if (write) x[i] = c;
            agg(k, c);
        }
    });
    return cur;
}

// every entry of [lo, hi) sets both of its fields
bool all_set(const Parsed& L, size_t lo, size_t hi, int threads) {
    if (hi <= lo) return true;
    const size_t step = 1 << 16, np = (hi - lo + step - 1) / step;
    std::vector<uint8_t> ok(np, 1);
    run_threads(threads, np, [&](size_t c) {
        for (size_t i = lo + c * step; i < std::min(hi, lo + (c + 1) * step); i++)
            if (L.set[i] != 3) {
                ok[c] = 0;
                return;
            }
    });
    return std::all_of(ok.begin(), ok.end(), [](uint8_t x) { return x != 0; });
}

}  // namespace

namespace pgd {

pgd_status read_raw_lines(const char* path, int threads, RawLines& out) {
    FileView f;
    if (!path || !f.open(path)) return PGD_ERR_INVALID;
    threads = thread_count(threads);
    const std::string p(path);
    out = RawLines();
    out.mtx = extension(p) == "mtx";
    // the leading comment block and the first data line
    size_t pos = 0, first = 0;
    const char* s = "";
    size_t len = 0;
    bool have = false;
    std::string delim;
    while (pos < f.n) {
        first = pos;
        pos = f.line(pos, s, len);
        if (!comment(s, len)) {
            have = true;
            break;
        }
        static const char tag[] = "MatrixMarket matrix coordinate pattern symmetric";
        if (!out.mtx && len >= sizeof(tag) - 1 && memmem(s, len, tag, sizeof(tag) - 1)) delim = " ";
    }
    if (!have) {
        s = "";
        len = 0;
        pos = f.n;
    }
    // `>>` on the first data line (both readers); v, u start at -1
    int32_t v0 = -1, u0 = -1;
    bool fix = true;
    if (len > 0) {
        parse_stream(s, len, v0, u0);
        if (v0 == 0 || u0 == 0) fix = false;
    }
    if (out.mtx) {
        // the header is not an edge; with a 0 in it no line is read (the
        // reference's read loop stops on fix_start_idx)
        if (!fix) return PGD_OK;
        Parsed L = parse_lines(f, pos, true, delim, threads);
        const size_t k = L.v.size();
        if (!all_set(L, 0, k, threads)) {
            // v, u are decremented in place (graphlet_core.cpp:505-506): an
            // unset field carries the previous line's decremented value
            int32_t v = v0, u = u0;
            int64_t maxid = -1;
            out.final_ids = true;
            for (size_t i = 0; i < k; i++) {
                if (L.set[i] & 1) v = L.v[i];
                if (L.set[i] & 2) u = L.u[i];
                v--;
                u--;
                if (v == u) { out.loops++; continue; }
                if (v < 0 || u < 0) return PGD_ERR_INVALID;  // the reference never terminates on these
                out.v.push_back(v);
                out.u.push_back(u);
                maxid = std::max<int64_t>(maxid, std::max(v, u));
            }
            out.n = maxid + 1;
            return PGD_OK;
        }
        out.v = std::move(L.v);
        out.u = std::move(L.u);
        run_threads(threads, (k + 65535) / 65536, [&](size_t c) {
            for (size_t i = c * 65536; i < std::min(k, (c + 1) * 65536); i++) {  // ids are decremented
                out.v[i]--;
                out.u[i]--;
            }
        });
        return PGD_OK;
    }
    if (delim.empty()) {  // detect_delim, graphlet_core.cpp:196-226
        if (std::memchr(s, ',', len)) delim = ",";
        else if (std::memchr(s, '\t', len)) delim = "\t";
        else if (memmem(s, len, "   ", 3)) delim = "   ";
        else if (std::memchr(s, ' ', len)) delim = " ";
    }
    if (delim.empty()) {
        std::string ext = extension(p);
        delim = ext == "csv" ? "," : ext == "tab" ? "\t" : " ";
    }
    // every data line from the first one on, parsed with get_token; entry 0
    // is the first data line when there is one
    Parsed L = parse_lines(f, have ? first : f.n, false, delim, threads);
    const size_t k = L.v.size();
    // the first pass starts after the first line (an empty first line is
    // not an entry: it ended the comment block and is skipped later)
    const size_t b1 = have && len > 0 ? 1 : 0;
    // first pass: any 0 id (then ids start at 0) and the largest id
    const size_t parts = (size_t)threads * 4 + 1;
    std::vector<int32_t> mx(parts, 0);
    std::vector<uint8_t> zero(parts, 0);
    auto agg = [&](size_t slot, int32_t x) {
        if (x == 0) zero[slot] = 1;
        if (x > mx[slot]) mx[slot] = x;
    };
    const int32_t v1 = resolve(L.v, L.set, 1, b1, k, v0, false, threads, agg);
    const int32_t u1 = resolve(L.u, L.set, 2, b1, k, u0, false, threads, agg);
    int32_t maxv = 0;
    for (size_t i = 0; i < parts; i++) {
        if (zero[i]) fix = false;
        maxv = std::max(maxv, mx[i]);
    }
    out.maxv = maxv;
    // second pass: the same lines from the first one on, seeded with the
    // first pass's last values
    if (!all_set(L, 0, k, threads)) {
        // v, u are decremented and renumbered in place
        // (graphlet_core.cpp:413-424): an unset field carries the previous
        // line's values after both
        std::vector<int32_t> lookup((size_t)maxv + 1, -1);
        int32_t v = v1, u = u1, next_id = 0;
        out.final_ids = true;
        for (size_t i = 0; i < k; i++) {
            if (L.set[i] & 1) v = L.v[i];
            if (L.set[i] & 2) u = L.u[i];
            if (fix) {
                v--;
                u--;
            }
            if (v == u) { out.loops++; continue; }
            if (v < 0 || u < 0 || v > maxv || u > maxv) return PGD_ERR_INVALID;  // out of the reference's table
            if (lookup[(size_t)v] < 0) lookup[(size_t)v] = next_id++;
            v = lookup[(size_t)v];
            if (lookup[(size_t)u] < 0) lookup[(size_t)u] = next_id++;
            u = lookup[(size_t)u];
            out.v.push_back(v);
            out.u.push_back(u);
        }
        out.n = next_id;
        return PGD_OK;
    }
    if (fix)
        run_threads(threads, (k + 65535) / 65536, [&](size_t c) {
            for (size_t i = c * 65536; i < std::min(k, (c + 1) * 65536); i++) {
                L.v[i]--;
                L.u[i]--;
            }
        });
    out.v = std::move(L.v);
    out.u = std::move(L.u);
    return PGD_OK;
}

}  // namespace pgd

extern "C" pgd_status pgd_read_graph(const char* path, int64_t* n, int64_t* nlines, int32_t** src, int32_t** dst,
                                     int32_t* dedup, int64_t* self_loops) {
    if (!path || !n || !nlines || !src || !dst) return PGD_ERR_INVALID;
    pgd::RawLines R;
    pgd_status st = pgd::read_raw_lines(path, 0, R);
    if (st != PGD_OK) return st;
    // self loops dropped; edge-list ids renumbered in order of first
    // appearance (graphlet_core.cpp:413-428); .mtx: n = 1 + largest endpoint
    std::vector<int32_t> a, b;
    int64_t loops = 0, nv = 0;
    if (R.final_ids) {
        a = std::move(R.v);
        b = std::move(R.u);
        loops = R.loops;
        nv = R.n;
    } else if (R.mtx) {
        a.reserve(R.v.size());
        b.reserve(R.v.size());
        int64_t maxid = -1;
        for (size_t i = 0; i < R.v.size(); i++) {
            const int32_t v = R.v[i], u = R.u[i];
            if (v == u) { loops++; continue; }
            if (v < 0 || u < 0) return PGD_ERR_INVALID;  // the reference never terminates on these
            a.push_back(v);
            b.push_back(u);
            maxid = std::max<int64_t>(maxid, std::max(v, u));
        }
        nv = maxid + 1;
    } else {
        a.reserve(R.v.size());
        b.reserve(R.v.size());
        std::vector<int32_t> lookup((size_t)R.maxv + 1, -1);
        int32_t next_id = 0;
        for (size_t i = 0; i < R.v.size(); i++) {
            const int32_t v = R.v[i], u = R.u[i];
            if (v == u) { loops++; continue; }
            if (v < 0 || u < 0 || v > R.maxv || u > R.maxv) return PGD_ERR_INVALID;  // out of the reference's table
            if (lookup[(size_t)v] < 0) lookup[(size_t)v] = next_id++;
            const int32_t x = lookup[(size_t)v];
            if (lookup[(size_t)u] < 0) lookup[(size_t)u] = next_id++;
            a.push_back(x);
            b.push_back(lookup[(size_t)u]);
        }
        nv = next_id;
    }
    const size_t k = a.size();
    *src = (int32_t*)std::malloc((k ? k : 1) * sizeof(int32_t));
    *dst = (int32_t*)std::malloc((k ? k : 1) * sizeof(int32_t));
    if (!*src || !*dst) {
        std::free(*src);
        std::free(*dst);
        return PGD_ERR_OOM;
    }
    if (k) {
        std::memcpy(*src, a.data(), k * sizeof(int32_t));
        std::memcpy(*dst, b.data(), k * sizeof(int32_t));
    }
    *n = nv;
    *nlines = (int64_t)k;
    if (dedup) *dedup = R.mtx ? 0 : 1;
    if (self_loops) *self_loops = loops;
    return PGD_OK;
}

extern "C" void pgd_free(void* p) { std::free(p); }
