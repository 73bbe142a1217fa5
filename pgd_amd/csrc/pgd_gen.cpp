// pgd_gen.cpp -- seeded synthetic graph generators and a Matrix Market writer.
//
// Inputs for the bench and the tests: Erdos-Renyi G(n,p), Barabasi-Albert and
// R-MAT/Kronecker graphs.  Every generator is a pure function of its
// parameters and seed (counter-based RNG), so the same call yields the same
// edge list on any machine and thread count.  Edge lists are undirected
// simple graphs (no self loops, no duplicates), written as (src, dst) pairs
// with 0-based ids in a fixed order, which is the order a Matrix Market file
// written from them lists the edges.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pgd_hip.h"

namespace {

inline uint64_t splitmix(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}
// uniform double in [0,1) from a 64-bit word
inline double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(splitmix(seed ^ 0x5851f42d4c957f2dULL)) {}
    uint64_t next() { s = splitmix(s); return s; }
    double uni() { return u01(next()); }
};

}  // namespace

extern "C" {

// G(n,p) by geometric skipping over the pairs (i<j) in row-major order.
// Returns the number of edges (written when cap allows).
int64_t pgd_gen_gnp(int32_t n, double p, uint64_t seed, int32_t* src, int32_t* dst, int64_t cap) {
    if (n < 2 || p <= 0.0) return 0;
    Rng r(seed);
    int64_t m = 0;
    if (p >= 1.0) {
        for (int32_t i = 0; i < n; i++)
            for (int32_t j = i + 1; j < n; j++) {
                if (m < cap) { src[m] = i; dst[m] = j; }
                m++;
            }
        return m;
    }
    const double lq = std::log1p(-p);
    int64_t total = (int64_t)n * (n - 1) / 2;
    int64_t k = -1;
    int32_t i = 0;
    int64_t row_start = 0;   // linear index of pair (i, i+1)
    while (true) {
        double u = r.uni();
        int64_t skip = (int64_t)std::floor(std::log1p(-u) / lq);
        k += 1 + skip;
        if (k >= total) break;
        while (k >= row_start + (n - 1 - i)) { row_start += n - 1 - i; i++; }
        int32_t j = (int32_t)(i + 1 + (k - row_start));
        if (m < cap) { src[m] = i; dst[m] = j; }
        m++;
    }
    return m;
}

// Barabasi-Albert: start from a clique on m0 = mm+1 vertices, then each new
// vertex attaches to mm distinct targets chosen proportionally to degree.
int64_t pgd_gen_ba(int32_t n, int32_t mm, uint64_t seed, int32_t* src, int32_t* dst, int64_t cap) {
    if (n < 2 || mm < 1) return 0;
    Rng r(seed);
    std::vector<int32_t> rep;  // each endpoint once per incident edge
    int64_t m = 0;
    int32_t m0 = std::min(n, mm + 1);
    for (int32_t i = 0; i < m0; i++)
        for (int32_t j = i + 1; j < m0; j++) {
            if (m < cap) { src[m] = i; dst[m] = j; }
            m++;
            rep.push_back(i); rep.push_back(j);
        }
    std::vector<int32_t> picked;
    for (int32_t v = m0; v < n; v++) {
        picked.clear();
        while ((int32_t)picked.size() < mm) {
            int32_t t = rep[(size_t)(r.next() % rep.size())];
            if (std::find(picked.begin(), picked.end(), t) == picked.end()) picked.push_back(t);
        }
        for (int32_t t : picked) {
            if (m < cap) { src[m] = t; dst[m] = v; }
            m++;
            rep.push_back(t); rep.push_back(v);
        }
    }
    return m;
}

// R-MAT: `ef * 2^scale` samples, each descending `scale` quadrant levels with
// probabilities (a, b, c, 1-a-b-c); symmetrised, with self loops and
// duplicates removed; edges are returned sorted by (min, max) endpoint.
int64_t pgd_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed,
                     int32_t* src, int32_t* dst, int64_t cap) {
    if (scale < 1 || scale > 30) return -1;
    const int64_t nsamp = (int64_t)ef << scale;
    std::vector<uint64_t> ed((size_t)nsamp);
    const double ab = a + b, abc = a + b + c;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nsamp; i++) {
        uint64_t s = splitmix(seed * 0x100000001b3ULL + (uint64_t)i);
        uint64_t x = 0, y = 0;
        for (int l = 0; l < scale; l++) {
            s = splitmix(s);
            double q = u01(s);
            int bx, by;
            if (q < a) { bx = 0; by = 0; }
            else if (q < ab) { bx = 0; by = 1; }
            else if (q < abc) { bx = 1; by = 0; }
            else { bx = 1; by = 1; }
            x = (x << 1) | (uint64_t)bx;
            y = (y << 1) | (uint64_t)by;
        }
        if (x > y) std::swap(x, y);
        ed[(size_t)i] = (x == y) ? ~0ULL : ((x << 32) | y);
    }
    std::sort(ed.begin(), ed.end());
    ed.erase(std::unique(ed.begin(), ed.end()), ed.end());
    if (!ed.empty() && ed.back() == ~0ULL) ed.pop_back();
    int64_t m = (int64_t)ed.size();
    int64_t lim = std::min(m, cap);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < lim; i++) {
        src[i] = (int32_t)(ed[(size_t)i] >> 32);
        dst[i] = (int32_t)(ed[(size_t)i] & 0xffffffffULL);
    }
    return m;
}

// Symmetric pattern Matrix Market file, 1-based, one line per edge in the
// given order; the header's n is `n` (the reader ignores it).
int32_t pgd_write_mtx(const char* path, int32_t n, int64_t m, const int32_t* src, const int32_t* dst) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    std::vector<char> buf(1 << 20);
    std::setvbuf(f, buf.data(), _IOFBF, buf.size());
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate pattern symmetric\n%d %d %lld\n", n, n, (long long)m);
    for (int64_t i = 0; i < m; i++) std::fprintf(f, "%d %d\n", src[i] + 1, dst[i] + 1);
    std::fclose(f);
    return 0;
}

}  // extern "C"
