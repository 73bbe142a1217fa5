// host_checks.cpp -- the host code of the library (reader, derivation,
// generators, .mtx writer) driven under the compiler's sanitizers: built by
// tests/test_sanitize.py once with AddressSanitizer + UndefinedBehaviorSanitizer
// and once with ThreadSanitizer, and run on the reader's test files.  Every
// check here is also a result check, so a sanitizer build that runs clean has
// also reproduced the plain build's answers.
//
// usage: host_checks [--no-gen] file...
//   each file is read by read_raw_lines with 1, 2, 3 and 8 threads (results
//   must agree) and by pgd_read_graph; then the derivation runs on edge-value
//   and random inputs, and (unless --no-gen) the generators and the .mtx
//   writer round trip run.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "pgd_reader.hpp"

namespace {

int failures = 0;

void fail(const char* what, const std::string& ctx) {
    std::fprintf(stderr, "FAIL %s: %s\n", what, ctx.c_str());
    failures++;
}

bool same(const pgd::RawLines& a, const pgd::RawLines& b) {
    return a.mtx == b.mtx && a.maxv == b.maxv && a.v == b.v && a.u == b.u && a.final_ids == b.final_ids &&
           a.n == b.n && a.loops == b.loops;
}

void check_file(const char* path) {
    pgd::RawLines ref;
    if (pgd::read_raw_lines(path, 1, ref) != PGD_OK) {
        fail("read_raw_lines", path);
        return;
    }
    for (int t : {2, 3, 8}) {
        pgd::RawLines r;
        if (pgd::read_raw_lines(path, t, r) != PGD_OK || !same(ref, r))
            fail("thread counts disagree", std::string(path) + " threads " + std::to_string(t));
    }
    int64_t n = 0, nl = 0, loops = 0;
    int32_t *s = nullptr, *d = nullptr, dedup = 0;
    if (pgd_read_graph(path, &n, &nl, &s, &d, &dedup, &loops) != PGD_OK) {
        fail("pgd_read_graph", path);
        return;
    }
    uint64_t h = 1469598103934665603ull;
    for (int64_t i = 0; i < nl; i++) {
        if (s[i] < 0 || s[i] >= n || d[i] < 0 || d[i] >= n || s[i] == d[i]) fail("id out of range", path);
        h = (h ^ (uint64_t)(uint32_t)s[i]) * 1099511628211ull;
        h = (h ^ (uint64_t)(uint32_t)d[i]) * 1099511628211ull;
    }
    std::printf("%s n=%" PRId64 " lines=%" PRId64 " loops=%" PRId64 " dedup=%d hash=%016" PRIx64 "\n", path, n, nl,
                loops, dedup, h);
    pgd_free(s);
    pgd_free(d);
}

void check_derive() {
    // edge values of the 15 sums (0, 1, powers of two around 2^53 and 2^63,
    // the top of the range) and random ones; n, m over small and large values
    std::vector<uint64_t> vals = {0, 1, 2, 3, 5, 6, 7, (1ull << 24) - 1, 1ull << 24, (1ull << 53) - 1, 1ull << 53,
                                  (1ull << 53) + 1, (1ull << 63) - 1, 1ull << 63, ~0ull - 1, ~0ull};
    std::mt19937_64 rng(7);
    uint64_t sums[PGD_NSUMS], tot[PGD_NTOTALS];
    const int64_t ns[] = {0, 1, 2, 3, 1000, 1 << 24, (int64_t)1 << 31, ((int64_t)1 << 40) + 3, INT64_MAX / 2};
    int runs = 0;
    for (int64_t n : ns)
        for (int rep = 0; rep < 400; rep++) {
            for (int k = 0; k < PGD_NSUMS; k++) sums[k] = rep < 200 ? vals[(rep + 3 * k) % vals.size()] : rng();
            const int64_t m = rep % 3 == 0 ? 0 : (int64_t)(rng() >> 2);
            const pgd_status st = pgd_derive(n, m, sums, tot);
            if (st != PGD_OK) fail("pgd_derive", std::to_string(n));
            runs++;
        }
    std::printf("derive runs=%d\n", runs);
}

void check_generators(const char* tmpdir) {
    std::vector<int32_t> s(1 << 20), d(1 << 20);
    const int64_t g = pgd_gen_gnp(2000, 0.01, 1, s.data(), d.data(), (int64_t)s.size());
    const int64_t b = pgd_gen_ba(3000, 4, 1, s.data(), d.data(), (int64_t)s.size());
    const int64_t r = pgd_gen_rmat(14, 16, 0.57, 0.19, 0.19, 1, s.data(), d.data(), (int64_t)s.size());
    if (g <= 0 || b <= 0 || r <= 0) fail("generators", "sizes");
    // a capacity too small is reported, not overrun
    std::vector<int32_t> s2(10), d2(10);
    const int64_t small = pgd_gen_rmat(14, 16, 0.57, 0.19, 0.19, 1, s2.data(), d2.data(), 10);
    if (small <= 10 && small >= 0) fail("generators", "a small capacity was not reported");
    const std::string path = std::string(tmpdir) + "/rt.mtx";
    int32_t n = 0;
    for (int64_t i = 0; i < r; i++) n = std::max(n, std::max(s[i], d[i]) + 1);
    if (pgd_write_mtx(path.c_str(), n, r, s.data(), d.data()) != 0) fail("pgd_write_mtx", path);
    int64_t rn = 0, rl = 0, loops = 0;
    int32_t *rs = nullptr, *rd = nullptr, dedup = 0;
    if (pgd_read_graph(path.c_str(), &rn, &rl, &rs, &rd, &dedup, &loops) != PGD_OK) {
        fail("read back", path);
        return;
    }
    int64_t kept = 0;
    for (int64_t i = 0; i < r; i++) kept += s[i] != d[i];
    if (rl != kept || rn != n) fail("round trip", path);
    for (int64_t i = 0, j = 0; i < r && j < rl; i++)
        if (s[i] != d[i]) {
            if (rs[j] != s[i] || rd[j] != d[i]) {
                fail("round trip ids", path);
                break;
            }
            j++;
        }
    pgd_free(rs);
    pgd_free(rd);
    std::printf("generators gnp=%" PRId64 " ba=%" PRId64 " rmat=%" PRId64 " roundtrip=%" PRId64 "\n", g, b, r, rl);
}

}  // namespace

int main(int argc, char** argv) {
    bool gen = true;
    const char* tmpdir = std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp";
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--no-gen")) {
            gen = false;
            continue;
        }
        check_file(argv[i]);
    }
    check_derive();
    if (gen) check_generators(tmpdir);
    std::printf("failures=%d\n", failures);
    return failures ? 1 : 0;
}
