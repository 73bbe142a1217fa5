// ref_arith.cpp -- the reference's floating-point expressions, evaluated with
// the reference's types under the reference's compiler flags (oracle/Makefile
// REF_CXXFLAGS, as the reference Makefile:21-34), for checking the integer
// emulation of them in pgd_derive and in the device epilogue.  Test tool only.
//
// stdin lines:
//   D n m s0 .. s14   -> the 17 totals of graphlet_core.cpp:1449-1472 and
//                        graphlet_core.h:548-557 (s = the 15 reduced sums in
//                        the order PGD_NSUMS documents)
//   M T K4 C4 dv du   -> the micro columns chordal, 3-star, 4-path of
//                        graphlet_core.cpp:1566-1575
// Written in C++98 to build with the reference's own flags.
#include <cstdio>
#include <cstring>

typedef unsigned long long ull;

// keep the compiler from folding the inputs (they come from stdin anyway)
static void totals(long long n, long long m, const ull* s) {
    ull tri = s[0], star = s[1], e31 = s[2], k4s = s[3], c4s = s[4];
    ull N3 = s[7], N4 = s[8], N5 = s[9], N6 = s[10], N7 = s[11], N8 = s[12], N9 = s[13], N10 = s[14];
    ull t[17];
    t[0] = m;
    t[1] = (n * (n - 1) / 2.0) - m;
    t[2] = tri / 3.0;
    t[3] = star / 2.0;
    t[4] = e31;
    t[5] = (n * (n - 1) * (n - 2)) / (3 * 2);
    t[5] = t[5] - (t[2] + t[3] + t[4]);
    t[6] = k4s / 6.0;
    t[7] = N3 - (6 * t[6]);
    t[9] = c4s / 4.0;
    t[11] = N5 - (4 * t[9]);
    t[8] = (N4 - (4 * t[7])) / 2.0;
    t[10] = (N6 - t[8]) / 3.0;
    t[13] = N7;
    t[13] = t[13] - (6 * t[6]) - (4 * t[9]) - (4 * t[7]) - (2 * t[11]) - (2 * t[8]);
    t[13] = t[13] / 2.0;
    t[12] = N8 - (2 * t[13]);
    t[14] = N9 - (2 * t[11]);
    t[14] = t[14] / 2.0;
    t[15] = N10 - t[8];
    t[15] = t[15] / 3.0;
    ull un = (ull)n;
    t[16] = (un * (un - 1) * (un - 2) * (un - 3)) / (4 * 3 * 2);
    t[16] = t[16] - (t[6] + t[7] + t[9] + t[11] + t[8] + t[10]);
    t[16] = t[16] - (t[13] + t[12] + t[14] + t[15]);
    // printed order: 2_1edge 2_indep 3_tris 2_star 3_1edge 3_indep 4_clique
    // 4_chordcycle 4_tailed_tris 4_cycle 3_star 4_path 4_1edge 4_2edge
    // 4_2star 4_tri 4_indep
    for (int i = 0; i < 17; i++) printf("%llu%c", t[i], i == 16 ? '\n' : ' ');
}

static void micro(ull tri_count, ull k4, ull c4, long long deg_v, long long deg_u) {
    ull chordal, star3, path4;
    chordal = (tri_count * (tri_count - 1) / 2.0) - k4;
    deg_v = deg_v - tri_count - 1;
    deg_u = deg_u - tri_count - 1;
    path4 = (deg_v * deg_u) - c4;
    star3 = (deg_v * (deg_v - 1) / 2.0);
    star3 = star3 + (deg_u * (deg_u - 1) / 2.0);
    star3 = star3 - 0ULL;
    printf("%llu %llu %llu\n", chordal, star3, path4);
}

int main() {
    char kind[8];
    while (scanf("%7s", kind) == 1) {
        if (!strcmp(kind, "D")) {
            long long n, m;
            ull s[15];
            if (scanf("%lld %lld", &n, &m) != 2) return 1;
            for (int i = 0; i < 15; i++)
                if (scanf("%llu", &s[i]) != 1) return 1;
            totals(n, m, s);
        } else if (!strcmp(kind, "M")) {
            ull T, K, C;
            long long dv, du;
            if (scanf("%llu %llu %llu %lld %lld", &T, &K, &C, &dv, &du) != 5) return 1;
            micro(T, K, C, dv, du);
        } else {
            return 1;
        }
    }
    return 0;
}
