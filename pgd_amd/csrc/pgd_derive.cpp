// pgd_derive.cpp -- the closed-form derivation of the 17 graphlet totals in
// integer arithmetic only.
//
// The reference computes the totals from the 15 reduced sums with a few
// binary64 operations (graphlet_core.cpp:1449-1472; graphlet_core.h:548-557).
// Its build (g++ -O3 -ffast-math, Makefile:21-34) compiles them as:
//   uint64 -> double  : round to nearest, ties to even
//   x / 2.0, x / 4.0  : multiplication by 0.5, 0.25 (exact)
//   x / 3.0, x / 6.0  : multiplication by fl(1/3) = 0x3fd5555555555555 and
//                       fl(1/6) = 0x3fc5555555555555, rounded to nearest even
//   double -> uint64  : truncation toward zero (x86-64 cvttsd2si sequence)
//   n*(n-1)*(n-2)/6   : signed 64-bit product (wraps) and truncating division
//   n*(n-1)*(n-2)*(n-3)/24 : unsigned 64-bit product (wraps) and division
// This file reproduces each of those operations exactly with integers, so the
// totals are bit-identical to the reference's whatever their magnitude.
#include <cstdint>

#include "pgd_hip.h"

namespace {

typedef unsigned __int128 u128;
typedef __int128 i128;

// A finite binary64 value: (-1)^neg * mant * 2^exp with mant < 2^53.
struct F64 {
    bool neg;
    uint64_t mant;
    int exp;
};

int bitlen(u128 x) {
    int n = 0;
    while (x) { x >>= 1; n++; }
    return n;
}

// Round mag * 2^e to 53 significant bits, ties to even.
F64 round53(bool neg, u128 mag, int e) {
    if (mag == 0) return F64{neg, 0, 0};
    int L = bitlen(mag);
    if (L <= 53) return F64{neg, (uint64_t)mag, e};
    int sh = L - 53;
    u128 q = mag >> sh;
    u128 rem = mag - (q << sh);
    u128 half = (u128)1 << (sh - 1);
    if (rem > half || (rem == half && (q & 1))) q++;
    if (q >> 53) { q >>= 1; sh++; }
    return F64{neg, (uint64_t)q, e + sh};
}

F64 from_u64(uint64_t x) { return round53(false, x, 0); }
F64 from_i64(int64_t x) {
    bool neg = x < 0;
    uint64_t mag = neg ? (uint64_t)0 - (uint64_t)x : (uint64_t)x;
    return round53(neg, mag, 0);
}

// a * (mant_c * 2^exp_c), correctly rounded
F64 mul(F64 a, uint64_t mant_c, int exp_c) {
    u128 p = (u128)a.mant * (u128)mant_c;
    return round53(a.neg, p, a.exp + exp_c);
}

// a - b, correctly rounded (both operands are integers or halves here, so the
// exact difference fits comfortably in 128 bits after alignment)
F64 sub(F64 a, F64 b) {
    int e = a.exp < b.exp ? a.exp : b.exp;
    i128 va = (i128)((u128)a.mant << (a.exp - e));
    i128 vb = (i128)((u128)b.mant << (b.exp - e));
    if (a.neg) va = -va;
    if (b.neg) vb = -vb;
    i128 d = va - vb;
    bool neg = d < 0;
    return round53(neg, (u128)(neg ? -d : d), e);
}

// trunc(|x|) as an unsigned 128-bit integer (x integer-valued or not)
u128 trunc_mag(F64 x) {
    if (x.exp >= 0) return (u128)x.mant << x.exp;
    if (x.exp <= -64) return 0;
    return (u128)(x.mant >> (-x.exp));
}

// double -> unsigned long long as the x86-64 build converts it:
// x >= 2^63: cvttsd2si(x - 2^63) ^ 2^63 (0 when x >= 2^64);
// otherwise cvttsd2si(x), whose out-of-range result is 2^63.
uint64_t to_u64(F64 x) {
    const u128 two63 = (u128)1 << 63;
    u128 t = trunc_mag(x);
    if (!x.neg) {
        if (t >= two63) {
            u128 y = t - two63;          // x - 2^63 is exact for these x
            if (y >= two63) return 0;    // 0x8000000000000000 ^ 2^63
            return (uint64_t)y ^ 0x8000000000000000ULL;
        }
        return (uint64_t)t;
    }
    if (t > two63) return 0x8000000000000000ULL;
    return (uint64_t)0 - (uint64_t)t;
}

const uint64_t kThirdMant = 0x15555555555555ULL;  // fl(1/3) = kThirdMant * 2^-54
const uint64_t kSixthMant = 0x15555555555555ULL;  // fl(1/6) = kSixthMant * 2^-55

uint64_t half(uint64_t x) { return to_u64(mul(from_u64(x), 1, -1)); }
uint64_t quarter(uint64_t x) { return to_u64(mul(from_u64(x), 1, -2)); }
uint64_t third(uint64_t x) { return to_u64(mul(from_u64(x), kThirdMant, -54)); }
uint64_t sixth(uint64_t x) { return to_u64(mul(from_u64(x), kSixthMant, -55)); }

}  // namespace

extern "C" pgd_status pgd_derive(int64_t n, int64_t m, const uint64_t* s, uint64_t* t) {
    if (!s || !t) return PGD_ERR_INVALID;
    const uint64_t ST = s[0], SS = s[1], SD = s[2], SK = s[3], SC = s[4];
    const uint64_t N3 = s[7], N4 = s[8], N5 = s[9], N6 = s[10], N7 = s[11], N8 = s[12], N9 = s[13],
                   N10 = s[14];
    const uint64_t un = (uint64_t)n;
    // 1449-1450
    uint64_t e1 = (uint64_t)m;
    int64_t nn1 = (int64_t)(un * (un - 1));
    uint64_t i2 = to_u64(sub(mul(from_i64(nn1), 1, -1), from_i64(m)));
    // 1451-1455
    uint64_t tris = third(ST);
    uint64_t s2 = half(SS);
    uint64_t e31 = SD;
    int64_t p3 = (int64_t)(un * (un - 1) * (un - 2));
    uint64_t i3 = (uint64_t)(p3 / 6);
    i3 = i3 - (tris + s2 + e31);
    // 1458-1471
    uint64_t k4 = sixth(SK);
    uint64_t cc = N3 - 6 * k4;
    uint64_t c4 = quarter(SC);
    uint64_t p4 = N5 - 4 * c4;
    uint64_t tt = half(N4 - 4 * cc);
    uint64_t st3 = third(N6 - tt);
    uint64_t e2 = half(N7 - 6 * k4 - 4 * c4 - 4 * cc - 2 * p4 - 2 * tt);
    uint64_t e14 = N8 - 2 * e2;
    uint64_t st24 = half(N9 - 2 * p4);
    uint64_t tri4 = third(N10 - tt);
    // graphlet_core.h:548-557
    uint64_t i4 = (un * (un - 1) * (un - 2) * (un - 3)) / 24;
    i4 = i4 - (k4 + cc + c4 + p4 + tt + st3);
    i4 = i4 - (e2 + e14 + st24 + tri4);
    const uint64_t out[PGD_NTOTALS] = {e1, i2, tris, s2, e31, i3, k4, cc, tt, c4, st3, p4, e14, e2, st24, tri4, i4};
    for (int i = 0; i < PGD_NTOTALS; i++) t[i] = out[i];
    return PGD_OK;
}
