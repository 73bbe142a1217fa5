// pgd_epilogue.hpp -- per-edge equations of the counting epilogue, shared by
// the plain and the batched paths: solve_graphlet_equations
// (graphlet_core.cpp:1325-1364), the micro columns (1559-1575) and the terms
// of the 15 sums (1429-1448), in uint64 wrap-around arithmetic.
#pragma once
#include "pgd_internal.hpp"

namespace pgd {

// (double)p * 0.5 converted back to uint64, for even p: exact below 2^53,
// otherwise p rounded to 53 significant bits (nearest, ties to even) halved.
__device__ __forceinline__ uint64_t half_rne(uint64_t p) {
    if (p < (1ULL << 53)) return p >> 1;
    int L = 64 - __clzll((long long)p);
    int sh = L - 53;
    uint64_t q = p >> sh;
    uint64_t rem = p & ((1ULL << sh) - 1);
    uint64_t hf = 1ULL << (sh - 1);
    if (rem > hf || (rem == hf && (q & 1))) q++;
    return q << (sh - 1);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

constexpr uint64_t kExact = 1ULL << 53;

// The binary64 value nearest p (ties to even), as an integer: what the
// reference's (double)p holds.  For p < 2^63 the result fits 64 bits.
__device__ __forceinline__ uint64_t rne53(uint64_t p) {
    if (p < kExact) return p;
    const int L = 64 - __clzll((long long)p);
    const int sh = L - 53;
    uint64_t q = p >> sh;
    const uint64_t rem = p & ((1ULL << sh) - 1), hf = 1ULL << (sh - 1);
    if (rem > hf || (rem == hf && (q & 1))) q++;
    return q << sh;
}

// One edge at E_ordered position e of a graph with n vertices and m edges:
// adds its terms to acc[15]; with micro != nullptr writes its 8 columns at
// micro[col * mstride + e].
//
// The micro columns follow graphlet_core.cpp:1566-1575 value for value,
// including the binary64 steps of the reference above 2^53 (vertex ids are
// int32, so T, Sv, Su < 2^31 and every product below is < 2^62):
//   chordal = (double)(T(T-1)) / 2.0 - (double)K4, stored as uint64: both
//     operands are integers, so the difference is rounded once and exact
//     below 2^53; C(T,2) >= K4 keeps it non-negative;
//   3-star  = (uint64)((double)(Sv(Sv-1)) / 2.0), then that value plus
//     (double)(Su(Su-1)) / 2.0, rounded once (the first term has at most 53
//     significant bits, so its conversion back to double is exact).
__device__ __forceinline__ void edge_terms(int64_t dv, int64_t du, uint64_t T, uint64_t K, uint64_t C, int64_t n,
                                           int64_t m, uint64_t* acc, uint64_t* __restrict__ micro, int64_t mstride,
                                           int64_t e) {
    uint64_t sv = (uint64_t)dv - T - 1, su = (uint64_t)du - T - 1;
    uint64_t S = sv + su;
    uint64_t D = (uint64_t)n - (S + T) - 2;
    acc[0] += T;
    acc[1] += S;
    acc[2] += D;
    acc[3] += K;
    acc[4] += C;
    acc[5] += half_rne((T + S) * (T + S - 1));
    acc[6] += half_rne(S * (S - 1));
    acc[7] += half_rne(T * (T - 1));
    acc[8] += T * S;
    acc[9] += sv * su;
    acc[10] += sv * (sv - 1) / 2 + su * (su - 1) / 2;
    acc[11] += (uint64_t)(m - dv - du + 1);
    acc[12] += half_rne(D * (D - 1));
    acc[13] += S * D;
    acc[14] += T * D;
    if (micro) {
        const uint64_t tt = T * (T - 1);
        const int64_t svs = (int64_t)sv, sus = (int64_t)su;
        const uint64_t a3 = (uint64_t)(svs * (svs - 1)), b3 = (uint64_t)(sus * (sus - 1));
        micro[(size_t)PGD_MICRO_TRI * mstride + e] = T;
        micro[(size_t)PGD_MICRO_CLIQUE * mstride + e] = K;
        micro[(size_t)PGD_MICRO_CHORDAL * mstride + e] = rne53(half_rne(tt) - rne53(K));
        micro[(size_t)PGD_MICRO_TAILED * mstride + e] = 0;
        micro[(size_t)PGD_MICRO_CYCLE * mstride + e] = C;
        micro[(size_t)PGD_MICRO_STAR3 * mstride + e] = rne53(half_rne(a3) + half_rne(b3));
        micro[(size_t)PGD_MICRO_PATH * mstride + e] = (uint64_t)(svs * sus) - C;
        micro[(size_t)PGD_MICRO_STAR2 * mstride + e] = S;
    }
}

// block-wide sum of acc[15] (blockDim.x a multiple of 64, at most 1024);
// thread k < 15 adds the block's k-th sum to out[k]
__device__ __forceinline__ void block_add_sums(const uint64_t* acc, uint64_t* __restrict__ out) {
    __shared__ uint64_t red[PGD_NSUMS][16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < PGD_NSUMS; k++) {
        uint64_t s = wave_sum_u64(acc[k]);
        if (lane == 0) red[k][w] = s;
    }
    __syncthreads();
    if (threadIdx.x < PGD_NSUMS) {
        uint64_t s = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[threadIdx.x][i];
        atomicAdd((unsigned long long*)&out[threadIdx.x], (unsigned long long)s);
    }
    __syncthreads();
}

}  // namespace pgd
