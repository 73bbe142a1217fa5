// pgd_engine_dense.hip -- dense-graph tier on the int8 matrix cores.
//
// For a dense graph the per-edge counts are contractions of the adjacency
// matrix A (0/1, held as int8, zero-padded to N = 128*ceil(n/128)):
//
//   S    = A.A                  codegrees; T(e) = S_vu for the edge e = (v,u)
//   A3   = A.S                  P(e) = A3_vu - d_v - d_u + 1 (4-cycles through e)
//   AT   = A.T, T = S o A       Q(e) = AT_vu + AT_uv  (sum over triangles {v,u,w}
//                               of T(v,w) + T(u,w))
//   K4(e) = edges among the common neighbours of v and u (bit rows, popcounts)
//   C4(e) = P - Q + 2T + 2K4    (the oriented engine's identity)
//
// Every product is an int8 x int8 -> int32 MFMA (v_mfma_i32_32x32x32_i8).
// S and T are at most n-2 < 2^13, so they enter the second products as two
// base-127 digits (s0 = S mod 127 in [0, 126], s1 = S div 127 <= 64).  The
// two second products are one launch over K = 2N:
//   [A3 | AT] = [A | 127 A] . [[s0 | s1] ; [t0 | t1]]^T
// (the left operand N x 2N, the right 2N x 2N, both read along rows because
// every factor is symmetric); each int32 sum stays below N (126 + 127 * 64)
// < 2^31 and is the exact value.
#include <mutex>

#include "pgd_orient.hpp"

namespace pgd {
namespace {

constexpr int kDenseMaxN = 8192;  // int32 products of N^2 entries, bit rows of N/32 words
constexpr int kTile = 128;        // output tile of one GEMM block (2 x 2 waves of 64 x 64); N is a multiple

typedef int v16i __attribute__((ext_vector_type(16)));
typedef int v4i __attribute__((ext_vector_type(4)));

__global__ void d_fill_u32(uint32_t* __restrict__ p, int64_t n) {
    GRID_STRIDE(i, n) p[i] = 0;
}

// [A | 127 A] (int8, N x 2N) and the bit rows (N x Wd words) from the CSR
__global__ void k_dense_adj(const int64_t* __restrict__ row, const int32_t* __restrict__ col, int64_t n, int64_t N,
                            int Wd, int8_t* __restrict__ A2, uint32_t* __restrict__ bits) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v = wave; v < n; v += nw)
        for (int64_t p = row[v] + lane; p < row[v + 1]; p += 64) {
            const int32_t u = col[p];
            A2[v * 2 * N + u] = 1;
            A2[v * 2 * N + N + u] = 127;
            atomicOr(&bits[v * Wd + (u >> 5)], 1u << (u & 31));
        }
}

// C = A . X^T over K: C[i][j] = sum_{k < K} A[i lda + k] X[j ldx + k] for int8
// row-major A (M x K used) and X (cols x K used), int32 C (row stride ldc);
// M, cols multiples of 128, K of 64.  One block per 128 x 128 output tile,
// four wavefronts of 64 x 64 (2 x 2 MFMA tiles of 32 x 32).  Each 64-deep
// step of both operands is staged in LDS (rows padded to 80 bytes, so the
// 16-byte fragment reads of 32 consecutive rows spread over the banks),
// double-buffered: the next step's global loads are in flight during the
// current step's products, and one barrier per step separates the buffers.
// MFMA operands: lane l feeds row (l mod 32) and bytes 16 (l / 32) .. + 16 of
// the 32-deep k slice, for both operands; accumulator r of lane l holds row
// (r mod 4) + 8 (r / 4) + 4 (l / 32), column l mod 32.
constexpr int kGK = 64;       // k per staged step
constexpr int kGStride = 80;  // LDS bytes per staged row
__global__ __launch_bounds__(256) void k_gemm_i8(const int8_t* __restrict__ A, int64_t lda,
                                                 const int8_t* __restrict__ X, int64_t ldx,
                                                 int32_t* __restrict__ C, int64_t ldc, int64_t K) {
    __shared__ __align__(16) int8_t As[2][kTile * kGStride];
    __shared__ __align__(16) int8_t Xs[2][kTile * kGStride];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t bi = (int64_t)blockIdx.y * kTile, bj = (int64_t)blockIdx.x * kTile;
    // staging: thread t moves 16-byte chunks t and t + 256 of each 128 x 64 tile
    // (chunk c: row c / 4, bytes 16 (c mod 4) .. + 16)
    const int r0 = tid >> 2, c0 = (tid & 3) * 16;
    const int8_t* ga0 = A + (bi + r0) * lda + c0;
    const int8_t* ga1 = A + (bi + r0 + 64) * lda + c0;
    const int8_t* gx0 = X + (bj + r0) * ldx + c0;
    const int8_t* gx1 = X + (bj + r0 + 64) * ldx + c0;
    const int so0 = r0 * kGStride + c0, so1 = (r0 + 64) * kGStride + c0;
    v4i ra0 = *(const v4i*)ga0, ra1 = *(const v4i*)ga1, rx0 = *(const v4i*)gx0, rx1 = *(const v4i*)gx1;
    const int wm = w >> 1, wn = w & 1;
    const int fo = (lane & 31) * kGStride + 16 * (lane >> 5);  // this lane's fragment offset in a 32-row slice
    v16i acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = v16i{};
    int buf = 0;
    for (int64_t k0 = 0; k0 < K; k0 += kGK) {
        *(v4i*)(As[buf] + so0) = ra0;
        *(v4i*)(As[buf] + so1) = ra1;
        *(v4i*)(Xs[buf] + so0) = rx0;
        *(v4i*)(Xs[buf] + so1) = rx1;
        __syncthreads();
        if (k0 + kGK < K) {
            ra0 = *(const v4i*)(ga0 + k0 + kGK);
            ra1 = *(const v4i*)(ga1 + k0 + kGK);
            rx0 = *(const v4i*)(gx0 + k0 + kGK);
            rx1 = *(const v4i*)(gx1 + k0 + kGK);
        }
        const int8_t* as = As[buf] + (wm * 64) * kGStride + fo;
        const int8_t* xs = Xs[buf] + (wn * 64) * kGStride + fo;
#pragma unroll
        for (int kk = 0; kk < kGK; kk += 32) {
            const v4i a0 = *(const v4i*)(as + kk), a1 = *(const v4i*)(as + 32 * kGStride + kk);
            const v4i x0 = *(const v4i*)(xs + kk), x1 = *(const v4i*)(xs + 32 * kGStride + kk);
            acc[0][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, x0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, x1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, x0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, x1, acc[1][1], 0, 0, 0);
        }
        buf ^= 1;
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int64_t i = bi + wm * 64 + a * 32 + (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5);
                const int64_t j = bj + wn * 64 + b * 32 + (lane & 31);
                C[i * ldc + j] = acc[a][b][r];
            }
}

// the right operand of the second products: row j < N is [s0 | s1] of row j
// of S, row N + j is [t0 | t1] of row j of T = S o A (base-127 digits)
__global__ void k_dense_digits(const int32_t* __restrict__ S, const int8_t* __restrict__ A2, int64_t N,
                               int8_t* __restrict__ X) {
    GRID_STRIDE(i, N * N) {
        const int64_t r = i / N, c = i - r * N;
        const int32_t sv = S[i];
        const int32_t tv = A2[r * 2 * N + c] ? sv : 0;
        X[r * 2 * N + c] = (int8_t)(sv % 127);
        X[r * 2 * N + N + c] = (int8_t)(sv / 127);
        X[(N + r) * 2 * N + c] = (int8_t)(tv % 127);
        X[(N + r) * 2 * N + N + c] = (int8_t)(tv / 127);
    }
}

// per-edge counts.  K4(e) for e = (v,u) is the number of edges among
// C = N(v) & N(u): half the sum over the members w of C of
// popcount(N(w) & C).  A block keeps a slab of SW words of every bit row in
// LDS (column words [q SW, q SW + SW) of rows 0..N-1) and works through
// groups of kK4Waves x 64 edges, each group once per slab.  A wavefront takes
// its 64 edges one after another (their ids held one per lane, the next
// edge's bit rows loaded while the current one is counted), lists the
// members of C in LDS in ascending order (16-bit ids, N <= 8192), and the
// lanes take the members below the slab's end, four at a time: each adds the
// popcounts of the member's slab row AND the slab of C, twice for a member
// below the slab (whose pairs with the slab's members are met once) and once
// for a member inside it (whose pairs inside the slab are met from both
// ends); members above the slab add nothing, their pairs being met from the
// lower end.  Lane i keeps the sum of the wavefront's i-th edge across the
// slabs and, after the last one, finishes that edge's T, K4 and C4 from the
// products.
constexpr int kMaxWordsPerLane = kDenseMaxN / 32 / 64;
constexpr int kK4Waves = 8;
constexpr int kK4Threads = 64 * kK4Waves;
constexpr size_t kLdsBytes = 160 * 1024;

template <int SW>
__global__ __launch_bounds__(kK4Threads) void k_dense_k4(
    const int32_t* __restrict__ ev, const int32_t* __restrict__ eu, const int64_t* __restrict__ row, int64_t m,
    int64_t N, int Wd, const int32_t* __restrict__ S, const int32_t* __restrict__ PQ, const uint32_t* __restrict__ bits,
    uint32_t* __restrict__ T, uint64_t* __restrict__ K4, uint64_t* __restrict__ C4) {
    extern __shared__ __align__(16) uint32_t lds[];
    uint32_t* slab = lds;  // N x SW words
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint16_t* list = (uint16_t*)(lds + N * SW) + (size_t)wv * N;
    const int nslab = (Wd + SW - 1) / SW;
    const int nwl = (Wd + 63) / 64;  // words of C per lane (<= kMaxWordsPerLane)
    const int64_t ngroups = (m + kK4Threads - 1) / kK4Threads;
    for (int64_t gi = blockIdx.x; gi < ngroups; gi += gridDim.x) {
        const int64_t eb = gi * kK4Threads + (int64_t)wv * 64;  // this wavefront's 64 edges
        const int ne = (int)max<int64_t>(0, min<int64_t>(64, m - eb));
        const int32_t myv = lane < ne ? ev[eb + lane] : 0, myu = lane < ne ? eu[eb + lane] : 0;
        uint32_t mine = 0;  // twice K4 of edge eb + lane, summed over the slabs
        for (int q = 0; q < nslab; q++) {
            __syncthreads();  // no wavefront still reads the previous slab
            for (int64_t i = threadIdx.x; i < N * SW; i += kK4Threads) {
                const int64_t r = i / SW;
                const int wj = q * SW + (int)(i % SW);
                slab[i] = wj < Wd ? bits[r * Wd + wj] : 0u;
            }
            __syncthreads();
            const int q0 = q * SW, qe = min(q0 + SW, Wd);  // the slab's words of C
            // the bit rows of N(v) and N(u), one edge ahead
            uint32_t rv[kMaxWordsPerLane], ru[kMaxWordsPerLane];
            auto load_rows = [&](int i) {
                const int64_t v = __shfl(myv, i, 64), u = __shfl(myu, i, 64);
#pragma unroll
                for (int k = 0; k < kMaxWordsPerLane; k++) {
                    const int x = lane + 64 * k;
                    const bool ok = k < nwl && x < Wd;
                    rv[k] = ok ? bits[v * Wd + x] : 0u;
                    ru[k] = ok ? bits[u * Wd + x] : 0u;
                }
            };
            if (ne > 0) load_rows(0);
            for (int i = 0; i < ne; i++) {
                uint32_t cw[kMaxWordsPerLane];
#pragma unroll
                for (int k = 0; k < kMaxWordsPerLane; k++) cw[k] = rv[k] & ru[k];
                if (i + 1 < ne) load_rows(i + 1);
                // the slab's words of C: each lane stores its words in this
                // wavefront's list area, and every lane reads the slab's words
                // back (the same address for all lanes); the member list then
                // overwrites the area
                uint32_t* cwl = (uint32_t*)list;
#pragma unroll
                for (int k = 0; k < kMaxWordsPerLane; k++)
                    if (k < nwl && lane + 64 * k < Wd) cwl[lane + 64 * k] = cw[k];
                __builtin_amdgcn_wave_barrier();
                uint32_t cs[SW];
                if constexpr (SW >= 4) {
#pragma unroll
                    for (int j = 0; j < SW; j += 4) {
                        const uint4 t = *(const uint4*)(cwl + q0 + j);
                        cs[j] = t.x;
                        cs[j + 1] = t.y;
                        cs[j + 2] = t.z;
                        cs[j + 3] = t.w;
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < SW; j++) cs[j] = cwl[q0 + j];
                }
#pragma unroll
                for (int j = 0; j < SW; j++)
                    if (q0 + j >= Wd) cs[j] = 0u;
                __builtin_amdgcn_wave_barrier();
                // members in ascending order: word 64k + lane's members after
                // those of every lower word; nlo / nhi = members below the
                // slab's first / past its last word (members past the slab
                // are never read, so only those below it are written)
                uint32_t base = 0, nlo = 0, nhi = 0;
#pragma unroll
                for (int k = 0; k < kMaxWordsPerLane; k++) {
                    if (k >= nwl) break;
                    const uint32_t cnt = (uint32_t)__popc(cw[k]);
                    const uint32_t incl = wave_incl_sum(cnt);
                    uint32_t pos = base + incl - cnt;
                    uint32_t x = lane + 64 * k < qe ? cw[k] : 0u;
                    const uint32_t wbase = (uint32_t)(lane + 64 * k) * 32u;
                    while (x) {
                        list[pos++] = (uint16_t)(wbase + (uint32_t)(__ffs(x) - 1));
                        x &= x - 1;
                    }
                    const int rl = q0 - 64 * k, rh = qe - 64 * k;
                    const uint32_t at_lo = (uint32_t)__shfl((int)incl, max(rl - 1, 0), 64);
                    const uint32_t at_hi = (uint32_t)__shfl((int)incl, max(rh - 1, 0), 64);
                    if (rl > 0) nlo = base + (rl >= 64 ? (uint32_t)__shfl((int)incl, 63, 64) : at_lo);
                    if (rh > 0) nhi = base + (rh >= 64 ? (uint32_t)__shfl((int)incl, 63, 64) : at_hi);
                    base += (uint32_t)__shfl((int)incl, 63, 64);
                }
                __builtin_amdgcn_wave_barrier();
                // twice the pairs {w, x} of C with x in the slab: a member below
                // the slab meets each such x above it (weight 2); the members
                // inside meet each pair inside twice (weight 1).  A lane takes
                // four consecutive members per step (one 8-byte list read, the
                // four rows' loads issued together); ids past nhi weigh 0
                uint32_t acc = 0;
                for (uint32_t r0 = 0; r0 < nhi; r0 += 256) {
                    const uint32_t r = r0 + 4u * (uint32_t)lane;
                    uint2 ids = make_uint2(0u, 0u);
                    if (r < nhi) ids = *(const uint2*)(list + r);
                    uint32_t w[4], wt[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        const uint32_t id = ((t < 2 ? ids.x : ids.y) >> (16 * (t & 1))) & 0xffffu;
                        const bool ok = r + t < nhi;
                        w[t] = ok ? id : 0u;
                        wt[t] = ok ? (r + t < nlo ? 2u : 1u) : 0u;
                    }
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        const uint32_t* rw = slab + (size_t)w[t] * SW;
                        uint32_t pc = 0;
#pragma unroll
                        for (int j = 0; j < SW; j++) pc += __popc(rw[j] & cs[j]);
                        acc += wt[t] * pc;
                    }
                }
                __builtin_amdgcn_wave_barrier();  // the list is read before the next edge rewrites it
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) acc += (uint32_t)__shfl_xor((int)acc, off, 64);
                if (lane == i) mine += acc;
            }
        }
        if (lane < ne) {
            const int64_t e = eb + lane;
            const int64_t v = myv, u = myu;
            const uint64_t t = (uint64_t)S[v * N + u];
            // [A3 | AT]: row stride 2N, A3 in columns [0, N), AT in [N, 2N)
            const uint64_t a3 = (uint64_t)PQ[v * 2 * N + u];
            const uint64_t dv = (uint64_t)(row[v + 1] - row[v]), du = (uint64_t)(row[u + 1] - row[u]);
            const uint64_t p = a3 - dv - du + 1;
            const uint64_t qq = (uint64_t)PQ[v * 2 * N + N + u] + (uint64_t)PQ[u * 2 * N + N + v];
            const uint64_t k4 = (uint64_t)(mine >> 1);
            T[e] = (uint32_t)t;
            K4[e] = k4;
            C4[e] = p - qq + 2 * t + 2 * k4;
        }
    }
}

// slab width (words) of k_dense_k4 for padded size N: the largest power of
// two up to 16 whose slab and member lists fit the LDS
int k4_slab_words(int64_t N) {
    const int64_t lists = (int64_t)kK4Waves * N * 2;
    int sw = 16;
    while (sw > 1 && N * sw * 4 + lists > (int64_t)kLdsBytes) sw >>= 1;
    return sw;
}
size_t k4_lds(int64_t N, int sw) { return (size_t)N * sw * 4 + (size_t)kK4Waves * N * 2; }

// The engine's arrays for one graph: [A | 127 A], its bit rows, S, the
// digits operand, and the products [A3 | AT]
struct DenseBufs {
    int8_t *A2, *X;
    int32_t *S, *PQ;
    uint32_t* bits;
};

// bytes of DenseBufs for padded size N, each array 256-byte aligned
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t dense_bytes(int64_t N) {
    return align256(2 * (size_t)N * N) + align256(4 * (size_t)N * N) + align256(4 * (size_t)N * N) +
           align256(8 * (size_t)N * N) + align256((size_t)N * N / 8);
}
DenseBufs carve(unsigned char* base, int64_t N) {
    const size_t nn = (size_t)N * N;
    DenseBufs d;
    unsigned char* p = base;
    d.A2 = (int8_t*)p; p += align256(2 * nn);
    d.X = (int8_t*)p; p += align256(4 * nn);
    d.S = (int32_t*)p; p += align256(4 * nn);
    d.PQ = (int32_t*)p; p += align256(8 * nn);
    d.bits = (uint32_t*)p;
    return d;
}

// k_dense_k4 with slab width SW: dynamic LDS above 64 KiB is allowed for the
// kernel (an attribute of the current device), and the grid fills every CU
// with as many blocks as the occupancy at this LDS size allows
template <int SW>
void launch_k4(pgd_ctx* ctx, const GraphView& g, int64_t N, int Wd, const DenseBufs& d, EdgeCounts out) {
    const size_t lds = k4_lds(N, SW);
    PGD_HIP(hipFuncSetAttribute((const void*)k_dense_k4<SW>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)kLdsBytes));
    int per_cu = 0;
    PGD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dense_k4<SW>, kK4Threads, lds));
    PGD_CHECK(per_cu > 0, PGD_ERR_UNSUPPORTED, "dense 4-clique pass: slab and lists exceed the LDS of a CU");
    const int64_t ngroups = (g.m + kK4Threads - 1) / kK4Threads;
    hipLaunchKernelGGL(k_dense_k4<SW>, dim3(grid_cap(ctx, ngroups, per_cu)), dim3(kK4Threads), lds, ctx->stream,
                       g.ev, g.eu, g.row, g.m, N, Wd, d.S, d.PQ, d.bits, out.T, out.K4, out.C4);
}

// the launch sequence; every array of d is written before it is read (A and
// the bit rows are cleared first), so no state carries over between calls
void dense_launch(pgd_ctx* ctx, const GraphView& g, const DenseBufs& d, EdgeCounts out, Timer* timer) {
    hipStream_t s = ctx->stream;
    const int64_t n = g.n, m = g.m;
    const int64_t N = (n + kTile - 1) / kTile * kTile, NN = N * N;
    const int Wd = (int)(N / 32);
    if (timer) timer->begin("k_dense_adj");
    hipLaunchKernelGGL(d_fill_u32, dim3(grid_cap(ctx, (2 * NN / 4 + 255) / 256, 16)), dim3(256), 0, s,
                       (uint32_t*)d.A2, 2 * NN / 4);
    hipLaunchKernelGGL(d_fill_u32, dim3(grid_cap(ctx, (N * Wd + 255) / 256, 16)), dim3(256), 0, s, d.bits, N * Wd);
    hipLaunchKernelGGL(k_dense_adj, dim3(grid_cap(ctx, (n * 64 + 255) / 256, 16)), dim3(256), 0, s, g.row, g.col, n,
                       N, Wd, d.A2, d.bits);
    PGD_HIP(hipGetLastError());
    if (timer) timer->end();
    const unsigned nt = (unsigned)(N / kTile);
    // S = A . A (the first N columns of [A | 127 A]); then [A3 | AT] over K = 2N
    if (timer) timer->begin("k_gemm_i8");
    hipLaunchKernelGGL(k_gemm_i8, dim3(nt, nt), dim3(256), 0, s, d.A2, 2 * N, d.A2, 2 * N, d.S, N, N);
    PGD_HIP(hipGetLastError());
    if (timer) timer->end();
    if (timer) timer->begin("k_dense_digits");
    hipLaunchKernelGGL(k_dense_digits, dim3(grid_cap(ctx, (NN + 255) / 256, 16)), dim3(256), 0, s, d.S, d.A2, N,
                       d.X);
    PGD_HIP(hipGetLastError());
    if (timer) timer->end();
    if (timer) timer->begin("k_gemm_i8_pq");
    hipLaunchKernelGGL(k_gemm_i8, dim3(2 * nt, nt), dim3(256), 0, s, d.A2, 2 * N, d.X, 2 * N, d.PQ, 2 * N, 2 * N);
    PGD_HIP(hipGetLastError());
    if (timer) timer->end();
    if (timer) timer->begin("k_dense_k4");
    if (m) {
        switch (k4_slab_words(N)) {
            case 16: launch_k4<16>(ctx, g, N, Wd, d, out); break;
            case 8: launch_k4<8>(ctx, g, N, Wd, d, out); break;
            case 4: launch_k4<4>(ctx, g, N, Wd, d, out); break;
            case 2: launch_k4<2>(ctx, g, N, Wd, d, out); break;
            default: launch_k4<1>(ctx, g, N, Wd, d, out); break;
        }
    }
    PGD_HIP(hipGetLastError());
    if (timer) timer->end();
}

}  // namespace

bool engine_dense_fits(const GraphView& g) { return g.n <= kDenseMaxN; }

bool engine_dense_ok(const GraphView& g) {
    // the reference's own dense criterion (create_adj_mat,
    // graphlet_core.cpp:602-633): more than 80% of all pairs are edges; the
    // N x N products stay modest.  Below that density the oriented engine is
    // faster (G(2000, 0.5): 10 ms against 17 ms for this tier).
    return g.n >= 64 && g.n <= kDenseMaxN && 10 * g.m > 4 * g.n * (g.n - 1);
}

void engine_dense(pgd_ctx* ctx, const GraphView& g, EdgeCounts out, Timing* tm) {
    PGD_CHECK(g.n <= kDenseMaxN, PGD_ERR_UNSUPPORTED, "graph too large for the dense engine");
    const int64_t n = g.n, m = g.m;
    const int64_t N = (n + kTile - 1) / kTile * kTile, NN = N * N;
    DBuf<unsigned char> buf;
    buf.alloc(ctx, dense_bytes(N));
    Timer timer{ctx};
    dense_launch(ctx, g, carve(buf.get(), N), out, &timer);
    PGD_HIP(hipStreamSynchronize(ctx->stream));
    // work models: the MFMA products are 2 N^3 (S) and 8 N^3 ([A3 | AT] over
    // K = 2N) int8 operations; the adjacency pass writes 2 N^2 + N^2/8 bytes;
    // the edge pass reads four int32 products and writes T, K4 and C4 per edge
    const double n3 = (double)N * N * N;
    tm->phases.clear();
    double best = 0;
    for (size_t i = 0; i < timer.ev.size(); i++) {
        float ms = 0;
        PGD_HIP(hipEventElapsedTime(&ms, timer.ev[i].first, timer.ev[i].second));
        double bytes = 0, units = 0;
        // for the MFMA passes, bytes holds operations
        if (i == 0) { bytes = 2.0 * NN + NN / 8.0; units = (double)NN; }
        if (i == 1) { bytes = 2.0 * n3; units = 1; }
        if (i == 2) { bytes = 9.0 * NN; units = (double)NN; }  // S and A read, four digit bytes written
        if (i == 3) { bytes = 8.0 * n3; units = 1; }
        if (i == 4) { bytes = 36.0 * m; units = (double)m; }
        tm->phases.push_back(Phase{timer.names[i], ms, bytes, units});
        if (ms > best) best = ms;
    }
    tm->dominant_ms = best;
    tm->dominant_launches = 1;
}

}  // namespace pgd

// ---------------------------------------------------------------------------
// Allocation-free decomposition of small graphs (pgd_decompose_noalloc).
// Every intermediate array of the dense engine and the per-edge counts live
// in one static device array per GPU, sized for n <= PGD_NOALLOC_MAX_N.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t kNoAllocMaxN = PGD_NOALLOC_MAX_N;
constexpr int64_t kNoAllocMaxM = kNoAllocMaxN * (kNoAllocMaxN - 1) / 2;
constexpr size_t kNoAllocCounts =
    pgd::align256(4 * (size_t)kNoAllocMaxM) + 2 * pgd::align256(8 * (size_t)kNoAllocMaxM);
constexpr size_t kNoAllocBytes = pgd::dense_bytes(kNoAllocMaxN) + kNoAllocCounts;
static_assert(kNoAllocBytes == PGD_NOALLOC_SCRATCH_BYTES, "pgd_hip.h states the scratch size");
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU
}  // namespace

namespace pgd {

// once per device, with that device current: the address of this GPU's copy
// of `scratch` (host code never passes the symbol itself to a kernel)
void noalloc_scratch_init(int device) {
    PGD_CHECK(device >= 0 && device < kMaxDevices, PGD_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return;
    int cur = 0;
    PGD_HIP(hipGetDevice(&cur));
    PGD_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    PGD_HIP(e);
    PGD_CHECK(ptr != nullptr, PGD_ERR_HIP, "no device address for the scratch array");
    g_scratch[device] = (unsigned char*)ptr;
}

void decompose_noalloc(pgd_ctx* ctx, const GraphView& g, uint64_t* d_sums, uint64_t* d_micro) {
    PGD_CHECK(g.n <= kNoAllocMaxN, PGD_ERR_UNSUPPORTED,
              "pgd_decompose_noalloc: more than PGD_NOALLOC_MAX_N vertices");
    PGD_CHECK(g.m <= kNoAllocMaxM, PGD_ERR_UNSUPPORTED, "pgd_decompose_noalloc: too many edges");
    unsigned char* base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch[ctx->device];
    }
    PGD_CHECK(base != nullptr, PGD_ERR_STATE, "scratch not initialised on this device");
    const int64_t N = (g.n + kTile - 1) / kTile * kTile;
    const DenseBufs d = carve(base, N);
    unsigned char* cp = base + dense_bytes(kNoAllocMaxN);
    EdgeCounts c{(uint32_t*)cp, (uint64_t*)(cp + align256(4 * (size_t)kNoAllocMaxM)),
                 (uint64_t*)(cp + align256(4 * (size_t)kNoAllocMaxM) + align256(8 * (size_t)kNoAllocMaxM))};
    if (g.m > 0) dense_launch(ctx, g, d, c, nullptr);
    epilogue(ctx, g, c, d_sums, d_micro);
    PGD_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace pgd
