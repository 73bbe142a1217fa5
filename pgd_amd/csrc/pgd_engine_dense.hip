// pgd_engine_dense.hip -- dense-graph tier on the int8 matrix cores.
//
// For a dense graph the per-edge counts are contractions of the adjacency
// matrix A (0/1, held as int8, zero-padded to N = 128*ceil(n/128)):
//
//   S    = A.A                  codegrees; T(e) = S_vu for the edge e = (v,u)
//   A3   = A.S                  P(e) = A3_vu - d_v - d_u + 1 (4-cycles through e)
//   AT   = A.T, T = S o A       Q(e) = AT_vu + AT_uv  (sum over triangles {v,u,w}
//                               of T(v,w) + T(u,w))
//   K4(e) = edges among the common neighbours of v and u: for e = (v, u),
//           v < u, half of rowsum((B.B) o B) at u's row of B = A[N(v), N(v)],
//           v's local graph (int8 products of compressed bit rows, below)
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

constexpr int kDenseMaxN = PGD_DENSE_MAX_N;  // int32 products of N^2 entries, bit rows of N/32 words
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
// M, cols and K multiples of 128.  One block per 128 x 128 output tile, four
// wavefronts of 64 x 64 (2 x 2 MFMA tiles of 32 x 32).  Each 128-deep step
// of both operands is staged in LDS, double-buffered: the next step's global
// loads are in flight during the current step's products, and one barrier
// per step separates the buffers.  LDS rows are 144 bytes: the 16-byte
// fragment reads of 16 consecutive rows then fall on distinct banks, and the
// staging stores, eight lanes per row (a row's 128 bytes), too.  MFMA
// operands: lane l feeds row (l mod 32) and bytes 16 (l / 32) .. + 16 of
// the 32-deep k slice, for both operands; accumulator r of lane l holds row
// (r mod 4) + 8 (r / 4) + 4 (l / 32), column l mod 32.
constexpr int kGK = 128;       // k per staged step
constexpr int kGStride = 144;  // LDS bytes per staged row
__global__ __launch_bounds__(256) void k_gemm_i8(const int8_t* __restrict__ A, int64_t lda,
                                                 const int8_t* __restrict__ X, int64_t ldx,
                                                 int32_t* __restrict__ C, int64_t ldc, int64_t K) {
    __shared__ __align__(16) int8_t As[2][kTile * kGStride];
    __shared__ __align__(16) int8_t Xs[2][kTile * kGStride];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t bi = (int64_t)blockIdx.y * kTile, bj = (int64_t)blockIdx.x * kTile;
    // staging: thread t moves 16-byte chunk t mod 8 of rows t / 8 + 32 j
    // (j = 0..3) of each 128 x 128 tile
    const int r0 = tid >> 3, c0 = (tid & 7) * 16;
    const int8_t* ga = A + (bi + r0) * lda + c0;
    const int8_t* gx = X + (bj + r0) * ldx + c0;
    v4i ra[4], rx[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        ra[j] = *(const v4i*)(ga + 32 * j * lda);
        rx[j] = *(const v4i*)(gx + 32 * j * ldx);
    }
    const int wm = w >> 1, wn = w & 1;
    const int fo = (lane & 31) * kGStride + 16 * (lane >> 5);  // this lane's fragment offset in a 32-row slice
    v16i acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = v16i{};
    int buf = 0;
    for (int64_t k0 = 0; k0 < K; k0 += kGK) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            *(v4i*)(As[buf] + (r0 + 32 * j) * kGStride + c0) = ra[j];
            *(v4i*)(Xs[buf] + (r0 + 32 * j) * kGStride + c0) = rx[j];
        }
        __syncthreads();
        if (k0 + kGK < K) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                ra[j] = *(const v4i*)(ga + 32 * j * lda + k0 + kGK);
                rx[j] = *(const v4i*)(gx + 32 * j * ldx + k0 + kGK);
            }
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

// ---- 4-cliques per edge on the matrix cores.  For a vertex v with the
// ascending neighbour list L (d = |L|), B = A[L, L] is its local graph, and
// for u = L[i]
//   2 K4(v,u) = sum_j B[i][j] (B.B)[i][j]
// (ordered pairs of adjacent members of N(v) & N(u)).  Each edge is taken at
// its lower end: only the rows i >= i0 (L[i] > v) are formed.  The products
// run per vertex on compressed bit rows (column x of A becomes column j of B
// when x = L[j]), which cuts the operations by (N/d)^2 against masking A.
//
// k_dense_compress writes, for each vertex of a chunk, B's rows as bits,
// words past d zero (row stride `stride` words, stride = dpad / 32 with dpad
// the largest degree rounded up to 128), L, (v, d, i0), and appends the
// vertex's row tiles (slot << 6 | tile, 128 rows each) to the tile list.
// Bits are moved with the compress of Hacker's Delight (7-4): five masked
// shift steps whose masks depend only on the mask word, so they are computed
// once per vertex and word and shared by every row.
constexpr int kK4MaxW = kDenseMaxN / 32;  // words of a bit row at most
constexpr int kK4Rows = 128;              // rows (and columns) of one product tile

struct K4Info {
    int32_t v, d, i0, pad;
};
constexpr int kCB = 8;  // rows per lane in flight in k_dense_compress
// lanes of k_dense_compress per row: the mask words Wd rounded up to a power
// of two, at most 64 (a wavefront then takes 64 / that rows at once)
__host__ __device__ inline int k4_lanes_per_row(int Wd) {
    int l = 1;
    while (l < Wd && l < 64) l <<= 1;
    return l;
}
// row-buffer words per wavefront of k_dense_compress: kCB batches of rows of
// `stride` words (stride <= Wd <= lanes per row when Wd <= 64)
inline int k4_rowbuf_words(int Wd, int64_t stride) { return kCB * (64 / k4_lanes_per_row(Wd)) * (int)stride; }

__global__ __launch_bounds__(256) void k_dense_compress(const uint32_t* __restrict__ bits, int Wd, int64_t v0,
                                                        int64_t v1, int64_t stride, int64_t vwords, int64_t dpad,
                                                        uint32_t* __restrict__ G, uint16_t* __restrict__ Lg,
                                                        K4Info* __restrict__ info, uint32_t* __restrict__ tiles,
                                                        uint32_t* __restrict__ count,
                                                        unsigned long long* __restrict__ ops, int rbw, int64_t N,
                                                        int32_t* __restrict__ K4M) {
    __shared__ uint32_t msk[kK4MaxW], pre[kK4MaxW], mv[5][kK4MaxW];
    extern __shared__ uint32_t rowbuf[];  // per wavefront: kCB rpp rows of `stride` words (k4_rowbuf_words)
    __shared__ uint32_t wsum[4];
    __shared__ uint16_t L[kDenseMaxN];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int64_t v = v0 + blockIdx.x; v < v1; v += gridDim.x) {
        const int64_t slot = v - v0;
        __syncthreads();  // the previous vertex is done with the LDS
        uint32_t m = 0, c = 0;
        if (tid < Wd) {
            m = bits[v * Wd + tid];
            c = (uint32_t)__popc(m);
            msk[tid] = m;
            uint32_t mm = m, mk = ~m << 1;
#pragma unroll
            for (int s = 0; s < 5; s++) {
                uint32_t mp = mk ^ (mk << 1);
                mp ^= mp << 2;
                mp ^= mp << 4;
                mp ^= mp << 8;
                mp ^= mp << 16;
                const uint32_t t = mp & mm;
                mv[s][tid] = t;
                mm = (mm ^ t) | (t >> (1 << s));
                mk &= ~mp;
            }
        }
        const uint32_t inc = wave_incl_sum(c);
        if (lane == 63) wsum[wv] = inc;
        for (int i = tid; i < 4 * rbw; i += 256) rowbuf[i] = 0u;
        __syncthreads();
        uint32_t before = 0;
        for (int k = 0; k < wv; k++) before += wsum[k];
        const uint32_t d = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (tid < Wd) {
            uint32_t pos = before + inc - c;
            pre[tid] = pos;
            uint32_t x = m;
            while (x) {
                L[pos++] = (uint16_t)(32 * tid + __ffs(x) - 1);
                x &= x - 1;
            }
        }
        __syncthreads();
        const int vw = (int)(v >> 5);
        const uint32_t i0 = pre[vw] + (uint32_t)__popc(msk[vw] & ((1u << (v & 31)) - 1u));
        for (uint32_t i = tid; i < d; i += 256) {
            Lg[slot * dpad + i] = L[i];
            if (i >= i0) K4M[v * N + L[i]] = 0;  // the products add into these
        }
        if (tid == 0) {
            info[slot] = K4Info{(int32_t)v, (int32_t)d, (int32_t)i0, 0};
            const uint32_t nt = (d - i0 + kK4Rows - 1) / kK4Rows;
            if (nt) {
                const uint32_t base = atomicAdd(count, nt);
                for (uint32_t t = 0; t < nt; t++) tiles[base + t] = (uint32_t)slot << 6 | t;
                // row tile t multiplies the column tiles below i0 and those from t on
                const unsigned long long dp = (d + kK4Rows - 1) / kK4Rows * kK4Rows;
                const unsigned long long nlow = (i0 + kK4Rows - 1) / kK4Rows;
                const unsigned long long ncols = nt * nlow + (unsigned long long)nt * (nt + 1) / 2;
                atomicAdd(ops, 2ull * kK4Rows * kK4Rows * dp * ncols);
            }
        }
        // B's rows, in batches of kCB rpp consecutive rows per wavefront: a row
        // gets lpr lanes (Wd rounded up to a power of two, at most 64), lane
        // wl of a row compresses words wl, wl + lpr, ... of row L[i] under the
        // mask and ors them in at bit pre[w]; the kCB loads of a lane are
        // issued together.  The batch's rows are then stored contiguously.
        const int lpr = k4_lanes_per_row(Wd), rpp = 64 / lpr, sub = lane / lpr, wl = lane % lpr;
        uint32_t* rb = rowbuf + wv * rbw;
        uint32_t* gv = G + slot * vwords;
        const int nb = kCB * rpp;
        for (uint32_t ib = (uint32_t)(wv * nb); ib < d; ib += (uint32_t)(4 * nb)) {
            for (int w = wl; w < Wd; w += lpr) {
                const uint32_t mw = msk[w];
                uint32_t x[kCB];
#pragma unroll
                for (int b2 = 0; b2 < kCB; b2++) {
                    const uint32_t i = ib + (uint32_t)(b2 * rpp + sub);
                    x[b2] = (mw && i < d) ? bits[(int64_t)L[i] * Wd + w] : 0u;
                }
                if (!mw) continue;
                const uint32_t off = pre[w], sh = off & 31;
                const bool spill = sh && sh + (uint32_t)__popc(mw) > 32;
#pragma unroll
                for (int b2 = 0; b2 < kCB; b2++) {
                    uint32_t y = x[b2] & mw;
#pragma unroll
                    for (int s2 = 0; s2 < 5; s2++) {
                        const uint32_t t = y & mv[s2][w];
                        y = (y ^ t) | (t >> (1 << s2));
                    }
                    uint32_t* row = rb + (b2 * rpp + sub) * stride;
                    if (y) {
                        atomicOr(&row[off >> 5], y << sh);
                        if (spill) atomicOr(&row[(off >> 5) + 1], y >> (32 - sh));
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            const int nw = (int)min<int64_t>((int64_t)nb * stride, ((int64_t)d - ib) * stride);
            for (int f = lane; f < nb * stride; f += 64) {
                if (f < nw) gv[(int64_t)ib * stride + f] = rb[f];
                rb[f] = 0u;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// One 64-deep slice of an fp4 MFMA operand (v_mfma_scale_f32_32x32x64_f8f6f4
// with e2m1 operands: 32 values per lane, the lane's half of the slice, in 4
// registers) from a 32-bit word of a bit row: a set bit becomes the nibble 1
// (0.5 in e2m1), with register q holding bits q, q + 4, ..., q + 28 (two
// operations per eight values).  With a block scale of 2 for each operand
// (E8M0 128) every product of two set bits is 1 and the f32 sums are the
// exact counts (at most 8192).  The products sum over the depth, so this
// order, used alike for both operands, gives the same sums.
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
constexpr int kFp4 = 4;         // operand format code of e2m1
constexpr int kScaleTwo = 128;  // E8M0 2^(128 - 127)
__device__ __forceinline__ v8i slice_fp4(uint32_t x) {
    v8i f = v8i{};
    f[0] = (int)(x & 0x11111111u);
    f[1] = (int)((x >> 1) & 0x11111111u);
    f[2] = (int)((x >> 2) & 0x11111111u);
    f[3] = (int)((x >> 3) & 0x11111111u);
    return f;
}
__device__ __forceinline__ v16f mfma_fp4(v8i a, v8i b, v16f c) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, kFp4, kFp4, 0, kScaleTwo, 0, kScaleTwo);
}

// One block per row tile of one vertex's local graph (tiles from the list
// k_dense_compress built): for 128 rows i of B (from i0 + 128 t), the
// product (B.B) with 128-column tiles of B over the depth dp (d rounded up to
// 128), one 128-deep step at a time: both operands' bit rows are staged in
// LDS as bits (16 bytes per row and step, rows 24 bytes apart), double-
// buffered with one barrier per step, and each wavefront builds its fp4 MFMA
// operand slices from them in registers (one 32-bit LDS read per lane and
// 64-deep slice; lane l feeds row l mod 32 and half l / 32 of the slice, as
// for the int8 products).  Accumulators as in k_gemm_i8 (f32, exact
// integers).  The epilogue keeps the products where B[i][j] = 1, reading
// those bits from row j (B is
// symmetric: a lane's column j holds the bits of its rows in one 32-bit
// window of row j); rows and columns past d have zero products, so their
// bits are never needed.
// Column tiles are cut at i0 + 128 c like the row tiles.  (B.B) o B is
// symmetric, so among the rows and columns from i0 on, row tile t takes the
// column tiles c >= t only: the kept products add to the sums of their rows,
// and, for c > t, to the sums of their columns (which are rows of tile c);
// the column tiles below i0 add to the rows only.  Twice K4(v, L[i]) is the
// total, added into K4M[v N + L[i]] (cleared by k_dense_compress).
constexpr int kK4Step = 128;  // depth per staged step (4 words per row)
// LDS words per staged row: a slice read takes word 2s + hi of 32
// consecutive rows, and 6 r mod 64 is distinct and even for 32 rows, so the
// 64 addresses fall on 64 banks
__global__ __launch_bounds__(256) void k_dense_k4m(const uint32_t* __restrict__ G, const uint16_t* __restrict__ Lg,
                                                   const K4Info* __restrict__ info, const uint32_t* __restrict__ tiles,
                                                   const uint32_t* __restrict__ count, int64_t stride,
                                                   int64_t vwords, int64_t dpad, int64_t N,
                                                   int32_t* __restrict__ K4M) {
    constexpr int kRowW = 6;
    __shared__ __align__(16) uint32_t As[2][kK4Rows * kRowW];
    __shared__ __align__(16) uint32_t Xs[2][kK4Rows * kRowW];
    __shared__ int32_t rs[kK4Rows];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wm = w >> 1, wn = w & 1;
    const int r0 = tid & 127, h = tid >> 7;  // staging: words 2h, 2h + 1 of row r0 in each step
    const int hi = lane >> 5, lr = lane & 31;
    const uint32_t ntiles = *count;
    int buf = 0;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t tl = tiles[t];
        const int64_t slot = tl >> 6;
        const K4Info inf = info[slot];
        const int d = inf.d, rbase = inf.i0 + (int)(tl & 63) * kK4Rows;
        const int dp = (d + kK4Rows - 1) / kK4Rows * kK4Rows;
        const uint32_t* gv = G + slot * vwords;
        const int ia = rbase + r0;
        const uint2* pa = (const uint2*)(gv + (int64_t)min(ia, d - 1) * stride) + h;
        if (tid < kK4Rows) rs[tid] = 0;
        int32_t part[2][16];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int r = 0; r < 16; r++) part[a][r] = 0;
        const int rt = (int)(tl & 63), nlow = (inf.i0 + kK4Rows - 1) / kK4Rows;
        const int ntile = (d - inf.i0 + kK4Rows - 1) / kK4Rows;
        for (int c = -nlow; c < ntile; c++) {
            if (c >= 0 && c < rt) continue;  // met from row tile c
            const int ct = inf.i0 + c * kK4Rows;  // first column (negative below 0)
            const int jx = ct + r0;
            const bool vx = jx >= 0 && jx < d;
            const uint2* px = (const uint2*)(gv + (int64_t)min(max(jx, 0), d - 1) * stride) + h;
            const uint2 zero2 = make_uint2(0u, 0u);
            uint2 wa = ia < d ? pa[0] : zero2, wx = vx ? px[0] : zero2;
            v16f acc[2][2];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) acc[a][b] = v16f{};
            for (int ks = 0; ks < dp; ks += kK4Step) {
                *(uint2*)(As[buf] + r0 * kRowW + 2 * h) = wa;
                *(uint2*)(Xs[buf] + r0 * kRowW + 2 * h) = wx;
                __syncthreads();
                if (ks + kK4Step < dp) {
                    const int q = (ks + kK4Step) / 64;  // uint2 index of the next step's words
                    wa = ia < d ? pa[q] : zero2;
                    wx = vx ? px[q] : zero2;
                }
                const uint32_t* as = As[buf] + (wm * 64 + lr) * kRowW;
                const uint32_t* xs = Xs[buf] + (wn * 64 + lr) * kRowW;
#pragma unroll
                for (int sl = 0; sl < 2; sl++) {
                    const int wd = 2 * sl + hi;  // this lane's word of the slice
                    const v8i a0 = slice_fp4(as[wd]), a1 = slice_fp4(as[32 * kRowW + wd]);
                    const v8i x0 = slice_fp4(xs[wd]), x1 = slice_fp4(xs[32 * kRowW + wd]);
                    acc[0][0] = mfma_fp4(a0, x0, acc[0][0]);
                    acc[0][1] = mfma_fp4(a0, x1, acc[0][1]);
                    acc[1][0] = mfma_fp4(a1, x0, acc[1][0]);
                    acc[1][1] = mfma_fp4(a1, x1, acc[1][1]);
                }
                buf ^= 1;
            }
            // accumulator r of (a, b) is row wm 64 + a 32 + k with
            // k = (r mod 4) + 8 (r / 4) + 4 (lane / 32), column cl: its bit is
            // bit k of row cl's 32-bit window at rbase + wm 64 + a 32
            // (sums within one column tile stay below 128 x 8192 < 2^24, so
            // they are exact in f32; each is converted once)
            float tile[2][16];
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int r = 0; r < 16; r++) tile[a][r] = 0.f;
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const int cl = ct + wn * 64 + b * 32 + (lane & 31);
                const uint32_t* gr = gv + (int64_t)min(max(cl, 0), dp - 1) * stride;
                float col = 0.f;
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    const int o = rbase + wm * 64 + a * 32;
                    const int q0 = min(o >> 5, (int)stride - 1), q1 = min((o >> 5) + 1, (int)stride - 1);
                    const uint32_t win = __builtin_amdgcn_alignbit(gr[q1], gr[q0], (uint32_t)(o & 31));
#pragma unroll
                    for (int r = 0; r < 16; r++) {
                        const int k = (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5);
                        const float kept = ((win >> k) & 1u) ? acc[a][b][r] : 0.f;
                        tile[a][r] += kept;
                        col += kept;
                    }
                }
                if (c > rt) {  // the column sums of a tile above the diagonal
                    int32_t ci = (int32_t)col;
                    ci += __shfl_xor(ci, 32, 64);
                    if (lane < 32 && cl < d) atomicAdd(&K4M[(int64_t)inf.v * N + Lg[slot * dpad + cl]], ci);
                }
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int r = 0; r < 16; r++) part[a][r] += (int32_t)tile[a][r];
        }
        // each row's sum over the 32 lanes of its half wavefront, then over the
        // two wavefronts that hold its two column halves
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                int32_t x = part[a][r];
#pragma unroll
                for (int off = 16; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
                if ((lane & 31) == 0) atomicAdd(&rs[wm * 64 + a * 32 + (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5)], x);
            }
        __syncthreads();
        if (tid < kK4Rows) {
            const int i = rbase + tid;
            if (i < d) atomicAdd(&K4M[(int64_t)inf.v * N + Lg[slot * dpad + i]], rs[tid]);
        }
        __syncthreads();  // rs is cleared for the next tile
    }
}

// per-edge T, K4 and C4 from the products (S, [A3 | AT], K4M at (min, max))
__global__ void k_dense_finish(const int32_t* __restrict__ ev, const int32_t* __restrict__ eu,
                               const int64_t* __restrict__ row, int64_t m, int64_t N, const int32_t* __restrict__ S,
                               const int32_t* __restrict__ PQ, const int32_t* __restrict__ K4M,
                               uint32_t* __restrict__ T, uint64_t* __restrict__ K4, uint64_t* __restrict__ C4) {
    GRID_STRIDE(e, m) {
        const int64_t a = ev[e], b = eu[e];
        const int64_t v = min(a, b), u = max(a, b);
        const uint64_t t = (uint64_t)S[v * N + u];
        // [A3 | AT]: row stride 2N, A3 in columns [0, N), AT in [N, 2N)
        const uint64_t a3 = (uint64_t)PQ[v * 2 * N + u];
        const uint64_t dv = (uint64_t)(row[v + 1] - row[v]), du = (uint64_t)(row[u + 1] - row[u]);
        const uint64_t p = a3 - dv - du + 1;
        const uint64_t qq = (uint64_t)PQ[v * 2 * N + N + u] + (uint64_t)PQ[u * 2 * N + N + v];
        const uint64_t k4 = (uint64_t)(K4M[v * N + u] >> 1);
        T[e] = (uint32_t)t;
        K4[e] = k4;
        C4[e] = p - qq + 2 * t + 2 * k4;
    }
}

// The per-vertex work area of the 4-clique products for graphs of largest
// degree dmax: slots of compressed rows, lists, (v, d, i0) and row tiles for
// `chunk` vertices at a time, then the tile counter and the operation count
struct K4Plan {
    int64_t dpad, stride, vwords, rtmax, chunk;
    size_t oL, oInfo, oTiles, oCount;
};
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
inline size_t k4_per_vertex(int64_t dpad) {
    return (size_t)dpad * (dpad / 32) * 4 + (size_t)dpad * 2 + sizeof(K4Info) + (size_t)(dpad / kK4Rows) * 4;
}
K4Plan k4_plan(int64_t dmax, size_t wbytes) {
    K4Plan p;
    p.dpad = std::max<int64_t>(1, (dmax + kK4Rows - 1) / kK4Rows) * kK4Rows;
    p.stride = p.dpad / 32;
    p.vwords = p.dpad * p.stride;
    p.rtmax = p.dpad / kK4Rows;
    const size_t slack = 5 * 256 + 256;  // alignment of the four arrays, the counters
    PGD_CHECK(wbytes > slack + k4_per_vertex(p.dpad), PGD_ERR_STATE, "dense 4-clique work area too small");
    p.chunk = (int64_t)((wbytes - slack) / k4_per_vertex(p.dpad));
    p.oL = align256((size_t)p.chunk * p.vwords * 4);
    p.oInfo = p.oL + align256((size_t)p.chunk * p.dpad * 2);
    p.oTiles = p.oInfo + align256((size_t)p.chunk * sizeof(K4Info));
    p.oCount = p.oTiles + align256((size_t)p.chunk * p.rtmax * 4);
    PGD_CHECK(p.oCount + 256 <= wbytes, PGD_ERR_STATE, "dense 4-clique work area layout");
    return p;
}
// the work area the engine allocates: every vertex in one chunk, within
// max(16 N^2, 512 MiB) bytes
size_t k4_work_bytes(int64_t n, int64_t dmax, int64_t N) {
    const int64_t dpad = std::max<int64_t>(1, (dmax + kK4Rows - 1) / kK4Rows) * kK4Rows;
    const size_t per = k4_per_vertex(dpad), slack = 5 * 256 + 256;
    const size_t cap = std::max<size_t>(16 * (size_t)N * N, (size_t)512 << 20);
    return std::max(per + slack + 1, std::min(per * (size_t)std::max<int64_t>(n, 1) + slack, cap));
}

// The engine's arrays for one graph: [A | 127 A], its bit rows, S, the
// digits operand, and the products [A3 | AT]
struct DenseBufs {
    int8_t *A2, *X;
    int32_t *S, *PQ;
    uint32_t* bits;
    unsigned char* W = nullptr;  // work area of the 4-clique products (wbytes)
    size_t wbytes = 0;
};

// bytes of DenseBufs for padded size N, each array 256-byte aligned
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
    if (m) {
        // twice K4 per edge at (min, max) of K4M, which takes X's place (X is
        // read only by the second product)
        int32_t* K4M = (int32_t*)d.X;
        const K4Plan kp = k4_plan(g.max_deg, d.wbytes);
        uint32_t* G = (uint32_t*)d.W;
        uint16_t* Lg = (uint16_t*)(d.W + kp.oL);
        K4Info* info = (K4Info*)(d.W + kp.oInfo);
        uint32_t* tiles = (uint32_t*)(d.W + kp.oTiles);
        uint32_t* count = (uint32_t*)(d.W + kp.oCount);
        unsigned long long* ops = (unsigned long long*)(d.W + kp.oCount + 8);
        hipLaunchKernelGGL(d_fill_u32, dim3(1), dim3(64), 0, s, count, (int64_t)4);
        PGD_HIP(hipGetLastError());
        const int rbw = k4_rowbuf_words(Wd, kp.stride);  // <= kCB kK4MaxW: 32 KiB per block at most
        int per_cu = 0;
        PGD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dense_k4m, 256, 0));
        per_cu = std::max(per_cu, 1);
        for (int64_t v0 = 0; v0 < n; v0 += kp.chunk) {
            const int64_t v1 = std::min<int64_t>(n, v0 + kp.chunk);
            if (v0) hipLaunchKernelGGL(d_fill_u32, dim3(1), dim3(64), 0, s, count, (int64_t)1);
            if (timer) timer->begin("k_dense_compress");
            hipLaunchKernelGGL(k_dense_compress, dim3(grid_cap(ctx, v1 - v0, 8)), dim3(256),
                               (size_t)4 * rbw * sizeof(uint32_t), s, d.bits, Wd, v0, v1, kp.stride, kp.vwords,
                               kp.dpad, G, Lg, info, tiles, count, ops, rbw, N, K4M);
            PGD_HIP(hipGetLastError());
            if (timer) timer->end();
            if (timer) timer->begin("k_dense_k4m");
            hipLaunchKernelGGL(k_dense_k4m, dim3(grid_cap(ctx, (v1 - v0) * kp.rtmax, per_cu)), dim3(256), 0, s, G,
                               Lg, info, tiles, count, kp.stride, kp.vwords, kp.dpad, N, K4M);
            PGD_HIP(hipGetLastError());
            if (timer) timer->end();
        }
        if (timer) timer->begin("k_dense_finish");
        hipLaunchKernelGGL(k_dense_finish, dim3(grid_cap(ctx, (m + 255) / 256, 16)), dim3(256), 0, s, g.ev, g.eu,
                           g.row, m, N, d.S, d.PQ, K4M, out.T, out.K4, out.C4);
        PGD_HIP(hipGetLastError());
        if (timer) timer->end();
    }
}

}  // namespace

bool engine_dense_fits(const GraphView& g) { return g.n <= kDenseMaxN; }

size_t engine_dense_bytes(const GraphView& g) {
    const int64_t N = (g.n + kTile - 1) / kTile * kTile;
    return dense_bytes(N) + (g.m ? k4_work_bytes(g.n, g.max_deg, N) : 0);
}

bool engine_dense_ok(const GraphView& g) {
    // where this tier is the faster engine.  Measured on G(n, p), one
    // decomposition each way: at n <= 4000 it wins at every density from 0.01
    // to 0.5 (the oriented engine's launches and host synchronisations cost
    // about 0.6 ms even for a few hundred edges; G(4000, 0.5): 19 against 167
    // ms); at n = 8000 the oriented engine wins up to density 0.05 (1.6
    // against 5.0 ms at 0.01, 4.6 against 6.4 ms at 0.05) and this tier from
    // 0.1 on (10 against 15 ms).  So: every graph up to 4096 vertices, and up
    // to 8192 vertices those with at least 8% of all pairs as edges (at most
    // 18 N^2 bytes plus a 512 MiB work area).
    if (g.n < 2 || g.n > kDenseMaxN) return false;
    return g.n <= 4096 || 25 * g.m >= g.n * (g.n - 1);
}

struct DenseRun {
    explicit DenseRun(pgd_ctx* c) : timer{c} {}
    Timer timer;
    DBuf<unsigned char> buf, work;
    const unsigned long long* d_ops = nullptr;  // the 4-clique products' operation count, on the device
    int64_t n = 0, m = 0, N = 0;
};
void DenseRunDeleter::operator()(DenseRun* r) const { delete r; }

DenseRunPtr engine_dense(pgd_ctx* ctx, const GraphView& g, EdgeCounts out) {
    PGD_CHECK(g.n <= kDenseMaxN, PGD_ERR_UNSUPPORTED, "graph too large for the dense engine");
    DenseRunPtr r(new DenseRun(ctx));
    r->n = g.n;
    r->m = g.m;
    r->N = (g.n + kTile - 1) / kTile * kTile;
    r->buf.alloc(ctx, dense_bytes(r->N));
    DenseBufs d = carve(r->buf.get(), r->N);
    if (g.m) {
        d.wbytes = k4_work_bytes(g.n, g.max_deg, r->N);
        r->work.alloc(ctx, d.wbytes);
        d.W = r->work.get();
    }
    dense_launch(ctx, g, d, out, &r->timer);
    if (g.m) r->d_ops = (const unsigned long long*)(d.W + k4_plan(g.max_deg, d.wbytes).oCount + 8);
    return r;
}

void dense_timing(DenseRun& r, Timing* tm) {
    // work models: the MFMA products are 2 N^3 (S), 8 N^3 ([A3 | AT] over
    // K = 2N) and, for the 4-clique products, 2 x 128 x 128 x dp per column
    // tile a row tile multiplies (dp: the vertex's degree rounded up to 128;
    // counted on the device) int8 operations; the adjacency pass writes
    // 2 N^2 + N^2/8 bytes; the edge pass reads four int32 products and writes
    // T, K4 and C4 per edge.  A pass run once per chunk of vertices is one
    // phase (times summed).
    const double N = (double)r.N, NN = N * N, n3 = NN * N;
    unsigned long long k4_ops = 0;
    if (r.d_ops) PGD_HIP(hipMemcpy(&k4_ops, r.d_ops, sizeof(k4_ops), hipMemcpyDeviceToHost));
    Timer& timer = r.timer;
    tm->phases.clear();
    for (size_t i = 0; i < timer.ev.size(); i++) {
        float ms = 0;
        PGD_HIP(hipEventElapsedTime(&ms, timer.ev[i].first, timer.ev[i].second));
        const std::string name = timer.names[i];
        Phase* ph = nullptr;
        for (auto& q : tm->phases)
            if (q.name == name) ph = &q;
        if (ph) {
            ph->ms += ms;
            continue;
        }
        double bytes = 0, units = 0;  // for the MFMA passes, bytes holds operations
        if (name == "k_dense_adj") { bytes = 2.0 * NN + NN / 8.0; units = NN; }
        if (name == "k_gemm_i8") { bytes = 2.0 * n3; units = 1; }
        if (name == "k_dense_digits") { bytes = 9.0 * NN; units = NN; }  // S and A read, four digit bytes
        if (name == "k_gemm_i8_pq") { bytes = 8.0 * n3; units = 1; }
        if (name == "k_dense_compress") { units = (double)r.n; }
        if (name == "k_dense_k4m") { bytes = (double)k4_ops; units = 1; }
        if (name == "k_dense_finish") { bytes = 36.0 * r.m; units = (double)r.m; }
        tm->phases.push_back(Phase{name, ms, bytes, units});
    }
    double best = 0;
    for (auto& q : tm->phases) best = std::max(best, q.ms);
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
    DenseBufs d = carve(base, N);
    // the 4-clique products' work area takes [A | 127 A]'s place (read only by
    // the products before them), in chunks of vertices
    d.W = (unsigned char*)d.A2;
    d.wbytes = align256(2 * (size_t)N * N);
    unsigned char* cp = base + dense_bytes(kNoAllocMaxN);
    EdgeCounts c{(uint32_t*)cp, (uint64_t*)(cp + align256(4 * (size_t)kNoAllocMaxM)),
                 (uint64_t*)(cp + align256(4 * (size_t)kNoAllocMaxM) + align256(8 * (size_t)kNoAllocMaxM))};
    if (g.m > 0) dense_launch(ctx, g, d, c, nullptr);
    epilogue(ctx, g, c, d_sums, d_micro);
    PGD_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace pgd
