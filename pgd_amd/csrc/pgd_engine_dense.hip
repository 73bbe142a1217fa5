// pgd_engine_dense.hip -- dense-graph tier on the int8 matrix cores.
//
// For a dense graph the per-edge counts are contractions of the adjacency
// matrix A (0/1, held as int8, zero-padded to N = 64*ceil(n/64)):
//
//   S    = A.A                  codegrees; T(e) = S_vu for the edge e = (v,u)
//   A3   = A.S                  P(e) = A3_vu - d_v - d_u + 1 (4-cycles through e)
//   AT   = A.T, T = S o A       Q(e) = AT_vu + AT_uv  (sum over triangles {v,u,w}
//                               of T(v,w) + T(u,w))
//   K4(e) = edges among the common neighbours of v and u (bit rows, popcounts)
//   C4(e) = P - Q + 2T + 2K4    (the oriented engine's identity)
//
// Every product is an int8 x int8 -> int32 MFMA (v_mfma_i32_32x32x32_i8).
// S and T are at most n-2 < 2^14, so they enter the second products as two
// base-128 digits (each an int8 in [0, 127]); every int32 accumulator stays
// below n * 127 * 1 < 2^31 and the digits recombine exactly in int64.  All
// operands are symmetric, so both MFMA operands are read along rows.
#include <mutex>

#include "pgd_orient.hpp"

namespace pgd {
namespace {

constexpr int kDenseMaxN = 8192;  // int32 products of N^2 entries, bit rows of N/32 words
constexpr int kTile = 64;         // output tile of one block (2 x 2 waves of 32 x 32)

typedef int v16i __attribute__((ext_vector_type(16)));
typedef int v4i __attribute__((ext_vector_type(4)));

__global__ void d_fill_u32(uint32_t* __restrict__ p, int64_t n) {
    GRID_STRIDE(i, n) p[i] = 0;
}

// A (int8, N x N) and the bit rows (N x Wd words) from the CSR
__global__ void k_dense_adj(const int64_t* __restrict__ row, const int32_t* __restrict__ col, int64_t n, int64_t N,
                            int Wd, int8_t* __restrict__ A, uint32_t* __restrict__ bits) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v = wave; v < n; v += nw)
        for (int64_t p = row[v] + lane; p < row[v + 1]; p += 64) {
            const int32_t u = col[p];
            A[v * N + u] = 1;
            atomicOr(&bits[v * Wd + (u >> 5)], 1u << (u & 31));
        }
}

// C = A . X^T for N x N int8 row-major A and X (N a multiple of 64): one
// block per 64 x 64 output tile, one wavefront per 32 x 32 quarter.  Lane l
// feeds row (l mod 32) of its A rows and of its X rows, 16 bytes of k each
// step (the same k positions for both operands).
__global__ __launch_bounds__(256) void k_gemm_i8(const int8_t* __restrict__ A, const int8_t* __restrict__ X,
                                                 int32_t* __restrict__ C, int64_t N) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t ti = (int64_t)blockIdx.y * kTile + (w >> 1) * 32;
    const int64_t tj = (int64_t)blockIdx.x * kTile + (w & 1) * 32;
    const int8_t* pa = A + (ti + (lane & 31)) * N + 16 * (lane >> 5);
    const int8_t* px = X + (tj + (lane & 31)) * N + 16 * (lane >> 5);
    v16i acc0 = {}, acc1 = {};
    int64_t k = 0;
    for (; k + 64 <= N; k += 64) {  // two independent accumulators in flight
        const v4i a0 = *(const v4i*)(pa + k), x0 = *(const v4i*)(px + k);
        const v4i a1 = *(const v4i*)(pa + k + 32), x1 = *(const v4i*)(px + k + 32);
        acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, x0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, x1, acc1, 0, 0, 0);
    }
    for (; k < N; k += 32) {
        const v4i a0 = *(const v4i*)(pa + k), x0 = *(const v4i*)(px + k);
        acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, x0, acc0, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int64_t i = ti + (r & 3) + 8 * (r >> 2) + 4 * (lane >> 5), j = tj + (lane & 31);
        C[i * N + j] = acc0[r] + acc1[r];
    }
}

// base-128 digits of S and of T = S o A
__global__ void k_dense_digits(const int32_t* __restrict__ S, const int8_t* __restrict__ A, int64_t NN,
                               int8_t* __restrict__ s0, int8_t* __restrict__ s1, int8_t* __restrict__ t0,
                               int8_t* __restrict__ t1) {
    GRID_STRIDE(i, NN) {
        const int32_t s = S[i];
        const int32_t t = A[i] ? s : 0;
        s0[i] = (int8_t)(s & 127);
        s1[i] = (int8_t)(s >> 7);
        t0[i] = (int8_t)(t & 127);
        t1[i] = (int8_t)(t >> 7);
    }
}

// per-edge counts.  K4(e) for e = (v,u) is the number of edges among
// C = N(v) & N(u), half of the sum over the members w of C of
// popcount(N(w) & C).  A block keeps a slab of SW words of every bit row in
// LDS (column words [q SW, q SW + SW) of rows 0..N-1) and works through groups
// of kK4Waves x 64 edges, each group once per slab: a wavefront takes 64
// edges of the group one after another, lists the members of C in LDS
// (16-bit ids, N <= 8192), and every lane adds, for the members lane,
// lane + 64, ..., the popcounts of the member's slab row AND the slab of C.
// Lane i keeps the running sum of the wavefront's i-th edge across the slabs
// and, after the last one, finishes that edge's T, K4 and C4 from the
// products.  Bit rows are re-read from L2 once per edge and slab; the member
// rows come from LDS.
constexpr int kMaxWordsPerLane = kDenseMaxN / 32 / 64;
constexpr int kK4Waves = 8;
constexpr int kK4Threads = 64 * kK4Waves;
constexpr size_t kLdsBytes = 160 * 1024;

template <int SW>
__global__ __launch_bounds__(kK4Threads) void k_dense_k4(
    const int32_t* __restrict__ ev, const int32_t* __restrict__ eu, const int64_t* __restrict__ row, int64_t m,
    int64_t N, int Wd, const int32_t* __restrict__ S, const int32_t* __restrict__ P0, const int32_t* __restrict__ P1,
    const int32_t* __restrict__ Q0, const int32_t* __restrict__ Q1, const uint32_t* __restrict__ bits,
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
        uint32_t mine = 0;  // twice K4 of edge eb + lane, summed over the slabs
        for (int q = 0; q < nslab; q++) {
            __syncthreads();  // no wavefront still reads the previous slab
            for (int64_t i = threadIdx.x; i < N * SW; i += kK4Threads) {
                const int64_t r = i / SW;
                const int wj = q * SW + (int)(i % SW);
                slab[i] = wj < Wd ? bits[r * Wd + wj] : 0u;
            }
            __syncthreads();
            for (int i = 0; i < ne; i++) {
                const int64_t e = eb + i;
                const int64_t v = ev[e], u = eu[e];
                uint32_t cw[kMaxWordsPerLane];
                uint32_t cnt = 0;
#pragma unroll
                for (int k = 0; k < kMaxWordsPerLane; k++) {
                    const int x = lane + 64 * k;
                    cw[k] = (k < nwl && x < Wd) ? (bits[v * Wd + x] & bits[u * Wd + x]) : 0u;
                    cnt += __popc(cw[k]);
                }
                const uint32_t incl = wave_incl_sum(cnt);
                const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
                uint32_t pos = incl - cnt;
#pragma unroll
                for (int k = 0; k < kMaxWordsPerLane; k++) {
                    uint32_t x = cw[k];
                    const uint32_t base = (uint32_t)(lane + 64 * k) * 32u;
                    while (x) {
                        list[pos++] = (uint16_t)(base + (uint32_t)(__ffs(x) - 1));
                        x &= x - 1;
                    }
                }
                // the slab's words of C, from the lanes that hold them
                uint32_t cs[SW];
#pragma unroll
                for (int j = 0; j < SW; j++) {
                    const int wj = q * SW + j;
                    uint32_t src = 0;
#pragma unroll
                    for (int k = 0; k < kMaxWordsPerLane; k++)
                        if (k == (wj >> 6)) src = cw[k];
                    const uint32_t got = (uint32_t)__shfl((int)src, wj & 63, 64);
                    cs[j] = wj < Wd ? got : 0u;
                }
                __builtin_amdgcn_wave_barrier();
                uint32_t acc = 0;
                for (uint32_t r = lane; r < total; r += 64) {
                    const uint32_t* rw = slab + (size_t)list[r] * SW;
#pragma unroll
                    for (int j = 0; j < SW; j++) acc += __popc(rw[j] & cs[j]);
                }
                __builtin_amdgcn_wave_barrier();  // the list is read before the next edge rewrites it
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) acc += (uint32_t)__shfl_xor((int)acc, off, 64);
                if (lane == i) mine += acc;
            }
        }
        if (lane < ne) {
            const int64_t e = eb + lane;
            const int64_t v = ev[e], u = eu[e];
            const int64_t vu = v * N + u, uv = u * N + v;
            const uint64_t t = (uint64_t)S[vu];
            const uint64_t a3 = (uint64_t)P0[vu] + 128 * (uint64_t)P1[vu];
            const uint64_t dv = (uint64_t)(row[v + 1] - row[v]), du = (uint64_t)(row[u + 1] - row[u]);
            const uint64_t p = a3 - dv - du + 1;
            const uint64_t qq = (uint64_t)Q0[vu] + 128 * (uint64_t)Q1[vu] + (uint64_t)Q0[uv] + 128 * (uint64_t)Q1[uv];
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

// The engine's arrays for one graph: A, its bit rows, the products S, P (two
// digits), Q (two digits), the digits of S and T, and the per-edge counts
// (the caller's, or scratch)
struct DenseBufs {
    int8_t *A, *s0, *s1, *t0, *t1;
    int32_t *S, *P0, *P1, *Q0, *Q1;
    uint32_t* bits;
};

// bytes of DenseBufs for padded size N, each array 256-byte aligned
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t dense_bytes(int64_t N) {
    return 5 * align256((size_t)N * N) + 5 * align256(4 * (size_t)N * N) + align256((size_t)N * N / 8);
}
DenseBufs carve(unsigned char* base, int64_t N) {
    const size_t b1 = align256((size_t)N * N), b4 = align256(4 * (size_t)N * N);
    DenseBufs d;
    unsigned char* p = base;
    d.A = (int8_t*)p; p += b1;
    d.s0 = (int8_t*)p; p += b1;
    d.s1 = (int8_t*)p; p += b1;
    d.t0 = (int8_t*)p; p += b1;
    d.t1 = (int8_t*)p; p += b1;
    d.S = (int32_t*)p; p += b4;
    d.P0 = (int32_t*)p; p += b4;
    d.P1 = (int32_t*)p; p += b4;
    d.Q0 = (int32_t*)p; p += b4;
    d.Q1 = (int32_t*)p; p += b4;
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
                       g.ev, g.eu, g.row, g.m, N, Wd, d.S, d.P0, d.P1, d.Q0, d.Q1, d.bits, out.T, out.K4, out.C4);
}

// the launch sequence; every array of d is written before it is read (A and
// the bit rows are cleared first), so no state carries over between calls
void dense_launch(pgd_ctx* ctx, const GraphView& g, const DenseBufs& d, EdgeCounts out, Timer* timer) {
    hipStream_t s = ctx->stream;
    const int64_t n = g.n, m = g.m;
    const int64_t N = (n + kTile - 1) / kTile * kTile, NN = N * N;
    const int Wd = (int)(N / 32);
    if (timer) timer->begin("k_dense_adj");
    const int gf = grid_cap(ctx, (NN / 4 + 255) / 256, 16);
    hipLaunchKernelGGL(d_fill_u32, dim3(gf), dim3(256), 0, s, (uint32_t*)d.A, NN / 4);
    hipLaunchKernelGGL(d_fill_u32, dim3(grid_cap(ctx, (N * Wd + 255) / 256, 16)), dim3(256), 0, s, d.bits, N * Wd);
    hipLaunchKernelGGL(k_dense_adj, dim3(grid_cap(ctx, (n * 64 + 255) / 256, 16)), dim3(256), 0, s, g.row, g.col, n,
                       N, Wd, d.A, d.bits);
    PGD_HIP(hipGetLastError());
    if (timer) timer->end();
    const dim3 gg((unsigned)(N / kTile), (unsigned)(N / kTile));
    if (timer) timer->begin("k_gemm_i8");
    hipLaunchKernelGGL(k_gemm_i8, gg, dim3(256), 0, s, d.A, d.A, d.S, N);
    PGD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_dense_digits, dim3(grid_cap(ctx, (NN + 255) / 256, 16)), dim3(256), 0, s, d.S, d.A, NN,
                       d.s0, d.s1, d.t0, d.t1);
    hipLaunchKernelGGL(k_gemm_i8, gg, dim3(256), 0, s, d.A, d.s0, d.P0, N);
    hipLaunchKernelGGL(k_gemm_i8, gg, dim3(256), 0, s, d.A, d.s1, d.P1, N);
    hipLaunchKernelGGL(k_gemm_i8, gg, dim3(256), 0, s, d.A, d.t0, d.Q0, N);
    hipLaunchKernelGGL(k_gemm_i8, gg, dim3(256), 0, s, d.A, d.t1, d.Q1, N);
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
    // work models: the MFMA products are 5 x 2 N^3 int8 operations; the adjacency
    // pass writes N^2 + N^2/8 bytes; the edge pass reads its five int32 entries
    // and, per common neighbour, three bit rows
    const double flops = 5.0 * 2.0 * (double)N * N * N;
    tm->phases.clear();
    double best = 0;
    for (size_t i = 0; i < timer.ev.size(); i++) {
        float ms = 0;
        PGD_HIP(hipEventElapsedTime(&ms, timer.ev[i].first, timer.ev[i].second));
        double bytes = 0, units = 0;
        if (i == 0) { bytes = (double)NN + NN / 8.0; units = (double)NN; }
        if (i == 1) { bytes = flops; units = 5; }  // for the MFMA pass: operations, not bytes
        if (i == 2) { bytes = 28.0 * m; units = (double)m; }
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
