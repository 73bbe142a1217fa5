// pgd_engine_edge.hip -- edge-centric engine: one wavefront per edge.
//
// The per-edge algorithm of graphlet_core.cpp:1408-1427 (mark N(v), split
// N(u) into triangle and star sets, 4-cycles through the star set's
// neighbour lists, 4-cliques through the triangle set's lists), mapped to
// CDNA4 as follows:
//   * a wavefront owns one edge at a time and pulls the next E_ordered
//     position from a device counter (heaviest edges first, dynamic balance);
//   * the perfect-hash array `ind` becomes two bitmaps per wavefront:
//     A = N(x)\{y} and B = T, in LDS when n fits (else in a per-wavefront HBM
//     slice), so membership tests are single LDS reads;
//   * the 2-hop walk is flattened: the 64 lanes hold 64 members of N(y), an
//     inclusive scan of their degrees maps every neighbour-list entry to one
//     lane-item, so hub lists and leaf lists are processed at full width and
//     all neighbour-list reads are contiguous;
//   * the walked side y is the endpoint with the smaller sum of neighbour
//     degrees (T, K4 and C4 are symmetric in the two endpoints), which is the
//     cheaper of the two 2-hop walks.
// Results per edge id: T, K4 = (ordered adjacent pairs inside T)/2 and C4 =
// edges between the two star sets; they equal the reference's counters
// (clique counts each adjacent pair of T once, cycle counts x in N(w) with
// ind[x]==1, i.e. x in N(v)\N[u]).
#include "pgd_internal.hpp"

using namespace pgd;

namespace {

constexpr int kWaves = 4;
constexpr int kLdsBudget = 60 * 1024;  // dynamic LDS per block for the bitmaps

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <bool kLds>
__device__ __forceinline__ uint32_t bm_load(const uint32_t* p) {
    if constexpr (kLds) return *(volatile const uint32_t*)p;
    else return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool kLds>
__device__ __forceinline__ void bm_set(uint32_t* p, uint32_t bit) {
    if constexpr (kLds) atomicOr(p, bit);
    else __hip_atomic_fetch_or(p, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool kLds>
__device__ __forceinline__ void bm_zero(uint32_t* p) {
    if constexpr (kLds) *(volatile uint32_t*)p = 0;
    else __hip_atomic_store(p, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ int32_t wave_incl_scan(int32_t x, int lane) {
    for (int off = 1; off < 64; off <<= 1) {
        int32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}

template <bool kLds>
__global__ __launch_bounds__(256) void k_edge_engine(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                                                     const int32_t* __restrict__ ev, const int32_t* __restrict__ eu,
                                                     const int64_t* __restrict__ svol,
                                                     const int32_t* __restrict__ order, int64_t m, int32_t words,
                                                     uint32_t* __restrict__ gbits,
                                                     unsigned long long* __restrict__ counter,
                                                     uint32_t* __restrict__ outT, uint64_t* __restrict__ outK,
                                                     uint64_t* __restrict__ outC) {
    extern __shared__ uint32_t smem[];
    __shared__ int64_t s_start[kWaves][64];
    __shared__ int32_t s_incl[kWaves][64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    uint32_t* A = kLds ? smem + (size_t)w * 2 * words
                       : gbits + ((size_t)blockIdx.x * kWaves + w) * 2 * (size_t)words;
    uint32_t* B = A + words;
    if constexpr (kLds) {
        for (int i = lane; i < 2 * words; i += 64) A[i] = 0;
    }
    wave_sync();
    for (;;) {
        unsigned long long pos = 0;
        if (lane == 0) pos = atomicAdd(counter, 1ULL);
        pos = __shfl(pos, 0, 64);
        if (pos >= (unsigned long long)m) break;
        const int32_t id = order[pos];
        const int32_t v = ev[id], u = eu[id];
        int32_t x, y;  // x is marked, y's neighbourhood is walked
        if (svol[u] <= svol[v]) { y = u; x = v; } else { y = v; x = u; }
        const int64_t xb = row[x], xe = row[x + 1], yb = row[y], ye = row[y + 1];
        // mark A = N(x) \ {y}
        for (int64_t j = xb + lane; j < xe; j += 64) {
            int32_t z = col[j];
            if (z != y) bm_set<kLds>(&A[z >> 5], 1u << (z & 31));
        }
        wave_sync();
        // T = N(y) & A, marked in B
        uint32_t tc = 0;
        for (int64_t j = yb + lane; j < ye; j += 64) {
            int32_t z = col[j];
            if (z != x && ((bm_load<kLds>(&A[z >> 5]) >> (z & 31)) & 1u)) {
                bm_set<kLds>(&B[z >> 5], 1u << (z & 31));
                tc++;
            }
        }
        wave_sync();
        // flattened 2-hop walk over N(y)\{x}: triangle members count B hits
        // (ordered pairs), star members count hits in A\B
        uint64_t k4x2 = 0, c4 = 0;
        for (int64_t cb = yb; cb < ye; cb += 64) {
            int64_t j = cb + lane;
            int32_t len = 0;
            int64_t st = 0;
            bool tri = false;
            if (j < ye) {
                int32_t z = col[j];
                if (z != x) {
                    st = row[z];
                    len = (int32_t)(row[z + 1] - st);
                    tri = (bm_load<kLds>(&A[z >> 5]) >> (z & 31)) & 1u;
                }
            }
            const uint64_t tmask = __ballot(tri);
            int32_t incl = wave_incl_scan(len, lane);
            const int32_t total = __shfl(incl, 63, 64);
            s_start[w][lane] = st;
            s_incl[w][lane] = incl;
            wave_sync();
            for (int32_t t = lane; t < total; t += 64) {
                int lo = 0, hi = 63;
                while (lo < hi) {
                    int mid = (lo + hi) >> 1;
                    if (s_incl[w][mid] > t) hi = mid;
                    else lo = mid + 1;
                }
                const int32_t excl = lo ? s_incl[w][lo - 1] : 0;
                const int32_t z2 = col[s_start[w][lo] + (t - excl)];
                const uint32_t a = (bm_load<kLds>(&A[z2 >> 5]) >> (z2 & 31)) & 1u;
                const uint32_t b = (bm_load<kLds>(&B[z2 >> 5]) >> (z2 & 31)) & 1u;
                if ((tmask >> lo) & 1ULL) k4x2 += b;
                else c4 += a & (b ^ 1u);
            }
            wave_sync();
        }
        // clear the bitmaps (only N(x) bits can be set)
        if ((int64_t)2 * words <= 4 * (xe - xb) + 256) {
            for (int i = lane; i < 2 * words; i += 64) bm_zero<kLds>(&A[i]);
        } else {
            for (int64_t j = xb + lane; j < xe; j += 64) {
                int32_t z = col[j];
                bm_zero<kLds>(&A[z >> 5]);
                bm_zero<kLds>(&B[z >> 5]);
            }
        }
        const uint64_t T = wave_sum64(tc), K2 = wave_sum64(k4x2), C = wave_sum64(c4);
        wave_sync();
        if (lane == 0) {
            outT[id] = (uint32_t)T;
            outK[id] = K2 >> 1;
            outC[id] = C;
        }
    }
}

// s(v) = sum of neighbour degrees (cost of walking v's 2-hop neighbourhood)
__global__ void k_svol(const int64_t* __restrict__ row, const int32_t* __restrict__ col, int64_t n,
                       int64_t* __restrict__ svol) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v = wave; v < n; v += nw) {
        int64_t s = 0;
        for (int64_t j = row[v] + lane; j < row[v + 1]; j += 64) {
            int32_t z = col[j];
            s += row[z + 1] - row[z];
        }
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) svol[v] = s;
    }
}

}  // namespace

namespace pgd {

void engine_edge(pgd_ctx* ctx, const GraphView& g, EdgeCounts out, Timing* t) {
    hipStream_t s = ctx->stream;
    const int32_t words = (int32_t)((g.n + 31) / 32);
    DBuf<int64_t> svol;
    DBuf<unsigned long long> counter;
    svol.alloc(ctx, std::max<int64_t>(g.n, 1));
    counter.alloc(ctx, 1);
    {
        int grid = std::min(div_up(g.n * 64, 256), ctx->num_cus * 16);
        hipLaunchKernelGGL(k_svol, dim3(std::max(grid, 1)), dim3(256), 0, s, g.row, g.col, g.n, svol.get());
        PGD_HIP(hipGetLastError());
    }
    PGD_HIP(hipMemsetAsync(counter.get(), 0, sizeof(unsigned long long), s));
    const size_t lds = (size_t)kWaves * 2 * words * sizeof(uint32_t);
    hipEvent_t e0, e1;
    PGD_HIP(hipEventCreate(&e0));
    PGD_HIP(hipEventCreate(&e1));
    if (lds <= (size_t)kLdsBudget) {
        int blocks = ctx->num_cus * std::max(1, std::min(8, (int)(kLdsBudget * 2 / std::max<size_t>(lds, 1))));
        blocks = std::min<int64_t>(blocks, div_up(g.m, kWaves));
        PGD_HIP(hipEventRecord(e0, s));
        hipLaunchKernelGGL(k_edge_engine<true>, dim3(std::max(blocks, 1)), dim3(64 * kWaves), lds, s, g.row, g.col,
                           g.ev, g.eu, svol.get(), g.order, g.m, words, nullptr, counter.get(), out.T, out.K4, out.C4);
        PGD_HIP(hipGetLastError());
        PGD_HIP(hipEventRecord(e1, s));
    } else {
        // per-wavefront bitmaps in HBM, zeroed once and kept clean by the kernel
        const size_t per_wave = (size_t)2 * words;
        int64_t want = (int64_t)ctx->num_cus * 4;
        int64_t cap_waves = std::max<int64_t>(kWaves, ((int64_t)4 << 30) / (int64_t)(per_wave * 4));
        int blocks = (int)std::max<int64_t>(1, std::min(want, cap_waves / kWaves));
        DBuf<uint32_t> bits;
        bits.alloc(ctx, (size_t)blocks * kWaves * per_wave);
        PGD_HIP(hipMemsetAsync(bits.get(), 0, (size_t)blocks * kWaves * per_wave * sizeof(uint32_t), s));
        PGD_HIP(hipEventRecord(e0, s));
        hipLaunchKernelGGL(k_edge_engine<false>, dim3(blocks), dim3(64 * kWaves), 0, s, g.row, g.col, g.ev, g.eu,
                           svol.get(), g.order, g.m, words, bits.get(), counter.get(), out.T, out.K4, out.C4);
        PGD_HIP(hipGetLastError());
        PGD_HIP(hipEventRecord(e1, s));
        PGD_HIP(hipStreamSynchronize(s));
    }
    PGD_HIP(hipEventSynchronize(e1));
    float ms = 0;
    PGD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    t->dominant_ms = ms;
    t->dominant_launches = 1;
}

}  // namespace pgd
