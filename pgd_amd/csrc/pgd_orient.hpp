// pgd_orient.hpp -- shared pieces of the degree-oriented engine: launch and
// rocPRIM helpers, block-level device helpers, and the engine's persistent
// per-graph state (relabelled CSR, vertex bins, the tiled wedge plan).
#pragma once
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <vector>

#include "pgd_internal.hpp"

namespace pgd {

#define GRID_STRIDE(i, n) \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
inline int grid_cap(pgd_ctx* ctx, int64_t items, int per_cu) {
    int64_t g = std::min<int64_t>(items, (int64_t)ctx->num_cus * per_cu);
    return (int)std::max<int64_t>(g, 1);
}

template <class T, class U>
void excl_sum(pgd_ctx* ctx, const T* in, U* out, size_t n) {
    if (!n) return;
    size_t bytes = 0;
    PGD_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, U(0), n, rocprim::plus<U>(), ctx->stream));
    void* t = tmp_storage(ctx, bytes);
    PGD_HIP(rocprim::exclusive_scan(t, bytes, in, out, U(0), n, rocprim::plus<U>(), ctx->stream));
}

template <class K, class V>
void sort_kv(pgd_ctx* ctx, K* kin, K* kout, V* vin, V* vout, size_t n, unsigned b1) {
    if (!n) return;
    size_t bytes = 0;
    PGD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0u, b1, ctx->stream));
    void* t = tmp_storage(ctx, bytes);
    PGD_HIP(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, n, 0u, b1, ctx->stream));
}

template <class K>
void sort_k(pgd_ctx* ctx, K* kin, K* kout, size_t n, unsigned b1) {
    if (!n) return;
    size_t bytes = 0;
    PGD_HIP(rocprim::radix_sort_keys(nullptr, bytes, kin, kout, n, 0u, b1, ctx->stream));
    void* t = tmp_storage(ctx, bytes);
    PGD_HIP(rocprim::radix_sort_keys(t, bytes, kin, kout, n, 0u, b1, ctx->stream));
}

// maximum of in[0..n) read back to the host (synchronises the stream)
template <class T>
T device_max(pgd_ctx* ctx, const T* in, size_t n) {
    if (!n) return T(0);
    DBuf<T> mx;
    mx.alloc(ctx, 1);
    size_t bytes = 0;
    PGD_HIP(rocprim::reduce(nullptr, bytes, in, mx.get(), T(0), n, rocprim::maximum<T>(), ctx->stream));
    void* t = tmp_storage(ctx, bytes);
    PGD_HIP(rocprim::reduce(t, bytes, in, mx.get(), T(0), n, rocprim::maximum<T>(), ctx->stream));
    T h = 0;
    PGD_HIP(hipMemcpyAsync(&h, mx.get(), sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    PGD_HIP(hipStreamSynchronize(ctx->stream));
    return h;
}

// sum of in[0..n) (as type S) read back to the host (synchronises the stream)
template <class S, class T>
S device_sum(pgd_ctx* ctx, const T* in, size_t n) {
    if (!n) return S(0);
    DBuf<S> r;
    r.alloc(ctx, 1);
    size_t bytes = 0;
    PGD_HIP(rocprim::reduce(nullptr, bytes, in, r.get(), S(0), n, rocprim::plus<S>(), ctx->stream));
    void* t = tmp_storage(ctx, bytes);
    PGD_HIP(rocprim::reduce(t, bytes, in, r.get(), S(0), n, rocprim::plus<S>(), ctx->stream));
    S h = 0;
    PGD_HIP(hipMemcpyAsync(&h, r.get(), sizeof(S), hipMemcpyDeviceToHost, ctx->stream));
    PGD_HIP(hipStreamSynchronize(ctx->stream));
    return h;
}

inline int bitsfor(int64_t v) {
    int b = 1;
    while (b < 63 && ((int64_t)1 << b) <= v) b++;
    return b;
}

// capacity-reusing allocation: keeps the buffer when it is already large enough
template <class T>
void ensure(DBuf<T>& b, pgd_ctx* ctx, size_t n) {
    if (n == 0) n = 1;
    if (b.get() == nullptr || b.n < n) b.alloc(ctx, n);
}

// ---------------------------------------------------------------------------
// block-level device helpers (blockDim.x == NT)
// ---------------------------------------------------------------------------
// inclusive scan of s[0..k) in place, NT threads, k arbitrary
template <int NT, class T>
__device__ void block_incl_scan(T* s, int k, T* part) {
    const int tid = threadIdx.x;
    const int per = (k + NT - 1) / NT;
    const int b = tid * per, e = min(b + per, k);
    T acc = 0;
    for (int i = b; i < e; i++) { acc += s[i]; s[i] = acc; }
    part[tid] = acc;
    __syncthreads();
    for (int off = 1; off < NT; off <<= 1) {
        T y = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += y;
        __syncthreads();
    }
    T base = tid ? part[tid - 1] : 0;
    for (int i = b; i < e; i++) s[i] += base;
    __syncthreads();
}

// first index i in [0,k) with incl[i] > t (incl nondecreasing, incl[k-1] > t)
template <class T>
__device__ __forceinline__ int owner_of(const T* incl, int k, T t) {
    int lo = 0, hi = k - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (incl[mid] > t) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// index of c in L[lo..hi) (ascending) or -1
__device__ __forceinline__ int find_in(const int32_t* L, int lo, int hi, int32_t c) {
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        int32_t x = L[mid];
        if (x < c) lo = mid + 1;
        else if (x > c) hi = mid;
        else return mid;
    }
    return -1;
}

// inclusive sum over the lanes 0..lane of a wavefront: within rows of 16
// (DPP row shifts), then the row totals carried by DPP row broadcasts
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); // row_bcast:15 -> rows 1, 3
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); // row_bcast:31 -> rows 2, 3
    return x;
}

// Runs of equal keys in lane order (the whole wavefront takes part; a lane
// with nothing to contribute passes a key of its own, e.g. ~lane): `start` is
// the first lane of this lane's run, `last` marks the run's last lane.  Used
// to turn the same-address LDS atomics of a run into one.  The run starts
// come from one ballot, and the sums below are DPP scans (row shifts, then
// the row broadcasts of lanes 15 and 31), so no LDS instruction is spent.
struct LaneRun {
    int start;
    bool last;
};
__device__ __forceinline__ LaneRun lane_run(uint32_t key) {
    const int lane = threadIdx.x & 63;
    const uint32_t kp = (uint32_t)__shfl_up((int)key, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || kp != key);
    const int start = 63 - __clzll((long long)(heads & ((2ULL << lane) - 1)));
    return LaneRun{start, lane == 63 || ((heads >> (lane + 1)) & 1) != 0};
}
// one step of a segmented inclusive scan: the value `off` lanes below
// (within the row), or a row broadcast, when it lies in this lane's run
#define PGD_RUN_STEPS(OP, T, MOV)                                                                              \
    {                                                                                                          \
        const int lane_ = threadIdx.x & 63, r_ = lane_ & 15;                                                   \
        T y_;                                                                                                  \
        y_ = MOV(x, 0x111, 0xf); if (r_ >= 1 && lane_ - 1 >= start) x = x OP y_;                               \
        y_ = MOV(x, 0x112, 0xf); if (r_ >= 2 && lane_ - 2 >= start) x = x OP y_;                               \
        y_ = MOV(x, 0x114, 0xf); if (r_ >= 4 && lane_ - 4 >= start) x = x OP y_;                               \
        y_ = MOV(x, 0x118, 0xf); if (r_ >= 8 && lane_ - 8 >= start) x = x OP y_;                               \
        y_ = MOV(x, 0x142, 0xa); if (((lane_ >> 4) & 1) && start <= (lane_ & ~15) - 1) x = x OP y_;            \
        y_ = MOV(x, 0x143, 0xc); if ((lane_ >> 5) && start <= 31) x = x OP y_;                                 \
    }
#define PGD_DPP32(v, ctrl, rmask) (uint32_t) __builtin_amdgcn_update_dpp(0, (int)(v), ctrl, rmask, 0xf, true)
#define PGD_DPP64(v, ctrl, rmask)                                                                          \
    ((uint64_t)(uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)((v) >> 32), ctrl, rmask, 0xf, true) \
         << 32 |                                                                                           \
     (uint64_t)(uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v), ctrl, rmask, 0xf, true))
// inclusive sum / or of x over the lanes of this lane's run up to it (the
// run's last lane holds the run's total)
__device__ __forceinline__ uint32_t run_sum32(uint32_t x, int start) {
    PGD_RUN_STEPS(+, uint32_t, PGD_DPP32)
    return x;
}
__device__ __forceinline__ uint64_t run_sum64(uint64_t x, int start) {
    PGD_RUN_STEPS(+, uint64_t, PGD_DPP64)
    return x;
}
__device__ __forceinline__ uint32_t run_or(uint32_t x, int start) {
    PGD_RUN_STEPS(|, uint32_t, PGD_DPP32)
    return x;
}

// first position in col[lo..hi) (ascending) holding a value >= x
__device__ __forceinline__ int64_t lower_bound_g(const int32_t* __restrict__ col, int64_t lo, int64_t hi, int32_t x) {
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (col[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// per-pass HIP event timer on the context's stream
// ---------------------------------------------------------------------------
struct Timer {
    pgd_ctx* ctx;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::vector<const char*> names;
    void begin(const char* name) {
        hipEvent_t a, b;
        PGD_HIP(hipEventCreate(&a));
        PGD_HIP(hipEventCreate(&b));
        PGD_HIP(hipEventRecord(a, ctx->stream));
        ev.push_back({a, b});
        names.push_back(name);
    }
    void end() { PGD_HIP(hipEventRecord(ev.back().second, ctx->stream)); }
    ~Timer() {
        for (auto& p : ev) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    }
};

// ---------------------------------------------------------------------------
// tiled wedge plan (heavy anchors of the 4-cycle pass)
// ---------------------------------------------------------------------------
// One segment: the run of N(b) (ascending CSR slots [start, start+len)) whose
// endpoints a fall in one endpoint tile, above max(b, c), for the neighbour
// b of an anchor c; sab is the slot of b in N(c) (the edge (c, b)).  The top bit of len marks
// a wide list (d(b) * max degree >= 2^32: its slots take 64-bit sums).
// Slots are below 2^31 (the plan requires 2m < 2^31): 12 bytes a segment.
struct WedgeSeg {
    uint32_t start;
    int32_t len;
    int32_t sab;
};

// one unit of work of the tiled kernel: anchor (index into the heavy list)
// and its endpoint tiles [t0, t1)
struct WedgeItem {
    int32_t hub, t0, t1, pad;
};

struct WedgePlan {
    int log2r = 15;                 // endpoint tile = 2^log2r consecutive ranks (LDS counters)
    int tmax = 128;                 // tiles per item
    int64_t nitems = 0, segcap = 0, wedges = 0;  // segcap: segments per item (unless one cell has more)
    int grid = 0;
    DBuf<int32_t> sfx;              // per slot (c, b) of a heavy anchor: offset in N(b) of the first endpoint above max(b, c)
    DBuf<uint64_t> tt;              // tile table entries: (tile << 32) | offset in N(b)
    DBuf<int64_t> ttrow;            // first tile-table entry of each vertex (n+1)
    DBuf<int32_t> hbase;            // first endpoint tile of each heavy anchor
    DBuf<int64_t> htoff;            // first (anchor, tile) cell of each heavy anchor (+1)
    DBuf<uint32_t> htseg;           // segments per (anchor, tile) cell (+1, zero)
    DBuf<int64_t> cseg;             // first segment of each cell in seg (+1: the total)
    DBuf<WedgeSeg> seg;             // every cell's segments, in cell order
    DBuf<uint32_t> gcur;            // per-cell cursors of the plan's segment walk (cells outside LDS)
    DBuf<uint64_t> htw;             // wedges per (anchor, tile) cell
    DBuf<WedgeItem> items;          // work queue, largest first
    DBuf<uint64_t> ps;              // per-slot P accumulator (ascending CSR slots), 64-bit sums
    DBuf<uint32_t> ps32;            // per-slot P accumulator of the lists that are not wide
    int32_t wide_deg = 0;           // lists of at least this degree are wide
    DBuf<uint32_t> qctr;            // queue head
    DBuf<unsigned long long> pacc;  // totals of the plan's cells: segments, wedges, finished blocks
};

// Persistent state of the oriented engine for one graph: the relabelled CSR,
// the vertex bins, scratch and accumulators.  Buffers are kept across calls
// (their contents are rebuilt every call), so the counting launches can be
// captured once in a HIP graph and replayed.
struct OrientState {
    OrientParams prm;
    int64_t n = 0, m = 0;
    DBuf<int32_t> rank, rdeg, rcol, reid, outdeg, v0, flags;
    DBuf<int64_t> rrow, ostart, off, ub;
    DBuf<uint64_t> k0, k1;
    DBuf<int32_t> b_small, b_mid, b_large, b_huge, b_light, b_heavy;
    int64_t n_small = 0, n_mid = 0, n_large = 0, n_huge = 0, n_light = 0, n_heavy = 0;
    int kmax_lds = 0, maxk = 0, ghuge = 0, kq = 0, max_deg = 0;  // kq: local-graph size the large bin is carved for
    size_t small_lds = 0, mid_lds = 0, large_lds = 0;              // clique pass, per block
    size_t small_lds_tri = 0, mid_lds_tri = 0, large_lds_tri = 0;  // triangle pass, per block
    size_t huge_stride = 0;                                         // HBM slice of a huge-bin block
    int64_t mw_small = 0, mw_mid = 0, mw_large = 0, mw_huge = 0;  // bit-matrix words per block
    DBuf<char> hscratch;
    DBuf<uint64_t> elist;  // local-edge lists of the clique pass (per block of each LDS bin)
    uint64_t *el_small = nullptr, *el_mid = nullptr, *el_large = nullptr, *el_huge = nullptr;
    int64_t ecap_small = 0, ecap_mid = 0, ecap_large = 0, ecap_huge = 0;
    DBuf<uint64_t> tlist, tcur;  // local-edge lists kept from the triangle pass for the clique pass
    DBuf<int64_t> toff;
    DBuf<int32_t> tlen;
    int64_t tcap = 0;
    DBuf<uint64_t> P, Q;
    DBuf<uint32_t> TS;  // per-slot triangle counts of the triangle pass
    DBuf<unsigned long long> work;
    DBuf<uint32_t> lq;  // work queues of the mid, large and HBM bins of the two local passes
    DBuf<uint64_t> ohtab;  // out-list hash tables of the vertices of large out-degree
    DBuf<int64_t> ohoff;   // per vertex rank: (first entry << 6) | log2 size, or -1
    bool has_oh = false;
    hipEvent_t fork[6] = {};  // fork and join of the side stream of the local passes
    WedgePlan wp;
    ~OrientState() {
        for (hipEvent_t e : fork)
            if (e) (void)hipEventDestroy(e);
    }
};

// heavy-vertex 4-cycle pass (pgd_engine_wedge.hip)
void wedge_plan_build(pgd_ctx* ctx, OrientState& st);
void wedge_plan_count(pgd_ctx* ctx, OrientState& st, unsigned long long* work);

}  // namespace pgd
