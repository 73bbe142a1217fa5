// pgd_core.hip -- context, device CSR construction, preprocessing (neighbour
// order, orientation, k-core, edge order), the counting epilogue and the
// C-ABI glue for the plain (one graph) path.
//
// Device data layout (HBM), per graph:
//   row   int64[n+1]   CSR offsets            (graphlet_core.h:84-86 `vertices`)
//   col   int32[2m]    neighbour ids           (`edges`, reference order)
//   src   int32[2m]    owning vertex of each CSR slot
//   deg   int32[n]
//   ev/eu int32[m]     edge endpoints by edge id (graphlet_core.cpp:929-949)
//   order int32[m]     E_ordered               (graphlet_core.cpp:854-867)
//   T u32[m], K4 u64[m], C4 u64[m]             per-edge counts by edge id
//   micro u64[8][m]    per-edge columns at E_ordered position (SoA)
// Sorting uses rocPRIM's stable LSD radix sort; scans/reductions use rocPRIM.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <climits>
#include <new>

#include "pgd_epilogue.hpp"
#include "pgd_internal.hpp"
#include "pgd_reader.hpp"

using namespace pgd;

// ===========================================================================
// memory
// ===========================================================================
namespace pgd {

// Size class of a request: the next value m * 2^k with m in [8, 15] (at most
// 12.5% above the request), at least 256 bytes.
size_t size_class(size_t bytes) {
    if (bytes <= 256) return 256;
    int k = 0;
    while ((bytes - 1) >> (k + 3) >= 8) k++;  // mantissa of bytes-1 in [8, 15] at scale 2^k
    return (((bytes - 1) >> k) + 1) << k;
}

// cached blocks beyond this many bytes go back to the runtime pool
constexpr size_t kCacheLimit = (size_t)48 << 30;

void* dmalloc(pgd_ctx* ctx, size_t bytes) {
    if (ctx->mem_cap > 0 && ctx->mem_used.load() + (int64_t)bytes > ctx->mem_cap)
        throw Error{PGD_ERR_OOM, "device memory cap exceeded (" + std::to_string(ctx->mem_used.load() + bytes) +
                                     " > " + std::to_string(ctx->mem_cap) + " bytes)"};
    const size_t cls = size_class(bytes);
    void* p = nullptr;
    {
        // a cached block of this class: released earlier in the order of
        // this context's stream, so work enqueued on it from now on follows
        // every use of the block's previous owner
        std::lock_guard<std::mutex> lk(ctx->cache_mu);
        auto it = ctx->cache.find(cls);
        if (it != ctx->cache.end()) {
            p = it->second;
            ctx->cache.erase(it);
            ctx->cached_bytes -= cls;
        }
    }
    if (!p) {
        hipError_t e = hipMallocAsync(&p, cls, ctx->stream);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            // out of memory with blocks parked in the cache: give them back and retry
            release_cached(ctx);
            e = hipMallocAsync(&p, cls, ctx->stream);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error{PGD_ERR_OOM, "hipMallocAsync(" + std::to_string(cls) + ") failed: " + hipGetErrorString(e)};
        }
    }
    ctx->mem_used += (int64_t)bytes;
    return p;
}

void dfree(pgd_ctx* ctx, void* p, size_t bytes) {
    if (!p) return;
    const size_t cls = size_class(bytes);
    ctx->mem_used -= (int64_t)bytes;
    std::vector<void*> spill;
    {
        std::lock_guard<std::mutex> lk(ctx->cache_mu);
        ctx->cache.emplace(cls, p);
        ctx->cached_bytes += cls;
        // over the limit: the largest blocks go back first
        while (ctx->cached_bytes > kCacheLimit && !ctx->cache.empty()) {
            auto last = std::prev(ctx->cache.end());
            ctx->cached_bytes -= last->first;
            spill.push_back(last->second);
            ctx->cache.erase(last);
        }
    }
    for (void* q : spill) (void)hipFreeAsync(q, ctx->stream);
}

void release_cached(pgd_ctx* ctx) {
    std::vector<void*> all;
    {
        std::lock_guard<std::mutex> lk(ctx->cache_mu);
        for (auto& kv : ctx->cache) all.push_back(kv.second);
        ctx->cache.clear();
        ctx->cached_bytes = 0;
    }
    for (void* q : all) (void)hipFreeAsync(q, ctx->stream);
    if (!all.empty()) (void)hipStreamSynchronize(ctx->stream);
}

unsigned long long* pinned_words(pgd_ctx* ctx) {
    if (!ctx->pin) {
        void* h = nullptr;
        PGD_HIP(hipHostMalloc(&h, 8 * sizeof(unsigned long long), hipHostMallocMapped));
        void* d = nullptr;
        if (hipHostGetDevicePointer(&d, h, 0) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipHostFree(h);
            throw Error{PGD_ERR_HIP, "hipHostGetDevicePointer failed"};
        }
        ctx->pin = (unsigned long long*)h;
        ctx->pin_dev = (unsigned long long*)d;
    }
    return ctx->pin_dev;
}

void* tmp_storage(pgd_ctx* ctx, size_t bytes) {
    if (bytes > ctx->tmp_bytes) {
        if (ctx->tmp) dfree(ctx, ctx->tmp, ctx->tmp_bytes);
        ctx->tmp = nullptr;
        ctx->tmp_bytes = 0;
        size_t b = std::max(bytes, (size_t)1 << 20);
        ctx->tmp = dmalloc(ctx, b);
        ctx->tmp_bytes = b;
    }
    return ctx->tmp;
}

}  // namespace pgd

namespace {

int bits_for(int64_t maxval) {  // bits needed to represent values in [0, maxval]
    int b = 0;
    while (b < 63 && ((int64_t)1 << b) <= maxval) b++;
    return std::max(b, 1);
}

template <class K, class V>
void sort_pairs(pgd_ctx* ctx, K* kin, K* kout, V* vin, V* vout, size_t n, unsigned b0, unsigned b1) {
    if (n == 0) return;
    size_t bytes = 0;
    PGD_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, b0, b1, ctx->stream));
    void* t = tmp_storage(ctx, bytes);
    PGD_HIP(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, n, b0, b1, ctx->stream));
}

template <class K>
void sort_keys(pgd_ctx* ctx, K* kin, K* kout, size_t n, unsigned b0, unsigned b1) {
    if (n == 0) return;
    size_t bytes = 0;
    PGD_HIP(rocprim::radix_sort_keys(nullptr, bytes, kin, kout, n, b0, b1, ctx->stream));
    void* t = tmp_storage(ctx, bytes);
    PGD_HIP(rocprim::radix_sort_keys(t, bytes, kin, kout, n, b0, b1, ctx->stream));
}

template <class T, class U>
void exclusive_sum(pgd_ctx* ctx, const T* in, U* out, size_t n) {
    if (n == 0) return;
    size_t bytes = 0;
    PGD_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, U(0), n, rocprim::plus<U>(), ctx->stream));
    void* t = tmp_storage(ctx, bytes);
    PGD_HIP(rocprim::exclusive_scan(t, bytes, in, out, U(0), n, rocprim::plus<U>(), ctx->stream));
}

template <class T>
void reduce_max(pgd_ctx* ctx, const T* in, T* out, size_t n) {
    size_t bytes = 0;
    PGD_HIP(rocprim::reduce(nullptr, bytes, in, out, T(0), n, rocprim::maximum<T>(), ctx->stream));
    void* t = tmp_storage(ctx, bytes);
    PGD_HIP(rocprim::reduce(t, bytes, in, out, T(0), n, rocprim::maximum<T>(), ctx->stream));
}

constexpr int kBlock = 256;

inline int grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 1 << 20) g = 1 << 20;
    return (int)g;
}

// ===========================================================================
// CSR construction kernels
// ===========================================================================
// keep[i]: line i is an edge (not a self loop).  With drop_neg, a line whose
// first id is negative was removed by the reader (its self loops) and is not
// an error.
__global__ void k_line_keep(const int32_t* __restrict__ s, const int32_t* __restrict__ d, int64_t L, int64_t n,
                            int drop_neg, int32_t* __restrict__ keep, uint32_t* __restrict__ bad) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t a = s[i], b = d[i];
        if (drop_neg && a < 0) { keep[i] = 0; continue; }
        if (a < 0 || b < 0 || a >= n || b >= n) { atomicOr(bad, 1u); keep[i] = 0; continue; }
        keep[i] = a != b;
    }
}

// ---- renumbering of an edge list by first appearance (read_edge_list,
// graphlet_core.cpp:413-428): position p = 2 * line + (0 for v, 1 for u);
// a raw id's new id is the number of raw ids first seen before it.
// first[x] = the first position of raw id x over the non-loop lines
__global__ void k_first_pos(const int32_t* __restrict__ v, const int32_t* __restrict__ u, int64_t L, int64_t maxv,
                            uint32_t* __restrict__ first, unsigned long long* __restrict__ stat) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = v[i], b = u[i];
        if (a == b) { atomicAdd(&stat[0], 1ull); continue; }  // a self loop: no id
        if (a < 0 || b < 0 || a > maxv || b > maxv) { atomicOr(&stat[1], 1ull); continue; }
        atomicMin(&first[a], (uint32_t)(2 * i));
        atomicMin(&first[b], (uint32_t)(2 * i + 1));
    }
}
// mark[p] = 1 when position p is the first appearance of its raw id
__global__ void k_first_mark(const int32_t* __restrict__ v, const int32_t* __restrict__ u, int64_t L,
                             const uint32_t* __restrict__ first, uint32_t* __restrict__ mark) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < 2 * L; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = p >> 1;
        const int32_t a = v[i], b = u[i];
        mark[p] = a != b && first[(p & 1) ? b : a] == (uint32_t)p;
    }
}
// the new ids in place; self-loop lines become (-1, -1), dropped later
__global__ void k_renumber(int32_t* __restrict__ v, int32_t* __restrict__ u, int64_t L,
                           const uint32_t* __restrict__ first, const uint32_t* __restrict__ rank) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = v[i], b = u[i];
        if (a == b) {
            v[i] = -1;
            u[i] = -1;
            continue;
        }
        v[i] = (int32_t)rank[first[a]];
        u[i] = (int32_t)rank[first[b]];
    }
}
__global__ void k_fill_u32_max(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0xffffffffu;
}

__global__ void k_pair_keys(const int32_t* __restrict__ s, const int32_t* __restrict__ d, const int32_t* __restrict__ keep,
                            int64_t L, uint64_t* __restrict__ key, int32_t* __restrict__ idx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t a = (uint32_t)s[i], b = (uint32_t)d[i];
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        key[i] = keep[i] ? (((uint64_t)lo << 32) | hi) : ~0ULL;
        idx[i] = (int32_t)i;
    }
}

// sorted by pair key, stable by line: every later copy of a pair is dropped
__global__ void k_dedup_mark(const uint64_t* __restrict__ key, const int32_t* __restrict__ idx, int64_t L,
                             int32_t* __restrict__ keep) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < L; j += (int64_t)gridDim.x * blockDim.x) {
        if (j > 0 && key[j] != ~0ULL && key[j] == key[j - 1]) keep[idx[j]] = 0;
    }
}

// number of repeated pairs in the sorted pair keys
__global__ void k_count_repeats(const uint64_t* __restrict__ key, int64_t L, uint32_t* __restrict__ reps) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < L; j += (int64_t)gridDim.x * blockDim.x)
        if (j > 0 && key[j] != ~0ULL && key[j] == key[j - 1]) atomicAdd(reps, 1u);
}

// A caller's CSR, checked after k_csr_owner has passed its ids: the directed
// pairs (v, u) sorted, and the transposed pairs (u, v) sorted, must be the same
// sequence (a symmetric adjacency) with no pair twice (no repeated neighbour).
__global__ void k_dir_keys(const int32_t* __restrict__ src, const int32_t* __restrict__ col, int64_t nnz,
                           uint64_t* __restrict__ kf, uint64_t* __restrict__ kt) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t v = (uint32_t)src[j], u = (uint32_t)col[j];
        kf[j] = v << 32 | u;
        kt[j] = u << 32 | v;
    }
}
__global__ void k_csr_sym_check(const uint64_t* __restrict__ kf, const uint64_t* __restrict__ kt, int64_t nnz,
                                uint32_t* __restrict__ bad) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * blockDim.x) {
        if (kf[j] != kt[j]) atomicOr(bad, 2u);
        if (j > 0 && kf[j] == kf[j - 1]) atomicOr(bad, 4u);
    }
}

// line i (kept, compacted index o) -> slots 2o (src->dst) and 2o+1 (dst->src)
__global__ void k_make_slots(const int32_t* __restrict__ s, const int32_t* __restrict__ d,
                             const int32_t* __restrict__ keep, const int64_t* __restrict__ off, int64_t L,
                             uint32_t* __restrict__ owner, int32_t* __restrict__ nb) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x) {
        if (!keep[i]) continue;
        int64_t o = off[i];
        int32_t a = s[i], b = d[i];
        owner[2 * o] = (uint32_t)a;
        nb[2 * o] = b;
        owner[2 * o + 1] = (uint32_t)b;
        nb[2 * o + 1] = a;
    }
}

// CSR offsets from the slots' owners sorted ascending: row[v] = the first
// slot whose owner is >= v (row[n] = nnz), by binary search; no atomics, so
// a hub's many slots do not all update one degree counter
__global__ void k_row_from_owners(const int32_t* __restrict__ owner, int64_t nnz, int64_t n,
                                  int64_t* __restrict__ row) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v <= n; v += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = nnz;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (owner[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        row[v] = lo;
    }
}
__global__ void k_deg_from_row(const int64_t* __restrict__ row, int64_t n, int32_t* __restrict__ deg) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        deg[v] = (int32_t)(row[v + 1] - row[v]);
}

// owner of every CSR slot and the degrees, from the offsets (one wavefront per vertex)
__global__ void k_csr_owner(const int64_t* __restrict__ row, const int32_t* __restrict__ col, int64_t n,
                            int32_t* __restrict__ src, int32_t* __restrict__ deg, uint32_t* __restrict__ bad) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v = wave; v < n; v += nw) {
        const int64_t b = row[v], e = row[v + 1];
        if (lane == 0) {
            deg[v] = (int32_t)(e - b);
            if (e < b) atomicOr(bad, 1u);
        }
        for (int64_t j = b + lane; j < e; j += 64) {
            src[j] = (int32_t)v;
            int32_t u = col[j];
            if (u < 0 || u >= n || u == v) atomicOr(bad, 1u);
        }
    }
}

// ===========================================================================
// preprocessing kernels
// ===========================================================================
// neighbour-order key of vertex x (order_vertex_neighbors, graphlet_core.cpp:1910-1953)
enum NeighKey { NK_NONE = 0, NK_DEG, NK_KCORE, NK_DEG_KCORE, NK_DEG_VOL, NK_KCORE_VOL, NK_DEG_KCORE_VOL };

// the reference's kcore[] is 1-based: kcore[x] holds the core of vertex x-1
__device__ __forceinline__ int32_t kc1(const int32_t* core, int64_t x) { return x >= 1 ? core[x - 1] : 0; }

__global__ void k_vertex_feature(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                                 const int32_t* __restrict__ deg, const int32_t* __restrict__ core, int64_t n,
                                 int kind, int32_t* __restrict__ feat) {
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        uint32_t f = 0;
        switch (kind) {
            case NK_DEG: f = (uint32_t)deg[x]; break;
            case NK_KCORE: f = (uint32_t)kc1(core, x); break;
            case NK_DEG_KCORE: f = (uint32_t)deg[x] + (uint32_t)kc1(core, x); break;
            default:
                for (int64_t j = row[x]; j < row[x + 1]; j++) {
                    int32_t y = col[j];
                    if (kind == NK_DEG_VOL) f += (uint32_t)deg[y];
                    else if (kind == NK_KCORE_VOL) f += (uint32_t)kc1(core, y);
                    else f += (uint32_t)deg[y] + (uint32_t)kc1(core, y);
                }
        }
        feat[x] = (int32_t)f;
    }
}

// composite key (owner, key) for the stable per-vertex sort; the large->small
// order reverses the input first so equal keys end in reverse input order
// (bucket_sort_neighbors_parallel, graphlet_core.cpp:1846-1877)
// The feature enters the key as its offset from the smallest (small->large)
// or from the largest (large->small) feature, in fb bits, so the sort runs
// over fb + log2 n bits only.
__global__ void k_neigh_keys(const int32_t* __restrict__ src, const int32_t* __restrict__ col,
                             const int32_t* __restrict__ feat, int64_t nnz, int s2l, const int32_t* __restrict__ mm,
                             int fb, uint64_t* __restrict__ key, int32_t* __restrict__ val) {
    const int64_t fmin = mm[0], fmax = mm[1];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t p = s2l ? i : nnz - 1 - i;
        int32_t x = col[p];
        const uint64_t k = s2l ? (uint64_t)((int64_t)feat[x] - fmin) : (uint64_t)(fmax - (int64_t)feat[x]);
        key[i] = ((uint64_t)(uint32_t)src[p] << fb) | k;
        val[i] = x;
    }
}

// smallest and largest feature (mm[0], mm[1]; mm set to INT32_MAX, INT32_MIN)
__global__ void k_feat_minmax(const int32_t* __restrict__ feat, int64_t n, int32_t* __restrict__ mm) {
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        lo = min(lo, feat[x]);
        hi = max(hi, feat[x]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off, 64));
        hi = max(hi, __shfl_xor(hi, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}
__global__ void k_minmax_init(int32_t* mm) {
    if (threadIdx.x == 0) {
        mm[0] = INT32_MAX;
        mm[1] = INT32_MIN;
    }
}

__global__ void k_edge_flags(const int32_t* __restrict__ src, const int32_t* __restrict__ col, int64_t nnz,
                             int32_t* __restrict__ f) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x)
        f[p] = src[p] < col[p];
}

// create_edge_list_arrays, graphlet_core.cpp:929-949: edge id = scan order of
// (v<u) slots; e_v is the strictly lower-degree endpoint, ties go to u.
__global__ void k_edge_arrays(const int32_t* __restrict__ src, const int32_t* __restrict__ col,
                              const int32_t* __restrict__ deg, const int64_t* __restrict__ eid, int64_t nnz,
                              int32_t* __restrict__ ev, int32_t* __restrict__ eu) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        int32_t v = src[p], u = col[p];
        if (v < u) {
            int64_t e = eid[p];
            if (deg[v] < deg[u]) { ev[e] = v; eu[e] = u; }
            else { ev[e] = u; eu[e] = v; }
        }
    }
}

// ---- k-core: level-synchronous peel -------------------------------------
__global__ void k_core_init(const int32_t* __restrict__ deg, int64_t n, int32_t* __restrict__ cur,
                            uint8_t* __restrict__ alive) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        cur[v] = deg[v];
        alive[v] = 1;
    }
}

__global__ void k_core_select(int32_t k, int64_t n, const int32_t* __restrict__ cur, uint8_t* __restrict__ alive,
                              int32_t* __restrict__ core, int32_t* __restrict__ frontier, int32_t* __restrict__ stat) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        if (!alive[v]) continue;
        int32_t c = cur[v];
        if (c <= k) {
            alive[v] = 0;
            core[v] = k;
            frontier[atomicAdd(&stat[0], 1)] = (int32_t)v;
        } else {
            atomicMin(&stat[1], c);
        }
    }
}

// one wavefront per removed vertex
__global__ void k_core_update(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                              const int32_t* __restrict__ frontier, int32_t cnt, const uint8_t* __restrict__ alive,
                              int32_t* __restrict__ cur) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t i = wave; i < cnt; i += nw) {
        int32_t v = frontier[i];
        for (int64_t j = row[v] + lane; j < row[v + 1]; j += 64) {
            int32_t u = col[j];
            if (alive[u]) atomicSub(&cur[u], 1);
        }
    }
}

// ---- edge order keys (get_ordering_value, graphlet_core.cpp:773-834) -------
enum OrderKind { OK_NATURAL = 0, OK_DEG, OK_KCORE, OK_KCORE_DEG, OK_DEG_VOL, OK_KCORE_VOL, OK_DEG_KCORE_VOL, OK_DEFAULT };

// Keys of the positions p < cnt of the edge list: edge p itself (sort_edges
// over all edges, graphlet_core.cpp:854-867) or, with ids, the edge ids[p] of
// a sampled set E_s (837-851), whose natural key is the position p.
__global__ void k_order_keys(const int32_t* __restrict__ ev, const int32_t* __restrict__ eu,
                             const int32_t* __restrict__ deg, const int32_t* __restrict__ core,
                             const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                             const int32_t* __restrict__ ids, int64_t cnt, int kind, int s2l,
                             uint64_t* __restrict__ key) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < cnt; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e = ids ? ids[p] : (int32_t)p;
        int32_t v = ev[e], u = eu[e];
        long long val = p;
        switch (kind) {
            case OK_NATURAL: break;
            case OK_DEG: val = (long long)deg[u] + deg[v]; break;
            case OK_KCORE: val = (long long)kc1(core, u) + kc1(core, v); break;
            case OK_KCORE_DEG:
                val = (int32_t)((uint32_t)(deg[u] + deg[v]) * (uint32_t)(kc1(core, u) + kc1(core, v)));
                break;
            case OK_DEG_VOL:
            case OK_KCORE_VOL:
            case OK_DEG_KCORE_VOL: {
                val = 0;
                int32_t ends[2] = {u, v};
                for (int t = 0; t < 2; t++)
                    for (int64_t j = row[ends[t]]; j < row[ends[t] + 1]; j++) {
                        int32_t y = col[j];
                        if (kind == OK_DEG_VOL) val += deg[y];
                        else if (kind == OK_KCORE_VOL) val += kc1(core, y);
                        else val += (long long)deg[y] * kc1(core, y);
                    }
                break;
            }
            default: val = deg[u]; break;
        }
        uint32_t k = (uint32_t)(int32_t)val ^ 0x80000000u;  // Vertex keeps an int bound
        if (!s2l) k = ~k;                                   // decr_bound: larger bound first
        key[p] = ((uint64_t)k << 32) | (uint32_t)~(uint32_t)e;  // ties: larger id first
    }
}

__global__ void k_order_from_keys(const uint64_t* __restrict__ key, int64_t m, int32_t* __restrict__ order) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x)
        order[e] = (int32_t)~(uint32_t)(key[e] & 0xffffffffULL);
}

// ===========================================================================
// epilogue: per-edge equations (solve_graphlet_equations,
// graphlet_core.cpp:1325-1364), micro columns (1559-1575) and the reduction
// of the 15 sums (1429-1448) in uint64 wrap-around arithmetic.
// ===========================================================================
__global__ __launch_bounds__(256) void k_epilogue(const int32_t* __restrict__ order, const int32_t* __restrict__ ev,
                                                  const int32_t* __restrict__ eu, const int32_t* __restrict__ deg,
                                                  const uint32_t* __restrict__ Tc, const uint64_t* __restrict__ K4c,
                                                  const uint64_t* __restrict__ C4c, int64_t n, int64_t m,
                                                  int64_t nord, uint64_t* __restrict__ sums,
                                                  uint64_t* __restrict__ micro) {
    uint64_t acc[PGD_NSUMS];
#pragma unroll
    for (int k = 0; k < PGD_NSUMS; k++) acc[k] = 0;
    // the loop runs over the E_ordered positions (all m edges, a sampled set
    // E_s, or any list of edge ids); m stays the graph's edge count in the
    // equations.  With micro, nord <= m.
    const int64_t np = nord > m ? nord : m;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < np; e += (int64_t)gridDim.x * blockDim.x) {
        if (e >= nord) {
            // positions past |E_ordered| keep the 0 of the reference's
            // freshly resized per-edge vectors
            if (micro)
                for (int k = 0; k < PGD_NMICRO; k++) micro[(size_t)k * m + e] = 0;
            continue;
        }
        int32_t id = order[e];
        int32_t v = ev[id], u = eu[id];
        edge_terms(deg[v], deg[u], Tc[id], K4c[id], C4c[id], n, m, acc, micro, m, e);
    }
    block_add_sums(acc, sums);
}

__global__ void k_zero_sums(uint64_t* __restrict__ sums) {
    if (threadIdx.x < PGD_NSUMS) sums[threadIdx.x] = 0;
}

}  // namespace

namespace pgd {
void epilogue(pgd_ctx* ctx, const GraphView& g, EdgeCounts c, uint64_t* d_sums, uint64_t* d_micro) {
    // the sums start at zero (a fill kernel: the allocation-free entry point
    // hands in a caller's buffer)
    hipLaunchKernelGGL(k_zero_sums, dim3(1), dim3(64), 0, ctx->stream, d_sums);
    PGD_HIP(hipGetLastError());
    if (g.m == 0 && g.nord == 0) return;
    int grid = std::min(grid_for(std::max(g.m, g.nord)), ctx->num_cus * 8);
    hipLaunchKernelGGL(k_epilogue, dim3(grid), dim3(256), 0, ctx->stream, g.order, g.ev, g.eu, g.deg, c.T, c.K4,
                       c.C4, g.n, g.m, g.nord, d_sums, d_micro);
    PGD_HIP(hipGetLastError());
}
}  // namespace pgd

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

template <class F>
pgd_status guarded(pgd_ctx* ctx, F&& f) {
    try {
        f();
        return PGD_OK;
    } catch (const Error& e) {
        if (ctx) ctx->last_err = e.what;
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->last_err = "host allocation failed";
        return PGD_ERR_OOM;
    } catch (...) {
        if (ctx) ctx->last_err = "unexpected exception";
        return PGD_ERR_INVALID;
    }
}

void sync(pgd_ctx* ctx) { PGD_HIP(hipStreamSynchronize(ctx->stream)); }

int neigh_kind(const char* m) {
    std::string s(m ? m : "");
    if (s == "deg" || s == "degree") return NK_DEG;
    if (s == "kcore") return NK_KCORE;
    if (s == "degree_kcore" || s == "deg_kcore") return NK_DEG_KCORE;
    if (s == "degree_vol" || s == "deg_vol") return NK_DEG_VOL;
    if (s == "kcore_vol") return NK_KCORE_VOL;
    if (s == "degree_kcore_vol" || s == "deg_kcore_vol") return NK_DEG_KCORE_VOL;
    return NK_NONE;
}

// get_ordering_enum (graphlet_core.cpp:713-763); -1: rand/var (unsupported)
int order_kind(const char* o) {
    std::string s(o ? o : "");
    if (s == "natural") return OK_NATURAL;
    if (s == "rand" || s == "random" || s == "var") return -1;
    if (s == "deg" || s == "degree") return OK_DEG;
    if (s == "kcore" || s == "kcores") return OK_KCORE;
    if (s == "kcore_deg" || s == "kcore_degree") return OK_KCORE_DEG;
    if (s == "deg_vol" || s == "degree_vol") return OK_DEG_VOL;
    if (s == "kcore_vol" || s == "kcore_volume") return OK_KCORE_VOL;
    if (s == "kcore_deg_vol" || s == "kcore_degree_vol") return OK_DEG_KCORE_VOL;
    static const char* rest[] = {
        "triangles", "tri", "triangle", "wedges", "tcores", "triangle_cores", "triangle_core", "tcore", "coloring",
        "greedy_coloring", "tri_vol", "triangle_vol", "tcore_vol", "tcore_volume", "triangle_core_vol",
        "triangles_only", "triangle_only", "dynamic_largest_first", "lfo", "LFO", "triangle_core_max", "tcore_max",
        "tcores_max", "degree_triangles", "degree_tri", "deg_tri", "deg_triangles", "kcore_tri", "kcore_triangles",
        "kcore_triangle", "kcore_deg_tri", "kcore_degree_triangle", "kcore_degree_tri", "kcore_tri_vol",
        "kcore_triangle_vol", "kcore_deg_tri_vol", "kcore_degree_tri_vol", "dist_two_lfo", "dist_two_largest_first",
        "dist_two_dlfo", "dist_two_dynamic_lfo", "dist_two_slo", "dist_two_smallest_last", "ido", "incidence",
        "incidence_deg", "dist_two_ido", "dist_two_incidence", "dist_two_tcore", "dist_two_triangle_core",
        "tcore_min", "triangle_core_min", "tri_core_min", "common", "common_neighbors"};
    for (const char* r : rest)
        if (s == r) return OK_DEFAULT;
    return OK_NATURAL;  // unknown names fall back to natural order
}

void compute_cores(pgd_graph* g) {
    if (g->have_core) return;
    pgd_ctx* ctx = g->ctx;
    const int64_t n = g->n;
    g->core.alloc(ctx, (size_t)std::max<int64_t>(n, 1));
    g->K = 0;
    if (n == 0) { g->have_core = true; return; }
    DBuf<int32_t> cur, frontier, stat;
    DBuf<uint8_t> alive;
    cur.alloc(ctx, n);
    frontier.alloc(ctx, n);
    stat.alloc(ctx, 2);
    alive.alloc(ctx, n);
    int grid = std::min(grid_for(n), ctx->num_cus * 16);
    hipLaunchKernelGGL(k_core_init, dim3(grid), dim3(kBlock), 0, ctx->stream, g->deg.get(), n, cur.get(),
                       alive.get());
    PGD_HIP(hipGetLastError());
    int64_t remaining = n;
    int32_t k = 0;
    int32_t h[2];
    while (remaining > 0) {
        h[0] = 0;
        h[1] = INT_MAX;
        PGD_HIP(hipMemcpyAsync(stat.get(), h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_core_select, dim3(grid), dim3(kBlock), 0, ctx->stream, k, n, cur.get(), alive.get(),
                           g->core.get(), frontier.get(), stat.get());
        PGD_HIP(hipGetLastError());
        PGD_HIP(hipMemcpyAsync(h, stat.get(), sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
        if (h[0] == 0) {
            k = std::max(k + 1, h[1]);
            continue;
        }
        int ugrid = std::min(grid_for((int64_t)h[0] * 64), ctx->num_cus * 16);
        hipLaunchKernelGGL(k_core_update, dim3(ugrid), dim3(kBlock), 0, ctx->stream, g->row.get(), g->col.get(),
                           frontier.get(), h[0], alive.get(), cur.get());
        PGD_HIP(hipGetLastError());
        remaining -= h[0];
    }
    // printed K: maximum core over vertices 0..n-2 (graphlet_core.cpp:1011-1015)
    if (n >= 2) {
        DBuf<int32_t> mx;
        mx.alloc(ctx, 1);
        reduce_max(ctx, g->core.get(), mx.get(), (size_t)(n - 1));
        PGD_HIP(hipMemcpyAsync(&g->K, mx.get(), sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
    }
    g->have_core = true;
}

}  // namespace

extern "C" {

const char* pgd_status_str(pgd_status s) {
    switch (s) {
        case PGD_OK: return "ok";
        case PGD_ERR_INVALID: return "invalid argument";
        case PGD_ERR_NO_DEVICE: return "no HIP device";
        case PGD_ERR_HIP: return "HIP runtime error";
        case PGD_ERR_OOM: return "out of device memory";
        case PGD_ERR_UNSUPPORTED: return "unsupported input";
        case PGD_ERR_STATE: return "call out of order";
        case PGD_ERR_BUSY: return "busy";
        default: return "unknown status";
    }
}

pgd_status pgd_open(int32_t device, pgd_ctx** out) {
    if (!out) return PGD_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return PGD_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) return PGD_ERR_NO_DEVICE;
    pgd_ctx* ctx = new (std::nothrow) pgd_ctx();
    if (!ctx) return PGD_ERR_OOM;
    ctx->device = device;
    pgd_status st = guarded(ctx, [&] {
        PGD_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        PGD_HIP(hipGetDeviceProperties(&prop, device));
        ctx->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        ctx->lds_per_block = prop.sharedMemPerBlock;
        PGD_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        // keep released blocks in the device's pool instead of returning them
        // to the driver at every synchronisation
        hipMemPool_t pool = nullptr;
        PGD_HIP(hipDeviceGetDefaultMemPool(&pool, device));
        uint64_t keep = UINT64_MAX;
        PGD_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
        noalloc_scratch_init(device);
    });
    if (st != PGD_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return PGD_OK;
}

pgd_status pgd_close(pgd_ctx* ctx) {
    if (!ctx) return PGD_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    if (ctx->tmp) dfree(ctx, ctx->tmp, ctx->tmp_bytes);
    release_cached(ctx);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    if (ctx->aux) {
        (void)hipStreamSynchronize(ctx->aux);
        (void)hipStreamDestroy(ctx->aux);
    }
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamDestroy(ctx->side);
    }
    delete ctx;
    return PGD_OK;
}

pgd_status pgd_set_mem_cap(pgd_ctx* ctx, int64_t bytes) {
    if (!ctx || bytes < 0) return PGD_ERR_INVALID;
    ctx->mem_cap = bytes;
    return PGD_OK;
}

namespace {
__global__ void k_atomic_probe(uint32_t* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&p[i], 1u);
}
__global__ void k_atomic_check(const uint32_t* __restrict__ p, int64_t n, uint32_t want, uint32_t* __restrict__ bad) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (p[i] != want) atomicOr(bad, 1u);
}
}  // namespace

pgd_status pgd_atomic_rate(pgd_ctx* ctx, int64_t words, int32_t reps, double* requests_per_s, double* ms_per_launch) {
    if (!ctx || !requests_per_s || words < ((int64_t)1 << 20) || words >= ((int64_t)1 << 31) || reps < 1 ||
        reps > 100)
        return PGD_ERR_INVALID;
    return guarded(ctx, [&] {
        PGD_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DBuf<uint32_t> cnt, bad;
        cnt.alloc(ctx, words);
        bad.alloc(ctx, 1);
        PGD_HIP(hipMemsetAsync(cnt.get(), 0, (size_t)words * 4, s));
        PGD_HIP(hipMemsetAsync(bad.get(), 0, 4, s));
        const int grid = std::min(grid_for(words), ctx->num_cus * 16);
        // one untimed launch first: the first launch of a kernel of the
        // library also loads its code
        hipLaunchKernelGGL(k_atomic_probe, dim3(grid), dim3(kBlock), 0, s, cnt.get(), words);
        PGD_HIP(hipGetLastError());
        hipEvent_t e0, e1;
        PGD_HIP(hipEventCreate(&e0));
        PGD_HIP(hipEventCreate(&e1));
        float ms = 0;
        hipError_t err = hipEventRecord(e0, s);
        for (int r = 0; r < reps && err == hipSuccess; r++) {
            hipLaunchKernelGGL(k_atomic_probe, dim3(grid), dim3(kBlock), 0, s, cnt.get(), words);
            err = hipGetLastError();
        }
        if (err == hipSuccess) err = hipEventRecord(e1, s);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        PGD_HIP(err);
        hipLaunchKernelGGL(k_atomic_check, dim3(grid), dim3(kBlock), 0, s, cnt.get(), words, (uint32_t)reps + 1u,
                           bad.get());
        PGD_HIP(hipGetLastError());
        uint32_t hbad = 0;
        PGD_HIP(hipMemcpyAsync(&hbad, bad.get(), 4, hipMemcpyDeviceToHost, s));
        sync(ctx);
        PGD_CHECK(hbad == 0, PGD_ERR_STATE, "atomic probe: a counter does not hold the number of launches (timed ones and the warm-up)");
        *requests_per_s = (double)reps * (double)words / 16.0 / ((double)ms / 1e3);
        if (ms_per_launch) *ms_per_launch = (double)ms / reps;
    });
}

pgd_status pgd_mem_in_use(pgd_ctx* ctx, int64_t* bytes) {
    if (!ctx || !bytes) return PGD_ERR_INVALID;
    *bytes = ctx->mem_used.load();
    return PGD_OK;
}

const char* pgd_last_error(pgd_ctx* ctx) { return ctx ? ctx->last_err.c_str() : "no context"; }

namespace {
// The CSR of a graph from its COO in device memory (ds, dd: L lines in file
// order, ids in [0, n); with drop_neg, lines with a negative first id are
// skipped): self loops dropped, repeated pairs dropped keeping the first
// (PGD_DEDUP) or refused, each adjacency list in line order, degrees.
void coo_build(pgd_ctx* ctx, pgd_graph* g, int64_t n, int64_t L, DBuf<int32_t>& ds, DBuf<int32_t>& dd,
               uint32_t flags, bool drop_neg) {
        PGD_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        DBuf<int32_t> keep;
        DBuf<int64_t> off;
        DBuf<uint32_t> bad;
        keep.alloc(ctx, L);
        off.alloc(ctx, L + 1);
        bad.alloc(ctx, 1);
        PGD_HIP(hipMemsetAsync(bad.get(), 0, sizeof(uint32_t), s));
        int grid = std::min(grid_for(L), ctx->num_cus * 16);
        if (L) {
            hipLaunchKernelGGL(k_line_keep, dim3(grid), dim3(kBlock), 0, s, ds.get(), dd.get(), L, n,
                               drop_neg ? 1 : 0, keep.get(), bad.get());
            PGD_HIP(hipGetLastError());
        }
        uint32_t hbad = 0;
        PGD_HIP(hipMemcpyAsync(&hbad, bad.get(), sizeof(hbad), hipMemcpyDeviceToHost, s));
        sync(ctx);
        PGD_CHECK(hbad == 0, PGD_ERR_INVALID, "vertex id outside [0, n)");
        if (L) {
            // sort the undirected pairs (stable by line): with PGD_DEDUP later copies
            // are dropped; without it (.mtx) a repeated pair is a multi-edge, which
            // the reference counts with multiplicity and this build rejects
            DBuf<uint64_t> k0, k1;
            DBuf<int32_t> i0, i1;
            k0.alloc(ctx, L); k1.alloc(ctx, L); i0.alloc(ctx, L); i1.alloc(ctx, L);
            hipLaunchKernelGGL(k_pair_keys, dim3(grid), dim3(kBlock), 0, s, ds.get(), dd.get(), keep.get(), L,
                               k0.get(), i0.get());
            PGD_HIP(hipGetLastError());
            sort_pairs(ctx, k0.get(), k1.get(), i0.get(), i1.get(), (size_t)L, 0, 64);
            if (flags & PGD_DEDUP) {
                hipLaunchKernelGGL(k_dedup_mark, dim3(grid), dim3(kBlock), 0, s, k1.get(), i1.get(), L, keep.get());
            } else {
                hipLaunchKernelGGL(k_count_repeats, dim3(grid), dim3(kBlock), 0, s, k1.get(), L, bad.get());
            }
            PGD_HIP(hipGetLastError());
            uint32_t reps = 0;
            PGD_HIP(hipMemcpyAsync(&reps, bad.get(), sizeof(reps), hipMemcpyDeviceToHost, s));
            sync(ctx);
            PGD_CHECK(reps == 0, PGD_ERR_UNSUPPORTED,
                      "the edge list repeats an undirected pair (a multi-edge); only simple graphs are supported");
        }
        exclusive_sum(ctx, keep.get(), off.get(), (size_t)L);
        int64_t kept = 0;
        if (L) {
            int64_t last_off = 0;
            int32_t last_keep = 0;
            PGD_HIP(hipMemcpyAsync(&last_off, off.get() + (L - 1), sizeof(int64_t), hipMemcpyDeviceToHost, s));
            PGD_HIP(hipMemcpyAsync(&last_keep, keep.get() + (L - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
            sync(ctx);
            kept = last_off + last_keep;
        }
        const int64_t nnz = 2 * kept;
        g->n = n;
        g->nnz = nnz;
        g->m = nnz / 2;
        g->row.alloc(ctx, n + 1);
        g->col.alloc(ctx, nnz);
        g->src.alloc(ctx, nnz);
        g->deg.alloc(ctx, std::max<int64_t>(n, 1));
        if (nnz) {
            DBuf<uint32_t> own;
            DBuf<int32_t> nb;
            own.alloc(ctx, nnz);
            nb.alloc(ctx, nnz);
            hipLaunchKernelGGL(k_make_slots, dim3(grid), dim3(kBlock), 0, s, ds.get(), dd.get(), keep.get(),
                               off.get(), L, own.get(), nb.get());
            PGD_HIP(hipGetLastError());
            // stable by owner: every adjacency list stays in line order
            sort_pairs(ctx, own.get(), (uint32_t*)g->src.get(), nb.get(), g->col.get(), (size_t)nnz, 0,
                       (unsigned)bits_for(n - 1));
        }
        {
            // owners ascending (src, n <= INT32_MAX): offsets by search, degrees from them
            const int gv = std::min(grid_for(n + 1), ctx->num_cus * 16);
            hipLaunchKernelGGL(k_row_from_owners, dim3(gv), dim3(kBlock), 0, s, g->src.get(), nnz, n, g->row.get());
            if (n) hipLaunchKernelGGL(k_deg_from_row, dim3(gv), dim3(kBlock), 0, s, g->row.get(), n, g->deg.get());
            PGD_HIP(hipGetLastError());
        }
        int32_t md = 0;
        if (n) {
            DBuf<int32_t> mx;
            mx.alloc(ctx, 1);
            reduce_max(ctx, g->deg.get(), mx.get(), (size_t)n);
            PGD_HIP(hipMemcpyAsync(&md, mx.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
        }
        sync(ctx);
        // vertex_degrees: max_degree starts from vertex 0's degree; avg is an int
        g->max_deg = md;
        g->avg_deg = n > 0 ? (int32_t)((double)nnz / (double)n) : 0;
}
}  // namespace

extern "C++" namespace pgd {
void graph_from_device_coo(pgd_ctx* ctx, pgd_graph* g, int64_t n, int64_t L, DBuf<int32_t>& ds,
                           DBuf<int32_t>& dd, uint32_t flags) {
    coo_build(ctx, g, n, L, ds, dd, flags, false);
}
}  // namespace pgd

pgd_status pgd_graph_from_coo(pgd_ctx* ctx, int64_t n, int64_t L, const int32_t* src, const int32_t* dst,
                              uint32_t flags, pgd_graph** out) {
    if (!ctx || !out || n < 0 || L < 0 || (L > 0 && (!src || !dst))) return PGD_ERR_INVALID;
    if (n >= INT32_MAX || L >= ((int64_t)1 << 31)) return PGD_ERR_UNSUPPORTED;
    *out = nullptr;
    pgd_graph* g = new (std::nothrow) pgd_graph();
    if (!g) return PGD_ERR_OOM;
    g->ctx = ctx;
    pgd_status st = guarded(ctx, [&] {
        PGD_HIP(hipSetDevice(ctx->device));
        DBuf<int32_t> ds, dd;
        ds.alloc(ctx, L);
        dd.alloc(ctx, L);
        if (L) {
            PGD_HIP(hipMemcpyAsync(ds.get(), src, L * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
            PGD_HIP(hipMemcpyAsync(dd.get(), dst, L * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        }
        coo_build(ctx, g, n, L, ds, dd, flags, false);
    });
    if (st != PGD_OK) {
        delete g;
        return st;
    }
    *out = g;
    return PGD_OK;
}

pgd_status pgd_graph_from_file(pgd_ctx* ctx, const char* path, int32_t threads, pgd_graph** out,
                               int64_t* self_loops) {
    if (!ctx || !out || !path) return PGD_ERR_INVALID;
    *out = nullptr;
    RawLines R;
    pgd_status st = read_raw_lines(path, threads, R);
    if (st != PGD_OK) {
        ctx->last_err = std::string("cannot read ") + path;
        return st;
    }
    const int64_t L = (int64_t)R.v.size();
    if (L >= ((int64_t)1 << 31)) return PGD_ERR_UNSUPPORTED;
    pgd_graph* g = new (std::nothrow) pgd_graph();
    if (!g) return PGD_ERR_OOM;
    g->ctx = ctx;
    int64_t loops = R.loops;
    st = guarded(ctx, [&] {
        PGD_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        const uint32_t flags = R.mtx ? 0u : PGD_DEDUP;
        DBuf<int32_t> ds, dd;
        ds.alloc(ctx, L);
        dd.alloc(ctx, L);
        if (L) {
            PGD_HIP(hipMemcpyAsync(ds.get(), R.v.data(), L * sizeof(int32_t), hipMemcpyHostToDevice, s));
            PGD_HIP(hipMemcpyAsync(dd.get(), R.u.data(), L * sizeof(int32_t), hipMemcpyHostToDevice, s));
        }
        if (R.final_ids) {  // replayed on the host: final ids, self loops removed
            PGD_CHECK(R.n < INT32_MAX, PGD_ERR_UNSUPPORTED, "more than 2^31 - 1 vertices");
            coo_build(ctx, g, R.n, L, ds, dd, flags, false);
            return;
        }
        const int grid = std::min(grid_for(L), ctx->num_cus * 16);
        if (R.mtx) {
            // n = 1 + the largest endpoint of a non-loop line; a negative id
            // is an error (the reference never terminates on it)
            int64_t maxid = -1;
            bool neg = false;
            for (int64_t i = 0; i < L; i++) {
                const int32_t a = R.v[(size_t)i], b = R.u[(size_t)i];
                if (a == b) { loops++; continue; }
                if (a < 0 || b < 0) neg = true;
                maxid = std::max<int64_t>(maxid, std::max(a, b));
            }
            PGD_CHECK(!neg, PGD_ERR_INVALID, "negative vertex id");
            PGD_CHECK(maxid + 1 < INT32_MAX, PGD_ERR_UNSUPPORTED, "more than 2^31 - 1 vertices");
            coo_build(ctx, g, maxid + 1, L, ds, dd, flags, false);
            return;
        }
        // edge list: renumbering by first appearance on the device
        const int64_t T = R.maxv + 1;
        DBuf<uint32_t> first, mark, rank;
        DBuf<unsigned long long> stat;
        first.alloc(ctx, T);
        stat.alloc(ctx, 2);
        PGD_HIP(hipMemsetAsync(stat.get(), 0, 2 * sizeof(unsigned long long), s));
        hipLaunchKernelGGL(k_fill_u32_max, dim3(std::min(grid_for(T), ctx->num_cus * 16)), dim3(kBlock), 0, s,
                           first.get(), T);
        unsigned long long hs[2] = {0, 0};
        if (L) {
            hipLaunchKernelGGL(k_first_pos, dim3(grid), dim3(kBlock), 0, s, ds.get(), dd.get(), L, R.maxv,
                               first.get(), stat.get());
            PGD_HIP(hipGetLastError());
        }
        PGD_HIP(hipMemcpyAsync(hs, stat.get(), sizeof(hs), hipMemcpyDeviceToHost, s));
        sync(ctx);
        PGD_CHECK(hs[1] == 0, PGD_ERR_INVALID, "vertex id outside the reader's table");
        loops = (int64_t)hs[0];
        int64_t nv = 0;
        if (L) {
            mark.alloc(ctx, 2 * L);
            rank.alloc(ctx, 2 * L);
            hipLaunchKernelGGL(k_first_mark, dim3(std::min(grid_for(2 * L), ctx->num_cus * 16)), dim3(kBlock), 0,
                               s, ds.get(), dd.get(), L, first.get(), mark.get());
            PGD_HIP(hipGetLastError());
            exclusive_sum(ctx, mark.get(), rank.get(), (size_t)(2 * L));
            uint32_t tail[2] = {0, 0};
            PGD_HIP(hipMemcpyAsync(&tail[0], rank.get() + (2 * L - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            PGD_HIP(hipMemcpyAsync(&tail[1], mark.get() + (2 * L - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            hipLaunchKernelGGL(k_renumber, dim3(grid), dim3(kBlock), 0, s, ds.get(), dd.get(), L, first.get(),
                               rank.get());
            PGD_HIP(hipGetLastError());
            sync(ctx);
            nv = (int64_t)tail[0] + tail[1];
        }
        coo_build(ctx, g, nv, L, ds, dd, flags, true);
    });
    if (st != PGD_OK) {
        delete g;
        return st;
    }
    if (self_loops) *self_loops = loops;
    *out = g;
    return PGD_OK;
}

pgd_status pgd_graph_from_csr(pgd_ctx* ctx, int64_t n, const int64_t* row, const int32_t* col, pgd_graph** out) {
    if (!ctx || !out || n < 0 || !row) return PGD_ERR_INVALID;
    if (n >= INT32_MAX) return PGD_ERR_UNSUPPORTED;
    const int64_t nnz = row[n];
    if (row[0] != 0 || nnz < 0 || (nnz & 1) || (nnz > 0 && !col)) return PGD_ERR_INVALID;
    // offsets non-decreasing (so every list lies inside [0, nnz)) before any upload
    for (int64_t v = 0; v < n; v++)
        if (row[v + 1] < row[v]) {
            ctx->last_err = "CSR offsets decrease";
            return PGD_ERR_INVALID;
        }
    *out = nullptr;
    pgd_graph* g = new (std::nothrow) pgd_graph();
    if (!g) return PGD_ERR_OOM;
    g->ctx = ctx;
    pgd_status st = guarded(ctx, [&] {
        PGD_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        g->n = n;
        g->nnz = nnz;
        g->m = nnz / 2;
        g->row.alloc(ctx, n + 1);
        g->col.alloc(ctx, nnz);
        g->src.alloc(ctx, nnz);
        g->deg.alloc(ctx, std::max<int64_t>(n, 1));
        DBuf<uint32_t> bad;
        bad.alloc(ctx, 1);
        PGD_HIP(hipMemsetAsync(bad.get(), 0, sizeof(uint32_t), s));
        PGD_HIP(hipMemcpyAsync(g->row.get(), row, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
        if (nnz) PGD_HIP(hipMemcpyAsync(g->col.get(), col, nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
        int32_t md = 0;
        if (n) {
            int grid = std::min(grid_for(n * 64), ctx->num_cus * 16);
            hipLaunchKernelGGL(k_csr_owner, dim3(grid), dim3(kBlock), 0, s, g->row.get(), g->col.get(), n,
                               g->src.get(), g->deg.get(), bad.get());
            PGD_HIP(hipGetLastError());
            DBuf<int32_t> mx;
            mx.alloc(ctx, 1);
            reduce_max(ctx, g->deg.get(), mx.get(), (size_t)n);
            PGD_HIP(hipMemcpyAsync(&md, mx.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s));
        }
        uint32_t hbad = 0;
        PGD_HIP(hipMemcpyAsync(&hbad, bad.get(), sizeof(hbad), hipMemcpyDeviceToHost, s));
        sync(ctx);
        PGD_CHECK(hbad == 0, PGD_ERR_INVALID, "CSR with an id outside [0, n) or a self loop");
        if (nnz) {
            DBuf<uint64_t> kf, kt, sf, st;
            kf.alloc(ctx, nnz); kt.alloc(ctx, nnz); sf.alloc(ctx, nnz); st.alloc(ctx, nnz);
            int grid = std::min(grid_for(nnz), ctx->num_cus * 16);
            hipLaunchKernelGGL(k_dir_keys, dim3(grid), dim3(kBlock), 0, s, g->src.get(), g->col.get(), nnz, kf.get(),
                               kt.get());
            PGD_HIP(hipGetLastError());
            const unsigned kb = 32u + (unsigned)bits_for(n - 1);
            sort_keys(ctx, kf.get(), sf.get(), (size_t)nnz, 0, kb);
            sort_keys(ctx, kt.get(), st.get(), (size_t)nnz, 0, kb);
            hipLaunchKernelGGL(k_csr_sym_check, dim3(grid), dim3(kBlock), 0, s, sf.get(), st.get(), nnz, bad.get());
            PGD_HIP(hipGetLastError());
            PGD_HIP(hipMemcpyAsync(&hbad, bad.get(), sizeof(hbad), hipMemcpyDeviceToHost, s));
            sync(ctx);
            PGD_CHECK(!(hbad & 2u), PGD_ERR_INVALID, "CSR is not symmetric");
            PGD_CHECK(!(hbad & 4u), PGD_ERR_UNSUPPORTED,
                      "CSR repeats a neighbour (a multi-edge); only simple graphs are supported");
        }
        g->max_deg = md;
        g->avg_deg = n > 0 ? (int32_t)((double)nnz / (double)n) : 0;
    });
    if (st != PGD_OK) {
        delete g;
        return st;
    }
    *out = g;
    return PGD_OK;
}

pgd_status pgd_graph_free(pgd_graph* g) {
    if (!g) return PGD_ERR_INVALID;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->orient) orient_state_free(g->ctx, g->orient);
    delete g;
    return PGD_OK;
}

pgd_status pgd_graph_info(pgd_graph* g, int64_t* n, int64_t* m, int32_t* max_deg, int32_t* avg_deg) {
    if (!g) return PGD_ERR_INVALID;
    if (n) *n = g->n;
    if (m) *m = g->m;
    if (max_deg) *max_deg = g->max_deg;
    if (avg_deg) *avg_deg = g->avg_deg;
    return PGD_OK;
}

pgd_status pgd_graph_csr(pgd_graph* g, int64_t* row, int32_t* col) {
    if (!g || !row || (!col && g->nnz)) return PGD_ERR_INVALID;
    return guarded(g->ctx, [&] {
        PGD_HIP(hipMemcpyAsync(row, g->row.get(), (g->n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                               g->ctx->stream));
        if (g->nnz)
            PGD_HIP(hipMemcpyAsync(col, g->col.get(), g->nnz * sizeof(int32_t), hipMemcpyDeviceToHost,
                                   g->ctx->stream));
        sync(g->ctx);
    });
}

pgd_status pgd_preprocess(pgd_graph* g, const char* method, int32_t s2l) {
    if (!g) return PGD_ERR_INVALID;
    return guarded(g->ctx, [&] {
        pgd_ctx* ctx = g->ctx;
        hipStream_t s = ctx->stream;
        PGD_HIP(hipSetDevice(ctx->device));
        const int64_t n = g->n, nnz = g->nnz;
        int kind = neigh_kind(method);
        int grid = std::min(grid_for(std::max(n, nnz)), ctx->num_cus * 16);
        if (kind != NK_NONE && nnz) {
            if (kind == NK_KCORE || kind == NK_DEG_KCORE || kind == NK_KCORE_VOL || kind == NK_DEG_KCORE_VOL)
                compute_cores(g);
            DBuf<int32_t> feat, v0;
            DBuf<uint64_t> k0, k1;
            feat.alloc(ctx, std::max<int64_t>(n, 1));
            hipLaunchKernelGGL(k_vertex_feature, dim3(grid), dim3(kBlock), 0, s, g->row.get(), g->col.get(),
                               g->deg.get(), g->have_core ? g->core.get() : nullptr, n, kind, feat.get());
            PGD_HIP(hipGetLastError());
            DBuf<int32_t> mm;
            mm.alloc(ctx, 2);
            hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, s, mm.get());
            hipLaunchKernelGGL(k_feat_minmax, dim3(grid), dim3(kBlock), 0, s, feat.get(), n, mm.get());
            PGD_HIP(hipGetLastError());
            int32_t hmm[2] = {0, 0};
            PGD_HIP(hipMemcpyAsync(hmm, mm.get(), sizeof(hmm), hipMemcpyDeviceToHost, s));
            PGD_HIP(hipStreamSynchronize(s));
            const int fb = bits_for((int64_t)hmm[1] - (int64_t)hmm[0]);
            k0.alloc(ctx, nnz); k1.alloc(ctx, nnz); v0.alloc(ctx, nnz);
            hipLaunchKernelGGL(k_neigh_keys, dim3(grid), dim3(kBlock), 0, s, g->src.get(), g->col.get(), feat.get(),
                               nnz, s2l ? 1 : 0, mm.get(), fb, k0.get(), v0.get());
            PGD_HIP(hipGetLastError());
            sort_pairs(ctx, k0.get(), k1.get(), v0.get(), g->col.get(), (size_t)nnz, 0,
                       (unsigned)(fb + bits_for(n - 1)));
        }
        // orientation and edge ids
        g->ev.alloc(ctx, std::max<int64_t>(g->m, 1));
        g->eu.alloc(ctx, std::max<int64_t>(g->m, 1));
        if (nnz) {
            DBuf<int32_t> f;
            DBuf<int64_t> eid;
            f.alloc(ctx, nnz);
            eid.alloc(ctx, nnz);
            hipLaunchKernelGGL(k_edge_flags, dim3(grid), dim3(kBlock), 0, s, g->src.get(), g->col.get(), nnz, f.get());
            PGD_HIP(hipGetLastError());
            exclusive_sum(ctx, f.get(), eid.get(), (size_t)nnz);
            hipLaunchKernelGGL(k_edge_arrays, dim3(grid), dim3(kBlock), 0, s, g->src.get(), g->col.get(),
                               g->deg.get(), eid.get(), nnz, g->ev.get(), g->eu.get());
            PGD_HIP(hipGetLastError());
        }
        sync(ctx);
        g->prepped = true;
        g->have_order = false;
        g->have_counts = false;
        // (the oriented engine's buffers stay: every count rebuilds their
        // contents, and releasing and re-mapping its multi-GB lists costs
        // about 8 ms per call on a 64M-edge graph)
    });
}

pgd_status pgd_edge_arrays(pgd_graph* g, int32_t* e_v, int32_t* e_u) {
    if (!g || (!e_v && !e_u)) return PGD_ERR_INVALID;
    if (!g->prepped) return PGD_ERR_STATE;
    return guarded(g->ctx, [&] {
        if (g->m && e_v)
            PGD_HIP(hipMemcpyAsync(e_v, g->ev.get(), g->m * sizeof(int32_t), hipMemcpyDeviceToHost, g->ctx->stream));
        if (g->m && e_u)
            PGD_HIP(hipMemcpyAsync(e_u, g->eu.get(), g->m * sizeof(int32_t), hipMemcpyDeviceToHost, g->ctx->stream));
        sync(g->ctx);
    });
}

pgd_status pgd_compute_cores(pgd_graph* g, int32_t* core, int32_t* max_core) {
    if (!g) return PGD_ERR_INVALID;
    return guarded(g->ctx, [&] {
        PGD_HIP(hipSetDevice(g->ctx->device));
        compute_cores(g);
        if (core && g->n)
            PGD_HIP(hipMemcpyAsync(core, g->core.get(), g->n * sizeof(int32_t), hipMemcpyDeviceToHost,
                                   g->ctx->stream));
        sync(g->ctx);
        if (max_core) *max_core = g->K;
    });
}

namespace {
// E_ordered from the keys of positions [0, cnt): all edges (ids == nullptr)
// or the device copy of a sampled set E_s
void build_order(pgd_graph* g, int kind, int32_t s2l, const int32_t* ids, int64_t cnt) {
    pgd_ctx* ctx = g->ctx;
    PGD_HIP(hipSetDevice(ctx->device));
    if (kind == OK_KCORE || kind == OK_KCORE_DEG || kind == OK_KCORE_VOL || kind == OK_DEG_KCORE_VOL)
        compute_cores(g);
    g->have_order = false;
    g->have_counts = false;
    g->have_micro = false;
    const int64_t need = std::max<int64_t>(std::max(g->m, cnt), 1);
    if (g->order.n < (size_t)need) g->order.alloc(ctx, need);
    if (cnt) {
        DBuf<uint64_t> k0, k1;
        k0.alloc(ctx, cnt);
        k1.alloc(ctx, cnt);
        int grid = std::min(grid_for(cnt), ctx->num_cus * 16);
        hipLaunchKernelGGL(k_order_keys, dim3(grid), dim3(kBlock), 0, ctx->stream, g->ev.get(), g->eu.get(),
                           g->deg.get(), g->have_core ? g->core.get() : nullptr, g->row.get(), g->col.get(), ids,
                           cnt, kind, s2l ? 1 : 0, k0.get());
        PGD_HIP(hipGetLastError());
        sort_keys(ctx, k0.get(), k1.get(), (size_t)cnt, 0, 64);
        hipLaunchKernelGGL(k_order_from_keys, dim3(grid), dim3(kBlock), 0, ctx->stream, k1.get(), cnt,
                           g->order.get());
        PGD_HIP(hipGetLastError());
    }
    sync(ctx);
    g->nord = cnt;
    g->have_order = true;
}

// an edge list given by the caller: ids in [0, m), repeats allowed
pgd_status check_edge_ids(pgd_graph* g, const int32_t* ids, int64_t count) {
    if (count < 0 || count >= ((int64_t)1 << 31) || (count > 0 && !ids)) return PGD_ERR_INVALID;
    for (int64_t i = 0; i < count; i++)
        if (ids[i] < 0 || ids[i] >= g->m) {
            g->ctx->last_err = "edge id outside [0, m)";
            return PGD_ERR_INVALID;
        }
    return PGD_OK;
}
}  // namespace

pgd_status pgd_sort_edges(pgd_graph* g, const char* ordering, int32_t s2l) {
    if (!g) return PGD_ERR_INVALID;
    if (!g->prepped) return PGD_ERR_STATE;
    int kind = order_kind(ordering);
    if (kind < 0) {
        g->ctx->last_err = "ordering has no reproducible key in the reference (rand/var)";
        return PGD_ERR_UNSUPPORTED;
    }
    return guarded(g->ctx, [&] { build_order(g, kind, s2l, nullptr, g->m); });
}

pgd_status pgd_sort_edges_subset(pgd_graph* g, const char* ordering, int32_t s2l, const int32_t* ids,
                                 int64_t count) {
    if (!g) return PGD_ERR_INVALID;
    if (!g->prepped) return PGD_ERR_STATE;
    pgd_status st = check_edge_ids(g, ids, count);
    if (st != PGD_OK) return st;
    int kind = order_kind(ordering);
    if (kind < 0) {
        g->ctx->last_err = "ordering has no reproducible key in the reference (rand/var)";
        return PGD_ERR_UNSUPPORTED;
    }
    return guarded(g->ctx, [&] {
        DBuf<int32_t> d;
        d.alloc(g->ctx, std::max<int64_t>(count, 1));
        if (count)
            PGD_HIP(hipMemcpyAsync(d.get(), ids, count * sizeof(int32_t), hipMemcpyHostToDevice, g->ctx->stream));
        build_order(g, kind, s2l, d.get(), count);
    });
}

pgd_status pgd_set_e_ordered(pgd_graph* g, const int32_t* ids, int64_t count) {
    if (!g) return PGD_ERR_INVALID;
    if (!g->prepped) return PGD_ERR_STATE;
    pgd_status st = check_edge_ids(g, ids, count);
    if (st != PGD_OK) return st;
    return guarded(g->ctx, [&] {
        pgd_ctx* ctx = g->ctx;
        PGD_HIP(hipSetDevice(ctx->device));
        g->have_order = false;
        g->have_counts = false;
        g->have_micro = false;
        const int64_t need = std::max<int64_t>(std::max(g->m, count), 1);
        if (g->order.n < (size_t)need) g->order.alloc(ctx, need);
        if (count)
            PGD_HIP(hipMemcpyAsync(g->order.get(), ids, count * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        sync(ctx);
        g->nord = count;
        g->have_order = true;
    });
}

pgd_status pgd_e_ordered_count(pgd_graph* g, int64_t* count) {
    if (!g || !count) return PGD_ERR_INVALID;
    if (!g->have_order) return PGD_ERR_STATE;
    *count = g->nord;
    return PGD_OK;
}

pgd_status pgd_e_ordered(pgd_graph* g, int32_t* ids) {
    if (!g || (!ids && g->nord)) return PGD_ERR_INVALID;
    if (!g->have_order) return PGD_ERR_STATE;
    return guarded(g->ctx, [&] {
        if (g->nord)
            PGD_HIP(hipMemcpyAsync(ids, g->order.get(), g->nord * sizeof(int32_t), hipMemcpyDeviceToHost,
                                   g->ctx->stream));
        sync(g->ctx);
    });
}

pgd_status pgd_decompose(pgd_graph* g, uint32_t flags, uint64_t* sums) {
    if (!g || !sums) return PGD_ERR_INVALID;
    if (flags & ~(PGD_MICRO | PGD_ENGINE_ORIENT | PGD_ENGINE_DENSE)) {
        g->ctx->last_err = "unknown decomposition flag";
        return PGD_ERR_INVALID;
    }
    if (!g->prepped || !g->have_order) return PGD_ERR_STATE;
    if ((flags & PGD_MICRO) && g->nord > g->m) {
        // the reference's micro vectors hold m positions (graphlet_core.cpp:1516-1522)
        g->ctx->last_err = "E_ordered has more than m positions; per-edge columns hold m";
        return PGD_ERR_INVALID;
    }
    return guarded(g->ctx, [&] {
        pgd_ctx* ctx = g->ctx;
        PGD_HIP(hipSetDevice(ctx->device));
        g->have_counts = false;
        g->have_micro = false;
        const int64_t m = std::max<int64_t>(g->m, 1);
        if (g->T.n != (size_t)m) {
            g->T.alloc(ctx, m);
            g->K4.alloc(ctx, m);
            g->C4.alloc(ctx, m);
        }
        if (!g->sums.get()) g->sums.alloc(ctx, PGD_NSUMS);
        const bool micro = (flags & PGD_MICRO) != 0;
        if (micro && g->micro.n != (size_t)PGD_NMICRO * m) g->micro.alloc(ctx, (size_t)PGD_NMICRO * m);
        GraphView v = g->view();
        EdgeCounts c{g->T.get(), g->K4.get(), g->C4.get()};
        hipEvent_t e0, e1;
        PGD_HIP(hipEventCreate(&e0));
        PGD_HIP(hipEventCreate(&e1));
        PGD_HIP(hipEventRecord(e0, ctx->stream));
        Timing t;
        int engine;
        if (flags & PGD_ENGINE_ORIENT) engine = 2;
        else if (flags & PGD_ENGINE_DENSE) engine = 3;
        else {
            // the faster engine, unless the dense tier would pass the context's memory cap
            const bool room = ctx->mem_cap <= 0 ||
                              ctx->mem_used.load() + (int64_t)engine_dense_bytes(v) <= ctx->mem_cap;
            engine = engine_dense_ok(v) && room ? 3 : 2;
        }
        DenseRunPtr drun;  // the dense tier's timers, read after the final synchronisation
        if (g->m > 0) {
            if (engine == 2) engine_orient(ctx, v, c, &t, &g->orient, g->prm);
            else {
                PGD_CHECK(engine_dense_fits(v), PGD_ERR_UNSUPPORTED, "graph too large for the dense engine");
                drun = engine_dense(ctx, v, c);
            }
        }
        epilogue(ctx, v, c, g->sums.get(), micro ? g->micro.get() : nullptr);
        PGD_HIP(hipEventRecord(e1, ctx->stream));
        PGD_HIP(hipMemcpyAsync(sums, g->sums.get(), PGD_NSUMS * sizeof(uint64_t), hipMemcpyDeviceToHost,
                               ctx->stream));
        sync(ctx);
        if (drun) dense_timing(*drun, &t);
        float ms = 0;
        PGD_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        t.total_ms = ms;
        g->timing = t;
        g->last_engine = engine;
        g->have_counts = true;
        g->have_micro = micro;
    });
}

pgd_status pgd_decompose_noalloc(pgd_graph* g, uint32_t flags, uint64_t* d_sums, uint64_t* d_micro) {
    if (!g || !d_sums) return PGD_ERR_INVALID;
    if (flags & ~PGD_MICRO) {
        g->ctx->last_err = "unknown decomposition flag";
        return PGD_ERR_INVALID;
    }
    if ((flags & PGD_MICRO) && !d_micro) return PGD_ERR_INVALID;
    if (!g->prepped || !g->have_order) return PGD_ERR_STATE;
    if (g->nord > g->m && (flags & PGD_MICRO)) {
        g->ctx->last_err = "E_ordered has more than m positions; per-edge columns hold m";
        return PGD_ERR_INVALID;
    }
    return guarded(g->ctx, [&] {
        PGD_HIP(hipSetDevice(g->ctx->device));
        decompose_noalloc(g->ctx, g->view(), d_sums, (flags & PGD_MICRO) ? d_micro : nullptr);
    });
}

pgd_status pgd_micro(pgd_graph* g, int32_t column, uint64_t* out) {
    if (!g || !out || column < 0 || column >= PGD_NMICRO) return PGD_ERR_INVALID;
    if (!g->have_micro) return PGD_ERR_STATE;
    return guarded(g->ctx, [&] {
        if (g->m)
            PGD_HIP(hipMemcpyAsync(out, g->micro.get() + (size_t)column * g->m, g->m * sizeof(uint64_t),
                                   hipMemcpyDeviceToHost, g->ctx->stream));
        sync(g->ctx);
    });
}

pgd_status pgd_write_micro(pgd_graph* g, int32_t fd, int32_t output_id, const char* prefix, int64_t* bytes) {
    if (!g || fd < 0) return PGD_ERR_INVALID;
    if (!g->have_micro) return PGD_ERR_STATE;
    return guarded(g->ctx, [&] {
        PGD_HIP(hipSetDevice(g->ctx->device));
        write_micro_text(g->ctx, g->view(), g->micro.get(), fd, output_id != 0, prefix, bytes);
    });
}

pgd_status pgd_edge_counts(pgd_graph* g, uint64_t* tri, uint64_t* k4, uint64_t* c4) {
    if (!g) return PGD_ERR_INVALID;
    if (!g->have_counts) return PGD_ERR_STATE;
    return guarded(g->ctx, [&] {
        const int64_t m = g->m;
        if (m && tri) {
            std::vector<uint32_t> t(m);
            PGD_HIP(hipMemcpyAsync(t.data(), g->T.get(), m * sizeof(uint32_t), hipMemcpyDeviceToHost, g->ctx->stream));
            sync(g->ctx);
            for (int64_t i = 0; i < m; i++) tri[i] = t[i];
        }
        if (m && k4)
            PGD_HIP(hipMemcpyAsync(k4, g->K4.get(), m * sizeof(uint64_t), hipMemcpyDeviceToHost, g->ctx->stream));
        if (m && c4)
            PGD_HIP(hipMemcpyAsync(c4, g->C4.get(), m * sizeof(uint64_t), hipMemcpyDeviceToHost, g->ctx->stream));
        sync(g->ctx);
    });
}

pgd_status pgd_tuning_default(pgd_tuning* t) {
    if (!t) return PGD_ERR_INVALID;
    *t = tuning_from(OrientParams{});
    return PGD_OK;
}

pgd_status pgd_graph_set_tuning(pgd_graph* g, const pgd_tuning* t) {
    if (!g || !t) return PGD_ERR_INVALID;
    return guarded(g->ctx, [&] {
        const char* why = tuning_problem(*t);
        PGD_CHECK(why[0] == 0, PGD_ERR_INVALID, why);
        g->prm = params_from(*t);
    });
}

pgd_status pgd_graph_get_tuning(pgd_graph* g, pgd_tuning* t) {
    if (!g || !t) return PGD_ERR_INVALID;
    *t = tuning_from(g->prm);
    return PGD_OK;
}

int32_t pgd_last_engine(pgd_graph* g) { return g ? g->last_engine : 0; }

pgd_status pgd_last_phase(pgd_graph* g, int32_t idx, char* name, int32_t cap, double* ms, double* bytes,
                          double* units) {
    if (!g || idx < 0 || idx >= (int32_t)g->timing.phases.size()) return PGD_ERR_INVALID;
    const Phase& p = g->timing.phases[(size_t)idx];
    if (name && cap > 0) {
        size_t k = std::min<size_t>(p.name.size(), (size_t)cap - 1);
        std::memcpy(name, p.name.data(), k);
        name[k] = 0;
    }
    if (ms) *ms = p.ms;
    if (bytes) *bytes = p.bytes;
    if (units) *units = p.units;
    return PGD_OK;
}

pgd_status pgd_last_timing(pgd_graph* g, double* total_ms, double* dominant_ms, int32_t* dominant_launches) {
    if (!g) return PGD_ERR_INVALID;
    if (total_ms) *total_ms = g->timing.total_ms;
    if (dominant_ms) *dominant_ms = g->timing.dominant_ms;
    if (dominant_launches) *dominant_launches = g->timing.dominant_launches;
    return PGD_OK;
}

}  // extern "C"
