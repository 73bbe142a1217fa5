// pgd_internal.hpp -- shared host/device declarations of the MI355X
// implementation (not part of the public ABI; see include/pgd_hip.h).
#pragma once
#include <cstring>
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <vector>

#include "pgd_hip.h"

namespace pgd {

// ---------------------------------------------------------------------------
// error plumbing: every HIP call goes through PGD_HIP, which records the
// failing call in the context and returns PGD_ERR_HIP from the entry point.
// ---------------------------------------------------------------------------
struct Error {
    pgd_status code;
    std::string what;
};

#define PGD_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e__ = (call);                                                          \
        if (e__ != hipSuccess)                                                            \
            throw ::pgd::Error{PGD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e__)}; \
    } while (0)

#define PGD_CHECK(cond, code, msg)                                  \
    do {                                                            \
        if (!(cond)) throw ::pgd::Error{(code), std::string(msg)}; \
    } while (0)

}  // namespace pgd

struct pgd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;  // second stream of the plain oriented engine (its 4-cycle plan), made on first use
    hipStream_t side = nullptr;  // third: the small local-graph bins beside the large ones, made on first use
    int num_cus = 256;
    size_t lds_per_block = 65536;  // hipDeviceProp_t::sharedMemPerBlock, read once at open
    int tiled_per_cu[3][9] = {};   // occupancy of the tiled wedge kernel by (block size, tile), 0 = not asked yet
    int64_t mem_cap = 0;   // 0: no cap
    std::atomic<int64_t> mem_used{0};  // also released from other threads (served results)
    // released device blocks kept for reuse by size class (see dmalloc):
    // reuse follows the context's stream as the runtime pool's would, without
    // a runtime call, whose pool lock stalls concurrent contexts for
    // milliseconds at a time under the served form's allocation rate
    std::mutex cache_mu;
    std::multimap<size_t, void*> cache;
    size_t cached_bytes = 0;
    std::string last_err;
    // rocPRIM temporary storage, grown on demand
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    // 8 words of pinned host memory that kernels write and the host reads
    // after a stream synchronisation (made on first use)
    unsigned long long* pin = nullptr;
    unsigned long long* pin_dev = nullptr;
};

namespace pgd {

void* dmalloc(pgd_ctx* ctx, size_t bytes);
void dfree(pgd_ctx* ctx, void* p, size_t bytes);
void release_cached(pgd_ctx* ctx);  // return every cached block to the runtime pool
void* tmp_storage(pgd_ctx* ctx, size_t bytes);
unsigned long long* pinned_words(pgd_ctx* ctx);  // device address of ctx->pin

// RAII device buffer charged against the context's cap
template <typename T>
struct DBuf {
    pgd_ctx* ctx = nullptr;
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { reset(); }
    void alloc(pgd_ctx* c, size_t count) {
        reset();
        ctx = c;
        n = count;
        p = (T*)dmalloc(c, (count ? count : 1) * sizeof(T));
    }
    void reset() {
        if (p) dfree(ctx, p, (n ? n : 1) * sizeof(T));
        p = nullptr;
        n = 0;
    }
    T* get() const { return p; }
};

// Timing of the counting stage of the last decomposition
struct Phase {
    std::string name;    // kernel (or kernel family) of one counting pass
    double ms = 0;       // device time, HIP events on the context's stream
    double bytes = 0;    // algorithmic HBM bytes of the pass (work model)
    double units = 0;    // work units the pass processed (wedges, list entries)
};
struct Timing {
    double total_ms = 0, dominant_ms = 0;
    int dominant_launches = 0;
    std::vector<Phase> phases;
};

// Per-edge counts shared by every engine: indexed by edge id.
struct EdgeCounts {
    uint32_t* T;   // triangles on the edge
    uint64_t* K4;  // 4-cliques on the edge
    uint64_t* C4;  // induced 4-cycles on the edge
};

// Read-only view of a preprocessed graph as the engines see it.
struct GraphView {
    int64_t n, m, nnz;
    int32_t max_deg;
    const int64_t* row;   // n+1
    const int32_t* col;   // nnz, reference neighbour order
    const int32_t* src;   // nnz, vertex owning each CSR position
    const int32_t* deg;   // n
    const int32_t* ev;    // m
    const int32_t* eu;    // m
    const int32_t* order; // nord, E_ordered
    const int32_t* group = nullptr;  // n, graph of each vertex (batched union), or null
    int64_t ngroups = 0;
    int64_t nord = 0;     // E_ordered positions: m, or |E_s| <= m for a sampled edge set
};

// engines (each writes EdgeCounts for every edge id)
struct OrientState;
struct Timer;
// tunables of the oriented engine (bin thresholds and per-CU grid factors)
struct OrientParams {
    int small_k = 64;       // out-degree up to which one wavefront owns a local graph
    int mid_k = 256;        // up to which a 256-thread block does (1024 threads above)
    int grid_small = 64;    // resident blocks per CU for the wavefront bins
    int grid_mid = 32;      // resident blocks per CU for the 256-thread bin
    int wedge_log2r = 15;   // endpoint tile of the heavy 4-cycle pass: 2^log2r 16-bit LDS counters
    int wedge_nt = 512;     // threads per block of the heavy 4-cycle pass (256, 512 or 1024)
    int wedge_split = 4;    // work items per block the heavy wedges are cut into (load balance)
    int tri_lists = 64;     // kept local-edge lists, entries per edge (0: off)
    int tri_search = 4;     // look members up in N+(b)'s table when this x their number < |N+(b)| (0: off)
};
OrientState* orient_state_new();
OrientParams& orient_params(OrientState* s);
void orient_build(pgd_ctx* ctx, const GraphView& g, OrientState& st, bool plan = true);
void orient_count(pgd_ctx* ctx, OrientState& st, EdgeCounts out, Timer* timer);
void engine_orient(pgd_ctx* ctx, const GraphView& g, EdgeCounts out, Timing* t, OrientState** cache,
                   const OrientParams& prm);
inline OrientParams params_from(const pgd_tuning& t) {
    OrientParams p;
    p.small_k = t.small_k;
    p.mid_k = t.mid_k;
    p.grid_small = t.grid_small;
    p.grid_mid = t.grid_mid;
    p.wedge_log2r = t.wedge_log2r;
    p.wedge_nt = t.wedge_nt;
    p.wedge_split = t.wedge_split;
    p.tri_lists = t.tri_lists;
    p.tri_search = t.tri_search;
    return p;
}
inline pgd_tuning tuning_from(const OrientParams& p) {
    return pgd_tuning{p.small_k,     p.mid_k,    p.grid_small,  p.grid_mid,
                      p.wedge_log2r, p.wedge_nt, p.wedge_split, p.tri_lists, p.tri_search};
}
// empty string when t is usable, else what is wrong
const char* tuning_problem(const pgd_tuning& t);
void orient_state_free(pgd_ctx* ctx, OrientState* s);
bool engine_dense_ok(const GraphView& g);    // chosen automatically (n <= 4096, or n <= 8192 and density >= 0.08)
bool engine_dense_fits(const GraphView& g);  // may be requested (small enough)
size_t engine_dense_bytes(const GraphView& g);  // the tier's device memory for the graph
// the dense tier launches its passes without a host synchronisation; the run
// keeps its buffers and timers until dense_timing reads them, once the
// context's stream has finished the run (freeing it earlier is safe: its
// memory returns to the stream-ordered cache)
struct DenseRun;
struct DenseRunDeleter {
    void operator()(DenseRun* r) const;
};
using DenseRunPtr = std::unique_ptr<DenseRun, DenseRunDeleter>;
DenseRunPtr engine_dense(pgd_ctx* ctx, const GraphView& g, EdgeCounts out);
void dense_timing(DenseRun& r, Timing* t);
// allocation-free path (pgd_decompose_noalloc): the dense engine and the
// epilogue over the static scratch array of this GPU
void noalloc_scratch_init(int device);
void decompose_noalloc(pgd_ctx* ctx, const GraphView& g, uint64_t* d_sums, uint64_t* d_micro);

// epilogue: the 15 sums (and the micro columns when micro != nullptr),
// processed in E_ordered order over its nord positions.
void epilogue(pgd_ctx* ctx, const GraphView& g, EdgeCounts c, uint64_t* d_sums, uint64_t* d_micro);

// the .micro text of a decomposition's micro columns, written to fd
void write_micro_text(pgd_ctx* ctx, const GraphView& g, const uint64_t* micro, int fd, bool output_id,
                      const char* prefix, int64_t* written);

// the CSR of g (a new, empty pgd_graph) from a COO already in device memory
// (L lines, ids in [0, n)), as pgd_graph_from_coo builds it from host lines;
// ds and dd are only read
void graph_from_device_coo(pgd_ctx* ctx, pgd_graph* g, int64_t n, int64_t L, DBuf<int32_t>& ds,
                           DBuf<int32_t>& dd, uint32_t flags);

// launch-geometry helpers
inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace pgd

struct pgd_graph {
    pgd_ctx* ctx = nullptr;
    int64_t n = 0, m = 0, nnz = 0;
    int32_t max_deg = 0, avg_deg = 0;
    pgd::DBuf<int64_t> row;
    pgd::DBuf<int32_t> col, src, deg;
    bool prepped = false;
    pgd::DBuf<int32_t> ev, eu;
    bool have_core = false;
    int32_t K = 0;
    pgd::DBuf<int32_t> core;
    bool have_order = false;
    int64_t nord = 0;  // E_ordered positions (m after pgd_sort_edges; |E_s| for a sampled set)
    pgd::DBuf<int32_t> order;
    bool have_counts = false, have_micro = false;
    pgd::DBuf<uint32_t> T;
    pgd::DBuf<uint64_t> K4, C4, micro, sums;
    int last_engine = 0;
    pgd::Timing timing;
    pgd::OrientState* orient = nullptr;
    pgd::OrientParams prm;
    pgd::GraphView view() const {
        pgd::GraphView v{n, m, nnz, max_deg, row.get(), col.get(), src.get(), deg.get(), ev.get(), eu.get(),
                         order.get()};
        v.nord = nord;
        return v;
    }
};
