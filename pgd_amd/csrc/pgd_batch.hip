// pgd_batch.hip -- batched form: thousands of independent graphs counted
// together.
//
// A batch is the disjoint union of its graphs: graph g's vertices are the
// union ids [voff_g, voff_g + n_g).  Everything the plain path computes per
// graph (CSR, neighbour order, e_v/e_u, edge ids, E_ordered, per-edge counts)
// is a per-component property, so running the plain pipeline on the union
// gives every graph's own result: its edge ids are the union ids
// [eoff_g, eoff_g + m_g) in the same order, and its E_ordered is the union
// E_ordered restricted to its edges (stable partition by graph).  Two things
// are per graph: the ranks of the oriented engine are grouped by graph (so a
// graph's vertices form one rank range and its wedge endpoints one or two
// LDS tiles), and the epilogue, which needs each graph's n and m, is a
// persistent kernel pulling (graph, edge-range) tiles from a device queue.
//
// The counting sequence (the oriented engine's passes, the segmented
// epilogue) has no host synchronisation, allocation or memset, and is
// captured once in a HIP graph that pgd_batch_run replays.  Launch
// parameters come from pgd_batch_set_tuning or the built-in autotuner.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pgd_epilogue.hpp"
#include "pgd_orient.hpp"

using namespace pgd;

namespace {

constexpr int kEpiTile = 4096;  // E_ordered positions per epilogue tile

struct EpiItem {
    int32_t graph, pad;
    int64_t p0, p1;
};

__global__ void b_fill_u64(uint64_t* __restrict__ p, int64_t n) {
    GRID_STRIDE(i, n) p[i] = 0;
}
__global__ void b_fill_u32(uint32_t* __restrict__ p, int64_t n) {
    GRID_STRIDE(i, n) p[i] = 0;
}

// union ids of the packed edge lists, in place: line i belongs to the graph g
// with line_off[g] <= i < line_off[g + 1] (the last such g, so graphs without
// lines are passed over) and its ids are shifted by voff[g]; an id outside
// [0, n_g) sets *bad and the line gets -1, -1
__global__ void k_union_ids(int32_t* __restrict__ s, int32_t* __restrict__ d, int64_t L,
                            const int64_t* __restrict__ line_off, const int64_t* __restrict__ voff, int32_t ng,
                            uint32_t* __restrict__ bad) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < L; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = ng - 1;  // line_off[0] = 0 <= i < line_off[ng] = L
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (line_off[mid] <= i) lo = mid; else hi = mid - 1;
        }
        const int64_t off = voff[lo], n = voff[lo + 1] - off;
        const int32_t x = s[i], y = d[i];
        if (x < 0 || y < 0 || x >= n || y >= n) {
            atomicOr(bad, 1u);
            s[i] = -1;
            d[i] = -1;
        } else {
            s[i] = (int32_t)(x + off);
            d[i] = (int32_t)(y + off);
        }
    }
}

__global__ void k_vertex_group(const int64_t* __restrict__ voff, int32_t ng, int32_t* __restrict__ group) {
    for (int64_t g = blockIdx.x; g < ng; g += gridDim.x)
        for (int64_t v = voff[g] + threadIdx.x; v < voff[g + 1]; v += blockDim.x) group[v] = (int32_t)g;
}

// row offsets at the graph boundaries (edge offsets = row / 2)
__global__ void k_graph_rows(const int64_t* __restrict__ row, const int64_t* __restrict__ voff, int32_t ng,
                             int64_t* __restrict__ out) {
    GRID_STRIDE(g, (int64_t)ng + 1) out[g] = row[voff[g]];
}

// key of E_ordered position p: the graph of its edge
__global__ void k_order_group(const int32_t* __restrict__ order, const int32_t* __restrict__ ev,
                              const int32_t* __restrict__ group, int64_t m, uint32_t* __restrict__ key) {
    GRID_STRIDE(p, m) key[p] = (uint32_t)group[ev[order[p]]];
}

// segmented epilogue: per-graph sums and micro columns, tiles from a queue
__global__ __launch_bounds__(256) void k_epilogue_batch(
    const EpiItem* __restrict__ items, int64_t nitems, uint32_t* __restrict__ qctr, const int32_t* __restrict__ order,
    const int32_t* __restrict__ ev, const int32_t* __restrict__ eu, const int32_t* __restrict__ deg,
    const uint32_t* __restrict__ Tc, const uint64_t* __restrict__ K4c, const uint64_t* __restrict__ C4c,
    const int64_t* __restrict__ gn, const int64_t* __restrict__ gm, uint64_t* __restrict__ sums,
    uint64_t* __restrict__ micro, int64_t mstride) {
    __shared__ int64_t it_sh;
    for (;;) {
        if (threadIdx.x == 0) it_sh = (int64_t)atomicAdd(qctr, 1u);
        __syncthreads();
        const int64_t it = it_sh;
        __syncthreads();
        if (it >= nitems) break;
        const EpiItem I = items[it];
        const int64_t n = gn[I.graph], m = gm[I.graph];
        uint64_t acc[PGD_NSUMS];
#pragma unroll
        for (int k = 0; k < PGD_NSUMS; k++) acc[k] = 0;
        for (int64_t p = I.p0 + threadIdx.x; p < I.p1; p += blockDim.x) {
            const int32_t id = order[p];
            const int32_t v = ev[id], u = eu[id];
            edge_terms(deg[v], deg[u], Tc[id], K4c[id], C4c[id], n, m, acc, micro,
                       mstride, p);
        }
        block_add_sums(acc, sums + (size_t)I.graph * PGD_NSUMS);
    }
}

}  // namespace

struct pgd_batch {
    pgd_ctx* ctx = nullptr;
    int32_t ng = 0;
    uint32_t flags = 0;
    int64_t N = 0, M = 0;
    std::vector<int64_t> n, voff, m, eoff;
    pgd_graph* U = nullptr;
    DBuf<int32_t> group;
    DBuf<int64_t> gn, gm;
    DBuf<EpiItem> items;
    int64_t nitems = 0;
    DBuf<uint64_t> sums;
    DBuf<uint32_t> ctl;  // [0] epilogue queue head
    OrientState* st = nullptr;
    OrientParams prm;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double last_ms = 0;
    std::vector<uint64_t> host_sums;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~pgd_batch() {
        drop_graph();
        if (st) orient_state_free(ctx, st);
        if (U) pgd_graph_free(U);
    }
};

namespace {

template <class F>
pgd_status batch_guard(pgd_ctx* ctx, F&& f) {
    try {
        f();
        return PGD_OK;
    } catch (const Error& e) {
        ctx->last_err = e.what;
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->last_err = "host allocation failed";
        return PGD_ERR_OOM;
    }
}

void check_status(pgd_ctx* ctx, pgd_status s, const char* what) {
    if (s != PGD_OK) throw Error{s, std::string(what) + ": " + ctx->last_err};
}

GraphView union_view(pgd_batch* b) {
    GraphView v = b->U->view();
    v.group = b->group.get();
    v.ngroups = b->ng;
    return v;
}

// the capturable counting sequence
void batch_count(pgd_batch* b) {
    pgd_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    pgd_graph* U = b->U;
    if (b->M) orient_count(ctx, *b->st, EdgeCounts{U->T.get(), U->K4.get(), U->C4.get()}, nullptr);
    const int64_t ns = (int64_t)b->ng * PGD_NSUMS;
    hipLaunchKernelGGL(b_fill_u64, dim3(grid_cap(ctx, (ns + 255) / 256, 4)), dim3(256), 0, s, b->sums.get(), ns);
    hipLaunchKernelGGL(b_fill_u32, dim3(1), dim3(64), 0, s, b->ctl.get(), (int64_t)1);
    PGD_HIP(hipGetLastError());
    if (b->nitems) {
        hipLaunchKernelGGL(k_epilogue_batch, dim3(grid_cap(ctx, b->nitems, 8)), dim3(256), 0, s, b->items.get(),
                           b->nitems, b->ctl.get(), U->order.get(), U->ev.get(), U->eu.get(), U->deg.get(),
                           U->T.get(), U->K4.get(), U->C4.get(), b->gn.get(), b->gm.get(), b->sums.get(),
                           (b->flags & PGD_MICRO) ? U->micro.get() : nullptr, b->M);
        PGD_HIP(hipGetLastError());
    }
}

// plan the engine for the current parameters and capture the sequence
void batch_plan(pgd_batch* b) {
    pgd_ctx* ctx = b->ctx;
    b->drop_graph();
    if (b->M) {
        if (!b->st) b->st = orient_state_new();
        orient_params(b->st) = b->prm;
        orient_build(ctx, union_view(b), *b->st);
    }
    PGD_HIP(hipStreamSynchronize(ctx->stream));
    PGD_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    try {
        batch_count(b);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    PGD_HIP(hipStreamEndCapture(ctx->stream, &b->graph));
    PGD_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
}

// one replay; sums to host; returns device ms
double batch_replay(pgd_batch* b) {
    pgd_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    if (!b->exec) batch_plan(b);
    hipEvent_t e0, e1;
    PGD_HIP(hipEventCreate(&e0));
    PGD_HIP(hipEventCreate(&e1));
    PGD_HIP(hipEventRecord(e0, s));
    PGD_HIP(hipGraphLaunch(b->exec, s));
    PGD_HIP(hipEventRecord(e1, s));
    b->host_sums.resize((size_t)b->ng * PGD_NSUMS);
    PGD_HIP(hipMemcpyAsync(b->host_sums.data(), b->sums.get(), b->host_sums.size() * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, s));
    PGD_HIP(hipStreamSynchronize(s));
    float ms = 0;
    PGD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->last_ms = ms;
    return ms;
}

}  // namespace

namespace {

// the union, its preprocessing, grouped E_ordered and per-graph tables
void batch_build(pgd_batch* b, int32_t ngraphs, const int32_t* n, const int64_t* line_off, const int32_t* src,
                 const int32_t* dst) {
    pgd_ctx* ctx = b->ctx;
    PGD_HIP(hipSetDevice(ctx->device));
    const int32_t ng = ngraphs;
    b->n.assign(n, n + ng);
    b->voff.assign(ng + 1, 0);
    PGD_CHECK(line_off[0] == 0, PGD_ERR_INVALID, "line_off[0] must be 0");
    for (int32_t g = 0; g < ng; g++) {
        PGD_CHECK(n[g] >= 0, PGD_ERR_INVALID, "negative vertex count");
        PGD_CHECK(line_off[g + 1] >= line_off[g], PGD_ERR_INVALID, "line_off must be non-decreasing");
        b->voff[g + 1] = b->voff[g] + n[g];
    }
    const int64_t L = line_off[ng];
    PGD_CHECK(L == 0 || (src && dst), PGD_ERR_INVALID, "missing edge lists");
    b->N = b->voff[ng];
    PGD_CHECK(b->N < INT32_MAX, PGD_ERR_UNSUPPORTED, "batch has 2^31 vertices or more");
    PGD_CHECK(L < ((int64_t)1 << 31), PGD_ERR_UNSUPPORTED, "batch has 2^31 lines or more");
    // union COO on the device: graph-local ids shifted by voff, checked against n_g
    {
        hipStream_t s = ctx->stream;
        DBuf<int32_t> ds, dd;
        DBuf<int64_t> dlo, dvo;
        DBuf<uint32_t> dbad;
        ds.alloc(ctx, L);
        dd.alloc(ctx, L);
        dlo.alloc(ctx, ng + 1);
        dvo.alloc(ctx, ng + 1);
        dbad.alloc(ctx, 1);
        uint32_t bad = 0;
        if (L) {
            PGD_HIP(hipMemcpyAsync(ds.get(), src, L * sizeof(int32_t), hipMemcpyHostToDevice, s));
            PGD_HIP(hipMemcpyAsync(dd.get(), dst, L * sizeof(int32_t), hipMemcpyHostToDevice, s));
            PGD_HIP(hipMemcpyAsync(dlo.get(), line_off, (ng + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
            PGD_HIP(hipMemcpyAsync(dvo.get(), b->voff.data(), (ng + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
            PGD_HIP(hipMemsetAsync(dbad.get(), 0, sizeof(uint32_t), s));
            hipLaunchKernelGGL(k_union_ids, dim3(grid_cap(ctx, (L + 255) / 256, 16)), dim3(256), 0, s, ds.get(),
                               dd.get(), L, dlo.get(), dvo.get(), ng, dbad.get());
            PGD_HIP(hipGetLastError());
            PGD_HIP(hipMemcpyAsync(&bad, dbad.get(), sizeof(bad), hipMemcpyDeviceToHost, s));
            PGD_HIP(hipStreamSynchronize(s));
        }
        PGD_CHECK(bad == 0, PGD_ERR_INVALID, "vertex id outside [0, n) of its graph");
        pgd_graph* U = new pgd_graph();
        U->ctx = ctx;
        b->U = U;
        graph_from_device_coo(ctx, U, b->N, L, ds, dd, b->flags & PGD_DEDUP);
    }
    check_status(ctx, pgd_preprocess(b->U, "deg", 0), "preprocess");
    check_status(ctx, pgd_sort_edges(b->U, "deg", 0), "sort_edges");
    pgd_graph* U = b->U;
    hipStream_t s = ctx->stream;
    b->M = U->m;
    // vertex groups, per-graph edge ranges
    DBuf<int64_t> dvoff, grow;
    dvoff.alloc(ctx, ng + 1);
    grow.alloc(ctx, ng + 1);
    b->group.alloc(ctx, std::max<int64_t>(b->N, 1));
    PGD_HIP(hipMemcpyAsync(dvoff.get(), b->voff.data(), (ng + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_vertex_group, dim3(grid_cap(ctx, ng, 8)), dim3(256), 0, s, dvoff.get(), ng,
                       b->group.get());
    hipLaunchKernelGGL(k_graph_rows, dim3(grid_cap(ctx, (ng + 256) / 256, 4)), dim3(256), 0, s, U->row.get(),
                       dvoff.get(), ng, grow.get());
    PGD_HIP(hipGetLastError());
    std::vector<int64_t> hrow(ng + 1);
    PGD_HIP(hipMemcpyAsync(hrow.data(), grow.get(), (ng + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PGD_HIP(hipStreamSynchronize(s));
    b->eoff.resize(ng + 1);
    b->m.resize(ng);
    for (int32_t g = 0; g <= ng; g++) b->eoff[g] = hrow[g] / 2;
    for (int32_t g = 0; g < ng; g++) b->m[g] = b->eoff[g + 1] - b->eoff[g];
    PGD_CHECK(b->eoff[ng] == b->M, PGD_ERR_STATE, "graph edge ranges do not cover the union");
    // E_ordered grouped by graph (stable: each graph keeps its own order)
    if (b->M) {
        DBuf<uint32_t> k0, k1;
        DBuf<int32_t> o1;
        k0.alloc(ctx, b->M);
        k1.alloc(ctx, b->M);
        o1.alloc(ctx, b->M);
        hipLaunchKernelGGL(k_order_group, dim3(grid_cap(ctx, (b->M + 255) / 256, 16)), dim3(256), 0, s,
                           U->order.get(), U->ev.get(), b->group.get(), b->M, k0.get());
        PGD_HIP(hipGetLastError());
        sort_kv(ctx, k0.get(), k1.get(), U->order.get(), o1.get(), (size_t)b->M, (unsigned)bitsfor(ng - 1));
        PGD_HIP(hipMemcpyAsync(U->order.get(), o1.get(), b->M * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    }
    // per-graph sizes and epilogue tiles
    std::vector<EpiItem> items;
    for (int32_t g = 0; g < ng; g++)
        for (int64_t p = b->eoff[g]; p < b->eoff[g + 1]; p += kEpiTile)
            items.push_back(EpiItem{g, 0, p, std::min<int64_t>(p + kEpiTile, b->eoff[g + 1])});
    b->nitems = (int64_t)items.size();
    b->items.alloc(ctx, std::max<int64_t>(b->nitems, 1));
    if (b->nitems)
        PGD_HIP(hipMemcpyAsync(b->items.get(), items.data(), items.size() * sizeof(EpiItem),
                               hipMemcpyHostToDevice, s));
    b->gn.alloc(ctx, ng);
    b->gm.alloc(ctx, ng);
    PGD_HIP(hipMemcpyAsync(b->gn.get(), b->n.data(), ng * sizeof(int64_t), hipMemcpyHostToDevice, s));
    PGD_HIP(hipMemcpyAsync(b->gm.get(), b->m.data(), ng * sizeof(int64_t), hipMemcpyHostToDevice, s));
    b->sums.alloc(ctx, (size_t)ng * PGD_NSUMS);
    b->ctl.alloc(ctx, 1);
    const int64_t M1 = std::max<int64_t>(b->M, 1);
    U->T.alloc(ctx, M1);
    U->K4.alloc(ctx, M1);
    U->C4.alloc(ctx, M1);
    if (b->flags & PGD_MICRO) U->micro.alloc(ctx, (size_t)PGD_NMICRO * M1);
    PGD_HIP(hipStreamSynchronize(s));
}

}  // namespace

extern "C" {

pgd_status pgd_batch_create(pgd_ctx* ctx, int32_t ngraphs, const int32_t* n, const int64_t* line_off,
                            const int32_t* src, const int32_t* dst, uint32_t flags, pgd_batch** out) {
    if (!ctx || !out || ngraphs < 1 || !n || !line_off) return PGD_ERR_INVALID;
    *out = nullptr;
    std::unique_ptr<pgd_batch> b(new (std::nothrow) pgd_batch());
    if (!b) return PGD_ERR_OOM;
    b->ctx = ctx;
    b->ng = ngraphs;
    b->flags = flags & (PGD_DEDUP | PGD_MICRO);
    pgd_status st = batch_guard(ctx, [&] { batch_build(b.get(), ngraphs, n, line_off, src, dst); });
    if (st != PGD_OK) return st;
    *out = b.release();
    return PGD_OK;
}

pgd_status pgd_batch_run(pgd_batch* b, uint64_t* sums, uint64_t* totals) {
    if (!b) return PGD_ERR_INVALID;
    return batch_guard(b->ctx, [&] {
        PGD_HIP(hipSetDevice(b->ctx->device));
        batch_replay(b);
        const size_t ns = (size_t)b->ng * PGD_NSUMS;
        if (sums) std::copy(b->host_sums.begin(), b->host_sums.begin() + ns, sums);
        if (totals)
            for (int32_t g = 0; g < b->ng; g++)
                check_status(b->ctx,
                             pgd_derive(b->n[g], b->m[g], b->host_sums.data() + (size_t)g * PGD_NSUMS,
                                        totals + (size_t)g * PGD_NTOTALS),
                             "derive");
    });
}

pgd_status pgd_batch_micro(pgd_batch* b, int32_t gi, int32_t column, uint64_t* out) {
    if (!b || !out || gi < 0 || gi >= b->ng || column < 0 || column >= PGD_NMICRO) return PGD_ERR_INVALID;
    if (!(b->flags & PGD_MICRO) || b->host_sums.empty()) return PGD_ERR_STATE;
    return batch_guard(b->ctx, [&] {
        if (b->m[gi])
            PGD_HIP(hipMemcpy(out, b->U->micro.get() + (size_t)column * b->M + b->eoff[gi],
                              b->m[gi] * sizeof(uint64_t), hipMemcpyDeviceToHost));
    });
}

pgd_status pgd_batch_info(pgd_batch* b, int32_t* ngraphs, int64_t* total_vertices, int64_t* total_edges) {
    if (!b) return PGD_ERR_INVALID;
    if (ngraphs) *ngraphs = b->ng;
    if (total_vertices) *total_vertices = b->N;
    if (total_edges) *total_edges = b->M;
    return PGD_OK;
}

pgd_status pgd_batch_free(pgd_batch* b) {
    if (!b) return PGD_ERR_INVALID;
    delete b;
    return PGD_OK;
}

pgd_status pgd_batch_set_tuning(pgd_batch* b, const pgd_tuning* t) {
    if (!b || !t) return PGD_ERR_INVALID;
    return batch_guard(b->ctx, [&] {
        const char* why = tuning_problem(*t);
        PGD_CHECK(why[0] == 0, PGD_ERR_INVALID, why);
        b->prm = params_from(*t);
        b->drop_graph();
    });
}

pgd_status pgd_batch_get_tuning(pgd_batch* b, pgd_tuning* t) {
    if (!b || !t) return PGD_ERR_INVALID;
    *t = tuning_from(b->prm);
    return PGD_OK;
}

pgd_status pgd_batch_last_ms(pgd_batch* b, double* ms) {
    if (!b || !ms) return PGD_ERR_INVALID;
    *ms = b->last_ms;
    return PGD_OK;
}

// Coordinate search: starting from the current parameters, each parameter in
// turn takes every value of its list with the others held; the fastest
// (median replay time) is kept before moving to the next parameter.
pgd_status pgd_batch_autotune(pgd_batch* b, int32_t trials, pgd_tuning* best, double* best_ms) {
    if (!b || trials < 1) return PGD_ERR_INVALID;
    return batch_guard(b->ctx, [&] {
        PGD_HIP(hipSetDevice(b->ctx->device));
        std::vector<uint64_t> ref;
        auto measure = [&](const OrientParams& p) {
            b->prm = p;
            b->drop_graph();
            std::vector<double> t;
            for (int i = 0; i < trials; i++) t.push_back(batch_replay(b));
            if (ref.empty()) ref = b->host_sums;
            PGD_CHECK(b->host_sums == ref, PGD_ERR_STATE, "autotuner candidates disagree on the counts");
            std::sort(t.begin(), t.end());
            return t[t.size() / 2];
        };
        OrientParams cur = b->prm;
        double cur_ms = measure(cur);
        struct Axis {
            int OrientParams::*field;
            std::vector<int> values;
        };
        const std::vector<Axis> axes = {
            {&OrientParams::wedge_log2r, {12, 13, 14, 15}}, {&OrientParams::wedge_nt, {256, 512, 1024}},
            {&OrientParams::small_k, {32, 64}},             {&OrientParams::mid_k, {128, 256}},
            {&OrientParams::grid_small, {16, 32}},          {&OrientParams::wedge_split, {2, 4, 8}},
        };
        for (const Axis& ax : axes) {
            for (int v : ax.values) {
                if (cur.*ax.field == v) continue;
                OrientParams cand = cur;
                cand.*ax.field = v;
                if (tuning_problem(tuning_from(cand))[0]) continue;
                const double ms = measure(cand);
                if (ms < cur_ms) {
                    cur_ms = ms;
                    cur = cand;
                }
            }
        }
        b->prm = cur;
        b->drop_graph();
        if (best) *best = tuning_from(cur);
        if (best_ms) *best_ms = cur_ms;
    });
}

}  // extern "C"

// ===========================================================================
// serving: continuous batching over two private streams
// ===========================================================================
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    int32_t n = 0;
    uint32_t flags = 0;
    std::vector<int32_t> src, dst;
    int64_t footprint = 0;
    Clock::time_point t_submit, t_done;
    bool done = false;
    pgd_status status = PGD_OK;
    std::string err;
    int64_t m = 0;
    uint64_t sums[PGD_NSUMS] = {};
    uint64_t totals[PGD_NTOTALS] = {};
    // micro columns when requested: kept on the device in the batch's block
    // (column c at block + c * stride + offset), shared by the batch's requests
    std::shared_ptr<DBuf<uint64_t>> micro;
    int64_t micro_stride = 0, micro_offset = 0;
};

// launch the counting of a built batch without capture (one use)
void batch_launch_eager(pgd_batch* b) {
    if (b->M) {
        if (!b->st) b->st = orient_state_new();
        orient_params(b->st) = b->prm;
        orient_build(b->ctx, union_view(b), *b->st);
    }
    batch_count(b);
}

// wait for the batch's stream; per-request sums and totals to the host, and
// the micro columns handed to the requests on the device
void batch_collect(pgd_batch* b, std::vector<Request*>& reqs) {
    pgd_ctx* ctx = b->ctx;
    hipStream_t s = ctx->stream;
    b->host_sums.resize((size_t)b->ng * PGD_NSUMS);
    PGD_HIP(hipMemcpyAsync(b->host_sums.data(), b->sums.get(), b->host_sums.size() * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, s));
    PGD_HIP(hipStreamSynchronize(s));
    std::shared_ptr<DBuf<uint64_t>> block;
    if ((b->flags & PGD_MICRO) && b->M) {
        block = std::make_shared<DBuf<uint64_t>>();
        std::swap(block->ctx, b->U->micro.ctx);
        std::swap(block->p, b->U->micro.p);
        std::swap(block->n, b->U->micro.n);
    }
    for (int32_t g = 0; g < b->ng; g++) {
        Request* r = reqs[g];
        r->m = b->m[g];
        std::copy(b->host_sums.begin() + (size_t)g * PGD_NSUMS, b->host_sums.begin() + (size_t)(g + 1) * PGD_NSUMS,
                  r->sums);
        check_status(ctx, pgd_derive(b->n[g], b->m[g], r->sums, r->totals), "derive");
        if (block) {
            r->micro = block;
            r->micro_stride = b->M;
            r->micro_offset = b->eoff[g];
        }
    }
}

}  // namespace

// Requests are served by two workers by size class (small graphs; graphs of
// kLargeLines lines or more), so a large graph never holds small requests
// back; each worker alternates two private streams, building one batch while
// the previous one counts.  Both share the HBM budget and the statistics.
// A large graph nearly fills the device by itself, so large requests are not
// gathered into one big union: a batch of them stops at kLargeBatchLines
// lines (at least one request), which keeps the two streams alternating.
// On R-MAT scale-18 requests (2^22 lines each) a bound of two per batch
// served 5620 graphs/s of the 1% mix, against 4681 with one, 4764 with four
// and 2063 with no bound.
constexpr int64_t kLargeLines = (int64_t)1 << 18;
constexpr int64_t kLargeBatchLines = (int64_t)1 << 23;

struct ServerWorker {
    pgd_ctx* lane[2] = {nullptr, nullptr};  // private contexts: one stream each
    std::deque<Request*> pending;
    std::thread th;
};

struct pgd_server {
    pgd_ctx* owner = nullptr;
    ServerWorker cls[2];
    int64_t budget = 0;
    int32_t max_inflight = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<int64_t> unreported;
    int64_t next_id = 0, inflight_bytes = 0, inflight_reqs = 0, outstanding = 0;
    bool stop = false;
    // statistics
    std::vector<double> latencies;
    int64_t finished = 0, edges = 0, batches = 0;
    bool any_submit = false;
    Clock::time_point first_submit, last_done;

    void run(int c);
    // the head of class c's queue can be admitted now (caller holds mu)
    bool head_admissible(int c) const;
    // take queued requests of class c for one batch (caller holds mu)
    std::vector<Request*> admit(int c);
};

bool pgd_server::head_admissible(int c) const {
    const std::deque<Request*>& q = cls[c].pending;
    if (q.empty()) return false;
    if (inflight_reqs == 0) return true;
    return inflight_reqs < max_inflight && inflight_bytes + q.front()->footprint <= budget;
}

std::vector<Request*> pgd_server::admit(int c) {
    std::deque<Request*>& q = cls[c].pending;
    std::vector<Request*> out;
    int64_t bytes = 0, lines = 0;
    while (!q.empty()) {
        Request* r = q.front();
        const bool idle = out.empty() && inflight_reqs == 0;  // the head alone always fits an empty device
        if (!out.empty() && r->flags != out[0]->flags) break;
        if (c == 1 && !out.empty() && lines + (int64_t)r->src.size() > kLargeBatchLines) break;
        if ((int64_t)out.size() + inflight_reqs >= max_inflight && !idle) break;
        if (inflight_bytes + bytes + r->footprint > budget && !idle) break;
        out.push_back(r);
        bytes += r->footprint;
        lines += (int64_t)r->src.size();
        q.pop_front();
    }
    inflight_bytes += bytes;
    inflight_reqs += (int64_t)out.size();
    return out;
}

void pgd_server::run(int c) {
    struct InFlight {
        std::unique_ptr<pgd_batch> b;
        std::vector<Request*> reqs;
        int64_t bytes = 0;
        pgd_status status = PGD_OK;
        std::string err;
    };
    std::unique_ptr<InFlight> prev;
    int turn = 0;
    // build a batch of the given requests on ctx and launch its counting
    auto start = [&](pgd_batch* b, pgd_ctx* ctx, const std::vector<Request*>& rs) {
        b->ctx = ctx;
        b->ng = (int32_t)rs.size();
        b->flags = rs[0]->flags & (PGD_DEDUP | PGD_MICRO);
        return batch_guard(ctx, [&] {
            PGD_HIP(hipSetDevice(ctx->device));
            std::vector<int32_t> n(rs.size());
            std::vector<int64_t> off(rs.size() + 1, 0);
            for (size_t i = 0; i < rs.size(); i++) {
                n[i] = rs[i]->n;
                off[i + 1] = off[i] + (int64_t)rs[i]->src.size();
            }
            std::vector<int32_t> src(std::max<int64_t>(off.back(), 1)), dst(std::max<int64_t>(off.back(), 1));
            for (size_t i = 0; i < rs.size(); i++) {
                std::copy(rs[i]->src.begin(), rs[i]->src.end(), src.begin() + off[i]);
                std::copy(rs[i]->dst.begin(), rs[i]->dst.end(), dst.begin() + off[i]);
            }
            batch_build(b, b->ng, n.data(), off.data(), src.data(), dst.data());
            batch_launch_eager(b);
        });
    };
    auto finish = [&](InFlight& f) {
        if (f.status == PGD_OK && f.b) {
            f.status = batch_guard(f.b->ctx, [&] { batch_collect(f.b.get(), f.reqs); });
            if (f.status != PGD_OK) f.err = f.b->ctx->last_err;
        }
        pgd_ctx* lctx = f.b ? f.b->ctx : cls[c].lane[0];
        if (f.b) (void)hipStreamSynchronize(lctx->stream);  // nothing of it still runs
        f.b.reset();
        // a failed batch of several requests: count each alone, so that one
        // request's failure (an invalid graph) does not fail the others
        std::vector<pgd_status> own(f.reqs.size(), f.status);
        std::vector<std::string> own_err(f.reqs.size(), f.err);
        if (f.status != PGD_OK && f.reqs.size() > 1) {
            for (size_t i = 0; i < f.reqs.size(); i++) {
                std::vector<Request*> one{f.reqs[i]};
                std::unique_ptr<pgd_batch> b1(new pgd_batch());
                own[i] = start(b1.get(), lctx, one);
                if (own[i] == PGD_OK) own[i] = batch_guard(lctx, [&] { batch_collect(b1.get(), one); });
                own_err[i] = own[i] == PGD_OK ? std::string() : lctx->last_err;
                (void)hipStreamSynchronize(lctx->stream);
            }
        }
        const Clock::time_point now = Clock::now();
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i = 0; i < f.reqs.size(); i++) {
            Request* r = f.reqs[i];
            r->status = own[i];
            r->err = own_err[i];
            r->t_done = now;
            r->done = true;
            r->src.clear();
            r->src.shrink_to_fit();
            r->dst.clear();
            r->dst.shrink_to_fit();
            unreported.push_back(r->id);
            if (r->status == PGD_OK) {
                latencies.push_back(std::chrono::duration<double, std::milli>(now - r->t_submit).count());
                finished++;
                edges += r->m;
            }
            last_done = now;
            outstanding--;
        }
        inflight_bytes -= f.bytes;
        inflight_reqs -= (int64_t)f.reqs.size();
        cv_done.notify_all();
        cv_work.notify_all();  // the budget freed may admit the other class's head
    };
    for (;;) {
        std::vector<Request*> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (!prev) cv_work.wait(lk, [&] { return (stop && cls[c].pending.empty()) || head_admissible(c); });
            if (stop && cls[c].pending.empty() && !prev) return;
            batch = admit(c);
        }
        std::unique_ptr<InFlight> cur;
        if (!batch.empty()) {
            cur.reset(new InFlight());
            cur->reqs = batch;
            for (Request* r : batch) cur->bytes += r->footprint;
            pgd_ctx* ctx = cls[c].lane[turn];
            turn ^= 1;
            cur->b.reset(new pgd_batch());
            cur->status = start(cur->b.get(), ctx, batch);  // counting runs while the previous batch is collected
            if (cur->status != PGD_OK) cur->err = ctx->last_err;
            std::lock_guard<std::mutex> lk(mu);
            batches++;
        }
        if (prev) finish(*prev);
        prev = std::move(cur);
    }
}

extern "C" {

int64_t pgd_server_footprint(int32_t n, int64_t nlines) {
    if (n < 0 || nlines < 0) return -1;
    return 8 * nlines + 8 * ((int64_t)n + 1) + 64 * nlines + 4096;
}

pgd_status pgd_server_create(pgd_ctx* ctx, int64_t hbm_budget_bytes, int32_t max_inflight, pgd_server** out) {
    if (!ctx || !out || hbm_budget_bytes <= 0 || max_inflight < 1) return PGD_ERR_INVALID;
    *out = nullptr;
    std::unique_ptr<pgd_server> s(new (std::nothrow) pgd_server());
    if (!s) return PGD_ERR_OOM;
    s->owner = ctx;
    s->budget = hbm_budget_bytes;
    s->max_inflight = max_inflight;
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < 2; i++) {
            pgd_status st = pgd_open(ctx->device, &s->cls[c].lane[i]);
            if (st != PGD_OK) {
                ctx->last_err = "cannot open a private stream for the server";
                for (int c2 = 0; c2 < 2; c2++)
                    for (int j = 0; j < 2; j++)
                        if (s->cls[c2].lane[j]) pgd_close(s->cls[c2].lane[j]);
                return st;
            }
            pgd_set_mem_cap(s->cls[c].lane[i], ctx->mem_cap);
        }
    pgd_server* p = s.get();
    try {
        for (int c = 0; c < 2; c++) s->cls[c].th = std::thread([p, c] { p->run(c); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (int c = 0; c < 2; c++)
            if (s->cls[c].th.joinable()) s->cls[c].th.join();
        for (int c = 0; c < 2; c++)
            for (int i = 0; i < 2; i++) pgd_close(s->cls[c].lane[i]);
        return PGD_ERR_OOM;
    }
    *out = s.release();
    return PGD_OK;
}

pgd_status pgd_server_submit(pgd_server* s, int32_t n, int64_t nlines, const int32_t* src, const int32_t* dst,
                             uint32_t flags, int64_t* req_id) {
    if (!s || !req_id || n < 0 || nlines < 0 || (nlines > 0 && (!src || !dst))) return PGD_ERR_INVALID;
    const int64_t fp = pgd_server_footprint(n, nlines);
    if (fp > s->budget) return PGD_ERR_BUSY;
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return PGD_ERR_OOM;
    try {
        r->src.assign(src, src + nlines);
        r->dst.assign(dst, dst + nlines);
    } catch (const std::bad_alloc&) {
        return PGD_ERR_OOM;
    }
    r->n = n;
    r->flags = flags & (PGD_DEDUP | PGD_MICRO);
    r->footprint = fp;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        *req_id = r->id;
        if (!s->any_submit) {
            s->any_submit = true;
            s->first_submit = r->t_submit;
        }
        s->cls[nlines >= kLargeLines ? 1 : 0].pending.push_back(r.get());
        s->outstanding++;
        s->reqs.emplace(r->id, std::move(r));
    }
    s->cv_work.notify_all();
    return PGD_OK;
}

pgd_status pgd_server_poll(pgd_server* s, int64_t* done, int32_t cap, int32_t* ndone) {
    if (!s || !ndone || cap < 0 || (cap > 0 && !done)) return PGD_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    const int32_t k = (int32_t)std::min<size_t>((size_t)cap, s->unreported.size());
    std::copy(s->unreported.begin(), s->unreported.begin() + k, done);
    s->unreported.erase(s->unreported.begin(), s->unreported.begin() + k);
    *ndone = k;
    return PGD_OK;
}

pgd_status pgd_server_drain(pgd_server* s, int64_t timeout_ms) {
    if (!s) return PGD_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto idle = [&] { return s->outstanding == 0; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, idle);
        return PGD_OK;
    }
    return s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), idle) ? PGD_OK : PGD_ERR_BUSY;
}

pgd_status pgd_server_result(pgd_server* s, int64_t req_id, uint64_t* sums, uint64_t* totals, int64_t* m,
                             double* latency_ms) {
    if (!s) return PGD_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return PGD_ERR_INVALID;
    Request* r = it->second.get();
    if (!r->done) return PGD_ERR_STATE;
    if (r->status != PGD_OK) {
        s->owner->last_err = r->err;
        return r->status;
    }
    if (sums) std::copy(r->sums, r->sums + PGD_NSUMS, sums);
    if (totals) std::copy(r->totals, r->totals + PGD_NTOTALS, totals);
    if (m) *m = r->m;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    return PGD_OK;
}

pgd_status pgd_server_micro(pgd_server* s, int64_t req_id, int32_t column, uint64_t* out) {
    if (!s || !out || column < 0 || column >= PGD_NMICRO) return PGD_ERR_INVALID;
    std::shared_ptr<DBuf<uint64_t>> block;
    int64_t off = 0, m = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(req_id);
        if (it == s->reqs.end()) return PGD_ERR_INVALID;
        Request* r = it->second.get();
        if (!r->done || r->status != PGD_OK || !(r->flags & PGD_MICRO)) return PGD_ERR_STATE;
        block = r->micro;
        off = column * r->micro_stride + r->micro_offset;
        m = r->m;
    }
    if (!m) return PGD_OK;
    hipError_t e = hipMemcpy(out, block->get() + off, m * sizeof(uint64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        s->owner->last_err = std::string("hipMemcpy: ") + hipGetErrorString(e);
        return PGD_ERR_HIP;
    }
    return PGD_OK;
}

pgd_status pgd_server_release(pgd_server* s, int64_t req_id) {
    if (!s) return PGD_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req_id);
    if (it == s->reqs.end()) return PGD_ERR_INVALID;
    if (!it->second->done) return PGD_ERR_STATE;
    s->reqs.erase(it);
    return PGD_OK;
}

pgd_status pgd_server_stats(pgd_server* s, int64_t* finished, int64_t* edges, double* elapsed_s, double* p50_ms,
                            double* p99_ms, int64_t* batches) {
    if (!s) return PGD_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<double> l = s->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
        if (l.empty()) return 0.0;
        size_t i = (size_t)std::min<double>((double)l.size() - 1, std::ceil(q * (double)l.size()) - 1);
        return l[i];
    };
    if (finished) *finished = s->finished;
    if (edges) *edges = s->edges;
    if (elapsed_s)
        *elapsed_s = s->finished ? std::chrono::duration<double>(s->last_done - s->first_submit).count() : 0.0;
    if (p50_ms) *p50_ms = pct(0.50);
    if (p99_ms) *p99_ms = pct(0.99);
    if (batches) *batches = s->batches;
    return PGD_OK;
}

pgd_status pgd_server_free(pgd_server* s) {
    if (!s) return PGD_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (int c = 0; c < 2; c++)
        if (s->cls[c].th.joinable()) s->cls[c].th.join();
    s->reqs.clear();
    for (int c = 0; c < 2; c++)
        for (int i = 0; i < 2; i++)
            if (s->cls[c].lane[i]) pgd_close(s->cls[c].lane[i]);
    delete s;
    return PGD_OK;
}

}  // extern "C"
