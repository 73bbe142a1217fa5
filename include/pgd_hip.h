/*
 * pgd_hip.h -- C ABI of the MI355X implementation of PGD's exact edge-centric
 * 2/3/4-node graphlet decomposition.
 *
 * Plain C types only: pointers, sizes, status codes.  Host buffers belong to
 * the caller; device memory belongs to the context.  Calls on one context are
 * serialised on that context's HIP stream; different contexts are
 * independent.  Errors are returned as pgd_status codes and never cross the
 * ABI as exceptions.
 *
 * Each entry point names the reference member function it replaces
 * (graphlet_core.{h,cpp} of the reference, file:line).
 */
#ifndef PGD_HIP_H
#define PGD_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pgd_status;
#define PGD_OK               0
#define PGD_ERR_INVALID     -1  /* bad argument: null pointer, out-of-range id, unknown name */
#define PGD_ERR_NO_DEVICE   -2  /* no HIP device, or the device index is out of range */
#define PGD_ERR_HIP         -3  /* a HIP runtime call failed */
#define PGD_ERR_OOM         -4  /* device allocation failed, or the context's memory cap was hit */
#define PGD_ERR_UNSUPPORTED -5  /* input outside what the device path supports (e.g. n >= 2^31) */
#define PGD_ERR_STATE       -6  /* call order: e.g. counting before the edge order exists */
#define PGD_ERR_BUSY        -7  /* serving queue full / request not finished */

typedef struct pgd_ctx pgd_ctx;
typedef struct pgd_graph pgd_graph;

/* Number of per-edge sums the decomposition reduces, in this order:
 *   0 sum T  1 sum S  2 sum D  3 sum K4  4 sum C4  5..14 N1..N10
 * (graphlet_core.cpp:1325-1364 and the reduction at 1429-1448). */
#define PGD_NSUMS   15
/* The 17 graphlet totals, in the reference's printed order
 * (graphlet_core.cpp:2061-2084): 2_1edge 2_indep 3_tris 2_star 3_1edge
 * 3_indep 4_clique 4_chordcycle 4_tailed_tris 4_cycle 3_star 4_path 4_1edge
 * 4_2edge 4_2star 4_tri 4_indep. */
#define PGD_NTOTALS 17
/* Per-edge (micro) columns, stored at E_ordered position e
 * (graphlet_core.cpp:1498-1504, 1559-1575). */
#define PGD_MICRO_TRI       0  /* tri */
#define PGD_MICRO_CLIQUE    1  /* local_4_clique */
#define PGD_MICRO_CHORDAL   2  /* local_4_chordal_cycle */
#define PGD_MICRO_TAILED    3  /* local_4_tailed_tris (never written: 0) */
#define PGD_MICRO_CYCLE     4  /* local_4_cycle */
#define PGD_MICRO_STAR3     5  /* local_3_star */
#define PGD_MICRO_PATH      6  /* local_4_path */
#define PGD_MICRO_STAR2     7  /* get_2_star (graphlet_core.h:353-355) */
#define PGD_NMICRO          8

/* graph construction flags */
#define PGD_DEDUP   1u  /* drop repeated undirected pairs, keeping the first
                           occurrence (remove_multiple_edges, graphlet_core.cpp:530-556) */

/* decomposition flags */
#define PGD_MICRO        1u   /* also produce the per-edge columns */
/* 2u is not a flag: any bit other than those below is PGD_ERR_INVALID */
#define PGD_ENGINE_ORIENT 4u  /* force the degree-oriented engine */
#define PGD_ENGINE_DENSE 8u   /* force the dense (int8 and fp4 MFMA) engine */
#define PGD_DENSE_MAX_N 8192  /* the dense engine's largest vertex count */

/* ---- context -------------------------------------------------------- */
pgd_status  pgd_open(int32_t device, pgd_ctx **out);
pgd_status  pgd_close(pgd_ctx *ctx);
const char *pgd_status_str(pgd_status s);
/* Cap on device bytes this context may allocate (0 = no cap). */
pgd_status  pgd_set_mem_cap(pgd_ctx *ctx, int64_t bytes);
pgd_status  pgd_mem_in_use(pgd_ctx *ctx, int64_t *bytes);
/* Last error detail for this context (static storage owned by the context). */
const char *pgd_last_error(pgd_ctx *ctx);
/* The device's rate of global atomic adds, for the roofline of the passes
 * whose per-slot sums are atomics (no counterpart in the reference: a
 * measurement helper).  `words` 32-bit counters in device memory each get
 * one atomic add per launch (consecutive lanes on consecutive words, so a
 * wavefront's adds form 64-byte requests of 16 words), `reps` launches are
 * timed on the context's stream after one untimed launch, and every counter
 * is checked to equal reps + 1; *requests_per_s = reps * words / 16 /
 * seconds.  words in
 * [2^20, 2^31), reps in [1, 100]. */
pgd_status  pgd_atomic_rate(pgd_ctx *ctx, int64_t words, int32_t reps, double *requests_per_s,
                            double *ms_per_launch);

/* ---- graph construction -----------------------------------------------
 * Replaces the CSR assembly of read_mtx (graphlet_core.cpp:509-517) and
 * read_edge_list (446-452), and remove_multiple_edges (530-556) with
 * PGD_DEDUP.  Input: `nlines` undirected pairs (src[i], dst[i]) with 0-based
 * ids in [0, n), in the order the file lists them; self loops are dropped.
 * Each vertex's adjacency list is built in line order, as the reference's
 * reader appends.  n must be the reader's vertex count (1 + largest id for
 * .mtx; number of distinct endpoints for edge lists). */
pgd_status pgd_graph_from_coo(pgd_ctx *ctx, int64_t n, int64_t nlines,
                              const int32_t *src, const int32_t *dst,
                              uint32_t flags, pgd_graph **out);
/* Upload a CSR as-is: `row` (n+1 offsets) and `col` (row[n] ids), the
 * symmetric simple adjacency held by graphlet_core's `vertices`/`edges`
 * (graphlet_core.h:84-86) after its reader and, if already run, its neighbour
 * ordering.  Neighbour order is kept, so a later pgd_preprocess with a method
 * that orders nothing ("natural") reproduces the object's edge ids.
 * Checked before use: row[0] == 0, offsets non-decreasing, ids in [0, n), no
 * self loop and a symmetric adjacency (PGD_ERR_INVALID otherwise); a repeated
 * neighbour (the multi-edges read_mtx keeps) is PGD_ERR_UNSUPPORTED. */
pgd_status pgd_graph_from_csr(pgd_ctx *ctx, int64_t n, const int64_t *row, const int32_t *col,
                              pgd_graph **out);
/* read_graph (graphlet_core.cpp:259-283) straight to a device graph: the
 * same result as pgd_read_graph followed by pgd_graph_from_coo (PGD_DEDUP
 * for edge lists), with the text parsed by `threads` host threads (0: all
 * hardware threads) and, for edge lists, the renumbering by first appearance
 * (graphlet_core.cpp:413-428) done on the device.  self_loops (may be NULL)
 * receives the number of self-loop lines dropped.  PGD_ERR_INVALID for a
 * missing file or an id the reference reader cannot handle. */
pgd_status pgd_graph_from_file(pgd_ctx *ctx, const char *path, int32_t threads, pgd_graph **out,
                               int64_t *self_loops);
pgd_status pgd_graph_free(pgd_graph *g);
/* n = num_vertices(), m = num_edges(), max degree, avg degree
 * (graphlet_core.h:220-221, 262, 276; vertex_degrees graphlet_core.cpp:877-891) */
pgd_status pgd_graph_info(pgd_graph *g, int64_t *n, int64_t *m, int32_t *max_deg, int32_t *avg_deg);
/* Download the CSR: row holds n+1, col holds 2m (get_vertices/get_edges). */
pgd_status pgd_graph_csr(pgd_graph *g, int64_t *row, int32_t *col);

/* ---- preprocessing ----------------------------------------------------
 * order_vertex_neighbors + create_edge_list_arrays (preprocess,
 * graphlet_core.cpp:77-93, 929-949, 1846-1953).  method: "deg"/"degree",
 * "kcore", "degree_kcore", "degree_vol", "kcore_vol", "degree_kcore_vol";
 * any other name keeps the input order. */
pgd_status pgd_preprocess(pgd_graph *g, const char *neigh_method, int32_t small_to_large);
/* e_v, e_u: m ints each, indexed by edge id. */
pgd_status pgd_edge_arrays(pgd_graph *g, int32_t *e_v, int32_t *e_u);

/* compute_cores (graphlet_core.cpp:958-1018): core numbers (n ints, may be
 * NULL) and the printed K (maximum core over vertices 0..n-2). */
pgd_status pgd_compute_cores(pgd_graph *g, int32_t *core, int32_t *max_core);

/* sort_edges (graphlet_core.cpp:854-867): builds E_ordered on the device.
 * ordering: "natural", "deg"/"degree", "kcore", "kcore_deg", "deg_vol",
 * "kcore_vol", "kcore_deg_vol", or any other name (key = degree of e_u, as
 * get_ordering_value's default).  "rand"/"var" are PGD_ERR_UNSUPPORTED
 * (their reference keys are time-seeded / undefined). */
pgd_status pgd_sort_edges(pgd_graph *g, const char *ordering, int32_t small_to_large);
/* sort_edges over a sampled edge set E_s (graphlet_core.cpp:836-851):
 * E_ordered becomes the `count` edge ids `ids` sorted by the same keys, except
 * that the "natural" key is the position in E_s.  Every id must be in [0, m)
 * (PGD_ERR_INVALID otherwise); ids may repeat, as in the reference.  The
 * decomposition then runs over these count positions only
 * (graphlet_core.cpp:1380, 1513: loops to E_ordered.size()), with n and m
 * still the graph's; the micro columns hold m values whose positions past
 * count stay 0. */
pgd_status pgd_sort_edges_subset(pgd_graph *g, const char *ordering, int32_t small_to_large,
                                 const int32_t *ids, int64_t count);
/* Adopt a caller's E_ordered (graphlet_core.h:207, `vector<Vertex>
 * E_ordered`, its ids in order): `count` edge ids, each in [0, m).  Replaces
 * the order pgd_sort_edges would build, so an object whose E_ordered came from
 * any ordering (a time-seeded "rand", a sampled E_s, or the duplicates a
 * second sort_edges call appends) is decomposed in exactly that order.  A
 * later pgd_decompose with PGD_MICRO needs count <= m (the reference's
 * per-edge vectors hold m positions): PGD_ERR_INVALID otherwise. */
pgd_status pgd_set_e_ordered(pgd_graph *g, const int32_t *ids, int64_t count);
/* Number of E_ordered positions (E_ordered.size()). */
pgd_status pgd_e_ordered_count(pgd_graph *g, int64_t *count);
/* E_ordered edge ids (pgd_e_ordered_count ints). */
pgd_status pgd_e_ordered(pgd_graph *g, int32_t *ids);

/* ---- decomposition ----------------------------------------------------
 * graphlet_decomposition / graphlet_decomposition_micro
 * (graphlet_core.cpp:1366-1474, 1483-1625): the per-edge counting and the
 * reduction of the 15 sums.  Counting is exact uint64 integer arithmetic.
 * sums receives PGD_NSUMS values.  With PGD_MICRO the per-edge columns are
 * kept on the device (fetch with pgd_micro). */
pgd_status pgd_decompose(pgd_graph *g, uint32_t flags, uint64_t *sums);
/* Allocation-free decomposition of a small graph.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same result as pgd_decompose(g, flags | PGD_ENGINE_DENSE, ...) -- the
 * dense matrix-core engine and the epilogue -- with every intermediate array
 * (adjacency, products, per-edge counts) in one static device array of
 * PGD_NOALLOC_SCRATCH_BYTES per GPU, whose address is taken once per device
 * at pgd_open.  It calls no allocating routine.
 *   g:       after pgd_preprocess and an edge order (pgd_sort_edges,
 *            pgd_sort_edges_subset or pgd_set_e_ordered); at most
 *            PGD_NOALLOC_MAX_N vertices, else PGD_ERR_UNSUPPORTED
 *   flags:   0 or PGD_MICRO (anything else is PGD_ERR_INVALID)
 *   d_sums:  device memory for PGD_NSUMS uint64 (cleared, then written)
 *   d_micro: with PGD_MICRO, device memory for PGD_NMICRO x m uint64, column
 *            c at d_micro + c*m, in E_ordered order as pgd_micro returns it
 * Returns after the results are in place (the context's stream is
 * synchronised).  Every call on one GPU shares that GPU's scratch array,
 * which keeps the last call's contents: calls that overlap on one GPU (from
 * several threads or contexts) race and must be serialised by the caller.
 * The graph's own per-edge buffers (pgd_micro, pgd_edge_counts) are not
 * touched. */
#define PGD_NOALLOC_MAX_N 1024
#define PGD_NOALLOC_SCRATCH_BYTES 29480960  /* 28.1 MiB per GPU, held until the process ends */
pgd_status pgd_decompose_noalloc(pgd_graph *g, uint32_t flags, uint64_t *d_sums, uint64_t *d_micro);
/* Download one micro column (m uint64 values in E_ordered order). */
pgd_status pgd_micro(pgd_graph *g, int32_t column, uint64_t *out);
/* The per-edge table as text, formatted on the device and written to the
 * file descriptor fd (write_micro_stats / get_micro_stats,
 * graphlet_core.cpp:1668-1761, 1780-1827): `prefix` (the file's "% ", or
 * "" for the stdout dump), the column header, then one row per E_ordered
 * position:  [e_v+1,e_u+1,]tri,2-star,4-clique,4-chordal-cycle,
 * 4-tailed-triangle,4-cycle,3-star,4-path  (the ids with output_id).
 * Needs the micro columns of the last pgd_decompose (PGD_ERR_STATE
 * otherwise).  *bytes (may be NULL) receives the bytes written; a failed
 * write(2) is PGD_ERR_INVALID.  Device memory: about 450 MB of chunk
 * buffers; host: two pinned chunk buffers of about 220 MB. */
pgd_status pgd_write_micro(pgd_graph *g, int32_t fd, int32_t output_id, const char *prefix, int64_t *bytes);
/* Raw per-edge counts by edge id: T, K4 (4-cliques on the edge) and C4
 * (induced 4-cycles on the edge); any pointer may be NULL. */
pgd_status pgd_edge_counts(pgd_graph *g, uint64_t *tri, uint64_t *k4, uint64_t *c4);
/* Which engine the last pgd_decompose used: 2 oriented, 3 dense. */
int32_t    pgd_last_engine(pgd_graph *g);
/* Device time (ms) of the last pgd_decompose's counting kernels, and of its
 * dominant kernel alone, with its launch count (HIP events on the context's
 * stream). */
pgd_status pgd_last_timing(pgd_graph *g, double *total_ms, double *dominant_ms, int32_t *dominant_launches);
/* Per-pass profile of the last pgd_decompose: pass `idx` (0-based) gives its
 * kernel name (NUL-terminated, truncated to cap bytes), device time, the
 * algorithmic HBM bytes of its work model and the work units it processed.
 * PGD_ERR_INVALID once idx is past the last pass.  No reference
 * counterpart: the reference only prints one wall-clock time. */
pgd_status pgd_last_phase(pgd_graph *g, int32_t idx, char *name, int32_t cap, double *ms, double *bytes,
                          double *units);

/* ---- tuning -------------------------------------------------------------
 * Launch parameters of the oriented engine.  No reference counterpart (the
 * reference's -w/-b flags size its OpenMP schedule); results never depend on
 * them, only the time does. */
typedef struct pgd_tuning {
    int32_t small_k;     /* out-degree up to which one wavefront owns a local graph (64) */
    int32_t mid_k;       /* out-degree up to which a 256-thread block does, 1024 threads above (256) */
    int32_t grid_small;  /* wavefront blocks per CU launched for the local passes (64) */
    int32_t grid_mid;    /* 256-thread blocks per CU launched (32) */
    int32_t wedge_log2r; /* endpoint tile of the heavy 4-cycle pass, 2^r 16-bit LDS counters, r in 8..16 (15) */
    int32_t wedge_nt;    /* threads per block of that pass: 256, 512 or 1024 (512) */
    int32_t wedge_split; /* work items per block the heavy wedges are cut into (4) */
    int32_t tri_lists;   /* local-edge lists kept from the triangle pass for the clique pass, in
                            entries per edge (0: off; the clique pass searches again) (64) */
    int32_t tri_search;  /* the triangle pass looks the later members of N+(a) up in the hash table of
                            N+(b) (out-degree above 32) instead of probing N+(b), when tri_search x
                            their number is below |N+(b)| (0: always probe) (4) */
} pgd_tuning;
pgd_status pgd_tuning_default(pgd_tuning *t);
/* Parameters of later pgd_decompose calls on g. */
pgd_status pgd_graph_set_tuning(pgd_graph *g, const pgd_tuning *t);
pgd_status pgd_graph_get_tuning(pgd_graph *g, pgd_tuning *t);

/* ---- derivation -------------------------------------------------------
 * The closed-form derivation of the 17 totals (graphlet_core.cpp:1449-1472,
 * graphlet_core.h:548-557) in integer arithmetic that reproduces the
 * reference build's binary64 steps bit for bit (round-to-nearest-even
 * conversions, multiplication by fl(1/3), fl(1/6), truncation, and the
 * signed/unsigned wrap of the n^3 and n^4 terms).  Host only. */
pgd_status pgd_derive(int64_t n, int64_t m, const uint64_t *sums, uint64_t *totals);

/* ---- batched form -----------------------------------------------------
 * Thousands of small graphs counted together (no reference counterpart).
 * The caller packs the graphs' COO lists back to back (ids local to each
 * graph); graph g owns lines [line_off[g], line_off[g+1]) and has n[g]
 * vertices; flags: PGD_DEDUP, PGD_MICRO.  The batch builds each graph's CSR,
 * preprocessing and E_ordered as the plain path does (deg/deg defaults) on
 * the disjoint union of the graphs; each counting pass is one launch over all
 * graphs (the 4-cycle pass and the epilogue pull their tiles from device
 * queues, the epilogue's tiles being (graph, edge-range) pairs), and the
 * whole counting sequence is captured in a HIP graph that pgd_batch_run
 * replays.  sums receives PGD_NSUMS values per graph and totals PGD_NTOTALS
 * per graph (either may be NULL); each equals the plain path's for that
 * graph alone. */
typedef struct pgd_batch pgd_batch;
pgd_status pgd_batch_create(pgd_ctx *ctx, int32_t ngraphs, const int32_t *n,
                            const int64_t *line_off, const int32_t *src, const int32_t *dst,
                            uint32_t flags, pgd_batch **out);
pgd_status pgd_batch_run(pgd_batch *b, uint64_t *sums, uint64_t *totals);
/* Per-edge micro column of graph gi (m_gi values, E_ordered order). */
pgd_status pgd_batch_micro(pgd_batch *b, int32_t gi, int32_t column, uint64_t *out);
/* Sizes of the batch: graphs, total vertices and edges (after the reader's
 * dedup rules). */
pgd_status pgd_batch_info(pgd_batch *b, int32_t *ngraphs, int64_t *total_vertices, int64_t *total_edges);
pgd_status pgd_batch_free(pgd_batch *b);
/* Launch parameters for later runs (re-plans and re-captures the launch). */
pgd_status pgd_batch_set_tuning(pgd_batch *b, const pgd_tuning *t);
pgd_status pgd_batch_get_tuning(pgd_batch *b, pgd_tuning *t);
/* Time each candidate of a built-in search space on this batch (replays of
 * the captured launch), keep the fastest and return it.  Every candidate's
 * sums are compared with the first one's: PGD_ERR_STATE if any differs.
 * trials: replays per candidate (>= 1). */
pgd_status pgd_batch_autotune(pgd_batch *b, int32_t trials, pgd_tuning *best, double *best_ms);
/* Device time (ms) of the last pgd_batch_run's replay. */
pgd_status pgd_batch_last_ms(pgd_batch *b, double *ms);

/* ---- serving ----------------------------------------------------------
 * A request loop with continuous batching (no reference counterpart).  A
 * worker thread owns two private streams on the context's device.  Requests
 * (one graph each, ids local to the graph) queue in arrival order; the
 * worker admits queued requests while their estimated HBM footprint
 * (pgd_server_footprint) fits what the in-flight batches leave of the budget
 * and fewer than max_inflight requests are in flight, builds them into one
 * batch on one stream while the previous batch counts on the other, and
 * stores each finished request's counts on the host until released.
 * Requests with different flags go to different batches.  Latency is
 * submission to completion, on the host's steady clock. */
typedef struct pgd_server pgd_server;
pgd_status pgd_server_create(pgd_ctx *ctx, int64_t hbm_budget_bytes, int32_t max_inflight, pgd_server **out);
/* Estimated HBM bytes of a request: CSR 8m + 8(n+1), per-edge scratch and
 * outputs 64m (m = nlines), and its share of the tile queues. */
int64_t    pgd_server_footprint(int32_t n, int64_t nlines);
/* Queue a graph (flags: PGD_DEDUP, PGD_MICRO); returns its request id.
 * PGD_ERR_BUSY if its footprint exceeds the whole budget. */
pgd_status pgd_server_submit(pgd_server *s, int32_t n, int64_t nlines, const int32_t *src,
                             const int32_t *dst, uint32_t flags, int64_t *req_id);
/* Ids of requests finished since the last poll (up to cap; *ndone of them).
 * Never blocks. */
pgd_status pgd_server_poll(pgd_server *s, int64_t *done, int32_t cap, int32_t *ndone);
/* Block until every queued request has finished (timeout_ms < 0: no limit;
 * PGD_ERR_BUSY on timeout). */
pgd_status pgd_server_drain(pgd_server *s, int64_t timeout_ms);
/* Result of a finished request: sums (PGD_NSUMS), totals (PGD_NTOTALS), its
 * edge count and latency in ms (any pointer may be NULL).  PGD_ERR_STATE if
 * it has not finished; its own failure status if it failed. */
pgd_status pgd_server_result(pgd_server *s, int64_t req_id, uint64_t *sums, uint64_t *totals, int64_t *m,
                             double *latency_ms);
/* Micro column of a finished PGD_MICRO request (m values, E_ordered order). */
pgd_status pgd_server_micro(pgd_server *s, int64_t req_id, int32_t column, uint64_t *out);
pgd_status pgd_server_release(pgd_server *s, int64_t req_id);
/* Totals since creation: finished requests, their edges, seconds from the
 * first submission to the last completion, latency percentiles (ms), and
 * the number of batches launched. */
pgd_status pgd_server_stats(pgd_server *s, int64_t *finished, int64_t *edges, double *elapsed_s, double *p50_ms,
                            double *p99_ms, int64_t *batches);
pgd_status pgd_server_free(pgd_server *s);

/* ---- synthetic inputs (host) ------------------------------------------ */
int64_t pgd_gen_gnp(int32_t n, double p, uint64_t seed, int32_t *src, int32_t *dst, int64_t cap);
int64_t pgd_gen_ba(int32_t n, int32_t m, uint64_t seed, int32_t *src, int32_t *dst, int64_t cap);
int64_t pgd_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed,
                     int32_t *src, int32_t *dst, int64_t cap);
int32_t pgd_write_mtx(const char *path, int32_t n, int64_t m, const int32_t *src, const int32_t *dst);

/* ---- file reader (host) -----------------------------------------------
 * read_graph (graphlet_core.cpp:259-283): parses a .mtx or edge-list file
 * with the reference reader's rules and returns the COO in line order plus
 * the reader's vertex count; *src/*dst are malloc'd (free with pgd_free).
 * Returns PGD_ERR_INVALID for a missing file or an id the reference reader
 * cannot handle.  *dedup is set when the reference would remove duplicates
 * (edge lists). */
pgd_status pgd_read_graph(const char *path, int64_t *n, int64_t *nlines, int32_t **src, int32_t **dst,
                          int32_t *dedup, int64_t *self_loops);
void       pgd_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* PGD_HIP_H */
