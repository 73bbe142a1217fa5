/*
 * pgd_oracle.h -- CPU oracle for PGD's exact 2/3/4-node graphlet decomposition.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a clean-room C restatement of the
 * reference's CPU path, used as the checker for the HIP implementation and as
 * the timed host-core baseline.  Only tests/, the smoke check and the bench's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).
 *
 * Parity pinning (what checks each part independently of this file):
 *   or_read_graph      : compiled reference (-m output ids/rows) on fixture and
 *                        hand-made files; sample_graph.csv / DD21.mtx globals.
 *   or_order_neighbors,
 *   or_edge_arrays,
 *   or_sort_edges      : compiled reference .micro row order (src,dst columns).
 *   or_kcore           : brute-force peeling on small graphs; compiled
 *                        reference "K =" line.
 *   or_decompose       : brute-force per-edge sets on small graphs;
 *                        sample_graph.csv per-edge columns; compiled reference.
 *   or_derive          : brute-force 4-subset enumeration (n <= 40); closed
 *                        forms; compiled reference on large-n graphs where the
 *                        uint64 / signed wrap and double rounding occur.
 */
#ifndef PGD_ORACLE_H
#define PGD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of per-edge sums reduced by the decomposition loop, in this order:
 *   0 sum T, 1 sum S(=Sv+Su), 2 sum D, 3 sum K4, 4 sum C4, 5..14 = N1..N10
 * (graphlet_core.cpp:1325-1364, 1431-1448). */
#define OR_NSUMS 15
/* The 17 totals, in the printed order (graphlet_core.cpp:2061-2084). */
#define OR_NTOTALS 17

typedef struct {
    int64_t  n;        /* vertices.size()-1 */
    int64_t  nnz;      /* edges.size() = 2m */
    int64_t *row;      /* n+1 offsets (vertices) */
    int32_t *col;      /* nnz neighbour ids (edges) */
} or_csr;

/* Reader (graphlet_core.cpp:259-283, 291-457, 459-522, 530-556).
 * Returns 0 on success, <0 on an input the reference cannot process
 * (missing file: -1; negative or out-of-table vertex id: -2). */
int  or_read_graph(const char *path, or_csr *out, int64_t *self_loops);
void or_free(or_csr *g);
/* Build a CSR from an in-memory directed adjacency (used by tests). */
int  or_csr_from_arrays(int64_t n, int64_t nnz, const int64_t *row, const int32_t *col, or_csr *out);

/* Degrees, max degree, avg degree (graphlet_core.cpp:877-891). */
void or_degrees(const or_csr *g, int32_t *deg, int32_t *max_deg, int32_t *avg_deg);

/* k-core numbers of every vertex (0-based), and the printed "K" value
 * (graphlet_core.cpp:958-1018: the 1-based arrays make it the maximum core
 * over vertices 0..n-2).  core must hold n ints. */
int32_t or_kcore(const or_csr *g, int32_t *core);

/* Per-vertex neighbour ordering (graphlet_core.cpp:1846-1877, 1910-1953).
 * method: "deg"/"degree", "kcore", "degree_kcore"/"deg_kcore",
 * "degree_vol"/"deg_vol", "kcore_vol", "degree_kcore_vol"/"deg_kcore_vol";
 * anything else leaves the input order.  Permutes g->col in place. */
void or_order_neighbors(or_csr *g, const char *method, int small_to_large);

/* Edge arrays (graphlet_core.cpp:929-949).  Returns m; e_v/e_u hold m ints. */
int64_t or_edge_arrays(const or_csr *g, int32_t *e_v, int32_t *e_u);

/* Edge processing order (graphlet_core.cpp:686-867; graphlet_vertex.h:39-76).
 * Returns 0, or -1 for an ordering whose key the reference leaves undefined
 * ("rand", "var").  e_ordered holds m edge ids. */
int or_sort_edges(const or_csr *g, const int32_t *e_v, const int32_t *e_u, int64_t m,
                  const char *ordering, int small_to_large, int32_t *e_ordered);
/* sort_edges over a sampled edge set E_s (graphlet_core.cpp:836-851):
 * E_ordered of the count ids, natural key = position in E_s. */
int or_sort_edges_subset(const or_csr *g, const int32_t *e_v, const int32_t *e_u, const int32_t *ids,
                         int64_t count, const char *o, int small_to_large, int32_t *e_ordered);

/* The edge-parallel decomposition over positions [0,count) of `order`
 * (graphlet_core.cpp:1366-1448 macro, 1483-1599 micro, CSR branch).
 * sums[OR_NSUMS] receives the reduced sums.  When the per-edge pointers are
 * non-NULL they receive values at position p (E_ordered indexing, quirk
 * graphlet_core.cpp:1518): T, K4, C4 (raw) and the micro columns chordal,
 * 4-path, 3-star.  threads<=0 uses all OpenMP threads. */
void or_decompose(const or_csr *g, const int32_t *e_v, const int32_t *e_u,
                  const int32_t *order, int64_t count, int64_t m_total, int threads,
                  uint64_t *sums,
                  uint64_t *tri, uint64_t *k4, uint64_t *c4,
                  uint64_t *chordal, uint64_t *path4, uint64_t *star3);

/* Closed-form derivation of the 17 totals (graphlet_core.cpp:1449-1472 and
 * graphlet_core.h:548-557), restated with the floating-point operations the
 * reference build performs: x/3.0 and x/6.0 are multiplications by fl(1/3)
 * and fl(1/6) (what g++ -ffast-math emits for Makefile:21-34). */
void or_derive(int64_t n, int64_t m, const uint64_t *sums, uint64_t *totals);

/* Brute-force per-edge sets for tiny graphs: T, Sv, Su, K4, C4 for edge
 * (v,u) by direct set algebra, independent of the perfect-hash loop. */
void or_brute_edge(const or_csr *g, int32_t v, int32_t u, uint64_t out[5]);

#ifdef __cplusplus
}
#endif
#endif
