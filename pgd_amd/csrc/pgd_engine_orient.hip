// pgd_engine_orient.hip -- degree-oriented engine.
//
// Computes the same per-edge counts as the reference's edge loop
// (graphlet_core.cpp:1408-1427: T, K4 = 4-cliques on the edge, C4 = induced
// 4-cycles on the edge) without walking 2-hop neighbourhoods edge by edge.
// Vertices are relabelled by rank (degree, id); every edge points from the
// lower to the higher rank.  Then, for an edge e = (v,u):
//
//   T(e)  = #triangles on e
//   K4(e) = #4-cliques on e
//   P(e)  = #4-cycles (as subgraphs) through e = (A^3)_vu - d_v - d_u + 1
//   Q(e)  = sum over triangles {v,u,w} of T(v,w) + T(u,w)
//   C4(e) = P(e) - Q(e) + 2 T(e) + 2 K4(e)
//
// The last identity splits the 4-cycles v-x-w-u through e by their chords
// (x~u, w~v): no chord is an induced 4-cycle, two chords a 4-clique (each
// counted twice), one chord a diamond; summing |N(w) & (Sv u Su)| over w in
// T(e) gives Q(e) - 2T(e) - 4K4(e) for the diamonds.
//
// Work, per pass:
//   * triangles (T): each triangle {a<b<c} once, from its lowest vertex a,
//     as an edge of a's local graph on N+(a) (the out-neighbours);
//   * 4-cliques and Q: a's local graph as a bit matrix in LDS; the 4-cliques
//     with lowest vertex a are its triangles, found with popcounts;
//   * P: each 4-cycle once, from its highest vertex a, as a pair of wedges
//     a-b-c with b,c < a (Chiba-Nishizeki); edge (a,b) gains cnt_a[c]-1 and
//     edge (b,c) gains cnt_a[c]-1 per wedge.
// With the degree order the out-degree stays small (about 10^3 on R-MAT
// scale 22 against a maximum degree of about 1.6*10^5) and the wedge count
// is O(m * arboricity), orders of magnitude below the reference's per-edge
// 2-hop work.  All counts are exact uint64 (uint32 for T), accumulated with
// atomics, so the result is independent of scheduling.
#include <cmath>
#include <memory>

#include "pgd_orient.hpp"

using namespace pgd;


namespace {

constexpr int kLightBound = 256;  // wedge bound handled by a wavefront with an LDS hash
constexpr int kHashSlots = 512;


// ---------------------------------------------------------------------------
// relabelled, sorted CSR
// ---------------------------------------------------------------------------
// rank key (group, degree, id): the rank order is by degree then id, within
// the vertex's group when groups are given (the graphs of a batch), so each
// group occupies a contiguous rank range
__global__ void k_rank_keys(const int32_t* __restrict__ deg, const int32_t* __restrict__ group, int64_t n, int db,
                            int vb, uint64_t* __restrict__ key) {
    GRID_STRIDE(v, n) {
        uint64_t k = ((uint64_t)(uint32_t)deg[v] << vb) | (uint64_t)v;
        if (group) k |= (uint64_t)(uint32_t)group[v] << (db + vb);
        key[v] = k;
    }
}

__global__ void k_rank_scatter(const uint64_t* __restrict__ sorted, int64_t n, int db, int vb,
                               int32_t* __restrict__ rank, int32_t* __restrict__ rdeg) {
    GRID_STRIDE(i, n) {
        uint64_t k = sorted[i];
        rank[k & ((1ULL << vb) - 1)] = (int32_t)i;
        rdeg[i] = (int32_t)((k >> vb) & ((1ULL << db) - 1));
    }
}

// two slots per edge id: (rank(ev) -> rank(eu)) and the reverse, value = edge id
__global__ void k_slot_keys(const int32_t* __restrict__ ev, const int32_t* __restrict__ eu,
                            const int32_t* __restrict__ rank, int64_t m, int vb, uint64_t* __restrict__ key,
                            int32_t* __restrict__ val) {
    GRID_STRIDE(e, m) {
        uint64_t a = (uint32_t)rank[ev[e]], b = (uint32_t)rank[eu[e]];
        key[2 * e] = (a << vb) | b;
        key[2 * e + 1] = (b << vb) | a;
        val[2 * e] = (int32_t)e;
        val[2 * e + 1] = (int32_t)e;
    }
}

__global__ void k_slot_cols(const uint64_t* __restrict__ key, int64_t nnz, int vb, int32_t* __restrict__ col) {
    GRID_STRIDE(p, nnz) col[p] = (int32_t)(key[p] & ((1ULL << vb) - 1));
}

__global__ void k_row_tail(int64_t* row, int64_t n, int64_t nnz) {
    if (threadIdx.x == 0 && blockIdx.x == 0) row[n] = nnz;
}

// ostart[a]: first slot of N(a) holding a higher rank (lists are ascending)
__global__ void k_out_start(const int64_t* __restrict__ row, const int32_t* __restrict__ col, int64_t n,
                            int64_t* __restrict__ ostart, int32_t* __restrict__ outdeg) {
    GRID_STRIDE(a, n) {
        int64_t lo = row[a], hi = row[a + 1];
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (col[mid] <= a) lo = mid + 1;
            else hi = mid;
        }
        ostart[a] = lo;
        outdeg[a] = (int32_t)(row[a + 1] - lo);
    }
}

// upper bound of the wedges c-b-a with a > max(b, c) anchored at c: the sum
// of |N+(b)| over the neighbours b of c (for b < c only the part of N+(b)
// above c counts, which |N+(b)| bounds)
__global__ void k_wedge_bound(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                              const int32_t* __restrict__ outdeg, int64_t n, int64_t* __restrict__ ub) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t c = wave; c < n; c += nw) {
        int64_t s = 0;
        for (int64_t p = row[c] + lane; p < row[c + 1]; p += 64) s += outdeg[col[p]];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) ub[c] = s;
    }
}

// bin flags: 1 if lo < key <= hi
__global__ void k_bin_flags_i32(const int32_t* __restrict__ key, int64_t n, int64_t lo, int64_t hi,
                                int32_t* __restrict__ f) {
    GRID_STRIDE(a, n) f[a] = key[a] > lo && key[a] <= hi;
}
__global__ void k_bin_flags_i64(const int64_t* __restrict__ key, int64_t n, int64_t lo, int64_t hi,
                                int32_t* __restrict__ f) {
    GRID_STRIDE(a, n) f[a] = key[a] > lo && key[a] <= hi;
}
__global__ void k_bin_scatter(const int32_t* __restrict__ f, const int64_t* __restrict__ off, int64_t n,
                              int32_t* __restrict__ out) {
    GRID_STRIDE(a, n) if (f[a]) out[off[a]] = (int32_t)a;
}


// ---------------------------------------------------------------------------
// local-graph kernels (triangles; 4-cliques + Q).  One block per vertex a of
// the list; the block owns the local graph on N+(a), k = |N+(a)| <= KMAX.
// Storage for the per-vertex arrays is either LDS (kGlobal == false) or a
// per-block slice of HBM scratch (kGlobal == true, for k > KMAX).  These are
// the clique pass's arrays (the triangle pass has its own, TriArrays).
// ---------------------------------------------------------------------------
struct LocalArrays {
    int32_t* L;        // N+(a), ascending          [kmax]
    int32_t* Le;       // edge id of a -> L[i]; then the degree of L[i] [kmax]
    int64_t* obeg;     // first out-slot of L[i]     [kmax]
    int64_t* pref;     // inclusive scan of |N+(L[i])| [kmax]
    uint32_t* cnt;     // per-local-vertex counter   [kmax]
    uint32_t* ta;      // T of the edge a -> L[i]    [kmax]
    uint64_t* acc;     // per-local-vertex u64 accumulator [kmax]
    int32_t* hk;       // hash of N+(a): vertex id (-1 empty) [H]
    int32_t* hv;       //                local index          [H]
    uint32_t* M;       // bit matrix, k rows of up to mw/k words [mw]
    int hmask;
};

// hash slots for kmax local vertices: a power of two, at most half full
__host__ __device__ inline int hash_slots(int kmax) {
    int h = 64;
    while (h < 2 * kmax) h <<= 1;
    return h;
}

// per-vertex arrays of a local graph of at most kmax vertices; the bit
// matrix (clique pass only) follows them
__device__ __forceinline__ LocalArrays carve(char* base, int kmax) {
    LocalArrays r;
    char* p = base;
    const int H = hash_slots(kmax);
    r.pref = (int64_t*)p; p += (size_t)kmax * 8;
    r.obeg = (int64_t*)p; p += (size_t)kmax * 8;
    r.acc = (uint64_t*)p; p += (size_t)kmax * 8;
    r.L = (int32_t*)p; p += (size_t)kmax * 4;
    r.Le = (int32_t*)p; p += (size_t)kmax * 4;
    r.cnt = (uint32_t*)p; p += (size_t)kmax * 4;
    r.ta = (uint32_t*)p; p += (size_t)kmax * 4;
    r.hk = (int32_t*)p; p += (size_t)H * 4;
    r.hv = (int32_t*)p; p += (size_t)H * 4;
    r.M = (uint32_t*)p;
    r.hmask = H - 1;
    return r;
}
// rounded to 256 B so every per-block HBM slice keeps 8-byte alignment for the u64 atomics
__host__ __device__ inline size_t carve_bytes(int kmax, int64_t mw) {
    return ((size_t)kmax * 40 + (size_t)hash_slots(kmax) * 8 + (size_t)mw * 4 + 255) & ~(size_t)255;
}
// bit-matrix words that hold all of a local graph of kmax vertices
__host__ __device__ inline int64_t full_matrix_words(int kmax) { return (int64_t)kmax * ((kmax + 31) / 32); }

__device__ __forceinline__ uint32_t vhash(int32_t c) { return (uint32_t)c * 2654435761u; }

// local index of vertex c in N+(a), or -1
__device__ __forceinline__ int find_local(const LocalArrays& A, int32_t c) {
    uint32_t h = (vhash(c) >> 7) & (uint32_t)A.hmask;
    for (;;) {
        const int32_t key = A.hk[h];
        if (key == c) return A.hv[h];
        if (key < 0) return -1;
        h = (h + 1) & (uint32_t)A.hmask;
    }
}

// load N+(a), its hash and the flattened out-list prefix; returns total local work
template <int NT>
__device__ int64_t load_local(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                              const int32_t* __restrict__ eid, const int64_t* __restrict__ ostart, int32_t a,
                              int k, LocalArrays& A, int64_t* part) {
    const int64_t s0 = ostart[a];
    for (int h = threadIdx.x; h <= A.hmask; h += NT) A.hk[h] = -1;
    for (int i = threadIdx.x; i < k; i += NT) {
        int32_t b = col[s0 + i];
        A.L[i] = b;
        A.Le[i] = eid[s0 + i];
        int64_t ob = ostart[b];
        A.obeg[i] = ob;
        A.pref[i] = row[b + 1] - ob;
        A.cnt[i] = 0;
        A.acc[i] = 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += NT) {
        const int32_t b = A.L[i];
        uint32_t h = (vhash(b) >> 7) & (uint32_t)A.hmask;
        while (atomicCAS(&A.hk[h], -1, b) != -1) h = (h + 1) & (uint32_t)A.hmask;
        A.hv[h] = i;
    }
    __syncthreads();
    block_incl_scan<NT>(A.pref, k, part);
    return A.pref[k - 1];
}

constexpr int kU = 4;  // flattened positions per thread per round (independent searches and loads)
constexpr int kListChunk = 2048;  // kept-list entries a block reserves at a time

// owner i and CSR slot of the flattened positions t0, t0+NT, ... (i = -1 past the end)
template <int NT>
__device__ __forceinline__ void locate(const LocalArrays& A, int k, int64_t total, int64_t t0, int* i,
                                       int64_t* slot) {
#pragma unroll
    for (int u = 0; u < kU; u++) {
        const int64_t t = t0 + (int64_t)NT * u;
        i[u] = -1;
        if (t < total) {
            i[u] = owner_of(A.pref, k, t);
            const int64_t ex = i[u] ? A.pref[i[u] - 1] : 0;
            slot[u] = A.obeg[i[u]] + (t - ex);
        }
    }
}

// Record the hits of one round (kU positions per thread) in the block's
// local edge list: one LDS reservation per wavefront per round.  Entry =
// (out-slot of the edge (L[i], L[j]) << 32) | (j << 16) | i.
// kFull: every lane of the wavefront is active (the scan is then a DPP scan)
template <int KU, bool kFull>
__device__ __forceinline__ void record_hits(uint64_t* __restrict__ el, int* ecount, const int* i, const int* j,
                                            const uint32_t* e) {
    const int lane = threadIdx.x & 63;
    int mine = 0;
#pragma unroll
    for (int u = 0; u < KU; u++) mine += j[u] >= 0;
    // exclusive scan of the lanes' hit counts
    int x, wave_total, leader;
    if (kFull) {
        x = (int)wave_incl_sum((uint32_t)mine);
        wave_total = __shfl(x, 63, 64);
        leader = 0;
    } else {
        x = mine;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        const uint64_t active = __ballot(1);
        const int last = 63 - __clzll((long long)active);
        wave_total = __shfl(x, last, 64);
        leader = __ffsll((unsigned long long)active) - 1;
    }
    if (!wave_total) return;
    int basepos = 0;
    if (lane == leader) basepos = atomicAdd(ecount, wave_total);
    basepos = __shfl(basepos, leader, 64);
    int pos = basepos + x - mine;
#pragma unroll
    for (int u = 0; u < KU; u++)
        if (j[u] >= 0) el[pos++] = ((uint64_t)e[u] << 32) | ((uint32_t)j[u] << 16) | (uint32_t)i[u];
}

// The next local graph of a block: from the bin's work queue, highest rank
// (heaviest) first, when the launch has one; otherwise round-robin by block.
// Returns -1 when the bin is done.  Uniform over the block.
__device__ __forceinline__ int64_t next_vertex(uint32_t* queue, int64_t cnt, int64_t prev, int64_t* sh) {
    if (!queue) {
        const int64_t li = prev < 0 ? (int64_t)blockIdx.x : prev + gridDim.x;
        return li < cnt ? li : -1;
    }
    __syncthreads();  // every thread has read the previous index
    if (threadIdx.x == 0) {
        const uint32_t q = atomicAdd(queue, 1u);
        *sh = (int64_t)q < cnt ? cnt - 1 - (int64_t)q : -1;
    }
    __syncthreads();
    return *sh;
}

// Lists of local edges kept for the clique pass: the triangle pass appends
// vertex a's list (from its block scratch) to `G` when it fits and records
// its offset and length (length -1: not kept, the clique pass searches again).
struct TriLists {
    uint64_t* G;              // all kept lists
    int64_t gcap;             // entries of G
    unsigned long long* cur;  // next free entry
    int64_t* off;             // per vertex rank: first entry of its list
    int32_t* len;             // per vertex rank: entries, or -1
};

// ---------------------------------------------------------------------------
// pass 1: triangles.  Local edge (i,j), i<j <=> triangle {a, L[i], L[j]}.
// Every edge is counted in TS at its out-slot (the slot in the list of its
// lower-ranked end): the far edge (L[i], L[j]) per hit, the edges (a, L[i])
// from the block's counters at the end.  TS is then T by out-slot, as the
// clique pass reads it, and the fold into T by edge id needs no atomics.
//
// The out-lists N+(L[i]) are walked in groups of 64 lists: a wavefront takes
// chunks of 256 consecutive positions of a group's flattened lists (chunks
// of all groups are dealt round-robin to the block's wavefronts).  The list
// holding a position is found from the list starts flagged in LDS (a ballot
// and a popcount), and membership of the endpoint x in N+(a) is first tested
// in a bitmap of hashed members (one LDS word; about 3% false positives) and
// only then looked up in the hash of (member, local index) pairs.
// ---------------------------------------------------------------------------
constexpr int kChunk = 64 * kU;  // positions per wavefront chunk

struct TriArrays {
    int64_t* obeg;   // first out-slot of L[i]                      [kmax]
    int64_t* pref;   // inclusive scan of |N+(L[i])|                [kmax]
    uint64_t* hkv;   // hash of N+(a): (vertex << 32) | local index, ~0 empty [2^hbits]
    int32_t* L;      // N+(a), ascending                            [kmax]
    int32_t* Le;     // edge id of a -> L[i]                        [kmax]
    uint32_t* cnt;   // triangles on the edge a -> L[i]             [kmax]
    uint32_t* bm;    // bitmap of hashed members                    [2^bbits bits]
    int32_t* gch;    // first chunk of each group of 64 lists       [groups + 1]
    int hbits, bbits;
};

__host__ __device__ inline int tri_hash_bits(int kmax) {  // at most half full
    int b = 6;
    while ((1 << b) < 2 * kmax) b++;
    return b;
}
__host__ __device__ inline int tri_bm_bits(int kmax) {  // at least 32 bits per member
    int b = 11;
    while ((1 << b) < 32 * kmax) b++;
    return b;
}
__host__ __device__ inline size_t tri_carve_bytes(int kmax) {
    const size_t groups = (size_t)(kmax + 63) / 64 + 1;
    return ((size_t)kmax * 28 + ((size_t)8 << tri_hash_bits(kmax)) + ((size_t)1 << tri_bm_bits(kmax)) / 8 +
            groups * 4 + 255) & ~(size_t)255;
}
__device__ __forceinline__ TriArrays tri_carve(char* base, int kmax) {
    TriArrays r;
    r.hbits = tri_hash_bits(kmax);
    r.bbits = tri_bm_bits(kmax);
    char* p = base;
    r.obeg = (int64_t*)p; p += (size_t)kmax * 8;
    r.pref = (int64_t*)p; p += (size_t)kmax * 8;
    r.hkv = (uint64_t*)p; p += (size_t)8 << r.hbits;
    r.L = (int32_t*)p; p += (size_t)kmax * 4;
    r.Le = (int32_t*)p; p += (size_t)kmax * 4;
    r.cnt = (uint32_t*)p; p += (size_t)kmax * 4;
    r.bm = (uint32_t*)p; p += ((size_t)1 << r.bbits) / 8;
    r.gch = (int32_t*)p;
    return r;
}

__device__ __forceinline__ uint32_t mhash(int32_t c) { return (uint32_t)c * 0x9E3779B1u; }

// local index of x in N+(a) or -1 (x hashed to h)
__device__ __forceinline__ int tri_find(const TriArrays& A, int32_t x, uint32_t h) {
    const uint32_t mask = (1u << A.hbits) - 1;
    uint32_t s = h >> (32 - A.hbits);
    for (;;) {
        const uint64_t kv = A.hkv[s];
        if (kv == ~0ull) return -1;
        if ((int32_t)(kv >> 32) == x) return (int)(uint32_t)kv;
        s = (s + 1) & mask;
    }
}

// Hash tables of the out-lists N+(b) of the vertices of large out-degree,
// in HBM: entry = (member << 32) | position in N+(b), ~0 empty; off[b] =
// (first entry << 6) | log2(table size), or -1 without a table.  factor: a
// local vertex's later members are looked up in its table when factor x
// their number is below its out-degree (0: never).
struct OutHash {
    const uint64_t* tab;
    const int64_t* off;
    int factor;
};
constexpr int64_t kSearchBit = (int64_t)1 << 62;  // in obeg: the list is searched, not probed
constexpr int kSearchOwner = 1 << 30;             // in the owner index: likewise
constexpr int kOutHashMin = 32;                   // out-degree above which a vertex has a table

// position of x in N+(b), or -1 (h = mhash(x))
__device__ __forceinline__ int32_t out_hash_find(const OutHash& oh, int32_t b, int32_t x, uint32_t h) {
    const int64_t t = oh.off[b];
    const int bits = (int)(t & 63);
    const uint64_t* tab = oh.tab + (t >> 6);
    const uint32_t mask = (1u << bits) - 1;
    uint32_t s = h >> (32 - bits);
    for (;;) {
        const uint64_t kv = tab[s];
        if (kv == ~0ull) return -1;
        if ((int32_t)(kv >> 32) == x) return (int32_t)(uint32_t)kv;
        s = (s + 1) & mask;
    }
}

// table sizes: 2^bits >= 2 out-degree for out-degrees above kOutHashMin
__global__ void k_out_hash_size(const int32_t* __restrict__ outdeg, int64_t n, int32_t* __restrict__ sz) {
    GRID_STRIDE(b, n) {
        const int d = outdeg[b];
        int bits = 0;
        if (d > kOutHashMin)
            while ((1 << bits) < 2 * d) bits++;
        sz[b] = bits ? 1 << bits : 0;
    }
}
__global__ void k_out_hash_off(const int32_t* __restrict__ sz, const int64_t* __restrict__ ex, int64_t n,
                               int64_t* __restrict__ off) {
    GRID_STRIDE(b, n) off[b] = sz[b] ? (ex[b] << 6) | (int64_t)(31 - __clz(sz[b])) : -1;
}
// one wavefront per vertex: insert N+(b) into its table (cleared to ~0)
__global__ void k_out_hash_fill(const int64_t* __restrict__ rrow, const int32_t* __restrict__ rcol,
                                const int64_t* __restrict__ ostart, const int64_t* __restrict__ off, int64_t n,
                                uint64_t* __restrict__ tab) {
    const int lane = threadIdx.x & 63;
    for (int64_t b = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6; b < n;
         b += ((int64_t)gridDim.x * blockDim.x) >> 6) {
        const int64_t t = off[b];
        if (t < 0) continue;
        const int bits = (int)(t & 63);
        uint64_t* tb = tab + (t >> 6);
        const uint32_t mask = (1u << bits) - 1;
        const int64_t s0 = ostart[b], d = rrow[b + 1] - s0;
        for (int64_t p = lane; p < d; p += 64) {
            const int32_t x = rcol[s0 + p];
            const unsigned long long kv = ((unsigned long long)(uint32_t)x << 32) | (uint32_t)p;
            uint32_t s = mhash(x) >> (32 - bits);
            while (atomicCAS((unsigned long long*)&tb[s], ~0ull, kv) != ~0ull) s = (s + 1) & mask;
        }
    }
}

template <int NT, bool kGlobal>
__global__ __launch_bounds__(NT) void k_local_tri(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                                                  const int32_t* __restrict__ eid, const int64_t* __restrict__ ostart,
                                                  const int32_t* __restrict__ outdeg, const int32_t* __restrict__ list,
                                                  int64_t cnt, int kmax, size_t stride, char* __restrict__ gscratch,
                                                  uint64_t* __restrict__ elist, int64_t ecap, TriLists tl,
                                                  uint32_t* __restrict__ TS, uint32_t* __restrict__ T,
                                                  OutHash oh, uint32_t* __restrict__ queue,
                                                  unsigned long long* __restrict__ work) {
    constexpr int NW = NT / 64;
    __shared__ int64_t li_sh;
    extern __shared__ char lds[];
    __shared__ int64_t part[NT];
    __shared__ int ecount;
    __shared__ int64_t gpos;
    __shared__ int64_t wbase[NW][64];   // per wavefront: slot of position q of list r = wbase[r] + q
    __shared__ int32_t wown[NW][64];    //                local index of the r-th non-empty list
    __shared__ uint32_t wflag[NW][64];  //                list starts of the chunk (byte p>>6 of word p&63)
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // stride: bytes of a block's HBM slice (the HBM bin)
    char* base = kGlobal ? gscratch + (size_t)blockIdx.x * stride : lds;
    TriArrays A = tri_carve(base, kmax);
    uint64_t* el = (elist && tl.G) ? elist + (size_t)blockIdx.x * (size_t)ecap : nullptr;
    unsigned long long done = 0;       // local work of this block, added once at the end
    __shared__ int64_t chunk_pos, chunk_end;  // the block's reserved part of the kept lists
    if (tid == 0) chunk_pos = chunk_end = 0;
    const uint64_t upto = (2ULL << lane) - 1;  // lanes 0..lane
    uint8_t* fl = (uint8_t*)wflag[w];
    for (int64_t li = next_vertex(queue, cnt, -1, &li_sh); li >= 0; li = next_vertex(queue, cnt, li, &li_sh)) {
        const int32_t a = list[li];
        const int k = outdeg[a];
        const int ng = (k + 63) / 64;
        if (tid == 0) {
            ecount = 0;
            // LDS bins write the list in place, in the block's reserved part
            // of G, when the part holds k(k-1)/2 (the most local edges):
            // parts of twice that are reserved as needed
            gpos = -1;
            if (!kGlobal && el) {
                const int64_t bound = (int64_t)k * (k - 1) / 2;
                if (chunk_end - chunk_pos < bound) {
                    const int64_t want = max(2 * bound, (int64_t)kListChunk);
                    const int64_t c = (int64_t)atomicAdd(tl.cur, (unsigned long long)want);
                    chunk_pos = c;
                    chunk_end = c + want <= tl.gcap ? c + want : c;  // an empty part once G is full
                }
                if (chunk_end - chunk_pos >= bound) gpos = chunk_pos;
            }
        }
        // N+(a), its hash and bitmap, the flattened out-list prefix
        {
            const int64_t s0 = ostart[a];
            for (int h = tid; h < (1 << A.hbits); h += NT) A.hkv[h] = ~0ull;
            for (int h = tid; h < (1 << A.bbits) / 32; h += NT) A.bm[h] = 0;
            for (int i = tid; i < k; i += NT) {
                const int32_t b = col[s0 + i];
                A.L[i] = b;
                const int64_t ob = ostart[b];
                const int64_t od = row[b + 1] - ob;
                // the later members L[j], j > i, looked up in N+(b)'s table,
                // when that is cheaper than probing all of N+(b) here
                const int64_t later = k - 1 - i;
                const bool srch = oh.factor > 0 && oh.off[b] >= 0 && (int64_t)oh.factor * later < od;
                A.obeg[i] = srch ? (ob | kSearchBit) : ob;
                A.pref[i] = srch ? later : od;
                A.cnt[i] = 0;
            }
            __syncthreads();
            const uint32_t mask = (1u << A.hbits) - 1;
            for (int i = tid; i < k; i += NT) {
                const int32_t b = A.L[i];
                const uint32_t h = mhash(b);
                atomicOr(&A.bm[h >> (32 - A.bbits + 5) ], 1u << ((h >> (32 - A.bbits)) & 31));
                const unsigned long long kv = ((unsigned long long)(uint32_t)b << 32) | (uint32_t)i;
                uint32_t sl = h >> (32 - A.hbits);
                while (atomicCAS((unsigned long long*)&A.hkv[sl], ~0ull, kv) != ~0ull) sl = (sl + 1) & mask;
            }
            __syncthreads();
            block_incl_scan<NT>(A.pref, k, part);
            // chunks per group of 64 lists, then their exclusive prefix (one wavefront)
            if (w == 0) {
                int carry = 0;
                for (int g0 = 0; g0 < ng; g0 += 64) {
                    const int g = g0 + lane;
                    int nc = 0;
                    if (g < ng) {
                        const int64_t lo = g ? A.pref[64 * g - 1] : 0;
                        const int64_t hi = A.pref[min(64 * g + 64, k) - 1];
                        nc = (int)((hi - lo + kChunk - 1) / kChunk);
                    }
                    int x = nc;
#pragma unroll
                    for (int off = 1; off < 64; off <<= 1) {
                        const int y = __shfl_up(x, off, 64);
                        if (lane >= off) x += y;
                    }
                    if (g < ng) A.gch[g] = carry + x - nc;
                    carry += __shfl(x, 63, 64);
                }
                if (lane == 0) A.gch[ng] = carry;
            }
            __syncthreads();
        }
        const int64_t total = A.pref[k - 1];
        if (tid == 0) done += (unsigned long long)total;
        const int nch = A.gch[ng];
        // where the hits go: G in place, or the block's list (copied to G below)
        uint64_t* const ev = (!kGlobal && el && gpos >= 0) ? tl.G + gpos : el;
        for (int cid = w; cid < nch; cid += NW) {
            // the chunk's group: the last g with gch[g] <= cid
            int glo = 0, ghi = ng - 1;
            while (glo < ghi) {
                const int mid = (glo + ghi + 1) >> 1;
                if (A.gch[mid] <= cid) glo = mid;
                else ghi = mid - 1;
            }
            const int g = glo;
            const int32_t R = (cid - A.gch[g]) * kChunk;  // chunk start within the group
            const int64_t G0 = g ? A.pref[64 * g - 1] : 0;
            const int32_t glen = (int32_t)(A.pref[min(64 * g + 64, k) - 1] - G0);
            const int i = 64 * g + lane;
            int32_t es = INT32_MAX, len = 0;
            if (i < k) {
                const int64_t ex = i ? A.pref[i - 1] : 0;
                len = (int32_t)(A.pref[i] - ex);
                es = (int32_t)(ex - G0);
            }
            // the group's non-empty lists, ranked (a searched list's "slots"
            // are the local offsets j - i - 1)
            const uint64_t nz = __ballot(len > 0);
            if (len > 0) {
                const int r = __popcll(nz & (upto >> 1));
                const int64_t ob = A.obeg[i];
                const bool srch = (ob & kSearchBit) != 0;
                wbase[w][r] = (srch ? 0 : ob) - es;
                wown[w][r] = i | (srch ? kSearchOwner : 0);
            } else {
                es = INT32_MAX;
            }
            wflag[w][lane] = 0;
            __builtin_amdgcn_wave_barrier();
            if (es >= R && es < R + kChunk) fl[((es - R) & 63) * 4 + ((es - R) >> 6)] = 1;
            __builtin_amdgcn_wave_barrier();
            const uint32_t f4 = wflag[w][lane];
            uint32_t f[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) f[u] = (f4 >> (8 * u)) & 0xffu;
            int cum = __popcll(__ballot(es < R)) - 1;  // last list starting before the chunk
            int ii[kU], j[kU];
            bool ok[kU], sr[kU];
            int64_t slot[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const uint64_t starts = __ballot(f[u] != 0);
                const int32_t q = R + 64 * u + lane;
                ok[u] = q < glen;
                // past the group's end: clamp to its last position (a valid slot)
                const int o = cum + __popcll(starts & upto);
                cum += __popcll(starts);
                slot[u] = wbase[w][o] + (ok[u] ? q : glen - 1);
                const int wo = wown[w][o];
                ii[u] = wo & ~kSearchOwner;
                sr[u] = (wo & kSearchOwner) != 0;
            }
            int32_t x[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) x[u] = sr[u] ? A.L[ii[u] + 1 + (int)slot[u]] : col[slot[u]];
            uint32_t h[kU];
            bool maybe[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                h[u] = mhash(x[u]);
                const uint32_t bw = A.bm[h[u] >> (32 - A.bbits + 5)];
                maybe[u] = ok[u] && !sr[u] && ((bw >> ((h[u] >> (32 - A.bbits)) & 31)) & 1);
            }
#pragma unroll
            for (int u = 0; u < kU; u++) {
                j[u] = maybe[u] ? tri_find(A, x[u], h[u]) : -1;
                if (ok[u] && sr[u]) {
                    // L[j] in N+(L[i])?  Its position there gives the out-slot
                    const int32_t pos = out_hash_find(oh, A.L[ii[u]], x[u], h[u]);
                    if (pos >= 0) {
                        j[u] = ii[u] + 1 + (int)slot[u];
                        slot[u] = (A.obeg[ii[u]] & ~kSearchBit) + pos;
                    }
                }
            }
            uint32_t e[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                if (j[u] >= 0) {
                    atomicAdd(&TS[slot[u]], 1u);  // the edge (L[i], L[j]) by its out-slot: consecutive lanes
                    atomicAdd(&A.cnt[ii[u]], 1u);
                    atomicAdd(&A.cnt[j[u]], 1u);
                    e[u] = (uint32_t)slot[u];
                }
            }
            if (ev) record_hits<kU, true>(ev, &ecount, ii, j, e);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        {
            const int64_t s0 = ostart[a];  // the out-slots of a
            for (int i = tid; i < k; i += NT)
                if (A.cnt[i]) atomicAdd(&TS[s0 + i], A.cnt[i]);
        }
        if (!kGlobal && el) {
            if (tid == 0) {
                // written in place, or not kept once G is full (the clique
                // pass then searches again)
                const int64_t p = gpos;
                tl.off[a] = p;
                tl.len[a] = p < 0 ? -1 : ecount;
                if (p >= 0) chunk_pos += ecount;
            }
        } else if (el) {
            const int ne = ecount;
            if (tid == 0) {
                // lists are placed in chunks the block reserves (one global
                // atomic per chunk, not per vertex)
                if (chunk_pos + ne > chunk_end) {
                    const int64_t want = max((int64_t)ne, (int64_t)kListChunk);
                    const int64_t c = (int64_t)atomicAdd(tl.cur, (unsigned long long)want);
                    chunk_pos = c;
                    chunk_end = c + want <= tl.gcap ? c + want : c;  // an empty chunk once G is full
                }
                int64_t p = -1;
                if (chunk_pos + ne <= chunk_end) {
                    p = chunk_pos;
                    chunk_pos += ne;
                }
                gpos = p;
                tl.off[a] = p;
                tl.len[a] = p < 0 ? -1 : ne;
            }
            __syncthreads();
            const int64_t p = gpos;
            if (p >= 0)
                for (int q = tid; q < ne; q += NT) tl.G[p + q] = el[q];
        }
        __syncthreads();
    }
    if (done) atomicAdd(work, done);
}

// pass 2: 4-cliques and Q, with the local bit matrix.  The pass that builds
// the bit matrix also records every local edge (i, j, out-slot of (L[i],
// L[j])) in the block's slice of `elist` (capacity k(k-1)/2 for the bin's
// largest k, so it cannot overflow); the scan then reads that list instead
// of searching again.  Without a list (the HBM-scratch bin) the scan
// searches again.
//
// The far edge (L[i], L[j]) of a local edge is addressed by its out-slot:
// its T is read from TSL (T by slot) and its K4 and Q terms go to per-slot
// accumulators, folded into K4 and Q by edge id after the pass.  When an
// edge's lower endpoint has degree at most pk.deg, its K4 (at most
// C(pk.deg - 1, 2) < 2^(64 - pk.qb)) and its Q (at most 2 pk.deg max-degree
// < 2^pk.qb) share one 64-bit sum PK = K4 << pk.qb | Q, one atomic per local
// edge; other edges use K4S and QS.  pk is chosen per graph from its maximum
// degree (pack_split).  Local edges
// come in list order, which follows the out-lists, so consecutive lanes
// touch nearby slots instead of scattered edge ids.
//
// When the full k x k bit matrix exceeds the block's mw words, it is held in
// column blocks of wc words per row: the K4 terms of a local edge are summed
// over the blocks (popcounts of partial rows), the Q terms are added once.
struct PackSplit {
    int32_t deg;  // largest lower-end degree of a packed edge (0: no packing)
    int qb;       // low bits of PK that hold Q
};

// (narrow: the lower endpoint L[i] of the far edge has degree <= pk.deg,
// kept by the caller in A.Le[i]; tij: T of the far edge, read by the caller
// in the first column block)
__device__ __forceinline__ void edge_terms_k4q(const LocalArrays& A, int wn, int ws, bool first, bool act, int i,
                                               int j, uint32_t sij, uint32_t tij, uint64_t* __restrict__ K4S,
                                               uint64_t* __restrict__ QS, uint64_t* __restrict__ PK, PackSplit pk) {
    // called by every lane of the wavefront (act: the lane holds a local
    // edge); rows are ws words apart, wn of them held
    const int lane = threadIdx.x & 63;
    uint32_t kij = 0;
    if (act) {
        // the rows' words four at a time: the eight LDS reads are in flight together
        const uint32_t* ri = A.M + i * ws;
        const uint32_t* rj = A.M + j * ws;
        int w = 0;
        for (; w + 4 <= wn; w += 4) {
            const uint32_t a0 = ri[w], a1 = ri[w + 1], a2 = ri[w + 2], a3 = ri[w + 3];
            const uint32_t b0 = rj[w], b1 = rj[w + 1], b2 = rj[w + 2], b3 = rj[w + 3];
            kij += __popc(a0 & b0) + __popc(a1 & b1) + __popc(a2 & b2) + __popc(a3 & b3);
        }
        for (; w < wn; w++) kij += __popc(ri[w] & rj[w]);
    }
    // lanes in list order share the local vertex i: one atomic per run of i
    const LaneRun r = lane_run(act ? (uint32_t)i : ~(uint32_t)lane);
    const uint32_t ksum = run_sum32(kij, r.start);
    if (act && r.last && ksum) atomicAdd(&A.cnt[i], ksum);
    if (act && kij) atomicAdd(&A.cnt[j], kij);
    uint64_t q = 0;
    if (first) {
        const uint64_t tai = act ? A.ta[i] : 0, taj = act ? A.ta[j] : 0;
        q = tai + taj;
        const uint64_t asum = run_sum64(act ? taj + tij : 0, r.start);
        if (act && r.last && asum) atomicAdd((unsigned long long*)&A.acc[i], (unsigned long long)asum);
        if (act) atomicAdd((unsigned long long*)&A.acc[j], (unsigned long long)(tai + tij));
    }
    if (!act) return;
    if (A.Le[i] <= pk.deg) {
        if (kij | q) atomicAdd((unsigned long long*)&PK[sij], (unsigned long long)(((uint64_t)kij << pk.qb) | q));
    } else {
        if (kij) atomicAdd((unsigned long long*)&K4S[sij], (unsigned long long)kij);
        if (q) atomicAdd((unsigned long long*)&QS[sij], (unsigned long long)q);
    }
}

// set the bits of the local edge (i, j) that fall in columns [jlo, jlo + 32 wn)
// (rows ws words apart)
__device__ __forceinline__ void set_pair(uint32_t* M, int wn, int ws, int jlo, int i, int j) {
    const int ji = j - jlo, ii = i - jlo;
    if (ji >= 0 && ji < 32 * wn) atomicOr(&M[i * ws + (ji >> 5)], 1u << (ji & 31));
    if (ii >= 0 && ii < 32 * wn) atomicOr(&M[j * ws + (ii >> 5)], 1u << (ii & 31));
}
// the same, called by every lane of the wavefront (act: the lane holds a local
// edge): lanes in list order share i and often the word of row i, so the
// bits of a run of equal words are or-ed across lanes and set by one atomic
__device__ __forceinline__ void set_pair_wave(uint32_t* M, int wn, int ws, int jlo, int i, int j, bool act) {
    const int lane = threadIdx.x & 63;
    const int ji = j - jlo, ii = i - jlo;
    const bool a1 = act && ji >= 0 && ji < 32 * wn;
    const uint32_t key = a1 ? (uint32_t)(i * ws + (ji >> 5)) : ~(uint32_t)lane;
    const LaneRun r = lane_run(key);
    const uint32_t bits = run_or(a1 ? 1u << (ji & 31) : 0u, r.start);
    if (a1 && r.last) atomicOr(&M[key], bits);
    if (act && ii >= 0 && ii < 32 * wn) atomicOr(&M[j * ws + (ii >> 5)], 1u << (ii & 31));
}

template <int NT, bool kGlobal>
__global__ __launch_bounds__(NT) void k_local_k4q(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                                                  const int32_t* __restrict__ eid, const int64_t* __restrict__ ostart,
                                                  const int32_t* __restrict__ outdeg, const int32_t* __restrict__ list,
                                                  int64_t cnt, int kmax, int64_t mw, char* __restrict__ gscratch,
                                                  uint64_t* __restrict__ elist, int64_t ecap, TriLists tl,
                                                  const uint32_t* __restrict__ TSL, uint64_t* __restrict__ K4S,
                                                  uint64_t* __restrict__ QS, uint64_t* __restrict__ PK, PackSplit pk,
                                                  uint32_t* __restrict__ queue, unsigned long long* __restrict__ work) {
    __shared__ int64_t li_sh;
    extern __shared__ char lds[];
    __shared__ int64_t part[NT];
    __shared__ int ecount;
    char* base = kGlobal ? gscratch + (size_t)blockIdx.x * carve_bytes(kmax, mw) : lds;
    LocalArrays A = carve(base, kmax);
    uint64_t* el = elist ? elist + (size_t)blockIdx.x * (size_t)ecap : nullptr;
    unsigned long long done = 0;
    for (int64_t li = next_vertex(queue, cnt, -1, &li_sh); li >= 0; li = next_vertex(queue, cnt, li, &li_sh)) {
        const int32_t a = list[li];
        const int k = outdeg[a];
        const int W = (k + 31) / 32;
        const int wc = (int)min<int64_t>(W, mw / k);  // words per row held at once (>= 1)
        if (threadIdx.x == 0) ecount = 0;
        const int64_t total = load_local<NT>(row, col, eid, ostart, a, k, A, part);
        // the edges a -> L[i] are a's out-slots ostart[a] + i
        const int64_t s0 = ostart[a];
        // (Le, the edge ids, is not needed here: it holds the degree of L[i])
        for (int i = threadIdx.x; i < k; i += NT) {
            A.ta[i] = TSL[s0 + i];
            A.Le[i] = (int32_t)(row[A.L[i] + 1] - row[A.L[i]]);
        }
        // the triangle pass's list of a's local edges, when it kept one
        const int kept = tl.G ? tl.len[a] : -1;
        const uint64_t* gl = kept >= 0 ? tl.G + tl.off[a] : nullptr;
        for (int cw = 0; cw < W; cw += wc) {
            const int wn = min(wc, W - cw), jlo = 32 * cw;
            // an odd row stride when it fits: the rows read together by a
            // wavefront start in different banks
            const int ws = (int64_t)k * (wn | 1) <= mw ? (wn | 1) : wn;
            const bool first = cw == 0;
            for (int w = threadIdx.x; w < k * ws; w += NT) A.M[w] = 0;
            __syncthreads();
            // local edges: the kept list, this block's list (recorded while
            // the first column block is built), or a new search
            const uint64_t* src = gl ? gl : (first ? nullptr : el);
            const int ne = gl ? kept : ecount;
            if (src) {
                const int lane = threadIdx.x & 63;
                for (int pb = threadIdx.x - lane; pb < ne; pb += NT) {  // uniform per wavefront
                    const int p = pb + lane;
                    const uint64_t e = src[min(p, ne - 1)];
                    set_pair_wave(A.M, wn, ws, jlo, (int)(e & 0xffff), (int)((e >> 16) & 0xffff), p < ne);
                }
            } else {
                for (int64_t t0 = threadIdx.x; t0 < total; t0 += NT * kU) {
                    int i[kU], j[kU];
                    int64_t slot[kU];
                    int32_t x[kU];
                    locate<NT>(A, k, total, t0, i, slot);
#pragma unroll
                    for (int u = 0; u < kU; u++)
                        if (i[u] >= 0) x[u] = col[slot[u]];
#pragma unroll
                    for (int u = 0; u < kU; u++) j[u] = i[u] >= 0 ? find_local(A, x[u]) : -1;
                    uint32_t e[kU];
#pragma unroll
                    for (int u = 0; u < kU; u++) {
                        if (j[u] >= 0) {
                            set_pair(A.M, wn, ws, jlo, i[u], j[u]);
                            e[u] = (uint32_t)slot[u];
                        }
                    }
                    if (first && el) record_hits<kU, false>(el, &ecount, i, j, e);
                }
            }
            __syncthreads();
            const uint64_t* lst = gl ? gl : el;
            if (lst) {
                const int nl = gl ? kept : ecount;
                if (threadIdx.x == 0) done += (unsigned long long)nl;
                // kU list entries per thread at once: their loads, and then
                // their T loads, are in flight together (unconditional:
                // entries past the end are read as the last one)
                const int lane = threadIdx.x & 63;
                for (int pb = threadIdx.x - lane; pb < nl; pb += NT * kU) {  // uniform per wavefront
                    const int p = pb + lane;
                    uint64_t e[kU];
                    uint32_t t[kU] = {};
#pragma unroll
                    for (int u = 0; u < kU; u++) e[u] = lst[min(p + u * NT, nl - 1)];
                    if (first) {
#pragma unroll
                        for (int u = 0; u < kU; u++) t[u] = TSL[(uint32_t)(e[u] >> 32)];
                    }
#pragma unroll
                    for (int u = 0; u < kU; u++)
                        edge_terms_k4q(A, wn, ws, first, p + u * NT < nl, (int)(e[u] & 0xffff),
                                       (int)((e[u] >> 16) & 0xffff), (uint32_t)(e[u] >> 32), t[u], K4S, QS, PK, pk);
                }
            } else {
                // no list: wc == W here (the HBM bin holds the full matrix)
                if (threadIdx.x == 0) done += (unsigned long long)total;
                const int64_t lane = threadIdx.x & 63;
                for (int64_t tb = threadIdx.x - lane; tb < total; tb += NT * kU) {  // uniform per wavefront
                    const int64_t t0 = tb + lane;
                    int iu[kU], ju[kU];
                    int64_t su[kU];
                    int32_t x[kU];
                    locate<NT>(A, k, total, t0, iu, su);
#pragma unroll
                    for (int u = 0; u < kU; u++)
                        if (iu[u] >= 0) x[u] = col[su[u]];
#pragma unroll
                    for (int u = 0; u < kU; u++) ju[u] = iu[u] >= 0 ? find_local(A, x[u]) : -1;
                    uint32_t tu[kU];
#pragma unroll
                    for (int u = 0; u < kU; u++) tu[u] = first && ju[u] >= 0 ? TSL[su[u]] : 0u;
#pragma unroll
                    for (int u = 0; u < kU; u++)
                        edge_terms_k4q(A, wn, ws, first, ju[u] >= 0, iu[u], max(ju[u], 0), (uint32_t)su[u], tu[u],
                                       K4S, QS, PK, pk);
                }
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < k; i += NT) {
            // each local triangle through i is seen from its two edges at i
            // the edge (a, L[i]): its lower endpoint is a
            const uint64_t k4 = A.cnt[i] >> 1, q = A.acc[i];
            if (row[a + 1] - row[a] <= pk.deg) {
                if (k4 | q) atomicAdd((unsigned long long*)&PK[s0 + i], (unsigned long long)((k4 << pk.qb) | q));
            } else {
                if (k4) atomicAdd((unsigned long long*)&K4S[s0 + i], (unsigned long long)k4);
                if (q) atomicAdd((unsigned long long*)&QS[s0 + i], (unsigned long long)q);
            }
        }
        __syncthreads();
    }
    if (done) atomicAdd(work, done);
}

// ---------------------------------------------------------------------------
// P: every 4-cycle is found once, from the vertex c opposite its
// highest-ranked vertex a, as two wedges c-b-a and c-b'-a with
// a > max(b, c) and a > max(b', c).  For the anchor c the endpoints a of
// the wedges through its neighbour b are the suffix of the ascending N(b)
// above max(b, c).
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t hslot(int32_t c) { return ((uint32_t)c * 2654435761u) >> 23; }  // 9 bits

// light anchors: one wavefront (block of 64) per anchor c, an LDS hash of
// (endpoint a -> wedge count); distinct a <= bound <= 256 < 512 slots.  The
// wedge c-b-a adds cnt[a]-1 to the edges (b, a) and (c, b).
__global__ __launch_bounds__(64) void k_wedge_light(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                                                    const int32_t* __restrict__ eid,
                                                    const int32_t* __restrict__ list, int64_t cnt,
                                                    uint64_t* __restrict__ P, unsigned long long* __restrict__ work) {
    __shared__ int32_t hkey[kHashSlots];
    __shared__ uint32_t hcnt[kHashSlots];
    __shared__ int64_t pref[64];
    __shared__ int64_t bbeg[64];
    __shared__ uint64_t psum[64];
    __shared__ int64_t part[64];
    const int tid = threadIdx.x;
    unsigned long long done = 0;
    for (int64_t li = blockIdx.x; li < cnt; li += gridDim.x) {
        const int32_t c = list[li];
        const int64_t ib = row[c], ie = row[c + 1];
        for (int s = tid; s < kHashSlots; s += 64) { hkey[s] = -1; hcnt[s] = 0; }
        __syncthreads();
        for (int pass = 0; pass < 2; pass++) {
            for (int64_t cb = ib; cb < ie; cb += 64) {
                const int kc = (int)min<int64_t>(64, ie - cb);
                if (tid < kc) {
                    const int32_t b = col[cb + tid];
                    const int64_t be = row[b + 1];
                    const int64_t q = lower_bound_g(col, row[b], be, max(b, c) + 1);  // first endpoint above max(b, c)
                    bbeg[tid] = q;
                    pref[tid] = be - q;
                    psum[tid] = 0;
                }
                __syncthreads();
                block_incl_scan<64>(pref, kc, part);
                const int64_t total = pref[kc - 1];
                if (pass == 0 && tid == 0) done += (unsigned long long)total;
                for (int64_t t = tid; t < total; t += 64) {
                    const int i = owner_of(pref, kc, t);
                    const int64_t ex = i ? pref[i - 1] : 0;
                    const int64_t slot = bbeg[i] + (t - ex);
                    const int32_t a = col[slot];
                    uint32_t h = hslot(a);
                    if (pass == 0) {
                        for (;;) {
                            int32_t old = atomicCAS(&hkey[h], -1, a);
                            if (old == -1 || old == a) { atomicAdd(&hcnt[h], 1u); break; }
                            h = (h + 1) & (kHashSlots - 1);
                        }
                    } else {
                        while (hkey[h] != a) h = (h + 1) & (kHashSlots - 1);
                        const uint32_t v = hcnt[h] - 1;
                        if (v) {
                            atomicAdd((unsigned long long*)&P[eid[slot]], (unsigned long long)v);
                            atomicAdd((unsigned long long*)&psum[i], (unsigned long long)v);
                        }
                    }
                }
                __syncthreads();
                if (pass == 1 && tid < kc && psum[tid])
                    atomicAdd((unsigned long long*)&P[eid[cb + tid]], (unsigned long long)psum[tid]);
                __syncthreads();
            }
        }
    }
    if (done) atomicAdd(work, done);
}

// C4 = P - Q + 2T + 2K4 (uint64 wrap-around; the true value is >= 0)
__global__ void k_combine(const uint64_t* __restrict__ P, const uint64_t* __restrict__ Q,
                          const uint32_t* __restrict__ T, const uint64_t* __restrict__ K4, int64_t m,
                          uint64_t* __restrict__ C4) {
    GRID_STRIDE(e, m) C4[e] = P[e] - Q[e] + 2 * (uint64_t)T[e] + 2 * K4[e];
}

}  // namespace


namespace {


// compaction of the vertices whose key falls in (lo, hi] into `out` (capacity n)
template <class K>
int64_t make_bin(pgd_ctx* ctx, const K* key, int64_t n, int64_t lo, int64_t hi, DBuf<int32_t>& out,
                 DBuf<int32_t>& flags, DBuf<int64_t>& off) {
    hipStream_t s = ctx->stream;
    int g = grid_cap(ctx, (n + 255) / 256, 16);
    if constexpr (sizeof(K) == 4)
        hipLaunchKernelGGL(k_bin_flags_i32, dim3(g), dim3(256), 0, s, (const int32_t*)key, n, lo, hi, flags.get());
    else
        hipLaunchKernelGGL(k_bin_flags_i64, dim3(g), dim3(256), 0, s, (const int64_t*)key, n, lo, hi, flags.get());
    PGD_HIP(hipGetLastError());
    excl_sum(ctx, flags.get(), off.get(), (size_t)n);
    int64_t last_off = 0;
    int32_t last_f = 0;
    PGD_HIP(hipMemcpyAsync(&last_off, off.get() + (n - 1), sizeof(int64_t), hipMemcpyDeviceToHost, s));
    PGD_HIP(hipMemcpyAsync(&last_f, flags.get() + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost, s));
    PGD_HIP(hipStreamSynchronize(s));
    int64_t c = last_off + last_f;
    ensure(out, ctx, (size_t)n);
    if (c) {
        hipLaunchKernelGGL(k_bin_scatter, dim3(g), dim3(256), 0, s, flags.get(), off.get(), n, out.get());
        PGD_HIP(hipGetLastError());
    }
    return c;
}

__global__ void k_fill_ones_u64(uint64_t* __restrict__ p, int64_t n) {
    GRID_STRIDE(i, n) p[i] = ~0ull;
}
__global__ void k_fill_u64(uint64_t* __restrict__ p, int64_t n) {
    GRID_STRIDE(i, n) p[i] = 0;
}
__global__ void k_fill_u32(uint32_t* __restrict__ p, int64_t n) {
    GRID_STRIDE(i, n) p[i] = 0;
}
__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
    GRID_STRIDE(i, n) p[i] = v;
}
// TSL[s] = T of the edge at slot s
// K4[eid[s]] += K4S[s] + PK[s] >> qb, Q[eid[s]] += QS[s] + PK[s] mod 2^qb:
// the clique pass adds only at out-slots, one per edge, so the adds need no
// atomics
__global__ void k_fold_k4q(const uint64_t* __restrict__ K4S, const uint64_t* __restrict__ QS,
                           const uint64_t* __restrict__ PK, int qb, const int32_t* __restrict__ eid, int64_t nnz,
                           uint64_t* __restrict__ K4, uint64_t* __restrict__ Q) {
    GRID_STRIDE(s, nnz) {
        const uint64_t pk = PK[s];
        const uint64_t k = K4S[s] + (pk >> qb), q = QS[s] + (pk & ((1ULL << qb) - 1));
        if (k | q) {
            const int32_t e = eid[s];
            K4[e] += k;
            Q[e] += q;
        }
    }
}
// The packed (K4, Q) split for a graph of maximum degree dmax: with Q in the
// low qb bits, an edge whose lower end has degree d packs safely when
// C(d - 1, 2) < 2^(64 - qb) (K4 <= C(T, 2), T <= d - 1) and
// 2 (d - 1)(dmax - 1) < 2^qb (Q sums at most d - 1 terms of two T's, each
// below dmax).  The qb that allows the largest d is chosen.
PackSplit pack_split(int64_t dmax) {
    PackSplit best{0, 40};
    for (int qb = 20; qb <= 44; qb++) {
        const double kcap = std::ldexp(1.0, 64 - qb), qcap = std::ldexp(1.0, qb);
        // largest d with (d-1)(d-2)/2 < kcap, and with 2 (d-1)(dmax-1) < qcap
        // (all products below 2^53: exact in double)
        double dk = std::floor((3.0 + std::sqrt(1.0 + 8.0 * kcap)) / 2.0);
        while (dk > 2 && (dk - 1) * (dk - 2) / 2 >= kcap) dk -= 1;
        double dq = dmax > 1 ? std::floor(qcap / (2.0 * (double)(dmax - 1))) : 1e18;
        while (dq > 1 && 2.0 * (dq - 1) * (double)(dmax - 1) >= qcap) dq -= 1;
        const double d = std::min(std::min(dk, dq), (double)INT32_MAX);
        if (d > best.deg) best = PackSplit{(int32_t)d, qb};
    }
    return best;
}

// T[eid[s]] += TS[s] for every slot s with a count
// T[eid[s]] = TS[s]: only out-slots hold counts, one per edge (T is zeroed)
__global__ void k_fold_u32(const uint32_t* __restrict__ TS, const int32_t* __restrict__ eid, int64_t nnz,
                           uint32_t* __restrict__ T) {
    GRID_STRIDE(s, nnz) {
        const uint32_t v = TS[s];
        if (v) T[eid[s]] = v;
    }
}

}  // namespace

namespace pgd {


void orient_state_free(pgd_ctx*, OrientState* s) { delete s; }

OrientParams& orient_params(OrientState* s) { return s->prm; }

// relabel, ascending CSR, bins and scratch, and (with plan) the 4-cycle
// plan (host synchronisations; not capturable)
void orient_build(pgd_ctx* ctx, const GraphView& g, OrientState& st, bool plan) {
    hipStream_t s = ctx->stream;
    const int64_t n = g.n, m = g.m, nnz = 2 * m;
    const OrientParams& prm = st.prm;
    st.n = n;
    st.m = m;
    st.max_deg = g.max_deg;
    ensure(st.rank, ctx, n);
    ensure(st.rdeg, ctx, n);
    ensure(st.rrow, ctx, n + 1);
    ensure(st.ostart, ctx, n);
    ensure(st.outdeg, ctx, n);
    ensure(st.rcol, ctx, nnz);
    ensure(st.reid, ctx, nnz);
    ensure(st.k0, ctx, 2 * std::max(n, nnz));  // (also the clique pass's per-slot K4 and Q sums)
    ensure(st.k1, ctx, std::max(n, nnz));
    ensure(st.v0, ctx, nnz);
    ensure(st.flags, ctx, n);
    ensure(st.off, ctx, n);
    ensure(st.ub, ctx, n);
    ensure(st.P, ctx, m);
    ensure(st.Q, ctx, m);
    ensure(st.TS, ctx, nnz);
    ensure(st.work, ctx, 5);
    ensure(st.lq, ctx, 6);
    int gn = grid_cap(ctx, (n + 255) / 256, 16), gm = grid_cap(ctx, (m + 255) / 256, 16);
    const int vb = bitsfor(n - 1), db = bitsfor(g.max_deg), gb = g.group ? bitsfor(g.ngroups - 1) : 0;
    PGD_CHECK(vb + db + gb <= 64, PGD_ERR_UNSUPPORTED, "rank key of (group, degree, id) exceeds 64 bits");
    PGD_CHECK(nnz < ((int64_t)1 << 32), PGD_ERR_UNSUPPORTED, "2m >= 2^32: local-edge lists hold 32-bit slots");
    hipLaunchKernelGGL(k_rank_keys, dim3(gn), dim3(256), 0, s, g.deg, g.group, n, db, vb, st.k0.get());
    PGD_HIP(hipGetLastError());
    sort_k(ctx, st.k0.get(), st.k1.get(), (size_t)n, (unsigned)(vb + db + gb));
    hipLaunchKernelGGL(k_rank_scatter, dim3(gn), dim3(256), 0, s, st.k1.get(), n, db, vb, st.rank.get(),
                       st.rdeg.get());
    PGD_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_slot_keys, dim3(gm), dim3(256), 0, s, g.ev, g.eu, st.rank.get(), m, vb, st.k0.get(),
                       st.v0.get());
    PGD_HIP(hipGetLastError());
    sort_kv(ctx, st.k0.get(), st.k1.get(), st.v0.get(), st.reid.get(), (size_t)nnz, 2u * (unsigned)vb);
    int gz = grid_cap(ctx, (nnz + 255) / 256, 16);
    hipLaunchKernelGGL(k_slot_cols, dim3(gz), dim3(256), 0, s, st.k1.get(), nnz, vb, st.rcol.get());
    PGD_HIP(hipGetLastError());
    excl_sum(ctx, st.rdeg.get(), st.rrow.get(), (size_t)n);
    hipLaunchKernelGGL(k_row_tail, dim3(1), dim3(1), 0, s, st.rrow.get(), n, nnz);
    hipLaunchKernelGGL(k_out_start, dim3(gn), dim3(256), 0, s, st.rrow.get(), st.rcol.get(), n, st.ostart.get(),
                       st.outdeg.get());
    PGD_HIP(hipGetLastError());
    st.maxk = device_max(ctx, st.outdeg.get(), (size_t)n);
    // the largest local graph whose arrays fit one block's LDS (the triangle
    // pass); the clique pass gives the rest of the LDS to its bit matrix,
    // held in column blocks when the whole matrix does not fit.  Local
    // indices are 16-bit in the local-edge lists.
    // (static LDS: 8 KiB of scan partials in the clique pass; 24 KiB of scan
    // partials and per-wavefront tables in the triangle pass)
    const size_t lds_cap = std::min<size_t>(ctx->lds_per_block, 160 * 1024) - 8192;
    const size_t lds_cap_tri = std::min<size_t>(ctx->lds_per_block, 160 * 1024) - 26 * 1024;
    int kl = 32;
    while (kl + 32 < 65536 && carve_bytes(kl + 32, (int64_t)kl + 32) <= lds_cap &&
           tri_carve_bytes(kl + 32) <= lds_cap_tri)
        kl += 32;
    st.kmax_lds = std::max(kl, prm.mid_k);
    st.n_small = make_bin(ctx, st.outdeg.get(), n, 1, prm.small_k, st.b_small, st.flags, st.off);
    st.n_mid = make_bin(ctx, st.outdeg.get(), n, prm.small_k, prm.mid_k, st.b_mid, st.flags, st.off);
    st.n_large = make_bin(ctx, st.outdeg.get(), n, prm.mid_k, st.kmax_lds, st.b_large, st.flags, st.off);
    st.n_huge = make_bin(ctx, st.outdeg.get(), n, st.kmax_lds, INT32_MAX, st.b_huge, st.flags, st.off);
    {
        int gw = grid_cap(ctx, (n * 64 + 255) / 256, 16);
        hipLaunchKernelGGL(k_wedge_bound, dim3(gw), dim3(256), 0, s, st.rrow.get(), st.rcol.get(), st.outdeg.get(),
                           n, st.ub.get());
        PGD_HIP(hipGetLastError());
    }
    st.n_light = make_bin(ctx, st.ub.get(), n, 0, kLightBound, st.b_light, st.flags, st.off);
    st.n_heavy = make_bin(ctx, st.ub.get(), n, kLightBound, INT64_MAX, st.b_heavy, st.flags, st.off);
    st.mw_small = full_matrix_words(prm.small_k);
    st.mw_mid = full_matrix_words(prm.mid_k);
    st.small_lds = carve_bytes(prm.small_k, st.mw_small);
    st.mid_lds = carve_bytes(prm.mid_k, st.mw_mid);
    // the large bin's arrays are sized for its largest member
    st.kq = std::max(prm.mid_k, std::min(st.maxk, st.kmax_lds));
    {
        const int64_t spare = ((int64_t)lds_cap - (int64_t)carve_bytes(st.kq, 0)) / 4 - 64;
        st.mw_large = std::min<int64_t>(full_matrix_words(st.kq), std::max<int64_t>(spare, st.kq));
    }
    st.small_lds_tri = tri_carve_bytes(prm.small_k);
    st.mid_lds_tri = tri_carve_bytes(prm.mid_k);
    st.large_lds_tri = tri_carve_bytes(st.kq);
    st.large_lds = carve_bytes(st.kq, st.mw_large);
    if (st.n_large) {
        PGD_HIP(hipFuncSetAttribute((const void*)k_local_tri<1024, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)st.large_lds_tri));
        PGD_HIP(hipFuncSetAttribute((const void*)k_local_k4q<1024, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)st.large_lds));
    }
    if (st.n_huge) {
        st.ghuge = (int)std::min<int64_t>(st.n_huge, ctx->num_cus);
        // its local-edge lists (kept for the clique pass) take maxk(maxk-1)/2
        // entries per block: fewer blocks when that exceeds 4 GiB in all
        const int64_t per = std::max<int64_t>(1, (int64_t)st.maxk * (st.maxk - 1) / 2) * 8;
        st.ghuge = (int)std::max<int64_t>(1, std::min<int64_t>(st.ghuge, ((int64_t)4 << 30) / per));
        st.mw_huge = full_matrix_words(st.maxk);
        // one HBM slice per block, shared by the triangle and the clique pass
        st.huge_stride = std::max(carve_bytes(st.maxk, st.mw_huge), tri_carve_bytes(st.maxk));
        ensure(st.hscratch, ctx, (size_t)st.ghuge * st.huge_stride);
    }
    {
        // local-edge lists of the clique pass: per block, the bin's k(k-1)/2 at most
        auto cap = [](int64_t k) { return std::max<int64_t>(1, k * (k - 1) / 2); };
        st.ecap_small = cap(prm.small_k);
        st.ecap_mid = cap(prm.mid_k);
        st.ecap_large = cap(st.kq);
        const int64_t bs = st.n_small ? grid_cap(ctx, st.n_small, prm.grid_small) : 0;
        const int64_t bm = st.n_mid ? grid_cap(ctx, st.n_mid, prm.grid_mid) : 0;
        const int64_t bl = st.n_large ? grid_cap(ctx, st.n_large, 1) : 0;
        st.ecap_huge = st.n_huge ? cap(st.maxk) : 0;
        const int64_t bh = st.n_huge ? st.ghuge : 0;
        ensure(st.elist,
               ctx, (size_t)(bs * st.ecap_small + bm * st.ecap_mid + bl * st.ecap_large + bh * st.ecap_huge));
        st.el_small = st.elist.get();
        st.el_mid = st.el_small + bs * st.ecap_small;
        st.el_large = st.el_mid + bm * st.ecap_mid;
        st.el_huge = st.el_large + bl * st.ecap_large;
    }
    {
        // global local-edge lists (triangle pass -> clique pass), as many
        // entries as a third of the free memory holds, at most tri_lists per
        // edge; a vertex whose list does not fit is searched again
        st.tcap = 0;
        if (prm.tri_lists > 0 && m >= ((int64_t)1 << 12)) {
            size_t fr = 0, tot = 0;
            PGD_HIP(hipMemGetInfo(&fr, &tot));
            int64_t avail = (int64_t)(fr / 3);
            if (ctx->mem_cap > 0) avail = std::min<int64_t>(avail, (ctx->mem_cap - ctx->mem_used.load()) / 3);
            st.tcap = std::min<int64_t>(avail / 8, (int64_t)prm.tri_lists * m);
        }
        if (st.tcap > 0) {
            ensure(st.tlist, ctx, (size_t)st.tcap);
            ensure(st.toff, ctx, n);
            ensure(st.tlen, ctx, n);
            ensure(st.tcur, ctx, 1);
        }
    }
    {
        // out-list hash tables of the vertices of large out-degree (the
        // triangle pass looks members up in them instead of probing the
        // whole list)
        st.has_oh = false;
        if (prm.tri_search > 0 && n > 0) {
            DBuf<int32_t> sz;
            DBuf<int64_t> ex;
            sz.alloc(ctx, n);
            ex.alloc(ctx, n);
            const int gn = grid_cap(ctx, (n + 255) / 256, 16);
            hipLaunchKernelGGL(k_out_hash_size, dim3(gn), dim3(256), 0, s, st.outdeg.get(), n, sz.get());
            PGD_HIP(hipGetLastError());
            {
                size_t bytes = 0;
                PGD_HIP(rocprim::exclusive_scan(nullptr, bytes, sz.get(), ex.get(), (int64_t)0, (size_t)n,
                                                rocprim::plus<int64_t>(), s));
                void* t = tmp_storage(ctx, bytes);
                PGD_HIP(rocprim::exclusive_scan(t, bytes, sz.get(), ex.get(), (int64_t)0, (size_t)n,
                                                rocprim::plus<int64_t>(), s));
            }
            int64_t last_ex = 0;
            int32_t last_sz = 0;
            PGD_HIP(hipMemcpyAsync(&last_ex, ex.get() + (n - 1), sizeof(last_ex), hipMemcpyDeviceToHost, s));
            PGD_HIP(hipMemcpyAsync(&last_sz, sz.get() + (n - 1), sizeof(last_sz), hipMemcpyDeviceToHost, s));
            PGD_HIP(hipStreamSynchronize(s));
            const int64_t entries = last_ex + last_sz;
            if (entries > 0) {
                ensure(st.ohoff, ctx, n);
                ensure(st.ohtab, ctx, (size_t)entries);
                hipLaunchKernelGGL(k_out_hash_off, dim3(gn), dim3(256), 0, s, sz.get(), ex.get(), n, st.ohoff.get());
                hipLaunchKernelGGL(k_fill_ones_u64, dim3(grid_cap(ctx, (entries + 255) / 256, 16)), dim3(256), 0, s,
                                   st.ohtab.get(), entries);
                hipLaunchKernelGGL(k_out_hash_fill, dim3(grid_cap(ctx, (n * 64 + 255) / 256, 16)), dim3(256), 0, s,
                                   st.rrow.get(), st.rcol.get(), st.ostart.get(), st.ohoff.get(), n, st.ohtab.get());
                PGD_HIP(hipGetLastError());
                st.has_oh = true;
            }
        }
    }
    if (plan && st.n_heavy) wedge_plan_build(ctx, st);
}

// the counting launches (no host synchronisation, no allocation, no memset:
// capturable in a HIP graph).  With a timer, each pass is bracketed by events.
// Local passes (triangles, 4-cliques, Q), then the 4-cycle passes, which
// need the plan.
static void orient_count_local(pgd_ctx* ctx, OrientState& st, EdgeCounts out, Timer* timer, hipStream_t side);
static void orient_count_wedges(pgd_ctx* ctx, OrientState& st, EdgeCounts out, Timer* timer,
                                hipEvent_t folded = nullptr, hipEvent_t planned = nullptr);

void orient_count(pgd_ctx* ctx, OrientState& st, EdgeCounts out, Timer* timer) {
    orient_count_local(ctx, st, out, timer, nullptr);
    orient_count_wedges(ctx, st, out, timer);
}

// fork the side stream off the context's stream at this point (event i),
// and join it back (event i): no-ops without a side stream
static void fork_begin(pgd_ctx* ctx, OrientState& st, hipStream_t side, int i) {
    if (!side) return;
    if (!st.fork[i]) PGD_HIP(hipEventCreateWithFlags(&st.fork[i], hipEventDisableTiming));
    PGD_HIP(hipEventRecord(st.fork[i], ctx->stream));
    PGD_HIP(hipStreamWaitEvent(side, st.fork[i], 0));
}
static void fork_end(pgd_ctx* ctx, OrientState& st, hipStream_t side, int i) {
    if (!side) return;
    if (!st.fork[i]) PGD_HIP(hipEventCreateWithFlags(&st.fork[i], hipEventDisableTiming));
    PGD_HIP(hipEventRecord(st.fork[i], side));
    PGD_HIP(hipStreamWaitEvent(ctx->stream, st.fork[i], 0));
}

// side: a second stream for the wavefront and 256-thread bins, which then
// run beside the 1024-thread bins (those hold one block per CU and leave
// wave slots free); nullptr runs every bin on the context's stream
static void orient_count_local(pgd_ctx* ctx, OrientState& st, EdgeCounts out, Timer* timer, hipStream_t side) {
    hipStream_t s = ctx->stream;
    const int64_t m = st.m, n = st.n;
    const OrientParams& prm = st.prm;
    int gm = grid_cap(ctx, (m + 255) / 256, 16);
    hipLaunchKernelGGL(k_fill_u32, dim3(gm), dim3(256), 0, s, out.T, m);
    hipLaunchKernelGGL(k_fill_u64, dim3(gm), dim3(256), 0, s, out.K4, m);
    hipLaunchKernelGGL(k_fill_u64, dim3(gm), dim3(256), 0, s, st.P.get(), m);
    hipLaunchKernelGGL(k_fill_u64, dim3(gm), dim3(256), 0, s, st.Q.get(), m);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, (uint64_t*)st.work.get(), (int64_t)5);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, st.lq.get(), (int64_t)6);  // bin work queues
    PGD_HIP(hipGetLastError());
    unsigned long long* w = st.work.get();
    TriLists tl{st.tlist.get(), st.tcap, (unsigned long long*)st.tcur.get(), st.toff.get(), st.tlen.get()};
    if (st.tcap > 0) {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, st.tcur.get(), (int64_t)1);
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_cap(ctx, (n + 255) / 256, 16)), dim3(256), 0, s, st.tlen.get(), n,
                           -1);  // not kept, until the triangle pass writes a list
        PGD_HIP(hipGetLastError());
    } else {
        tl.G = nullptr;
    }
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_cap(ctx, (2 * m + 255) / 256, 16)), dim3(256), 0, s, st.TS.get(),
                       2 * m);
    const OutHash oh{st.ohtab.get(), st.ohoff.get(), st.has_oh ? prm.tri_search : 0};
    if (timer) timer->begin("k_local_tri");
    auto tri_small_mid = [&](hipStream_t q) {
        if (st.n_small)
            hipLaunchKernelGGL((k_local_tri<64, false>), dim3(grid_cap(ctx, st.n_small, prm.grid_small)), dim3(64),
                               st.small_lds_tri, q, st.rrow.get(), st.rcol.get(), st.reid.get(), st.ostart.get(),
                               st.outdeg.get(), st.b_small.get(), st.n_small, prm.small_k, (size_t)0,
                               (char*)nullptr, st.el_small, st.ecap_small, tl, st.TS.get(), out.T, oh,
                               (uint32_t*)nullptr, w);
        if (st.n_mid)
            hipLaunchKernelGGL((k_local_tri<256, false>), dim3(grid_cap(ctx, st.n_mid, prm.grid_mid)), dim3(256),
                               st.mid_lds_tri, q, st.rrow.get(), st.rcol.get(), st.reid.get(), st.ostart.get(),
                               st.outdeg.get(), st.b_mid.get(), st.n_mid, prm.mid_k, (size_t)0, (char*)nullptr,
                               st.el_mid, st.ecap_mid, tl, st.TS.get(), out.T, oh, st.lq.get() + 4, w);
    };
    fork_begin(ctx, st, side, 0);
    if (!side) tri_small_mid(s);
    if (st.n_large)
        hipLaunchKernelGGL((k_local_tri<1024, false>), dim3(grid_cap(ctx, st.n_large, 1)), dim3(1024),
                           st.large_lds_tri, s, st.rrow.get(), st.rcol.get(), st.reid.get(), st.ostart.get(),
                           st.outdeg.get(), st.b_large.get(), st.n_large, st.kq, (size_t)0, (char*)nullptr,
                           st.el_large, st.ecap_large, tl, st.TS.get(), out.T, oh, st.lq.get(), w);
    if (st.n_huge)
        hipLaunchKernelGGL((k_local_tri<1024, true>), dim3(st.ghuge), dim3(1024), 0, s, st.rrow.get(), st.rcol.get(),
                           st.reid.get(), st.ostart.get(), st.outdeg.get(), st.b_huge.get(), st.n_huge, st.maxk,
                           st.huge_stride, st.hscratch.get(), st.el_huge, st.ecap_huge, tl, st.TS.get(), out.T,
                           oh, st.lq.get() + 2, w);
    if (side) tri_small_mid(side);
    fork_end(ctx, st, side, 1);
    const int gz = grid_cap(ctx, (2 * m + 255) / 256, 16);
    if (timer) timer->end();
    if (timer) timer->begin("k_local_k4q");
    // the clique pass addresses edges by out-slot: T by out-slot (TS, as the
    // triangle pass left it) and per-slot K4/Q accumulators (k0/k1, free
    // after the build)
    uint32_t* TSL = st.TS.get();
    uint64_t *K4S = st.k0.get(), *QS = st.k0.get() + 2 * m, *PK = st.k1.get();
    const PackSplit pk = pack_split(st.max_deg);
    hipLaunchKernelGGL(k_fill_u64, dim3(gz), dim3(256), 0, s, K4S, 4 * m);
    hipLaunchKernelGGL(k_fill_u64, dim3(gz), dim3(256), 0, s, PK, 2 * m);
    PGD_HIP(hipGetLastError());
    auto k4q_small_mid = [&](hipStream_t q) {
        if (st.n_small)
            hipLaunchKernelGGL((k_local_k4q<64, false>), dim3(grid_cap(ctx, st.n_small, prm.grid_small)), dim3(64),
                               st.small_lds, q, st.rrow.get(), st.rcol.get(), st.reid.get(), st.ostart.get(),
                               st.outdeg.get(), st.b_small.get(), st.n_small, prm.small_k, st.mw_small,
                               (char*)nullptr, st.el_small, st.ecap_small, tl, TSL, K4S, QS, PK, pk,
                               (uint32_t*)nullptr, w + 1);
        if (st.n_mid)
            hipLaunchKernelGGL((k_local_k4q<256, false>), dim3(grid_cap(ctx, st.n_mid, prm.grid_mid)), dim3(256),
                               st.mid_lds, q, st.rrow.get(), st.rcol.get(), st.reid.get(), st.ostart.get(),
                               st.outdeg.get(), st.b_mid.get(), st.n_mid, prm.mid_k, st.mw_mid, (char*)nullptr,
                               st.el_mid, st.ecap_mid, tl, TSL, K4S, QS, PK, pk, st.lq.get() + 5, w + 1);
    };
    fork_begin(ctx, st, side, 2);
    if (!side) k4q_small_mid(s);
    if (st.n_large)
        hipLaunchKernelGGL((k_local_k4q<1024, false>), dim3(grid_cap(ctx, st.n_large, 1)), dim3(1024),
                           st.large_lds, s, st.rrow.get(), st.rcol.get(), st.reid.get(), st.ostart.get(),
                           st.outdeg.get(), st.b_large.get(), st.n_large, st.kq, st.mw_large, (char*)nullptr,
                           st.el_large, st.ecap_large, tl, TSL, K4S, QS, PK, pk, st.lq.get() + 1, w + 1);
    if (st.n_huge)
        hipLaunchKernelGGL((k_local_k4q<1024, true>), dim3(st.ghuge), dim3(1024), 0, s, st.rrow.get(), st.rcol.get(),
                           st.reid.get(), st.ostart.get(), st.outdeg.get(), st.b_huge.get(), st.n_huge, st.maxk,
                           st.mw_huge, st.hscratch.get(), (uint64_t*)nullptr, (int64_t)0, tl, TSL, K4S, QS, PK, pk,
                           st.lq.get() + 3, w + 1);
    if (side) k4q_small_mid(side);
    fork_end(ctx, st, side, 3);
    // with a side stream the folds of the per-slot sums run there, beside
    // the 4-cycle passes (which use neither their inputs nor their outputs);
    // fork[5], recorded after them, is the join before k_combine.  The
    // triangle fold (T by edge id, from TS; the clique pass reads TS only)
    // comes here too: launched ahead of the clique bins it raced the
    // 1024-thread bin for the CUs, and when it lost, the side stream's bins
    // waited behind it (the clique pass then ended about 18 ms later)
    fork_begin(ctx, st, side, 4);
    hipLaunchKernelGGL(k_fold_u32, dim3(gz), dim3(256), 0, side ? side : s, st.TS.get(), st.reid.get(), 2 * m,
                       out.T);
    hipLaunchKernelGGL(k_fold_k4q, dim3(gz), dim3(256), 0, side ? side : s, K4S, QS, PK, pk.qb, st.reid.get(), 2 * m,
                       out.K4, st.Q.get());
    PGD_HIP(hipGetLastError());
    if (side) {
        if (!st.fork[5]) PGD_HIP(hipEventCreateWithFlags(&st.fork[5], hipEventDisableTiming));
        PGD_HIP(hipEventRecord(st.fork[5], side));
    }
    if (timer) timer->end();
}

// folded (optional): recorded where K4 and Q are complete; waited for
// before k_combine.  planned (optional): recorded where the 4-cycle plan is
// complete; waited for after the light wedges, before the tiled pass.
static void orient_count_wedges(pgd_ctx* ctx, OrientState& st, EdgeCounts out, Timer* timer, hipEvent_t folded,
                                hipEvent_t planned) {
    hipStream_t s = ctx->stream;
    const int64_t m = st.m;
    const OrientParams& prm = st.prm;
    const int gm = grid_cap(ctx, (m + 255) / 256, 16);
    unsigned long long* w = st.work.get();
    if (timer) timer->begin("k_wedge_light");
    if (st.n_light)
        hipLaunchKernelGGL(k_wedge_light, dim3(grid_cap(ctx, st.n_light, prm.grid_small)), dim3(64), 0, s,
                           st.rrow.get(), st.rcol.get(), st.reid.get(), st.b_light.get(), st.n_light, st.P.get(),
                           w + 2);
    PGD_HIP(hipGetLastError());
    if (timer) timer->end();
    if (planned) PGD_HIP(hipStreamWaitEvent(s, planned, 0));
    if (timer) timer->begin("k_wedge_tiled");
    if (st.n_heavy) wedge_plan_count(ctx, st, w + 3);
    if (timer) timer->end();
    if (folded) PGD_HIP(hipStreamWaitEvent(s, folded, 0));
    if (timer) timer->begin("k_combine");
    hipLaunchKernelGGL(k_combine, dim3(gm), dim3(256), 0, s, st.P.get(), st.Q.get(), out.T, out.K4, m, out.C4);
    PGD_HIP(hipGetLastError());
    if (timer) timer->end();
}

void orient_profile(pgd_ctx* ctx, OrientState& st, Timer& timer, Timing* tm) {
    unsigned long long wk[5] = {0, 0, 0, 0, 0};
    PGD_HIP(hipMemcpyAsync(wk, st.work.get(), sizeof(wk), hipMemcpyDeviceToHost, ctx->stream));
    PGD_HIP(hipStreamSynchronize(ctx->stream));
    // algorithmic HBM bytes per pass (work models):
    //   triangle pass: 4 B per out-list entry read;
    //   clique pass: per local edge and column block, its 8 B list entry and
    //     its two 8 B per-slot sums (an out-list entry searched counts as one
    //     local edge);
    //   light wedges: the 4 B endpoint id of every wedge, read by the count
    //     and the accumulate pass;
    //   tiled wedges: the 4 B endpoint id read by the count pass, and for
    //     the wedges of tiles with a repeated endpoint (the only ones the
    //     accumulate pass visits) a second 4 B read and the read-modify-write
    //     of the 4 B per-slot sum (4 B read and 4 B written);
    //   combine: 3 x 8 B + 4 B read and 8 B written per edge.
    const double bytes_of[5] = {4.0 * wk[0], 24.0 * wk[1], 8.0 * wk[2], 4.0 * wk[3] + 12.0 * wk[4], 36.0 * st.m};
    const double units_of[5] = {(double)wk[0], (double)wk[1], (double)wk[2], (double)wk[3], (double)st.m};
    double best = 0;
    tm->phases.clear();
    for (size_t i = 0; i < timer.ev.size(); i++) {
        float ms = 0;
        PGD_HIP(hipEventElapsedTime(&ms, timer.ev[i].first, timer.ev[i].second));
        tm->phases.push_back(Phase{timer.names[i], ms, i < 5 ? bytes_of[i] : 0, i < 5 ? units_of[i] : 0});
        if (ms > best) best = ms;
    }
    tm->dominant_ms = best;
    tm->dominant_launches = 1;
}

void engine_orient(pgd_ctx* ctx, const GraphView& g, EdgeCounts out, Timing* tm, OrientState** cache,
                   const OrientParams& prm) {
    OrientState* st = cache ? *cache : nullptr;
    std::unique_ptr<OrientState> local;
    if (!st) {
        st = new OrientState();
        if (cache) *cache = st;
        else local.reset(st);
    }
    st->prm = prm;
    // the 4-cycle plan is built on a second stream while the local passes
    // run: its host synchronisations wait for that stream only
    orient_build(ctx, g, *st, false);
    Timer timer{ctx};
    hipEvent_t built = nullptr, planned = nullptr;
    if (st->n_heavy) {
        if (!ctx->aux) PGD_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
        PGD_HIP(hipEventCreateWithFlags(&built, hipEventDisableTiming));
        PGD_HIP(hipEventCreateWithFlags(&planned, hipEventDisableTiming));
        PGD_HIP(hipEventRecord(built, ctx->stream));
    }
    if (!ctx->side) PGD_HIP(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
    orient_count_local(ctx, *st, out, &timer, ctx->side);
    if (st->n_heavy) {
        hipStream_t main = ctx->stream;
        PGD_HIP(hipStreamWaitEvent(ctx->aux, built, 0));
        ctx->stream = ctx->aux;
        try {
            wedge_plan_build(ctx, *st);
        } catch (...) {
            ctx->stream = main;
            (void)hipStreamSynchronize(ctx->aux);
            (void)hipEventDestroy(built);
            (void)hipEventDestroy(planned);
            throw;
        }
        ctx->stream = main;
        (void)hipEventDestroy(built);
        try {
            PGD_HIP(hipEventRecord(planned, ctx->aux));
        } catch (...) {
            (void)hipEventDestroy(planned);
            throw;
        }
    }
    // the light wedges run before the wait for the plan
    try {
        orient_count_wedges(ctx, *st, out, &timer, st->fork[5], planned);
    } catch (...) {
        if (planned) (void)hipEventDestroy(planned);
        throw;
    }
    if (planned) (void)hipEventDestroy(planned);
    orient_profile(ctx, *st, timer, tm);
}

OrientState* orient_state_new() { return new OrientState(); }

const char* tuning_problem(const pgd_tuning& t) {
    if (t.small_k < 2 || t.small_k > 64) return "small_k must be in 2..64 (one wavefront)";
    if (t.mid_k < t.small_k || t.mid_k > 256) return "mid_k must be in small_k..256";
    if (t.grid_small < 1 || t.grid_small > 64 || t.grid_mid < 1 || t.grid_mid > 64)
        return "grid factors must be in 1..64 blocks per CU";
    if (t.wedge_log2r < 8 || t.wedge_log2r > 16) return "wedge_log2r must be in 8..16";
    if (t.wedge_nt != 256 && t.wedge_nt != 512 && t.wedge_nt != 1024) return "wedge_nt must be 256, 512 or 1024";
    if (t.wedge_split < 1 || t.wedge_split > 64) return "wedge_split must be in 1..64";
    if (t.tri_lists < 0 || t.tri_lists > 1024) return "tri_lists must be in 0..1024";
    if (t.tri_search < 0 || t.tri_search > 1024) return "tri_search must be in 0..1024";
    return "";
}

}  // namespace pgd
