// pgd_engine_wedge.hip -- the 4-cycle pass (P) of the oriented engine for
// heavy anchors, with the wedge endpoints counted in LDS tiles.
//
// A 4-cycle is found once, from the vertex c opposite its highest-ranked
// vertex a, as two wedges c-b-a and c-b'-a with a > max(b, c) and
// a > max(b', c).  For the anchor c, cnt_c[a] = number of wedges c-.-a; the
// wedge c-b-a adds cnt_c[a]-1 to the edges (b, a) and (c, b).  The endpoints
// of the wedges through b are the suffix of the ascending list N(b) above
// max(b, c); on a power-law graph b is mostly a hub and the suffix its
// out-list, so these runs are long (about 190 wedges on R-MAT scale 22,
// against 20 for the prefixes below a of the same wedges counted from a).
// cnt_c spans the ranks above c, so it is cut into endpoint tiles of
// 2^log2r ranks whose counters live in LDS.  For one anchor, the wedges of
// one tile are a set of segments: runs of the suffixes whose endpoints fall
// in the tile.  A per-vertex tile table (one entry per run of equal tiles in
// N(b)) lets a block bucket an anchor's segments by tile with a counting
// sort, then count (LDS atomics) and accumulate tile by tile.  The update of
// the edge (b, a) is added to a per-slot accumulator PS (consecutive lanes
// hit consecutive slots of N(b)), that of (c, b) once per segment, and PS is
// folded into P by edge id at the end.
//
// The work is cut into items (anchor, tile range) bounded in segments (the
// per-block buffer) and in wedges (load balance), and a persistent grid pulls
// them largest first from a device queue.
#include "pgd_orient.hpp"

namespace pgd {
namespace {

constexpr int kTmax = 128;       // tiles per item (LDS bucket table)
constexpr int kPlanLds = 2048;   // (hub, tile) cells counted in LDS by the planner

__global__ void w_fill_u64(uint64_t* __restrict__ p, int64_t n) {
    GRID_STRIDE(i, n) p[i] = 0;
}
__global__ void w_fill_u32(uint32_t* __restrict__ p, int64_t n) {
    GRID_STRIDE(i, n) p[i] = 0;
}

// outslot[e] = the slot of edge e in the list of its lower-ranked end (the
// out-slots of v are [ostart[v], row[v+1])); one wavefront per vertex
__global__ void k_out_slots(const int64_t* __restrict__ row, const int64_t* __restrict__ ostart,
                            const int32_t* __restrict__ eid, int64_t n, int32_t* __restrict__ outslot) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v = wave; v < n; v += nw)
        for (int64_t s = ostart[v] + lane; s < row[v + 1]; s += 64) outslot[eid[s]] = (int32_t)s;
}

// sfx[s] = the offset in N(b) of the first endpoint above max(b, c), for
// every slot s = (c, b) of a heavy anchor c: |N+(b)| from its end when
// b > c, the twin slot b -> c plus one when b < c.  Also the anchor's tile
// range: from the lowest tile of a suffix's first endpoint to the highest
// tile of a suffix's last one (nt = 0 when every suffix is empty); one
// wavefront per anchor
__global__ void k_pair_wedges(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                              const int64_t* __restrict__ ostart, const int32_t* __restrict__ eid,
                              const int32_t* __restrict__ outslot, const int32_t* __restrict__ heavy, int64_t nheavy,
                              int log2r, int32_t* __restrict__ sfx, int32_t* __restrict__ hbase,
                              int64_t* __restrict__ nt) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t h = wave; h < nheavy; h += nw) {
        const int32_t c = heavy[h];
        int32_t lo = INT32_MAX, hi = -1;
        for (int64_t s = row[c] + lane; s < row[c + 1]; s += 64) {
            const int32_t b = col[s];
            const int64_t rb = row[b], re = row[b + 1];
            const int64_t q = b > c ? ostart[b] : (int64_t)outslot[eid[s]] + 1;
            sfx[s] = (int32_t)(q - rb);
            if (q < re) {
                lo = min(lo, col[q]);
                hi = max(hi, col[re - 1]);
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            lo = min(lo, __shfl_xor(lo, off, 64));
            hi = max(hi, __shfl_xor(hi, off, 64));
        }
        if (lane == 0) {
            hbase[h] = hi >= 0 ? lo >> log2r : 0;
            nt[h] = hi >= 0 ? (int64_t)(hi >> log2r) - (lo >> log2r) + 1 : 0;
        }
    }
}

// tile-table flags: slot p starts a run of equal endpoint tiles in its list
__global__ void k_tt_flags(const int64_t* __restrict__ row, const int32_t* __restrict__ col, int64_t n, int log2r,
                           int32_t* __restrict__ flag) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t b = wave; b < n; b += nw) {
        const int64_t r0 = row[b], r1 = row[b + 1];
        for (int64_t p = r0 + lane; p < r1; p += 64)
            flag[p] = (p == r0) || ((col[p] >> log2r) != (col[p - 1] >> log2r));
    }
}

__global__ void k_tt_fill(const int64_t* __restrict__ row, const int32_t* __restrict__ col, int64_t n, int64_t nnz,
                          int log2r, const int32_t* __restrict__ flag, const int64_t* __restrict__ pos,
                          uint64_t* __restrict__ tt, int64_t* __restrict__ ttrow) {
    const int lane = threadIdx.x & 63;
    const int64_t total = nnz ? pos[nnz - 1] + flag[nnz - 1] : 0;  // entries
    const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t b = wave; b < n; b += nw) {
        const int64_t r0 = row[b], r1 = row[b + 1];
        if (lane == 0) ttrow[b] = r0 < nnz ? pos[r0] : total;
        for (int64_t p = r0 + lane; p < r1; p += 64)
            if (flag[p]) tt[pos[p]] = ((uint64_t)(uint32_t)(col[p] >> log2r) << 32) | (uint32_t)(p - r0);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ttrow[n] = total;
}

// first tile-table entry of b with tile >= t (entries ascending by tile)
__device__ __forceinline__ int64_t tt_seek(const uint64_t* __restrict__ tt, int64_t lo, int64_t hi, uint32_t t) {
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((uint32_t)(tt[mid] >> 32) < t) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// first tile-table entry of b, in [lo, hi), whose offset is >= w (entries
// ascending by offset)
__device__ __forceinline__ int64_t tt_seek_off(const uint64_t* __restrict__ tt, int64_t lo, int64_t hi, uint32_t w) {
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((uint32_t)tt[mid] < w) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// The tile-table entries of 64 pairs (c, b), flattened across a wavefront.
// Lane i describes pair i: entries [jb, je) of b's table (possibly empty;
// entry jb holds the suffix start q), the end of b's table, |N(b)| and q.
// fn(pair, tile, offset, len) is called once per entry with the part of the
// entry's run at or above q, consecutive lanes on consecutive entries of a
// pair; pair is the rank of the lane among the lanes with entries.  flag: 64 words of the wavefront's LDS (the start flag of position
// p of a round is byte p >> 6 of word p & 63, so each lane clears and reads
// its own word); pj, pe, pw, pd: 64 entries each, indexed by that rank.
template <class F>
__device__ __forceinline__ void walk_entries(int64_t jb, int64_t je, int64_t tend, uint32_t deg, uint32_t q0,
                                             int64_t* pj, int64_t* pe, uint32_t* pw, uint32_t* pd, uint32_t* flag,
                                             const uint64_t* __restrict__ tt, int lane, F&& fn) {
    const int32_t cnt = (int32_t)(je > jb ? je - jb : 0);
    int32_t x = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    const int32_t total = __shfl(x, 63, 64);
    const uint64_t nz = __ballot(cnt > 0);
    if (cnt > 0) {
        // owners are counted over the lanes with entries only
        const int r = __popcll(nz & ((1ULL << lane) - 1));
        pj[r] = jb - (x - cnt);  // entry index = pj[owner] + flattened position
        pe[r] = tend;
        pw[r] = q0;
        pd[r] = deg;
    }
    const int32_t es = cnt ? x - cnt : INT32_MAX;
    __builtin_amdgcn_wave_barrier();
    constexpr int kU = 4;
    for (int32_t rb0 = 0; rb0 < total; rb0 += 64 * kU) {
        flag[lane] = 0;
        __builtin_amdgcn_wave_barrier();
        if (es >= rb0 && es < rb0 + 64 * kU) ((uint8_t*)flag)[((es - rb0) & 63) * 4 + ((es - rb0) >> 6)] = 1;
        __builtin_amdgcn_wave_barrier();
        const uint32_t f4 = flag[lane];
        int cum = __popcll(__ballot(es < rb0)) - 1;
        int o[kU];
        int64_t j[kU];
        uint64_t e[kU], en[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const uint64_t starts = __ballot(((f4 >> (8 * u)) & 0xffu) != 0);
            const int32_t q = rb0 + 64 * u + lane;
            o[u] = -1;
            if (q < total) {
                o[u] = cum + __popcll(starts & ((2ULL << lane) - 1));
                j[u] = pj[o[u]] + q;
            }
            cum += __popcll(starts);
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
            if (o[u] >= 0) {
                e[u] = tt[j[u]];
                en[u] = j[u] + 1 < pe[o[u]] ? tt[j[u] + 1] : (uint64_t)pd[o[u]];
            }
#pragma unroll
        for (int u = 0; u < kU; u++)
            if (o[u] >= 0) {
                const uint32_t off = max((uint32_t)e[u], pw[o[u]]);
                fn(o[u], (uint32_t)(e[u] >> 32), off, (uint32_t)en[u] - off);
            }
        __builtin_amdgcn_wave_barrier();
    }
}

// segments and wedges per (anchor, tile) cell; one block per anchor, its
// pairs 64 per wavefront
template <int NT>
__global__ __launch_bounds__(NT) void k_plan_cells(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                                                   const int32_t* __restrict__ heavy, int64_t nheavy,
                                                   const int32_t* __restrict__ sfx, const uint64_t* __restrict__ tt,
                                                   const int64_t* __restrict__ ttrow, const int32_t* __restrict__ hbase,
                                                   const int64_t* __restrict__ htoff, uint32_t* __restrict__ htseg,
                                                   uint64_t* __restrict__ htw, unsigned long long* __restrict__ acc,
                                                   unsigned long long* hout, int32_t* __restrict__ pjb) {
    constexpr int NW = NT / 64;
    unsigned long long mseg = 0, mw = 0;  // this thread's segments and wedges
    __shared__ uint32_t lseg[kPlanLds];
    __shared__ unsigned long long lw[kPlanLds];
    __shared__ int64_t pj[NW][64], pe[NW][64];
    __shared__ uint32_t pw[NW][64], pd[NW][64];
    __shared__ uint32_t flag[NW][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int64_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
        const int32_t c = heavy[h];
        const int64_t c0 = htoff[h], T = htoff[h + 1] - c0;
        const uint32_t base = (uint32_t)hbase[h];
        const bool in_lds = T <= kPlanLds;
        if (in_lds)
            for (int i = threadIdx.x; i < T; i += NT) { lseg[i] = 0; lw[i] = 0; }
        __syncthreads();
        const int64_t ib = row[c], ie = row[c + 1];
        for (int64_t g0 = ib + (int64_t)wv * 64; g0 < ie; g0 += NT) {
            const int64_t s = g0 + lane;
            int64_t jb = 0, je = 0, tend = 0;
            uint32_t q = 0, deg = 0;
            if (s < ie) {
                const int32_t b = col[s];
                q = (uint32_t)sfx[s];
                deg = (uint32_t)(row[b + 1] - row[b]);
                if (q < deg) {
                    tend = ttrow[b + 1];
                    jb = tt_seek_off(tt, ttrow[b], tend, q + 1) - 1;  // the entry holding q
                    je = tend;
                    pjb[s] = (int32_t)jb;  // for the segment walk (k_plan_fill)
                }
            }
            walk_entries(jb, je, tend, deg, q, pj[wv], pe[wv], pw[wv], pd[wv], flag[wv], tt, lane,
                         [&](int, uint32_t tile, uint32_t, uint32_t len) {
                             const uint32_t t = tile - base;
                             mseg += 1;
                             mw += len;
                             if (in_lds) {
                                 atomicAdd(&lseg[t], 1u);
                                 atomicAdd(&lw[t], (unsigned long long)len);
                             } else {
                                 atomicAdd(&htseg[c0 + t], 1u);
                                 atomicAdd((unsigned long long*)&htw[c0 + t], (unsigned long long)len);
                             }
                         });
        }
        __syncthreads();
        if (in_lds)
            for (int i = threadIdx.x; i < T; i += NT) { htseg[c0 + i] = lseg[i]; htw[c0 + i] = lw[i]; }
        __syncthreads();
    }
    // the totals of all cells (acc[0], acc[1]; acc[2] counts finished
    // blocks): the last block to finish copies them to host memory, so the
    // host reads them after a stream synchronisation without a copy
    if (mseg) atomicAdd(&acc[0], mseg);
    if (mw) atomicAdd(&acc[1], mw);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(&acc[2], 1ull) == gridDim.x - 1) {
        __threadfence();
        hout[0] = atomicAdd(&acc[0], 0ull);
        hout[1] = atomicAdd(&acc[1], 0ull);
        __threadfence_system();
    }
}

// the segments of every (anchor, tile) cell, written from the cell's offset
// in the plan's segment array (the walk of k_plan_cells again, from the tile
// entries it found; within a cell the order is that of the adds, on which no
// sum depends).  A segment is the part of N(b)'s suffix in one tile: its
// first CSR position, its length (top bit: N(b) is a wide list) and the slot
// of the pair (anchor, b).
template <int NT>
__global__ __launch_bounds__(NT) void k_plan_fill(const int64_t* __restrict__ row, const int32_t* __restrict__ col,
                                                  const int32_t* __restrict__ heavy, int64_t nheavy,
                                                  const int32_t* __restrict__ sfx, const uint64_t* __restrict__ tt,
                                                  const int64_t* __restrict__ ttrow, const int32_t* __restrict__ pjb,
                                                  const int32_t* __restrict__ hbase,
                                                  const int64_t* __restrict__ htoff, const int64_t* __restrict__ cseg,
                                                  uint32_t* __restrict__ gcur, int32_t wide_deg,
                                                  WedgeSeg* __restrict__ gseg) {
    constexpr int NW = NT / 64;
    __shared__ uint32_t lcur[kPlanLds];
    __shared__ int64_t pj[NW][64], pe[NW][64];
    __shared__ uint32_t pw[NW][64], pd[NW][64];
    __shared__ uint32_t flag[NW][64];
    __shared__ int64_t prow[NW][64];   // row[b] of the wavefront's pairs, by rank
    __shared__ int32_t pslot[NW][64];  // their slots in the anchor's list
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int64_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
        const int32_t c = heavy[h];
        const int64_t c0 = htoff[h], T = htoff[h + 1] - c0;
        const uint32_t base = (uint32_t)hbase[h];
        const bool in_lds = T <= kPlanLds;
        if (in_lds)
            for (int i = threadIdx.x; i < T; i += NT) lcur[i] = 0;
        __syncthreads();
        const int64_t ib = row[c], ie = row[c + 1];
        for (int64_t g0 = ib + (int64_t)wv * 64; g0 < ie; g0 += NT) {
            const int64_t s = g0 + lane;
            int64_t jb = 0, je = 0, tend = 0, rb = 0;
            uint32_t q = 0, deg = 0;
            if (s < ie) {
                const int32_t b = col[s];
                q = (uint32_t)sfx[s];
                rb = row[b];
                deg = (uint32_t)(row[b + 1] - rb);
                if (q < deg) {
                    tend = ttrow[b + 1];
                    jb = pjb[s];  // the entry holding q, found by k_plan_cells
                    je = tend;
                }
            }
            // the pairs' ranks as walk_entries numbers them: lanes with entries
            const uint64_t nz = __ballot(je > jb);
            if (je > jb) {
                const int r = __popcll(nz & ((1ULL << lane) - 1));
                prow[wv][r] = rb;
                pslot[wv][r] = (int32_t)s;
            }
            walk_entries(jb, je, tend, deg, q, pj[wv], pe[wv], pw[wv], pd[wv], flag[wv], tt, lane,
                         [&](int pr, uint32_t tile, uint32_t off, uint32_t len) {
                             const uint32_t t = tile - base;
                             const uint32_t k = in_lds ? atomicAdd(&lcur[t], 1u) : atomicAdd(&gcur[c0 + t], 1u);
                             const uint32_t lw = len | (pd[wv][pr] >= (uint32_t)wide_deg ? 0x80000000u : 0u);
                             gseg[cseg[c0 + t] + k] = WedgeSeg{(uint32_t)(prow[wv][pr] + off), (int32_t)lw, pslot[wv][pr]};
                         });
        }
        __syncthreads();
    }
}

// cut each anchor's tiles into items: at most segcap segments, kTmax tiles
// and (unless a single tile exceeds it) wcap wedges; empty items are dropped.
// With out == nullptr only the number of items per anchor is written, and
// nitems[nheavy] is cleared (the scan's last term).  zero128, when given, is
// cleared by the first block (the item sort's bucket counters): these two
// words would otherwise each take a launch, and a launch of the plan waits
// for a free wave slot beside the persistent local passes.
__global__ void k_plan_items(const int64_t* __restrict__ htoff, const int32_t* __restrict__ hbase,
                             const uint32_t* __restrict__ htseg, const uint64_t* __restrict__ htw, int64_t nheavy,
                             int64_t segcap, uint64_t wcap,
                             int64_t* __restrict__ nitems, const int64_t* __restrict__ ioff,
                             WedgeItem* __restrict__ out, uint64_t* __restrict__ key, uint32_t* __restrict__ zero128) {
    if (blockIdx.x == 0) {
        if (!out && threadIdx.x == 0) nitems[nheavy] = 0;
        if (zero128)
            for (int i = threadIdx.x; i < 128; i += blockDim.x) zero128[i] = 0;
    }
    GRID_STRIDE(h, nheavy) {
        const int64_t c0 = htoff[h], T = htoff[h + 1] - c0;
        int64_t cnt = 0, t0 = 0, segs = 0;
        uint64_t wed = 0;
        auto close = [&](int64_t t1) {
            if (wed) {
                if (out) {
                    const int64_t i = ioff[h] + cnt;
                    out[i] = WedgeItem{(int32_t)h, (int32_t)(hbase[h] + t0), (int32_t)(hbase[h] + t1), 0};
                    key[i] = ~wed;  // ascending sort of ~w: largest first
                }
                cnt++;
            }
        };
        for (int64_t t = 0; t < T; t++) {
            const uint32_t s = htseg[c0 + t];
            const uint64_t w = htw[c0 + t];
            if (t > t0 && (segs + s > segcap || (wed && wed + w > wcap) || t - t0 >= kTmax)) {
                close(t);
                t0 = t;
                segs = 0;
                wed = 0;
            }
            segs += s;
            wed += w;
        }
        close(T);
        if (!out) nitems[h] = cnt;
    }
}

// item order by wedge count, largest first to a factor of two: bucket
// 63 - floor(log2 w) (key = ~w, w >= 1); cnt[64] counts, then cnt[64 + d]
// is bucket d's next position.  A wavefront adds once per distinct bucket
// of its 64 items (the items of one hub mostly share one), so the few
// bucket counters are not contended item by item.
__device__ __forceinline__ int item_bucket(uint64_t key) { return __clzll((long long)~key); }
__global__ void k_item_buckets(const uint64_t* __restrict__ key, int64_t n, uint32_t* __restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); i0 < n; i0 += stride) {
        const int64_t i = i0 + lane;
        const bool ok = i < n;
        const int b = ok ? item_bucket(key[i]) : -1;
        uint64_t pending = __ballot(ok);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int lb = __shfl(b, leader, 64);
            const uint64_t same = __ballot(ok && b == lb);
            if (lane == leader) atomicAdd(&cnt[lb], (uint32_t)__popcll(same));
            pending &= ~same;
        }
    }
}
__global__ void k_bucket_starts(uint32_t* __restrict__ cnt) {  // one wavefront
    const int lane = threadIdx.x & 63;
    const uint32_t c = cnt[lane];
    uint32_t x = c;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    cnt[64 + lane] = x - c;
}
__global__ void k_item_scatter(const uint64_t* __restrict__ key, const WedgeItem* __restrict__ in, int64_t n,
                               uint32_t* __restrict__ next, WedgeItem* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); i0 < n; i0 += stride) {
        const int64_t i = i0 + lane;
        const bool ok = i < n;
        const int b = ok ? item_bucket(key[i]) : -1;
        uint64_t pending = __ballot(ok);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int lb = __shfl(b, leader, 64);
            const uint64_t same = __ballot(ok && b == lb);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&next[lb], (uint32_t)__popcll(same));
            base = __shfl(base, leader, 64);
            if ((same >> lane) & 1) out[base + __popcll(same & ((1ull << lane) - 1))] = in[i];
            pending &= ~same;
        }
    }
}

// add v to the per-slot accumulator of the slot: 32-bit unless its list
// (segment o of the group) is wide.  A slot (b, c) gains at most d(c) - 1
// from each of at most d(b) hubs, below 2^32 unless d(b) * max degree
// reaches it.
__device__ __forceinline__ void slot_add(uint64_t* __restrict__ PS, uint32_t* __restrict__ PS32, uint64_t wmask,
                                         int o, int64_t slot, uint32_t v) {
    if ((wmask >> o) & 1) atomicAdd((unsigned long long*)&PS[slot], (unsigned long long)v);
    else atomicAdd(&PS32[slot], v);
}

// One pass of the block over the segments of a tile, 64 segments (a group)
// at a time.  pass 0 counts endpoints, pass 1 adds cnt-1 per wedge to the
// slot's accumulator (PS32, or PS for a wide list) and, per segment, to the
// anchor's edge (c, b) (PS).  mode 0: 16-bit counters over the whole tile;
// with chk, a count reaching 65536 sets *ovf (without chk the anchor has
// fewer than 65536 neighbours, so no count can); mode 1/2: 32-bit counters
// over the lower/upper half of the tile, other endpoints skipped.
//
// Every wavefront scans the group's lengths in registers; the group's
// flattened wedges are walked in rounds of 64 kU positions, round r by
// wavefront r mod NW (a group's segments are few and long: the rounds, not
// the segments, are shared out).  Every lane computes kU positions per
// round.  Positions past the end are clamped to the last one (same owner,
// same slot), so the owner lookups, the endpoint loads and the counter reads
// are issued for all lanes without branches and overlap; only the updates
// are predicated.  gbase/gslot (the group's slot bases and anchor slots) are
// written by wavefront 0; gpsum collects the per-segment sums of pass 1,
// flushed by wavefront 0 after the group's closing barrier.
template <int NT>
__device__ __forceinline__ void tile_pass(int pass, int mode, bool chk, uint32_t sb, uint32_t se, int32_t c0,
                                          int log2r, const WedgeSeg* __restrict__ seg,
                                          const int32_t* __restrict__ col, uint32_t* cnt,
                                          uint64_t* __restrict__ PS, uint32_t* __restrict__ PS32,
                                          int64_t* gbase, int2* gbnd, int32_t* gslot, unsigned long long* gpsum,
                                          uint32_t* wflag, int* ovf, int* rep, int lane, int w,
                                          unsigned long long* done) {
    constexpr int NW = NT / 64;
    constexpr int kU = 4;
    const int32_t half = 1 << (log2r - 1);
    const int32_t lo = mode == 2 ? half : 0;
    bool again = false;  // mode 0, pass 0: some endpoint of this wavefront's wedges was seen before
    // mode 0: the counter of endpoint c is the low (c < half) or high half of
    // word c mod half, so consecutive endpoints use consecutive words (banks)
    const int hb = log2r - 1;
    const int32_t hm = half - 1;
    for (uint32_t gb = sb; gb < se; gb += 64) {
        const int kc = (int)min<uint32_t>(64, se - gb);
        int32_t len = 0, sl = 0;
        int64_t st = 0;
        bool wide = false;
        if (lane < kc) {
            const WedgeSeg g = seg[gb + lane];
            len = g.len & 0x7fffffff;
            wide = g.len < 0;
            st = g.start;
            sl = g.sab;
        }
        const uint64_t wmask = __ballot(wide);  // segments of wide lists
        // pass 1 lays each segment out on the 16-slot (64-byte) lines of its
        // slots: its positions start at a multiple of 16, with st mod 16
        // unused positions before its first slot and padding after its last
        // up to the next multiple of 16.  The 16 lanes of an add instruction
        // that share a line then take that line's slots only, and each line
        // of a segment is one atomic request (unaligned, a 64-lane run of
        // slots spans five lines and shares one with the next run).
        const int32_t front = pass ? (int32_t)(st & 15) : 0;
        const int32_t plen = pass && lane < kc ? (front + len + 15) & ~15 : len;
        const int32_t x = (int32_t)wave_incl_sum((uint32_t)plen);
        if (w == 0) {
            // slot of flattened position q in segment i: gbase[i] + q, for q in
            // [gbnd[i].x, gbnd[i].y)
            gbase[lane] = st - front - (int64_t)(x - plen);
            gbnd[lane] = make_int2(x - plen + front, x - plen + front + len);
            gslot[lane] = sl;
            gpsum[lane] = 0;
        }
        __syncthreads();
        const int32_t total = __shfl(x, 63, 64);
        // wedges counted (once per tile) and wedges accumulated (tiles with
        // a repeated endpoint; an overflowing tile once)
        if (pass ? mode < 2 : mode == 0) {
            const int32_t nw = pass ? __shfl((int32_t)wave_incl_sum((uint32_t)len), 63, 64) : total;
            if (w == 0 && lane == 0) done[pass] += (unsigned long long)nw;
        }
        // (the owner of position q, the segment holding it, is the last segment
        // starting at or before q: the segment starts inside a round are
        // flagged in LDS and read back as one 64-bit ballot per step, so the
        // owner is a popcount, not a search -- segments are non-empty, so their
        // starts are distinct)
        const int32_t es = lane < kc ? x - plen : INT32_MAX;  // exclusive start of segment `lane`
        const uint64_t upto = (2ULL << lane) - 1;             // lanes 0..lane
        if (pass == 0 && mode == 0) {
            // the count pass takes two rounds per step: the owners of both
            // are looked up first, so each lane has 2 kU endpoint loads in
            // flight before the counter adds that need them
            constexpr int kB = 2 * kU;
            for (int32_t rb0 = w * 64 * kB; rb0 < total; rb0 += NW * 64 * kB) {
                bool ok[kB];
                int64_t slot[kB];
                int32_t c[kB];
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int32_t r0 = rb0 + h * 64 * kU;
                    wflag[lane] = 0;
                    __builtin_amdgcn_wave_barrier();
                    if (es >= r0 && es < r0 + 64 * kU) ((uint8_t*)wflag)[((es - r0) & 63) * 4 + ((es - r0) >> 6)] = 1;
                    __builtin_amdgcn_wave_barrier();
                    const uint32_t f4 = wflag[lane];
                    int cum = __popcll(__ballot(es < r0)) - 1;
#pragma unroll
                    for (int u = 0; u < kU; u++) {
                        const uint64_t st4 = __ballot(((f4 >> (8 * u)) & 0xffu) != 0);
                        const int32_t q = r0 + 64 * u + lane;
                        const int o = cum + __popcll(st4 & upto);
                        cum += __popcll(st4);
                        ok[h * kU + u] = q < total;
                        slot[h * kU + u] = gbase[o] + (q < total ? q : total - 1);
                    }
                    __builtin_amdgcn_wave_barrier();
                }
#pragma unroll
                for (int u = 0; u < kB; u++) c[u] = col[slot[u]] - c0;
                // the counts before the add tell whether an endpoint repeats
                // (a tile without repeats closes no 4-cycle: its accumulate
                // pass is skipped) and, with chk, whether a count would pass
                // 65535
                uint32_t old[kB];
#pragma unroll
                for (int u = 0; u < kB; u++)
                    old[u] = ok[u] ? atomicAdd(&cnt[c[u] & hm], 1u << ((c[u] >> hb) * 16)) : 0u;
                bool hit = false;
#pragma unroll
                for (int u = 0; u < kB; u++) {
                    const uint32_t h = (old[u] >> ((c[u] >> hb) * 16)) & 0xffffu;
                    hit |= h == 0xffffu;
                    again |= h != 0;
                }
                if (chk && hit) *ovf = 1;
            }
        } else
        for (int32_t rb0 = w * 64 * kU; rb0 < total; rb0 += NW * 64 * kU) {
            int o[kU];
            bool ok[kU];
            int64_t slot[kU];
            int32_t c[kU];
            // the start flag of position p is byte p >> 6 of word p & 63: each
            // lane clears and reads one word
            wflag[lane] = 0;
            __builtin_amdgcn_wave_barrier();
            if (es >= rb0 && es < rb0 + 64 * kU) ((uint8_t*)wflag)[((es - rb0) & 63) * 4 + ((es - rb0) >> 6)] = 1;
            __builtin_amdgcn_wave_barrier();
            const uint32_t f4 = wflag[lane];
            const int cum0 = __popcll(__ballot(es < rb0)) - 1;  // last segment starting before this round
            int cum = cum0;
            uint64_t starts[kU];
            int cend[kU];  // the owner of the slice's last lane
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const uint32_t own = ((f4 >> (8 * u)) & 0xffu) != 0;
                starts[u] = __ballot(own);
                const int32_t q = rb0 + 64 * u + lane;
                // past the end: the owner is the last segment; clamp to its
                // last position.  Owner = starts at lanes below, plus this one.
                o[u] = cum + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(starts[u] >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)starts[u], own));
                cum += __popcll(starts[u]);
                cend[u] = cum;
                if (pass) {  // padding positions: clamped to the segment's slots, no update
                    const int2 bd = gbnd[o[u]];
                    ok[u] = q >= bd.x && q < bd.y;
                    slot[u] = gbase[o[u]] + min(max(q, bd.x), bd.y - 1);
                } else {
                    ok[u] = q < total;
                    slot[u] = gbase[o[u]] + (ok[u] ? q : total - 1);
                }
            }
#pragma unroll
            for (int u = 0; u < kU; u++) c[u] = col[slot[u]] - c0;
            if (mode == 0) {  // (the count pass of mode 0 is the two-round loop above)
                uint32_t v[kU];
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    v[u] = ((cnt[c[u] & hm] >> ((c[u] >> hb) * 16)) & 0xffffu) - 1;
                    v[u] = ok[u] ? v[u] : 0u;
                }
                if (wmask == 0) {  // no wide list in the group: 32-bit sums only
#pragma unroll
                    for (int u = 0; u < kU; u++)
                        if (v[u]) atomicAdd(&PS32[slot[u]], v[u]);
                } else {
#pragma unroll
                    for (int u = 0; u < kU; u++)
                        if (v[u]) slot_add(PS, PS32, wmask, o[u], slot[u], v[u]);
                }
                // the edge (c, b) of each segment.  The lanes of one segment
                // form runs delimited by segment starts.  A slice without a
                // start continues the open run (its owner oo); a slice with
                // starts closes it, adds the runs between its first and last
                // start (a wavefront scan; the last lane of each run adds),
                // and opens the run of its last start.  v < 2^16, so the sums
                // of a round fit 32 bits.
                uint32_t open = 0;
                int oo = cum0;
                auto flush = [&](uint32_t x, int owner) {
                    const uint32_t t = __builtin_amdgcn_readlane(wave_incl_sum(x), 63);
                    if (lane == 0 && t) atomicAdd(&gpsum[owner], (unsigned long long)t);
                };
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    const uint64_t sm = starts[u];
                    if (sm == 0) {
                        open += v[u];
                        continue;
                    }
                    const int f = (int)__builtin_ctzll(sm);       // first start
                    const int l = 63 - (int)__builtin_clzll(sm);  // last start
                    if (lane < f) open += v[u];
                    if (oo >= 0) flush(open, oo);
                    if (l > f) {
                        const uint32_t val = lane >= f && lane < l ? v[u] : 0u;
                        const uint32_t P = wave_incl_sum(val);
                        const uint64_t below = sm & upto;
                        const int rs = 63 - (int)__builtin_clzll(below | 1ull);  // run start (lane; >= f here)
                        const uint32_t before = __shfl(P, rs > 0 ? rs - 1 : 0, 64);
                        const uint32_t run = P - (rs > 0 ? before : 0u);
                        const bool last = lane < l && ((sm >> (lane + 1)) & 1);
                        if (last && run) atomicAdd(&gpsum[o[u]], (unsigned long long)run);
                    }
                    open = lane >= l ? v[u] : 0u;
                    oo = cend[u];
                }
                if (oo >= 0) flush(open, oo);
            } else {
                bool in[kU];
                int32_t ci[kU];
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    ci[u] = c[u] - lo;
                    in[u] = ok[u] && ci[u] >= 0 && ci[u] < half;
                    ci[u] = min(max(ci[u], 0), half - 1);
                }
                if (pass == 0) {
#pragma unroll
                    for (int u = 0; u < kU; u++)
                        if (in[u]) atomicAdd(&cnt[ci[u]], 1u);
                    continue;
                }
                uint32_t v[kU];
#pragma unroll
                for (int u = 0; u < kU; u++) v[u] = cnt[ci[u]] - 1;
#pragma unroll
                for (int u = 0; u < kU; u++)
                    if (in[u] && v[u]) {
                        slot_add(PS, PS32, wmask, o[u], slot[u], v[u]);
                        atomicAdd(&gpsum[o[u]], (unsigned long long)v[u]);
                    }
            }
        }
        // the group's closing barrier: every wavefront is done with gbase,
        // gslot and gpsum before wavefront 0 flushes and rewrites them
        __syncthreads();
        if (pass == 1 && w == 0 && lane < kc && gpsum[lane])
            atomicAdd((unsigned long long*)&PS[gslot[lane]], gpsum[lane]);
    }
    if (pass == 0 && mode == 0 && __ballot(again) && lane == 0) *rep = 1;
}

// A tile of at most 64 segments and few wedges, taken by one wavefront on
// its own: the endpoints are counted in a private open-addressing table of
// tw words (entry = endpoint << 16 | count, 0 = empty) instead of the
// block's counter array, so the block neither clears that array nor
// synchronises for the tile.  Returns false, having changed nothing, when
// the tile has more than min(64 kU, tw / 2) wedges (one round per lane, the
// table at most half full).
template <int kU>
__device__ __forceinline__ bool small_tile(uint32_t sb, uint32_t se, int32_t c0, const WedgeSeg* __restrict__ seg,
                                        const int32_t* __restrict__ col, uint32_t* tab, int tbits,
                                        uint64_t* __restrict__ PS, uint32_t* __restrict__ PS32, int64_t* wbase,
                                        int32_t* wslot, unsigned long long* wpsum, uint32_t* wflag, int lane,
                                        unsigned long long* done) {
    const int kc = (int)(se - sb);
    int32_t len = 0, sl = 0;
    int64_t st = 0;
    bool wide = false;
    if (lane < kc) {
        const WedgeSeg g = seg[sb + lane];
        len = g.len & 0x7fffffff;
        wide = g.len < 0;
        st = g.start;
        sl = g.sab;
    }
    const uint64_t wmask = __ballot(wide);
    const int32_t x = wave_incl_sum((uint32_t)len);
    const int32_t total = __shfl(x, 63, 64);
    if (total > min(64 * kU, 1 << (tbits - 1))) return false;
    const int tw = 1 << tbits;
    for (int i = lane; i < tw / 4; i += 64) ((uint4*)tab)[i] = make_uint4(0u, 0u, 0u, 0u);
    wbase[lane] = st - (int64_t)(x - len);
    wslot[lane] = sl;
    wpsum[lane] = 0;
    wflag[lane] = 0;
    __builtin_amdgcn_wave_barrier();
    // owners of the positions: as in the block passes, with one round
    const int32_t es = lane < kc ? x - len : INT32_MAX;
    const uint64_t upto = (2ULL << lane) - 1;
    if (es < 64 * kU) ((uint8_t*)wflag)[(es & 63) * 4 + (es >> 6)] = 1;
    __builtin_amdgcn_wave_barrier();
    const uint32_t f4 = wflag[lane];
    int o[kU];
    bool ok[kU];
    int64_t slot[kU];
    int32_t c[kU];
    uint64_t starts[kU];
    int cum = -1;
#pragma unroll
    for (int u = 0; u < kU; u++) {
        starts[u] = __ballot(((f4 >> (8 * u)) & 0xffu) != 0);
        const int32_t q = 64 * u + lane;
        ok[u] = q < total;
        o[u] = cum + __popcll(starts[u] & upto);
        cum += __popcll(starts[u]);
        slot[u] = wbase[o[u]] + (ok[u] ? q : total - 1);
    }
#pragma unroll
    for (int u = 0; u < kU; u++) c[u] = col[slot[u]] - c0;
    // count: claim the endpoint's entry (or find it) and add one
    const uint32_t tmask = (uint32_t)tw - 1;
    uint32_t pos[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
        pos[u] = 0;
        if (ok[u]) {
            const uint32_t key = (uint32_t)c[u] << 16;
            uint32_t sp = ((uint32_t)c[u] * 0x9E3779B1u) >> (32 - tbits);
            for (;;) {
                const uint32_t e = atomicCAS(&tab[sp], 0u, key | 1u);
                if (e == 0u) break;
                if ((e & 0xffff0000u) == key) {
                    atomicAdd(&tab[sp], 1u);
                    break;
                }
                sp = (sp + 1) & tmask;
            }
            pos[u] = sp;
        }
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t v[kU];
    bool again = false;
#pragma unroll
    for (int u = 0; u < kU; u++) {
        v[u] = ok[u] ? (tab[pos[u]] & 0xffffu) - 1 : 0u;
        again |= v[u] != 0;
    }
    if (lane == 0) done[0] += (unsigned long long)total;
    if (!__ballot(again)) return true;  // every count 1: no 4-cycle closes here
    if (lane == 0) done[1] += (unsigned long long)total;
#pragma unroll
    for (int u = 0; u < kU; u++)
        if (v[u]) slot_add(PS, PS32, wmask, o[u], slot[u], v[u]);
#pragma unroll
    for (int u = 0; u < kU; u++) {
        const uint32_t P = wave_incl_sum(v[u]);
        const uint64_t sm = starts[u] & upto;
        const int rs = sm ? 63 - __clzll((long long)sm) : 0;
        const uint32_t before = __shfl(P, rs > 0 ? rs - 1 : 0, 64);
        const uint32_t run = P - (rs > 0 ? before : 0u);
        const bool last = lane == 63 || ((starts[u] >> (lane + 1)) & 1);
        if (last && run) atomicAdd(&wpsum[o[u]], (unsigned long long)run);
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < kc && wpsum[lane]) atomicAdd((unsigned long long*)&PS[wslot[lane]], wpsum[lane]);
    __builtin_amdgcn_wave_barrier();
    return true;
}

// zero n words of LDS (n a multiple of 4, p 16-byte aligned) with 16-byte stores
template <int NT>
__device__ __forceinline__ void clear_lds(uint32_t* p, int n) {
    uint4* q = (uint4*)p;
    for (int i = threadIdx.x; i < n / 4; i += NT) q[i] = make_uint4(0u, 0u, 0u, 0u);
}

// the tiled count/accumulate kernel (persistent; items from the queue).
// Within a tile each wavefront takes 64 segments at a time, scans their
// lengths in registers and walks the flattened wedges 64 lanes at a time
// (consecutive lanes read consecutive slots of a list); the block only
// synchronises between the count and the accumulate pass of a tile.
template <int NT>
__global__ __launch_bounds__(NT) void k_wedge_tiled(
    const int64_t* __restrict__ row, const int32_t* __restrict__ col,
    const int32_t* __restrict__ heavy, const int32_t* __restrict__ hbase, const int64_t* __restrict__ htoff,
    const int64_t* __restrict__ cseg, const WedgeSeg* __restrict__ gseg, const WedgeItem* __restrict__ items,
    int64_t nitems, int log2r, uint32_t* __restrict__ qctr, uint64_t* __restrict__ PS,
    uint32_t* __restrict__ PS32, unsigned long long* __restrict__ work) {
    constexpr int NW = NT / 64;
    // 2^log2r 16-bit endpoint counters in 2^(log2r-1) words (c and c + 2^(log2r-1) share one)
    extern __shared__ __align__(16) uint32_t cnt[];
    __shared__ uint32_t base[kTmax + 1];
    __shared__ uint32_t cur[kTmax];
    __shared__ uint32_t wflag[NW][64];  // segment-start flags of one round of a wavefront (4 bytes per lane)
    // a wavefront's small-tile segment bases, slots and sums.  With 1024
    // threads they live in the counter array beside the small-tile tables
    // (the block passes clear it before use), so that two blocks of 16
    // wavefronts fit a CU's LDS with 2^15-rank tiles.
    constexpr bool kMetaInCnt = NT == 1024;
    constexpr int kMetaWords = 64 * 2 + 64 + 64 * 2;
    __shared__ int64_t wbase_s[kMetaInCnt ? 1 : NW][64];
    __shared__ int32_t wslot_s[kMetaInCnt ? 1 : NW][64];
    __shared__ unsigned long long wpsum_s[kMetaInCnt ? 1 : NW][64];
    __shared__ int64_t gbase[64];  // the block passes' current group of segments
    __shared__ int2 gbnd[64];
    __shared__ int32_t gslot[64];
    __shared__ unsigned long long gpsum[64];
    __shared__ int64_t item_sh;
    __shared__ int ovf, rep;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int R = 1 << log2r;
    // a wavefront's small-tile table: 2^tbits words of the counter array's R/2
    // (followed by the wavefronts' small-tile segment data when it is there)
    const int meta = kMetaInCnt ? kMetaWords : 0;
    int tbits = 9;
    while (tbits > 2 && NW * ((1 << tbits) + meta) > R / 2) tbits--;
    const bool small_fits = NW * ((1 << tbits) + meta) <= R / 2;
    int64_t* wbase;
    int32_t* wslot;
    unsigned long long* wpsum;
    if constexpr (kMetaInCnt) {
        uint32_t* m = cnt + (NW << tbits) + w * kMetaWords;  // 8-byte aligned: both terms are even
        wbase = (int64_t*)m;
        wslot = (int32_t*)(m + 128);
        wpsum = (unsigned long long*)(m + 192);
    } else {
        wbase = wbase_s[w];
        wslot = wslot_s[w];
        wpsum = wpsum_s[w];
    }
    unsigned long long done[2] = {0, 0};  // wedges counted, wedges accumulated
    for (;;) {
        if (tid == 0) item_sh = (int64_t)atomicAdd(qctr, 1u);
        __syncthreads();
        const int64_t it = item_sh;
        __syncthreads();
        if (it >= nitems) break;
        const WedgeItem W = items[it];
        const int32_t a = heavy[W.hub];  // the anchor
        const int64_t ib = row[a], ie = row[a + 1];
        const int T = W.t1 - W.t0;
        // the item's segments: cells [first, first + T) of the plan's segment
        // array, in tile order
        const int64_t first = htoff[W.hub] + (W.t0 - hbase[W.hub]);
        const int64_t s0 = cseg[first];
        for (int i = tid; i <= T; i += NT) base[i] = (uint32_t)(cseg[first + i] - s0);
        const WedgeSeg* seg = gseg + s0;
        __syncthreads();
        // tile by tile: count the endpoints, then accumulate cnt-1 per wedge.
        // Counters are 16-bit, two per word (mode 0); a count that would pass
        // 65535 (an anchor with that many neighbours on one endpoint) flags
        // the tile, which is then recounted as two half tiles with 32-bit
        // counters in the same LDS (modes 1 and 2).
        // a count is at most the anchor's degree: below 65536 no 16-bit counter can overflow
        const bool chk = ie - ib > 65535;
        // small tiles first, one per wavefront in the wavefront's slice of
        // the counter array; cur[t] then marks the tiles left for the block
        for (int t = w; t < T; t += NW) {
            const uint32_t sb = base[t], se = base[t + 1];
            uint32_t left = se - sb > 64u || !small_fits;
            if (!left && sb < se)
                left = !small_tile<4>(sb, se, (W.t0 + t) << log2r, seg, col, cnt + w * (1 << tbits), tbits, PS,
                                      PS32, wbase, wslot, wpsum, wflag[w], lane, done);
            if (lane == 0) cur[t] = left;
        }
        __syncthreads();
        for (int t = 0; t < T; t++) {
            if (!cur[t]) continue;
            const uint32_t sb = base[t], se = base[t + 1];
            const int32_t c0 = (W.t0 + t) << log2r;
            clear_lds<NT>(cnt, R / 2);
            if (tid == 0) {
                ovf = 0;
                rep = 0;
            }
            __syncthreads();
            tile_pass<NT>(0, 0, chk, sb, se, c0, log2r, seg, col, cnt, PS, PS32, gbase, gbnd, gslot, gpsum, wflag[w], &ovf, &rep,
                          lane, w, done);
            __syncthreads();
            if (!ovf) {
                // every count 1: no 4-cycle closes in this tile
                if (rep) {
                    tile_pass<NT>(1, 0, chk, sb, se, c0, log2r, seg, col, cnt, PS, PS32, gbase, gbnd, gslot, gpsum,
                                  wflag[w], &ovf, &rep, lane, w, done);
                    __syncthreads();
                }
            } else {
                for (int mode = 1; mode <= 2; mode++) {
                    clear_lds<NT>(cnt, R / 2);
                    __syncthreads();
                    tile_pass<NT>(0, mode, chk, sb, se, c0, log2r, seg, col, cnt, PS, PS32, gbase, gbnd, gslot, gpsum,
                                  wflag[w], &ovf, &rep, lane, w, done);
                    __syncthreads();
                    tile_pass<NT>(1, mode, chk, sb, se, c0, log2r, seg, col, cnt, PS, PS32, gbase, gbnd, gslot, gpsum,
                                  wflag[w], &ovf, &rep, lane, w, done);
                    __syncthreads();
                }
            }
        }
    }
    if (done[0]) atomicAdd(work, done[0]);
    if (done[1]) atomicAdd(work + 1, done[1]);
}

// fold the per-slot accumulators into P by edge id
__global__ void k_ps_fold(const uint64_t* __restrict__ PS, const uint32_t* __restrict__ PS32,
                          const int32_t* __restrict__ eid, int64_t nnz, uint64_t* __restrict__ P) {
    GRID_STRIDE(s, nnz) {
        const uint64_t v = PS[s] + PS32[s];
        if (v) atomicAdd((unsigned long long*)&P[eid[s]], (unsigned long long)v);
    }
}

// Exclusive prefix sums of n integer values (into int64) without LDS, for the plan's scans
// that run beside the local passes (whose blocks hold the LDS, so an
// LDS-using scan waits for them): one wavefront per chunk of kScanChunk
// values sums it, one wavefront scans the chunk sums, and one wavefront per
// chunk rescans it from its base.
constexpr int kScanChunk = 4096;
__device__ __forceinline__ long long wave_incl_sum64(long long x, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const long long y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    return x;
}
template <class T>
__global__ void k_scan_parts(const T* __restrict__ in, int64_t n, int64_t* __restrict__ part) {
    const int lane = threadIdx.x;  // one wavefront per block, block b = chunk b
    const int64_t c0 = (int64_t)blockIdx.x * kScanChunk, c1 = min(n, c0 + kScanChunk);
    long long x = 0;
    for (int64_t i = c0 + lane; i < c1; i += 64) x += in[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) part[blockIdx.x] = x;
}
__global__ void k_scan_tops(int64_t* __restrict__ part, int64_t np) {  // one wavefront, in place
    const int lane = threadIdx.x;
    long long carry = 0;
    for (int64_t b = 0; b < np; b += 64) {
        const int64_t i = b + lane;
        const long long v = i < np ? part[i] : 0;
        const long long x = wave_incl_sum64(v, lane);
        if (i < np) part[i] = carry + x - v;
        carry += __shfl(x, 63, 64);
    }
}
template <class T>
__global__ void k_scan_apply(const T* __restrict__ in, int64_t n, const int64_t* __restrict__ part,
                             int64_t* __restrict__ out) {
    const int lane = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * kScanChunk, c1 = min(n, c0 + kScanChunk);
    long long carry = part[blockIdx.x];
    for (int64_t b = c0; b < c1; b += 64) {
        const int64_t i = b + lane;
        const long long v = i < c1 ? in[i] : 0;
        const long long x = wave_incl_sum64(v, lane);
        if (i < c1) out[i] = carry + x - v;
        carry += __shfl(x, 63, 64);
    }
}
template <class T>
void excl_sum_nolds(pgd_ctx* ctx, const T* in, int64_t* out, int64_t n) {
    if (n <= 0) return;
    hipStream_t s = ctx->stream;
    const int64_t np = (n + kScanChunk - 1) / kScanChunk;
    PGD_CHECK(np <= INT32_MAX, PGD_ERR_UNSUPPORTED, "scan too large");
    DBuf<int64_t> part;
    part.alloc(ctx, np);
    hipLaunchKernelGGL(k_scan_parts<T>, dim3((unsigned)np), dim3(64), 0, s, in, n, part.get());
    hipLaunchKernelGGL(k_scan_tops, dim3(1), dim3(64), 0, s, part.get(), np);
    hipLaunchKernelGGL(k_scan_apply<T>, dim3((unsigned)np), dim3(64), 0, s, in, n, part.get(), out);
    PGD_HIP(hipGetLastError());
}

template <int NT>
void launch_tiled(pgd_ctx* ctx, OrientState& st, unsigned long long* work) {
    WedgePlan& wp = st.wp;
    const size_t dyn = (size_t)2 << wp.log2r;
    hipLaunchKernelGGL(k_wedge_tiled<NT>, dim3(wp.grid), dim3(NT), dyn, ctx->stream, st.rrow.get(), st.rcol.get(),
                       st.b_heavy.get(), wp.hbase.get(), wp.htoff.get(), wp.cseg.get(), wp.seg.get(),
                       wp.items.get(), wp.nitems, wp.log2r, wp.qctr.get(), wp.ps.get(), wp.ps32.get(), work);
}

template <int NT>
int tiled_blocks_per_cu(pgd_ctx* ctx, int log2r) {
    int& slot = ctx->tiled_per_cu[NT == 256 ? 0 : NT == 512 ? 1 : 2][log2r - 8];
    if (slot) return slot;  // the attribute and the occupancy are per process and device
    const size_t dyn = (size_t)2 << log2r;
    PGD_HIP(hipFuncSetAttribute((const void*)k_wedge_tiled<NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)dyn));
    int nb = 0;
    PGD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_wedge_tiled<NT>, NT, dyn));
    slot = nb;
    return nb;
}

}  // namespace

// plan of the heavy 4-cycle pass: suffix starts and tile ranges of the
// anchors' pairs, tile table, per-cell sizes, items sorted largest first,
// per-block buffers (host synchronisations)
void wedge_plan_build(pgd_ctx* ctx, OrientState& st) {
    hipStream_t s = ctx->stream;
    WedgePlan& wp = st.wp;
    const OrientParams& prm = st.prm;
    const int64_t n = st.n, nnz = 2 * st.m, nh = st.n_heavy;
    PGD_CHECK(prm.wedge_log2r >= 8 && prm.wedge_log2r <= 16, PGD_ERR_INVALID, "wedge tile must be 2^8..2^16");
    PGD_CHECK(prm.wedge_nt == 256 || prm.wedge_nt == 512 || prm.wedge_nt == 1024, PGD_ERR_INVALID,
              "wedge block size must be 256, 512 or 1024");
    PGD_CHECK(nnz < INT32_MAX, PGD_ERR_UNSUPPORTED, "2m >= 2^31 slots: the wedge plan indexes slots as int32");
    wp.log2r = prm.wedge_log2r;
    // suffix starts of the anchors' pairs and the anchors' tile ranges
    ensure(wp.sfx, ctx, nnz);
    ensure(wp.hbase, ctx, nh);
    DBuf<int64_t> nt;
    nt.alloc(ctx, nh + 1);
    {
        // st.v0 (2m int32, free after the relabel) holds the out-slot of every edge
        int32_t* outslot = st.v0.get();
        hipLaunchKernelGGL(k_out_slots, dim3(grid_cap(ctx, (n * 64 + 255) / 256, 16)), dim3(256), 0, s,
                           st.rrow.get(), st.ostart.get(), st.reid.get(), n, outslot);
        hipLaunchKernelGGL(k_pair_wedges, dim3(grid_cap(ctx, (nh * 64 + 255) / 256, 16)), dim3(256), 0, s,
                           st.rrow.get(), st.rcol.get(), st.ostart.get(), st.reid.get(), outslot, st.b_heavy.get(),
                           nh, wp.log2r, wp.sfx.get(), wp.hbase.get(), nt.get());
        PGD_HIP(hipGetLastError());
    }
    PGD_HIP(hipGetLastError());
    // tile table (st.v0 is free scratch of 2m entries after the relabel; the
    // prefix gets its own buffer, since the plan may run beside the clique
    // pass, which uses st.k0)
    {
        int32_t* flag = st.v0.get();
        DBuf<int64_t> posb;
        posb.alloc(ctx, nnz);
        int64_t* pos = posb.get();
        const int gw = grid_cap(ctx, (n * 64 + 255) / 256, 16);
        hipLaunchKernelGGL(k_tt_flags, dim3(gw), dim3(256), 0, s, st.rrow.get(), st.rcol.get(), n, wp.log2r, flag);
        PGD_HIP(hipGetLastError());
        excl_sum_nolds(ctx, flag, pos, nnz);
        // at most one entry per slot: sized without reading the total back
        ensure(wp.tt, ctx, nnz);
        ensure(wp.ttrow, ctx, n + 1);
        hipLaunchKernelGGL(k_tt_fill, dim3(gw), dim3(256), 0, s, st.rrow.get(), st.rcol.get(), n, nnz, wp.log2r,
                           flag, pos, wp.tt.get(), wp.ttrow.get());
        PGD_HIP(hipGetLastError());
    }
    // (anchor, tile) cells
    ensure(wp.htoff, ctx, nh + 1);
    int64_t ncell = 0;
    {
        hipLaunchKernelGGL(w_fill_u64, dim3(1), dim3(64), 0, s, (uint64_t*)nt.get() + nh, (int64_t)1);
        PGD_HIP(hipGetLastError());
        excl_sum_nolds(ctx, nt.get(), wp.htoff.get(), nh + 1);
        PGD_HIP(hipMemcpyAsync(&ncell, wp.htoff.get() + nh, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        PGD_HIP(hipStreamSynchronize(s));
    }
    // (one count past the last cell stays 0: the scan of the counts then ends
    // with their total)
    ensure(wp.htseg, ctx, ncell + 1);
    ensure(wp.htw, ctx, ncell);
    {
        const int g = grid_cap(ctx, (ncell + 255) / 256, 16);
        hipLaunchKernelGGL(w_fill_u32, dim3(g), dim3(256), 0, s, wp.htseg.get(), ncell + 1);
        hipLaunchKernelGGL(w_fill_u64, dim3(g), dim3(256), 0, s, wp.htw.get(), ncell);
        ensure(wp.pacc, ctx, 3);
        hipLaunchKernelGGL(w_fill_u64, dim3(1), dim3(64), 0, s, (uint64_t*)wp.pacc.get(), (int64_t)3);
        ensure(wp.gcur, ctx, std::max<int64_t>(ncell, 1));  // cursors of the segment walk
        hipLaunchKernelGGL(w_fill_u32, dim3(g), dim3(256), 0, s, wp.gcur.get(), ncell);
        hipLaunchKernelGGL(k_plan_cells<256>, dim3(grid_cap(ctx, nh, 8)), dim3(256), 0, s, st.rrow.get(),
                           st.rcol.get(), st.b_heavy.get(), nh, wp.sfx.get(), wp.tt.get(),
                           wp.ttrow.get(), wp.hbase.get(), wp.htoff.get(), wp.htseg.get(), wp.htw.get(),
                           wp.pacc.get(), pinned_words(ctx), st.v0.get());
        PGD_HIP(hipGetLastError());
    }
    // the cells' segment offsets: a scan of the counts, queued behind the
    // count walk (so it runs on the CU slots that walk leaves)
    ensure(wp.cseg, ctx, ncell + 1);
    excl_sum_nolds(ctx, wp.htseg.get(), wp.cseg.get(), ncell + 1);
    PGD_HIP(hipStreamSynchronize(s));
    const uint64_t segtot = ((volatile unsigned long long*)ctx->pin)[0];
    wp.wedges = (int64_t)((volatile unsigned long long*)ctx->pin)[1];
    // smallest degree d with d * (max degree) >= 2^32
    wp.wide_deg = st.max_deg > 0
                      ? (int32_t)std::min<int64_t>(INT32_MAX, (((int64_t)1 << 32) + st.max_deg - 1) / st.max_deg)
                      : INT32_MAX;
    // every cell's segments, in cell order (the walk again)
    ensure(wp.seg, ctx, std::max<uint64_t>(segtot, 1));
    {
        hipLaunchKernelGGL(k_plan_fill<256>, dim3(grid_cap(ctx, nh, 8)), dim3(256), 0, s, st.rrow.get(),
                           st.rcol.get(), st.b_heavy.get(), nh, wp.sfx.get(), wp.tt.get(), wp.ttrow.get(),
                           st.v0.get(), wp.hbase.get(), wp.htoff.get(), wp.cseg.get(), wp.gcur.get(), wp.wide_deg,
                           wp.seg.get());
        PGD_HIP(hipGetLastError());
    }
    // a cell holds at most one segment per neighbour of its anchor: the
    // maximum degree bounds it (read from the cells when that bound is large)
    const uint32_t maxcell = st.max_deg <= (1 << 20) ? (uint32_t)st.max_deg
                                                      : device_max(ctx, wp.htseg.get(), (size_t)ncell);
    // grid from the kernel's occupancy at this tile size
    int per_cu = 0;
    if (prm.wedge_nt == 256) per_cu = tiled_blocks_per_cu<256>(ctx, wp.log2r);
    else if (prm.wedge_nt == 512) per_cu = tiled_blocks_per_cu<512>(ctx, wp.log2r);
    else per_cu = tiled_blocks_per_cu<1024>(ctx, wp.log2r);
    PGD_CHECK(per_cu > 0, PGD_ERR_INVALID, "wedge tile and block size exceed the LDS of a CU");
    wp.grid = ctx->num_cus * per_cu;
    // segments per item: the largest cell, or twice an even share of all
    // segments, at most 2^18 (items are cut to fit it)
    wp.segcap = (int64_t)std::max<uint64_t>(
        maxcell, std::min<uint64_t>((uint64_t)1 << 18, 2 * (segtot + (uint64_t)wp.grid - 1) / (uint64_t)wp.grid));
    wp.segcap = std::max<int64_t>(wp.segcap, 1);
    const uint64_t wcap =
        std::max<uint64_t>((uint64_t)wp.wedges / ((uint64_t)wp.grid * std::max(1, prm.wedge_split)), 1 << 16);
    // items: count, offsets, write, sort by size
    {
        DBuf<int64_t> per, ioff;
        per.alloc(ctx, nh + 1);
        ioff.alloc(ctx, nh + 1);
        const int g = grid_cap(ctx, (nh + 255) / 256, 16);
        hipLaunchKernelGGL(k_plan_items, dim3(g), dim3(256), 0, s, wp.htoff.get(), wp.hbase.get(), wp.htseg.get(),
                           wp.htw.get(), nh, wp.segcap, wcap, per.get(), (const int64_t*)nullptr, (WedgeItem*)nullptr,
                           (uint64_t*)nullptr, (uint32_t*)nullptr);
        PGD_HIP(hipGetLastError());
        excl_sum_nolds(ctx, per.get(), ioff.get(), nh + 1);
        PGD_HIP(hipMemcpyAsync(&wp.nitems, ioff.get() + nh, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        PGD_HIP(hipStreamSynchronize(s));
        const int64_t ni = wp.nitems;
        DBuf<WedgeItem> raw;
        DBuf<uint64_t> key;
        DBuf<uint32_t> bk;
        raw.alloc(ctx, ni);
        key.alloc(ctx, ni);
        bk.alloc(ctx, 128);
        hipLaunchKernelGGL(k_plan_items, dim3(g), dim3(256), 0, s, wp.htoff.get(), wp.hbase.get(), wp.htseg.get(),
                           wp.htw.get(), nh, wp.segcap, wcap, per.get(), ioff.get(), raw.get(), key.get(), bk.get());
        PGD_HIP(hipGetLastError());
        ensure(wp.items, ctx, ni);
        if (ni) {
            // largest first, by powers of two of the wedges: a counting sort
            // of 64 buckets with global atomics (no LDS: these kernels find
            // CU slots beside the local passes)
            const int gi = grid_cap(ctx, (ni + 255) / 256, 16);
            hipLaunchKernelGGL(k_item_buckets, dim3(gi), dim3(256), 0, s, key.get(), ni, bk.get());
            hipLaunchKernelGGL(k_bucket_starts, dim3(1), dim3(64), 0, s, bk.get());
            hipLaunchKernelGGL(k_item_scatter, dim3(gi), dim3(256), 0, s, key.get(), raw.get(), ni, bk.get() + 64,
                               wp.items.get());
            PGD_HIP(hipGetLastError());
        }
    }
    ensure(wp.ps, ctx, nnz);
    ensure(wp.ps32, ctx, nnz);
    ensure(wp.qctr, ctx, 1);
}

// the heavy 4-cycle pass (capturable: fills, the persistent kernel, the fold)
void wedge_plan_count(pgd_ctx* ctx, OrientState& st, unsigned long long* work) {
    hipStream_t s = ctx->stream;
    WedgePlan& wp = st.wp;
    const int64_t nnz = 2 * st.m;
    const int g = grid_cap(ctx, (nnz + 255) / 256, 16);
    hipLaunchKernelGGL(w_fill_u64, dim3(g), dim3(256), 0, s, wp.ps.get(), nnz);
    hipLaunchKernelGGL(w_fill_u32, dim3(g), dim3(256), 0, s, wp.ps32.get(), nnz);
    hipLaunchKernelGGL(w_fill_u32, dim3(1), dim3(64), 0, s, wp.qctr.get(), (int64_t)1);
    PGD_HIP(hipGetLastError());
    if (wp.nitems) {
        if (st.prm.wedge_nt == 256) launch_tiled<256>(ctx, st, work);
        else if (st.prm.wedge_nt == 512) launch_tiled<512>(ctx, st, work);
        else launch_tiled<1024>(ctx, st, work);
        PGD_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_ps_fold, dim3(g), dim3(256), 0, s, wp.ps.get(), wp.ps32.get(), st.reid.get(), nnz,
                       st.P.get());
    PGD_HIP(hipGetLastError());
}

}  // namespace pgd
