/*
 * pgd_oracle.c -- CPU oracle (test infrastructure only; see pgd_oracle.h).
 *
 * A clean-room C restatement of the reference's counting path.  It mirrors the
 * reference's observable behaviour (reader semantics, neighbour/edge order,
 * per-edge perfect-hash loop, uint64 reductions and the floating-point
 * derivation as built) so it can serve as the parity checker and the timed
 * host-core baseline.  Nothing in the product links against it.
 */
#define _GNU_SOURCE
#include "pgd_oracle.h"

#include <ctype.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------------ */
/* small growable vectors                                                    */
/* ------------------------------------------------------------------------ */
typedef struct { int32_t *a; int64_t n, cap; } ivec;
static void iv_push(ivec *v, int32_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 4;
        v->a = (int32_t *)realloc(v->a, (size_t)v->cap * sizeof(int32_t));
    }
    v->a[v->n++] = x;
}

/* ------------------------------------------------------------------------ */
/* line reader with std::getline semantics (no '\r' stripping; a trailing    */
/* segment without '\n' is a line; nothing after the last '\n' is not)       */
/* ------------------------------------------------------------------------ */
typedef struct { char *buf; size_t len, pos; } lines_t;
static int load_file(const char *path, lines_t *L) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    L->buf = (char *)malloc((size_t)sz + 1);
    L->len = fread(L->buf, 1, (size_t)sz, f);
    L->buf[L->len] = 0;
    L->pos = 0;
    fclose(f);
    return 0;
}
/* returns 1 and sets [*s,*e) for the next line, 0 at end of input */
static int next_line(lines_t *L, const char **s, size_t *n) {
    if (L->pos >= L->len) return 0;
    const char *b = L->buf + L->pos;
    const char *nl = memchr(b, '\n', L->len - L->pos);
    size_t k = nl ? (size_t)(nl - b) : L->len - L->pos;
    *s = b; *n = k;
    L->pos += k + (nl ? 1 : 0);
    return 1;
}

/* std::istream >> int, as libstdc++ does it: skip whitespace; at end of the
 * stream the target is left unchanged; a field with no digits stores 0; an
 * out-of-range value stores INT_MAX/INT_MIN; any of these sets `fail`, after
 * which further extractions leave their targets unchanged. */
typedef struct { const char *p, *e; int fail; } istr;
static void is_int(istr *s, int *out) {
    if (s->fail) return;
    while (s->p < s->e && isspace((unsigned char)*s->p)) s->p++;
    if (s->p >= s->e) { s->fail = 1; return; }
    int neg = 0;
    if (*s->p == '+' || *s->p == '-') { neg = (*s->p == '-'); s->p++; }
    long long acc = 0; int nd = 0, ovf = 0;
    while (s->p < s->e && *s->p >= '0' && *s->p <= '9') {
        acc = acc * 10 + (*s->p - '0');
        if (acc > (long long)INT_MAX + 1) { ovf = 1; acc = (long long)INT_MAX + 1; }
        nd++; s->p++;
    }
    if (nd == 0) { *out = 0; s->fail = 1; return; }
    if (neg) acc = -acc;
    if (ovf || acc > INT_MAX || acc < INT_MIN) {
        *out = neg ? INT_MIN : INT_MAX; s->fail = 1; return;
    }
    *out = (int)acc;
}

/* atoi of [b, b+n): glibc atoi is (int)strtol(s, 0, 10) */
static int atoi_n(const char *b, size_t n) {
    char tmp[64];
    if (n >= sizeof(tmp)) n = sizeof(tmp) - 1;
    memcpy(tmp, b, n); tmp[n] = 0;
    return (int)strtol(tmp, NULL, 10);
}
/* std::string::find(delim, prev) */
static long str_find(const char *s, size_t n, const char *d, size_t dn, size_t from) {
    if (dn == 0) return from <= n ? (long)from : -1;
    for (size_t i = from; i + dn <= n; i++)
        if (memcmp(s + i, d, dn) == 0) return (long)i;
    return -1;
}
/* get_token(int&,...) graphlet_core.cpp:229-238 */
static void get_token(int *v, const char *line, size_t n, const char *delim, size_t dn, size_t *prev) {
    long pos = str_find(line, n, delim, dn, *prev);
    if (pos >= 0) {
        if ((size_t)pos > *prev) *v = atoi_n(line + *prev, (size_t)pos - *prev);
        *prev = (size_t)pos + 1;
    } else if (*prev < n) {
        *v = atoi_n(line + *prev, n - *prev);
    }
}
static int is_comment(const char *s, size_t n) { return n > 0 && (s[0] == '#' || s[0] == '%'); }

static const char *file_ext(const char *path) {   /* graphlet_core.cpp:900-907 */
    const char *d = strrchr(path, '.');
    return d ? d + 1 : "";
}

static void build_csr(ivec *lists, int64_t n, or_csr *out) {
    out->n = n;
    out->row = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
    out->row[0] = 0;
    for (int64_t i = 0; i < n; i++) out->row[i + 1] = out->row[i] + lists[i].n;
    out->nnz = out->row[n];
    out->col = (int32_t *)malloc((size_t)(out->nnz ? out->nnz : 1) * sizeof(int32_t));
    for (int64_t i = 0; i < n; i++) {
        if (lists[i].n) memcpy(out->col + out->row[i], lists[i].a, (size_t)lists[i].n * sizeof(int32_t));
        free(lists[i].a);
    }
}

/* read_mtx, graphlet_core.cpp:459-522.  The header's n is ignored: n is one
 * plus the largest endpoint id; no duplicate removal. */
static int read_mtx(const char *path, or_csr *out, int64_t *self_loops) {
    lines_t L;
    if (load_file(path, &L)) return -1;
    const char *s = NULL; size_t n = 0; int have = 0;
    while ((have = next_line(&L, &s, &n)) && is_comment(s, n)) {}
    if (!have) { s = ""; n = 0; }
    int v = -1, u = -1, fix = 1;
    if (n > 0) {
        istr st = { s, s + n, 0 };
        is_int(&st, &v); is_int(&st, &u);
        if (v == 0 || u == 0) fix = 0;
    }
    ivec *lists = NULL; int64_t cap = 0, maxid = -1; int64_t self = 0; int bad = 0;
    while (fix && next_line(&L, &s, &n)) {
        if (n == 0) continue;
        if (is_comment(s, n)) continue;
        istr st = { s, s + n, 0 };
        is_int(&st, &v); is_int(&st, &u);
        v--; u--;
        if (v == u) { self++; continue; }
        if (v < 0 || u < 0) { bad = 1; break; }   /* the reference loops forever */
        int64_t need = (v > u ? v : u) + 1;
        if (need > cap) {
            int64_t nc = cap ? cap : 1024;
            while (nc < need) nc *= 2;
            lists = (ivec *)realloc(lists, (size_t)nc * sizeof(ivec));
            memset(lists + cap, 0, (size_t)(nc - cap) * sizeof(ivec));
            cap = nc;
        }
        iv_push(&lists[v], u);
        iv_push(&lists[u], v);
        if (need - 1 > maxid) maxid = need - 1;
    }
    free(L.buf);
    if (bad) { for (int64_t i = 0; i < cap; i++) free(lists[i].a); free(lists); return -2; }
    build_csr(lists, maxid + 1, out);
    for (int64_t i = maxid + 1; i < cap; i++) free(lists[i].a);
    free(lists);
    if (self_loops) *self_loops = self;
    return 0;
}

/* remove_multiple_edges, graphlet_core.cpp:530-556: keep the first
 * occurrence of each neighbour in every adjacency list. */
static void remove_multiple_edges(or_csr *g) {
    uint8_t *seen = (uint8_t *)calloc((size_t)g->n + 1, 1);
    int64_t w = 0;
    int64_t *nr = (int64_t *)malloc((size_t)(g->n + 1) * sizeof(int64_t));
    nr[0] = 0;
    for (int64_t i = 0; i < g->n; i++) {
        int64_t b = g->row[i], e = g->row[i + 1];
        for (int64_t j = b; j < e; j++) {
            int32_t x = g->col[j];
            if (!seen[x]) { seen[x] = 1; g->col[w++] = x; }
        }
        /* the kept entries are exactly the marked ids */
        for (int64_t j = nr[i]; j < w; j++) seen[g->col[j]] = 0;
        nr[i + 1] = w;
    }
    free(seen);
    free(g->row);
    g->row = nr;
    g->nnz = w;
}

/* read_edge_list, graphlet_core.cpp:291-457 */
static int read_edge_list(const char *path, or_csr *out, int64_t *self_loops) {
    lines_t L;
    if (load_file(path, &L)) return -1;
    const char *s = NULL; size_t n = 0; int have;
    char delim[4] = ""; size_t dn = 0;
    while ((have = next_line(&L, &s, &n)) && is_comment(s, n)) {
        if (memmem(s, n, "MatrixMarket matrix coordinate pattern symmetric", 48)) { strcpy(delim, " "); dn = 1; }
    }
    if (!have) { s = ""; n = 0; }
    /* detect_delim, graphlet_core.cpp:196-226 */
    if (dn == 0) {
        if (memchr(s, ',', n)) { strcpy(delim, ","); dn = 1; }
        else if (memchr(s, '\t', n)) { strcpy(delim, "\t"); dn = 1; }
        else if (memmem(s, n, "   ", 3)) { strcpy(delim, "   "); dn = 3; }
        else if (memchr(s, ' ', n)) { strcpy(delim, " "); dn = 1; }
    }
    if (dn == 0) {
        const char *ext = file_ext(path);
        if (!strcmp(ext, "csv")) strcpy(delim, ",");
        else if (!strcmp(ext, "tab")) strcpy(delim, "\t");
        else strcpy(delim, " ");
        dn = 1;
    }
    int v = -1, u = -1, fix = 1;
    if (n > 0) {
        istr st = { s, s + n, 0 };
        is_int(&st, &v); is_int(&st, &u);
        if (v == 0 || u == 0) fix = 0;
    }
    int maxv = 0;
    while (next_line(&L, &s, &n)) {           /* first pass, from the line after */
        if (n == 0 || is_comment(s, n)) continue;
        size_t prev = 0;
        get_token(&v, s, n, delim, dn, &prev);
        get_token(&u, s, n, delim, dn, &prev);
        if (v == 0 || u == 0) fix = 0;
        if (v > maxv) maxv = v;
        if (u > maxv) maxv = u;
    }
    /* second pass over every line */
    L.pos = 0;
    int32_t *lookup = (int32_t *)malloc(((size_t)maxv + 1) * sizeof(int32_t));
    for (int64_t i = 0; i <= maxv; i++) lookup[i] = -1;
    ivec *lists = NULL; int64_t cap = 0; int32_t vid = 0; int64_t self = 0; int bad = 0;
    while (next_line(&L, &s, &n)) {
        if (n == 0 || is_comment(s, n)) continue;
        size_t prev = 0;
        get_token(&v, s, n, delim, dn, &prev);
        get_token(&u, s, n, delim, dn, &prev);
        if (fix) { v--; u--; }
        if (v == u) { self++; continue; }
        if (v < 0 || u < 0 || v > maxv || u > maxv) { bad = 1; break; }   /* reference: out-of-bounds */
        /* v and u take their new ids in place (graphlet_core.cpp:417-424), so a
         * field the next line leaves unset carries the new id */
        if (lookup[v] == -1) lookup[v] = vid++;
        v = lookup[v];
        if (lookup[u] == -1) lookup[u] = vid++;
        u = lookup[u];
        int32_t a = v, b = u;
        if (vid > cap) {
            int64_t nc = cap ? cap * 2 : 1024;
            lists = (ivec *)realloc(lists, (size_t)nc * sizeof(ivec));
            memset(lists + cap, 0, (size_t)(nc - cap) * sizeof(ivec));
            cap = nc;
        }
        iv_push(&lists[a], b);
        iv_push(&lists[b], a);
    }
    free(lookup);
    free(L.buf);
    if (bad) { for (int64_t i = 0; i < cap; i++) free(lists[i].a); free(lists); return -2; }
    build_csr(lists, vid, out);
    for (int64_t i = vid; i < cap; i++) free(lists[i].a);
    free(lists);
    remove_multiple_edges(out);
    if (self_loops) *self_loops = self;
    return 0;
}

int or_read_graph(const char *path, or_csr *out, int64_t *self_loops) {
    memset(out, 0, sizeof(*out));
    const char *ext = file_ext(path);     /* graphlet_core.cpp:259-283 */
    if (!strcmp(ext, "mtx")) return read_mtx(path, out, self_loops);
    return read_edge_list(path, out, self_loops);
}

void or_free(or_csr *g) { free(g->row); free(g->col); memset(g, 0, sizeof(*g)); }

int or_csr_from_arrays(int64_t n, int64_t nnz, const int64_t *row, const int32_t *col, or_csr *out) {
    out->n = n; out->nnz = nnz;
    out->row = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
    out->col = (int32_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
    memcpy(out->row, row, (size_t)(n + 1) * sizeof(int64_t));
    if (nnz) memcpy(out->col, col, (size_t)nnz * sizeof(int32_t));
    return 0;
}

/* ------------------------------------------------------------------------ */
void or_degrees(const or_csr *g, int32_t *deg, int32_t *max_deg, int32_t *avg_deg) {
    int32_t md = g->n > 0 ? (int32_t)(g->row[1] - g->row[0]) : 0;
    for (int64_t v = 0; v < g->n; v++) {
        deg[v] = (int32_t)(g->row[v + 1] - g->row[v]);
        if (deg[v] > md) md = deg[v];
    }
    if (max_deg) *max_deg = md;
    if (avg_deg) *avg_deg = g->n > 0 ? (int32_t)((double)g->nnz / (double)g->n) : 0;
}

/* Core numbers by repeated minimum-degree peeling with buckets (the
 * Batagelj-Zaversnik order; core numbers are unique, so any correct peel
 * matches graphlet_core.cpp:958-1009).  The printed K skips the last vertex
 * because of the 1-based arrays (graphlet_core.cpp:1011-1015). */
int32_t or_kcore(const or_csr *g, int32_t *core) {
    int64_t n = g->n;
    if (n == 0) return 0;
    int32_t md = 0;
    for (int64_t v = 0; v < n; v++) {
        core[v] = (int32_t)(g->row[v + 1] - g->row[v]);
        if (core[v] > md) md = core[v];
    }
    int64_t *bin = (int64_t *)calloc((size_t)md + 2, sizeof(int64_t));
    int64_t *pos = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *vert = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int64_t v = 0; v < n; v++) bin[core[v]]++;
    int64_t start = 0;
    for (int32_t d = 0; d <= md; d++) { int64_t c = bin[d]; bin[d] = start; start += c; }
    for (int64_t v = 0; v < n; v++) { pos[v] = bin[core[v]]; vert[pos[v]] = v; bin[core[v]]++; }
    for (int32_t d = md; d >= 1; d--) bin[d] = bin[d - 1];
    bin[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t v = vert[i];
        for (int64_t j = g->row[v]; j < g->row[v + 1]; j++) {
            int64_t u = g->col[j];
            if (core[u] > core[v]) {
                int32_t du = core[u];
                int64_t pu = pos[u], pw = bin[du], w = vert[pw];
                if (u != w) { pos[u] = pw; vert[pu] = w; pos[w] = pu; vert[pw] = u; }
                bin[du]++;
                core[u]--;
            }
        }
    }
    int32_t K = 0;
    for (int64_t v = 0; v + 1 < n; v++) if (core[v] > K) K = core[v];
    free(bin); free(pos); free(vert);
    return K;
}

/* per-vertex stable counting sort by key, graphlet_core.cpp:1846-1877 */
static void bucket_sort_neighbors(or_csr *g, const int64_t *key, int small_to_large) {
    int32_t *tmp = (int32_t *)malloc((size_t)(g->nnz ? g->nnz : 1) * sizeof(int32_t));
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t v = 0; v < g->n; v++) {
        int64_t b = g->row[v], d = g->row[v + 1] - b;
        if (d == 0) continue;
        /* stable ascending order by key (ties keep input order) */
        int64_t *idx = (int64_t *)malloc((size_t)d * sizeof(int64_t));
        for (int64_t i = 0; i < d; i++) idx[i] = i;
        /* insertion-merge sort: stable mergesort */
        int64_t *aux = (int64_t *)malloc((size_t)d * sizeof(int64_t));
        for (int64_t w = 1; w < d; w *= 2) {
            for (int64_t lo = 0; lo < d; lo += 2 * w) {
                int64_t mid = lo + w < d ? lo + w : d, hi = lo + 2 * w < d ? lo + 2 * w : d;
                int64_t i = lo, j = mid, k = lo;
                while (i < mid && j < hi) {
                    if (key[g->col[b + idx[j]]] < key[g->col[b + idx[i]]]) aux[k++] = idx[j++];
                    else aux[k++] = idx[i++];
                }
                while (i < mid) aux[k++] = idx[i++];
                while (j < hi) aux[k++] = idx[j++];
            }
            memcpy(idx, aux, (size_t)d * sizeof(int64_t));
        }
        if (small_to_large) for (int64_t i = 0; i < d; i++) tmp[b + i] = g->col[b + idx[i]];
        else for (int64_t i = 0; i < d; i++) tmp[b + i] = g->col[b + idx[d - 1 - i]];
        free(idx); free(aux);
    }
    memcpy(g->col, tmp, (size_t)g->nnz * sizeof(int32_t));
    free(tmp);
}

void or_order_neighbors(or_csr *g, const char *m, int small_to_large) {
    int64_t n = g->n;
    int64_t *key = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t *deg = (int32_t *)malloc((size_t)(n ? n : 1) * sizeof(int32_t));
    int32_t *core = (int32_t *)malloc((size_t)(n ? n : 1) * sizeof(int32_t));
    or_degrees(g, deg, NULL, NULL);
    or_kcore(g, core);
    /* the reference's kcore[] is 1-based: kcore[x] is the core of x-1 */
    #define KC(x) ((x) >= 1 ? (int64_t)core[(x) - 1] : 0)
    int ok = 1;
    if (!strcmp(m, "deg") || !strcmp(m, "degree")) {
        for (int64_t x = 0; x < n; x++) key[x] = deg[x];
    } else if (!strcmp(m, "kcore")) {
        for (int64_t x = 0; x < n; x++) key[x] = KC(x);
    } else if (!strcmp(m, "degree_kcore") || !strcmp(m, "deg_kcore")) {
        for (int64_t x = 0; x < n; x++) key[x] = (int32_t)((uint32_t)deg[x] + (uint32_t)KC(x));
    } else if (!strcmp(m, "degree_vol") || !strcmp(m, "deg_vol")) {
        for (int64_t x = 0; x < n; x++) { uint32_t s = 0; for (int64_t j = g->row[x]; j < g->row[x + 1]; j++) s += (uint32_t)deg[g->col[j]]; key[x] = (int32_t)s; }
    } else if (!strcmp(m, "kcore_vol")) {
        for (int64_t x = 0; x < n; x++) { uint32_t s = 0; for (int64_t j = g->row[x]; j < g->row[x + 1]; j++) s += (uint32_t)KC(g->col[j]); key[x] = (int32_t)s; }
    } else if (!strcmp(m, "degree_kcore_vol") || !strcmp(m, "deg_kcore_vol")) {
        for (int64_t x = 0; x < n; x++) { uint32_t s = 0; for (int64_t j = g->row[x]; j < g->row[x + 1]; j++) s += (uint32_t)deg[g->col[j]] + (uint32_t)KC(g->col[j]); key[x] = (int32_t)s; }
    } else ok = 0;
    #undef KC
    if (ok) bucket_sort_neighbors(g, key, small_to_large);
    free(key); free(deg); free(core);
}

int64_t or_edge_arrays(const or_csr *g, int32_t *e_v, int32_t *e_u) {
    int64_t m = 0;
    for (int64_t v = 0; v < g->n; v++) {
        for (int64_t j = g->row[v]; j < g->row[v + 1]; j++) {
            int64_t u = g->col[j];
            if (v < u) {
                int64_t dv = g->row[v + 1] - g->row[v], du = g->row[u + 1] - g->row[u];
                if (dv < du) { e_v[m] = (int32_t)v; e_u[m] = (int32_t)u; }
                else { e_v[m] = (int32_t)u; e_u[m] = (int32_t)v; }
                m++;
            }
        }
    }
    return m;
}

typedef struct { int32_t id, b; } vtx;
static int cmp_decr(const void *x, const void *y) {   /* decr_bound, graphlet_vertex.h:61-64 */
    const vtx *a = (const vtx *)x, *c = (const vtx *)y;
    if (a->b != c->b) return a->b > c->b ? -1 : 1;
    return a->id > c->id ? -1 : (a->id < c->id);
}
static int cmp_incr(const void *x, const void *y) {   /* incr_bound, graphlet_vertex.h:73-76 */
    const vtx *a = (const vtx *)x, *c = (const vtx *)y;
    if (a->b != c->b) return a->b < c->b ? -1 : 1;
    return a->id > c->id ? -1 : (a->id < c->id);
}

/* sort_edges over all m edges (graphlet_core.cpp:854-867) when ids is NULL,
 * else over the sampled set E_s = ids[0..m) (836-851), whose natural key is
 * the position in E_s */
static int sort_edge_list(const or_csr *g, const int32_t *e_v, const int32_t *e_u, const int32_t *ids, int64_t m,
                          const char *o, int small_to_large, int32_t *e_ordered);

int or_sort_edges(const or_csr *g, const int32_t *e_v, const int32_t *e_u, int64_t m,
                  const char *o, int small_to_large, int32_t *e_ordered) {
    return sort_edge_list(g, e_v, e_u, NULL, m, o, small_to_large, e_ordered);
}

int or_sort_edges_subset(const or_csr *g, const int32_t *e_v, const int32_t *e_u, const int32_t *ids,
                         int64_t count, const char *o, int small_to_large, int32_t *e_ordered) {
    return sort_edge_list(g, e_v, e_u, ids, count, o, small_to_large, e_ordered);
}

static int sort_edge_list(const or_csr *g, const int32_t *e_v, const int32_t *e_u, const int32_t *ids, int64_t m,
                          const char *o, int small_to_large, int32_t *e_ordered) {
    int64_t n = g->n;
    int32_t *deg = (int32_t *)malloc((size_t)(n ? n : 1) * sizeof(int32_t));
    int32_t *core = (int32_t *)malloc((size_t)(n ? n : 1) * sizeof(int32_t));
    or_degrees(g, deg, NULL, NULL);
    or_kcore(g, core);
    #define KC(x) ((x) >= 1 ? (int64_t)core[(x) - 1] : 0)
    /* get_ordering_enum, graphlet_core.cpp:713-763 */
    int ord = 0;
    if (!strcmp(o, "natural")) ord = 0;
    else if (!strcmp(o, "rand") || !strcmp(o, "random")) ord = 1;
    else if (!strcmp(o, "deg") || !strcmp(o, "degree")) ord = 2;
    else if (!strcmp(o, "kcore") || !strcmp(o, "kcores")) ord = 3;
    else if (!strcmp(o, "kcore_deg") || !strcmp(o, "kcore_degree")) ord = 4;
    else if (!strcmp(o, "deg_vol") || !strcmp(o, "degree_vol")) ord = 5;
    else if (!strcmp(o, "kcore_vol") || !strcmp(o, "kcore_volume")) ord = 6;
    else if (!strcmp(o, "kcore_deg_vol") || !strcmp(o, "kcore_degree_vol")) ord = 7;
    else if (!strcmp(o, "var")) ord = 8;
    else {
        static const char *rest[] = { "triangles", "tri", "triangle", "wedges", "tcores", "triangle_cores",
            "triangle_core", "tcore", "coloring", "greedy_coloring", "tri_vol", "triangle_vol", "tcore_vol",
            "tcore_volume", "triangle_core_vol", "triangles_only", "triangle_only", "dynamic_largest_first",
            "lfo", "LFO", "triangle_core_max", "tcore_max", "tcores_max", "degree_triangles", "degree_tri",
            "deg_tri", "deg_triangles", "kcore_tri", "kcore_triangles", "kcore_triangle", "kcore_deg_tri",
            "kcore_degree_triangle", "kcore_degree_tri", "kcore_tri_vol", "kcore_triangle_vol",
            "kcore_deg_tri_vol", "kcore_degree_tri_vol", "dist_two_lfo", "dist_two_largest_first",
            "dist_two_dlfo", "dist_two_dynamic_lfo", "dist_two_slo", "dist_two_smallest_last", "ido",
            "incidence", "incidence_deg", "dist_two_ido", "dist_two_incidence", "dist_two_tcore",
            "dist_two_triangle_core", "tcore_min", "triangle_core_min", "tri_core_min", "common",
            "common_neighbors", NULL };
        for (int i = 0; rest[i]; i++) if (!strcmp(o, rest[i])) { ord = 9; break; }
    }
    if (ord == 1 || ord == 8) { free(deg); free(core); return -1; }
    vtx *E = (vtx *)malloc((size_t)(m ? m : 1) * sizeof(vtx));
    for (int64_t p = 0; p < m; p++) {
        const int32_t e = ids ? ids[p] : (int32_t)p;   /* `int edge_id = E_s[e]` (843) */
        int64_t v = e_v[e], u = e_u[e];
        long long val = p;                  /* get_ordering_value, graphlet_core.cpp:773-834 */
        switch (ord) {
        case 0: break;
        case 2: val = (long long)deg[u] + deg[v]; break;
        case 3: val = KC(u) + KC(v); break;
        case 4: val = (int32_t)((uint32_t)(deg[u] + deg[v]) * (uint32_t)(KC(u) + KC(v))); break;
        case 5: val = 0;
            for (int64_t j = g->row[u]; j < g->row[u + 1]; j++) val += deg[g->col[j]];
            for (int64_t j = g->row[v]; j < g->row[v + 1]; j++) val += deg[g->col[j]];
            break;
        case 6: val = 0;
            for (int64_t j = g->row[u]; j < g->row[u + 1]; j++) val += KC(g->col[j]);
            for (int64_t j = g->row[v]; j < g->row[v + 1]; j++) val += KC(g->col[j]);
            break;
        case 7: val = 0;
            for (int64_t j = g->row[u]; j < g->row[u + 1]; j++) val += (long long)deg[g->col[j]] * KC(g->col[j]);
            for (int64_t j = g->row[v]; j < g->row[v + 1]; j++) val += (long long)deg[g->col[j]] * KC(g->col[j]);
            break;
        default: val = g->row[u + 1] - g->row[u]; break;
        }
        E[p].id = e;
        E[p].b = (int32_t)val;              /* Vertex stores an int bound */
    }
    #undef KC
    qsort(E, (size_t)m, sizeof(vtx), small_to_large ? cmp_incr : cmp_decr);
    for (int64_t e = 0; e < m; e++) e_ordered[e] = E[e].id;
    free(E); free(deg); free(core);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* decomposition loop, graphlet_core.cpp:1366-1448 / 1483-1599 (CSR branch) */
/* ------------------------------------------------------------------------ */
void or_decompose(const or_csr *g, const int32_t *e_v, const int32_t *e_u,
                  const int32_t *order, int64_t count, int64_t m_total, int threads,
                  uint64_t *sums_out,
                  uint64_t *tri, uint64_t *k4o, uint64_t *c4o,
                  uint64_t *chordal, uint64_t *path4, uint64_t *star3) {
    const int64_t n = g->n, m = m_total;
    const int64_t *row = g->row;
    const int32_t *col = g->col;
    int64_t maxd = 0;
    for (int64_t v = 0; v < n; v++) if (row[v + 1] - row[v] > maxd) maxd = row[v + 1] - row[v];
    int nt = threads > 0 ? threads : omp_get_max_threads();
    uint64_t *part = (uint64_t *)calloc((size_t)nt * OR_NSUMS, sizeof(uint64_t));
    #pragma omp parallel num_threads(nt)
    {
        int tid = omp_get_thread_num();
        uint64_t *S = part + (size_t)tid * OR_NSUMS;
        int64_t *ind = (int64_t *)calloc((size_t)(n ? n : 1), sizeof(int64_t));
        int64_t *Tvu = (int64_t *)calloc((size_t)maxd + 1, sizeof(int64_t));
        int64_t *Wu = (int64_t *)calloc((size_t)maxd + 1, sizeof(int64_t));
        #pragma omp for schedule(dynamic, 64)
        for (int64_t e = 0; e < count; e++) {
            int64_t id = order[e];
            int64_t v = e_v[id], u = e_u[id];
            long long dv = row[v + 1] - row[v], du = row[u + 1] - row[u];
            uint64_t tc = 0, wc = 0, k4 = 0, c4 = 0;
            /* mark_neighbors, 1068-1074 */
            for (int64_t i = row[v]; i < row[v + 1]; i++) if (col[i] != u) ind[col[i]] = 1;
            /* triangles_and_wedges, 1135-1152 */
            for (int64_t j = row[u]; j < row[u + 1]; j++) {
                int64_t w = col[j];
                if (w == v) continue;
                if (ind[w] == 1) { ind[w] = 3; Tvu[tc++] = w; }
                else { Wu[wc++] = w; ind[w] = 2; }
            }
            /* solve_graphlet_equations, 1325-1364 (uint64 arithmetic) */
            uint64_t T = tc;
            uint64_t st = (uint64_t)dv - T - 1;
            st = st + (uint64_t)du - T - 1;
            uint64_t D = (uint64_t)n - (st + T) - 2;
            S[0] += T; S[1] += st; S[2] += D;
            S[5] += (uint64_t)((double)((T + st) * (T + st - 1)) / 2.0);
            S[6] += (uint64_t)((double)(st * (st - 1)) / 2.0);
            S[7] += (uint64_t)((double)(T * (T - 1)) / 2.0);
            S[8] += T * st;
            uint64_t sv = (uint64_t)dv - T - 1, su = (uint64_t)du - T - 1;
            S[9] += sv * su;
            S[10] += sv * (sv - 1) / 2 + su * (su - 1) / 2;
            S[11] += (uint64_t)(m - dv - du + 1);
            S[12] += (uint64_t)((double)(D * (D - 1)) / 2.0);
            S[13] += st * D;
            S[14] += T * D;
            /* cycle, 1166-1176 */
            for (uint64_t j = 0; j < wc; j++) {
                int64_t w = Wu[j];
                for (int64_t i = row[w]; i < row[w + 1]; i++) if (ind[col[i]] == 1) c4++;
            }
            /* clique, 1229-1240 */
            for (uint64_t t = 0; t < tc; t++) {
                int64_t w = Tvu[t];
                for (int64_t i = row[w]; i < row[w + 1]; i++) if (ind[col[i]] == 3) k4++;
                ind[w] = 0;
            }
            S[3] += k4; S[4] += c4;
            if (tri) {
                tri[e] = T; k4o[e] = k4; c4o[e] = c4;
                /* micro columns, 1567-1575 */
                long long sv_s = dv - (long long)T - 1, su_s = du - (long long)T - 1;
                chordal[e] = (uint64_t)((double)(T * (T - 1)) / 2.0 - (double)k4);
                path4[e] = (uint64_t)(sv_s * su_s) - c4;
                uint64_t s3 = (uint64_t)((double)(sv_s * (sv_s - 1)) / 2.0);
                s3 = (uint64_t)((double)s3 + (double)(su_s * (su_s - 1)) / 2.0);
                star3[e] = s3;      /* minus local_4_tailed_tris[e], which is never set */
            }
            /* reset_perfect_hash, 1085-1088 */
            for (int64_t i = row[v]; i < row[v + 1]; i++) ind[col[i]] = 0;
        }
        free(ind); free(Tvu); free(Wu);
    }
    for (int k = 0; k < OR_NSUMS; k++) {
        uint64_t s = 0;
        for (int t = 0; t < nt; t++) s += part[(size_t)t * OR_NSUMS + k];
        sums_out[k] = s;
    }
    free(part);
}

/* ------------------------------------------------------------------------ */
/* derivation, graphlet_core.cpp:1449-1472 and graphlet_core.h:548-557       */
/* ------------------------------------------------------------------------ */
/* double -> unsigned long long as the x86-64 build converts it
 * (comisd 2^63 / cvttsd2si / btc 63) */
static uint64_t d2u(double x) {
    const double two63 = 9223372036854775808.0;
    if (x >= two63) {
        double y = x - two63;
        int64_t t = (y >= two63 || y != y) ? INT64_MIN : (int64_t)y;
        return (uint64_t)t ^ 0x8000000000000000ULL;
    }
    if (!(x > -two63)) return 0x8000000000000000ULL;   /* cvttsd2si overflow value */
    return (uint64_t)(int64_t)x;
}

void or_derive(int64_t n, int64_t m, const uint64_t *s, uint64_t *t) {
    const double third = 1.0 / 3.0, sixth = 1.0 / 6.0;
    uint64_t ST = s[0], SS = s[1], SD = s[2], SK = s[3], SC = s[4];
    uint64_t N3 = s[7], N4 = s[8], N5 = s[9], N6 = s[10], N7 = s[11], N8 = s[12], N9 = s[13], N10 = s[14];
    uint64_t e1 = (uint64_t)m;
    int64_t nn1 = (int64_t)((uint64_t)n * (uint64_t)(n - 1));
    uint64_t i2 = d2u((double)nn1 * 0.5 - (double)m);
    uint64_t tris = d2u((double)ST * third);
    uint64_t s2 = d2u((double)SS * 0.5);
    uint64_t e31 = SD;
    int64_t p3 = (int64_t)((uint64_t)n * (uint64_t)(n - 1) * (uint64_t)(n - 2));   /* wraps like imul */
    uint64_t i3 = (uint64_t)(p3 / 6);
    i3 = i3 - (tris + s2 + e31);
    uint64_t k4 = d2u((double)SK * sixth);
    uint64_t cc = N3 - 6 * k4;
    uint64_t c4 = d2u((double)SC * 0.25);
    uint64_t p4 = N5 - 4 * c4;
    uint64_t tt = d2u((double)(N4 - 4 * cc) * 0.5);
    uint64_t st3 = d2u((double)(N6 - tt) * third);
    uint64_t e2 = N7;
    e2 = e2 - 6 * k4 - 4 * c4 - 4 * cc - 2 * p4 - 2 * tt;
    e2 = d2u((double)e2 * 0.5);
    uint64_t e14 = N8 - 2 * e2;
    uint64_t st24 = N9 - 2 * p4;
    st24 = d2u((double)st24 * 0.5);
    uint64_t tri4 = N10 - tt;
    tri4 = d2u((double)tri4 * third);
    uint64_t un = (uint64_t)n;
    uint64_t i4 = (un * (un - 1) * (un - 2) * (un - 3)) / 24;
    i4 = i4 - (k4 + cc + c4 + p4 + tt + st3);
    i4 = i4 - (e2 + e14 + st24 + tri4);
    uint64_t out[OR_NTOTALS] = { e1, i2, tris, s2, e31, i3, k4, cc, tt, c4, st3, p4, e14, e2, st24, tri4, i4 };
    memcpy(t, out, sizeof(out));
}

/* ------------------------------------------------------------------------ */
static int has_nb(const or_csr *g, int64_t a, int64_t b) {
    for (int64_t j = g->row[a]; j < g->row[a + 1]; j++) if (g->col[j] == b) return 1;
    return 0;
}
void or_brute_edge(const or_csr *g, int32_t v, int32_t u, uint64_t out[5]) {
    uint64_t T = 0, Sv = 0, Su = 0, K4 = 0, C4 = 0;
    for (int64_t j = g->row[v]; j < g->row[v + 1]; j++) {
        int64_t x = g->col[j];
        if (x == u) continue;
        if (has_nb(g, u, x)) T++; else Sv++;
    }
    for (int64_t j = g->row[u]; j < g->row[u + 1]; j++) {
        int64_t x = g->col[j];
        if (x == v) continue;
        if (!has_nb(g, v, x)) Su++;
    }
    /* K4: adjacent pairs inside N(v) & N(u); C4: pairs (w in Su, x in Sv) adjacent */
    for (int64_t i = g->row[v]; i < g->row[v + 1]; i++) {
        int64_t a = g->col[i];
        if (a == u) continue;
        int a_in_t = has_nb(g, u, a);
        for (int64_t j = g->row[u]; j < g->row[u + 1]; j++) {
            int64_t b = g->col[j];
            if (b == v || b == a) continue;
            int b_in_t = has_nb(g, v, b);
            if (!has_nb(g, a, b)) continue;
            if (a_in_t && b_in_t && a < b) K4++;
            if (!a_in_t && !b_in_t) C4++;
        }
    }
    out[0] = T; out[1] = Sv; out[2] = Su; out[3] = K4; out[4] = C4;
}
