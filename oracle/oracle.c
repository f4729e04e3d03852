/*
 * CPU oracle for the ForkMatch edge-labelled subgraph matcher.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker or
 * the timed CPU baseline. The product (libforkmatch.so) never links or calls
 * it.
 *
 * It restates, in plain C and independently of the HIP engine:
 *  - the reference's matching order and per-depth requirements
 *    (genereateCCSRRequirements / preProcessQuery, src/ccsr.cuh:229-326,
 *    485-549, with the vertex ranking of src/hostparser.cuh:20-30,420);
 *  - the reference's matching semantics (src/smsm.cu:297-370 row
 *    intersection, src/smsm.cu:187-227 injectivity), as full enumeration
 *    without the reference's inexact symmetry breaking: an embedding is an
 *    injective map m with a data edge (m(u), m(w)) of label l for every
 *    query line (u, w, l); vertex labels are ignored;
 *  - the algorithmic-byte counters of the roofline: neighbour words of every
 *    requirement row of every partial embedding, and partial counts per
 *    depth.
 * The data structure is deliberately different from the engine's: rows are
 * found by binary search over the sorted keys of non-empty (label, vertex)
 * rows, and the candidates of a depth come from its first requirement.
 *
 * Pinning (tests/test_oracle.py): brute force over all injective maps on
 * tiny graphs, closed forms on single-label graphs (paths, stars, triangle,
 * 4-cycle counts from degrees and traces of A^k), hand-derived plans of the
 * eight reference query fixtures, and the table invariants.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define OMAXN 32

typedef struct {
    uint64_t nrows;
    uint64_t *keys;  /* sorted row keys label_index * V + vertex */
    uint64_t *off;   /* nrows + 1 */
    uint32_t *col;
} orows;

typedef struct {
    uint32_t V, L;
    uint32_t *labels; /* sorted distinct label values */
    orows dir[2];     /* 0: by source (out-rows), 1: by destination (in-rows) */
    int symmetric;
} ograph;

typedef struct {
    uint64_t key;
    uint32_t val;
} kv;

static int kv_cmp(const void *a, const void *b) {
    const kv *x = (const kv *)a, *y = (const kv *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->val != y->val) return x->val < y->val ? -1 : 1;
    return 0;
}

static int u32_cmp(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : (x > y);
}

static int64_t label_index(const ograph *g, uint32_t value) {
    uint32_t lo = 0, hi = g->L;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (g->labels[mid] < value) lo = mid + 1; else hi = mid;
    }
    if (lo < g->L && g->labels[lo] == value) return lo;
    return -1;
}

static void build_rows(orows *r, kv *e, uint64_t n) {
    qsort(e, n, sizeof(kv), kv_cmp);
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; i++)
        if (m == 0 || e[i].key != e[m - 1].key || e[i].val != e[m - 1].val) e[m++] = e[i];
    uint64_t nrows = 0;
    for (uint64_t i = 0; i < m; i++)
        if (i == 0 || e[i].key != e[i - 1].key) nrows++;
    r->nrows = nrows;
    r->keys = (uint64_t *)malloc(sizeof(uint64_t) * (nrows + 1));
    r->off = (uint64_t *)malloc(sizeof(uint64_t) * (nrows + 1));
    r->col = (uint32_t *)malloc(sizeof(uint32_t) * (m + 1));
    uint64_t k = 0;
    for (uint64_t i = 0; i < m; i++) {
        if (i == 0 || e[i].key != e[i - 1].key) {
            r->keys[k] = e[i].key;
            r->off[k] = i;
            k++;
        }
        r->col[i] = e[i].val;
    }
    r->off[nrows] = m;
}

/* Row of (label value, vertex) in direction d; returns its length. */
static uint32_t get_row(const ograph *g, int d, uint32_t label, uint32_t v, const uint32_t **row) {
    int64_t li = label_index(g, label);
    *row = NULL;
    if (li < 0) return 0;
    const orows *r = &g->dir[g->symmetric ? 0 : d];
    uint64_t key = (uint64_t)li * g->V + v;
    uint64_t lo = 0, hi = r->nrows;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (r->keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    if (lo == r->nrows || r->keys[lo] != key) return 0;
    *row = r->col + r->off[lo];
    return (uint32_t)(r->off[lo + 1] - r->off[lo]);
}

static int row_contains(const uint32_t *row, uint32_t len, uint32_t x) {
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (row[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < len && row[lo] == x;
}

void *oracle_graph_new(uint32_t V, uint64_t E, const uint32_t *src, const uint32_t *dst,
                       const uint32_t *lab) {
    ograph *g = (ograph *)calloc(1, sizeof(ograph));
    g->V = V;
    uint32_t *ls = (uint32_t *)malloc(sizeof(uint32_t) * (E + 1));
    memcpy(ls, lab, sizeof(uint32_t) * E);
    qsort(ls, E, sizeof(uint32_t), u32_cmp);
    uint32_t L = 0;
    for (uint64_t i = 0; i < E; i++)
        if (i == 0 || ls[i] != ls[i - 1]) ls[L++] = ls[i];
    g->labels = ls;
    g->L = L;
    kv *e = (kv *)malloc(sizeof(kv) * (E + 1));
    for (int d = 0; d < 2; d++) {
        for (uint64_t i = 0; i < E; i++) {
            uint32_t from = d == 0 ? src[i] : dst[i], to = d == 0 ? dst[i] : src[i];
            e[i].key = (uint64_t)label_index(g, lab[i]) * V + from;
            e[i].val = to;
        }
        build_rows(&g->dir[d], e, E);
    }
    free(e);
    /* symmetric iff the out-rows equal the in-rows */
    const orows *a = &g->dir[0], *b = &g->dir[1];
    g->symmetric = a->nrows == b->nrows && a->off[a->nrows] == b->off[b->nrows] &&
                   memcmp(a->keys, b->keys, sizeof(uint64_t) * a->nrows) == 0 &&
                   memcmp(a->off, b->off, sizeof(uint64_t) * (a->nrows + 1)) == 0 &&
                   memcmp(a->col, b->col, sizeof(uint32_t) * a->off[a->nrows]) == 0;
    return g;
}

void oracle_graph_free(void *p) {
    ograph *g = (ograph *)p;
    if (!g) return;
    for (int d = 0; d < 2; d++) {
        free(g->dir[d].keys);
        free(g->dir[d].off);
        free(g->dir[d].col);
    }
    free(g->labels);
    free(g);
}

int oracle_graph_symmetric(void *p) { return ((ograph *)p)->symmetric; }

/* ------------------------------------------------------------------------ */
/* Plan: restatement of the reference's ordering and requirement generation. */

typedef struct {
    uint32_t n;
    uint32_t order[OMAXN];
    uint32_t depth_of[OMAXN];
    uint32_t req_begin[OMAXN + 1];
    uint32_t anchor[OMAXN * OMAXN * 4];
    uint32_t label[OMAXN * OMAXN * 4];
    uint32_t dir[OMAXN * OMAXN * 4];
    uint32_t nreqs;
} oplan;

typedef struct {
    uint32_t a, l, d;
} oreq;

static int oreq_cmp(const void *x, const void *y) {
    const oreq *p = (const oreq *)x, *q = (const oreq *)y;
    /* climb ascending == anchor depth descending (src/ccsr.cuh:293-300) */
    if (p->a != q->a) return p->a > q->a ? -1 : 1;
    if (p->l != q->l) return p->l < q->l ? -1 : 1;
    if (p->d != q->d) return p->d < q->d ? -1 : 1;
    return 0;
}

/* Returns 0 on success, 1 for a disconnected query (the reference's
 * exit(1), src/ccsr.cuh:251-254), 2 for bad input. Duplicate lines are
 * collapsed first. */
static int make_plan(uint32_t n, uint32_t m, const uint32_t *qs, const uint32_t *qd,
                     const uint32_t *ql, const uint32_t *order_in, oplan *p) {
    if (n == 0 || n > OMAXN) return 2;
    uint32_t es[OMAXN * OMAXN * 4][3];
    uint32_t ne = 0;
    for (uint32_t i = 0; i < m; i++) {
        if (qs[i] >= n || qd[i] >= n) return 2;
        int dup = 0;
        for (uint32_t j = 0; j < ne; j++)
            if (es[j][0] == qs[i] && es[j][1] == qd[i] && es[j][2] == ql[i]) dup = 1;
        if (!dup) {
            if (ne >= OMAXN * OMAXN * 4) return 2;
            es[ne][0] = qs[i];
            es[ne][1] = qd[i];
            es[ne][2] = ql[i];
            ne++;
        }
    }
    /* rank: out-degree (number of lines leaving the vertex) descending, id
     * ascending (OrderPair, src/hostparser.cuh:20-30) */
    uint32_t deg[OMAXN] = {0}, rankv[OMAXN], rank_of[OMAXN];
    for (uint32_t i = 0; i < ne; i++) deg[es[i][0]]++;
    for (uint32_t i = 0; i < n; i++) rankv[i] = i;
    for (uint32_t i = 1; i < n; i++) { /* insertion sort: stable */
        uint32_t v = rankv[i];
        int j = (int)i - 1;
        while (j >= 0 && deg[rankv[j]] < deg[v]) {
            rankv[j + 1] = rankv[j];
            j--;
        }
        rankv[j + 1] = v;
    }
    for (uint32_t i = 0; i < n; i++) rank_of[rankv[i]] = i;
    /* mapping state as in genereateCCSRRequirements: -1 unseen, 0 frontier,
     * > 0 mapped */
    int state[OMAXN];
    for (uint32_t i = 0; i < n; i++) state[i] = -1;
    p->n = n;
    if (order_in && order_in[0] >= n) return 2;
    state[order_in ? order_in[0] : rankv[0]] = 0;
    p->nreqs = 0;
    for (uint32_t d = 0; d < n; d++) {
        int pick = -1;
        if (order_in) {
            /* a given order: each vertex must be on the frontier when taken */
            if (order_in[d] >= n || state[order_in[d]] != 0) return 2;
            pick = (int)order_in[d];
        } else {
            /* the frontier vertex of smallest rank */
            for (uint32_t r = 0; r < n; r++)
                if (state[rankv[r]] == 0) {
                    pick = (int)rankv[r];
                    break;
                }
        }
        if (pick < 0) return 1;
        uint32_t u = (uint32_t)pick;
        state[u] = 1;
        p->order[d] = u;
        p->depth_of[u] = d;
        oreq rs[OMAXN * 8];
        uint32_t nr = 0;
        for (uint32_t i = 0; i < ne; i++) {
            uint32_t a = es[i][0], b = es[i][1], l = es[i][2];
            if (a == u && b == u) {
                rs[nr++] = (oreq){d, l, 0};
                continue;
            }
            uint32_t w = a == u ? b : (b == u ? a : UINT32_MAX);
            if (w == UINT32_MAX) continue;
            if (state[w] < 0) state[w] = 0; /* discovered */
            if (state[w] > 0 && w != u) {
                /* w was mapped at an earlier depth */
                rs[nr++] = (oreq){p->depth_of[w], l, a == u ? 1u : 0u};
            }
        }
        qsort(rs, nr, sizeof(oreq), oreq_cmp);
        p->req_begin[d] = p->nreqs;
        for (uint32_t i = 0; i < nr; i++) {
            p->anchor[p->nreqs] = rs[i].a;
            p->label[p->nreqs] = rs[i].l;
            p->dir[p->nreqs] = rs[i].d;
            p->nreqs++;
        }
    }
    p->req_begin[n] = p->nreqs;
    (void)rank_of;
    return 0;
}

int oracle_plan(uint32_t n, uint32_t m, const uint32_t *qs, const uint32_t *qd, const uint32_t *ql,
                const uint32_t *order_in, uint32_t *order, uint32_t *req_begin, uint32_t *anchor,
                uint32_t *label, uint32_t *dir, uint32_t *nreqs) {
    oplan *p = (oplan *)malloc(sizeof(oplan));
    int rc = make_plan(n, m, qs, qd, ql, order_in, p);
    if (rc == 0) {
        memcpy(order, p->order, sizeof(uint32_t) * n);
        memcpy(req_begin, p->req_begin, sizeof(uint32_t) * (n + 1));
        memcpy(anchor, p->anchor, sizeof(uint32_t) * p->nreqs);
        memcpy(label, p->label, sizeof(uint32_t) * p->nreqs);
        memcpy(dir, p->dir, sizeof(uint32_t) * p->nreqs);
        *nreqs = p->nreqs;
    }
    free(p);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Backtracking enumeration. */

typedef struct {
    const ograph *g;
    const oplan *p;
    int sym;
    /* requirement list per depth with symmetric-graph de-duplication of
     * (anchor, label) pairs, and self-loop requirements apart */
    uint32_t nr[OMAXN], ra[OMAXN][OMAXN * 8], rl[OMAXN][OMAXN * 8], rd[OMAXN][OMAXN * 8];
    uint32_t ns[OMAXN], sl[OMAXN][OMAXN * 8], sd[OMAXN][OMAXN * 8];
    int want_table;
    uint64_t max_rows;
    uint32_t *table;
    int instrument;
    /* shared work distribution */
    uint64_t root_begin, root_end, root_next, root_stride;
    pthread_mutex_t lock;
    uint64_t rows_written;
    uint64_t count;
    uint64_t row_words;
    uint64_t partials[OMAXN];
    /* time budget (timed baselines only): past `deadline` (monotonic
     * seconds, 0 = none) the workers stop and the counts are partial */
    double deadline;
    int stop;
} octx;

static double now_s(void);

typedef struct {
    octx *c;
    uint64_t ticks;
    uint64_t count, row_words, partials[OMAXN];
    uint32_t *buf;
    uint64_t buf_rows, buf_cap;
} oworker;

static void emit(oworker *w, const uint32_t *m) {
    octx *c = w->c;
    uint32_t n = c->p->n;
    w->count++;
    if (!c->want_table) return;
    if (w->buf_rows == w->buf_cap) {
        w->buf_cap = w->buf_cap ? w->buf_cap * 2 : 1024;
        w->buf = (uint32_t *)realloc(w->buf, sizeof(uint32_t) * n * w->buf_cap);
    }
    /* columns in query-vertex id order */
    uint32_t *row = w->buf + w->buf_rows * n;
    for (uint32_t d = 0; d < n; d++) row[c->p->order[d]] = m[d];
    w->buf_rows++;
}

static void flush_rows(oworker *w) {
    octx *c = w->c;
    if (!c->want_table || w->buf_rows == 0) return;
    uint32_t n = c->p->n;
    pthread_mutex_lock(&c->lock);
    uint64_t room = c->rows_written < c->max_rows ? c->max_rows - c->rows_written : 0;
    uint64_t take = w->buf_rows < room ? w->buf_rows : room;
    memcpy(c->table + c->rows_written * n, w->buf, sizeof(uint32_t) * n * take);
    c->rows_written += take;
    pthread_mutex_unlock(&c->lock);
    w->buf_rows = 0;
}

static int check_self(const octx *c, uint32_t d, uint32_t x) {
    for (uint32_t i = 0; i < c->ns[d]; i++) {
        const uint32_t *row;
        uint32_t len = get_row(c->g, (int)c->sd[d][i], c->sl[d][i], x, &row);
        if (!row_contains(row, len, x)) return 0;
    }
    return 1;
}

static void extend(oworker *w, uint32_t d, uint32_t *m) {
    octx *c = w->c;
    uint32_t n = c->p->n;
    if (c->deadline > 0) {
        if (__atomic_load_n(&c->stop, __ATOMIC_RELAXED)) return;
        if ((++w->ticks & 4095) == 0 && now_s() > c->deadline) __atomic_store_n(&c->stop, 1, __ATOMIC_RELAXED);
    }
    if (d == n) {
        emit(w, m);
        return;
    }
    const uint32_t *rows[OMAXN * 8];
    uint32_t lens[OMAXN * 8];
    uint32_t k = c->nr[d];
    for (uint32_t i = 0; i < k; i++) {
        lens[i] = get_row(c->g, (int)c->rd[d][i], c->rl[d][i], m[c->ra[d][i]], &rows[i]);
        if (c->instrument) w->row_words += lens[i];
        else if (lens[i] == 0) return;
    }
    for (uint32_t i = 0; i < k; i++)
        if (lens[i] == 0) return;
    /* candidates: the first requirement's row */
    for (uint32_t j = 0; j < lens[0]; j++) {
        uint32_t x = rows[0][j];
        int ok = 1;
        for (uint32_t a = 0; a < d && ok; a++)
            if (m[a] == x) ok = 0;
        for (uint32_t i = 1; i < k && ok; i++)
            if (!row_contains(rows[i], lens[i], x)) ok = 0;
        if (ok && c->ns[d]) ok = check_self(c, d, x);
        if (!ok) continue;
        m[d] = x;
        w->partials[d]++;
        extend(w, d + 1, m);
    }
}

static void *worker_main(void *arg) {
    oworker *w = (oworker *)arg;
    octx *c = w->c;
    uint32_t m[OMAXN];
    const uint64_t chunk = 64;
    /* roots are root_begin + i * root_stride for i in [0, nroots) */
    uint64_t nroots = (c->root_end - c->root_begin + c->root_stride - 1) / c->root_stride;
    while (!__atomic_load_n(&c->stop, __ATOMIC_RELAXED)) {
        uint64_t b = __atomic_fetch_add(&c->root_next, chunk, __ATOMIC_RELAXED);
        if (b >= nroots) break;
        uint64_t e = b + chunk < nroots ? b + chunk : nroots;
        for (uint64_t i = b; i < e; i++) {
            uint64_t v = c->root_begin + i * c->root_stride;
            if (c->ns[0] && !check_self(c, 0, (uint32_t)v)) continue;
            m[0] = (uint32_t)v;
            w->partials[0]++;
            extend(w, 1, m);
        }
        if (w->buf_rows > 65536) flush_rows(w);
    }
    flush_rows(w);
    return NULL;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/*
 * Enumerates every embedding of the query (n vertices, m directed lines) in
 * graph gp, in the reference's matching order, or in order_in when it is not
 * NULL (each vertex adjacent to an earlier one; the embedding set is the
 * same, the partials per depth are those of that order). Roots are restricted to root_begin, root_begin + root_stride, ...
 * below root_end (pass 0, UINT64_MAX and 1 for all). threads <= 0 selects 1. Writes up to max_rows rows
 * (query-vertex column order) into table when want_table. Outputs the count,
 * the rows written, the neighbour words of all requirement rows of all
 * partial embeddings (row_words; instrument = 1 only), the partials per
 * depth (n entries) and the enumeration time in seconds.
 * Returns 0, or the planner's error (1 disconnected, 2 bad query).
 */
int oracle_match_budget(void *gp, uint32_t n, uint32_t m, const uint32_t *qs, const uint32_t *qd,
                        const uint32_t *ql, const uint32_t *order_in, int threads, uint64_t root_begin,
                        uint64_t root_end, uint64_t root_stride, int want_table, uint64_t max_rows,
                        uint32_t *table, int instrument, double max_seconds, uint64_t *count, uint64_t *rows,
                        uint64_t *row_words, uint64_t *partials, double *seconds, int *stopped);

int oracle_match(void *gp, uint32_t n, uint32_t m, const uint32_t *qs, const uint32_t *qd,
                 const uint32_t *ql, const uint32_t *order_in, int threads, uint64_t root_begin,
                 uint64_t root_end, uint64_t root_stride, int want_table, uint64_t max_rows, uint32_t *table,
                 int instrument, uint64_t *count, uint64_t *rows, uint64_t *row_words, uint64_t *partials,
                 double *seconds) {
    return oracle_match_budget(gp, n, m, qs, qd, ql, order_in, threads, root_begin, root_end, root_stride,
                               want_table, max_rows, table, instrument, 0.0, count, rows, row_words, partials,
                               seconds, NULL);
}

/*
 * oracle_match with a time budget, for the timed CPU baseline: after
 * max_seconds (0 = none) of enumeration the workers stop; *stopped is then 1
 * and the outputs count what was enumerated until then (a throughput
 * sample, not a result).
 */
int oracle_match_budget(void *gp, uint32_t n, uint32_t m, const uint32_t *qs, const uint32_t *qd,
                        const uint32_t *ql, const uint32_t *order_in, int threads, uint64_t root_begin,
                        uint64_t root_end, uint64_t root_stride, int want_table, uint64_t max_rows,
                        uint32_t *table, int instrument, double max_seconds, uint64_t *count, uint64_t *rows,
                        uint64_t *row_words, uint64_t *partials, double *seconds, int *stopped) {
    const ograph *g = (const ograph *)gp;
    octx *c = (octx *)calloc(1, sizeof(octx));
    oplan *p = (oplan *)malloc(sizeof(oplan));
    int rc = make_plan(n, m, qs, qd, ql, order_in, p);
    if (rc) {
        free(p);
        free(c);
        return rc;
    }
    c->g = g;
    c->p = p;
    c->sym = g->symmetric;
    for (uint32_t d = 0; d < n; d++) {
        c->nr[d] = 0;
        c->ns[d] = 0;
        for (uint32_t i = p->req_begin[d]; i < p->req_begin[d + 1]; i++) {
            uint32_t a = p->anchor[i], l = p->label[i], dd = p->dir[i];
            if (g->symmetric) dd = 0;
            if (a == d) {
                c->sl[d][c->ns[d]] = l;
                c->sd[d][c->ns[d]] = dd;
                c->ns[d]++;
                continue;
            }
            int dup = 0;
            for (uint32_t j = 0; j < c->nr[d]; j++)
                if (c->ra[d][j] == a && c->rl[d][j] == l && c->rd[d][j] == dd) dup = 1;
            if (dup) continue;
            c->ra[d][c->nr[d]] = a;
            c->rl[d][c->nr[d]] = l;
            c->rd[d][c->nr[d]] = dd;
            c->nr[d]++;
        }
    }
    c->want_table = want_table;
    c->max_rows = max_rows;
    c->table = table;
    c->instrument = instrument;
    c->root_begin = root_begin < g->V ? root_begin : g->V;
    c->root_end = root_end < g->V ? root_end : g->V;
    if (c->root_end < c->root_begin) c->root_end = c->root_begin;
    c->root_stride = root_stride ? root_stride : 1;
    pthread_mutex_init(&c->lock, NULL);
    if (threads <= 0) threads = 1;
    oworker *ws = (oworker *)calloc((size_t)threads, sizeof(oworker));
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    double t0 = now_s();
    c->deadline = max_seconds > 0 ? t0 + max_seconds : 0.0;
    for (int t = 0; t < threads; t++) {
        ws[t].c = c;
        if (threads == 1) worker_main(&ws[t]);
        else pthread_create(&th[t], NULL, worker_main, &ws[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    double t1 = now_s();
    uint64_t total = 0, words = 0;
    uint64_t part[OMAXN] = {0};
    for (int t = 0; t < threads; t++) {
        total += ws[t].count;
        words += ws[t].row_words;
        for (uint32_t d = 0; d < n; d++) part[d] += ws[t].partials[d];
        free(ws[t].buf);
    }
    if (count) *count = total;
    if (rows) *rows = c->rows_written;
    if (row_words) *row_words = words;
    if (partials) memcpy(partials, part, sizeof(uint64_t) * n);
    if (seconds) *seconds = t1 - t0;
    if (stopped) *stopped = c->stop;
    pthread_mutex_destroy(&c->lock);
    free(ws);
    free(th);
    free(p);
    free(c);
    return 0;
}
