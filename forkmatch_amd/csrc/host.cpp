// Host side of ForkMatch: .g parsing and writing, the label-partitioned CSR
// builder, the query planner, the binary graph cache and the seeded
// synthetic generators. No device code lives here.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "fm_internal.hpp"

namespace fm {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

bool read_file(const std::string &path, std::string &buf) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long long size = ftello(f);
    rewind(f);
    buf.resize((size_t)std::max<long long>(size, 0));
    size_t got = size > 0 ? fread(&buf[0], 1, (size_t)size, f) : 0;
    fclose(f);
    buf.resize(got);
    return true;
}

// ---------------------------------------------------------------------------
// RNG: splitmix64 seeding + xoshiro256** (public-domain algorithms by
// Vigna / Blackman).
static uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
Rng::Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &v : s) v = splitmix64(x);
}
uint64_t Rng::next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}
double Rng::uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
uint64_t Rng::below(uint64_t n) {
    // Lemire's multiply-shift with rejection for exact uniformity.
    if (n == 0) return 0;
    unsigned __int128 m = (unsigned __int128)next() * n;
    uint64_t l = (uint64_t)m;
    if (l < n) {
        uint64_t t = (0 - n) % n;
        while (l < t) {
            m = (unsigned __int128)next() * n;
            l = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

// ---------------------------------------------------------------------------
// Graph construction.

uint32_t HostGraph::label_id(uint32_t value) const {
    auto it = std::lower_bound(label_values.begin(), label_values.end(), value);
    if (it == label_values.end() || *it != value) return UINT32_MAX;
    return (uint32_t)(it - label_values.begin());
}

static unsigned hw_threads() {
    unsigned t = std::thread::hardware_concurrency();
    return std::max(1u, std::min(t ? t : 1u, 16u));
}

template <class F>
static void parallel_for(uint64_t n, F f) {
    unsigned T = hw_threads();
    if (n < 100000 || T == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    uint64_t step = (n + T - 1) / T;
    for (unsigned t = 0; t < T; t++) {
        uint64_t b = std::min<uint64_t>(n, t * step), e = std::min<uint64_t>(n, b + step);
        if (b < e) th.emplace_back(f, b, e);
    }
    for (auto &x : th) x.join();
}

// f(i) for every item i < n, the items spread over the threads (for few,
// large items such as labels).
template <class F>
static void parallel_for_items(uint64_t n, F f) {
    const unsigned T = std::max<unsigned>(1, std::min<uint64_t>(hw_threads(), n));
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&] {
            for (uint64_t i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto &x : th) x.join();
}

// out[r] = in[0] + ... + in[r - 1] for r <= n (out has n + 1 entries; out and
// in may alias): per-range sums on all threads, a scan of the range sums,
// then each range written from its offset. Returns the total.
static uint64_t parallel_exclusive_scan(const uint32_t *in, uint32_t *out, uint64_t n) {
    const unsigned T = hw_threads();
    const uint64_t step = (n + T - 1) / std::max(1u, T);
    std::vector<uint64_t> part(T + 1, 0);
    parallel_for(T, [&](uint64_t b, uint64_t e) {
        for (uint64_t t = b; t < e; t++) {
            uint64_t acc = 0;
            for (uint64_t r = t * step; r < std::min(n, (t + 1) * step); r++) acc += in[r];
            part[t + 1] = acc;
        }
    });
    for (unsigned t = 0; t < T; t++) part[t + 1] += part[t];
    // T ranges on T threads (parallel_for would run a short range list serially)
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t] {
            uint64_t acc = part[t];
            for (uint64_t r = t * step; r < std::min(n, (t + 1) * step); r++) {
                const uint32_t v = in[r];
                out[r] = (uint32_t)acc;
                acc += v;
            }
        });
    for (auto &x : th) x.join();
    out[n] = (uint32_t)part[T];
    return part[T];
}

// Counting sort of (row key -> value) pairs into a CSR with sorted,
// de-duplicated rows. key(i) = lid[i] * V + from[i].
static int build_dir(uint32_t V, uint32_t L, uint64_t n, const uint32_t *lid,
                     const uint32_t *from, const uint32_t *to, Csr &c) {
    uint64_t rows = (uint64_t)L * V;
    std::vector<uint32_t> cnt(rows + 1, 0);
    // counting and scattering run on all threads with atomic counters; the
    // order inside a row is then fixed by the row sort
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) __atomic_fetch_add(&cnt[(uint64_t)lid[i] * V + from[i]], 1u, __ATOMIC_RELAXED);
    });
    c.ptr.assign(rows + 1, 0);
    parallel_exclusive_scan(cnt.data(), c.ptr.data(), rows);
    std::vector<uint32_t> tmp(n);
    {
        // reuse cnt as the running write cursor
        parallel_for(rows, [&](uint64_t b, uint64_t e) {
            std::memcpy(cnt.data() + b, c.ptr.data() + b, (size_t)(e - b) * 4);
        });
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++)
                tmp[__atomic_fetch_add(&cnt[(uint64_t)lid[i] * V + from[i]], 1u, __ATOMIC_RELAXED)] = to[i];
        });
    }
    // sort + unique each row (in parallel over row ranges), then compact
    std::vector<uint32_t> keep(rows + 1, 0);
    parallel_for(rows, [&](uint64_t b, uint64_t e) {
        for (uint64_t r = b; r < e; r++) {
            uint32_t s = c.ptr[r], t = c.ptr[r + 1];
            if (t - s > 1) {
                std::sort(tmp.begin() + s, tmp.begin() + t);
                keep[r] = (uint32_t)(std::unique(tmp.begin() + s, tmp.begin() + t) - (tmp.begin() + s));
            } else {
                keep[r] = t - s;
            }
        }
    });
    // compacted offsets (keep[] becomes the new row starts), then the copy
    const uint64_t w = parallel_exclusive_scan(keep.data(), keep.data(), rows);
    c.col.resize(w);
    parallel_for(rows, [&](uint64_t b, uint64_t e) {
        for (uint64_t r = b; r < e; r++)
            std::memcpy(c.col.data() + keep[r], tmp.data() + c.ptr[r], (size_t)(keep[r + 1] - keep[r]) * 4);
    });
    c.ptr.swap(keep);
    // per-label root lists: counts per label, their offsets, then the lists
    // (labels on all threads)
    std::vector<uint32_t> nroot(L + 1, 0);
    parallel_for_items(L, [&](uint64_t l) {
        const uint32_t *p = c.ptr.data() + l * V;
        uint32_t k = 0;
        for (uint32_t v = 0; v < V; v++) k += p[v + 1] > p[v];
        nroot[l] = k;
    });
    c.root_ptr.assign(L + 1, 0);
    for (uint32_t l = 0; l < L; l++) c.root_ptr[l + 1] = c.root_ptr[l] + nroot[l];
    c.roots.assign(c.root_ptr[L], 0);
    parallel_for_items(L, [&](uint64_t l) {
        const uint32_t *p = c.ptr.data() + l * V;
        uint32_t *dst = c.roots.data() + c.root_ptr[l];
        for (uint32_t v = 0; v < V; v++)
            if (p[v + 1] > p[v]) *dst++ = v;
    });
    csr_stats(c, V, L);
    return FM_OK;
}

void csr_stats(Csr &c, uint32_t V, uint32_t L) {
    c.row_sq.assign(L, 0.0);
    parallel_for_items(L, [&](uint64_t l) {
        const uint32_t *p = c.ptr.data() + l * V;
        double sq = 0;
        for (uint32_t v = 0; v < V; v++) {
            double k = (double)(p[v + 1] - p[v]);
            sq += k * k;
        }
        c.row_sq[l] = sq;
    });
}

static bool row_has(const Csr &c, uint32_t V, uint32_t l, uint32_t v, uint32_t x) {
    uint64_t r = (uint64_t)l * V + v;
    const uint32_t *b = c.col.data() + c.ptr[r], *e = c.col.data() + c.ptr[r + 1];
    return std::binary_search(b, e, x);
}

int build_graph(uint32_t V, uint64_t n, const uint32_t *src, const uint32_t *dst,
                const uint32_t *label, HostGraph &g, bool symmetric_by_construction) {
    if (n >= (1ull << 32) - 1)
        return fail(FM_ERR_NOT_SUPPORTED, "graph has %llu directed edges; the CSR holds < 2^32",
                    (unsigned long long)n);
    // the first edge naming a vertex >= V, if any (ranges on all threads)
    std::atomic<uint64_t> bad{n};
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e && i < bad.load(std::memory_order_relaxed); i++)
            if (src[i] >= V || dst[i] >= V) {
                uint64_t cur = bad.load();
                while (i < cur && !bad.compare_exchange_weak(cur, i)) {
                }
                break;
            }
    });
    if (bad.load() < n) {
        const uint64_t i = bad.load();
        return fail(FM_ERR_DATA_BAD_BODY, "edge %llu (%u -> %u) names a vertex >= %u", (unsigned long long)i,
                    src[i], dst[i], V);
    }
    g.V = V;
    // dense label ids
    std::atomic<uint32_t> maxa{0};
    parallel_for(n, [&](uint64_t b, uint64_t e) {
        uint32_t m = 0;
        for (uint64_t i = b; i < e; i++) m = std::max(m, label[i]);
        uint32_t cur = maxa.load();
        while (m > cur && !maxa.compare_exchange_weak(cur, m)) {
        }
    });
    const uint32_t maxl = maxa.load();
    std::vector<uint32_t> lid(n);
    g.label_values.clear();
    if (maxl < (1u << 24)) {
        std::vector<uint32_t> map((size_t)maxl + 1, UINT32_MAX);
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++) __atomic_store_n(&map[label[i]], 0u, __ATOMIC_RELAXED);
        });
        for (uint32_t v = 0; v <= maxl && n; v++)
            if (map[v] == 0) {
                map[v] = (uint32_t)g.label_values.size();
                g.label_values.push_back(v);
            }
        parallel_for(n, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; i++) lid[i] = map[label[i]];
        });
    } else {
        g.label_values.assign(label, label + n);
        std::sort(g.label_values.begin(), g.label_values.end());
        g.label_values.erase(std::unique(g.label_values.begin(), g.label_values.end()),
                             g.label_values.end());
        for (uint64_t i = 0; i < n; i++) lid[i] = g.label_id(label[i]);
    }
    uint32_t L = g.L();
    if ((uint64_t)L * V > (1ull << 34))
        return fail(FM_ERR_NOT_SUPPORTED,
                    "labels x vertices = %llu exceeds the dense row index limit 2^34",
                    (unsigned long long)L * V);
    build_dir(V, L, n, lid.data(), src, dst, g.out);
    // symmetry: every (u, v, l) has (v, u, l)
    std::atomic<bool> sym{true};
    if (!symmetric_by_construction) parallel_for((uint64_t)L * V, [&](uint64_t b, uint64_t e) {
        for (uint64_t r = b; r < e && sym.load(std::memory_order_relaxed); r++) {
            uint32_t l = (uint32_t)(r / V), u = (uint32_t)(r % V);
            for (uint32_t k = g.out.ptr[r]; k < g.out.ptr[r + 1]; k++)
                if (!row_has(g.out, V, l, g.out.col[k], u)) {
                    sym = false;
                    break;
                }
        }
    });
    g.symmetric = sym.load();
    g.in = Csr();
    if (!g.symmetric) build_dir(V, L, n, lid.data(), dst, src, g.in);
    return FM_OK;
}

// ---------------------------------------------------------------------------
// .g data graph parser. Grammar and checks follow parseGraph/parseHeader
// (src/trie.cuh:999-1097) and parseCSR/parseNum (src/trie.cuh:652-664,
// 800-895): "t # 0\n", a header "V E vlabelmax elabelmax\n", then
// "v id label" and "e src dst label" lines, and a final "t # -1".

static bool parse_uint(const char *&p, const char *end, uint32_t &v) {
    if (p >= end || *p < '0' || *p > '9') return false;
    uint64_t x = 0;
    while (p < end && *p >= '0' && *p <= '9') {
        x = x * 10 + (uint64_t)(*p - '0');
        if (x > 0xFFFFFFFFull) return false;
        p++;
    }
    v = (uint32_t)x;
    return true;
}

int parse_data_frame(const std::string &buf, DataFrame &f) {
    size_t size = buf.size();
    if (size < 6 || std::memcmp(buf.data(), "t # 0\n", 6) != 0)
        return fail(FM_ERR_DATA_BAD_START, "data graph does not start with \"t # 0\\n\"");
    size_t end = size;
    // One trailing newline is accepted (the reference requires none).
    if (end > 0 && buf[end - 1] == '\n') end--;
    if (end < 12 || std::memcmp(buf.data() + end - 6, "t # -1", 6) != 0)
        return fail(FM_ERR_DATA_BAD_END, "data graph does not end with \"t # -1\"");
    const char *p = buf.data() + 6, *stop = buf.data() + end - 6;
    uint32_t header[4] = {0, 0, 0, 0};
    int nv = 0;
    {
        // parseHeader: digits accumulate, ' ' moves to the next value, '\n'
        // ends the header, anything else is an error.
        bool any = false;
        while (true) {
            if (p >= stop) return fail(FM_ERR_DATA_BAD_HEADER, "malformed data graph header");
            char ch = *p++;
            if (ch >= '0' && ch <= '9') {
                if (nv >= 4) return fail(FM_ERR_DATA_BAD_HEADER, "data header has more than 4 values");
                uint64_t x = (uint64_t)header[nv] * 10 + (uint64_t)(ch - '0');
                if (x > 0xFFFFFFFFull) return fail(FM_ERR_DATA_BAD_HEADER, "header value overflows u32");
                header[nv] = (uint32_t)x;
                any = true;
            } else if (ch == ' ') {
                nv++;
            } else if (ch == '\n') {
                break;
            } else {
                return fail(FM_ERR_DATA_BAD_HEADER, "malformed data graph header");
            }
        }
        if (!any) return fail(FM_ERR_DATA_BAD_HEADER, "empty data graph header");
    }
    f.V = header[0];
    f.header_edges = header[1];
    f.body_begin = (size_t)(p - buf.data());
    f.body_end = (size_t)(stop - buf.data());
    return FM_OK;
}

int parse_data_line(const char *&p, const char *stop, uint32_t V, uint64_t line, char &kind, uint32_t vals[3]) {
    kind = *p;
    if ((kind != 'v' && kind != 'e') || p + 1 >= stop || p[1] != ' ')
        return fail(FM_ERR_DATA_BAD_BODY, "data graph line %llu is not a v/e line", (unsigned long long)line);
    p += 2;
    uint32_t a, b, c = 0;
    bool ok = parse_uint(p, stop, a) && p < stop && *p++ == ' ' && parse_uint(p, stop, b);
    if (ok && kind == 'e') ok = p < stop && *p++ == ' ' && parse_uint(p, stop, c);
    if (ok) ok = p < stop && *p++ == '\n';
    if (!ok) return fail(FM_ERR_DATA_BAD_BODY, "malformed data graph line %llu", (unsigned long long)line);
    if (kind == 'v' && a >= V)
        return fail(FM_ERR_DATA_BAD_BODY, "vertex id %u >= %u on line %llu", a, V, (unsigned long long)line);
    vals[0] = a;
    vals[1] = b;
    vals[2] = c;
    return FM_OK;
}

int parse_data_g(const std::string &buf, HostGraph &g) {
    DataFrame f;
    int rc = parse_data_frame(buf, f);
    if (rc) return rc;
    const uint32_t V = f.V;
    g.header_edges = f.header_edges;
    const char *p = buf.data() + f.body_begin, *stop = buf.data() + f.body_end;
    std::vector<uint32_t> src, dst, lab;
    src.reserve(f.header_edges);
    dst.reserve(f.header_edges);
    lab.reserve(f.header_edges);
    uint64_t line = 2;
    while (p < stop) {
        line++;
        char kind;
        uint32_t vals[3];
        if ((rc = parse_data_line(p, stop, V, line, kind, vals))) return rc;
        if (kind == 'e') {
            src.push_back(vals[0]);
            dst.push_back(vals[1]);
            lab.push_back(vals[2]);
        }
    }
    return build_graph(V, src.size(), src.data(), dst.data(), lab.data(), g);
}

// ---------------------------------------------------------------------------
// .g query parser: fileParse/parseValues (src/hostparser.cuh:190-266,
// 286-386). Vertex lines are counted only; edge lines become
// (src, dst, label) lines; fewer lines than the header states is
// ERR_MISSING_DATA (234), anything malformed ERR_INVALID_DATA (233).

int parse_query_g(const std::string &buf, HostQuery &q) {
    const char *p = buf.data(), *end = buf.data() + buf.size();
    if (buf.size() < 6 || std::memcmp(p, "t # 0\n", 6) != 0)
        return fail(FM_ERR_QUERY_INVALID, "query does not start with \"t # 0\\n\"");
    p += 6;
    auto read_line_values = [&](uint32_t *vals, int maxv, int &nvals) -> bool {
        nvals = 0;
        while (p < end) {
            uint32_t v;
            if (!parse_uint(p, end, v)) return false;
            if (nvals < maxv) vals[nvals] = v;
            nvals++;
            if (p >= end) return true;
            char ch = *p++;
            if (ch == '\n') return true;
            if (ch != ' ') return false;
        }
        return nvals > 0;
    };
    uint32_t hdr[4] = {0, 0, 0, 0};
    int nh = 0;
    if (!read_line_values(hdr, 4, nh) || nh < 2)
        return fail(FM_ERR_QUERY_INVALID, "malformed query header");
    uint32_t n = hdr[0];
    uint32_t hdr_edges = hdr[1];
    uint32_t vseen = 0;
    bool ended = false;
    std::vector<uint32_t> s, d, l;
    while (p < end) {
        if (*p == 't') {
            if (end - p < 6 || std::memcmp(p, "t # -1", 6) != 0)
                return fail(FM_ERR_QUERY_INVALID, "malformed query trailer");
            ended = true;
            break;
        }
        char kind = *p;
        if (kind == 'v') {
            if (!s.empty()) return fail(FM_ERR_QUERY_INVALID, "vertex line after edge lines");
            if (p + 1 >= end || p[1] != ' ') return fail(FM_ERR_QUERY_INVALID, "malformed query vertex");
            p += 2;
            uint32_t vals[2];
            int nvals;
            if (!read_line_values(vals, 2, nvals) || nvals < 1)
                return fail(FM_ERR_QUERY_INVALID, "malformed query vertex");
            vseen++;
        } else if (kind == 'e') {
            if (vseen < n) return fail(FM_ERR_QUERY_MISSING_DATA, "query lists %u of %u vertices", vseen, n);
            if (p + 1 >= end || p[1] != ' ') return fail(FM_ERR_QUERY_INVALID, "malformed query edge");
            p += 2;
            uint32_t vals[3];
            int nvals;
            if (!read_line_values(vals, 3, nvals) || nvals != 3)
                return fail(FM_ERR_QUERY_INVALID, "malformed query edge");
            s.push_back(vals[0]);
            d.push_back(vals[1]);
            l.push_back(vals[2]);
        } else {
            return fail(FM_ERR_QUERY_INVALID, "unexpected query line");
        }
    }
    if (!ended) return fail(FM_ERR_QUERY_INVALID, "query has no \"t # -1\" trailer");
    if (vseen < n) return fail(FM_ERR_QUERY_MISSING_DATA, "query lists %u of %u vertices", vseen, n);
    if (s.size() < hdr_edges)
        return fail(FM_ERR_QUERY_MISSING_DATA, "query lists %zu of %u edges", s.size(), hdr_edges);
    int rc = build_query(n, (uint32_t)s.size(), s.data(), d.data(), l.data(), q);
    q.header_edges = hdr_edges;
    return rc;
}

int build_query(uint32_t n, uint32_t m, const uint32_t *src, const uint32_t *dst,
                const uint32_t *label, HostQuery &q) {
    if (n == 0) return fail(FM_ERR_QUERY_INVALID, "query has no vertices");
    if (n > FM_MAX_QUERY_VERTICES)
        return fail(FM_ERR_QUERY_TOO_LARGE, "query has %u vertices; the limit is %d", n,
                    FM_MAX_QUERY_VERTICES);
    q.n = n;
    q.header_edges = m;
    q.edges.clear();
    for (uint32_t i = 0; i < m; i++) {
        if (src[i] >= n || dst[i] >= n)
            return fail(FM_ERR_QUERY_INVALID, "query edge %u names a vertex >= %u", i, n);
        std::array<uint32_t, 3> e{src[i], dst[i], label[i]};
        if (std::find(q.edges.begin(), q.edges.end(), e) == q.edges.end()) q.edges.push_back(e);
    }
    return make_plan(q, q.plan);
}

// Matching order following genereateCCSRRequirements/preProcessQuery
// (src/ccsr.cuh:229-326, 485-549): rank vertices by (out-degree desc, id
// asc) (src/hostparser.cuh:20-30,420); depth 0 is rank 0; each later depth
// takes the lowest-ranked unmapped vertex adjacent to the mapped set.
int make_plan(const HostQuery &q, Plan &p) {
    uint32_t n = q.n;
    std::vector<uint32_t> outdeg(n, 0);
    for (auto &e : q.edges) outdeg[e[0]]++;
    std::vector<uint32_t> rank_order(n);
    std::iota(rank_order.begin(), rank_order.end(), 0);
    std::stable_sort(rank_order.begin(), rank_order.end(), [&](uint32_t a, uint32_t b) {
        if (outdeg[a] != outdeg[b]) return outdeg[a] > outdeg[b];
        return a < b;
    });
    std::vector<uint32_t> rank(n);
    for (uint32_t i = 0; i < n; i++) rank[rank_order[i]] = i;
    std::vector<uint32_t> order(n, 0);
    std::vector<char> seen(n, 0), mapped(n, 0);
    order[0] = rank_order[0];
    mapped[rank_order[0]] = 1;
    auto discover = [&](uint32_t u) {
        for (auto &e : q.edges) {
            if (e[0] == u) seen[e[1]] = 1;
            if (e[1] == u) seen[e[0]] = 1;
        }
    };
    discover(rank_order[0]);
    for (uint32_t d = 1; d < n; d++) {
        uint32_t best = UINT32_MAX;
        for (uint32_t v = 0; v < n; v++)
            if (seen[v] && !mapped[v] && (best == UINT32_MAX || rank[v] < rank[best])) best = v;
        if (best == UINT32_MAX) return fail(FM_ERR_GENERIC, "query graph is disconnected");
        order[d] = best;
        mapped[best] = 1;
        discover(best);
    }
    return plan_from_order(q, order, p);
}

// Requirements of a depth are its edges to earlier vertices, ordered by
// climb (most recent anchor first) as the reference sorts them
// (src/ccsr.cuh:293-300).
int plan_from_order(const HostQuery &q, const std::vector<uint32_t> &order, Plan &p) {
    const uint32_t n = q.n;
    if (order.size() != n) return fail(FM_ERR_INVALID_ARG, "order has %zu entries for %u vertices", order.size(), n);
    p.n = n;
    p.order = order;
    p.depth_of.assign(n, UINT32_MAX);
    for (uint32_t d = 0; d < n; d++) {
        if (order[d] >= n || p.depth_of[order[d]] != UINT32_MAX)
            return fail(FM_ERR_INVALID_ARG, "order is not a permutation");
        p.depth_of[order[d]] = d;
    }
    p.req_begin.assign(n + 1, 0);
    p.reqs.clear();
    uint32_t maxr = 0;
    for (uint32_t d = 0; d < n; d++) {
        uint32_t u = p.order[d];
        std::vector<Req> rs;
        for (auto &e : q.edges) {
            uint32_t a = e[0], b = e[1], l = e[2];
            if (a == u && b == u) {
                rs.push_back({d, l, 0});
            } else if (b == u && p.depth_of[a] < d) {
                rs.push_back({p.depth_of[a], l, 0});  // line anchor -> u: u in out-row
            } else if (a == u && p.depth_of[b] < d) {
                rs.push_back({p.depth_of[b], l, 1});  // line u -> anchor: u in in-row
            }
        }
        std::sort(rs.begin(), rs.end(), [](const Req &x, const Req &y) {
            if (x.anchor != y.anchor) return x.anchor > y.anchor;
            if (x.label != y.label) return x.label < y.label;
            return x.dir < y.dir;
        });
        p.req_begin[d] = (uint32_t)p.reqs.size();
        for (auto &r : rs) p.reqs.push_back(r);
        maxr = std::max<uint32_t>(maxr, (uint32_t)rs.size());
        if (d > 0) {
            bool has_anchor = false;
            for (auto &r : rs) has_anchor |= r.anchor < d;
            if (!has_anchor) return fail(FM_ERR_GENERIC, "query graph is disconnected (depth %u has no anchor)", d);
        }
    }
    p.req_begin[n] = (uint32_t)p.reqs.size();
    if (maxr > FM_MAX_REQS_PER_DEPTH)
        return fail(FM_ERR_QUERY_TOO_LARGE, "a query vertex has %u backward edges; the limit is %d",
                    maxr, FM_MAX_REQS_PER_DEPTH);
    if (p.reqs.size() > FM_MAX_REQS_TOTAL)
        return fail(FM_ERR_QUERY_TOO_LARGE, "query has %zu requirements; the limit is %d",
                    p.reqs.size(), FM_MAX_REQS_TOTAL);
    return FM_OK;
}

// Cost-based matching order (no reference counterpart: the reference orders
// by query degree only, which on skewed graphs expands hub rows of common
// labels first). The embedding set does not depend on the order, only the
// work does. Model, per direction d and label l of the data graph: E = edges,
// R = vertices with a non-empty row, Q = sum of squared row lengths.
//  - the root and the depth-1 vertex are joined by query edges; the root
//    list is the smallest R among them and the depth-1 partials number E of
//    that edge (every edge of the label out of a root);
//  - a later vertex reached from a matched anchor through label l draws its
//    candidates from a row of a vertex that was itself reached along edges,
//    so its expected length is the size-biased Q / E;
//  - each further edge to the mapped set keeps a candidate with probability
//    (its expected row length) / V.
// Every root is tried; from it the order grows greedily by the smallest
// estimated number of partials. The order with the fewest examined
// candidates (sum over depths of partials times the driver row length) wins;
// ties keep the reference's order.
int make_cost_plan(const HostQuery &q, const HostGraph &g, Plan &p) {
    const uint32_t n = q.n;
    Plan ref;
    int rc = make_plan(q, ref);
    if (rc || n <= 2) {
        if (!rc) p = ref;
        return rc;
    }
    const double V = std::max<double>(1.0, g.V);
    struct QE {
        uint32_t a, b;
        double E[2], R[2], fan[2];  // [0]: a's out-row holds b; [1]: b's in-row holds a
    };
    std::vector<QE> qe;
    for (auto &e : q.edges) {
        QE x{e[0], e[1], {0, 0}, {0, 0}, {0, 0}};
        uint32_t lid = g.label_id(e[2]);
        if (lid != UINT32_MAX) {
            for (int d = 0; d < 2; d++) {
                const Csr &c = g.dir(d);
                x.E[d] = (double)c.ptr[(uint64_t)(lid + 1) * g.V] - (double)c.ptr[(uint64_t)lid * g.V];
                x.R[d] = (double)(c.root_ptr[lid + 1] - c.root_ptr[lid]);
                x.fan[d] = x.E[d] > 0 ? c.row_sq[lid] / x.E[d] : 0.0;
            }
        }
        qe.push_back(x);
    }
    // row of anchor w that holds candidate u along edge x: out-row of a if
    // w == a (u == b), in-row of b if w == b (u == a)
    auto dir_from = [](const QE &x, uint32_t w) { return x.a == w ? 0 : 1; };
    double best_cost = 0;
    std::vector<uint32_t> best_order;
    std::vector<char> mapped(n);
    for (uint32_t ri = 0; ri < n; ri++) {
        const uint32_t r = ref.order[ri];  // try roots in the reference's order
        std::fill(mapped.begin(), mapped.end(), 0);
        mapped[r] = 1;
        std::vector<uint32_t> order{r};
        double card = 0, cost = 0;
        for (uint32_t d = 1; d < n; d++) {
            double best_next = 0, best_drv = 0, best_card0 = 0;
            uint32_t best_u = UINT32_MAX, best_links = 0;
            for (uint32_t u = 0; u < n; u++) {
                if (mapped[u]) continue;
                double drv = -1, keep = 1, card0 = -1;
                uint32_t links = 0;
                for (auto &x : qe) {
                    uint32_t w;
                    if (x.a == u && x.b == u) {
                        keep *= std::min(1.0, (x.E[0] / std::max(1.0, x.R[0])) / V);
                        continue;
                    } else if (x.a == u && mapped[x.b]) {
                        w = x.b;
                    } else if (x.b == u && mapped[x.a]) {
                        w = x.a;
                    } else {
                        continue;
                    }
                    links++;
                    const int dd = dir_from(x, w);
                    // depth 1: every edge of the label out of a root
                    const double f = d == 1 ? x.E[dd] : x.fan[dd];
                    if (d == 1 && (card0 < 0 || x.R[dd] < card0)) card0 = x.R[dd];
                    if (drv < 0 || f < drv) {
                        if (drv >= 0) keep *= std::min(1.0, drv / (d == 1 ? V * V : V));
                        drv = f;
                    } else {
                        keep *= std::min(1.0, f / (d == 1 ? V * V : V));
                    }
                }
                if (!links) continue;
                const double next = d == 1 ? drv * keep : card * drv * keep;
                if (best_u == UINT32_MAX || next < best_next || (next == best_next && links > best_links)) {
                    best_u = u;
                    best_next = next;
                    best_drv = drv;
                    best_card0 = card0;
                    best_links = links;
                }
            }
            if (best_u == UINT32_MAX) return fail(FM_ERR_GENERIC, "query graph is disconnected");
            if (d == 1) {
                card = best_card0;
                cost = best_card0;
                // the root's rows of the label: E candidates in all
                cost += best_drv;
            } else {
                cost += card * best_drv;
            }
            card = best_next;
            cost += card;
            mapped[best_u] = 1;
            order.push_back(best_u);
        }
        if (best_order.empty() || cost < best_cost) {
            best_cost = cost;
            best_order = order;
        }
    }
    return plan_from_order(q, best_order, p);
}

// Automorphisms by backtracking over the vertices in id order: a vertex may
// only go to a vertex with the same (label, direction) incidence signature,
// and every line between assigned vertices must map onto a line. The
// constraints follow Grochow and Kellis ("Network motif discovery using
// subgraph enumeration and symmetry-breaking", RECOMB 2007): while the group
// G is not trivial, take the smallest vertex v with a non-trivial orbit,
// require m(v) < m(u) for every other u in the orbit, and replace G by the
// stabiliser of v. Every orbit of embeddings {m o p : p in Aut} then has
// exactly one member that satisfies all constraints.
void query_symmetry(const HostQuery &q, uint32_t max_auts, Symmetry &s) {
    const uint32_t n = q.n;
    s.auts.clear();
    s.less.clear();
    s.complete = true;
    std::unordered_set<uint64_t> lines;
    auto key = [](uint32_t u, uint32_t w, uint32_t l) { return ((uint64_t)l << 32) | ((uint64_t)u << 16) | w; };
    std::vector<std::vector<uint64_t>> sig(n);
    for (auto &e : q.edges) {
        lines.insert(key(e[0], e[1], e[2]));
        sig[e[0]].push_back(((uint64_t)e[2] << 2) | (e[0] == e[1] ? 2u : 0u));
        if (e[0] != e[1]) sig[e[1]].push_back(((uint64_t)e[2] << 2) | 1u);
    }
    for (auto &x : sig) std::sort(x.begin(), x.end());
    std::vector<uint32_t> p(n, UINT32_MAX);
    std::vector<char> used(n, 0);
    bool over = false;
    std::function<void(uint32_t)> rec = [&](uint32_t v) {
        if (over) return;
        if (v == n) {
            if (s.auts.size() >= max_auts) {
                over = true;
                return;
            }
            s.auts.push_back(p);
            return;
        }
        for (uint32_t x = 0; x < n && !over; x++) {
            if (used[x] || sig[x] != sig[v]) continue;
            p[v] = x;
            bool ok = true;
            for (auto &e : q.edges) {
                const uint32_t a = e[0], b = e[1];
                if (a > v || b > v || (a != v && b != v)) continue;  // both assigned, one of them v
                if (!lines.count(key(p[a], p[b], e[2]))) {
                    ok = false;
                    break;
                }
            }
            if (ok) {
                used[x] = 1;
                rec(v + 1);
                used[x] = 0;
            }
        }
        p[v] = UINT32_MAX;
    };
    rec(0);
    if (over) {
        s.auts.clear();
        s.complete = false;
        return;
    }
    // the identity first (the search meets it first already: vertex order)
    std::vector<uint32_t> group(s.auts.size());
    std::iota(group.begin(), group.end(), 0);
    while (group.size() > 1) {
        uint32_t v = UINT32_MAX;
        std::vector<char> orbit(n, 0);
        for (uint32_t u = 0; u < n && v == UINT32_MAX; u++) {
            std::fill(orbit.begin(), orbit.end(), 0);
            uint32_t size = 0;
            for (uint32_t k : group)
                if (!orbit[s.auts[k][u]]) {
                    orbit[s.auts[k][u]] = 1;
                    size++;
                }
            if (size > 1) v = u;
        }
        if (v == UINT32_MAX) break;
        for (uint32_t u = 0; u < n; u++)
            if (orbit[u] && u != v) s.less.push_back({v, u});
        std::vector<uint32_t> stab;
        for (uint32_t k : group)
            if (s.auts[k][v] == v) stab.push_back(k);
        group.swap(stab);
    }
}

// ---------------------------------------------------------------------------
// Writers and binary cache.

static int write_graph_g(const HostGraph &g, const std::string &path) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return fail(FM_ERR_INVALID_ARG, "cannot write %s", path.c_str());
    std::vector<char> buf;
    buf.reserve(1 << 22);
    auto flush = [&]() {
        fwrite(buf.data(), 1, buf.size(), f);
        buf.clear();
    };
    char line[96];
    uint32_t maxl = g.label_values.empty() ? 0 : g.label_values.back();
    int k = snprintf(line, sizeof line, "t # 0\n%u %llu 1 %u\n", g.V, (unsigned long long)g.E(), maxl);
    buf.insert(buf.end(), line, line + k);
    for (uint32_t v = 0; v < g.V; v++) {
        k = snprintf(line, sizeof line, "v %u 0\n", v);
        buf.insert(buf.end(), line, line + k);
        if (buf.size() > (1 << 22)) flush();
    }
    uint32_t V = g.V;
    for (uint32_t l = 0; l < g.L(); l++)
        for (uint32_t v = 0; v < V; v++) {
            uint64_t r = (uint64_t)l * V + v;
            for (uint32_t i = g.out.ptr[r]; i < g.out.ptr[r + 1]; i++) {
                k = snprintf(line, sizeof line, "e %u %u %u\n", v, g.out.col[i], g.label_values[l]);
                buf.insert(buf.end(), line, line + k);
            }
            if (buf.size() > (1 << 22)) flush();
        }
    const char *tail = "t # -1";
    buf.insert(buf.end(), tail, tail + 6);
    flush();
    fclose(f);
    return FM_OK;
}

static const char kBinMagic[8] = {'F', 'M', 'G', 'R', 'A', 'P', 'H', '1'};

static int save_bin(const HostGraph &g, const std::string &path) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return fail(FM_ERR_INVALID_ARG, "cannot write %s", path.c_str());
    uint64_t hdr[6] = {g.V, g.L(), g.E(), g.symmetric ? 1u : 0u, g.header_edges, g.out.roots.size()};
    fwrite(kBinMagic, 1, 8, f);
    fwrite(hdr, 8, 6, f);
    fwrite(g.label_values.data(), 4, g.L(), f);
    auto wcsr = [&](const Csr &c) {
        fwrite(c.ptr.data(), 4, c.ptr.size(), f);
        fwrite(c.col.data(), 4, c.col.size(), f);
        fwrite(c.root_ptr.data(), 4, c.root_ptr.size(), f);
        uint64_t nr = c.roots.size();
        fwrite(&nr, 8, 1, f);
        fwrite(c.roots.data(), 4, c.roots.size(), f);
    };
    wcsr(g.out);
    if (!g.symmetric) wcsr(g.in);
    bool ok = ferror(f) == 0;
    fclose(f);
    return ok ? FM_OK : fail(FM_ERR_INVALID_ARG, "short write to %s", path.c_str());
}

static int load_bin(const std::string &path, HostGraph &g) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return fail(FM_ERR_DATA_MISSING_FILE, "missing file %s", path.c_str());
    char magic[8];
    uint64_t hdr[6];
    bool ok = fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kBinMagic, 8) == 0 &&
              fread(hdr, 8, 6, f) == 6;
    if (!ok) {
        fclose(f);
        return fail(FM_ERR_DATA_BAD_START, "%s is not a ForkMatch binary graph", path.c_str());
    }
    g.V = (uint32_t)hdr[0];
    uint64_t L = hdr[1], E = hdr[2];
    g.symmetric = hdr[3] != 0;
    g.header_edges = (uint32_t)hdr[4];
    g.label_values.resize(L);
    ok = fread(g.label_values.data(), 4, L, f) == L;
    auto rcsr = [&](Csr &c) {
        c.ptr.resize(L * g.V + 1);
        c.col.resize(E);
        c.root_ptr.resize(L + 1);
        uint64_t nr = 0;
        bool r = fread(c.ptr.data(), 4, c.ptr.size(), f) == c.ptr.size() &&
                 fread(c.col.data(), 4, E, f) == E &&
                 fread(c.root_ptr.data(), 4, L + 1, f) == L + 1 && fread(&nr, 8, 1, f) == 1;
        if (r) {
            c.roots.resize(nr);
            r = fread(c.roots.data(), 4, nr, f) == nr;
        }
        return r;
    };
    ok = ok && rcsr(g.out);
    if (ok && !g.symmetric) ok = rcsr(g.in);
    fclose(f);
    if (!ok) return fail(FM_ERR_DATA_BAD_BODY, "truncated binary graph %s", path.c_str());
    csr_stats(g.out, g.V, (uint32_t)L);
    if (!g.symmetric) csr_stats(g.in, g.V, (uint32_t)L);
    return FM_OK;
}

// ---------------------------------------------------------------------------
// Generators.

struct LabelSampler {
    std::vector<double> cdf;
    int dist;
    uint32_t n;
    LabelSampler(uint32_t nlabels, int dist_, double s) : dist(dist_), n(std::max(1u, nlabels)) {
        if (dist == 1) {
            cdf.resize(n);
            double acc = 0;
            for (uint32_t k = 0; k < n; k++) {
                acc += std::pow((double)(k + 1), -s);
                cdf[k] = acc;
            }
            for (auto &c : cdf) c /= acc;
            cdf.back() = 1.0;
        }
    }
    uint32_t operator()(Rng &r) const {
        if (dist != 1) return (uint32_t)r.below(n);
        double u = r.uniform();
        return (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    }
};

// Open-addressing set of undirected edge keys (min << 32 | max).
struct EdgeSet {
    std::vector<uint64_t> slots;
    uint64_t mask, count = 0;
    explicit EdgeSet(uint64_t expected) {
        uint64_t cap = 1;
        while (cap < expected * 2 + 16) cap <<= 1;
        slots.assign(cap, ~0ull);
        mask = cap - 1;
    }
    bool insert(uint64_t key) {
        uint64_t h = key * 0x9E3779B97F4A7C15ull;
        uint64_t i = (h ^ (h >> 29)) & mask;
        while (true) {
            if (slots[i] == ~0ull) {
                slots[i] = key;
                count++;
                return true;
            }
            if (slots[i] == key) return false;
            i = (i + 1) & mask;
        }
    }
};

static int finish_generated(uint32_t V, const std::vector<uint64_t> &und, uint32_t nlabels, int dist,
                            double zipf_s, Rng &rng, HostGraph &g) {
    LabelSampler ls(nlabels, dist, zipf_s);
    uint64_t m = und.size();
    std::vector<uint32_t> s(2 * m), d(2 * m), l(2 * m);
    for (uint64_t i = 0; i < m; i++) {
        uint32_t a = (uint32_t)(und[i] >> 32), b = (uint32_t)und[i];
        uint32_t lab = ls(rng);
        s[2 * i] = a;
        d[2 * i] = b;
        l[2 * i] = lab;
        s[2 * i + 1] = b;
        d[2 * i + 1] = a;
        l[2 * i + 1] = lab;
    }
    int rc = build_graph(V, 2 * m, s.data(), d.data(), l.data(), g, true);
    g.header_edges = (uint32_t)std::min<uint64_t>(2 * m, UINT32_MAX);
    return rc;
}

int gen_er(uint32_t V, uint64_t m, uint32_t nlabels, int dist, double zs, uint64_t seed,
           HostGraph &g) {
    if (V < 2 || m > (uint64_t)V * (V - 1) / 2)
        return fail(FM_ERR_INVALID_ARG, "G(n=%u, m=%llu) is impossible", V, (unsigned long long)m);
    Rng rng(seed);
    EdgeSet set(m);
    std::vector<uint64_t> und;
    und.reserve(m);
    while (und.size() < m) {
        uint32_t a = (uint32_t)rng.below(V), b = (uint32_t)rng.below(V);
        if (a == b) continue;
        uint64_t key = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
        if (set.insert(key)) und.push_back(key);
    }
    return finish_generated(V, und, nlabels, dist, zs, rng, g);
}

// 64-bit finaliser of splitmix64: a keyed hash for seeds, ranks and labels.
static uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Open-addressing set of undirected edge keys that several threads insert
// into at once (compare-and-swap on empty slots).
struct SharedEdgeSet {
    std::vector<uint64_t> slots;
    uint64_t mask;
    explicit SharedEdgeSet(uint64_t expected) {
        uint64_t cap = 1;
        while (cap < expected * 2 + 16) cap <<= 1;
        slots.resize(cap);
        mask = cap - 1;
        parallel_for(cap, [&](uint64_t b, uint64_t e) {
            std::fill(slots.begin() + b, slots.begin() + e, ~0ull);
        });
    }
    bool insert(uint64_t key) {
        uint64_t i = mix64(key) & mask;
        while (true) {
            uint64_t cur = __atomic_load_n(&slots[i], __ATOMIC_RELAXED);
            if (cur == key) return false;
            if (cur == ~0ull) {
                uint64_t expected = ~0ull;
                if (__atomic_compare_exchange_n(&slots[i], &expected, key, false, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED))
                    return true;
                if (expected == key) return false;
            }
            i = (i + 1) & mask;
        }
    }
};

// R-MAT: each edge picks one of the four quadrants (a, b, c, d) per bit of
// the vertex ids; self-loops are dropped and duplicates collapsed. The
// candidate stream is cut into fixed chunks, each with its own seeded
// generator, and chunks run on all host threads. The edge set is the union
// of the candidates (rounds are added until it holds m edges); when it holds
// more, the m edges of smallest keyed hash are kept. Labels are a keyed
// function of the edge. So the graph depends on the seed only, not on the
// thread count or timing.
int gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c, uint32_t nlabels, int dist,
             double zs, uint64_t seed, HostGraph &g) {
    if (scale == 0 || scale > 31) return fail(FM_ERR_INVALID_ARG, "R-MAT scale %u out of range", scale);
    uint32_t V = 1u << scale;
    if (m > (uint64_t)V * (V - 1) / 4)
        return fail(FM_ERR_INVALID_ARG, "R-MAT scale %u cannot hold %llu edges", scale,
                    (unsigned long long)m);
    const double ab = a + b, abc = a + b + c;
    std::unique_ptr<SharedEdgeSet> set(new SharedEdgeSet(m + m / 8));
    std::atomic<uint64_t> uniq{0};
    for (uint64_t round = 0; uniq.load() < m; round++) {
        const uint64_t need = m - uniq.load();
        const uint64_t gen = round == 0 ? m + m / 8 : need + need / 2 + 1024;
        // chunk size from the round size only, so the stream does not
        // depend on the thread count
        const uint64_t CH = std::min<uint64_t>(1ull << 16, std::max<uint64_t>(1024, gen / 64));
        const uint64_t nchunks = (gen + CH - 1) / CH;
        // room for every candidate of the round at load <= 1/2
        const uint64_t bound = uniq.load() + nchunks * CH;
        if (2 * bound + 16 > set->slots.size()) {
            std::unique_ptr<SharedEdgeSet> bigger(new SharedEdgeSet(bound));
            for (uint64_t k : set->slots)
                if (k != ~0ull) bigger->insert(k);
            set.swap(bigger);
        }
        std::atomic<uint64_t> next{0};
        auto worker = [&]() {
            for (uint64_t k; (k = next.fetch_add(1)) < nchunks;) {
                Rng rng(mix64(seed ^ mix64((round << 40) ^ k)));
                uint64_t added = 0;
                for (uint64_t i = 0; i < CH; i++) {
                    uint32_t u = 0, v = 0;
                    for (uint32_t bit = 0; bit < scale; bit++) {
                        const double r = rng.uniform();
                        const uint32_t du = r >= ab, dv = (r >= a && r < ab) || r >= abc;
                        u = (u << 1) | du;
                        v = (v << 1) | dv;
                    }
                    if (u == v) continue;
                    added += set->insert(((uint64_t)std::min(u, v) << 32) | std::max(u, v));
                }
                uniq.fetch_add(added);
            }
        };
        const unsigned T = nchunks > 1 ? hw_threads() : 1;
        std::vector<std::thread> th;
        for (unsigned t = 1; t < T; t++) th.emplace_back(worker);
        worker();
        for (auto &x : th) x.join();
    }
    // the set's edges (slot order varies with timing; nothing below depends on it)
    std::vector<uint64_t> und;
    und.reserve(uniq.load());
    for (uint64_t k : set->slots)
        if (k != ~0ull) und.push_back(k);
    set.reset();
    if (und.size() > m) {
        const uint64_t key2 = mix64(seed ^ 0x52414E4Bull);
        auto rank_less = [&](uint64_t x, uint64_t y) {
            const uint64_t hx = mix64(x ^ key2), hy = mix64(y ^ key2);
            return hx != hy ? hx < hy : x < y;
        };
        std::nth_element(und.begin(), und.begin() + m, und.end(), rank_less);
        und.resize(m);
    }
    // labels: a keyed draw per undirected edge, written on both lines
    LabelSampler ls(nlabels, dist, zs);
    std::vector<uint32_t> s(2 * m), d(2 * m), l(2 * m);
    const uint64_t key3 = mix64(seed ^ 0x4C4142454Cull);
    parallel_for(m, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; i++) {
            const uint32_t x = (uint32_t)(und[i] >> 32), y = (uint32_t)und[i];
            Rng lr(mix64(und[i] ^ key3));
            const uint32_t lab = ls(lr);
            s[2 * i] = x;
            d[2 * i] = y;
            l[2 * i] = lab;
            s[2 * i + 1] = y;
            d[2 * i + 1] = x;
            l[2 * i + 1] = lab;
        }
    });
    std::vector<uint64_t>().swap(und);
    int rc = build_graph(V, 2 * m, s.data(), d.data(), l.data(), g, true);
    g.header_edges = (uint32_t)std::min<uint64_t>(2 * m, UINT32_MAX);
    return rc;
}

// Query extraction on the data graph (undirected view over all labels),
// seeded. Every query keeps the labels of the data edges it was cut from,
// so it has at least one embedding.
//  shape 0 (tree): a random walk; the edge that first reached each vertex.
//  shape 1 (cyclic): the vertex set grows from a random edge endpoint by
//    drawing the next vertex among the neighbours of the set with weight
//    (edges into the set)^3, which favours vertices that close cycles; the
//    query is the growth tree plus `min_cycles` (or one more) induced chords.
//  shape 2 (near-clique): the same growth; every induced edge is kept.
//  shape 3 (tree-like): the growth tree plus exactly `min_cycles` chords.
int gen_query_walk(const HostGraph &g, uint32_t k, int shape, uint32_t min_cycles, uint64_t seed,
                   HostQuery &q, uint32_t max_attempts) {
    if (k == 0 || k > FM_MAX_QUERY_VERTICES) return fail(FM_ERR_INVALID_ARG, "walk size %u", k);
    if (!g.symmetric) return fail(FM_ERR_NOT_SUPPORTED, "walk extraction needs a symmetric graph");
    if (g.E() == 0) return fail(FM_ERR_INVALID_ARG, "graph has no edges");
    uint32_t V = g.V, L = g.L();
    Rng rng(seed);
    // undirected neighbour list of v across labels: (nbr, label value)
    auto neighbours = [&](uint32_t v, std::vector<std::pair<uint32_t, uint32_t>> &out) {
        out.clear();
        for (uint32_t l = 0; l < L; l++) {
            uint64_t r = (uint64_t)l * V + v;
            for (uint32_t i = g.out.ptr[r]; i < g.out.ptr[r + 1]; i++)
                out.push_back({g.out.col[i], g.label_values[l]});
        }
    };
    // a random edge endpoint: find the row holding edge index e
    auto random_endpoint = [&]() -> uint32_t {
        uint64_t e = rng.below(g.E());
        uint64_t lo = 0, hi = (uint64_t)L * V;
        while (lo + 1 < hi) {
            uint64_t mid = (lo + hi) / 2;
            if (g.out.ptr[mid] <= e) lo = mid;
            else hi = mid;
        }
        return (uint32_t)(lo % V);
    };
    std::vector<std::pair<uint32_t, uint32_t>> nb;
    if (shape == 4) {
        if (k < 3) return fail(FM_ERR_INVALID_ARG, "a cycle needs at least 3 vertices, not %u", k);
        // a closing edge is found with probability about degree / V per walk
        max_attempts = std::max<uint32_t>(max_attempts, 256u * 1000u);
    }
    for (uint32_t attempt = 0; attempt < max_attempts; attempt++) {
        std::vector<uint32_t> verts{random_endpoint()};
        std::vector<std::array<uint32_t, 3>> tree;  // (local a, local b, label)
        if (shape == 4) {
            // a self-avoiding walk v0 .. v(k-1), kept when an edge joins
            // v(k-1) back to v0: a k-cycle whose edges keep their labels
            bool stuck = false;
            while (verts.size() < k && !stuck) {
                neighbours(verts.back(), nb);
                nb.erase(std::remove_if(nb.begin(), nb.end(),
                                        [&](const std::pair<uint32_t, uint32_t> &p) {
                                            return std::find(verts.begin(), verts.end(), p.first) != verts.end();
                                        }),
                         nb.end());
                if (nb.empty()) {
                    stuck = true;
                    break;
                }
                auto pick = nb[rng.below(nb.size())];
                tree.push_back({(uint32_t)verts.size() - 1, (uint32_t)verts.size(), pick.second});
                verts.push_back(pick.first);
            }
            if (stuck) continue;
            neighbours(verts.back(), nb);
            std::vector<uint32_t> closing;
            for (auto &p : nb)
                if (p.first == verts[0]) closing.push_back(p.second);
            if (closing.empty()) continue;
            tree.push_back({0, k - 1, closing[rng.below(closing.size())]});
        } else if (shape == 0) {
            uint32_t cur = 0, steps = 0;
            while (verts.size() < k && steps < 64 * k) {
                steps++;
                neighbours(verts[cur], nb);
                if (nb.empty()) break;
                auto pick = nb[rng.below(nb.size())];
                uint32_t idx = (uint32_t)(std::find(verts.begin(), verts.end(), pick.first) - verts.begin());
                if (idx == verts.size()) {
                    verts.push_back(pick.first);
                    tree.push_back({std::min(cur, idx), std::max(cur, idx), pick.second});
                }
                cur = idx;
            }
        } else {
            // weighted growth: weight of a candidate = its edges into the set;
            // updated incrementally as members join. Members with more than
            // kScan neighbours contribute a random sample of kScan of them.
            constexpr size_t kScan = 2048;
            std::unordered_map<uint32_t, uint32_t> where;  // vertex -> index in cand
            std::vector<std::array<uint32_t, 4>> cand;      // (vertex, weight, member idx, label)
            auto add_member = [&](uint32_t idx) {
                neighbours(verts[idx], nb);
                size_t take = nb.size();
                if (take > kScan) {
                    for (size_t i = 0; i < kScan; i++) std::swap(nb[i], nb[i + rng.below(nb.size() - i)]);
                    take = kScan;
                }
                for (size_t i = 0; i < take; i++) {
                    auto &p = nb[i];
                    auto it = where.find(p.first);
                    if (it == where.end()) {
                        where.emplace(p.first, (uint32_t)cand.size());
                        cand.push_back({p.first, 1, idx, p.second});
                    } else {
                        cand[it->second][1]++;
                    }
                }
            };
            add_member(0);
            while (verts.size() < k) {
                double total = 0;
                for (auto &c : cand)
                    if (c[1]) total += (double)c[1] * c[1] * c[1];
                if (total == 0) break;
                double u = rng.uniform() * total;
                size_t j = 0, last = 0;
                for (; j < cand.size(); j++) {
                    if (!cand[j][1]) continue;
                    last = j;
                    u -= (double)cand[j][1] * cand[j][1] * cand[j][1];
                    if (u < 0) break;
                }
                if (j == cand.size()) j = last;
                uint32_t idx = (uint32_t)verts.size();
                verts.push_back(cand[j][0]);
                tree.push_back({cand[j][2], idx, cand[j][3]});
                cand[j][1] = 0;  // a member is no longer a candidate
                add_member(idx);
                // neighbours already in the set must not be drawn again
                for (uint32_t m = 0; m <= idx; m++) {
                    auto it = where.find(verts[m]);
                    if (it != where.end()) cand[it->second][1] = 0;
                }
            }
        }
        if (verts.size() < k) continue;
        std::vector<std::array<uint32_t, 3>> und = tree;
        if (shape != 0 && shape != 4) {
            std::vector<std::array<uint32_t, 3>> chords;
            for (uint32_t i = 0; i < k; i++)
                for (uint32_t l = 0; l < L; l++) {
                    uint64_t r = (uint64_t)l * V + verts[i];
                    const uint32_t *b = g.out.col.data() + g.out.ptr[r];
                    const uint32_t *e = g.out.col.data() + g.out.ptr[r + 1];
                    for (uint32_t j = i + 1; j < k; j++)
                        if (std::binary_search(b, e, verts[j])) {
                            std::array<uint32_t, 3> c{i, j, g.label_values[l]};
                            if (std::find(tree.begin(), tree.end(), c) == tree.end()) chords.push_back(c);
                        }
                }
            if (chords.size() < std::max<uint32_t>(min_cycles, shape == 2 ? 0u : 1u)) continue;
            if (shape == 1 || shape == 3) {
                // keep min_cycles chords (at least one); shape 1 sometimes
                // keeps one more, shape 3 exactly min_cycles
                uint32_t want = std::max<uint32_t>(min_cycles, 1);
                if (shape == 1 && chords.size() > want && (rng.next() & 1)) want++;
                for (uint32_t c = 0; c < want; c++) {
                    uint64_t pick = c + rng.below(chords.size() - c);
                    std::swap(chords[c], chords[pick]);
                    und.push_back(chords[c]);
                }
            } else {
                for (auto &c : chords) und.push_back(c);
            }
        }
        std::vector<uint32_t> s, d, l;
        for (auto &e : und) {
            s.push_back(e[0]);
            d.push_back(e[1]);
            l.push_back(e[2]);
            s.push_back(e[1]);
            d.push_back(e[0]);
            l.push_back(e[2]);
        }
        return build_query(k, (uint32_t)s.size(), s.data(), d.data(), l.data(), q);
    }
    return fail(FM_ERR_GENERIC, "no query of %u vertices (shape %d, %u cycles) after %u attempts", k, shape,
                min_cycles, max_attempts);
}

static int write_query_g(const HostQuery &q, const std::string &path) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return fail(FM_ERR_INVALID_ARG, "cannot write %s", path.c_str());
    uint32_t maxl = 0;
    for (auto &e : q.edges) maxl = std::max(maxl, e[2]);
    fprintf(f, "t # 0\n%u %zu 1 %u\n", q.n, q.edges.size(), maxl);
    for (uint32_t v = 0; v < q.n; v++) fprintf(f, "v %u 0\n", v);
    for (auto &e : q.edges) fprintf(f, "e %u %u %u\n", e[0], e[1], e[2]);
    fprintf(f, "t # -1\n");
    fclose(f);
    return FM_OK;
}

}  // namespace fm

// ---------------------------------------------------------------------------
// C ABI: host-only entry points.
using namespace fm;

extern "C" {

const char *fm_last_error(void) { return g_last_error.c_str(); }

const char *fm_version(void) { return "forkmatch-mi355x 0.1"; }

int fm_graph_parse_g(const char *path, fm_graph **out) {
    if (!path || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    std::string buf;
    if (!read_file(path, buf)) return fail(FM_ERR_DATA_MISSING_FILE, "missing file %s", path);
    auto *g = new fm_graph();
    int rc = parse_data_g(buf, g->h);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return FM_OK;
}

int fm_graph_from_edges(uint32_t vertices, uint64_t nedges, const uint32_t *src, const uint32_t *dst,
                        const uint32_t *label, fm_graph **out) {
    if (!out || (nedges && (!src || !dst || !label))) return fail(FM_ERR_INVALID_ARG, "null argument");
    auto *g = new fm_graph();
    int rc = build_graph(vertices, nedges, src, dst, label, g->h);
    if (rc) {
        delete g;
        return rc == FM_ERR_DATA_BAD_BODY ? FM_ERR_INVALID_ARG : rc;
    }
    g->h.header_edges = (uint32_t)std::min<uint64_t>(nedges, UINT32_MAX);
    *out = g;
    return FM_OK;
}

int fm_graph_info_get(const fm_graph *g, fm_graph_info *out) {
    if (!g || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    out->vertices = g->h.V;
    out->labels = g->h.L();
    out->edges = g->h.E();
    out->symmetric = g->h.symmetric ? 1 : 0;
    out->header_edges = g->h.header_edges;
    out->device_bytes = g->device_bytes;
    return FM_OK;
}

int fm_graph_copy_csr(const fm_graph *g, int dir, uint32_t *ptr, uint32_t *col, uint32_t *label_values) {
    if (!g || (dir != 0 && dir != 1)) return fail(FM_ERR_INVALID_ARG, "bad argument");
    const Csr &c = g->h.dir(dir);
    if (ptr) std::memcpy(ptr, c.ptr.data(), c.ptr.size() * 4);
    if (col) std::memcpy(col, c.col.data(), c.col.size() * 4);
    if (label_values) std::memcpy(label_values, g->h.label_values.data(), g->h.L() * 4);
    return FM_OK;
}

int fm_graph_save_bin(const fm_graph *g, const char *path) {
    if (!g || !path) return fail(FM_ERR_INVALID_ARG, "null argument");
    return save_bin(g->h, path);
}

int fm_graph_load_bin(const char *path, fm_graph **out) {
    if (!path || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    auto *g = new fm_graph();
    int rc = load_bin(path, g->h);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return FM_OK;
}

int fm_graph_write_g(const fm_graph *g, const char *path) {
    if (!g || !path) return fail(FM_ERR_INVALID_ARG, "null argument");
    return write_graph_g(g->h, path);
}

int fm_gen_erdos_renyi(uint32_t vertices, uint64_t undirected_edges, uint32_t nlabels, int label_dist,
                       double zipf_s, uint32_t vlabels, uint64_t seed, fm_graph **out) {
    (void)vlabels;  // vertex labels are ignored by the matcher and not stored
    if (!out) return fail(FM_ERR_INVALID_ARG, "null argument");
    auto *g = new fm_graph();
    int rc = gen_er(vertices, undirected_edges, nlabels, label_dist, zipf_s, seed, g->h);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return FM_OK;
}

int fm_gen_rmat(uint32_t scale, uint64_t undirected_edges, double a, double b, double c, uint32_t nlabels,
                int label_dist, double zipf_s, uint32_t vlabels, uint64_t seed, fm_graph **out) {
    (void)vlabels;
    if (!out) return fail(FM_ERR_INVALID_ARG, "null argument");
    auto *g = new fm_graph();
    int rc = gen_rmat(scale, undirected_edges, a, b, c, nlabels, label_dist, zipf_s, seed, g->h);
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return FM_OK;
}

int fm_gen_query_walk(const fm_graph *g, uint32_t vertices, int shape, uint32_t min_cycles, uint64_t seed,
                      fm_query **out) {
    if (!g || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    auto *q = new fm_query();
    int rc = gen_query_walk(g->h, vertices, shape, min_cycles, seed, q->h, 1000);
    if (rc) {
        delete q;
        return rc;
    }
    *out = q;
    return FM_OK;
}

int fm_query_set_count_only(fm_query *q, uint32_t count_only) {
    if (!q) return fail(FM_ERR_INVALID_ARG, "null argument");
    q->count_only = count_only ? 1u : 0u;
    return FM_OK;
}

int fm_query_load_g(const char *path, fm_query **out) {
    if (!path || !out) return fail(FM_ERR_INVALID_ARG, "null argument");
    std::string buf;
    if (!read_file(path, buf)) return fail(FM_ERR_QUERY_MISSING_FILE, "missing file %s", path);
    auto *q = new fm_query();
    int rc = parse_query_g(buf, q->h);
    if (rc) {
        delete q;
        return rc;
    }
    *out = q;
    return FM_OK;
}

int fm_query_from_edges(uint32_t vertices, uint32_t nedges, const uint32_t *src, const uint32_t *dst,
                        const uint32_t *label, fm_query **out) {
    if (!out || (nedges && (!src || !dst || !label))) return fail(FM_ERR_INVALID_ARG, "null argument");
    auto *q = new fm_query();
    int rc = build_query(vertices, nedges, src, dst, label, q->h);
    if (rc) {
        delete q;
        return rc;
    }
    *out = q;
    return FM_OK;
}

int fm_query_write_g(const fm_query *q, const char *path) {
    if (!q || !path) return fail(FM_ERR_INVALID_ARG, "null argument");
    return write_query_g(q->h, path);
}

int fm_query_info(const fm_query *q, uint32_t *vertices, uint32_t *nedges) {
    if (!q) return fail(FM_ERR_INVALID_ARG, "null argument");
    if (vertices) *vertices = q->h.n;
    if (nedges) *nedges = (uint32_t)q->h.edges.size();
    return FM_OK;
}

int fm_query_edges(const fm_query *q, uint32_t *src, uint32_t *dst, uint32_t *label) {
    if (!q) return fail(FM_ERR_INVALID_ARG, "null argument");
    for (size_t i = 0; i < q->h.edges.size(); i++) {
        if (src) src[i] = q->h.edges[i][0];
        if (dst) dst[i] = q->h.edges[i][1];
        if (label) label[i] = q->h.edges[i][2];
    }
    return FM_OK;
}

static void export_plan(const Plan &p, fm_plan_info *info, uint32_t *order, uint32_t *req_begin,
                        uint32_t *req_anchor, uint32_t *req_label, uint32_t *req_dir) {
    if (info) {
        info->vertices = p.n;
        info->nreqs = (uint32_t)p.reqs.size();
        uint32_t m = 0;
        for (uint32_t d = 0; d < p.n; d++) m = std::max(m, p.req_begin[d + 1] - p.req_begin[d]);
        info->max_reqs_per_depth = m;
    }
    for (uint32_t d = 0; d < p.n; d++) {
        if (order) order[d] = p.order[d];
        if (req_begin) req_begin[d] = p.req_begin[d];
    }
    if (req_begin) req_begin[p.n] = p.req_begin[p.n];
    for (size_t i = 0; i < p.reqs.size(); i++) {
        if (req_anchor) req_anchor[i] = p.reqs[i].anchor;
        if (req_label) req_label[i] = p.reqs[i].label;
        if (req_dir) req_dir[i] = p.reqs[i].dir;
    }
}

int fm_query_plan(const fm_query *q, fm_plan_info *info, uint32_t *order, uint32_t *req_begin,
                  uint32_t *req_anchor, uint32_t *req_label, uint32_t *req_dir) {
    if (!q) return fail(FM_ERR_INVALID_ARG, "null argument");
    export_plan(q->h.plan, info, order, req_begin, req_anchor, req_label, req_dir);
    return FM_OK;
}

int fm_query_plan_on(const fm_graph *g, const fm_query *q, uint32_t plan, fm_plan_info *info, uint32_t *order,
                     uint32_t *req_begin, uint32_t *req_anchor, uint32_t *req_label, uint32_t *req_dir) {
    if (!g || !q) return fail(FM_ERR_INVALID_ARG, "null argument");
    if (plan == FM_PLAN_REFERENCE) {
        export_plan(q->h.plan, info, order, req_begin, req_anchor, req_label, req_dir);
        return FM_OK;
    }
    if (plan != FM_PLAN_COST) return fail(FM_ERR_INVALID_ARG, "unknown plan %u", plan);
    Plan p;
    int rc = make_cost_plan(q->h, g->h, p);
    if (rc) return rc;
    export_plan(p, info, order, req_begin, req_anchor, req_label, req_dir);
    return FM_OK;
}

int fm_query_symmetry(const fm_query *q, uint32_t max_auts, uint32_t *n_aut, uint32_t *auts, uint32_t aut_cap,
                      uint32_t *n_less, uint32_t *less, uint32_t less_cap) {
    if (!q || !n_aut || !n_less) return fail(FM_ERR_INVALID_ARG, "null argument");
    Symmetry sym;
    query_symmetry(q->h, max_auts, sym);
    const uint32_t n = q->h.n;
    *n_aut = (uint32_t)sym.auts.size();
    *n_less = (uint32_t)sym.less.size();
    if (auts) {
        if ((uint64_t)aut_cap < (uint64_t)*n_aut * n) return fail(FM_ERR_INVALID_ARG, "auts holds %u entries", aut_cap);
        for (size_t k = 0; k < sym.auts.size(); k++)
            for (uint32_t u = 0; u < n; u++) auts[k * n + u] = sym.auts[k][u];
    }
    if (less) {
        if (less_cap < 2 * *n_less) return fail(FM_ERR_INVALID_ARG, "less holds %u entries", less_cap);
        for (size_t i = 0; i < sym.less.size(); i++) {
            less[2 * i] = sym.less[i].first;
            less[2 * i + 1] = sym.less[i].second;
        }
    }
    return FM_OK;
}

void fm_query_destroy(fm_query *q) { delete q; }

}  // extern "C"
