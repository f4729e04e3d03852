// Internal declarations shared by the host-side C++ (graph/query/plan/io)
// and the HIP engine. Nothing here crosses the C ABI.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "environment.h"
#include "forkmatch.h"

namespace fm {

// Records a failure for fm_last_error() and returns `code`.
int fail(int code, const char *fmt, ...);
void clear_error();

// Membership probes read fixed windows past their row: the device column
// array and its skip index are allocated (and zeroed) with this many words
// of padding.
constexpr uint32_t COL_PAD = 64;
// Skip index: skip[b] = col[b * 16], the first word of every 16-word block of
// the column array (1/16 of its size).
constexpr uint32_t SKIP_SHIFT = 4;

// One label-partitioned adjacency direction. Row (l, v) is
// col[ptr[l*V + v] .. ptr[l*V + v + 1]), destinations ascending. roots holds,
// per label, the ascending list of vertices whose row is non-empty
// (roots[root_ptr[l] .. root_ptr[l+1])): the candidate roots of a query
// whose first requirement uses label l.
struct Csr {
    std::vector<uint32_t> ptr;
    std::vector<uint32_t> col;
    std::vector<uint32_t> root_ptr;
    std::vector<uint32_t> roots;
    // per label: sum over vertices of the squared row length (the planner's
    // size-biased row length is row_sq[l] / edges of label l)
    std::vector<double> row_sq;
};

// Fills c.row_sq from c.ptr.
void csr_stats(Csr &c, uint32_t V, uint32_t L);

struct HostGraph {
    uint32_t V = 0;
    uint32_t header_edges = 0;
    std::vector<uint32_t> label_values;  // dense label id -> label value
    Csr out;                             // rows by source
    Csr in;                              // rows by destination (empty if symmetric)
    bool symmetric = false;
    uint32_t L() const { return (uint32_t)label_values.size(); }
    uint64_t E() const { return out.col.size(); }
    // dense id of a label value, or UINT32_MAX when the graph has no edge
    // with that label
    uint32_t label_id(uint32_t value) const;
    const Csr &dir(int d) const { return (d == 0 || symmetric) ? out : in; }
};

// Builds a HostGraph from directed (src, dst, label) lines. Duplicate lines
// are collapsed. Returns FM_OK or a status. A caller that wrote every line
// in both directions (the generators) skips the symmetry check.
// The .g data file's frame (trailing "t # -1", header): the body is
// buf[body_begin, body_end), lines of "v id label\n" / "e src dst label\n".
struct DataFrame {
    uint32_t V = 0, header_edges = 0;
    size_t body_begin = 0, body_end = 0;
};
int parse_data_frame(const std::string &buf, DataFrame &f);
// One body line starting at p (advanced past its newline); `line` is its
// 1-based line number in the file, for the error message.
int parse_data_line(const char *&p, const char *stop, uint32_t V, uint64_t line, char &kind, uint32_t vals[3]);

int parse_data_g(const std::string &buf, HostGraph &g);

int build_graph(uint32_t V, uint64_t n, const uint32_t *src, const uint32_t *dst,
                const uint32_t *label, HostGraph &g,
                bool symmetric_by_construction = false);

struct Req {
    uint32_t anchor;  // depth of the earlier query vertex (== own depth: self-loop)
    uint32_t label;   // label value
    uint32_t dir;     // 0: candidate in anchor's out-row, 1: in anchor's in-row
};

struct Plan {
    uint32_t n = 0;
    std::vector<uint32_t> order;      // depth -> query vertex
    std::vector<uint32_t> depth_of;   // query vertex -> depth
    std::vector<uint32_t> req_begin;  // n + 1
    std::vector<Req> reqs;
};

struct HostQuery {
    uint32_t n = 0;
    uint32_t header_edges = 0;
    std::vector<std::array<uint32_t, 3>> edges;  // unique (src, dst, label) lines
    Plan plan;
};

int build_query(uint32_t n, uint32_t m, const uint32_t *src, const uint32_t *dst,
                const uint32_t *label, HostQuery &q);
int make_plan(const HostQuery &q, Plan &p);
// Requirements of each depth for a given matching order (order[0] is the
// root; every later vertex must be adjacent to an earlier one).
int plan_from_order(const HostQuery &q, const std::vector<uint32_t> &order, Plan &p);
// Matching order chosen from the data graph's label statistics: the order
// whose estimated number of examined candidates is smallest.
int make_cost_plan(const HostQuery &q, const HostGraph &g, Plan &p);

// Symmetry of a query: its automorphisms (permutations p of the query
// vertices with (p(u), p(w), l) a line for every line (u, w, l)) and
// ordering constraints m(a) < m(b) that keep exactly one embedding of every
// orbit of the automorphism group acting on embeddings.
struct Symmetry {
    std::vector<std::vector<uint32_t>> auts;              // auts[k][u] = p_k(u); auts[0] = identity
    std::vector<std::pair<uint32_t, uint32_t>> less;      // (a, b): m(a) < m(b)
    bool complete = true;                                 // false: more than the cap, none listed
};
void query_symmetry(const HostQuery &q, uint32_t max_auts, Symmetry &s);

// Reads a whole file. Returns false if it cannot be opened.
bool read_file(const std::string &path, std::string &buf);

// Deterministic generator: xoshiro256** seeded through splitmix64.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed);
    uint64_t next();
    double uniform();                     // [0, 1)
    uint64_t below(uint64_t n);           // [0, n)
};

}  // namespace fm

// Opaque handle bodies (declared in forkmatch.h).
struct fm_graph {
    fm::HostGraph h;
    int device = -1;  // device holding the upload, -1 if host only
    // device CSR (dir 1 aliases dir 0 when the graph is symmetric)
    uint32_t *d_ptr[2] = {nullptr, nullptr};
    uint32_t *d_col[2] = {nullptr, nullptr};
    uint32_t *d_skip = nullptr;  // first word of every 16-word block of the column array
    uint32_t *d_root_ptr[2] = {nullptr, nullptr};
    uint32_t *d_roots[2] = {nullptr, nullptr};
    uint64_t device_bytes = 0;
};

struct fm_query {
    fm::HostQuery h;
    uint32_t count_only = 0;  // counted without rows in a table-mode launch
};

namespace fm {
// Builds g's out-rows on the current device from the device edge arrays
// d_edges = src[n] | dst[n] | label[n], with build_graph's checks and
// messages, as if every line had its reverse (the caller checks the
// symmetry), and keeps them as g's upload (rows, column array and skip
// index with their padding, root lists) besides the host copy. Leaves
// built = false, with nothing done, for a graph without edges, a label value
// of 2^24 or more, or (label, vertex, vertex) keys wider than 64 bits: the
// host build takes those.
int gpu_build_graph(uint32_t V, uint64_t n, const uint32_t *d_edges, fm_graph *g, bool &built);
}  // namespace fm
