"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

- closed forms on single-label undirected graphs for the reference's query
  fixtures (degrees d, adjacency A);
- brute force over all injective maps on tiny graphs (independent
  enumeration with itertools), including directed, multi-label and
  self-loop graphs;
- hand-derived matching orders of the eight reference fixtures
  (src/ccsr.cuh:229-249 rule: rank by out-degree desc / id asc, root = rank
  0, then the lowest-ranked frontier vertex);
- the algorithmic-byte counters against a brute-force enumeration of the
  partial embeddings of every depth;
- table invariants (distinct valid rows, automorphism divisibility,
  relabelling equivariance).
"""
import itertools

import numpy as np
import pytest

import forkmatch_amd as fm
import oracle as orc
from conftest import FIXTURE_NAMES, fixture_path


def single_label_graph(V, m, seed):
    g = fm.Graph.erdos_renyi(V, m, 1, seed=seed)
    s, d, _ = g.edge_list()
    return s, d, np.ones_like(s)


def fixture_edges(name, label=None):
    q = fm.Query.load_g(str(fixture_path(name)))
    e = q.edges().copy()
    if label is not None:
        e[:, 2] = label
    return q.vertices, e


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_closed_forms(seed):
    V = 150
    s, d, l = single_label_graph(V, 700, seed)
    og = orc.OracleGraph(V, s, d, l)
    A = np.zeros((V, V), np.int64)
    A[s, d] = 1
    deg = A.sum(1)
    A3 = A @ A @ A
    A4 = A3 @ A

    def cnt(name):
        n, e = fixture_edges(name, 1)
        return orc.match(og, n, e, table=False).count

    assert cnt("_2_road2") == deg.sum()
    assert cnt("_3_road3") == (deg * (deg - 1)).sum()
    assert cnt("_3_triangle") == np.trace(A3)
    assert cnt("_4_square") == np.trace(A4) - 2 * (deg ** 2).sum() + deg.sum()
    assert cnt("_4_roadY") == (deg * (deg - 1) * (deg - 2)).sum()
    assert cnt("_5_plus") == (deg * (deg - 1) * (deg - 2) * (deg - 3)).sum()
    p4 = sum((deg[u] - 1) * (deg[v] - 1) for u, v in zip(s, d)) - np.trace(A3)
    assert cnt("_4_road4") == p4


def brute_force(V, s, d, l, n, edges):
    """All injective maps V_q -> V_d such that each query line maps onto a data
    edge with its label, by exhaustive enumeration."""
    have = set(zip(s.tolist(), d.tolist(), l.tolist()))
    rows = []
    for m in itertools.permutations(range(V), n):
        if all((m[a], m[b], c) in have for a, b, c in edges):
            rows.append(m)
    return np.array(sorted(rows), dtype=np.uint32).reshape(-1, n)


def tiny_graphs():
    rng = np.random.default_rng(42)
    out = []
    for t in range(6):
        V = 7
        m = 14
        s = rng.integers(0, V, m)
        d = rng.integers(0, V, m)
        l = rng.integers(0, 2, m)
        if t % 2 == 0:  # symmetric
            s, d, l = np.concatenate([s, d]), np.concatenate([d, s]), np.concatenate([l, l])
        if t == 4:  # self loops
            s, d, l = np.concatenate([s, [1, 3]]), np.concatenate([d, [1, 3]]), np.concatenate([l, [0, 1]])
        keep = (s != d) | (t == 4)
        out.append((V, s[keep].astype(np.uint32), d[keep].astype(np.uint32), l[keep].astype(np.uint32)))
    return out


QUERIES = [
    (2, [(0, 1, 0)]),
    (2, [(0, 1, 0), (1, 0, 0)]),
    (3, [(0, 1, 0), (1, 2, 1)]),
    (3, [(0, 1, 0), (1, 2, 0), (2, 0, 1)]),
    (3, [(0, 1, 0), (1, 0, 0), (1, 2, 1), (2, 1, 1), (0, 2, 0), (2, 0, 0)]),
    (4, [(0, 1, 0), (1, 2, 0), (2, 3, 0), (3, 0, 0)]),
    (4, [(1, 0, 1), (2, 0, 0), (3, 0, 1)]),
    (2, [(0, 1, 0), (1, 1, 1)]),
    (1, [(0, 0, 0)]),
    (1, []),
]


@pytest.mark.parametrize("gi", range(6))
def test_brute_force(gi):
    V, s, d, l = tiny_graphs()[gi]
    og = orc.OracleGraph(V, s, d, l)
    for n, e in QUERIES:
        ref = brute_force(V, s, d, l, n, e)
        got = orc.match(og, n, np.array(e, np.uint32).reshape(-1, 3), threads=2)
        assert got.count == len(ref), (gi, n, e)
        assert np.array_equal(got.table, ref), (gi, n, e)


# order[d] = query vertex at depth d; reqs[d] = sorted set of anchor depths
HAND_PLANS = {
    "_2_road2": ([0, 1], [[], [0]]),
    "_3_road3": ([1, 0, 2], [[], [0], [0]]),
    "_3_triangle": ([0, 1, 2], [[], [0], [0, 1]]),
    "_4_road4": ([1, 2, 0, 3], [[], [0], [0], [1]]),
    "_4_roadY": ([0, 1, 2, 3], [[], [0], [0], [0]]),
    "_4_square": ([0, 1, 2, 3], [[], [0], [1], [0, 2]]),
    "_5_plus": ([1, 0, 2, 3, 4], [[], [0], [0], [0], [0]]),
    "_5_road5": ([1, 2, 3, 0, 4], [[], [0], [1], [0], [2]]),
}


@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_plans_match_hand_derivation(name):
    n, e = fixture_edges(name)
    order, rb, anchor, label, direction = orc.plan(n, e)
    want_order, want_anchors = HAND_PLANS[name]
    assert order == want_order
    for dpt in range(n):
        got = sorted(set(anchor[rb[dpt]:rb[dpt + 1]]))
        assert got == want_anchors[dpt], (name, dpt)
        # paired lines give one requirement per direction
        assert rb[dpt + 1] - rb[dpt] == 2 * len(want_anchors[dpt])
        # climb order: most recent anchor first (src/ccsr.cuh:293-300)
        assert list(anchor[rb[dpt]:rb[dpt + 1]]) == sorted(anchor[rb[dpt]:rb[dpt + 1]], reverse=True)


def test_disconnected_query_is_rejected():
    with pytest.raises(ValueError):
        orc.plan(4, np.array([(0, 1, 0), (1, 0, 0), (2, 3, 0), (3, 2, 0)], np.uint32))


def partials_brute(V, s, d, l, n, edges, order):
    """Partial embeddings of every depth of the plan, by brute force."""
    have = set(zip(s.tolist(), d.tolist(), l.tolist()))
    pos = {v: i for i, v in enumerate(order)}
    out = []
    for k in range(1, n + 1):
        sub = [(a, b, c) for a, b, c in edges if pos[a] < k and pos[b] < k]
        cnt = 0
        for m in itertools.permutations(range(V), k):
            if all((m[pos[a]], m[pos[b]], c) in have for a, b, c in sub):
                cnt += 1
        out.append(cnt)
    return out


def test_instrumented_counters_brute_force():
    V, s, d, l = tiny_graphs()[0]
    og = orc.OracleGraph(V, s, d, l)
    adj = {}
    for a, b, c in zip(s.tolist(), d.tolist(), l.tolist()):
        adj.setdefault((c, a), set()).add(b)
    have = set(zip(s.tolist(), d.tolist(), l.tolist()))
    for n, e in QUERIES[2:7]:
        e = np.array(e, np.uint32).reshape(-1, 3)
        order, rb, anchor, label, direction = orc.plan(n, e)
        got = orc.match(og, n, e, table=False, instrument=True)
        assert got.partials == partials_brute(V, s, d, l, n, e.tolist(), order)
        # neighbour words: every partial of depth k-1 reads the row of each
        # requirement of depth k (one per (anchor, label) on a symmetric graph)
        pos = {v: i for i, v in enumerate(order)}
        words = 0
        for k in range(1, n):
            reqs = sorted({(anchor[i], label[i]) for i in range(rb[k], rb[k + 1]) if anchor[i] != k})
            sub = [(a, b, c) for a, b, c in e.tolist() if pos[a] < k and pos[b] < k]
            for m in itertools.permutations(range(V), k):
                if all((m[pos[a]], m[pos[b]], c) in have for a, b, c in sub):
                    words += sum(len(adj.get((lab, m[a]), ())) for a, lab in reqs)
        assert got.row_words == words


def automorphisms(n, edges):
    es = set(map(tuple, edges))
    return sum(1 for p in itertools.permutations(range(n)) if {(p[a], p[b], c) for a, b, c in es} == es)


def test_table_invariants():
    g = fm.Graph.erdos_renyi(60, 240, 2, seed=5)
    s, d, l = g.edge_list()
    og = orc.OracleGraph(60, s, d, l)
    have = set(zip(s.tolist(), d.tolist(), l.tolist()))
    for name in FIXTURE_NAMES:
        n, e = fixture_edges(name, 0)
        r = orc.match(og, n, e)
        assert len(r.table) == r.count
        assert len({tuple(x) for x in r.table.tolist()}) == r.count
        for row in r.table.tolist():
            assert len(set(row)) == n
            assert all((row[a], row[b], c) in have for a, b, c in e.tolist())
        assert r.count % automorphisms(n, e.tolist()) == 0
        # relabelling the data vertices maps the table bijectively
        perm = np.random.default_rng(9).permutation(60).astype(np.uint32)
        og2 = orc.OracleGraph(60, perm[s], perm[d], l)
        r2 = orc.match(og2, n, e)
        mapped = perm[r.table]
        mapped = mapped[np.lexsort(mapped.T[::-1])] if len(mapped) else mapped
        assert np.array_equal(mapped, r2.table)


def test_threads_and_root_sampling_agree():
    g = fm.Graph.erdos_renyi(400, 2000, 2, seed=8)
    s, d, l = g.edge_list()
    og = orc.OracleGraph(400, s, d, l)
    n, e = fixture_edges("_4_square", 0)
    one = orc.match(og, n, e, threads=1)
    many = orc.match(og, n, e, threads=4)
    assert np.array_equal(one.table, many.table)
    # strided root samples partition the roots
    parts = sum(orc.match(og, n, e, table=False, roots=(k, 400), stride=3).count for k in range(3))
    assert parts == one.count


def _valid_orders(n, edges):
    """Every order in which each vertex after the first touches an earlier one."""
    nb = {u: set() for u in range(n)}
    for a, b, _ in edges:
        if a != b:
            nb[a].add(b)
            nb[b].add(a)
    for perm in itertools.permutations(range(n)):
        if all(any(w in nb[perm[k]] for w in perm[:k]) for k in range(1, n)):
            yield list(perm)


@pytest.mark.parametrize("gi", [0, 3])
def test_any_matching_order_gives_the_same_embeddings(gi):
    """The embedding set does not depend on the matching order; the partials
    of each depth are those of the order (brute force over every prefix)."""
    V, s, d, l = tiny_graphs()[gi]
    og = orc.OracleGraph(V, s, d, l)
    for n, e in QUERIES:
        if n < 2 or n > 4:
            continue
        ref = brute_force(V, s, d, l, n, e)
        ea = np.array(e, np.uint32).reshape(-1, 3)
        for order in _valid_orders(n, e):
            got = orc.match(og, n, ea, order=order)
            assert np.array_equal(got.table, ref), (gi, e, order)
            inst = orc.match(og, n, ea, table=False, instrument=True, order=order)
            assert inst.partials == partials_brute(V, s, d, l, n, e, order), (gi, e, order)
            assert orc.plan(n, ea, order=order)[0] == order


def test_invalid_matching_order_is_rejected():
    e = np.array([(0, 1, 0), (1, 2, 0), (2, 3, 0)], np.uint32)
    with pytest.raises(ValueError):
        orc.plan(4, e, order=[0, 2, 1, 3])  # 2 does not touch 0
    with pytest.raises(ValueError):
        orc.plan(4, e, order=[0, 1, 1, 3])


def test_timed_count_is_the_count_when_it_finishes_and_stops_when_it_does_not():
    """The CPU baseline's time-budgeted enumeration: with room to finish it
    counts exactly what the oracle counts; with a budget far below the work
    it stops early with a partial count."""
    V = 300
    s, d, l = single_label_graph(V, 3000, 5)
    og = orc.OracleGraph(V, s, d, l)
    path4 = np.array([(0, 1, 1), (1, 0, 1), (1, 2, 1), (2, 1, 1), (2, 3, 1), (3, 2, 1)], np.uint32)
    full = orc.match(og, 4, path4, table=False).count
    cnt, secs, stopped = orc.timed_count(og, 4, path4, 2, 60.0)
    assert (cnt, stopped) == (full, False) and secs >= 0
    star = np.array([(0, k, 1) for k in range(1, 7)] + [(k, 0, 1) for k in range(1, 7)], np.uint32)
    cnt, secs, stopped = orc.timed_count(og, 7, star, 1, 0.05)
    # sum over v of d(d-1)...(d-5): far more than 0.05 s of enumeration
    assert stopped and cnt > 0 and secs < 5
