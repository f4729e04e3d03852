"""Host side of the library (no device): .g parsing and its error codes,
the label-partitioned CSR, the binary cache, the generators, the planner
and the CLI's argument/parse exit codes."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import forkmatch_amd as fm
import oracle as orc
from conftest import FIXTURE_NAMES, ROOT, fixture_path

CLI = ROOT / "forkmatch_amd" / "ForkMatch"


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


GOOD_DATA = "t # 0\n4 8 1 2\nv 0 1\nv 1 1\nv 2 0\nv 3 0\ne 0 1 1\ne 1 0 1\ne 1 2 2\ne 2 1 2\ne 2 3 1\ne 3 2 1\ne 3 0 0\ne 0 3 0\nt # -1"


def test_parse_data_graph(tmp_path):
    g = fm.Graph.parse_g(write(tmp_path, "d.g", GOOD_DATA))
    gi = g.info()
    assert (gi.vertices, gi.edges, gi.labels, gi.symmetric, gi.header_edges) == (4, 8, 3, 1, 8)
    ptr, col, labels = g.csr(0)
    assert labels.tolist() == [0, 1, 2]
    V = 4
    row = lambda l, v: col[ptr[l * V + v]:ptr[l * V + v + 1]].tolist()  # noqa: E731
    assert row(1, 1) == [0] and row(2, 1) == [2] and row(0, 0) == [3] and row(1, 2) == [3]


def test_trailing_newline_is_accepted(tmp_path):
    assert fm.Graph.parse_g(write(tmp_path, "d.g", GOOD_DATA + "\n")).edges == 8


@pytest.mark.parametrize("text,code", [
    ("t # 1\n4 8 1 2\nt # -1", 76),
    (GOOD_DATA.replace("t # -1", "t # -2"), 77),
    (GOOD_DATA.replace("4 8 1 2", "4 8 x 2"), 78),
    (GOOD_DATA.replace("e 1 2 2", "e 1 2"), 79),
    (GOOD_DATA.replace("e 1 2 2", "x 1 2 2"), 79),
    (GOOD_DATA.replace("e 1 2 2", "e 1 9 2"), 79),
])
def test_data_graph_errors(tmp_path, text, code):
    with pytest.raises(fm.ForkMatchError) as ei:
        fm.Graph.parse_g(write(tmp_path, "d.g", text))
    assert ei.value.code == code


def test_missing_data_graph():
    with pytest.raises(fm.ForkMatchError) as ei:
        fm.Graph.parse_g("/nonexistent/data.g")
    assert ei.value.code == 12


@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_parse_query_fixtures(name):
    q = fm.Query.load_g(str(fixture_path(name)))
    text = fixture_path(name).read_text().splitlines()
    lines = [tuple(map(int, l.split()[1:])) for l in text if l.startswith("e ")]
    assert sorted(map(tuple, q.edges().tolist())) == sorted(set(lines))
    assert q.vertices == int(text[1].split()[0])


@pytest.mark.parametrize("text,code", [
    ("t # 1\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 0 1\nt # -1\n", 233),
    ("t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\nz 1 0 1\nt # -1\n", 233),
    ("t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 0 1\nt # -2\n", 233),
    ("t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 5 1\nt # -1\n", 233),
    ("t # 0\n3 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 0 1\nt # -1\n", 234),
    ("t # 0\n2 3 1 1\nv 0 1\nv 1 1\ne 0 1 1\ne 1 0 1\nt # -1\n", 234),
    ("t # 0\n4 4 1 1\nv 0 1\nv 1 1\nv 2 1\nv 3 1\ne 0 1 1\ne 1 0 1\ne 2 3 1\ne 3 2 1\nt # -1\n", 1),
])
def test_query_errors(tmp_path, text, code):
    with pytest.raises(fm.ForkMatchError) as ei:
        fm.Query.load_g(write(tmp_path, "q.g", text))
    assert ei.value.code == code


def test_query_limits():
    with pytest.raises(fm.ForkMatchError) as ei:
        fm.Query.from_edges(33, [(i, i + 1, 0) for i in range(32)])
    assert ei.value.code == 307
    # 17 backward edges on one vertex
    edges = [(i, 17, 0) for i in range(17)] + [(i, i + 1, 0) for i in range(16)]
    with pytest.raises(fm.ForkMatchError) as ei:
        fm.Query.from_edges(18, edges)
    assert ei.value.code == 307


def test_missing_query():
    with pytest.raises(fm.ForkMatchError) as ei:
        fm.Query.load_g("/nonexistent/q.g")
    assert ei.value.code == 232


def reference_rows(V, s, d, l):
    rows = {}
    for a, b, c in set(zip(s.tolist(), d.tolist(), l.tolist())):
        rows.setdefault((c, a), []).append(b)
    return {k: sorted(v) for k, v in rows.items()}


@pytest.mark.parametrize("directed", [False, True])
def test_csr_rows_equal_sorted_adjacency(directed):
    rng = np.random.default_rng(1)
    V = 300
    s = rng.integers(0, V, 3000).astype(np.uint32)
    d = rng.integers(0, V, 3000).astype(np.uint32)
    l = rng.choice(np.array([3, 9, 70000, 5], np.uint32), 3000)
    if not directed:
        s, d, l = np.concatenate([s, d]), np.concatenate([d, s]), np.concatenate([l, l])
    # duplicates are collapsed
    s, d, l = np.concatenate([s, s[:50]]), np.concatenate([d, d[:50]]), np.concatenate([l, l[:50]])
    g = fm.Graph.from_edges(V, s, d, l)
    gi = g.info()
    assert gi.symmetric == (0 if directed else 1)
    assert gi.edges == len(set(zip(s.tolist(), d.tolist(), l.tolist())))
    for direction in (0, 1):
        ptr, col, labels = g.csr(direction)
        assert labels.tolist() == sorted(set(l.tolist()))
        want = reference_rows(V, s, d, l) if direction == 0 else reference_rows(V, d, s, l)
        for li, lab in enumerate(labels.tolist()):
            for v in range(V):
                got = col[ptr[li * V + v]:ptr[li * V + v + 1]].tolist()
                assert got == want.get((lab, v), [])


def test_write_and_binary_round_trip(tmp_path):
    g = fm.Graph.rmat(10, 4000, 5, zipf_s=1.2, seed=3)
    p = str(tmp_path / "g.g")
    g.write_g(p)
    assert not Path(p).read_bytes().endswith(b"\n")
    g2 = fm.Graph.parse_g(p)
    b = str(tmp_path / "g.bin")
    g.save_bin(b)
    g3 = fm.Graph.load_bin(b)
    for other in (g2, g3):
        for direction in (0, 1):
            for x, y in zip(g.csr(direction), other.csr(direction)):
                assert np.array_equal(x, y)


def test_generators_are_deterministic_and_exact():
    a = fm.Graph.erdos_renyi(1000, 4000, 8, seed=11)
    b = fm.Graph.erdos_renyi(1000, 4000, 8, seed=11)
    c = fm.Graph.erdos_renyi(1000, 4000, 8, seed=12)
    sa, da, la = a.edge_list()
    assert a.edges == 8000 and a.info().symmetric
    assert not np.any(sa == da)
    assert all(np.array_equal(x, y) for x, y in zip(a.edge_list(), b.edge_list()))
    assert not np.array_equal(a.edge_list()[1], c.edge_list()[1])
    assert set(la.tolist()) == set(range(8))


def test_rmat_shape_and_zipf_labels():
    g = fm.Graph.rmat(14, 16 << 14, 87, zipf_s=1.2, seed=0xF0F2)
    s, d, l = g.edge_list()
    assert g.edges == 2 * (16 << 14) and g.info().symmetric
    deg = np.bincount(s, minlength=1 << 14)
    # skew: the max degree is far above the mean, vertex 0 (quadrant a) is a hub
    assert deg.max() > 20 * deg.mean() and deg[0] > 10 * deg.mean()
    freq = np.bincount(l, minlength=87) / len(l)
    w = (np.arange(87) + 1.0) ** -1.2
    w /= w.sum()
    assert abs(freq[0] - w[0]) < 0.01 and abs(freq[1] - w[1]) < 0.01


@pytest.mark.parametrize("shape,cycles", [(0, 0), (1, 1), (1, 2), (2, 0)])
def test_query_extraction(shape, cycles):
    g = fm.Graph.rmat(12, 16 << 12, 87, zipf_s=1.2, seed=1)
    s, d, l = g.edge_list()
    og = orc.OracleGraph(g.vertices, s, d, l)
    for seed in range(5):
        q = fm.Query.random_walk(g, 6, shape=shape, min_cycles=cycles, seed=seed)
        e = q.edges()
        und = len(e) // 2
        if shape == 0:
            assert und == 5
        else:
            assert und - 5 >= max(cycles, 1 if shape == 1 else 0)
        assert orc.match(og, 6, e, table=False).count >= 1


def test_product_plan_equals_oracle_plan():
    rng = np.random.default_rng(4)
    for t in range(200):
        n = int(rng.integers(1, 9))
        m = int(rng.integers(n - 1, 3 * n + 1)) if n > 1 else 0
        # a random spanning tree keeps the query connected
        edges = [(int(rng.integers(0, i)), i, int(rng.integers(0, 3))) for i in range(1, n)]
        edges += [(int(rng.integers(0, n)), int(rng.integers(0, n)), int(rng.integers(0, 3)))
                  for _ in range(max(0, m - len(edges)))]
        if t % 2:
            edges += [(b, a, c) for a, b, c in edges]
        q = fm.Query.from_edges(n, edges)
        p = q.plan()
        o = orc.plan(n, np.array(edges, np.uint32).reshape(-1, 3))
        assert p.order == o[0]
        assert p.req_begin == o[1]
        assert (p.anchor, p.label, p.direction) == (o[2], o[3], o[4])


def test_cli_argv_and_parse_exit_codes(tmp_path):
    assert subprocess.run([str(CLI)], capture_output=True).returncode == 99
    assert subprocess.run([str(CLI), "a"], capture_output=True).returncode == 99
    d = write(tmp_path, "d.g", GOOD_DATA)
    assert subprocess.run([str(CLI), "/nonexistent/q.g", "/nonexistent/d.g"], capture_output=True).returncode == 12
    assert subprocess.run([str(CLI), "/nonexistent/q.g", d], capture_output=True).returncode == 232
    bad = write(tmp_path, "b.g", "t # 0\n2 2 1 1\nv 0 1\nv 1 1\ne 0 1 1\nq 1 0 1\nt # -1\n")
    assert subprocess.run([str(CLI), bad, d], capture_output=True).returncode == 233


def _random_query(rng, n, labels):
    edges = [(int(rng.integers(0, i)), i, int(rng.integers(0, labels))) for i in range(1, n)]
    edges += [(int(rng.integers(0, n)), int(rng.integers(0, n)), int(rng.integers(0, labels)))
              for _ in range(int(rng.integers(0, n)))]
    return edges


@pytest.mark.parametrize("directed", [False, True])
def test_cost_plan_requirements_equal_oracle_for_its_order(directed):
    """The product's requirements for its cost-based order equal the oracle's
    requirements for the same order; the reference mode is the query's plan."""
    rng = np.random.default_rng(11)
    g = fm.Graph.rmat(10, 4000, 6, zipf_s=1.2, seed=3)
    if directed:
        s, d, l = g.edge_list()
        keep = s < d
        g = fm.Graph.from_edges(g.vertices, s[keep], d[keep], l[keep])
    for t in range(150):
        n = int(rng.integers(2, 10))
        edges = _random_query(rng, n, 7)
        q = fm.Query.from_edges(n, edges)
        p = q.plan(g)
        o = orc.plan(n, np.array(edges, np.uint32).reshape(-1, 3), order=p.order)
        assert sorted(p.order) == list(range(n))
        assert p.order == o[0]
        assert p.req_begin == o[1]
        assert (p.anchor, p.label, p.direction) == (o[2], o[3], o[4])
        r = q.plan(g, plan=fm.PLAN_REFERENCE)
        assert r == q.plan()


def test_cost_plan_starts_at_the_rare_label():
    """A path whose last edge has a label the graph holds twice starts there."""
    rng = np.random.default_rng(2)
    V = 500
    s = rng.integers(0, V, 6000).astype(np.uint32)
    d = rng.integers(0, V, 6000).astype(np.uint32)
    lab = np.zeros(6000, np.uint32)
    lab[:2] = 9
    S = np.concatenate([s, d])
    D = np.concatenate([d, s])
    L = np.concatenate([lab, lab])
    g = fm.Graph.from_edges(V, S, D, L)
    path = [(0, 1, 0), (1, 2, 0), (2, 3, 0), (3, 4, 9)]
    edges = path + [(b, a, c) for a, b, c in path]
    p = fm.Query.from_edges(5, edges).plan(g)
    assert set(p.order[:2]) == {3, 4}
    # the reference order starts at the lowest id of highest out-degree
    assert fm.Query.from_edges(5, edges).plan().order[0] == 1
