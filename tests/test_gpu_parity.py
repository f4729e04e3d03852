"""GPU parity: the HIP engine (through libforkmatch.so) against the CPU oracle.

Bit-exact bar: the same count and the same lexicographically sorted
embedding table (columns in query vertex order) on the same inputs.
"""
import numpy as np
import pytest

import forkmatch_amd as fm
import oracle as orc
from conftest import FIXTURE_NAMES, fixture_path

pytestmark = pytest.mark.gpu

TIMEOUT_MS = 30000


def _oracle_graph(g):
    s, d, l = g.edge_list()
    return orc.OracleGraph(g.vertices, s, d, l)


def _check(engine, g, og, q, label=""):
    n = q.vertices
    edges = q.edges()
    ref = orc.match(og, n, edges)
    r = engine.match(g, q, timeout_ms=TIMEOUT_MS)
    out = r.query(0)
    assert out.status == 0, (label, out)
    assert out.count == ref.count, (label, out.count, ref.count)
    assert out.rows == ref.count, (label, out.rows)
    tab = r.table(0)
    assert tab.shape == ref.table.shape, label
    assert np.array_equal(tab, ref.table), label
    return ref.count


@pytest.fixture(scope="module")
def er(er_graph):
    er_graph.upload()
    return er_graph, _oracle_graph(er_graph)


def test_device_csr_matches_host(er):
    g, _ = er
    for direction in (0, 1):
        hp, hc, _ = g.csr(direction)
        dp, dc, _ = g.csr(direction, device=True)
        assert np.array_equal(hp, dp)
        assert np.array_equal(hc, dc)


# The fixtures' label 1 closes no triangle and no 4-cycle on the 1,000-vertex
# graph with labels 0..7, so those two run on the single-label graph below.
@pytest.mark.parametrize("name", [n for n in FIXTURE_NAMES if n not in ("_3_triangle", "_4_square")])
def test_reference_fixture_queries(engine, er, name):
    g, og = er
    q = fm.Query.load_g(str(fixture_path(name)))
    assert _check(engine, g, og, q, name) > 0, name


def test_config1_walk_cycles(engine, er):
    """The 1,000-vertex configuration's query: a 4-cycle taken by seeded
    random walk on the data graph, keeping its edge labels (at least the
    walk itself matches), against the oracle's count and table."""
    g, og = er
    for seed in range(8):
        q = fm.Query.random_walk(g, 4, shape=4, seed=seed)
        assert len(q.edges()) == 8
        assert _check(engine, g, og, q, f"walk 4-cycle seed={seed}") > 0


@pytest.mark.parametrize("name", FIXTURE_NAMES)
def test_fixture_queries_single_label(engine, name):
    # every data edge labelled 1, like the fixtures: counts are the closed
    # forms of the unlabelled queries
    g = fm.Graph.erdos_renyi(300, 1500, 1, seed=7)
    s, d, _ = g.edge_list()
    g = fm.Graph.from_edges(300, s, d, np.ones_like(s)).upload()
    og = orc.OracleGraph(300, s, d, np.ones_like(s))
    q = fm.Query.load_g(str(fixture_path(name)))
    _check(engine, g, og, q, name)


def test_random_walk_queries(engine, er):
    g, og = er
    total = 0
    for seed in range(24):
        n = 3 + seed % 5
        shape = seed % 3
        q = fm.Query.random_walk(g, n, shape=shape, min_cycles=1 if shape == 1 else 0, seed=seed)
        total += _check(engine, g, og, q, f"walk seed={seed} n={n} shape={shape}")
    assert total > 0


def test_rmat_zipf_queries(engine):
    # R-MAT at scale 12 with Zipf(1.2) labels over 0..86 (the shape of the
    # 2^20-vertex benchmark graph)
    g = fm.Graph.rmat(12, 16 * 4096, 87, zipf_s=1.2, seed=0xF0F2).upload()
    og = _oracle_graph(g)
    for seed in range(8):
        n = 4 + seed % 3
        q = fm.Query.random_walk(g, n, shape=seed % 3, min_cycles=2 if seed % 3 == 1 else 0, seed=100 + seed)
        ref = orc.match(og, n, q.edges(), table=False)
        if ref.count > 2_000_000:
            continue
        _check(engine, g, og, q, f"rmat seed={seed}")


def test_count_mode_equals_table_mode(engine, er):
    g, og = er
    q = fm.Query.load_g(str(fixture_path("_4_square")))
    a = engine.match(g, q, table=False, timeout_ms=TIMEOUT_MS).query(0)
    b = engine.match(g, q, table=True, timeout_ms=TIMEOUT_MS).query(0)
    assert a.count == b.count
    assert a.rows == 0


def test_batch_equals_alone(engine, er):
    g, _ = er
    qs = [fm.Query.random_walk(g, 3 + i % 5, shape=i % 3, min_cycles=1 if i % 3 == 1 else 0, seed=500 + i)
          for i in range(16)]
    alone = []
    for q in qs:
        r = engine.match(g, q, timeout_ms=TIMEOUT_MS)
        alone.append((r.count(0), r.table(0)))
    rb = engine.match_batch(g, qs, timeout_ms=TIMEOUT_MS)
    for i, q in enumerate(qs):
        assert rb.query(i).status == 0
        assert rb.count(i) == alone[i][0]
        assert np.array_equal(rb.table(i), alone[i][1])


def test_directed_graph(engine):
    rng = np.random.default_rng(3)
    V = 200
    s = rng.integers(0, V, 1500, dtype=np.uint32)
    d = rng.integers(0, V, 1500, dtype=np.uint32)
    keep = s != d
    s, d = s[keep], d[keep]
    lab = rng.integers(0, 3, len(s), dtype=np.uint32)
    g = fm.Graph.from_edges(V, s, d, lab).upload()
    assert not g.info().symmetric
    og = orc.OracleGraph(V, s, d, lab)
    # directed path, directed triangle, 2-cycle with distinct labels, fan-in
    queries = [
        (3, [(0, 1, 0), (1, 2, 1)]),
        (3, [(0, 1, 0), (1, 2, 0), (2, 0, 0)]),
        (2, [(0, 1, 1), (1, 0, 2)]),
        (4, [(1, 0, 0), (2, 0, 1), (3, 0, 2)]),
        (4, [(0, 1, 0), (0, 2, 1), (1, 3, 2), (2, 3, 0)]),
    ]
    for n, e in queries:
        _check(engine, g, og, fm.Query.from_edges(n, e), str(e))


def test_absent_label_has_no_match(engine, er):
    g, _ = er
    q = fm.Query.from_edges(2, [(0, 1, 1000), (1, 0, 1000)])
    r = engine.match(g, q, timeout_ms=TIMEOUT_MS)
    assert r.count(0) == 0 and r.query(0).status == 0


def test_single_vertex_query(engine, er):
    g, og = er
    q = fm.Query.from_edges(1, [])
    r = engine.match(g, q, timeout_ms=TIMEOUT_MS)
    assert r.count(0) == g.vertices
    assert np.array_equal(r.table(0)[:, 0], np.arange(g.vertices, dtype=np.uint32))


def test_self_loops(engine):
    rng = np.random.default_rng(5)
    V = 100
    s = rng.integers(0, V, 600, dtype=np.uint32)
    d = rng.integers(0, V, 600, dtype=np.uint32)
    loops = rng.integers(0, V, 30, dtype=np.uint32)
    s = np.concatenate([s, d, loops])
    d = np.concatenate([d, s[:600], loops])
    lab = np.zeros(len(s), np.uint32)
    g = fm.Graph.from_edges(V, s, d, lab).upload()
    og = orc.OracleGraph(V, s, d, lab)
    for n, e in [(1, [(0, 0, 0)]), (2, [(0, 1, 0), (1, 0, 0), (0, 0, 0)]),
                 (3, [(0, 1, 0), (1, 0, 0), (1, 2, 0), (2, 1, 0), (2, 2, 0)])]:
        _check(engine, g, og, fm.Query.from_edges(n, e), str(e))


@pytest.mark.parametrize("plan", [fm.PLAN_REFERENCE, fm.PLAN_COST])
def test_instrumented_counters_match_oracle(engine, er, plan):
    g, og = er
    for name in ["_3_triangle", "_4_square", "_5_road5"]:
        q = fm.Query.load_g(str(fixture_path(name)))
        order = q.plan(g, plan=plan).order
        ref = orc.match(og, q.vertices, q.edges(), table=False, instrument=True, order=order)
        out = engine.match(g, q, table=False, instrument=True, timeout_ms=TIMEOUT_MS, plan=plan).query(0)
        assert out.count == ref.count
        assert out.alg_row_bytes == 4 * ref.row_words, name
        assert out.partials == ref.partials, name


def test_launch_knobs_keep_tables_and_counters(engine):
    """The autotuner's knobs change how the work is cut, never the result.
    A batch of walk queries on a skewed R-MAT graph gives the oracle's
    tables for every LDS tile (1, 5 or 64 parents staged per step) and for
    waves per query from 1 (a query's roots in one claim) to more claims
    than roots; one query's instrumented counters equal the oracle's for
    every tile. A tile above 64 parents is refused."""
    g = fm.Graph.rmat(11, 12000, 12, zipf_s=1.2, seed=21).upload()
    og = _oracle_graph(g)
    qs = [fm.Query.random_walk(g, n, shape=shape, min_cycles=1 if shape == 1 else 0, seed=seed)
          for n, shape, seed in [(5, 0, 401), (6, 0, 404), (6, 1, 401), (7, 1, 401), (7, 2, 401), (9, 2, 401),
                                 (4, 4, 3), (12, 2, 402)]]
    refs = [orc.match(og, q.vertices, q.edges(), threads=4) for q in qs]
    assert sum(1 for ref in refs if ref.count > 0) >= 6
    for knobs in ({"tile_parents": 1}, {"tile_parents": 5, "waves_per_query": 1},
                  {"tile_parents": 64, "waves_per_query": 1 << 20}, {"waves_per_query": 7, "task_rows": 16}):
        r = engine.match_batch(g, qs, timeout_ms=TIMEOUT_MS, **knobs)
        for i, ref in enumerate(refs):
            assert r.count(i) == ref.count, (knobs, i)
            assert np.array_equal(r.table(i), ref.table), (knobs, i)
        r.close()
    q = qs[3]
    ref = orc.match(og, q.vertices, q.edges(), table=False, instrument=True, order=q.plan(g).order)
    for tp in (1, 5, 64):
        out = engine.match(g, q, table=False, instrument=True, tile_parents=tp, timeout_ms=TIMEOUT_MS).query(0)
        assert out.count == ref.count, tp
        assert out.alg_row_bytes == 4 * ref.row_words, tp
        assert out.partials == ref.partials, tp
    with pytest.raises(fm.ForkMatchError) as err:
        engine.match(g, q, tile_parents=65, timeout_ms=TIMEOUT_MS)
    assert err.value.code == 300


def test_cost_and_reference_plans_give_the_same_table(engine):
    """Random-walk queries on a skewed R-MAT graph with Zipf labels: both
    matching orders give the oracle's sorted table."""
    g = fm.Graph.rmat(11, 12000, 12, zipf_s=1.2, seed=21).upload()
    og = _oracle_graph(g)
    # seeds whose cost-based order differs from the reference order, with
    # 768 .. 394,342 embeddings
    for i, (n, shape, seed) in enumerate([(5, 0, 401), (6, 0, 404), (6, 1, 401), (7, 1, 401), (7, 2, 401),
                                          (9, 2, 401)]):
        q = fm.Query.random_walk(g, n, shape=shape, min_cycles=1 if shape == 1 else 0, seed=seed)
        assert q.plan(g).order != q.plan().order
        ref = orc.match(og, n, q.edges(), threads=4)
        for plan in (fm.PLAN_COST, fm.PLAN_REFERENCE):
            r = engine.match(g, q, timeout_ms=TIMEOUT_MS, plan=plan)
            assert r.count(0) == ref.count, (i, plan)
            assert np.array_equal(r.table(0), ref.table), (i, plan)
            r.close()


def test_table_truncation_keeps_exact_count(engine, er):
    g, og = er
    q = fm.Query.load_g(str(fixture_path("_4_roadY")))
    ref = orc.match(og, q.vertices, q.edges())
    assert ref.count > 100
    r = engine.match(g, q, max_rows=100, timeout_ms=TIMEOUT_MS)
    out = r.query(0)
    assert out.count == ref.count
    assert out.rows == 100 and out.truncated
    tab = r.table(0, sorted=True)
    # every stored row is a real embedding
    ref_rows = {tuple(x) for x in ref.table.tolist()}
    assert all(tuple(x) in ref_rows for x in tab.tolist())
    assert len({tuple(x) for x in tab.tolist()}) == 100


def test_count_limit_stops_query(engine, er):
    g, _ = er
    q = fm.Query.load_g(str(fixture_path("_5_plus")))
    r = engine.match(g, q, table=False, max_count=10, timeout_ms=TIMEOUT_MS, allow_errors=True)
    out = r.query(0)
    assert out.status == 306  # FM_ERR_COUNT_LIMIT
    assert out.count > 10


def test_small_bank_pool_recycles_banks(er):
    g, og = er
    # 64 banks of 16 KiB (the smallest bank): the frontier must be recycled
    # through the free list
    with fm.Engine(soln_prealloc=64 * 16384, output_prealloc=1 << 26, bank_size=2048,
                   output_page_bytes=1 << 16) as e:
        for name in ["_4_road4", "_5_road5", "_4_square"]:
            q = fm.Query.load_g(str(fixture_path(name)))
            ref = orc.match(og, q.vertices, q.edges())
            r = e.match(g, q, timeout_ms=TIMEOUT_MS, grid_waves=16, task_rows=32)
            st = r.stats()
            assert r.count(0) == ref.count, name
            assert np.array_equal(r.table(0), ref.table), name
            assert st.banks_reused > 0 or st.banks_allocated <= 64


def test_large_queries_at_the_smallest_bank():
    """Walk queries of 12 to 32 vertices (the plan limit) with the smallest
    bank (2,048 8-byte units: one step of 2 x 64 children of 32 words), on
    a graph with 32 edge labels where long labelled walks have few matches:
    one launch each and all in one batch launch give the oracle's count and
    sorted table."""
    g = fm.Graph.erdos_renyi(2000, 8000, 32, seed=0xF0F6).upload()
    og = _oracle_graph(g)
    with fm.Engine(soln_prealloc=1 << 28, output_prealloc=1 << 28, bank_size=2048, output_page_bytes=1 << 16) as e:
        chosen = []
        for n, shape, cyc in [(12, 1, 2), (16, 1, 2), (18, 3, 3), (20, 3, 2), (24, 3, 4), (32, 3, 5), (32, 0, 0)]:
            for seed in range(20):
                q = fm.Query.random_walk(g, n, shape=shape, min_cycles=cyc, seed=1000 * n + seed)
                o = e.match(g, q, table=False, max_count=1 << 18, max_partials=1 << 22, timeout_ms=TIMEOUT_MS,
                            allow_errors=True).query(0)
                if o.status == 0:
                    chosen.append(q)
                    break
        assert [q.vertices for q in chosen] == [12, 16, 18, 20, 24, 32, 32]
        refs = [orc.match(og, q.vertices, q.edges(), threads=4, order=q.plan(g).order) for q in chosen]
        for q, ref in zip(chosen, refs):
            r = e.match(g, q, timeout_ms=TIMEOUT_MS)
            assert r.query(0).status == 0 and r.count(0) == ref.count > 0, q.vertices
            assert np.array_equal(r.table(0), ref.table), q.vertices
            r.close()
        rb = e.match_batch(g, chosen, timeout_ms=TIMEOUT_MS)
        for i, ref in enumerate(refs):
            assert rb.query(i).status == 0 and rb.count(i) == ref.count, i
            assert np.array_equal(rb.table(i), ref.table), i
        rb.close()


def test_wide_step_fits_the_smallest_bank():
    """A step of 2 x 64 children of 19 words (2,432 words) into a fresh bank
    of the smallest size (4,096 words): a labelled 17-edge path whose end is
    a hub of 1,000 leaves, each leaf with one tail edge. The 20-vertex query
    (the path, a leaf, its tail) reaches the hub at depth 18, where one
    parent has 1,000 candidates. Count and table equal the oracle's (1,000
    embeddings); with banks of half this size the step would not fit."""
    und = [(i, i + 1, i + 1) for i in range(17)]
    leaves = range(18, 1018)
    und += [(17, v, 50) for v in leaves] + [(v, v + 1000, 60) for v in leaves]
    s = np.array([a for a, b, _ in und] + [b for a, b, _ in und], np.uint32)
    d = np.array([b for a, b, _ in und] + [a for a, b, _ in und], np.uint32)
    lab = np.array([x for _, _, x in und] * 2, np.uint32)
    V = 2018
    g = fm.Graph.from_edges(V, s, d, lab).upload()
    og = orc.OracleGraph(V, s, d, lab)
    qe = [(i, i + 1, i + 1) for i in range(17)] + [(17, 18, 50), (18, 19, 60)]
    q = fm.Query.from_edges(20, [(a, b, x) for a, b, x in qe] + [(b, a, x) for a, b, x in qe])
    ref = orc.match(og, 20, q.edges())
    assert ref.count == 1000
    with fm.Engine(soln_prealloc=1 << 26, output_prealloc=1 << 26, bank_size=2048, output_page_bytes=1 << 16) as e:
        for plan in (fm.PLAN_COST, fm.PLAN_REFERENCE):
            r = e.match(g, q, timeout_ms=TIMEOUT_MS, plan=plan)
            assert r.query(0).status == 0 and r.count(0) == 1000, plan
            assert np.array_equal(r.table(0), ref.table), plan
            r.close()


def test_bank_size_below_one_step_is_refused():
    with pytest.raises(fm.ForkMatchError) as err:
        fm.Engine(soln_prealloc=1 << 24, output_prealloc=1 << 24, bank_size=2047)
    assert err.value.code == 300


def test_prealloc_budget_is_enforced(er):
    """PREALLOC bounds the arenas plus the graph: beyond it engine creation,
    a launch and a server fail with status 1 (the reference's exit status
    when its slab overflows); at exactly the budget they run."""
    g, _ = er
    kw = dict(soln_prealloc=1 << 24, output_prealloc=1 << 24, queue_capacity=1 << 12)
    with fm.Engine(prealloc=fm.PREALLOC_UNLIMITED, **kw) as e:
        arenas = e.info().arena_bytes
    need = arenas + g.info().device_bytes
    with pytest.raises(fm.ForkMatchError) as err:
        fm.Engine(prealloc=arenas - 1, **kw)
    assert err.value.code == 1 and "PREALLOC" in err.value.message
    q = fm.Query.load_g(str(fixture_path("_2_road2")))
    with fm.Engine(prealloc=need - 1, **kw) as e:
        with pytest.raises(fm.ForkMatchError) as err:
            e.match(g, q)
        assert err.value.code == 1 and "PREALLOC" in err.value.message
        with pytest.raises(fm.ForkMatchError) as err:
            fm.Server(e, g)
        assert err.value.code == 1
    with fm.Engine(prealloc=need, **kw) as e:
        assert e.match(g, q).count(0) > 0


def test_exhausted_bank_pool_fails_cleanly():
    # one 16 KiB bank (the smallest): the depth-1 rows of a 5-path on 8000
    # same-label edges fill it, and expanding them needs a second bank that
    # never comes
    g0 = fm.Graph.erdos_renyi(1000, 4000, 1, seed=9)
    s, d, _ = g0.edge_list()
    g = fm.Graph.from_edges(1000, s, d, np.ones_like(s)).upload()
    q = fm.Query.load_g(str(fixture_path("_5_road5")))
    with fm.Engine(soln_prealloc=16384, output_prealloc=1 << 24, bank_size=2048, output_page_bytes=1 << 16) as e:
        r = e.match(g, q, table=False, timeout_ms=TIMEOUT_MS, grid_waves=4, allow_errors=True)
        assert r.query(0).status == 303  # FM_ERR_POOL_EXHAUSTED
        # the engine recovers for the next launch
        q2 = fm.Query.load_g(str(fixture_path("_2_road2")))
        assert e.match(g, q2, table=False, timeout_ms=TIMEOUT_MS).count(0) > 0


def test_intersection_calls_of_different_pair_counts(engine):
    """One engine's work array serves calls with many pairs, then few, then
    many again (each call's first-segment table and status words reuse the
    array): every call's counts and intersections equal numpy's."""
    rng = np.random.default_rng(12)
    for npairs, max_len in [(3000, 3000), (7, 60000), (2500, 6000), (1, 200000), (3000, 3000), (40, 20000)]:
        lists, a_off, a_len, b_off, b_len, expected = [], [], [], [], [], []
        pos = 0
        for _ in range(npairs):
            na, nb = (int(x) for x in rng.integers(0, max_len, 2))
            span = 3 * max(na, nb, 1)
            a = np.unique(rng.integers(0, span, na, dtype=np.uint32))
            b = np.unique(rng.integers(0, span, nb, dtype=np.uint32))
            a_off.append(pos); a_len.append(len(a)); lists.append(a); pos += len(a)
            b_off.append(pos); b_len.append(len(b)); lists.append(b); pos += len(b)
            expected.append(np.intersect1d(a, b))
        data = np.concatenate(lists).astype(np.uint32)
        out_off = np.cumsum([0] + [min(x, y) for x, y in zip(a_len, b_len)]).astype(np.uint64)
        res = fm.intersect_pairs(engine, data, np.array(a_off, np.uint64), np.array(a_len, np.uint32),
                                 np.array(b_off, np.uint64), np.array(b_len, np.uint32), out_off[:-1],
                                 int(out_off[-1]))
        for i, ex in enumerate(expected):
            assert res["counts"][i] == len(ex), (npairs, i)
            assert np.array_equal(res["out"][out_off[i]: out_off[i] + len(ex)], ex), (npairs, i)


def test_intersection_microkernel(engine):
    rng = np.random.default_rng(11)
    lists, a_off, a_len, b_off, b_len = [], [], [], [], []
    pos = 0
    expected = []
    # one-segment and many-segment pairs (up to 129 merge segments of 16,384
    # path elements), skew pairs (longer list >= 32x the shorter), empty
    # pairs; dense segments (A part spanning < 32,768 values) at these spans,
    # sparse ones over the wide spans of the last three
    sizes = [(0, 5, 0), (1, 1, 0), (3, 3000, 0), (2500, 2600, 0), (5000, 70, 0), (64, 64, 0), (4096, 10000, 0),
             (1, 0, 0), (100000, 100000, 0), (600, 40000, 0), (30000, 25000, 0), (1000000, 1100000, 0),
             (20000, 30000, 10 ** 9), (5000, 6000, 10 ** 8), (700, 30000, 3 * 10 ** 6)]
    for na, nb, wide in sizes:
        span = wide or max(20000, 3 * max(na, nb))
        a = np.unique(rng.integers(0, span, na, dtype=np.uint32))
        b = np.unique(rng.integers(0, span, nb, dtype=np.uint32))
        a_off.append(pos); a_len.append(len(a)); lists.append(a); pos += len(a)
        b_off.append(pos); b_len.append(len(b)); lists.append(b); pos += len(b)
        expected.append(np.intersect1d(a, b))
    # a pair whose first segments are dense and later ones sparse
    a = np.concatenate([np.arange(40000, dtype=np.uint32),
                        np.arange(10 ** 8, 10 ** 8 + 30000 * 50000, 50000, dtype=np.int64).astype(np.uint32)])
    b = np.arange(0, 10 ** 8 + 30000 * 50000, 25000, dtype=np.int64).astype(np.uint32)
    b = np.union1d(b, np.arange(0, 40000, 3, dtype=np.uint32))
    a_off.append(pos); a_len.append(len(a)); lists.append(a); pos += len(a)
    b_off.append(pos); b_len.append(len(b)); lists.append(b); pos += len(b)
    expected.append(np.intersect1d(a, b))
    # equal lists (every element matches: every bit of the AND set) and
    # lists that hold the largest and smallest u32 values
    top = np.uint32(0xFFFFFFFF)
    for a, b in [(np.arange(9000, dtype=np.uint32), np.arange(9000, dtype=np.uint32)),
                 (np.array([0, 5, 7, top], np.uint32), np.array([0, 6, 7, top], np.uint32)),
                 (np.arange(top - np.uint32(6000), top, dtype=np.uint32).astype(np.uint32)[::2].copy(),
                  np.append(np.arange(top - np.uint32(7000), top, dtype=np.uint32)[::3], top).astype(np.uint32))]:
        a_off.append(pos); a_len.append(len(a)); lists.append(a); pos += len(a)
        b_off.append(pos); b_len.append(len(b)); lists.append(b); pos += len(b)
        expected.append(np.intersect1d(a, b))
    data = np.concatenate(lists).astype(np.uint32)
    out_off = np.cumsum([0] + [min(x, y) for x, y in zip(a_len, b_len)]).astype(np.uint64)
    res = fm.intersect_pairs(engine, data, np.array(a_off, np.uint64), np.array(a_len, np.uint32),
                             np.array(b_off, np.uint64), np.array(b_len, np.uint32), out_off[:-1],
                             int(out_off[-1]))
    counts, out = res["counts"], res["out"]
    for i, ex in enumerate(expected):
        assert counts[i] == len(ex)
        assert np.array_equal(out[out_off[i]: out_off[i] + counts[i]], ex)
    # counts only (no output arrays)
    res = fm.intersect_pairs(engine, data, np.array(a_off, np.uint64), np.array(a_len, np.uint32),
                             np.array(b_off, np.uint64), np.array(b_len, np.uint32))
    assert [int(c) for c in res["counts"]] == [len(ex) for ex in expected]


def test_hub_rows_are_split_across_waves(engine):
    # one hub joined to 2000 leaves plus a ring among the leaves: the hub's
    # driver rows are far longer than a step, so idle waves get halves
    V = 2001
    leaves = np.arange(1, V, dtype=np.uint32)
    ring_a = leaves
    ring_b = np.roll(leaves, 1)
    s = np.concatenate([np.zeros(V - 1, np.uint32), leaves, ring_a, ring_b])
    d = np.concatenate([leaves, np.zeros(V - 1, np.uint32), ring_b, ring_a])
    lab = np.zeros(len(s), np.uint32)
    g = fm.Graph.from_edges(V, s, d, lab).upload()
    og = orc.OracleGraph(V, s, d, lab)
    split = 0
    for n, e in [(3, [(0, 1, 0), (1, 0, 0), (0, 2, 0), (2, 0, 0)]),
                 (4, [(0, 1, 0), (1, 0, 0), (0, 2, 0), (2, 0, 0), (0, 3, 0), (3, 0, 0), (1, 2, 0), (2, 1, 0)])]:
        q = fm.Query.from_edges(n, e)
        ref = orc.match(og, n, q.edges())
        r = engine.match(g, q, timeout_ms=TIMEOUT_MS, task_rows=64)
        split += r.stats().tasks_split
        assert r.count(0) == ref.count
        assert np.array_equal(r.table(0), ref.table)
    assert split > 0


def test_count_only_queries_in_a_table_batch(engine, er):
    """Queries marked count-only in a table-mode batch are counted as in a
    count-mode launch (rows 0, not truncated), while the others in the same
    launch get their full tables: every count and table equals the
    oracle's, alone and batched."""
    g, og = er
    qs = [fm.Query.random_walk(g, 3 + i % 5, shape=i % 3, min_cycles=1 if i % 3 == 1 else 0, seed=700 + i)
          for i in range(12)]
    refs = [orc.match(og, q.vertices, q.edges()) for q in qs]
    for i, q in enumerate(qs):
        q.set_count_only(i % 2 == 0)
    rb = engine.match_batch(g, qs, timeout_ms=TIMEOUT_MS)
    for i, ref in enumerate(refs):
        out = rb.query(i)
        assert out.status == 0 and out.count == ref.count, i
        if i % 2 == 0:
            assert out.rows == 0 and not out.truncated, i
        else:
            assert out.rows == ref.count and np.array_equal(rb.table(i), ref.table), i
    rb.close()
    r = engine.match(g, qs[0], timeout_ms=TIMEOUT_MS)
    assert r.count(0) == refs[0].count and r.query(0).rows == 0
    r.close()
    qs[0].set_count_only(False)
    r = engine.match(g, qs[0], timeout_ms=TIMEOUT_MS)
    assert r.query(0).rows == refs[0].count and np.array_equal(r.table(0), refs[0].table)
    r.close()


def test_count_only_leaf_level_equals_oracle(engine):
    """Count-only queries whose last vertex is a leaf of the matching order
    run their last level in batches of up to 256 parents: counts equal the
    oracle's for trees and one-chord walks on a skewed R-MAT graph, for LDS
    tiles of 1, 5 and 64 parents (leaf batches of 4, 20 and 256), and for
    star queries on a hub graph whose long driver rows are split between
    waves (splits asserted)."""
    g = fm.Graph.rmat(11, 12000, 12, zipf_s=1.2, seed=21).upload()
    og = _oracle_graph(g)
    qs = [fm.Query.random_walk(g, n, shape=shape, seed=seed)
          for n, shape, seed in [(3, 0, 1), (5, 0, 401), (6, 0, 24), (7, 0, 33), (6, 3, 21), (7, 3, 24), (8, 3, 24),
                                 (5, 3, 20)]]
    refs = [orc.match(og, q.vertices, q.edges(), table=False, threads=4).count for q in qs]
    assert sum(1 for c in refs if c > 10000) >= 3
    for tp in (1, 5, 64):
        for q, ref in zip(qs, refs):
            r = engine.match(g, q, table=False, tile_parents=tp, timeout_ms=TIMEOUT_MS)
            assert r.query(0).status == 0 and r.count(0) == ref, (tp, q.vertices)
            r.close()
        rb = engine.match_batch(g, qs, table=False, tile_parents=tp, timeout_ms=TIMEOUT_MS)
        assert [rb.count(i) for i in range(len(qs))] == refs, tp
        rb.close()
    # stars on a hub joined to 3000 leaves (a ring among the leaves): a star
    # with k leaves has sum over v of deg(v) (deg(v) - 1) ... (deg(v) - k + 1)
    # embeddings (hub degree 3000, leaf degree 3)
    V = 3001
    leaves = np.arange(1, V, dtype=np.uint32)
    s = np.concatenate([np.zeros(V - 1, np.uint32), leaves, leaves, np.roll(leaves, 1)])
    d = np.concatenate([leaves, np.zeros(V - 1, np.uint32), np.roll(leaves, 1), leaves])
    lab = np.zeros(len(s), np.uint32)
    hg = fm.Graph.from_edges(V, s, d, lab).upload()
    split = 0
    for n in (3, 4):
        e = [x for v in range(1, n) for x in ((0, v, 0), (v, 0, 0))]
        q = fm.Query.from_edges(n, e)
        k = n - 1
        ref = int(np.prod(np.arange(3000 - k + 1, 3001, dtype=np.int64))) + 3000 * int(np.prod(np.arange(3 - k + 1, 4)))
        if n == 3:
            assert ref == orc.match(orc.OracleGraph(V, s, d, lab), n, q.edges(), table=False).count
        r = engine.match(hg, q, table=False, task_rows=64, timeout_ms=TIMEOUT_MS)
        split += r.stats().tasks_split
        assert r.count(0) == ref, n
    assert split > 0


@pytest.mark.parametrize("slice_ms,plan", [(0, fm.PLAN_COST), (1, fm.PLAN_REFERENCE)])
def test_serving_results_equal_alone_and_in_the_mix(er, slice_ms, plan):
    """Continuous batching: queries submitted while others run are admitted
    into the running launch; each result equals the query matched alone.
    With 1 ms slices the launch retires many times with work queued on the
    device, and the next poll relaunches it."""
    import time
    g, og = er
    qs = [fm.Query.random_walk(g, 3 + i % 5, shape=i % 3, min_cycles=1 if i % 3 == 1 else 0, seed=900 + i)
          for i in range(40)]
    refs = [orc.match(og, q.vertices, q.edges()) for q in qs]
    with fm.Engine(soln_prealloc=1 << 29, output_prealloc=1 << 29, bank_size=1 << 13, output_page_bytes=1 << 16) as e:
        with fm.Server(e, g, slots=16, slot_pages=64, idle_ms=5, max_launch_ms=60000, slice_ms=slice_ms,
                       plan=plan) as srv:
            pending = {}
            done = {}
            i = 0
            t_end = time.time() + 60
            while (i < len(qs) or pending) and time.time() < t_end:
                # submit in bursts while earlier queries are still running
                for _ in range(3):
                    if i < len(qs):
                        t = srv.submit(qs[i])
                        if t is None:
                            break
                        pending[t] = i
                        i += 1
                for c in srv.poll():
                    k = pending.pop(c.ticket)
                    assert c.status == 0, c.status
                    tab = srv.table(c.ticket, c.rows, c.vertices)
                    done[k] = (c.count, tab)
                    srv.release(c.ticket)
                time.sleep(0.001)
            st = srv.stats()
    assert len(done) == len(qs)
    assert st.launches >= 1
    for k, ref in enumerate(refs):
        assert done[k][0] == ref.count, k
        assert np.array_equal(done[k][1], ref.table), k


def test_small_tasks_fill_pages_instead_of_wasting_them():
    """511,472 rows of 16 bytes need 125 pages of 64 KiB. With 64-parent
    final tasks a wave keeps its page from task to task, so the table fits a
    256-page arena; a fresh page per task would need several hundred."""
    g = fm.Graph.erdos_renyi(1000, 4000, 1, seed=77).upload()
    og = _oracle_graph(g)
    q = fm.Query.from_edges(4, [(0, 1, 0), (1, 0, 0), (1, 2, 0), (2, 1, 0), (2, 3, 0), (3, 2, 0)])
    ref = orc.match(og, 4, q.edges())
    assert ref.count == 511472
    with fm.Engine(soln_prealloc=1 << 28, output_prealloc=256 << 16, output_page_bytes=1 << 16) as e:
        r = e.match(g, q, grid_waves=64, task_rows=64, timeout_ms=TIMEOUT_MS)
        out = r.query(0)
        assert out.status == 0 and out.count == ref.count
        assert out.rows == ref.count and not out.truncated
        assert np.array_equal(r.table(0), ref.table)
        r.close()


def _rows_subset(rows, table):
    """Every row of `rows` is a row of `table` and no row repeats."""
    v = np.dtype((np.void, rows.dtype.itemsize * rows.shape[1]))
    a = np.ascontiguousarray(rows).view(v).ravel()
    b = np.ascontiguousarray(table).view(v).ravel()
    return len(np.unique(a)) == len(a) and bool(np.isin(a, b).all())


def _serve_one(srv, q, max_rows=0):
    import time
    t = srv.submit(q, max_rows)
    assert t is not None
    t_end = time.time() + 60
    while time.time() < t_end:
        for c in srv.poll():
            assert c.ticket == t
            tab = srv.table(c.ticket, c.rows, c.vertices)
            srv.release(c.ticket)
            return c, tab
        time.sleep(0.001)
    raise AssertionError("query did not complete")


def test_serving_tables_are_dense():
    """A serving table holds ceil(rows / page_rows) pages however many waves
    emit its rows: 511,472 rows of 16 bytes (125 pages of 64 KiB) fit a
    130-page arena with every wave of the chip emitting. A row budget or an
    arena that runs out stores a prefix of whole pages (each row a distinct
    embedding) and flags the result truncated."""
    g = fm.Graph.erdos_renyi(1000, 4000, 1, seed=77).upload()
    og = _oracle_graph(g)
    q = fm.Query.from_edges(4, [(0, 1, 0), (1, 0, 0), (1, 2, 0), (2, 1, 0), (2, 3, 0), (3, 2, 0)])
    ref = orc.match(og, 4, q.edges())
    assert ref.count == 511472
    page_rows = (1 << 16) // 16
    for pages, max_rows, want_rows in [(130, 0, ref.count), (130, 10000, 10000), (64, 0, 64 * page_rows)]:
        with fm.Engine(soln_prealloc=1 << 28, output_prealloc=pages << 16, output_page_bytes=1 << 16) as e:
            with fm.Server(e, g, slots=4, slot_pages=256, max_launch_ms=60000) as srv:
                for _ in range(2):  # the second run reuses the recycled pages
                    c, tab = _serve_one(srv, q, max_rows)
                    assert c.status == 0 and c.count == ref.count, (pages, max_rows)
                    assert c.rows == want_rows, (pages, max_rows, c.rows)
                    assert bool(c.flags & fm.FM_RESULT_TRUNCATED) == (want_rows < ref.count)
                    if want_rows == ref.count:
                        assert np.array_equal(tab, ref.table)
                    else:
                        assert tab.shape == (want_rows, 4) and _rows_subset(tab, ref.table)


@pytest.mark.parametrize("table", [True, False])
def test_symmetry_breaking_gives_the_full_table(engine, table):
    """Symmetric patterns (including C5, C6 and the bowtie, where the
    reference's rule is inexact): one embedding per automorphism orbit is
    enumerated and expanded by the group; count and sorted table equal the
    oracle's full enumeration."""
    from test_symmetry import QUERIES
    V = 300
    g = fm.Graph.erdos_renyi(V, 1500, 1, seed=31)
    s, d, l = g.edge_list()
    og = orc.OracleGraph(V, s, d, l)
    g.upload()
    for name, (n, e) in QUERIES.items():
        if name.startswith("directed") or "loop" in name or "two labels" in name:
            continue
        q = fm.Query.from_edges(n, e)
        ref = orc.match(og, n, q.edges(), threads=4, table=table)
        r = engine.match(g, q, table=table, symmetry=True, timeout_ms=TIMEOUT_MS)
        out = r.query(0)
        assert out.status == 0 and out.count == ref.count, name
        if table:
            assert np.array_equal(r.table(0), ref.table), name
        r.close()


def test_serving_with_symmetry_breaking():
    """Serving with exact symmetry breaking: every embedding found is
    expanded by the query's automorphism group into the dense table (one
    claimed range per automorphism); each served query's count and sorted
    table equal the oracle's full enumeration, with all queries in flight
    together."""
    import time
    from test_symmetry import QUERIES
    V = 300
    g = fm.Graph.erdos_renyi(V, 1500, 1, seed=31)
    s, d, l = g.edge_list()
    og = orc.OracleGraph(V, s, d, l)
    g.upload()
    qs = [fm.Query.from_edges(n, e) for name, (n, e) in QUERIES.items()
          if not (name.startswith("directed") or "loop" in name or "two labels" in name)]
    refs = [orc.match(og, q.vertices, q.edges(), threads=4) for q in qs]
    with fm.Engine(soln_prealloc=1 << 29, output_prealloc=1 << 29, bank_size=1 << 13, output_page_bytes=1 << 16) as e:
        # 2,048 pages of 2,730 six-vertex rows hold the largest table (924k rows)
        with fm.Server(e, g, slots=32, slot_pages=2048, max_launch_ms=60000, symmetry=True) as srv:
            t2q = {}
            for i, q in enumerate(qs):
                t = srv.submit(q)
                assert t is not None
                t2q[t] = i
            done = {}
            t_end = time.time() + 60
            while t2q and time.time() < t_end:
                for c in srv.poll():
                    i = t2q.pop(c.ticket)
                    assert not (c.flags & fm.FM_RESULT_TRUNCATED), i
                    done[i] = (c.status, c.count, srv.table(c.ticket, c.rows, c.vertices))
                    srv.release(c.ticket)
                time.sleep(0.001)
    assert len(done) == len(qs)
    for i, ref in enumerate(refs):
        st, cnt, tab = done[i]
        assert st == 0 and cnt == ref.count, i
        assert np.array_equal(tab, ref.table), i


_GOOD_G = ("t # 0\n4 8 1 2\nv 0 1\nv 1 1\nv 2 0\nv 3 0\ne 0 1 1\ne 1 0 1\ne 1 2 2\ne 2 1 2\ne 2 3 1\ne 3 2 1\n"
           "e 3 0 0\ne 0 3 0\nt # -1")


def _same_graph(a, b, tmp_path=None):
    """a (loaded on the device) holds b's graph: header, rows of both
    directions, the device copy of the rows, and with tmp_path the binary
    cache files (root lists included) and a cost plan (row statistics)."""
    ia, ib = a.info(), b.info()
    assert (ia.vertices, ia.edges, ia.labels, ia.symmetric, ia.header_edges) == \
        (ib.vertices, ib.edges, ib.labels, ib.symmetric, ib.header_edges)
    for d in (0, 1):
        for x, y in zip(a.csr(d), b.csr(d)):
            assert np.array_equal(x, y)
        for x, y in zip(a.csr(d, device=True), b.csr(d)):
            assert np.array_equal(x, y)
    if tmp_path is not None:
        a.save_bin(str(tmp_path / "a.bin"))
        b.save_bin(str(tmp_path / "b.bin"))
        assert (tmp_path / "a.bin").read_bytes() == (tmp_path / "b.bin").read_bytes()
        labels = [int(x) for x in b.csr(0)[2][:3]]
        if len(labels) == 3:
            q = fm.Query.from_edges(4, [(u, v, labels[min(u, v)]) for u, v in
                                        ((0, 1), (1, 0), (1, 2), (2, 1), (2, 3), (3, 2))])
            assert q.plan(a) == q.plan(b)


def test_device_g_loader_equals_host_parser(tmp_path):
    """The .g body parsed on the device (line starts found per 4 KiB chunk,
    one thread per line) gives the host parser's graph: a symmetric
    multi-label graph, a directed one with vertex lines after edge lines,
    a file of ~120k lines crossing thousands of chunks, and that file with
    one edge's label changed (directed: the device symmetry check finds the
    one edge without its reverse)."""
    g1 = fm.Graph.erdos_renyi(1000, 4000, 8, seed=11)
    p1 = tmp_path / "er.g"
    g1.write_g(str(p1))
    _same_graph(fm.Graph.load_g(str(p1)), fm.Graph.parse_g(str(p1)), tmp_path)
    rng = np.random.default_rng(3)
    V, E = 3000, 60000
    s, d, lab = rng.integers(0, V, E), rng.integers(0, V, E), rng.integers(0, 40, E)
    lines = [f"e {a} {b} {c}" for a, b, c in zip(s, d, lab)] + [f"v {v} 0" for v in range(V)]
    p2 = tmp_path / "directed.g"
    p2.write_bytes(("t # 0\n" + f"{V} {E} 1 40\n" + "\n".join(lines) + "\nt # -1").encode())
    _same_graph(fm.Graph.load_g(str(p2)), fm.Graph.parse_g(str(p2)), tmp_path)
    g3 = fm.Graph.rmat(14, 60000, 16, zipf_s=1.2, seed=5)
    p3 = tmp_path / "rmat.g"
    g3.write_g(str(p3))
    _same_graph(fm.Graph.load_g(str(p3)), fm.Graph.parse_g(str(p3)), tmp_path)
    # symmetry is checked on the device: one edge among 120k whose reverse
    # has another label (one used elsewhere, then a new one) makes the graph
    # directed, with both directions built
    text = p3.read_text().split("\n")
    k = max(i for i, line in enumerate(text) if line.startswith("e "))
    for new_label in (int(text[k].split()[3]) ^ 1, 999):
        e = text[k].split()
        p4 = tmp_path / f"asym{new_label}.g"
        p4.write_text("\n".join(text[:k] + [f"e {e[1]} {e[2]} {new_label}"] + text[k + 1:]))
        loaded = fm.Graph.load_g(str(p4))
        assert not loaded.info().symmetric
        _same_graph(loaded, fm.Graph.parse_g(str(p4)), tmp_path)


@pytest.mark.parametrize("text", [
    _GOOD_G.replace("e 1 2 2", "e 1 2"),             # missing label
    _GOOD_G.replace("e 1 2 2", "x 1 2 2"),           # not a v/e line
    _GOOD_G.replace("e 1 2 2", "e 1 9 2"),           # dst out of range (found by the CSR build)
    _GOOD_G.replace("v 2 0", "v 7 0"),               # vertex id >= V
    _GOOD_G.replace("v 0 1", "v 0 99999999999"),     # number overflows u32, first body line
    _GOOD_G.replace("e 3 2 1\n", "e 3 2 1\n\n"),     # empty line
    _GOOD_G.replace("e 0 1 1", "e 0  1 1"),          # two spaces
    _GOOD_G.replace("e 2 1 2", "e 2 1 2 ") ,         # trailing space
    _GOOD_G.replace("e 0 3 0\nt # -1", "e 0 3 0t # -1"),  # last line unterminated
])
def test_device_g_loader_errors_equal_host_parser(tmp_path, text):
    """Malformed bodies: the device loader reports the host parser's code
    and message (the first bad line, with its line number)."""
    p = tmp_path / "bad.g"
    p.write_bytes(text.encode())
    with pytest.raises(fm.ForkMatchError) as host:
        fm.Graph.parse_g(str(p))
    with pytest.raises(fm.ForkMatchError) as dev:
        fm.Graph.load_g(str(p))
    assert (dev.value.code, dev.value.message) == (host.value.code, host.value.message)


@pytest.mark.parametrize("name", ["_3_triangle", "_4_square", "_5_road5"])
def test_cli_end_to_end_equals_oracle(tmp_path, name):
    """The ForkMatch CLI as the reference is run (`ForkMatch <query.g>
    <data.g>`: device .g load, one-query launch) prints the oracle's count
    and, with FM_TABLE_OUT, writes the oracle's table."""
    import os
    import re
    import subprocess
    from pathlib import Path
    g = fm.Graph.erdos_renyi(300, 1500, 1, seed=7)
    s, d, _ = g.edge_list()
    ones = np.ones_like(s)
    data = tmp_path / "d.g"
    fm.Graph.from_edges(300, s, d, ones).write_g(str(data))
    q = fm.Query.load_g(str(fixture_path(name)))
    ref = orc.match(orc.OracleGraph(300, s, d, ones), q.vertices, q.edges())
    cli = Path(fm.__file__).resolve().parent / "ForkMatch"
    table = tmp_path / "table.txt"
    r = subprocess.run([str(cli), str(fixture_path(name)), str(data)], capture_output=True, text=True,
                       timeout=120, cwd=tmp_path, env=dict(os.environ, FM_TABLE_OUT=str(table)))
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"^Soln Size: (\d+)$", r.stdout, re.M)
    assert m and int(m.group(1)) == ref.count, r.stdout
    rows = np.loadtxt(table, dtype=np.uint32, ndmin=2) if ref.count else np.zeros((0, q.vertices), np.uint32)
    rows = rows[np.lexsort(rows.T[::-1])] if len(rows) else rows
    assert np.array_equal(rows, ref.table.reshape(-1, q.vertices))
