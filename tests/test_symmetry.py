"""Exact symmetry breaking (host side): the automorphism group of a query and
the ordering constraints that keep one embedding per orbit. Pinned against
brute force: the group against all vertex permutations, and the constraints
against the oracle's full enumeration (the constrained rows times the group
must give back every embedding exactly once)."""
import itertools

import numpy as np
import pytest

import forkmatch_amd as fm
import oracle as orc


def und(es):
    return [(a, b, l) for a, b, l in es] + [(b, a, l) for a, b, l in es]


QUERIES = {
    "triangle": (3, und([(0, 1, 0), (1, 2, 0), (2, 0, 0)])),
    "C4": (4, und([(0, 1, 0), (1, 2, 0), (2, 3, 0), (3, 0, 0)])),
    "C5": (5, und([(i, (i + 1) % 5, 0) for i in range(5)])),
    "C6": (6, und([(i, (i + 1) % 6, 0) for i in range(6)])),
    "bowtie": (5, und([(0, 1, 0), (1, 2, 0), (2, 0, 0), (0, 3, 0), (3, 4, 0), (4, 0, 0)])),
    "K4": (4, und([(a, b, 0) for a in range(4) for b in range(a + 1, 4)])),
    "star4": (5, und([(0, i, 0) for i in range(1, 5)])),
    "P4": (4, und([(0, 1, 0), (1, 2, 0), (2, 3, 0)])),
    "C4 two labels": (4, und([(0, 1, 0), (1, 2, 1), (2, 3, 0), (3, 0, 1)])),
    "directed C3": (3, [(0, 1, 0), (1, 2, 0), (2, 0, 0)]),
    "path with loop": (3, und([(0, 1, 0), (1, 2, 0)]) + [(0, 0, 1), (2, 2, 1)]),
}


def brute_auts(n, edges):
    es = set(map(tuple, edges))
    return sorted(list(p) for p in itertools.permutations(range(n))
                  if {(p[a], p[b], c) for a, b, c in es} == es)


@pytest.mark.parametrize("name", list(QUERIES))
def test_automorphisms_equal_brute_force(name):
    n, e = QUERIES[name]
    auts, less = fm.Query.from_edges(n, e).symmetry()
    assert sorted(auts) == brute_auts(n, e)
    assert auts[0] == list(range(n))
    assert (len(less) == 0) == (len(auts) == 1)


def small_graph(seed, V=9, m=22, labels=2, directed=False):
    rng = np.random.default_rng(seed)
    pairs = set()
    while len(pairs) < m:
        a, b = (int(x) for x in rng.integers(0, V, 2))
        if a != b:
            pairs.add((a, b) if directed else (min(a, b), max(a, b)))
    s, d, l = [], [], []
    for a, b in sorted(pairs):
        lab = int(rng.integers(0, labels))
        s.append(a); d.append(b); l.append(lab)
        if not directed:
            s.append(b); d.append(a); l.append(lab)
    for v in range(0, V, 3):  # a few self-loops of label 1
        s.append(v); d.append(v); l.append(1)
    return V, np.array(s, np.uint32), np.array(d, np.uint32), np.array(l, np.uint32)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("name", list(QUERIES))
def test_constraints_keep_one_embedding_per_orbit(seed, name):
    n, e = QUERIES[name]
    V, s, d, l = small_graph(seed, labels=1 if seed % 2 == 0 else 2, directed=name.startswith("directed"))
    og = orc.OracleGraph(V, s, d, l)
    full = orc.match(og, n, np.array(e, np.uint32).reshape(-1, 3)).table
    auts, less = fm.Query.from_edges(n, e).symmetry()
    kept = [r for r in full.tolist() if all(r[a] < r[b] for a, b in less)]
    assert len(kept) * len(auts) == len(full)
    expanded = sorted(tuple(r[p[u]] for u in range(n)) for r in kept for p in auts)
    assert expanded == sorted(map(tuple, full.tolist()))


def test_group_above_the_cap_is_not_listed():
    n = 6
    e = und([(a, b, 0) for a in range(n) for b in range(a + 1, n)])  # K6: 720 automorphisms
    auts, less = fm.Query.from_edges(n, e).symmetry(max_auts=48)
    assert auts == [] and less == []
    auts, less = fm.Query.from_edges(n, e).symmetry(max_auts=1000)
    assert len(auts) == 720
