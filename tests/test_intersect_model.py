"""CPU model of the tiled intersection kernels (engine.hip:
fm_isect_cut_kernel, fm_isect_kernel): the host tile plan (merge tiles of
4864 merge-path elements, skew tiles of 256 elements of the shorter list when
the longer one is >= 32x longer), each tile's cut, the tile's LDS image (the
A slice from word 0 and the B slice from the next 256-word chunk, loaded a
chunk of one list per register at clamped indices), the per-thread 19-element
stretches found by binary search in the image, the branch-free steps that
every thread runs to the full 19 (past its stretch only B steps, which match
nothing), the match check against the next B element (staged one past the
tile), and the output offset of each tile as the sum of the counts of the
pair's earlier tiles (the running offset of the block that streams the pair).
Checked against numpy on random and edge-case pairs, including matches
straddling tile and thread boundaries."""
import numpy as np
import pytest

IB, XVT, ISKEW = 256, 19, 32
XT = IB * XVT


def merge_split(A, na, B, nb, diag):
    lo, hi = max(0, diag - nb), min(diag, na)
    while lo < hi:
        mid = (lo + hi) >> 1
        if A[mid] <= B[diag - 1 - mid]:
            lo = mid + 1
        else:
            hi = mid
    return lo


def plan(pairs):
    """tile_begin per pair and the pair of each tile (host plan)."""
    begin, tile_pair = [], []
    for p, (A, B) in enumerate(pairs):
        begin.append(len(tile_pair))
        na, nb = sorted((len(A), len(B)))
        if na == 0:
            continue
        n = -(-na // IB) if nb >= ISKEW * na else -(-(na + nb) // XT)
        tile_pair += [p] * n
    begin.append(len(tile_pair))
    return begin, tile_pair


def tile_matches(A, B, k, cuts, te_minus_tb):
    """Matches of tile k of a pair, in thread order (the block scan order)."""
    na, nb = len(A), len(B)
    if nb >= ISKEW * na:
        first, last = k * IB, min(k * IB + IB, na) - 1
        lo = int(np.searchsorted(B, A[first], "left"))
        hi = int(np.searchsorted(B, A[last], "right"))
        out = []
        for x in A[first:last + 1]:
            i = lo + int(np.searchsorted(B[lo:hi], x, "left"))
            if i < hi and B[i] == x:
                out.append(int(x))
        return out
    d0, d1 = k * XT, min(k * XT + XT, na + nb)
    a0 = cuts[k]
    a1 = cuts[k + 1] if k + 1 < te_minus_tb else na
    b0, b1 = d0 - a0, d1 - a1
    assert 0 <= a1 - a0 <= XT and 0 <= b1 - b0 <= XT
    sA = A[a0:a1]
    ta, tb = a1 - a0, b1 - b0
    tbx = tb + (1 if b1 < nb else 0)
    # the LDS image: register x of thread tid holds word x * IB + tid; the A
    # slice from word 0, the B slice (+ the next B element) from word
    # xa * IB; every load is a chunk of one list at a clamped index
    xa = -(-ta // IB)
    XR = XVT + 2
    assert xa + -(-tbx // IB) <= XR
    img = np.empty(XR * IB, dtype=np.int64)
    for x in range(XR):
        e = np.arange(IB)
        if x < xa:
            img[x * IB:(x + 1) * IB] = A[np.minimum(a0 + x * IB + e, na - 1)]
        else:
            img[x * IB:(x + 1) * IB] = B[np.minimum(b0 + (x - xa) * IB + e, nb - 1)]
    assert np.array_equal(img[:ta], sA) and np.array_equal(img[xa * IB:xa * IB + tbx], B[b0:b0 + tbx])
    sB_at = xa * IB
    out = []
    for tid in range(IB):
        dl = min(tid * XVT, ta + tb)
        i = merge_split(img, ta, img[sB_at:], tb, dl)
        j = dl - i
        # every thread runs XVT branch-free steps: one whose stretch ends at
        # the tile's end continues from (ta, tb) with B steps only
        av, bv = img[i], img[sB_at + j]
        for st in range(XVT):
            take_a = i < ta and (j >= tb or av <= bv)
            if take_a and j < tbx and bv == av:
                out.append(int(av))
            i += 1 if take_a else 0
            j += 0 if take_a else 1
            assert i <= ta and sB_at + j < XR * IB  # reads stay in the image
            if not take_a and j > tb:
                assert st >= min(XVT, ta + tb - dl)  # only past the stretch
            av, bv = img[i], img[sB_at + j]
    return out


def intersect_model(pairs):
    begin, tile_pair = plan(pairs)
    results = []
    for p, (A, B) in enumerate(pairs):
        if len(A) > len(B):
            A, B = B, A
        ntile = begin[p + 1] - begin[p]
        cuts = [merge_split(A, len(A), B, len(B), k * XT) for k in range(ntile)]
        out, pre = [], 0
        for k in range(ntile):
            m = tile_matches(A, B, k, cuts, ntile)
            assert len(out) == pre  # the running offset is the earlier tiles' total
            out += m
            pre += len(m)
        results.append(np.array(out, dtype=np.uint32))
    assert all(tile_pair[t] == p for p in range(len(pairs)) for t in range(begin[p], begin[p + 1]))
    return results


@pytest.mark.parametrize("seed", range(8))
def test_tiled_intersection_model(seed):
    rng = np.random.default_rng(seed)
    shapes = [(0, 10, 50), (1, 1, 2), (5000, 5200, 20000), (3000, 3000, 3000), (4096, 4096, 8192),
              (1, 9000, 9000), (8191, 8193, 18000), (300, 12000, 30000), (257, 9000, 9000), (20, 0, 5)]
    pairs = []
    for na, nb, span in shapes:
        A = np.unique(rng.integers(0, span, na)).astype(np.uint32)
        B = np.unique(rng.integers(0, span, nb)).astype(np.uint32)
        pairs.append((A, B) if seed % 2 else (B, A))
    got = intersect_model(pairs)
    for (A, B), g in zip(pairs, got):
        assert np.array_equal(g, np.intersect1d(A, B))


def test_tile_plan_counts():
    A = np.arange(10, dtype=np.uint32)
    pairs = [(A, np.arange(400, dtype=np.uint32)), (np.arange(5000, dtype=np.uint32), np.arange(5000, dtype=np.uint32)),
             (np.zeros(0, np.uint32), A)]
    begin, tile_pair = plan(pairs)
    # 400 >= 32 * 10: one skew tile; 10000 merge-path elements: 3 tiles; empty: none
    assert begin == [0, 1, 4, 4]
    assert tile_pair == [0, 1, 1, 1]


def size_class(nt):
    """isect_size_class: 16 * (bit length - 1) + the four bits after the
    leading one."""
    b = int(nt).bit_length() - 1
    frac = (nt >> (b - 4)) & 15 if b >= 4 else (nt << (4 - b)) & 15
    return 16 * b + frac


def test_size_classes_are_ordered():
    # monotone in the tile count, exact below 32, within 7% of each other
    # inside a class, and within 0..511 for every u32 count
    prev = -1
    for nt in list(range(1, 5000)) + [2**31 - 1, 2**31, 2**32 - 1]:
        c = size_class(nt)
        assert 0 <= c <= 511 and c >= prev
        assert nt >= 32 or c > prev
        prev = c
    for c in range(0, 16 * 14):
        members = [nt for nt in range(1, 1 << 14) if size_class(nt) == c]
        if members:
            assert max(members) < 1.07 * min(members)


def test_equal_pair_counts_per_block_balance_the_blocks():
    # the host's grid rule (every block takes the same number of pairs)
    # with blocks taking pairs largest first in size-class order, on the
    # benchmark's pair sizes: the busiest block ends within 5% of the mean
    import heapq
    rng = np.random.default_rng(7)
    nt = -(-(rng.integers(16384, 49152, 2048) + rng.integers(16384, 49152, 2048)) // XT)
    resident = 1280
    rounds = -(-len(nt) // resident)
    grid = -(-len(nt) // rounds)
    assert grid == 1024
    order = sorted(range(len(nt)), key=lambda p: -size_class(int(nt[p])))
    h = [(0, i) for i in range(grid)]
    for p in order:
        x, i = heapq.heappop(h)
        heapq.heappush(h, (x + int(nt[p]), i))
    assert max(x for x, _ in h) < 1.05 * nt.sum() / grid


def device_plan(lens, cap):
    """fm_isect_plan_kernel: tile_begin by a scan of the tile counts (each
    entry at most cap), the pairs with tiles ordered by the size class of
    their tile count, largest first, and the overflow flag (no work) when
    the tiles exceed cap."""
    counts = []
    for la, lb in lens:
        na, nb = sorted((la, lb))
        counts.append(0 if na == 0 else (-(-na // IB) if nb >= ISKEW * na else -(-(na + nb) // XT)))
    begin, acc = [], 0
    for c in counts:
        begin.append(min(acc, cap))
        acc += c
    begin.append(min(acc, cap))
    over = acc > cap
    buckets = {}
    for p, c in enumerate(counts):
        if c:
            buckets.setdefault(511 - size_class(c), []).append(p)
    order = [p for b in sorted(buckets) for p in buckets[b]]
    return begin, order, (0 if over else len(order)), over


def test_device_plan_equals_the_host_plan():
    rng = np.random.default_rng(9)
    pairs = []
    for _ in range(300):
        na, nb = (int(x) for x in rng.integers(0, 20000, 2))
        if rng.random() < 0.2:
            nb = na * 40
        pairs.append((np.zeros(na), np.zeros(nb)))
    lens = [(len(a), len(b)) for a, b in pairs]
    host_begin, host_tile_pair = plan(pairs)
    begin, order, nwork, over = device_plan(lens, len(host_tile_pair))
    assert begin == host_begin and not over and nwork == len(order)
    # every pair with tiles once, largest tile count (by size class) first
    assert sorted(order) == sorted(set(host_tile_pair))
    classes = [size_class(int(begin[p + 1] - begin[p])) for p in order]
    assert classes == sorted(classes, reverse=True)


def test_device_plan_flags_too_many_tiles_and_stays_in_the_array():
    lens = [(5000, 6000)] * 10
    begin, order, nwork, over = device_plan(lens, 20)
    assert over and nwork == 0
    assert all(b <= 20 for b in begin)  # cut[t] / cut2[t] are only read for t < cap
