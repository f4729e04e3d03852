"""CPU model of the segmented intersection kernels (isect.hip:
fm_isx_plan_kernel, fm_isx_kernel), checked against numpy.

Modelled as the device runs them: the segment counts and the plan's
first-segment table (clamped to the work array, overflow flagged); a wave's
pair lookup (64 samples per round); the two merge-path cuts of a segment
searched by half waves with 32 samples per round; a merge segment's A part
A[i0, i1) and B part B[j0, j1] (one B element past the cut); dense segments
(A part values spanning < 32,768 integers): chunks loaded as two 16-byte
loads per lane (lane l holds elements 4l..4l+3 and 256+4l..256+4l+3) through
a buffer range check (zeros past a chunk's end, masked by index in partial
chunks); each lane's group of four elements sets its bits with two ORs (the
word of the group's first element and the next one) and, when the group
spans further or its first element is out of range, one OR per element; the
A bitmap, and the bitmap of the B elements inside the A part's range, in
1,024-word regions whose words past the cleared range hold stale bits; the
count as the popcount of their AND, and the matches written from the AND
in 64-word rounds through a 512-word stage; sparse segments: one A chunk at
a time sorted in LDS (clamped copies of its last element after its end), B
windows searched there in 9 steps, chunk and window advanced by their last
values, counted first and searched again to write; skew segments (the
longer list >= 32x the shorter): the B range of a 512-element A chunk by two
32-ary searches and a branch-free binary search per element; and the
decoupled look-back that gives each segment the number of matches in its
pair's earlier segments, under random interleavings of the waves."""
import numpy as np
import pytest

W, K = 64, 8
CH = K * W
SEG = 16384
BMW = 1024
BMSPAN = BMW * 32
SKEW = 32
U32 = 0xFFFFFFFF


def segments(la, lb, seg=SEG, ch=CH):
    na, nb = min(la, lb), max(la, lb)
    if na == 0:
        return 0
    return -(-na // ch) if nb >= SKEW * na else -(-(na + nb) // seg)


def plan(lens, cap, seg=SEG, ch=CH):
    """fm_isx_plan_kernel: seg_begin (each entry at most cap), the total
    (0 when it exceeds cap) and the overflow flag."""
    begin, acc = [], 0
    for la, lb in lens:
        begin.append(min(acc, cap))
        acc += segments(la, lb, seg, ch)
    begin.append(min(acc, cap))
    over = acc > cap
    return begin, (0 if over else acc), over


def find_pair(seg_begin, npairs, g):
    """64 samples per round; the largest p < npairs with seg_begin[p] <= g."""
    lo, hi = 0, npairs
    while hi - lo > 1:
        step = -(-(hi - lo) // W)
        t = sum(1 for k in range(W) if k * step < hi - lo and seg_begin[lo + k * step] <= g)
        assert t >= 1
        lo, hi = lo + (t - 1) * step, min(lo + t * step, hi)
    return lo


def first_false(lo, hi, pred):
    """One half wave of first_false2: 32 samples per round."""
    rounds = 0
    while lo < hi:
        n, step = hi - lo, -(-(hi - lo) // 32)
        probes = [pred(lo + k * step) for k in range(32) if k * step < n]
        t = sum(probes)
        assert probes == [True] * t + [False] * (len(probes) - t)  # the predicate is monotone
        if t == 0:
            hi = lo
        else:
            lo, hi = lo + (t - 1) * step + 1, min(lo + t * step, hi)
        rounds += 1
    return lo, rounds


def merge_cut(A, B, d):
    """A elements among the first d of merge(A, B), A first on ties."""
    na, nb = len(A), len(B)
    return first_false(max(0, d - nb), min(d, na), lambda i: int(A[i]) <= int(B[d - 1 - i]))


def buffer_chunk(src, start, n):
    """load_chunk_x4: CH words from src[start:], zeros at or past n, as
    (lane, group, element) -> element index h*256 + 4*lane + e."""
    flat = np.zeros(CH, np.uint64)
    m = min(n, CH)
    if m > 0:
        flat[:m] = src[start:start + m]
    return flat


def set_bits(bm, flat, amin, span, n, check):
    """set_bits: per lane and group of four consecutive elements, two ORs
    at the group's first word and the next, single ORs for the rest."""
    for h in range(2):
        for l in range(W):
            idx = [h * (CH // 2) + 4 * l + e for e in range(4)]
            off = [(int(flat[i]) - amin) & U32 for i in idx]
            ok = [(not check or o <= span) and i < n for o, i in zip(off, idx)]
            wb = off[0] >> 5
            lo = hi = 0
            for e in range(4):
                rel = (off[e] - (wb << 5)) & U32
                near = ok[0] and rel < 64
                bit = 1 << (rel & 31)
                if ok[e] and near:
                    if rel < 32:
                        lo |= bit
                    else:
                        hi |= bit
                elif ok[e]:
                    bm[(off[e] >> 5) & (BMW - 1)] |= 1 << (off[e] & 31)
            if ok[0]:
                bm[wb & (BMW - 1)] |= lo
                bm[(wb + 1) & (BMW - 1)] |= hi


def clamped_chunk(src, start, n):
    """load_chunk: CH words, clamped to the last of n >= 1."""
    idx = np.minimum(np.arange(CH), n - 1)
    return src[start + idx].astype(np.uint64)


def dense_segment(pa, pb, amin, span, rng):
    """Matches of a dense segment as dense_emit writes them."""
    la, lb = len(pa), len(pb)
    bmA = [int(x) for x in rng.integers(0, 1 << 32, BMW, dtype=np.uint64)]  # stale bits
    bmB = [int(x) for x in rng.integers(0, 1 << 32, BMW, dtype=np.uint64)]
    nw = (span >> 5) + 1 if la else 0
    assert nw <= BMW
    for w in range(nw):
        bmA[w] = bmB[w] = 0
    for c in range(-(-la // CH)):
        set_bits(bmA, buffer_chunk(pa, c * CH, la - c * CH), amin, span, la - c * CH, False)
    if la == 0:
        return []
    for c in range(-(-lb // CH)):
        set_bits(bmB, buffer_chunk(pb, c * CH, lb - c * CH), amin, span, lb - c * CH, True)
    return [bmA[w] & bmB[w] for w in range(nw)]


def dense_emit(common, amin, o, lim):
    """dense_emit: 64-word rounds, each lane's values at its word's prefix
    offset in a CH-word stage, the stage flushed to o[base + lo, ...) in
    groups of four words (one 16-byte store per lane) and, for the group at
    the stage's end or at lim, word by word; nothing at or past lim is
    written. Returns the number of values."""
    nw = len(common)
    base = 0
    for w0 in range(0, nw, W):
        words = [common[w] if w < nw else 0 for w in range(w0, w0 + W)]
        counts = [bin(x).count("1") for x in words]
        tot = sum(counts)
        incl = np.cumsum(counts)
        for lo in range(0, tot, CH):
            stage = [None] * CH
            for l in range(W):
                pos = int(incl[l]) - counts[l]
                b = words[l]
                while b:
                    if lo <= pos < lo + CH:
                        stage[pos - lo] = amin + 32 * (w0 + l) + ((b & -b).bit_length() - 1)
                    pos += 1
                    b &= b - 1
            cnt, at0 = min(CH, tot - lo), base + lo
            assert None not in stage[:cnt]
            for l in range(W):
                for i in range(4 * l, cnt, 4 * W):
                    at = at0 + i
                    if i + 4 <= cnt and at + 4 <= lim:
                        o[at:at + 4] = stage[i:i + 4]
                    else:
                        for e in range(4):
                            if i + e < cnt and at + e < lim:
                                o[at + e] = stage[i + e]
        base += tot
    return base


def sparse_segment(pa, pb):
    la, lb = len(pa), len(pb)
    ia = jb = 0
    out = []
    a_has, b_has = la > 0, lb > 0
    sorted_chunk = amin = amax = bmax = b = None

    def next_a():
        ch = clamped_chunk(pa, ia, la - ia)
        return ch, int(ch[0]), int(ch[-1])

    def next_b():
        ch = clamped_chunk(pb, jb, lb - jb)
        return ch, int(ch[-1])

    if a_has:
        sorted_chunk, amin, amax = next_a()
    if b_has:
        b, bmax = next_b()
    while a_has or b_has:
        if a_has and b_has:
            n = lb - jb
            for k in range(CH):
                x = int(b[k])
                pos, h = 0, CH // 2
                while h >= 1:
                    if int(sorted_chunk[pos + h - 1]) < x:
                        pos += h
                    h >>= 1
                assert pos < CH
                if amin <= x <= amax and k < n and int(sorted_chunk[pos]) == x:
                    out.append(x)
        adv_a, adv_b = a_has, b_has
        if a_has and b_has:
            adv_b, adv_a = bmax <= amax, amax <= bmax
        if adv_b:
            jb += CH
            b_has = jb < lb
            if b_has:
                b, bmax = next_b()
        if adv_a:
            ia += CH
            a_has = ia < la
            if a_has:
                sorted_chunk, amin, amax = next_a()
    return out


def skew_segment(pa, B):
    la, nb = len(pa), len(B)
    a = clamped_chunk(pa, 0, la)
    amin, amax = int(a[0]), int(a[-1])
    lo, _ = first_false(0, nb, lambda j: int(B[j]) < amin)
    hi, _ = first_false(0, nb, lambda j: int(B[j]) <= amax)
    out = []
    if lo < hi:
        for k in range(CH):
            base, n = lo, hi - lo
            while n > 1:
                h = n >> 1
                assert base + h < hi
                if int(B[base + h]) <= int(a[k]):
                    base += h
                n -= h
            if k < la and int(B[base]) == int(a[k]):
                out.append(int(a[k]))
    return out


def run_segment(A, B, s, rng, stats):
    na, nb = len(A), len(B)
    if nb >= SKEW * na:
        first = s * CH
        stats["skew"] += 1
        return skew_segment(A[first:first + min(CH, na - first)], B)
    d0, d1 = s * SEG, min(s * SEG + SEG, na + nb)
    (i0, r0), (i1, r1) = merge_cut(A, B, d0), merge_cut(A, B, d1)
    stats["rounds"] = max(stats["rounds"], r0, r1)
    j0, j1 = d0 - i0, d1 - i1
    jx = min(j1 + 1, nb)
    pa, pb = A[i0:i1].astype(np.uint64), B[j0:jx].astype(np.uint64)
    amin = int(pa[0]) if len(pa) else 0
    span = int(pa[-1]) - amin if len(pa) else 0
    if span < BMSPAN:
        stats["dense"] += 1
        common = dense_segment(pa, pb, amin, span, rng)
        return ("bits", common, amin)
    stats["sparse"] += 1
    return sparse_segment(pa, pb)


def intersect_model(pairs, rng):
    """Per pair: the matches of its segments concatenated in segment order,
    each segment stored at the offset the look-back gives it."""
    lens = [(len(a), len(b)) for a, b in pairs]
    seg_begin, total, over = plan(lens, 1 << 30)
    assert not over
    stats = {"dense": 0, "sparse": 0, "skew": 0, "rounds": 0}
    per_seg = []
    for g in range(total):
        p = find_pair(seg_begin, len(pairs), g)
        assert seg_begin[p] <= g < seg_begin[p + 1]
        A, B = pairs[p]
        if len(A) > len(B):
            A, B = B, A
        per_seg.append((p, run_segment(A, B, g - seg_begin[p], rng, stats)))
    def n_of(m):
        return sum(bin(x).count("1") for x in m[1]) if isinstance(m, tuple) else len(m)

    prefix = look_back_model([n_of(m) for _, m in per_seg], seg_begin, rng, restart=bool(rng.integers(2)))
    results = []
    for p, (A, B) in enumerate(pairs):
        room = min(len(A), len(B))
        out = np.full(room, U32, np.uint64)
        count = 0
        for g in range(seg_begin[p], seg_begin[p + 1]):
            m = per_seg[g][1]
            lim = room - prefix[g] if prefix[g] < room else 0
            if isinstance(m, tuple):
                view = out[prefix[g]:]  # writes past lim would fail the room check below
                written = np.full(len(view), U32, np.uint64)
                n = dense_emit(m[1], m[2], written, lim)
                assert n == n_of(m) and np.all(written[lim:] == U32)
                view[:lim] = np.where(written[:lim] != U32, written[:lim], view[:lim])
            else:
                n = len(m)
                out[prefix[g]:prefix[g] + min(n, lim)] = m[:min(n, lim)]
            count = prefix[g] + n
        results.append(out[:count].astype(np.uint32))
    return results, stats


def look_back_model(counts, seg_begin, rng, restart=False):
    """Waves run segments concurrently in claim order; each publishes its
    count, then reads up to 64 predecessors per round until it meets an
    inclusive sum, then publishes its own. Random interleaving; returns every
    segment's exclusive prefix within its pair. With restart, an attempt
    that meets an unpublished predecessor starts again from g - 1 (the
    emitters' look-back that does not wait)."""
    n = len(counts)
    pair_first = {}
    for p in range(len(seg_begin) - 1):
        for g in range(seg_begin[p], seg_begin[p + 1]):
            pair_first[g] = seg_begin[p]
    status = [None] * n  # None, ("A", c) or ("I", sum)
    prefix = [None] * n
    # each wave: state 0 = counting, 1 = looking back, 2 = done
    state = [0] * n
    look = [None] * n  # (t, acc)
    started = 0
    active = []
    while started < n or active:
        if started < n and (not active or rng.random() < 0.4):
            active.append(started)
            started += 1
        g = active[rng.integers(len(active))]
        sb = pair_first[g]
        if state[g] == 0:
            if g == sb:
                status[g] = ("I", counts[g])
                prefix[g] = 0
                state[g] = 2
                active.remove(g)
            else:
                status[g] = ("A", counts[g])
                state[g] = 1
                look[g] = (g - 1, 0)
        elif state[g] == 1:
            t, acc = look[g]
            window = [status[t - l] if l <= t - sb else ("I", 0) for l in range(W)]
            incl = [i for i, s in enumerate(window) if s is not None and s[0] == "I"]
            waiting = [i for i, s in enumerate(window) if s is None]
            if incl:
                f = incl[0]
                if not [i for i in waiting if i <= f]:
                    prefix[g] = acc + sum(window[i][1] for i in range(f + 1))
                    status[g] = ("I", prefix[g] + counts[g])
                    state[g] = 2
                    active.remove(g)
                elif restart:
                    look[g] = (g - 1, 0)
            elif not waiting:
                look[g] = (t - W, acc + sum(s[1] for s in window))
            elif restart:
                look[g] = (g - 1, 0)
    for g in range(n):
        sb = pair_first[g]
        assert prefix[g] == sum(counts[sb:g])
    return prefix


def make_pairs(rng, shapes):
    pairs = []
    for na, nb, span in shapes:
        A = np.unique(rng.integers(0, span, na, dtype=np.int64)).astype(np.uint32)
        B = np.unique(rng.integers(0, span, nb, dtype=np.int64)).astype(np.uint32)
        pairs.append((A, B))
    return pairs


@pytest.mark.parametrize("seed", range(4))
def test_segmented_intersection_model(seed):
    rng = np.random.default_rng(seed)
    shapes = [(0, 10, 50), (1, 1, 2), (5000, 5200, 20000), (3000, 3000, 3000), (4096, 4096, 8192),
              (1, 9000, 9000), (8191, 8193, 18000), (300, 12000, 30000), (257, 9000, 9000), (20, 0, 5),
              (6000, 7000, 2_000_000), (2000, 2500, 10 ** 9), (700, 30000, 30000), (12000, 11000, 40000),
              (20000, 25000, 60000)]
    pairs = make_pairs(rng, shapes)
    if seed % 2:
        pairs = [(b, a) for a, b in pairs]
    got, stats = intersect_model(pairs, rng)
    for (A, B), g in zip(pairs, got):
        assert np.array_equal(g, np.intersect1d(A, B))
    # all three segment kinds ran, and the cut searches took few rounds
    assert stats["dense"] and stats["sparse"] and stats["skew"]
    assert stats["rounds"] <= 4


def test_edge_values_and_equal_lists():
    rng = np.random.default_rng(5)
    top = np.uint32(U32)
    pairs = [(np.arange(9000, dtype=np.uint32), np.arange(9000, dtype=np.uint32)),
             (np.array([0, 5, 7, top], np.uint32), np.array([0, 6, 7, top], np.uint32)),
             (np.arange(U32 - 6000, U32, 2, dtype=np.int64).astype(np.uint32),
              np.append(np.arange(U32 - 7000, U32, 3, dtype=np.int64), U32).astype(np.uint32)),
             # a value 0 in B against an A part starting at 0 (zeros read past a chunk's end are masked)
             (np.arange(0, 700, 1, dtype=np.uint32), np.arange(0, 1400, 2, dtype=np.uint32)),
             # dense A part next to a sparse one in the same pair
             (np.concatenate([np.arange(3000, dtype=np.uint32), np.arange(10 ** 8, 10 ** 8 + 3000 * 50000, 50000,
                                                                         dtype=np.int64).astype(np.uint32)]),
              np.arange(0, 10 ** 8 + 3000 * 50000, 25000, dtype=np.int64).astype(np.uint32))]
    got, stats = intersect_model(pairs, rng)
    for (A, B), g in zip(pairs, got):
        assert np.array_equal(g, np.intersect1d(A, B))


def test_benchmark_shape_segments_are_dense():
    """The benchmark's lists (gaps 1..4) make every merge segment dense: the
    A part of 16,384 path elements spans fewer than 32,768 values."""
    rng = np.random.default_rng(7)
    stats_total = {"dense": 0, "sparse": 0}
    for _ in range(3):
        la, lb = (int(x) for x in rng.integers(16384, 49152, 2))
        A = np.cumsum(rng.integers(1, 5, la)).astype(np.uint32)
        B = np.cumsum(rng.integers(1, 5, lb)).astype(np.uint32)
        got, stats = intersect_model([(A, B)], rng)
        assert np.array_equal(got[0], np.intersect1d(A, B))
        stats_total["dense"] += stats["dense"]
        stats_total["sparse"] += stats["sparse"]
    assert stats_total["sparse"] == 0 and stats_total["dense"] > 0


def test_segment_counts_and_plan():
    lens = [(10, 400), (9000, 9000), (0, 10), (4096, 0), (1, 31)]
    # 400 >= 32 * 10: one skew segment; 18,000 path elements: 2 segments; empty lists: none; 31 < 32: merge
    assert [segments(a, b) for a, b in lens] == [1, 2, 0, 0, 1]
    begin, total, over = plan(lens, 100)
    assert begin == [0, 1, 3, 3, 3, 4] and total == 4 and not over
    begin, total, over = plan(lens, 3)
    assert over and total == 0 and all(b <= 3 for b in begin)


def test_find_pair_equals_searchsorted():
    rng = np.random.default_rng(3)
    for npairs in (1, 2, 63, 64, 65, 2048, 5000):
        nseg = rng.integers(0, 5, npairs)
        begin = np.concatenate([[0], np.cumsum(nseg)]).tolist()
        for g in rng.integers(0, max(begin[-1], 1), 50):
            if begin[-1] == 0:
                break
            p = find_pair(begin, npairs, int(g))
            assert p == int(np.searchsorted(begin, g, side="right")) - 1
            assert begin[p] <= g < begin[p + 1]


def test_first_false_equals_bisect():
    rng = np.random.default_rng(4)
    for _ in range(300):
        lo = int(rng.integers(0, 1000))
        hi = lo + int(rng.integers(0, 200000))
        split = int(rng.integers(lo, hi + 1))
        got, rounds = first_false(lo, hi, lambda i: i < split)
        assert got == split
        assert rounds <= 4 or hi - lo > 32 ** 4


def test_merge_cuts_match_the_merge():
    rng = np.random.default_rng(6)
    for _ in range(50):
        A = np.unique(rng.integers(0, 5000, int(rng.integers(0, 3000))))
        B = np.unique(rng.integers(0, 5000, int(rng.integers(0, 3000))))
        merged = sorted([(int(x), 0) for x in A] + [(int(x), 1) for x in B])  # A first on ties
        for d in rng.integers(0, len(A) + len(B) + 1, 10):
            i, _ = merge_cut(A, B, int(d))
            assert i == sum(1 for _, src in merged[:int(d)] if src == 0)


def claim_seg(k, cls, ncls, run):
    """Segment of class cls's k-th claim: runs of `run` consecutive segments
    dealt to the classes in turn (fm_isx_kernel's claim mapping)."""
    return ((k // run) * ncls + cls) * run + k % run


def block_protocol_model(counts, seg_begin, rng, nblocks=3, nb=2, ns=2, ne=1, ncls=2, run=2):
    """Builder waves claim segments (home class = block % ncls; after the
    home class a class's counter is read before it is claimed from), build
    them for a random number of steps, publish the count (the inclusive sum
    for a pair's first segment), then hand the segment to a free slot of
    their block, skip it (a zero count that is not its pair's last), or find
    its offset themselves (waiting). Emitter waves take ready slots and try
    the look-back without waiting, making the slot ready again when it
    cannot finish, and end when their block's builders have ended with no
    slot ready. Every wave is scheduled at random; returns the offsets found
    (None for skipped segments) and fails if the waves stop making progress."""
    total = len(counts)
    sb_of, last_of = [0] * total, [False] * total
    for p in range(len(seg_begin) - 1):
        for g in range(seg_begin[p], seg_begin[p + 1]):
            sb_of[g], last_of[g] = seg_begin[p], g == seg_begin[p + 1] - 1
    ctr = [0] * ncls
    status = [None] * total
    prefix = [None] * total
    slots = [[None] * ns for _ in range(nblocks)]  # None (free) or ["ready" | "emitting", g]
    done = [0] * nblocks
    waves = []
    for b in range(nblocks):
        waves += [{"kind": "b", "block": b, "cls": b % ncls, "tried": 0, "state": "claim"} for _ in range(nb)]
        waves += [{"kind": "e", "block": b, "state": "run"} for _ in range(ne)]

    def attempt(g):
        acc = 0
        for t in range(g - 1, sb_of[g] - 1, -1):
            if status[t] is None:
                return None
            if status[t][0] == "I":
                return acc + status[t][1]
            acc += status[t][1]
        return acc

    def record(g, pre):
        assert prefix[g] is None
        prefix[g] = pre
        if g != sb_of[g]:
            status[g] = ("I", pre + counts[g])

    steps, limit = 0, 400 * (total + 1) * len(waves)
    while True:
        live = [w for w in waves if w["state"] != "end"]
        if not live:
            break
        steps += 1
        assert steps < limit, "no progress"
        w = live[rng.integers(len(live))]
        b = w["block"]
        if w["kind"] == "b":
            if w["state"] == "claim":
                if w["tried"] >= ncls:
                    w["state"] = "end"
                    done[b] += 1
                    continue
                c = w["cls"]
                spent = w["tried"] > 0 and claim_seg(ctr[c], c, ncls, run) >= total
                g = total
                if not spent:
                    g = claim_seg(ctr[c], c, ncls, run)
                    ctr[c] += 1
                if g >= total:
                    w["cls"], w["tried"] = (c + 1) % ncls, w["tried"] + 1
                    continue
                w.update(g=g, state="build", steps=int(rng.integers(1, 5)))
            elif w["state"] == "build":
                w["steps"] -= 1
                if w["steps"]:
                    continue
                g = w["g"]
                status[g] = ("I" if g == sb_of[g] else "A", counts[g])
                if counts[g] == 0 and not last_of[g]:
                    w["state"] = "claim"
                    continue
                free = [i for i in range(ns) if slots[b][i] is None]
                if free:
                    slots[b][free[0]] = ["ready", g]
                    w["state"] = "claim"
                else:
                    w["state"] = "fallback"
            elif w["state"] == "fallback":
                pre = attempt(w["g"])
                if pre is not None:
                    record(w["g"], pre)
                    w["state"] = "claim"
        else:
            ready = [i for i in range(ns) if slots[b][i] is not None and slots[b][i][0] == "ready"]
            if not ready:
                if done[b] == nb and all(s is None or s[0] != "ready" for s in slots[b]):
                    w["state"] = "end"
                continue
            i = ready[rng.integers(len(ready))]
            g = slots[b][i][1]
            slots[b][i][0] = "emitting"
            pre = attempt(g)
            if pre is None:
                slots[b][i][0] = "ready"
            else:
                record(g, pre)
                slots[b][i] = None
    assert all(s is None for blk in slots for s in blk)
    return prefix


@pytest.mark.parametrize("seed", range(12))
def test_builder_emitter_protocol_finds_every_offset(seed):
    rng = np.random.default_rng(100 + seed)
    nseg = rng.integers(1, 7, 9)  # segments per pair
    seg_begin = np.concatenate([[0], np.cumsum(nseg)]).tolist()
    total = seg_begin[-1]
    counts = [int(x) if rng.random() < 0.7 else 0 for x in rng.integers(0, 50, total)]
    params = dict(nblocks=int(rng.integers(1, 4)), nb=int(rng.integers(1, 3)), ns=int(rng.integers(1, 3)),
                  ne=int(rng.integers(1, 3)), run=int(rng.integers(1, 4)))
    params["ncls"] = int(rng.integers(1, params["nblocks"] + 1))  # every class has a home block
    prefix = block_protocol_model(counts, seg_begin, rng, **params)
    for p in range(len(seg_begin) - 1):
        for g in range(seg_begin[p], seg_begin[p + 1]):
            if counts[g] == 0 and g != seg_begin[p + 1] - 1:
                assert prefix[g] is None
            else:
                assert prefix[g] == sum(counts[seg_begin[p]:g])


def test_claim_mapping_covers_every_segment_once_in_class_order():
    for ncls, run, total in ((1, 16, 100), (8, 16, 9149), (3, 2, 50), (8, 1, 17)):
        seen = []
        for c in range(ncls):
            segs = []
            k = 0
            while claim_seg(k, c, ncls, run) < total:
                segs.append(claim_seg(k, c, ncls, run))
                k += 1
            assert segs == sorted(segs)
            seen += segs
        assert sorted(seen) == list(range(total))
