"""CPU model of the row membership probe of the matcher (engine.hip:
row_contains). The column array has a skip index skip[b] = col[16 b]. A row
of at most 16 words is read as one 20-word window starting at a 16-byte
boundary. A longer row first counts, in one 20-entry window of the skip
index (after binary steps on the skip index for rows of more than 17
blocks), the blocks starting inside the row whose first word is <= x; the
block after that count (or the row's first block) is read as the second
window. Checked against plain membership on random rows and edge cases
(empty rows, rows of 1..300 and thousands of words, x below, above, on and
between the row's values, rows starting at any alignment), and every word
the probe reads must lie inside the column array or the skip index plus
their 64-word padding. The paired probe (row_contains2, two candidates per
lane) is modelled step for step and must give the single probe's answers,
with its dummy reads (index 0 of either array) in bounds too. The window
tests are bit masks, as in the engine: a bit per window word (equal to x, or
at most x), masked once by the range's bits, which must fit the 20 words."""
import numpy as np
import pytest

COL_PAD = 64
SKIP_SHIFT = 4
SKIP_BLOCK = 1 << SKIP_SHIFT


def window(arr, lo, reads, name):
    w0 = lo & ~3
    for j in range(20):
        reads.append((name, w0 + j))
    return w0, [int(arr[w0 + j]) for j in range(20)]


def eq_bits(w, x):
    return sum(1 << j for j in range(20) if w[j] == x)


def le_bits(w, x):
    return sum(1 << j for j in range(20) if w[j] <= x)


def range_bits(r, n):
    """engine.hip range_bits: window positions [r, r + n), n <= 16, r <= 3."""
    assert 0 <= r <= 3 and 0 <= n <= 16
    return ((1 << n) - 1) << r


def probe(col, skip, st, n, x, reads):
    if n == 0:
        return False
    lo, end = st, st + n
    if n > SKIP_BLOCK:
        b0, b1 = st >> SKIP_SHIFT, (end - 1) >> SKIP_SHIFT
        blo, bn = b0 + 1, b1 - b0
        while bn > 16:
            half = bn >> 1
            reads.append(("skip", blo + half))
            if skip[blo + half] <= x:
                blo += half + 1
                bn -= half + 1
            else:
                bn = half
        s0, w = window(skip, blo, reads, "skip")
        assert blo - s0 == (blo & 3)
        c = bin(le_bits(w, x) & range_bits(blo & 3, bn)).count("1")
        bucket = blo - 1 + c
        lo = max(st, bucket << SKIP_SHIFT)
        end = min(end, (bucket + 1) << SKIP_SHIFT)
    w0, w = window(col, lo, reads, "col")
    return (eq_bits(w, x) & range_bits(lo & 3, end - lo)) != 0


def _arrays(rng, lens):
    """Strictly increasing rows back to back, padded; the skip index."""
    parts, starts, pos = [], [], 0
    for n in lens:
        parts.append(np.sort(rng.choice(8 * max(n, 1) + 40, size=n, replace=False)).astype(np.int64))
        starts.append(pos)
        pos += n
    ncol = pos
    col = np.concatenate(parts + [np.zeros(COL_PAD, np.int64)])
    nskip = (ncol + SKIP_BLOCK - 1) >> SKIP_SHIFT
    skip = np.concatenate([col[: nskip * SKIP_BLOCK: SKIP_BLOCK][:nskip], np.zeros(COL_PAD, np.int64)])
    return col, skip, starts, ncol, nskip


@pytest.mark.parametrize("seed", range(6))
def test_probe_matches_membership_and_stays_in_bounds(seed):
    rng = np.random.default_rng(seed)
    lens = [0, 1, 2, 3, 15, 16, 17, 18, 31, 32, 33, 255, 256, 257, 272, 273, 300] + \
        [int(v) for v in rng.integers(0, 4000, 12)]
    rng.shuffle(lens)
    col, skip, starts, ncol, nskip = _arrays(rng, lens)
    for st, n in zip(starts, lens):
        row = col[st:st + n]
        xs = {0, 10 ** 9}
        if n:
            xs |= {int(row[0]) - 1, int(row[-1]) + 1}
            xs |= {int(v) for v in row[:: max(1, n // 24)]} | {int(row[-1])}
            xs |= {int(v) + 1 for v in row[:: max(1, n // 12)]}
        rowset = set(int(v) for v in row)
        for x in xs:
            reads = []
            assert probe(col, skip, st, n, x, reads) == (x in rowset), (st, n, x)
            for name, i in reads:
                limit = ncol + COL_PAD if name == "col" else nskip + COL_PAD
                assert 0 <= i < limit, (name, i, limit, st, n, x)


def test_probe_planted_mistake_is_caught():
    """Picking the block one past the count (an off-by-one in `bucket`)
    misses values that the model finds."""
    rng = np.random.default_rng(3)
    col, skip, starts, ncol, nskip = _arrays(rng, [200])
    row = col[:200]

    def shifted(x):
        st, n = 0, 200
        b0, b1 = 0, (n - 1) >> SKIP_SHIFT
        blo, bn = 1, b1 - b0
        c = sum(1 for j in range(bn) if skip[blo + j] <= x)
        bucket = blo + c  # one too far
        lo, end = max(st, bucket << SKIP_SHIFT), min(n, (bucket + 1) << SKIP_SHIFT)
        return any(col[i] == x for i in range(lo, end))

    assert any(not shifted(int(x)) for x in row)
    assert all(probe(col, skip, 0, 200, int(x), []) for x in row)


def probe2(col, skip, sta, na, xa, stb, nb, xb, reads):
    """engine.hip row_contains2: the two probes of a lane in lockstep, both
    rows' windows requested together. A side without a skip step reads the
    skip index at 0 while the other steps; a row of length 0 (a lane with no
    probe on that side: st = 0) reads the column array at 0 and holds
    nothing."""
    loa, enda, lob, endb = sta, sta + na, stb, stb + nb
    sa, sb = na > SKIP_BLOCK, nb > SKIP_BLOCK
    if sa or sb:
        bloa = bna = blob = bnb = 0
        if sa:
            bloa, bna = (sta >> SKIP_SHIFT) + 1, ((enda - 1) >> SKIP_SHIFT) - (sta >> SKIP_SHIFT)
        if sb:
            blob, bnb = (stb >> SKIP_SHIFT) + 1, ((endb - 1) >> SKIP_SHIFT) - (stb >> SKIP_SHIFT)
        while bna > 16 or bnb > 16:
            hfa, hfb = bna >> 1, bnb >> 1
            ia, ib = (bloa + hfa if bna > 16 else 0), (blob + hfb if bnb > 16 else 0)
            reads += [("skip", ia), ("skip", ib)]
            va, vb = skip[ia], skip[ib]
            if bna > 16:
                if va <= xa:
                    bloa, bna = bloa + hfa + 1, bna - hfa - 1
                else:
                    bna = hfa
            if bnb > 16:
                if vb <= xb:
                    blob, bnb = blob + hfb + 1, bnb - hfb - 1
                else:
                    bnb = hfb
        sa0, wa = window(skip, bloa, reads, "skip")
        sb0, wb = window(skip, blob, reads, "skip")
        ca = bin(le_bits(wa, xa) & range_bits(bloa & 3, bna)).count("1")
        cb = bin(le_bits(wb, xb) & range_bits(blob & 3, bnb)).count("1")
        if sa:
            bucket = bloa - 1 + ca
            loa, enda = max(sta, bucket << SKIP_SHIFT), min(enda, (bucket + 1) << SKIP_SHIFT)
        if sb:
            bucket = blob - 1 + cb
            lob, endb = max(stb, bucket << SKIP_SHIFT), min(endb, (bucket + 1) << SKIP_SHIFT)
    a0, wa = window(col, loa, reads, "col")
    b0, wb = window(col, lob, reads, "col")
    ha = (eq_bits(wa, xa) & range_bits(loa & 3, enda - loa)) != 0
    hb = (eq_bits(wb, xb) & range_bits(lob & 3, endb - lob)) != 0
    return ha, hb


@pytest.mark.parametrize("seed", range(4))
def test_paired_probe_matches_membership_and_stays_in_bounds(seed):
    rng = np.random.default_rng(100 + seed)
    lens = [0, 1, 16, 17, 33, 272, 273, 300] + [int(v) for v in rng.integers(0, 3000, 10)]
    col, skip, starts, ncol, nskip = _arrays(rng, lens)
    rows = list(zip(starts, lens)) + [(0, 0)]  # (0, 0): a side with no probe

    def xs_of(st, n):
        row = col[st:st + n]
        xs = [0, 10 ** 9]
        if n:
            xs += [int(row[0]) - 1, int(row[-1]) + 1, int(row[-1])]
            xs += [int(v) for v in row[:: max(1, n // 6)]] + [int(v) + 1 for v in row[:: max(1, n // 4)]]
        return xs

    for _ in range(400):
        (sta, na), (stb, nb) = rows[rng.integers(len(rows))], rows[rng.integers(len(rows))]
        xa = int(rng.choice(xs_of(sta, na)))
        xb = int(rng.choice(xs_of(stb, nb)))
        reads = []
        ha, hb = probe2(col, skip, sta, na, xa, stb, nb, xb, reads)
        assert ha == (xa in set(int(v) for v in col[sta:sta + na])), (sta, na, xa)
        assert hb == (xb in set(int(v) for v in col[stb:stb + nb])), (stb, nb, xb)
        # the same answers as two single probes
        assert (ha, hb) == (probe(col, skip, sta, na, xa, []), probe(col, skip, stb, nb, xb, []))
        for name, i in reads:
            limit = ncol + COL_PAD if name == "col" else nskip + COL_PAD
            assert 0 <= i < limit, (name, i, limit)
