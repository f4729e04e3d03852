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
their 64-word padding."""
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
        c = sum(1 for j in range(20) if 0 <= s0 + j - blo < bn and w[j] <= x)
        bucket = blo - 1 + c
        lo = max(st, bucket << SKIP_SHIFT)
        end = min(end, (bucket + 1) << SKIP_SHIFT)
    w0, w = window(col, lo, reads, "col")
    return any(0 <= w0 + j - lo < end - lo and w[j] == x for j in range(20))


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
