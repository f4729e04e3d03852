"""CPU model of the row membership probe of the matcher (engine.hip:
row_contains): binary steps narrow the lower-bound position of x in the
sorted row col[st, st + len) to at most 13 places, then one 16-word window
starting at a 16-byte boundary is compared. Checked against plain
membership on random rows and the edge cases (empty rows, rows of 1..40
elements, x below, above and between the row's values, rows at the start
and the end of the column array), and every word the probe reads must lie
inside the column array plus its 16-word padding."""
import numpy as np
import pytest

COL_PAD = 16


def probe(col, st, n_len, x, reads):
    lo, n = st, n_len
    while n > 12:
        half = n >> 1
        reads.append(lo + half)
        if col[lo + half] < x:
            lo += half + 1
            n -= half + 1
        else:
            n = half
    w0 = lo & ~3
    hit = False
    for j in range(16):
        reads.append(w0 + j)
        idx = w0 + j
        if 0 <= idx - st < n_len and col[idx] == x:
            hit = True
    return hit


def _column(rng, rows):
    """Sorted rows of the given lengths laid out back to back, padded."""
    parts, starts = [], []
    pos = 0
    for n in rows:
        vals = np.sort(rng.choice(10 * max(n, 1) + 50, size=n, replace=False)).astype(np.int64)
        parts.append(vals)
        starts.append(pos)
        pos += n
    col = np.concatenate(parts + [np.zeros(COL_PAD, np.int64)]) if parts else np.zeros(COL_PAD, np.int64)
    return col, starts, pos


@pytest.mark.parametrize("seed", range(6))
def test_probe_matches_membership_and_stays_in_bounds(seed):
    rng = np.random.default_rng(seed)
    lens = [0, 1, 2, 3, 4, 11, 12, 13, 14, 15, 16, 17, 31, 32, 33, 40] + list(rng.integers(0, 3000, 20))
    rng.shuffle(lens)
    col, starts, ncol = _column(rng, lens)
    for st, n in zip(starts, lens):
        row = col[st:st + n]
        xs = set(int(v) for v in row[:: max(1, n // 16)]) if n else set()
        if n:
            xs |= {int(row[0]) - 1, int(row[-1]) + 1, int(row[0]), int(row[-1])}
            xs |= {int(v) + 1 for v in row[:: max(1, n // 8)]}
        xs |= {0, 10 ** 9}
        rowset = set(int(v) for v in row)
        for x in xs:
            reads = []
            assert probe(col, st, n, x, reads) == (x in rowset), (st, n, x)
            assert min(reads) >= 0 and max(reads) < ncol + COL_PAD, (st, n, x, max(reads), ncol)


def test_probe_planted_mistake_is_caught():
    """A window that stops one word short misses x at the last place."""
    rng = np.random.default_rng(1)
    col, starts, _ = _column(rng, [29])
    row = col[:29]

    def short_probe(x):
        lo, n = 0, 29
        while n > 12:
            half = n >> 1
            if col[lo + half] < x:
                lo += half + 1
                n -= half + 1
            else:
                n = half
        w0 = lo & ~3
        return any(0 <= w0 + j < 29 and col[w0 + j] == x for j in range(lo + n - w0))  # excludes lo + n

    misses = [int(x) for x in row if not short_probe(int(x))]
    assert misses, "the shortened window should miss some present value"
