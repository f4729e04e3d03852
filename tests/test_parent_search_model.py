"""CPU model of the matcher's candidate-to-parent search (engine.hip
process_task, the `prefetch` step): a batch holds up to 64 parents whose
driver-row lengths are scanned into cum[0..63] (inclusive; lanes past the
batch add 0, so their entries hold the batch total). Candidate tt belongs to
the first parent p with cum[p] > tt and starts at offset tt - cum[p - 1] of
its row. The kernel finds p with three rounds of four LDS reads; the model
replays exactly those reads and checks them against bisect over every tt,
including zero-length parents, a total shrunk by a candidate split, and
that every read index stays in [0, 64) with the 16-byte read aligned."""
import bisect
import random

import pytest

WAVE = 64


def search(cum, tt, reads, last_le=lambda v, t: v <= t):
    """The kernel's three-round search; returns (parent, exclusive start)."""
    def rd(i):
        assert 0 <= i < WAVE
        reads.append(i)
        return cum[i]
    lo = 0
    for j in range(4):
        lo += 16 if rd(16 * j + 15) <= tt else 0
    lo4 = lo
    for j in range(4):
        lo4 += 4 if rd(lo + 4 * j + 3) <= tt else 0
    assert lo4 % 4 == 0  # the uint4 read of cum[lo4 .. lo4 + 3]
    c4 = [rd(lo4 + i) for i in range(4)]
    excl = rd(lo4 - 1) if lo4 else 0
    lo = lo4
    for v in c4:
        if last_le(v, tt):
            lo += 1
            excl = v
    return lo, excl


def batch_cum(rng, nb):
    lens = []
    for _ in range(nb):
        r = rng.random()
        lens.append(0 if r < 0.25 else rng.randint(1, 3) if r < 0.6 else rng.randint(1, 200))
    cum, acc = [], 0
    for x in lens + [0] * (WAVE - nb):
        acc += x
        cum.append(acc)
    return lens, cum


@pytest.mark.parametrize("seed", range(40))
def test_search_matches_bisect(seed):
    rng = random.Random(seed)
    nb = rng.choice([1, 2, 3, 15, 16, 17, 31, 48, 63, 64, rng.randint(1, 64)])
    lens, cum = batch_cum(rng, nb)
    total = cum[-1]
    if total == 0:
        return
    # a candidate split may shrink the batch's total; cum keeps the original
    T = total if rng.random() < 0.5 else rng.randint(1, total)
    for tt in range(T):
        reads = []
        lo, excl = search(cum, tt, reads)
        want = bisect.bisect_right(cum, tt)
        assert lo == want < nb
        assert excl == (cum[want - 1] if want else 0)
        assert 0 <= tt - excl < lens[lo]
        assert all(0 <= i < WAVE for i in reads)


def test_planted_mistake_is_caught():
    # the model with its last round comparing < instead of <= must disagree
    # with bisect somewhere: the check above is not vacuous
    rng = random.Random(7)
    _, cum = batch_cum(rng, 64)
    wrong = sum(search(cum, tt, [], last_le=lambda v, t: v < t)[0] != bisect.bisect_right(cum, tt)
                for tt in range(cum[-1]))
    assert wrong > 0
