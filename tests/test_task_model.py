"""CPU model of the matcher's task control flow (engine.hip process_task):
batches of 64 parents, 64-candidate steps over the flattened driver rows,
work splitting (parent halves and candidate halves) and yields that requeue
the remainder, with `skip` / `last_end` bounding a task's first and last
parent. Fuzzed with random row lengths and random split / yield decisions,
the invariants are that every (parent, candidate) pair is processed exactly
once and that no u32 arithmetic wraps. This is where a split that left a
task ending inside its last batch used to underflow a parent count."""
import random

import pytest

WAVE = 64
ALL = 0xFFFFFFFF


def u32(x):
    assert 0 <= x <= ALL, f"u32 wrap: {x}"
    return x


def process(task, lens, rng, queue, seen, p_split, p_yield, counted=None, min_cand=32 * WAVE):
    first, count, skip, last_end = task
    cnt_local, last = count, last_end
    b0 = 0
    while b0 < cnt_local:
        nb = min(WAVE, u32(cnt_local - b0))
        dstart, dlen, acct = [], [], []
        for lane in range(nb):
            p = first + b0 + lane
            endv = lens[p]
            if b0 + lane == cnt_local - 1:
                endv = min(endv, last)
            first_par = b0 == 0 and lane == 0
            startv = min(skip, endv) if first_par else 0
            dstart.append(startv)
            dlen.append(u32(endv - startv))
            # the instrumented launch's flags: bit 0 the range starts at
            # candidate 0, bit 1 the driver row is empty
            acct.append((0 if (first_par and skip) else 1) | (2 if lens[p] == 0 else 0))
        cum = []
        acc = 0
        for x in dlen:
            acc += x
            cum.append(acc)
        T = acc

        def account(t_fin):
            # engine.hip process_task `account`: a parent is counted by the
            # task whose processed candidates [0, t_fin) hold its candidate 0
            if counted is None:
                return
            for lane in range(nb):
                hi, lo = cum[lane], (cum[lane - 1] if lane else 0)
                fl = acct[lane]
                if (fl & 1) and ((hi <= t_fin) if (fl & 2) else (lo < hi and lo < t_fin)):
                    p = first + b0 + lane
                    counted[p] = counted.get(p, 0) + 1

        def parent_of(c):
            lo, hi = 0, nb - 1
            while lo < hi:
                mid = (lo + hi) >> 1
                if cum[mid] > c:
                    hi = mid
                else:
                    lo = mid + 1
            return lo

        base, step = 0, 0
        while base < T:
            if (step & 3) == 0 and rng.random() < p_split:
                rem_par = cnt_local - (b0 + nb) if cnt_local > b0 + nb else 0
                if rem_par >= 2 * WAVE:
                    keep = rem_par // 2
                    f = b0 + nb + keep
                    queue.append((first + f, u32(rem_par - keep), 0, last))
                    cnt_local, last = f, ALL
                elif T - base >= min_cand:
                    mid = base + ((T - base) // (2 * WAVE)) * WAVE
                    lo = parent_of(mid)
                    excl = cum[lo - 1] if lo else 0
                    abs_mid = u32(mid - excl) + (skip if (b0 == 0 and lo == 0) else 0)
                    queue.append((first + b0 + lo, u32(cnt_local - b0 - lo), abs_mid, last))
                    cnt_local, last, T = b0 + lo + 1, abs_mid, mid
            if rng.random() < p_yield:
                lo = parent_of(base)
                excl = cum[lo - 1] if lo else 0
                queue.append((first + b0 + lo, u32(cnt_local - b0 - lo),
                              u32(base - excl) + (skip if (b0 == 0 and lo == 0) else 0), last))
                account(base)
                return
            for lane in range(WAVE):
                tt = base + lane
                if tt < T:
                    pi = parent_of(tt)
                    excl = cum[pi - 1] if pi else 0
                    key = (first + b0 + pi, dstart[pi] + (tt - excl))
                    assert key not in seen, f"candidate processed twice: {key}"
                    seen.add(key)
            base += WAVE
            step += 1
        account(T)
        b0 += WAVE


# min_cand: the fewest remaining candidates a candidate split donates from
# (SPLIT_MIN_CANDIDATES in the engine; any value of at least 2 x 64 keeps the
# split point inside the remaining range)
@pytest.mark.parametrize("min_cand", [2 * WAVE, 32 * WAVE])
@pytest.mark.parametrize("seed", range(40))
def test_every_candidate_exactly_once(seed, min_cand):
    rng = random.Random(seed)
    nparents = rng.choice([1, 5, 63, 64, 65, 300, 1024])
    lens = [rng.choice([0, 1, 3, 10, 64, 65, 200, 700, 5000]) for _ in range(nparents)]
    queue = [(0, nparents, 0, ALL)]
    seen = set()
    counted = {}
    p_split, p_yield = rng.choice([(0.0, 0.0), (0.5, 0.0), (0.0, 0.2), (0.7, 0.1), (1.0, 0.05)])
    guard = 0
    while queue:
        guard += 1
        assert guard < 1_000_000
        t = queue.pop(rng.randrange(len(queue)))
        assert t[1] > 0
        process(t, lens, rng, queue, seen, p_split, p_yield, counted, min_cand)
    want = {(p, c) for p in range(nparents) for c in range(lens[p])}
    assert seen == want
    # the algorithmic counters: every parent's row words and lookups are
    # counted exactly once however the work was split and requeued
    assert counted == {p: 1 for p in range(nparents)}


def claim_roots(grid, active, total_chunks, rng, first_before_queue=True):
    """Model of the launch start (engine.hip match_waves / claim_root): the
    waves below `active` start at once and own chunk (workgroup id) as their
    first claim, without an atomic; root_cursor starts at `active`, so later
    claims (every wave, dormant ones included) take the chunks past the
    first round. Waves run in a random interleaving and a wave may find a
    queued task before its first claim. Returns the claimed chunk ids."""
    cursor = active
    claimed = []
    first = [w < active for w in range(grid)]
    queued = 0
    live = list(range(grid))
    while live:
        w = rng.choice(live)
        took = False
        if first_before_queue and first[w]:
            first[w] = False
            if w < total_chunks:
                claimed.append(w)
                took = True
        if not took and queued > 0 and rng.random() < 0.5:
            queued -= 1  # a task pushed by an earlier wave
            took = True
        if not took and not first_before_queue and first[w]:
            first[w] = False
            if w < total_chunks:
                claimed.append(w)
                took = True
        # a claim is only attempted if no task was popped (the old order
        # cleared the first-claim flag before that)
        if not first_before_queue:
            first[w] = False
        if not took:
            if cursor < total_chunks:
                claimed.append(cursor)
                cursor += 1
                took = True
        if took:
            queued += rng.randrange(0, 3)  # the chunk's children
        else:
            live.remove(w)
    return claimed


@pytest.mark.parametrize("seed", range(30))
def test_every_root_chunk_claimed_exactly_once(seed):
    rng = random.Random(seed)
    grid = rng.choice([1, 4, 16, 64])
    active = rng.choice([grid, max(1, grid // 2)])
    total = rng.choice([0, 1, active - 1, active, active + 1, 5 * grid])
    total = max(0, total)
    claimed = claim_roots(grid, active, total, rng)
    assert sorted(claimed) == list(range(total))


def test_popping_before_the_first_claim_loses_a_chunk():
    """The order the claim must not have: a starting wave that pops a queued
    task first and then claims from the cursor never takes its own chunk."""
    lost = 0
    for seed in range(50):
        rng = random.Random(seed)
        claimed = claim_roots(16, 16, 64, rng, first_before_queue=False)
        lost += len(set(range(64)) - set(claimed))
    assert lost > 0
