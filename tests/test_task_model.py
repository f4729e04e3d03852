"""CPU model of the matcher's task control flow (engine.hip process_task):
batches of 64 parents, 64-candidate steps over the flattened driver rows,
work splitting (parent halves and candidate halves) and yields that requeue
the remainder, with `skip` / `last_end` bounding a task's first and last
parent. Fuzzed with random row lengths and random split / yield decisions,
the invariant is that every (parent, candidate) pair is processed exactly
once and that no u32 arithmetic wraps. This is where a split that left a
task ending inside its last batch used to underflow a parent count."""
import random

import pytest

WAVE = 64
ALL = 0xFFFFFFFF


def u32(x):
    assert 0 <= x <= ALL, f"u32 wrap: {x}"
    return x


def process(task, lens, rng, queue, seen, p_split, p_yield):
    first, count, skip, last_end = task
    cnt_local, last = count, last_end
    b0 = 0
    while b0 < cnt_local:
        nb = min(WAVE, u32(cnt_local - b0))
        dstart, dlen = [], []
        for lane in range(nb):
            p = first + b0 + lane
            endv = lens[p]
            if b0 + lane == cnt_local - 1:
                endv = min(endv, last)
            startv = min(skip, endv) if (b0 == 0 and lane == 0) else 0
            dstart.append(startv)
            dlen.append(u32(endv - startv))
        cum = []
        acc = 0
        for x in dlen:
            acc += x
            cum.append(acc)
        T = acc

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
                elif T - base >= 8 * WAVE:
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
        b0 += WAVE


@pytest.mark.parametrize("seed", range(40))
def test_every_candidate_exactly_once(seed):
    rng = random.Random(seed)
    nparents = rng.choice([1, 5, 63, 64, 65, 300, 1024])
    lens = [rng.choice([0, 1, 3, 10, 64, 65, 200, 700, 5000]) for _ in range(nparents)]
    queue = [(0, nparents, 0, ALL)]
    seen = set()
    p_split, p_yield = rng.choice([(0.0, 0.0), (0.5, 0.0), (0.0, 0.2), (0.7, 0.1), (1.0, 0.05)])
    guard = 0
    while queue:
        guard += 1
        assert guard < 1_000_000
        t = queue.pop(rng.randrange(len(queue)))
        assert t[1] > 0
        process(t, lens, rng, queue, seen, p_split, p_yield)
    want = {(p, c) for p in range(nparents) for c in range(lens[p])}
    assert seen == want
