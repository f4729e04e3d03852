"""CPU model of the merge-path tiling of fm_intersect_kernel (engine.hip):
tiles of 2048 merge-path elements, per-thread 8-element stretches found by
binary search, the match check against the next B element (staged one past
the tile), and the block-ordered output. Checked against numpy on random and
edge-case pairs, including matches straddling tile and thread boundaries."""
import numpy as np
import pytest

IB, IVT = 256, 8
ITILE = IB * IVT


def merge_split(A, na, B, nb, diag):
    lo, hi = max(0, diag - nb), min(diag, na)
    while lo < hi:
        mid = (lo + hi) >> 1
        if A[mid] <= B[diag - 1 - mid]:
            lo = mid + 1
        else:
            hi = mid
    return lo


def intersect_model(A, B):
    if len(A) > len(B):
        A, B = B, A
    na, nb = len(A), len(B)
    out = []
    a0 = b0 = 0
    for d0 in range(0, na + nb, ITILE):
        d1 = min(d0 + ITILE, na + nb)
        a1 = merge_split(A, na, B, nb, d1)
        b1 = d1 - a1
        assert 0 <= a1 - a0 <= ITILE and 0 <= b1 - b0 <= ITILE
        sA = A[a0:a1]
        sB = B[b0:b1 + (1 if b1 < nb else 0)]
        ta, tb = a1 - a0, b1 - b0
        tbx = len(sB)
        per = []
        for tid in range(IB):
            dl = min(tid * IVT, ta + tb)
            steps = min(IVT, ta + tb - dl)
            i = merge_split(sA, ta, sB, tb, dl)
            j = dl - i
            m = []
            for _ in range(steps):
                if i < ta and (j >= tb or sA[i] <= sB[j]):
                    if j < tbx and sB[j] == sA[i]:
                        m.append(int(sA[i]))
                    i += 1
                else:
                    j += 1
            per.append(m)
        for m in per:
            out += m
        a0, b0 = a1, b1
    return np.array(out, dtype=np.uint32)


@pytest.mark.parametrize("seed", range(12))
def test_merge_path_model(seed):
    rng = np.random.default_rng(seed)
    for na, nb, span in [(0, 10, 50), (1, 1, 2), (5000, 5200, 20000), (3000, 3000, 3000), (2048, 2048, 4096),
                         (1, 9000, 9000), (4095, 4097, 9000)]:
        A = np.unique(rng.integers(0, span, na)).astype(np.uint32)
        B = np.unique(rng.integers(0, span, nb)).astype(np.uint32)
        got = intersect_model(A, B)
        assert np.array_equal(got, np.intersect1d(A, B))
