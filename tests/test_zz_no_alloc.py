"""fm_intersect_pairs_noalloc (the entry point that allocates nothing: its
claim counters, status words and segment table live in the library's
fixed device array)
against fm_intersect_pairs, through the C ABI: the same counts and the same
intersections, bit for bit, on two calls in a row, the second with fewer
and shorter pairs (so state left in the fixed array by the first call must
not leak into the second), and the limit refused with a clear error."""
import ctypes as C

import numpy as np
import pytest

import forkmatch_amd as fm
from forkmatch_amd import _lib as L

pytestmark = pytest.mark.gpu


def _pairs(rng, npairs, max_len):
    lists, a_off, a_len, b_off, b_len = [], [], [], [], []
    pos = 0
    for i in range(npairs):
        kind = i % 4
        na = int(rng.integers(0 if i % 13 == 0 else 1, max_len))
        nb = int(rng.integers(1, max_len)) if kind != 3 else 40 * max(na, 1)  # kind 3: skewed
        span = max(16, int(2.5 * max(na, nb))) if kind != 2 else 10 ** 9     # kind 2: sparse values
        a = np.unique(rng.integers(0, span, na, dtype=np.int64)).astype(np.uint32)
        b = np.unique(rng.integers(0, span, nb, dtype=np.int64)).astype(np.uint32)
        a_off.append(pos), a_len.append(len(a)), lists.append(a)
        pos += len(a)
        b_off.append(pos), b_len.append(len(b)), lists.append(b)
        pos += len(b)
    data = np.concatenate(lists).astype(np.uint32)
    room = [min(x, y) for x, y in zip(a_len, b_len)]
    out_off = np.concatenate([[0], np.cumsum(room)]).astype(np.uint64)
    return data, [np.array(a_off, np.uint64), np.array(a_len, np.uint32), np.array(b_off, np.uint64),
                  np.array(b_len, np.uint32)], out_off


def _run(engine, fn, data, arrs, out_off):
    """One call of `fn` through the C ABI on device copies; (counts, out)."""
    npairs = len(arrs[1])
    bufs = [fm.DeviceBuffer.from_array(x) for x in [data] + arrs]
    counts = fm.DeviceBuffer(4 * max(npairs, 1))
    oo = fm.DeviceBuffer.from_array(out_off[:-1])
    words = max(int(out_off[-1]), 1)
    out = fm.DeviceBuffer.from_array(np.full(words, 0xDEADBEEF, np.uint32))  # stale words must stay stale
    ms = C.c_double()
    rc = fn(engine._h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, bufs[4].ptr, npairs, counts.ptr, oo.ptr,
            out.ptr, C.byref(ms))
    assert rc == 0, L.lib().fm_last_error()
    return counts.to_array(np.uint32, npairs), out.to_array(np.uint32, words)


def test_noalloc_equals_allocating_entry_point_on_two_calls():
    lib = L.lib()
    rng = np.random.default_rng(2024)
    with fm.Engine(soln_prealloc=1 << 24, output_prealloc=1 << 24) as e:
        for npairs, max_len in ((700, 9000), (90, 3000)):  # the second call is smaller
            data, arrs, out_off = _pairs(rng, npairs, max_len)
            ref_c, ref_o = _run(e, lib.fm_intersect_pairs, data, arrs, out_off)
            got_c, got_o = _run(e, lib.fm_intersect_pairs_noalloc, data, arrs, out_off)
            assert np.array_equal(got_c, ref_c)
            assert np.array_equal(got_o, ref_o)
            # and both against numpy
            for i in range(npairs):
                a = data[int(arrs[0][i]): int(arrs[0][i]) + int(arrs[1][i])]
                b = data[int(arrs[2][i]): int(arrs[2][i]) + int(arrs[3][i])]
                ex = np.intersect1d(a, b)
                assert got_c[i] == len(ex), i
                assert np.array_equal(got_o[int(out_off[i]): int(out_off[i]) + len(ex)], ex), i


def test_noalloc_refuses_more_pairs_than_its_array_holds():
    lib = L.lib()
    with fm.Engine(soln_prealloc=1 << 24, output_prealloc=1 << 24) as e:
        ms = C.c_double()
        one = fm.DeviceBuffer(64)
        rc = lib.fm_intersect_pairs_noalloc(e._h, one.ptr, one.ptr, one.ptr, one.ptr, one.ptr, (1 << 20) + 1, one.ptr,
                                            None, None, C.byref(ms))
        assert rc == 300
        assert b"no-allocation limit" in lib.fm_last_error()
