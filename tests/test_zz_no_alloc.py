"""trax_prefix_sum_noalloc against trax_prefix_sum through the C ABI: the same
exclusive scan and total, bit for bit, on two calls in a row (the second on a
smaller input, so that the shared status array holds the first call's
contents), in both forms of the scan (reduce-then-scan up to 2^25 elements,
look-back above), at the largest size it takes, and a clear error above it."""
import ctypes as C

import numpy as np
import pytest

from trax_amd.api import TraxError, check, lib, ptr

pytestmark = pytest.mark.gpu

TRAX_ERR_ARG = -1
NOALLOC_MAX = 1 << 28


def _scan(dev, values, entry):
    L = lib()
    v = np.ascontiguousarray(values, np.uint32)
    n = v.size
    buf, tot = C.c_void_p(), C.c_void_p()
    check(L.trax_buf_alloc(dev.h, max(n, 1) * 4, C.byref(buf)))
    check(L.trax_buf_alloc(dev.h, 4, C.byref(tot)))
    try:
        check(L.trax_h2d(dev.h, buf, ptr(v), n * 4, None))
        check(getattr(L, entry)(dev.h, buf, n, tot, None))
        out = np.empty(n, np.uint32)
        total = np.zeros(1, np.uint32)
        check(L.trax_d2h(dev.h, ptr(out), buf, n * 4, None))
        check(L.trax_d2h(dev.h, ptr(total), tot, 4, None))
        check(L.trax_stream_sync(dev.h, None))
        return out, int(total[0])
    finally:
        L.trax_buf_free(dev.h, buf)
        L.trax_buf_free(dev.h, tot)


@pytest.mark.parametrize("sizes", [((1 << 25) + 16384 * 3 + 5, 100_003),  # look-back, then reduce-then-scan
                                   (3_000_001, 4097),                      # reduce-then-scan twice
                                   ((1 << 26) + 7, (1 << 25) + 1)])        # look-back twice
def test_noalloc_scan_equals_scan_on_two_calls(dev, sizes):
    rng = np.random.default_rng(sizes[0])
    for n in sizes:
        v = rng.integers(0, 1 << 12, n, dtype=np.uint32)
        want, want_total = _scan(dev, v, "trax_prefix_sum")
        got, got_total = _scan(dev, v, "trax_prefix_sum_noalloc")
        assert np.array_equal(got, want) and got_total == want_total
        run = np.concatenate([np.zeros(1, np.uint64), np.cumsum(v.astype(np.uint64))[:-1]])
        assert np.array_equal(got, (run & np.uint64(0xFFFFFFFF)).astype(np.uint32))


def test_noalloc_scan_at_its_maximum_and_above(dev):
    L = lib()
    n = NOALLOC_MAX
    got, total = _scan(dev, np.ones(n, np.uint32), "trax_prefix_sum_noalloc")
    assert total == n and got[-1] == n - 1
    assert np.array_equal(got, np.arange(n, dtype=np.uint32))
    del got
    # one element more is refused before any launch (the buffer covers it anyway)
    buf = C.c_void_p()
    check(L.trax_buf_alloc(dev.h, (n + 1) * 4, C.byref(buf)))
    try:
        rc = L.trax_prefix_sum_noalloc(dev.h, buf, n + 1, None, None)
        assert rc == TRAX_ERR_ARG
        assert b"2^28" in L.trax_last_error()
        with pytest.raises(TraxError):
            check(L.trax_prefix_sum_noalloc(None, buf, 16, None, None))
    finally:
        L.trax_buf_free(dev.h, buf)
