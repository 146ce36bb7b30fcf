"""Device primitives through the C ABI: the float built-ins are bit-identical
on host and device, and the exclusive scan matches the closed form and the
reference's known-answer test (src/trax_test/trax_test.cpp:473-479)."""
import ctypes as C

import numpy as np
import pytest

from trax_amd.api import check, lib, math_host, ptr

pytestmark = pytest.mark.gpu


def test_builtins_bit_identical_host_device(dev):
    rng = np.random.default_rng(3)
    n = 1 << 20
    ang = rng.uniform(-np.pi, np.pi, n)
    r = rng.uniform(0.5, 120.0, n)
    x = (r * np.cos(ang)).astype(np.float32)
    y = (r * np.sin(ang)).astype(np.float32)
    special = np.array([0.0, -0.0, 1e-30, -1e-30, 1.0, -1.0, 3.0e38, np.inf, -np.inf], np.float32)
    xs = np.concatenate([x, np.repeat(special, special.size)])
    ys = np.concatenate([y, np.tile(special, special.size)])
    u = np.concatenate([rng.uniform(-1.0, 1.0, n), [0.5, -0.5, 1.0, -1.0, 0.0]]).astype(np.float32)
    w = np.concatenate([rng.uniform(-40.0, 40.0, n), np.geomspace(1e-8, 1e8, 4096)]).astype(np.float32)
    for fn, a, b in ((0, xs, ys), (1, u, None), (2, w, None), (3, w, None)):
        h = math_host(fn, a, b)
        d = dev.math_device(fn, a, b)
        assert np.array_equal(h.view(np.uint32), d.view(np.uint32)), f"function {fn}"


def _scan_device(dev, values, shift=0):
    """The scan of values placed `shift` u32 past a 256-byte aligned
    allocation (shift 1..3: the array is not 16-byte aligned)."""
    L = lib()
    v = np.ascontiguousarray(values, np.uint32)
    n = v.size
    buf, tot = C.c_void_p(), C.c_void_p()
    check(L.trax_buf_alloc(dev.h, (max(n, 1) + shift) * 4, C.byref(buf)))
    check(L.trax_buf_alloc(dev.h, 4, C.byref(tot)))
    data = C.c_void_p(buf.value + 4 * shift)
    try:
        check(L.trax_h2d(dev.h, data, ptr(v), n * 4, None))
        check(L.trax_prefix_sum(dev.h, data, n, tot, None))
        out = np.empty(n, np.uint32)
        total = np.zeros(1, np.uint32)
        check(L.trax_d2h(dev.h, ptr(out), data, n * 4, None))
        check(L.trax_d2h(dev.h, ptr(total), tot, 4, None))
        check(L.trax_stream_sync(dev.h, None))
        return out, int(total[0])
    finally:
        L.trax_buf_free(dev.h, buf)
        L.trax_buf_free(dev.h, tot)


# sizes up to 2^25 take the reduce-then-scan form, larger ones the look-back form
@pytest.mark.parametrize("n", [1, 2, 15, 16, 17, 63, 2047, 2048, 2049, 4095, 4096, 4097, 8191, 8192, 8193, 4096 * 64 + 17,
                               1 << 20, 3_000_001, 1 << 25, (1 << 25) + 16384 * 3 + 5])
@pytest.mark.parametrize("shift", [0, 1])
def test_scan_matches_closed_form(dev, n, shift):
    rng = np.random.default_rng(n)
    v = rng.integers(0, 1 << 12, n, dtype=np.uint32)
    out, total = _scan_device(dev, v, shift)
    run = np.concatenate([np.zeros(1, np.uint64), np.cumsum(v.astype(np.uint64))[:-1]])
    ref = (run & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(out, ref)
    assert total == int(v.astype(np.uint64).sum() & 0xFFFFFFFF)


def test_scan_reference_known_answer(dev):
    n = 1024 * 1024  # trax_test.cpp:447-479: 1 Mi ones, last element N - 1
    out, total = _scan_device(dev, np.ones(n, np.uint32))
    assert out[-1] == n - 1 and total == n
    assert np.array_equal(out, np.arange(n, dtype=np.uint32))


def test_scan_wraps_modulo_2_32(dev):
    v = np.full(5000, 0xF0000000, np.uint32)
    out, total = _scan_device(dev, v)
    ref = (np.arange(5000, dtype=np.uint64) * 0xF0000000) & 0xFFFFFFFF
    assert np.array_equal(out, ref.astype(np.uint32))


def test_timer_measures_stream_work(dev):
    """The device timer reads the time of the launches between start and stop
    on a stream: positive, and longer for eight scans than for one."""
    from trax_amd.api import Stream, Timer
    L = lib()
    n = 1 << 24
    buf = C.c_void_p()
    check(L.trax_buf_alloc(dev.h, 4 * n, C.byref(buf)))
    try:
        s = Stream(dev)
        t = Timer(dev)
        check(L.trax_buf_fill_u32(dev.h, buf, 1, n, s.h))
        check(L.trax_prefix_sum(dev.h, buf, n, None, s.h))
        s.sync()
        times = []
        for reps in (1, 8):
            t.start(s)
            for _ in range(reps):
                check(L.trax_prefix_sum(dev.h, buf, n, None, s.h))
            times.append(t.stop(s))
        assert 0.0 < times[0] < times[1]
    finally:
        L.trax_buf_free(dev.h, buf)
