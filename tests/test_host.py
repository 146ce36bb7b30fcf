"""Host-side checks that need no GPU: the float built-ins against float64,
HitCollection::addEvent ordering against a plain restatement, the configuration
constants, and the C ABI exports."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import small_workload
from trax_amd import make_events, workload
from trax_amd.api import Config, HitSet, LIB_PATH, math_host
from trax_amd.config import min_radius_curvature

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ulp_err(got, ref64):
    r32 = ref64.astype(np.float32)
    ulp = np.spacing(np.abs(r32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref64) / ulp


def test_atan2_accuracy():
    rng = np.random.default_rng(0)
    y = rng.uniform(-100, 100, 2_000_000).astype(np.float32)
    x = rng.uniform(-100, 100, 2_000_000).astype(np.float32)
    e = _ulp_err(math_host(0, x, y), np.arctan2(y.astype(np.float64), x.astype(np.float64)))
    assert e.max() <= 2.0


def test_atan2_special_values():
    sy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.inf, 1.0], np.float32)
    sx = np.array([0.0, 0.0, -0.0, -0.0, np.inf, -np.inf, np.inf, -np.inf, 1.0, 0.0], np.float32)
    got = math_host(0, sx, sy)
    ref = np.arctan2(sy, sx)
    assert np.array_equal(got.view(np.uint32), ref.astype(np.float32).view(np.uint32))
    assert np.isnan(math_host(0, np.array([np.nan], np.float32), np.array([1.0], np.float32)))[0]


def test_acos_atan_asinh_accuracy():
    a = np.linspace(-1, 1, 2_000_001).astype(np.float32)
    assert _ulp_err(math_host(1, a), np.arccos(a.astype(np.float64))).max() <= 2.0
    assert np.isnan(math_host(1, np.array([1.0000001, -1.5, np.nan], np.float32))).all()
    w = np.concatenate([np.linspace(-60, 60, 1_000_001), np.geomspace(1e-30, 1e30, 200_000),
                        -np.geomspace(1e-30, 1e30, 200_000)]).astype(np.float32)
    assert _ulp_err(math_host(2, w), np.arctan(w.astype(np.float64))).max() <= 2.0
    assert _ulp_err(math_host(3, w), np.arcsinh(w.astype(np.float64))).max() <= 4.0


def test_min_radius_curvature():
    # TripletConfiguration.cpp:36-43 at minPt = 1 GeV: 87.52 cm
    cfg = Config(workload(2))
    assert cfg.min_radius_curvature == pytest.approx(87.522, abs=1e-3)
    assert cfg.min_radius_curvature == min_radius_curvature(1.0)


def _add_event_restated(ev, triplets, min_pt, evt_in_group):
    """HitCollection::addEvent (src/datastructures/HitCollection.cpp:21-180),
    in plain Python loops."""
    used = set()
    max_layer = 0
    for t in triplets:
        used |= {t.layer1, t.layer2, t.layer3}
        max_layer = max(max_layer, t.layer3)
    tracks = {}
    for i in range(ev.n_hits):
        if ev.track_pt[i] < np.float32(min_pt):
            continue
        tracks.setdefault(int(ev.track_id[i]), []).append(i)
    layers, findable = {}, {}
    for tid in sorted(tracks):
        hits = sorted(tracks[tid], key=lambda i: int(ev.layer[i]))  # stable
        cov = {int(ev.layer[i]) for i in hits}
        missing = any(l not in cov for l in used)
        for i in hits:
            if not missing:
                findable.setdefault(tid, []).append(i)
            else:
                layers.setdefault(int(ev.layer[i]), []).append(i)
    for tid in sorted(findable):
        for i in findable[tid]:
            if ev.layer[i] <= max_layer:
                layers.setdefault(int(ev.layer[i]), []).append(i)
    order = [i for l in range(1, max_layer + 1) for i in layers.get(l, [])]
    counts = [len(layers.get(l, [])) for l in range(1, max_layer + 1)]
    return np.array(order, np.int64), counts


def test_add_event_matches_restatement():
    wl = small_workload(3, events=3, tracks_min=10, tracks_max=60)
    evs = make_events(wl, seed=9)
    # make some hits fall below minPt and drop some layers of a few tracks
    for ev in evs:
        ev.track_pt[::17] = 0.5
        keep = np.ones(ev.n_hits, bool)
        keep[(ev.track_id % 7 == 3) & (ev.layer == 4)] = False
        for k in ("x", "y", "z", "layer", "det_id", "track_id", "track_pt"):
            setattr(ev, k, getattr(ev, k)[keep])
    cfg = Config(wl)
    hs = HitSet(cfg, evs)
    cols = hs.columns()
    start = 0
    for e, ev in enumerate(evs):
        order, counts = _add_event_restated(ev, wl.triplets, wl.min_pt, e)
        n = order.size
        assert cols["event_nhits"][e] == n and cols["event_offset"][e] == start
        assert list(cols["layer_nhits"][e * wl.n_layers:(e + 1) * wl.n_layers]) == counts
        assert np.array_equal(cols["x"][start:start + n], ev.x[order])
        assert np.array_equal(cols["hit_id"][start:start + n], ev.track_id[order])
        assert np.all(cols["event"][start:start + n] == e)
        start += n
    assert start == hs.n_hits


def test_abi_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "trax_hip.h")).read()
    names = set(re.findall(r"\b(trax_[a-z0-9_]+)\s*\(", header))
    assert len(names) > 30
    lib = C.CDLL(LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_filter_phi_prefilter_never_rejects_an_accepted_candidate():
    """The filter's phi-cut prefilter (tx_phi_surely_outside) may only reject
    candidates that the reference's expression rejects: random angles around
    the cut, angles within a few ulp of it, wraps across +-pi, and tiny, huge
    and degenerate vectors."""
    from trax_amd.api import filter_phi_decide_host
    rng = np.random.default_rng(7)
    n = 400_000
    cuts = rng.choice(np.array([0.002, 0.01, 0.03, 0.05, 0.1, 0.3, 1.2], np.float32), n)
    a = rng.uniform(-np.pi, np.pi, n)
    # angles: half spread around the cut, half within 1e-5 rad of +-cut
    off = np.where(rng.random(n) < 0.5, rng.uniform(-0.5, 0.5, n),
                   np.sign(rng.random(n) - 0.5) * (cuts + rng.uniform(-1e-5, 1e-5, n)))
    r1 = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), n))
    r2 = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), n))
    ux, uy = (r1 * np.cos(a)).astype(np.float32), (r1 * np.sin(a)).astype(np.float32)
    vx, vy = (r2 * np.cos(a + off)).astype(np.float32), (r2 * np.sin(a + off)).astype(np.float32)
    so, ps = filter_phi_decide_host(ux, uy, vx, vy, cuts)
    assert not np.any(so & ps)
    # it decides almost everything clearly outside (by more than 1e-4 rad)
    clear = np.abs(off) > cuts + 1e-4
    assert so[clear].mean() > 0.999
    # a decision implies the test on the cross and dot products in float64
    # (which holds the products of floats exactly): d <= 0 or
    # |c| > tan(cut + 1e-5) * d; near-orthogonal pairs probe d around 0
    perp = rng.random(n) < 0.25
    off2 = np.where(perp, np.sign(rng.random(n) - 0.5) * (np.pi / 2 + rng.uniform(-1e-7, 1e-7, n)), off)
    vx2, vy2 = (r2 * np.cos(a + off2)).astype(np.float32), (r2 * np.sin(a + off2)).astype(np.float32)
    so2, ps2 = filter_phi_decide_host(ux, uy, vx2, vy2, cuts)
    assert not np.any(so2 & ps2)
    u_x, u_y, v_x, v_y = (np.float64(t) for t in (ux, uy, vx2, vy2))
    c64 = u_x * v_y - u_y * v_x
    d64 = u_x * v_x + u_y * v_y
    t64 = np.tan(cuts.astype(np.float64) + 1e-5)
    assert np.all(~so2 | (d64 <= 0) | (np.abs(c64) > t64 * d64))
    # degenerate and extreme vectors go to the reference expression
    special = np.array([0.0, -0.0, 1e-38, 1e-30, 1.0, 3e38, np.inf, -np.inf, np.nan], np.float32)
    g = np.array(np.meshgrid(special, special, special, special)).reshape(4, -1)
    c = np.full(g.shape[1], 0.05, np.float32)
    so, ps = filter_phi_decide_host(g[0], g[1], g[2], g[3], c)
    assert not np.any(so & ps)
    with np.errstate(invalid="ignore"):
        undecided = (~np.all(np.isfinite(g), 0) | (np.maximum(np.abs(g[0]), np.abs(g[1])) < 1e-30) |
                     (np.maximum(np.abs(g[2]), np.abs(g[3])) < 1e-30))
    assert not np.any(so & undecided)


def test_config_validate_rejects_what_the_kernels_cannot_take():
    """Layer 0 would index grid rows and radii at layer - 1 = 0xffffffff; more
    than 255 layer triplets do not fit the fused predictor's 8-bit triplet
    field; a missing geometry row leaves a radius unread."""
    import ctypes as C
    from trax_amd.api import P, TraxError, check, lib
    from trax_amd.config import geometry_arrays
    L = lib()
    layer, rmin, rmax = geometry_arrays()

    def make(triplets, geom_rows=None):
        h = P()
        check(L.trax_config_create(32, 128, 1.0, C.byref(h)))
        for t in triplets:
            rc = L.trax_config_add_triplet(h, *t, 0.1, 1.0, 0.1, 0.0, 0.1, 2, 10.0, 2, 100)
            if rc != 0:
                L.trax_config_destroy(h)
                return None
        n = layer.size if geom_rows is None else geom_rows
        l_, a_, b_ = (np.ascontiguousarray(v[:n]) for v in (layer, rmin, rmax))
        check(L.trax_config_set_geometry(h, n, l_.ctypes.data, a_.ctypes.data, b_.ctypes.data))
        return h

    good = make([(1, 2, 3), (2, 3, 4)])
    assert L.trax_config_validate(good) == 0
    L.trax_config_destroy(good)
    assert make([(0, 1, 2)]) is None  # layers are numbered from 1
    many = make([(1, 2, 3)] * 256)
    assert L.trax_config_validate(many) != 0 and b"255" in L.trax_last_error()
    L.trax_config_destroy(many)
    at_limit = make([(1, 2, 3)] * 255)
    assert L.trax_config_validate(at_limit) == 0
    L.trax_config_destroy(at_limit)
    short = make([(1, 2, 3), (2, 3, 4)], geom_rows=3)
    assert L.trax_config_validate(short) != 0
    L.trax_config_destroy(short)
    with pytest.raises(TraxError):
        check(L.trax_config_validate(None))


def _replay_scratch_allocator(chunk_totals_per_wave, chunk_range=256):
    """The wave-held scratch ranges of the fused predictor + filter and of the
    single-pass connectivity kernel: a wave whose chunk needs more than its
    range holds drops the rest and takes max(need, chunk_range) entries with
    one atomic.  Returns (entries handed out, entries used)."""
    handed = used = 0
    for totals in chunk_totals_per_wave:
        left = 0
        for t in totals:
            if t > left:
                size = max(t, chunk_range)
                handed += size
                left = size
            left -= t
            used += t
    return handed, used


def test_scratch_allocator_stays_within_twice_the_entries_plus_one_range_per_wave():
    """The scratch is sized 2 * capacity + one range per resident wave: every
    refill drops less than the refilling chunk's own entries, so the dropped
    parts add up to less than all entries.  Replayed on adversarial sequences
    (each chunk needs one more entry than its wave has left) and random ones."""
    rng = np.random.default_rng(7)
    cases = []
    # adversarial: 1 entry, then a chunk of just over what is left, repeatedly
    adv = []
    for _ in range(500):
        adv += [1, 256]
    cases.append([adv] * 16)
    cases.append([[257, 1] * 300] * 8)
    cases.append([list(rng.integers(0, 300, 400)) for _ in range(64)])
    cases.append([list(rng.integers(0, 40, 2000)) for _ in range(64)])
    cases.append([list(rng.geometric(0.01, 300)) for _ in range(32)])
    for waves in cases:
        handed, used = _replay_scratch_allocator(waves)
        assert handed <= 2 * used + 256 * len(waves), (handed, used, len(waves))
    # the old bound (capacity + one range per wave) is exceeded by the adversarial case
    handed, used = _replay_scratch_allocator(cases[0])
    assert handed > used + 256 * len(cases[0])
