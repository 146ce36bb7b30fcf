"""Serving loop: every event's results are the same whether it runs alone or in
whatever batch the loop forms for it (arrival pattern, slot count, overflow
re-runs), and the same as the oracle's."""
import numpy as np
import pytest

import oracle
from helpers import extract_tracks, hitset, small_workload
from trax_amd.api import Caps, HitSet, OutputBounds, Server, run_batch_fit

pytestmark = pytest.mark.gpu


def _stream(n, seed, **sizes):
    wl = small_workload(3, events=n, **sizes)
    return (wl, *hitset(wl, seed=seed))


def _alone(dev, cfg, evs):
    return [dev.run_plain(cfg, HitSet(cfg, [ev])) for ev in evs]


def _bounds(dev, cfg, hs, headroom=1.5):
    b, _ = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits))
    c = b.counts()
    b.close()
    return OutputBounds.measure([c], [hs.n_hits], headroom)


def _check_events(sv, alone):
    for e, a in enumerate(alone):
        for k in ("tracklets.h1", "tracklets.h2", "tracklets.h3"):
            got = sv.event_result(e, k)
            assert np.array_equal(got, a[k]), f"event {e} {k}"
        if "track_offsets" in a:  # the event has connections of its own
            assert np.array_equal(sv.event_result(e, "track_offsets"), a["track_offsets"]), f"event {e}"
            assert np.array_equal(sv.event_result(e, "tracks"), a["tracks"]), f"event {e}"
        off = sv.event_result(e, "track_hit_offsets")
        hits = sv.event_result(e, "track_hits")
        got_tracks = [list(map(int, hits[off[i]:off[i + 1]])) for i in range(off.size - 1)]
        assert got_tracks == extract_tracks(a), f"event {e} tracks"


@pytest.mark.parametrize("group", [False, True], ids=["arrival-order", "grouped-by-size"])
@pytest.mark.parametrize("pattern", ["all-at-once", "poisson", "one-slot"])
def test_served_events_match_alone_runs(dev, pattern, group):
    wl, cfg, hs, evs = _stream(16, seed=21, tracks_min=20, tracks_max=600)
    print(f"stream {pattern}: {hs.n_events} events, {hs.n_hits} hits", flush=True)
    alone = _alone(dev, cfg, evs)
    bounds = _bounds(dev, cfg, hs)
    rng = np.random.default_rng(3)
    if pattern == "poisson":
        arrival = np.cumsum(rng.exponential(1.0 / 2000.0, hs.n_events))
    else:
        arrival = np.zeros(hs.n_events)
    sv = Server(dev, cfg, bounds, n_slots=1 if pattern == "one-slot" else 2, max_batch_events=8,
                max_batch_hits=hs.n_hits, keep_results=True, group_by_size=group)
    lat, st = sv.replay(hs, arrival)
    assert st["failed"] == 0 and st["events"] == hs.n_events
    assert (lat > 0).all()
    assert st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["tracklets"] == sum(a["tracklets.h1"].size for a in alone)
    _check_events(sv, alone)
    # the same server replays a second stream with fresh results
    lat2, st2 = sv.replay(hs, arrival)
    assert st2["tracklets"] == st["tracklets"]
    _check_events(sv, alone)
    sv.close()


@pytest.mark.parametrize("group", [False, True], ids=["arrival-order", "grouped-by-size"])
def test_overflowing_batches_are_split(dev, group):
    # tiny events: every output of one event fits the capacities' fixed slack,
    # while bounds far below the real ratios make every multi-event batch overflow
    wl, cfg, hs, evs = _stream(12, seed=4, tracks_min=40, tracks_max=120)
    alone = _alone(dev, cfg, evs)
    pairs = [a["pairs"].size // 2 for a in alone]
    assert max(pairs) < 4096 < sum(pairs)
    tiny = OutputBounds(1e-3, 1e-3, 1e-3, 1e-3, 1e-3)
    sv = Server(dev, cfg, tiny, n_slots=2, max_batch_events=12, max_batch_hits=hs.n_hits, keep_results=True,
                group_by_size=group)
    _, st = sv.replay(hs, np.zeros(hs.n_events))
    assert st["failed"] == 0 and st["resubmits"] > 0
    _check_events(sv, alone)
    sv.close()


def test_served_tracklets_match_oracle(dev):
    wl, cfg, hs, evs = _stream(6, seed=8, tracks_min=50, tracks_max=400)
    bounds = _bounds(dev, cfg, hs)
    sv = Server(dev, cfg, bounds, n_slots=2, max_batch_events=4, max_batch_hits=hs.n_hits, keep_results=True)
    sv.replay(hs, np.zeros(hs.n_events))
    for e, ev in enumerate(evs):
        one = HitSet(cfg, [ev])
        want = oracle.run(one.columns(), wl, cfg.min_radius_curvature)
        for k in ("tracklets.h1", "tracklets.h2", "tracklets.h3"):
            assert np.array_equal(sv.event_result(e, k), want[k]), f"event {e} {k}"
        got_tracks = sv.event_result(e, "track_hits")
        want_tracks = [h for t in extract_tracks(want) for h in t]
        assert list(map(int, got_tracks)) == want_tracks, f"event {e}"
    sv.close()
