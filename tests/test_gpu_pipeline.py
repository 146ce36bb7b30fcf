"""Parity of the HIP path (through the C ABI) with the oracle, bit-exact on
every array of the path: grid, reordered hits, pairs, candidates, tracklets,
offsets, eta, connections, circle radii, CA states, best bases and tracks."""
import numpy as np
import pytest

import oracle
from helpers import PIPELINE_KEYS, compare, hitset, small_workload
from trax_amd.api import Batch, Caps, HitSet, run_batch_fit

pytestmark = pytest.mark.gpu

CASES = {
    # one event, 100 tracks, 3 layers, 8 x 32 grid
    "one-event-100-tracks": dict(cfg=1),
    # 2000 tracks + 20000 noise hits, 10 layers, 8 triplets, 32 x 128 grid
    "dense-2000-tracks-20000-noise": dict(cfg=2),
    # the same with a reference work-group size of 64 (a different pair order)
    "dense-2000-tracks-wg64": dict(cfg=2, n_threads=64),
    # four uneven events in one group (event boundaries inside every stage)
    "four-uneven-events": dict(cfg=3, events=4, tracks_min=20, tracks_max=400),
    # two events at the CA workload's density, reduced in size
    "two-events-ca-density": dict(cfg=5, events=2, tracks=500, noise=5000),
}


def make_case(name, seed=11):
    c = dict(CASES[name])
    cfg_no = c.pop("cfg")
    nt = c.pop("n_threads", None)
    if "tracks" in c:
        wl = small_workload(cfg_no, events=c["events"], tracks=c["tracks"], noise=c["noise"])
    elif "events" in c:
        wl = small_workload(cfg_no, events=c.pop("events"), **c)
    else:
        wl = small_workload(cfg_no)
    if nt:
        wl.n_threads = nt
    return (wl, *hitset(wl, seed=seed))


@pytest.mark.parametrize("name", list(CASES))
def test_plain_path_matches_oracle(dev, name):
    wl, cfg, hs, _ = make_case(name)
    print(f"case {name}: {hs.n_events} events, {hs.n_hits} hits", flush=True)
    got = dev.run_plain(cfg, hs)
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    bad = compare(got, want)
    assert not bad, "\n".join(bad)
    assert got["tracklets.h1"].size > 0


@pytest.mark.parametrize("name", ["one-event-100-tracks", "four-uneven-events", "two-events-ca-density"])
def test_batched_graph_matches_oracle_on_every_replay(dev, name):
    wl, cfg, hs, _ = make_case(name)
    print(f"case {name}: {hs.n_events} events, {hs.n_hits} hits", flush=True)
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    caps = Caps.for_hits(hs.n_events, hs.n_hits)
    b, got = run_batch_fit(dev, cfg, hs, caps)
    bad = compare(got, want)
    assert not bad, "\n".join(bad)
    # replays of the captured graphs, with host work in between
    for _ in range(3):
        b.submit(hs)
        assert b.wait() == 0
        again = b.result()
        bad = compare(again, want)
        assert not bad, "\n".join(bad)
    b.close()


def test_batched_eager_matches_graph(dev):
    wl, cfg, hs, _ = make_case("four-uneven-events")
    caps = Caps.for_hits(hs.n_events, hs.n_hits, use_graph=False)
    b, eager = run_batch_fit(dev, cfg, hs, caps)
    caps.use_graph = True
    g, graph = run_batch_fit(dev, cfg, hs, caps)
    assert not compare(eager, graph)
    b.close()
    g.close()


def _per_event_triplets(res, n_events):
    """Tracklets of each event as sorted (h1, h2, h3) with event-local hit indices."""
    ev = res["hits.event"]
    starts = np.searchsorted(ev, np.arange(n_events))
    out = []
    t = np.stack([res["tracklets.h1"], res["tracklets.h2"], res["tracklets.h3"]], 1).astype(np.int64)
    e_of = ev[t[:, 0]] if t.size else np.zeros(0, np.int64)
    for e in range(n_events):
        s = t[e_of == e] - starts[e]
        out.append(s[np.lexsort(s.T[::-1])] if s.size else s)
    return out


def test_batch_equals_per_event_runs(dev):
    wl = small_workload(3, events=6, tracks_min=20, tracks_max=600)
    from trax_amd import make_events
    from trax_amd.api import Config
    cfg = Config(wl)
    evs = make_events(wl, seed=5)
    hs = HitSet(cfg, evs)
    b, res = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits))
    batch_trip = _per_event_triplets(res, len(evs))
    for e, ev in enumerate(evs):
        single = dev.run_plain(cfg, HitSet(cfg, [ev]), stop_after=4)
        t = _per_event_triplets(single, 1)[0]
        assert np.array_equal(t, batch_trip[e]), f"event {e}"
    b.close()
