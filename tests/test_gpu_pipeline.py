"""Parity of the HIP path (through the C ABI) with the oracle, bit-exact on
every array of the path: grid, reordered hits, pairs, candidates, tracklets,
offsets, eta, connections, circle radii, CA states, best bases and tracks."""
import numpy as np
import pytest

import oracle
from helpers import PIPELINE_KEYS, compare, hitset, small_workload
from trax_amd.api import Batch, Caps, HitSet, TraxError, run_batch_fit

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


@pytest.mark.parametrize("fused", [False, True])
def test_one_graph_batch_matches_oracle_on_every_replay(dev, fused):
    """The whole sequence captured as one graph (use_graph 2): every replay,
    with the inputs uploaded again and eager runs of another batch between
    the replays, equals the oracle; only the total is timed."""
    wl, cfg, hs, _ = make_case("four-uneven-events")
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    keys = [k for k in PIPELINE_KEYS if k != "cands"] if fused else PIPELINE_KEYS
    caps = Caps.for_hits(hs.n_events, hs.n_hits, use_graph=2, fuse_filter=fused)
    b, got = run_batch_fit(dev, cfg, hs, caps)
    e, _ = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits, use_graph=0, fuse_filter=fused))
    bad = compare(got, want, keys)
    assert not bad, "\n".join(bad)
    for _ in range(3):
        e.rerun()
        assert e.wait() == 0
        b.submit(hs)
        assert b.wait() == 0
        bad = compare(b.result(), want, keys)
        assert not bad, "\n".join(bad)
        assert b.stage_ms("total") > 0.0
    with pytest.raises(TraxError):
        b.stage_ms("predict")
    b.close()
    e.close()


def test_batched_eager_matches_graph(dev):
    wl, cfg, hs, _ = make_case("four-uneven-events")
    caps = Caps.for_hits(hs.n_events, hs.n_hits, use_graph=False)
    b, eager = run_batch_fit(dev, cfg, hs, caps)
    caps.use_graph = True
    g, graph = run_batch_fit(dev, cfg, hs, caps)
    assert not compare(eager, graph)
    b.close()
    g.close()


@pytest.mark.parametrize("iterations", [1, 3])
def test_ca_iterations_match_layer_passes(dev, iterations):
    """The CA forward phase as fixed iterations on the device (with the host
    continuing while cells live) gives the states, tracks and iteration count
    of the default one pass per layer triplet, and both the oracle's."""
    wl, cfg, hs, _ = make_case("two-events-ca-density")
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    caps = Caps.for_hits(hs.n_events, hs.n_hits)
    b, layered = run_batch_fit(dev, cfg, hs, caps)
    b.close()
    caps.ca_iterations = iterations
    b, iterated = run_batch_fit(dev, cfg, hs, caps)
    b.close()
    assert int(want["ca.iterations"][0]) > iterations  # the host loop ran
    for got in (layered, iterated):
        bad = compare(got, want)
        assert not bad, "\n".join(bad)


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


def _event_slice(res, e):
    """Event e's arrays of a batch run with each event its own group, rebased to
    the event's first hit and first tracklet (the form a run of the event alone
    has)."""
    hev = res["hits.event"]
    h0 = int(np.searchsorted(hev, e))
    h1a = res["tracklets.h1"].astype(np.int64)
    tev = hev[h1a] if h1a.size else np.zeros(0, np.int64)
    t0, t1 = int(np.searchsorted(tev, e)), int(np.searchsorted(tev, e + 1))
    out = {k: res[k][t0:t1].astype(np.int64) - h0 for k in ("tracklets.h1", "tracklets.h2", "tracklets.h3")}
    out["eta"] = res["eta"][t0:t1]
    basis = res["conn.basis"].astype(np.int64)
    sel = (basis >= t0) & (basis < t1)
    out["conn.basis"] = basis[sel] - t0
    out["conn.followers"] = res["conn.followers"].astype(np.int64)[sel] - t0
    for k in ("conn.connectable", "conn.nonconnectable"):
        v = res[k].astype(np.int64)
        out[k] = v[(v >= t0) & (v < t1)] - t0
    if out["conn.basis"].size:
        out["pt"] = res["pt"][t0:t1]
        out["ca.states"] = res["ca.states"][t0:t1]
        bb = res["ca.best_basis"][t0:t1].astype(np.int64)
        out["ca.best_basis"] = np.where(bb == 0xFFFFFFFF, bb, bb - t0)
        to = res["track_offsets"].astype(np.int64)
        out["track_offsets"] = to[t0:t1 + 1] - to[t0]
        out["tracks"] = res["tracks"][to[t0]:to[t1]].astype(np.int64) - t0
    return out


def test_batch_with_per_event_groups_equals_per_event_runs(dev):
    """Each event its own group: every per-event array of the batch (tracklets,
    eta, connections, predicate lists, circle radii, CA states, best bases,
    tracks) equals the run of that event alone, bit for bit."""
    wl = small_workload(3, events=6, tracks_min=20, tracks_max=600)
    from trax_amd import make_events
    from trax_amd.api import Config
    cfg = Config(wl)
    evs = make_events(wl, seed=5)
    hs = HitSet(cfg, evs)
    print(f"per-event groups: {hs.n_events} events, {hs.n_hits} hits", flush=True)
    b, res = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits, per_event=True))
    b.close()
    n_conn = 0
    for e, ev in enumerate(evs):
        single = dev.run_plain(cfg, HitSet(cfg, [ev]))
        got = _event_slice(res, e)
        for k, v in got.items():
            w = single[k].astype(np.int64) if single[k].dtype == np.uint32 else single[k]
            g = v.view(np.uint32) if v.dtype == np.float32 else v
            w = w.view(np.uint32) if w.dtype == np.float32 else w
            assert np.array_equal(g, w), f"event {e} {k}"
        if single["conn.basis"].size == 0:
            assert "ca.states" not in got
        n_conn += single["conn.basis"].size
    assert n_conn > 0


def test_runtime_records_from_batch_phases(dev):
    """Runtime records filled from a batch's phase events (every phase name
    accepted by trax_batch_stage_ms) read back as the reference's plotting
    scripts read them; the Kernel columns sum to the run's event span."""
    import io

    from trax_amd import runtime_record as rr
    wl, cfg, hs, events = make_case("four-uneven-events")
    b, _ = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits))
    recs, totals = rr.RuntimeRecords(), []
    shape = (hs.n_events, wl.n_layers, len(wl.triplets), wl.n_threads, hs.n_hits, sum(e.n_tracks for e in events))
    for _ in range(3):
        b.rerun()
        assert b.wait() == 0
        recs.add_record(rr.RuntimeRecord(*shape).fill_runtimes(b))
        totals.append(b.stage_ms("total"))
    b.close()
    tab = np.genfromtxt(io.StringIO(recs.csv_dump()), delimiter=" ", names=True)
    assert tab["n"] == 3 and tab["events"] == hs.n_events and tab["hits"] == hs.n_hits
    assert tab["totalKernel"] == pytest.approx(np.mean(totals), rel=1e-3)
    assert tab["totalWalltime"] == pytest.approx(tab["totalKernel"], rel=1e-5)
    for s in ("buildGrid", "pairGen", "tripletPredict", "tripletConnectivityTight", "cellularAutomaton"):
        assert tab[f"{s}Kernel"] > 0.0, s
