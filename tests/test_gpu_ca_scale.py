"""The cellular-automaton workload at full scale: 64 events of 5000 helix
tracks and 50,000 noise hits, 10 layers, 8 layer triplets, CA iterated to
convergence on the device.  The events run as two batches of 32 (each event
its own group; 32 such events already hold about 2 G candidates).  Checked
through properties that do not depend on the size, plus exact equality of
sampled events with their runs alone (themselves pinned to the oracle at this
density by the golden digest of one such event)."""
import numpy as np
import pytest

from test_gpu_pipeline import _event_slice
from trax_amd import make_events, workload
from trax_amd.api import Batch, Caps, Config, HitSet

pytestmark = pytest.mark.gpu

# bounds per event at this density (one event: 4.34 M pairs, 60 M candidates,
# 1.39 M tracklets, 0.39 M connections), with headroom
PER_EVENT = dict(pairs=4_800_000, cands=66_000_000, tracklets=1_550_000, conn=450_000, track_entries=1_600_000)


def test_ca_scale_64_events(dev):
    wl = workload(5)
    events = make_events(wl, 64, seed=5)
    cfg = Config(wl)
    keys = ["hits.event", "tracklets.h1", "tracklets.h2", "tracklets.h3", "eta", "conn.basis", "conn.followers",
            "conn.connectable", "conn.nonconnectable", "pt", "ca.states", "ca.best_basis", "tracks",
            "track_offsets", "ca.iterations"]
    total_tracks = 0
    for half in range(2):
        evs = events[32 * half:32 * (half + 1)]
        hs = HitSet(cfg, evs)
        print(f"CA scale batch {half}: {hs.n_events} events, {hs.n_hits} hits", flush=True)
        n = hs.n_events
        caps = Caps(n, hs.n_hits, n * PER_EVENT["pairs"], n * PER_EVENT["cands"], n * PER_EVENT["tracklets"],
                    n * PER_EVENT["conn"], n * PER_EVENT["track_entries"], use_graph=True, per_event=True)
        assert caps.max_cands < 2 ** 32
        b = Batch(dev, cfg, caps)
        b.submit(hs)
        assert b.wait() == 0, "capacity overflow"
        res = b.result(keys)
        b.close()
        s, bb, ff = res["ca.states"].astype(np.int64), res["conn.basis"].astype(np.int64), \
            res["conn.followers"].astype(np.int64)
        assert bb.size > 0
        # converged: no connection joins two cells of equal state
        assert np.all(s[bb] != s[ff])
        # every chain steps from a cell to its best basis, one state down, along a connection
        off, trk = res["track_offsets"].astype(np.int64), res["tracks"].astype(np.int64)
        best = res["ca.best_basis"].astype(np.int64)
        lengths = np.diff(off)
        assert lengths.sum() == trk.size
        nxt_ok = np.ones(trk.size, bool)
        last_in_chain = np.zeros(trk.size, bool)
        last_in_chain[off[1:][lengths > 0] - 1] = True
        a, c = trk[:-1], trk[1:]
        inner = ~last_in_chain[:-1]
        nxt_ok[:-1] = ~inner | ((best[a] == c) & (s[c] + 1 == s[a]))
        assert np.all(nxt_ok)
        conn_keys = np.unique(bb * (1 << 32) + ff)
        steps = c[inner] * (1 << 32) + a[inner]
        assert np.all(np.isin(steps, conn_keys))
        total_tracks += int((lengths >= 3).sum())
        # sampled events equal their runs alone
        for e in (0, n - 1):
            single = dev.run_plain(cfg, HitSet(cfg, [evs[e]]))
            got = _event_slice(res, e)
            for k, v in got.items():
                w = single[k]
                g = v.view(np.uint32) if v.dtype == np.float32 else v
                w = w.view(np.uint32) if w.dtype == np.float32 else w.astype(np.int64)
                assert np.array_equal(g, w), f"batch {half} event {e} {k}"
    assert total_tracks > 64 * 500
