"""The recorded ttbar events shipped with the reference
(plotting/data/physics/hits.ttBar.pb, a serialized PEventContainer; kept here
as the data fixture tests/golden/ttbar/hits.ttBar.pb) through the reference's
ttbar configuration (configs/ttBar.reco.conf, configs/layerTriplets.xml):
the wire-format loader against an independent decoder, the physics of the
oracle's tracklets against the reference's published efficiency, and the HIP
path against the oracle on this real data."""
import os
import struct

import numpy as np
import pytest

import oracle
from helpers import compare
from trax_amd.api import Caps, Config, HitSet, physics_eval, run_batch_fit
from trax_amd.config import ttbar_workload
from trax_amd.generator import Event

PB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ttbar", "hits.ttBar.pb")

# plotting/data/physics/physics.ttBar.csv:2-5 (layerTriplet, efficiency, fake rate, clone rate)
PUBLISHED = {0: (0.788567, 0.129784, 0.0632926), 1: (0.811972, 0.536839, 0.0544185),
             2: (0.844733, 0.645111, 0.0568116), 3: (0.739065, 0.526192, 0.0638341)}


def _varint(b, i):
    r = s = 0
    while True:
        c = b[i]
        i += 1
        r |= (c & 0x7F) << s
        s += 7
        if c < 0x80:
            return r, i


def _fields(b):
    i = 0
    while i < len(b):
        key, i = _varint(b, i)
        f, wt = key >> 3, key & 7
        if wt == 0:
            v, i = _varint(b, i)
        elif wt == 2:
            n, i = _varint(b, i)
            v, i = b[i:i + n], i + n
        elif wt == 5:
            v, i = b[i:i + 4], i + 4
        elif wt == 1:
            v, i = b[i:i + 8], i + 8
        else:
            raise ValueError(wt)
        yield f, v


def _decode(data):
    """Event.proto:1-37, decoded field by field in Python."""
    events = []
    for f, ev in _fields(data):
        if f != 1:
            continue
        cols = {k: [] for k in ("x", "y", "z", "layer", "det", "tid", "pt")}
        for g, hit in _fields(ev):
            if g != 4:
                continue
            h = dict(_fields(hit))
            pos = {k: struct.unpack("<f", v)[0] for k, v in _fields(h[1])}
            for k, fk in (("x", 1), ("y", 2), ("z", 3)):
                cols[k].append(pos.get(fk, 0.0))
            cols["layer"].append(h.get(3, 0))
            cols["det"].append(h.get(4, 0) & 0xFFFFFFFF)
            cols["tid"].append(h.get(6, 0) & 0xFFFFFFFF)
            cols["pt"].append(struct.unpack("<f", h[7])[0] if 7 in h else 0.0)
        events.append(Event(np.array(cols["x"], np.float32), np.array(cols["y"], np.float32),
                            np.array(cols["z"], np.float32), np.array(cols["layer"], np.uint32),
                            np.array(cols["det"], np.uint32), np.array(cols["tid"], np.uint32),
                            np.array(cols["pt"], np.float32), 0))
    return events


@pytest.fixture(scope="module")
def ttbar():
    wl = ttbar_workload()
    cfg = Config(wl)
    hs = HitSet(cfg)
    with open(PB, "rb") as f:
        n = hs.add_pb(f.read(), only_tracks=True)
    return wl, cfg, hs, n


def test_pb_loader_matches_independent_decoder(ttbar):
    wl, cfg, hs, n = ttbar
    with open(PB, "rb") as f:
        events = _decode(f.read())
    assert n == len(events) == 10
    ref = HitSet(cfg)
    for ev in events:
        ref.add(ev, only_tracks=True)
    a, b = hs.columns(), ref.columns()
    for k in a:
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert hs.n_hits > 1000


def test_oracle_physics_on_ttbar_matches_published_efficiency(ttbar):
    wl, cfg, hs, _ = ttbar
    res = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    rows = physics_eval(cfg, res, hs.n_events, only_tracks=True)
    for p in range(len(wl.triplets)):
        r = [x for x in rows if x["layer_triplet"] == p]
        eff = np.mean([x["efficiency"] for x in r])
        assert 0.45 < eff <= 1.0, (p, eff)
        assert np.mean([x["fake_rate"] for x in r]) < 0.7
    eff0 = np.mean([x["efficiency"] for x in rows if x["layer_triplet"] == 0])
    assert abs(eff0 - PUBLISHED[0][0]) < 0.06, eff0


@pytest.mark.gpu
def test_hip_matches_oracle_on_ttbar(dev, ttbar):
    wl, cfg, hs, _ = ttbar
    print(f"ttbar: {hs.n_events} events, {hs.n_hits} hits", flush=True)
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    got = dev.run_plain(cfg, hs)
    bad = compare(got, want)
    assert not bad, "\n".join(bad)
    b, batched = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits))
    b.close()
    bad = compare(batched, want)
    assert not bad, "\n".join(bad)
    assert physics_eval(cfg, got, hs.n_events, True) == physics_eval(cfg, want, hs.n_events, True)
