"""Shared helpers of the tests: event sets, HitSets and array comparison."""
import numpy as np

from trax_amd import make_events, workload
from trax_amd.api import Config, HitSet

PIPELINE_KEYS = ["grid", "hits.x", "hits.y", "hits.z", "hits.layer", "hits.det_id", "hits.hit_id", "hits.event",
                 "pairs", "pair_offsets", "cands", "cand_offsets", "tracklets.h1", "tracklets.h2", "tracklets.h3",
                 "tracklet_offsets", "eta", "conn.basis", "conn.followers", "conn.event", "conn.connectable",
                 "conn.nonconnectable", "pt", "ca.states", "ca.best_basis", "tracks", "track_offsets",
                 "ca.iterations", "diag"]


def small_workload(cfg_no, events=None, tracks=None, noise=None, **extra):
    wl = workload(cfg_no)
    if tracks is not None:
        wl.extra = dict(tracks=tracks, noise=noise or 0, events=events or 1)
    elif events is not None:
        wl.extra = dict(wl.extra, events=events)
    wl.extra.update(extra)
    return wl


def hitset(wl, n_events=None, seed=11, first=0, of=None):
    cfg = Config(wl)
    evs = make_events(wl, n_events, seed=seed, first=first, of=of)
    return cfg, HitSet(cfg, evs), evs


def as_bits(a):
    a = np.asarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def compare(got, want, keys=PIPELINE_KEYS):
    """Bit-exact comparison of named arrays; returns a list of mismatch reports."""
    bad = []
    for k in keys:
        if k not in want and k not in got:
            continue
        if k not in want or k not in got:
            bad.append(f"{k}: present in only one result")
            continue
        g, w = as_bits(got[k]), as_bits(want[k])
        if g.shape != w.shape:
            bad.append(f"{k}: shape {g.shape} != {w.shape}")
            continue
        d = np.nonzero(g != w)[0]
        if d.size:
            i = d[0]
            bad.append(f"{k}: {d.size} of {g.size} differ, first at {i}: {g[i]} != {w[i]}")
    return bad


def extract_tracks(res):
    """extractTracks (TripletBuilder.cpp:955-988) over a result: CA chains of at
    least three tracklets as hit lists, innermost tracklet first."""
    if "track_offsets" not in res:
        return []
    off, trk = res["track_offsets"].astype(np.int64), res["tracks"].astype(np.int64)
    h1, h2, h3 = res["tracklets.h1"], res["tracklets.h2"], res["tracklets.h3"]
    out = []
    for i in range(off.size - 1):
        n = off[i + 1] - off[i]
        if n < 3:
            continue
        chain = trk[off[i]:off[i + 1]][::-1]
        out.append([int(h1[chain[0]]), int(h2[chain[0]])] + [int(h3[t]) for t in chain])
    return out
