"""The 256-event batch of the throughput measurement at full scale (log-uniform
50-5000 tracks per event, Poisson noise of 5 hits per track, seed 2024), as
bench.py runs it: each event its own group, the predictor and the filter
fused, the stage graphs captured and replayed.

- Two replays of the captured graphs give the same arrays.
- The fused batch equals the unfused batch (the candidates stored, no graphs)
  on every array but the candidates, scheduled by every wave on its own
  chunks or by the block's tiles.
- Eight events of it (its three largest, a median one, its smallest and three
  more) equal their runs alone through the plain path, every per-event array
  bit for bit; the largest, the median and the smallest also equal the
  oracle's digests of them (tests/golden, written by make_golden.py).
- The filter's float32 decisions on every candidate of the unfused batch
  differ from a float64 recomputation only where the reference's float32
  expression itself decides so (a bit-exact host replica of it agrees), within
  float32 rounding of the cut's quantity (trax_amd/adjudicate.py)."""
import hashlib
import json
import os

import numpy as np
import pytest

from helpers import PIPELINE_KEYS, compare
from test_gpu_pipeline import _event_slice
from trax_amd import make_events, workload
from trax_amd.adjudicate import adjudicate
from trax_amd.api import Caps, Config, HitSet, run_batch_fit

pytestmark = pytest.mark.gpu

N_EVENTS, SEED = 256, 2024
SAMPLED = [157, 71, 155, 166, 183, 0, 100, 255]
GOLDEN_EVENTS = [157, 166, 183]
FUSED_KEYS = [k for k in PIPELINE_KEYS if k != "cands"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _caps(hs, n_events, **kw):
    return Caps.for_hits(n_events, hs.n_hits, pairs_per_hit=26.0, cands_per_pair=8.0, trk_per_hit=5.0,
                         per_event=True, **kw)


@pytest.fixture(scope="module")
def bench_batch():
    wl = workload(3)
    events = make_events(wl, N_EVENTS, seed=SEED)
    cfg = Config(wl)
    hs = HitSet(cfg, events)
    print(f"batch: {hs.n_events} events, {hs.n_hits} hits", flush=True)
    return wl, cfg, hs, events


@pytest.fixture(scope="module")
def fused(dev, bench_batch):
    wl, cfg, hs, events = bench_batch
    b, res = run_batch_fit(dev, cfg, hs, _caps(hs, N_EVENTS, use_graph=True, fuse_filter=True),
                           tuning=dict(predict_wave=4))
    b.rerun()
    assert b.wait() == 0
    again = b.result(FUSED_KEYS)
    b.close()
    return res, again


@pytest.fixture(scope="module")
def unfused(dev, bench_batch):
    wl, cfg, hs, events = bench_batch
    b, _ = run_batch_fit(dev, cfg, hs, _caps(hs, N_EVENTS, use_graph=False, fuse_filter=False))
    res = b.result(FUSED_KEYS + ["cands", "filter.bits"])
    b.close()
    return res


def test_graph_replays_agree(fused):
    res, again = fused
    assert res["tracklets.h1"].size > 10_000_000
    bad = compare(again, res, FUSED_KEYS)
    assert not bad, "\n".join(bad)


def test_fused_equals_unfused(fused, unfused):
    bad = compare(fused[0], unfused, FUSED_KEYS)
    assert not bad, "\n".join(bad)


def test_fused_tiles_equal_wave_chunks(dev, bench_batch, fused):
    """The fused predictor + filter scheduled by the block's tiles gives the
    arrays it gives with every wave on its own chunks, at full scale."""
    wl, cfg, hs, events = bench_batch
    b, res = run_batch_fit(dev, cfg, hs, _caps(hs, N_EVENTS, use_graph=True, fuse_filter=True),
                           tuning=dict(predict_wave=128))
    b.close()
    bad = compare(res, fused[0], FUSED_KEYS)
    assert not bad, "\n".join(bad)


def _as_run(v, like):
    return v.astype(like.dtype) if like.dtype != np.float32 else v


def test_sampled_events_equal_their_plain_runs(dev, bench_batch, fused):
    wl, cfg, hs, events = bench_batch
    res = fused[0]
    with open(os.path.join(GOLDEN, "pipeline_digests.json")) as f:
        digests = json.load(f)
    for e in SAMPLED:
        single = dev.run_plain(cfg, HitSet(cfg, [events[e]]))
        got = _event_slice(res, e)
        print(f"event {e}: {events[e].n_tracks} tracks, {single['tracklets.h1'].size} tracklets", flush=True)
        for k, v in got.items():
            w = single[k]
            g = v.view(np.uint32) if v.dtype == np.float32 else v.astype(np.int64)
            w = w.view(np.uint32) if w.dtype == np.float32 else w.astype(np.int64)
            assert np.array_equal(g, w), f"event {e} {k}"
        if e in GOLDEN_EVENTS:
            d = digests[f"batch-256-event-{e}"]
            for k, v in got.items():
                a = np.ascontiguousarray(_as_run(v, single[k]))
                assert hashlib.sha256(a.tobytes()).hexdigest() == d[k]["sha256"], f"event {e} {k} vs the oracle"


def test_float64_adjudication_of_the_whole_batch(bench_batch, unfused):
    wl, cfg, hs, events = bench_batch
    print(f"adjudicating {unfused['cands'].size // 2} candidate slots", flush=True)
    a = adjudicate(unfused, cfg, workers=16, progress=lambda m: print(m, flush=True))
    print(json.dumps({k: v for k, v in a.items() if k != "details"}), flush=True)
    assert a["tracklets"] == unfused["tracklets.h1"].size  # the filter's bits are its tracklets
    assert a["unexamined"] == 0
    n = a["float32_float64_disagreements"]
    assert a["candidates"] > 250_000_000
    assert n <= a["candidates"] // 100_000
    # every disagreement is the reference's float32 expression's own decision
    # (a bit-exact host replica of it decides the same) and lies within the
    # float32 rounding of the cut's quantity: the half-spread of 256 Monte
    # Carlo samples underestimates the worst-case rounding range, so the bound
    # is two of its units for every disagreement and one for nearly all
    assert a["float32_replica_agrees"] == n, a["details"][:4]
    assert a["max_rounding_units"] <= 2.0, a["details"][:4]
    assert a["within_float32_rounding"] >= n - max(1, n // 100), a["details"][:4]
