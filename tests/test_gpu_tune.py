"""Autotuner and launch shapes: every supported launch shape gives the same
results as the oracle, and the autotuner's choice is one of them."""
import numpy as np
import pytest

import oracle
from helpers import PIPELINE_KEYS, compare, hitset, small_workload
from trax_amd.api import Batch, Caps, TraxError, autotune, run_batch_fit

pytestmark = pytest.mark.gpu

PREDICT_VARIANTS = [(128, 32, 0), (128, 64, 0), (256, 24, 0), (256, 32, 0), (256, 48, 0), (256, 64, 0),
                    (512, 24, 0), (512, 32, 0), (256, 32, 1)]


def _case():
    wl = small_workload(3, events=4, tracks_min=20, tracks_max=400)
    return (wl, *hitset(wl, seed=11))


@pytest.mark.parametrize("block,lst,srt", PREDICT_VARIANTS)
def test_predictor_variants_match_oracle(dev, block, lst, srt):
    wl, cfg, hs, _ = _case()
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    b, got = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits),
                           tuning=dict(predict_block=block, predict_list=lst, predict_sort=srt))
    b.close()
    bad = compare(got, want)
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("fine", [1, 2, 4, 8, 16])
def test_fine_walk_sub_rows_match_oracle(dev, fine):
    """The fused predictor walks z sub-rows of the grid rows: any sub-row count
    gives the grid walk's candidates in the grid walk's order."""
    wl, cfg, hs, _ = _case()
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    b, got = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits),
                           tuning=dict(predict_block=256, predict_list=24, predict_fine=fine))
    b.close()
    bad = compare(got, want)
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("fine,lst", [(16, 24), (4, 24), (8, 64)])
def test_fine_walk_dense_event_full_lists(dev, fine, lst):
    """A dense event, where many pairs match more hits than their LDS list
    holds: the list keeps the first matches in the grid walk's order and the
    rest come from the re-walk."""
    wl = small_workload(2, tracks=1500, noise=20000, events=1)  # 109 pairs match more than 24 hits
    cfg, hs, _ = hitset(wl, seed=5)
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    b, got = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits),
                           tuning=dict(predict_block=256, predict_list=lst, predict_fine=fine))
    b.close()
    bad = compare(got, want)
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("fine", [3, 32])
def test_fine_sub_rows_must_be_a_supported_power_of_two(dev, fine):
    wl, cfg, hs, _ = _case()
    with pytest.raises(TraxError):
        Batch(dev, cfg, Caps.for_hits(hs.n_events, hs.n_hits), dict(predict_fine=fine))


@pytest.mark.parametrize("shape", [dict(filter_block=64, pair_block=64, link_block=64, grid_store_block=128,
                                        grid_count_block=128, blocks_per_cu=2, pair_xcd=1),
                                   dict(filter_block=1024, pair_block=512, link_block=512, grid_store_block=1024,
                                        grid_count_block=1024, blocks_per_cu=16, pair_xcd=8),
                                   dict(pair_block=192, blocks_per_cu=3, pair_xcd=5)])
def test_block_shapes_match_oracle(dev, shape):
    wl, cfg, hs, _ = _case()
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    b, got = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits), tuning=shape)
    b.close()
    bad = compare(got, want)
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("stage_kib", [64, 80, 0])
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("block", [256, 1024])
def test_grid_build_variants_match_oracle(dev, stage_kib, order, block):
    """The batched grid build's variants: items staged whole in LDS (those
    that fit; larger ones take the global-memory scatter) or the column
    gather, the largest items first or in order, small and large blocks; on
    the uneven events and on a dense one (layers of up to 3,600 hits)."""
    shape = dict(grid_stage_kib=stage_kib, grid_order=order, grid_count_block=block)
    for wl, (cfg, hs, _) in ((w, hitset(w, seed=s)) for w, s in
                             ((small_workload(3, events=4, tracks_min=20, tracks_max=400), 11),
                              (small_workload(2, tracks=1500, noise=20000, events=1), 5))):
        print(f"grid variant {shape}: {wl.name}, {hs.n_hits} hits", flush=True)
        want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature, stop_after=1)
        b, got = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits), tuning=shape)
        b.close()
        keys = ["grid", "hits.x", "hits.y", "hits.z", "hits.layer", "hits.det_id", "hits.hit_id", "hits.event"]
        bad = compare(got, want, keys=keys)
        assert not bad, "\n".join(bad)


@pytest.mark.parametrize("stage_kib,block", [(64, 512), (80, 512), (24, 256)])
def test_grid_records_feed_the_whole_pipeline(dev, stage_kib, block):
    """Items the grid build does not stage in LDS leave 32-byte records whose
    phi becomes the (z, phi) column the pair and predictor walks read: the
    fused pipeline's results are the oracle's on a dense event (layers of up
    to 3,600 hits: staged, recorded with LDS lists, and recorded with the
    global scatter for the layers above the lists)."""
    wl = small_workload(2, tracks=1500, noise=20000, events=1)
    cfg, hs, _ = hitset(wl, seed=5)
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    b, got = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits, fuse_filter=True),
                           tuning=dict(grid_stage_kib=stage_kib, grid_count_block=block))
    b.close()
    bad = compare(got, want, [k for k in PIPELINE_KEYS if k != "cands"])
    assert not bad, "\n".join(bad)


def test_autotune_picks_supported_shapes_with_same_results(dev):
    wl, cfg, hs, _ = _case()
    caps = Caps.for_hits(hs.n_events, hs.n_hits, per_event=True)
    b, _ = run_batch_fit(dev, cfg, hs, caps)
    caps = b.caps
    b.close()
    tuning, trials = autotune(dev, cfg, hs, caps, repeats=1)
    assert len(trials) == 9 + 3 + 4 + 5 + 4 + 2 + 4 + 6 + 4
    assert all(t["same_results"] for t in trials)
    assert (tuning["predict_block"], tuning["predict_list"], tuning["predict_sort"]) in PREDICT_VARIANTS
    assert tuning["blocks_per_cu"] in (2, 4, 8, 16)
    assert tuning["predict_fine"] in (4, 8, 16)
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    tb = Batch(dev, cfg, caps, tuning)
    tb.submit(hs)
    assert tb.wait() == 0
    got = tb.result()
    tb.close()
    per_event = {k: got[k] for k in ("tracklets.h1", "tracklets.h2", "tracklets.h3", "pairs", "cands")}
    bad = compare(per_event, want, keys=list(per_event))
    assert not bad, "\n".join(bad)
