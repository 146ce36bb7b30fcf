"""The batched path with the predictor and the filter fused into one pass:
every result except the (unstored) candidates is bit-identical to the
oracle's and to the unfused batched path's, on every graph replay, for the
list sizes and sub-row counts of the launch shapes, and per event."""
import pytest

import oracle
from helpers import PIPELINE_KEYS, compare, hitset, small_workload
from test_gpu_pipeline import make_case
from trax_amd.api import Batch, Caps, autotune, run_batch_fit

pytestmark = pytest.mark.gpu

FUSED_KEYS = [k for k in PIPELINE_KEYS if k != "cands"]


@pytest.mark.parametrize("wave", [128, 1, 4])
@pytest.mark.parametrize("name", ["one-event-100-tracks", "dense-2000-tracks-20000-noise", "dense-2000-tracks-wg64",
                                  "four-uneven-events", "two-events-ca-density"])
def test_fused_batch_matches_oracle_on_every_replay(dev, name, wave):
    """wave: the block's tiles of pairs (128) or every wave on its own chunks,
    claimed one (1) or four (4) at a time from the regions of the XCDs."""
    wl, cfg, hs, _ = make_case(name)
    print(f"case {name} (wave {wave}): {hs.n_events} events, {hs.n_hits} hits", flush=True)
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    b, got = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits, fuse_filter=True),
                           tuning=dict(predict_wave=wave))
    assert "cands" not in got
    bad = compare(got, want, FUSED_KEYS)
    assert not bad, "\n".join(bad)
    for _ in range(2):
        b.submit(hs)
        assert b.wait() == 0
        bad = compare(b.result(), want, FUSED_KEYS)
        assert not bad, "\n".join(bad)
    b.close()


@pytest.mark.parametrize("wave", [128, 3])
@pytest.mark.parametrize("block,lst,fine", [(256, 24, 4), (512, 32, 8), (128, 64, 16), (256, 48, 1)])
def test_fused_launch_shapes_on_a_dense_event(dev, block, lst, fine, wave):
    """Pairs whose store window matches more hits than their LDS list holds,
    or more than their count-window slots, take the grid walk in the fused
    kernel; the result does not change, with the block's tiles or with every
    wave on its own chunks."""
    wl = small_workload(2, tracks=1500, noise=20000, events=1)  # 109 pairs match more than 24 hits
    cfg, hs, _ = hitset(wl, seed=5)
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    b, got = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits, fuse_filter=True),
                           tuning=dict(predict_block=block, predict_list=lst, predict_fine=fine, predict_wave=wave))
    b.close()
    bad = compare(got, want, FUSED_KEYS)
    assert not bad, "\n".join(bad)


def test_fused_equals_unfused_per_event_groups(dev):
    wl, cfg, hs, _ = make_case("four-uneven-events")
    caps = Caps.for_hits(hs.n_events, hs.n_hits, per_event=True)
    b, plain = run_batch_fit(dev, cfg, hs, caps)
    b.close()
    caps.fuse_filter = True
    f, fused = run_batch_fit(dev, cfg, hs, caps)
    f.close()
    bad = compare(fused, plain, FUSED_KEYS)
    assert not bad, "\n".join(bad)


def test_fused_autotune_keeps_results(dev):
    wl, cfg, hs, _ = make_case("four-uneven-events")
    caps = Caps.for_hits(hs.n_events, hs.n_hits, per_event=True, fuse_filter=True)
    b, want = run_batch_fit(dev, cfg, hs, caps)
    caps = b.caps
    b.close()
    tuning, trials = autotune(dev, cfg, hs, caps, repeats=1)
    # no sorted-walk predictor variant and no filter block size when fused
    assert len(trials) == 8 + 3 + 4 + 4 + 2 + 4 + 6 + 4
    assert all(t["same_results"] for t in trials)
    assert tuning["predict_sort"] == 0
    tb = Batch(dev, cfg, caps, tuning)
    tb.submit(hs)
    assert tb.wait() == 0
    bad = compare(tb.result(), want, FUSED_KEYS)
    tb.close()
    assert not bad, "\n".join(bad)
