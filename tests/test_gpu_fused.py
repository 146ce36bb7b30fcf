"""The batched path with the predictor and the filter fused into one pass:
every result except the (unstored) candidates is bit-identical to the
oracle's and to the unfused batched path's, on every graph replay, for the
list sizes and sub-row counts of the launch shapes, and per event."""
import numpy as np
import pytest

import oracle
from helpers import PIPELINE_KEYS, compare, hitset, small_workload
from test_gpu_pipeline import make_case
from trax_amd.api import Batch, Caps, HitSet, autotune, run_batch_fit

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


def _one_big_grid_row(n_noise, seed=3):
    """One event on three layers with a 1 z x 8 phi grid: clean tracks whose
    hits all lie at phi in [-0.2, 2.8], and n_noise hits on the third layer at
    phi in [-3.1, -1.2], so the third layer's single grid row holds about
    n_noise hits, all of them before the tracks' hits in grid order, while
    every window stays clear of the noise."""
    import dataclasses
    from trax_amd.api import Config
    from trax_amd.generator import NOISE_PT, Event, make_event
    wl = dataclasses.replace(small_workload(1), n_z=1, n_phi=8, extra=dict(tracks=0, noise=0, events=1))
    ev = make_event(wl, 300, 0, seed, 0)
    phi = np.arctan2(ev.y, ev.x)
    bad = np.unique(ev.track_id[(phi < -0.2) | (phi > 2.8)])
    keep = ~np.isin(ev.track_id, bad)
    keep &= np.isin(ev.track_id, np.unique(ev.track_id[keep])[:40])
    rng = np.random.default_rng(seed)
    r3, h3 = wl.radii[2], wl.half_lengths[2]
    nphi = rng.uniform(-3.1, -1.2, n_noise)
    cols = dict(x=r3 * np.cos(nphi), y=r3 * np.sin(nphi), z=rng.uniform(-0.99, 0.99, n_noise) * h3)
    big = Event(
        x=np.concatenate([ev.x[keep], cols["x"]]).astype(np.float32),
        y=np.concatenate([ev.y[keep], cols["y"]]).astype(np.float32),
        z=np.concatenate([ev.z[keep], cols["z"]]).astype(np.float32),
        layer=np.concatenate([ev.layer[keep], np.full(n_noise, 3)]).astype(np.uint32),
        det_id=np.concatenate([ev.det_id[keep], np.full(n_noise, 300_300_000)]).astype(np.uint32),
        track_id=np.concatenate([ev.track_id[keep], np.zeros(n_noise)]).astype(np.uint32),
        track_pt=np.concatenate([ev.track_pt[keep], np.full(n_noise, NOISE_PT)]).astype(np.float32),
        n_tracks=40)
    cfg = Config(wl)
    hs = HitSet(cfg, [big])
    return wl, cfg, hs


@pytest.mark.parametrize("n_noise", [65_000, 66_000])
def test_third_layer_at_and_beyond_the_list_offset_limit(dev, n_noise):
    """The fused predictor lists a pair's matches as 16-bit offsets from the
    first hit of the third layer's (event, layer) grid: with about 65,000 hits
    there (one grid row) the listed offsets run up to the limit; with 66,000
    the lists are off and every match goes through the overflow pool.  Both
    give the oracle's results."""
    wl, cfg, hs = _one_big_grid_row(n_noise)
    n3 = int((hs.column("layer") == 3).sum())
    assert (n3 <= 65_535) == (n_noise == 65_000)
    want = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    assert want["tracklets.h1"].size > 0
    b, got = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits, fuse_filter=True))
    b.close()
    bad = compare(got, want, FUSED_KEYS)
    assert not bad, "\n".join(bad)
