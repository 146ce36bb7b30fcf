"""The stage launchers of the C ABI chained in the reference driver's order
(trax_amd/stages.py, TripletBuilder.cpp:297-510) give what trax_run_plain
gives, bit for bit, on every array: the one-event 100-track case, the dense
2000-track event with 20,000 noise hits and a group of uneven events; and a
stage whose output buffer is too small reports it while the later stages'
kernels do nothing."""
import ctypes as C

import numpy as np
import pytest

from helpers import PIPELINE_KEYS, compare, hitset, small_workload
from trax_amd import workload
from trax_amd.api import TRAX_ERR_CAPACITY, P, check, lib
from trax_amd.stages import (DeviceVector, GridBuilder, PairGeneratorBeamspot, TripletThetaPhiPredictor,
                             build_triplets, prefix_sum)

pytestmark = pytest.mark.gpu

KEYS = [k for k in PIPELINE_KEYS if k != "diag"]


@pytest.mark.parametrize("case", ["one-event-100-tracks", "dense-2000-tracks", "group-6-uneven-events"])
def test_stage_chain_equals_plain_path(dev, case):
    if case == "one-event-100-tracks":
        wl = workload(1)
        cfg, hs, _ = hitset(wl, n_events=1, seed=2024)
    elif case == "dense-2000-tracks":
        wl = small_workload(2)
        cfg, hs, _ = hitset(wl, seed=3)
    else:
        wl = small_workload(2, events=6, tracks=400, noise=3000)
        cfg, hs, _ = hitset(wl, n_events=6, seed=9)
    print(f"case {case}: {hs.n_events} events, {hs.n_hits} hits", flush=True)
    want = dev.run_plain(cfg, hs)
    got = build_triplets(dev, cfg, hs)
    assert want["tracklets.h1"].size > 10
    bad = compare(got, want, KEYS)
    assert not bad, bad
    if case == "dense-2000-tracks":
        assert want["conn.basis"].size > 500 and want["tracks"].size > 500


def test_stage_reports_a_too_small_output_and_stops_the_chain(dev):
    """The pair store is given room for five pairs fewer than the scan
    counted: it writes the entries that fit, trax_stage_check returns
    TRAX_ERR_CAPACITY, and the predictor's kernels launched after it leave
    their output untouched."""
    L = lib()
    wl = small_workload(2, tracks=300, noise=2000)
    cfg, hs, _ = hitset(wl, seed=4)
    E, T = hs.n_events, len(wl.triplets)
    want = dev.run_plain(cfg, hs, stop_after=2)["pairs"]  # before the chain: a run resets the context's record
    dc = P()
    check(L.trax_device_config_create(dev.h, cfg.h, C.byref(dc)))
    try:
        check(L.trax_stage_begin(dev.h, E, hs.n_hits, None))
        grid, cols, zphi, gh = GridBuilder(dev).run(hs, cfg, cfg.max_layer)
        seg_base = DeviceVector(dev, E * T + 1)
        check(L.trax_pair_plan(dev.h, dc, E, grid.p, seg_base.p, None))
        n_slots = prefix_sum(dev, seg_base, E * T + 1)
        slot_counts = DeviceVector(dev, n_slots + 1)
        check(L.trax_pair_count(dev.h, dc, E, grid.p, C.byref(gh), seg_base.p, n_slots, slot_counts.p, None, None))
        n_pairs = prefix_sum(dev, slot_counts, n_slots + 1)
        assert n_pairs > 100
        short = n_pairs - 5
        pairs, offsets = DeviceVector(dev, 2 * short), DeviceVector(dev, E * T + 1)
        check(L.trax_pair_store(dev.h, dc, E, grid.p, C.byref(gh), seg_base.p, n_slots, slot_counts.p, short,
                                pairs.p, offsets.p, None, None))
        assert L.trax_stage_check(dev.h, None) == TRAX_ERR_CAPACITY
        # the entries that fit are the plain path's first ones
        assert np.array_equal(pairs.download(), want[:2 * short])
        # a later stage on the void chain does nothing
        counts = DeviceVector(dev, short + 1)
        check(L.trax_buf_fill_u32(dev.h, counts.p, 0xABCD, short + 1, None))
        check(L.trax_predict_count(dev.h, dc, E, grid.p, C.byref(gh), short, pairs.p, counts.p, None, None))
        assert np.all(counts.download() == 0xABCD)
        assert L.trax_stage_check(dev.h, None) == TRAX_ERR_CAPACITY
        # a new chain clears the error
        check(L.trax_stage_begin(dev.h, E, hs.n_hits, None))
        check(L.trax_stage_check(dev.h, None))
    finally:
        L.trax_device_config_destroy(dc)


def test_ca_best_basis_is_the_ascending_scan_with_nan_and_ties(dev):
    """followerBestBasisStore (CellularAutomaton.h:224-263) scans a follower's
    bases in ascending basis order, keeps the first and takes each later one
    of strictly smaller |pT difference|.  Random followers, each with one to
    eight bases of state 1, pT drawn from four values (equal differences: the
    lowest index stays) and 10 % NaN (a NaN difference never replaces, and
    stays when it comes first): the kernel's best basis is that scan's."""
    from trax_amd.stages import CellularAutomaton, DeviceVector
    rng = np.random.default_rng(7)
    n_b, n_f = 600, 200
    R = n_b + n_f
    conns = []
    for f in range(n_f):
        bases = sorted(int(x) for x in rng.choice(n_b, size=int(rng.integers(1, 9)), replace=False))
        conns += [(b, n_b + f) for b in bases]
    conns.sort()
    pt = rng.choice(np.array([1.0, 2.0, 2.5, 3.0], np.float32), size=R)
    pt[rng.random(R) < 0.1] = np.nan
    basis, foll, ptv = DeviceVector(dev, len(conns)), DeviceVector(dev, len(conns)), DeviceVector(dev, R, np.float32)
    basis.upload(np.array([c[0] for c in conns], np.uint32))
    foll.upload(np.array([c[1] for c in conns], np.uint32))
    ptv.upload(pt)
    b, _ = CellularAutomaton(dev).run(basis, foll, ptv)
    got = b["best_basis"].download()
    n_nan_first = 0
    for f in range(n_b, R):
        bases = sorted(c[0] for c in conns if c[1] == f)
        diffs = [np.abs(pt[x] - pt[f]) for x in bases]  # float32, as the kernel
        best, m = bases[0], diffs[0]
        n_nan_first += bool(np.isnan(m))
        for x, d in zip(bases[1:], diffs[1:]):
            if d < m:
                best, m = x, d
        assert got[f] == best, (f, bases, diffs, got[f])
    assert n_nan_first > 5  # the NaN-first case is exercised
