"""The reference's per-group call sequence (buildTriplets,
src/triplet_builder/TripletBuilder.cpp:297-510) driven stage by stage through
the stage launchers of the C ABI (include/trax_hip.h, section 2b).

The algorithm classes keep the reference's names and order, and each run()
does what its namesake does in the reference: launch count kernels, scan
their counts (PrefixSum), download the total (downloadScalar), allocate the
outputs to that size and launch the store kernels.  Device memory is the
library's (trax_buf_alloc); nothing here computes results on the host.
"""
import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import api as _api
from .api import CABufs, Config, Device, GridHits, HitCols, HitSet, LinkBufs, P, check, lib, ptr


class DeviceVector:
    """clever::vector<T, 1>: `count` elements of device memory (and `extra`
    elements of room past them)."""

    def __init__(self, dev: Device, count: int, dtype=np.uint32, extra: int = 0):
        self.dev, self.count, self.dtype = dev, int(count), np.dtype(dtype)
        self.h = P()
        check(lib().trax_buf_alloc(dev.h, max(self.count + extra, 1) * self.dtype.itemsize, C.byref(self.h)))

    @property
    def p(self):
        return self.h

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, self.dtype)
        if a.size:
            check(lib().trax_h2d(self.dev.h, self.h, ptr(a), a.nbytes, None))
            check(lib().trax_stream_sync(self.dev.h, None))

    def download(self) -> np.ndarray:
        out = np.empty(self.count, self.dtype)
        if self.count:
            check(lib().trax_d2h(self.dev.h, ptr(out), self.h, out.nbytes, None))
            check(lib().trax_stream_sync(self.dev.h, None))
        return out

    def scalar(self, i: int) -> int:
        """transfer::downloadScalar of element i."""
        v = np.zeros(1, self.dtype)
        check(lib().trax_d2h(self.dev.h, ptr(v), C.c_void_p(self.h.value + i * self.dtype.itemsize),
                             self.dtype.itemsize, None))
        check(lib().trax_stream_sync(self.dev.h, None))
        return int(v[0])

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _api is not None and _api._lib is not None and self.dev.h:
            _api._lib.trax_buf_free(self.dev.h, h)
            self.h = None


def prefix_sum(dev: Device, v: DeviceVector, n: int) -> int:
    """PrefixSum::run over the first n entries (the last one 0), then
    downloadScalar of the total at [n - 1] (PrefixSum.cpp:8-36)."""
    check(lib().trax_prefix_sum(dev.h, v.p, n, None, None))
    return v.scalar(n - 1)


class GridBuilder:  # src/algorithms/GridBuilder.h:25-46
    def __init__(self, dev: Device):
        self.dev = dev

    def run(self, hits: HitSet, cfg: Config, n_layers: int):
        """Grid of the event group and the grid-ordered hit columns with their
        (z, phi) column (GridBuilder.cpp:6-161)."""
        dev, L = self.dev, lib()
        col = hits.columns()
        N, E = hits.n_hits, hits.n_events
        wl = cfg.wl
        src = {k: DeviceVector(dev, N, np.float32 if k in "xyz" else np.uint32) for k in
               ("x", "y", "z", "layer", "det_id", "hit_id", "event")}
        for k, v in src.items():
            v.upload(col[k])
        out = {k: DeviceVector(dev, N, v.dtype) for k, v in src.items()}
        zphi = DeviceVector(dev, 2 * N, np.float32, extra=8)
        sup = {k: DeviceVector(dev, col[k].size) for k in ("event_offset", "layer_nhits", "layer_offset")}
        for k, v in sup.items():
            v.upload(col[k])
        grid = DeviceVector(dev, E * n_layers * (wl.n_z + 1) * (wl.n_phi + 1))
        hin = HitCols(*[src[k].p for k in ("x", "y", "z", "layer", "det_id", "hit_id", "event")])
        hout = HitCols(*[out[k].p for k in ("x", "y", "z", "layer", "det_id", "hit_id", "event")])
        check(L.trax_grid_build_zphi(dev.h, E, n_layers, N, wl.n_z, wl.n_phi, sup["event_offset"].p,
                                     sup["layer_nhits"].p, sup["layer_offset"].p, hin, hout, zphi.p, grid.p, None))
        gh = GridHits(out["x"].p, out["y"].p, out["z"].p, out["layer"].p, out["event"].p, zphi.p, N)
        return grid, out, zphi, gh


class PairGeneratorBeamspot:  # src/algorithms/PairGeneratorBeamspot.h:24-41
    def __init__(self, dev: Device):
        self.dev = dev

    def run(self, dc, E: int, T: int, grid: DeviceVector, gh: GridHits, tuning=None):
        dev, L = self.dev, lib()
        seg_base = DeviceVector(dev, E * T + 1)
        check(L.trax_pair_plan(dev.h, dc, E, grid.p, seg_base.p, None))
        n_slots = prefix_sum(dev, seg_base, E * T + 1)
        slot_counts = DeviceVector(dev, n_slots + 1)
        check(L.trax_pair_count(dev.h, dc, E, grid.p, C.byref(gh), seg_base.p, n_slots, slot_counts.p, tuning, None))
        n_pairs = prefix_sum(dev, slot_counts, n_slots + 1)  # PairGeneratorBeamspot.cpp:124
        pairs, offsets = DeviceVector(dev, 2 * n_pairs), DeviceVector(dev, E * T + 1)
        check(L.trax_pair_store(dev.h, dc, E, grid.p, C.byref(gh), seg_base.p, n_slots, slot_counts.p, n_pairs,
                                pairs.p, offsets.p, tuning, None))
        return pairs, offsets, n_pairs


class TripletThetaPhiPredictor:  # src/algorithms/TripletThetaPhiPredictor.h:24-47
    def __init__(self, dev: Device):
        self.dev = dev

    def run(self, dc, E: int, T: int, grid, gh, pairs, n_pairs, tuning=None):
        dev, L = self.dev, lib()
        counts = DeviceVector(dev, n_pairs + 1)
        check(L.trax_predict_count(dev.h, dc, E, grid.p, C.byref(gh), n_pairs, pairs.p, counts.p, tuning, None))
        n_cands = prefix_sum(dev, counts, n_pairs + 1)  # TripletThetaPhiPredictor.cpp:68
        cands, offsets = DeviceVector(dev, 2 * n_cands), DeviceVector(dev, E * T + 1)
        check(L.trax_predict_store(dev.h, dc, E, grid.p, C.byref(gh), n_pairs, pairs.p, counts.p, n_cands, cands.p,
                                   offsets.p, tuning, None))
        check(L.trax_offsets_fill(dev.h, offsets.p, E, T, None))  # predictOffsetMonotonizeStore
        return cands, offsets, n_cands


class TripletThetaPhiFilter:  # src/algorithms/TripletThetaPhiFilter.h:26-48
    def __init__(self, dev: Device):
        self.dev = dev

    def run(self, dc, E: int, T: int, gh, pairs, n_pairs, cands, n_cands, tuning=None):
        dev, L = self.dev, lib()
        words = (n_cands + 63) // 64
        bitmap, word_counts = DeviceVector(dev, 2 * max(words, 1)), DeviceVector(dev, words + 1)
        check(L.trax_filter_count(dev.h, dc, C.byref(gh), n_pairs, pairs.p, n_cands, cands.p, bitmap.p, word_counts.p,
                                  tuning, None))
        R = prefix_sum(dev, word_counts, words + 1)  # TripletThetaPhiFilter.cpp:85
        h = [DeviceVector(dev, R) for _ in range(3)]
        offsets = DeviceVector(dev, E * T + 1)
        check(L.trax_filter_store(dev.h, dc, E, C.byref(gh), n_pairs, pairs.p, n_cands, cands.p, bitmap.p,
                                  word_counts.p, R, h[0].p, h[1].p, h[2].p, offsets.p, tuning, None))
        check(L.trax_offsets_fill(dev.h, offsets.p, E, T, None))  # filterOffsetMonotonizeStore
        return h, offsets, R


class TripletConnectivityTight:  # src/algorithms/TripletConnectivityTight.h:25-44
    def __init__(self, dev: Device):
        self.dev = dev

    def run(self, dc, E: int, gh, tr, tr_offsets, R: int, tuning=None):
        """(basis, followers, connectable, nonconnectable) and their sizes, and
        the eta and event lists; None when no connection exists (.cpp:93-95)."""
        dev, L = self.dev, lib()
        N = gh.n_hits
        bufs = dict(eta=DeviceVector(dev, R, np.float32), hit_count=DeviceVector(dev, N + 1),
                    hit_cursor=DeviceVector(dev, N), csr=DeviceVector(dev, R), conn=DeviceVector(dev, R + 1),
                    counts=DeviceVector(dev, R + 1), flag_a=DeviceVector(dev, R + 1), flag_b=DeviceVector(dev, R + 1))
        lb = LinkBufs(*[bufs[k].p if k in bufs else None for k, _ in LinkBufs._fields_])
        check(L.trax_link_prepare(dev.h, dc, C.byref(gh), R, tr[0].p, tr[1].p, tr[2].p, C.byref(lb), tuning, None))
        check(L.trax_link_count(dev.h, dc, E, C.byref(gh), R, tr[0].p, tr[1].p, tr[2].p, tr_offsets.p, C.byref(lb),
                                tuning, None))
        K = prefix_sum(dev, bufs["counts"], R + 1)  # TripletConnectivityTight.cpp:89
        if K == 0:
            return None, bufs["eta"]
        for k in ("basis", "followers", "event"):
            bufs[k] = DeviceVector(dev, K)
        lb = LinkBufs(*[bufs[k].p if k in bufs else None for k, _ in LinkBufs._fields_])
        check(L.trax_link_store(dev.h, dc, E, C.byref(gh), R, tr[0].p, tr[1].p, tr[2].p, tr_offsets.p, K, C.byref(lb),
                                tuning, None))
        check(L.trax_link_flags(dev.h, C.byref(gh), R, K, tr[0].p, C.byref(lb), None))
        n_con = prefix_sum(dev, bufs["flag_a"], R + 1)      # .cpp:200
        n_non = prefix_sum(dev, bufs["flag_b"], R + 1)      # .cpp:211
        bufs["connectable"], bufs["nonconnectable"] = DeviceVector(dev, n_con), DeviceVector(dev, n_non)
        lb = LinkBufs(*[bufs[k].p if k in bufs else None for k, _ in LinkBufs._fields_])
        check(L.trax_link_compact(dev.h, R, n_con, n_non, C.byref(lb), None))
        return bufs, bufs["eta"]


class TrackletCircleFitter:  # src/algorithms/TrackletCircleFitter.h:17-29
    def __init__(self, dev: Device):
        self.dev = dev

    def run(self, gh, tr, R: int, connectable: DeviceVector):
        pt = DeviceVector(self.dev, R, np.float32)
        check(lib().trax_circle_fit(self.dev.h, C.byref(gh), R, tr[0].p, tr[1].p, tr[2].p, connectable.p,
                                    connectable.count, pt.p, None))
        return pt


class CellularAutomaton:  # src/algorithms/CellularAutomaton.h:22-49
    def __init__(self, dev: Device):
        self.dev = dev

    def run(self, basis: DeviceVector, followers: DeviceVector, pt: DeviceVector):
        dev, L = self.dev, lib()
        K, R = basis.count, pt.count  # CellularAutomaton.cpp:29-30
        b = dict(states=DeviceVector(dev, R), states_next=DeviceVector(dev, R), flags=DeviceVector(dev, 3),
                 basis_count=DeviceVector(dev, R + 1), cursor=DeviceVector(dev, R), best_basis=DeviceVector(dev, R),
                 is_best=DeviceVector(dev, R), track_offsets=DeviceVector(dev, R + 1))

        def cab():
            return CABufs(*[b[k].p if k in b else None for k, _ in CABufs._fields_])

        check(L.trax_ca_init(dev.h, R, C.byref(cab()), None))
        k = 0
        while True:  # forward phase, living cells read back every iteration (.cpp:40-76)
            check(L.trax_ca_iterate(dev.h, R, K, basis.p, followers.p, C.byref(cab()), k, None))
            alive = b["flags"].scalar(k % 3)
            k += 1
            if not alive:
                break
        check(L.trax_ca_basis_count(dev.h, R, K, basis.p, followers.p, C.byref(cab()), None))
        nb = prefix_sum(dev, b["basis_count"], R + 1)  # .cpp:138
        b["basis_index"], b["basis_diff"] = DeviceVector(dev, nb), DeviceVector(dev, nb, np.float32)
        check(L.trax_ca_basis_store(dev.h, R, K, basis.p, followers.p, pt.p, nb, C.byref(cab()), None))
        n_entries = prefix_sum(dev, b["track_offsets"], R + 1)  # .cpp:284
        b["tracks"] = DeviceVector(dev, n_entries)
        check(L.trax_ca_tracks(dev.h, R, n_entries, C.byref(cab()), None))
        return b, k


def build_triplets(dev: Device, cfg: Config, hits: HitSet, tuning=None) -> Dict[str, np.ndarray]:
    """The event group's whole path through the stage launchers, in the
    reference driver's order; results named as trax_run_plain's."""
    L = lib()
    wl = cfg.wl
    E, T = hits.n_events, len(wl.triplets)
    n_layers = cfg.max_layer
    dc = P()
    check(L.trax_device_config_create(dev.h, cfg.h, C.byref(dc)))
    try:
        check(L.trax_stage_begin(dev.h, E, hits.n_hits, None))
        out: Dict[str, np.ndarray] = {}
        grid, cols, zphi, gh = GridBuilder(dev).run(hits, cfg, n_layers)
        check(L.trax_stage_check(dev.h, None))
        out["grid"] = grid.download()
        for k in ("x", "y", "z", "layer", "det_id", "hit_id", "event"):
            out["hits." + k] = cols[k].download()
        pairs, pair_off, n_pairs = PairGeneratorBeamspot(dev).run(dc, E, T, grid, gh, tuning)
        out["pairs"], out["pair_offsets"] = pairs.download(), pair_off.download()
        cands, cand_off, n_cands = TripletThetaPhiPredictor(dev).run(dc, E, T, grid, gh, pairs, n_pairs, tuning)
        out["cands"], out["cand_offsets"] = cands.download(), cand_off.download()
        tr, tr_off, R = TripletThetaPhiFilter(dev).run(dc, E, T, gh, pairs, n_pairs, cands, n_cands, tuning)
        for i in range(3):
            out[f"tracklets.h{i + 1}"] = tr[i].download()
        out["tracklet_offsets"] = tr_off.download()
        check(L.trax_stage_check(dev.h, None))
        if R == 0:
            return out
        link, eta = TripletConnectivityTight(dev).run(dc, E, gh, tr, tr_off, R, tuning)
        out["eta"] = eta.download()
        if link is None:  # TripletBuilder.cpp:414-422
            for k in ("conn.basis", "conn.followers", "conn.event", "conn.connectable", "conn.nonconnectable"):
                out[k] = np.zeros(0, np.uint32)
            check(L.trax_stage_check(dev.h, None))
            return out
        for k, name in (("basis", "conn.basis"), ("followers", "conn.followers"), ("event", "conn.event"),
                        ("connectable", "conn.connectable"), ("nonconnectable", "conn.nonconnectable")):
            out[name] = link[k].download()
        pt = TrackletCircleFitter(dev).run(gh, tr, R, link["connectable"])
        ca, iterations = CellularAutomaton(dev).run(link["basis"], link["followers"], pt)
        check(L.trax_stage_check(dev.h, None))
        out["pt"] = pt.download()
        out["ca.states"] = ca["states"].download()
        out["ca.best_basis"] = ca["best_basis"].download()
        out["tracks"] = ca["tracks"].download()
        out["track_offsets"] = ca["track_offsets"].download()
        out["ca.iterations"] = np.array([iterations], np.uint32)
        return out
    finally:
        L.trax_device_config_destroy(dc)
