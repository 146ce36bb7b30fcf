"""Python host API over the C ABI of libtrax_hip.so (include/trax_hip.h).

The product path is the HIP library: this module only marshals numpy arrays
through ctypes.  It raises if the library has not been built.
"""
import ctypes as C
import os
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

from .config import Workload, geometry_arrays, triplet_arrays
from .generator import Event

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtrax_hip.so")

P = C.c_void_p
U32 = C.c_uint32
_lib = None


class TraxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"trax error {code}: {msg}")
        self.code = code


TRAX_ERR_CAPACITY = -4
TRAX_ERR_INPUT = -3


class HitCols(C.Structure):
    _fields_ = [(n, P) for n in ("x", "y", "z", "layer", "det_id", "hit_id", "event")]


class BatchCaps(C.Structure):
    _fields_ = [(n, U32) for n in ("max_events", "max_hits", "max_pairs", "max_cands", "max_tracklets",
                                   "max_conn", "max_track_entries", "use_graph", "ca_iterations", "per_event",
                                   "fuse_filter")]


class Tuning(C.Structure):
    _fields_ = [(n, U32) for n in ("grid_count_block", "grid_store_block", "pair_block", "predict_block",
                                   "filter_block", "link_block", "blocks_per_cu", "predict_list", "predict_sort",
                                   "predict_fine", "grid_order", "grid_stage_kib", "predict_wave", "pair_xcd")]


class ServeOpts(C.Structure):
    _fields_ = [("n_slots", U32), ("max_batch_events", U32), ("max_batch_hits", U32), ("device_budget", C.c_uint64),
                ("pairs_per_hit", C.c_float), ("cands_per_hit", C.c_float), ("tracklets_per_hit", C.c_float),
                ("conn_per_hit", C.c_float), ("track_entries_per_hit", C.c_float), ("keep_results", U32),
                ("use_graph", U32), ("fuse_filter", U32), ("group_by_size", U32)]


class ServeStats(C.Structure):
    _fields_ = [(n, U32) for n in ("events", "failed", "batches", "resubmits", "max_batch_events", "slot_hits")] + \
               [(n, C.c_uint64) for n in ("tracklets", "track_entries", "device_bytes")] + \
               [(n, C.c_double) for n in ("wall_s", "events_per_s", "tracklets_per_s", "mean_batch_events",
                                          "p50_ms", "p99_ms", "max_ms", "queue_ms", "pipeline_ms", "copy_ms",
                                          "split_ms")]


class PhysicsRow(C.Structure):
    _fields_ = [(n, U32) for n in ("event", "layer_triplet", "n_tracklets", "n_findable", "found", "fakes",
                                   "clones", "misc")] + [(n, C.c_double) for n in ("efficiency", "fake_rate",
                                                                                   "clone_rate")]


class TrackSummary(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("total", "filtered", "valids", "valid_unique", "fakes", "simulated")] + \
               [(n, C.c_float) for n in ("fake_rate", "clone_rate", "efficiency")]


class GridHits(C.Structure):
    _fields_ = [(n, P) for n in ("x", "y", "z", "layer", "event", "zphi")] + [("n_hits", U32)]


class LinkBufs(C.Structure):
    _fields_ = [(n, P) for n in ("eta", "hit_count", "hit_cursor", "csr", "conn", "counts", "basis", "followers",
                                 "event", "flag_a", "flag_b", "connectable", "nonconnectable")]


class CABufs(C.Structure):
    _fields_ = [(n, P) for n in ("states", "states_next", "flags", "basis_count", "cursor", "basis_index",
                                 "basis_diff", "best_basis", "is_best", "track_offsets", "tracks")]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("cu_count", U32), ("lds_bytes_per_block", U32), ("max_block", U32),
                ("wave_size", U32), ("global_mem_bytes", C.c_uint64)]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with trax_amd.build.build() "
                           "(there is no fallback path)")
    L = C.CDLL(LIB_PATH)
    sig = {
        "trax_last_error": (C.c_char_p, []),
        "trax_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
        "trax_ctx_destroy": (None, [P]),
        "trax_device_info_get": (C.c_int, [P, C.POINTER(DeviceInfo)]),
        "trax_stream_create": (C.c_int, [P, C.POINTER(P)]),
        "trax_stream_destroy": (C.c_int, [P]),
        "trax_stream_sync": (C.c_int, [P, P]),
        "trax_timer_create": (C.c_int, [P, C.POINTER(P)]),
        "trax_timer_destroy": (C.c_int, [P]),
        "trax_timer_start": (C.c_int, [P, P, P]),
        "trax_timer_stop": (C.c_int, [P, P, P, C.POINTER(C.c_float)]),
        "trax_buf_alloc": (C.c_int, [P, C.c_uint64, C.POINTER(P)]),
        "trax_buf_free": (C.c_int, [P, P]),
        "trax_buf_fill_u32": (C.c_int, [P, P, U32, C.c_uint64, P]),
        "trax_h2d": (C.c_int, [P, P, P, C.c_uint64, P]),
        "trax_d2h": (C.c_int, [P, P, P, C.c_uint64, P]),
        "trax_prefix_sum": (C.c_int, [P, P, U32, P, P]),
        "trax_prefix_sum_noalloc": (C.c_int, [P, P, U32, P, P]),
        "trax_grid_build": (C.c_int, [P, U32, U32, U32, U32, U32, P, P, P, HitCols, HitCols, P, P]),
        "trax_hits_create": (C.c_int, [U32, C.POINTER(P)]),
        "trax_hits_destroy": (None, [P]),
        "trax_hits_add_event": (C.c_int, [P, P, U32, P, P, P, P, P, P, P, C.c_int, C.c_int]),
        "trax_hits_add_pb": (C.c_int, [P, P, C.c_char_p, C.c_uint64, C.c_int, C.c_int, C.c_int, C.POINTER(U32)]),
        "trax_hits_add_pb_store": (C.c_int, [P, P, C.c_char_p, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.POINTER(U32), C.POINTER(U32)]),
        "trax_hits_count": (U32, [P]),
        "trax_hits_events": (U32, [P]),
        "trax_hits_column": (C.c_int, [P, C.c_char_p, C.POINTER(P), C.POINTER(C.c_uint64)]),
        "trax_config_create": (C.c_int, [U32, U32, C.c_float, C.POINTER(P)]),
        "trax_config_destroy": (None, [P]),
        "trax_config_add_triplet": (C.c_int, [P, U32, U32, U32, C.c_float, C.c_float, C.c_float, C.c_float,
                                              C.c_float, U32, C.c_float, U32, U32]),
        "trax_config_set_geometry": (C.c_int, [P, U32, P, P, P]),
        "trax_config_set_exec": (C.c_int, [P, U32, C.c_float]),
        "trax_config_min_radius_curvature": (C.c_float, [P]),
        "trax_config_validate": (C.c_int, [P]),
        "trax_run_plain": (C.c_int, [P, P, P, C.c_int, C.POINTER(P)]),
        "trax_autotune": (C.c_int, [P, P, P, C.POINTER(BatchCaps), U32, C.POINTER(Tuning), C.c_char_p, C.c_uint64]),
        "trax_batch_plan": (C.c_int, [P, C.POINTER(BatchCaps), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "trax_server_create": (C.c_int, [P, P, C.POINTER(ServeOpts), C.POINTER(P)]),
        "trax_server_destroy": (None, [P]),
        "trax_server_caps": (C.c_int, [P, C.POINTER(BatchCaps)]),
        "trax_server_replay": (C.c_int, [P, P, P, P, C.POINTER(ServeStats)]),
        "trax_server_event_result": (C.c_int, [P, U32, C.c_char_p, C.POINTER(P), C.POINTER(C.c_uint64)]),
        "trax_batch_create": (C.c_int, [P, P, C.POINTER(BatchCaps), C.POINTER(Tuning), P, C.POINTER(P)]),
        "trax_batch_destroy": (None, [P]),
        "trax_batch_submit": (C.c_int, [P, P]),
        "trax_batch_wait": (C.c_int, [P]),
        "trax_batch_rerun": (C.c_int, [P]),
        "trax_run_plain_bench": (C.c_int, [P, P, P, C.POINTER(C.c_double)]),
        "trax_batch_result": (C.c_int, [P, C.POINTER(P)]),
        "trax_batch_result_select": (C.c_int, [P, C.c_char_p, C.POINTER(P)]),
        "trax_batch_counts": (C.c_int, [P, C.POINTER(U32)]),
        "trax_batch_footprint": (C.c_uint64, [P]),
        "trax_batch_stage_ms": (C.c_int, [P, C.c_char_p, C.POINTER(C.c_float)]),
        "trax_result_get": (C.c_int, [P, C.c_char_p, C.POINTER(P), C.POINTER(C.c_uint64), C.POINTER(C.c_int)]),
        "trax_result_timing_ms": (C.c_double, [P, C.c_char_p]),
        "trax_result_destroy": (None, [P]),
        "trax_math_eval_host": (C.c_int, [C.c_int, P, P, P, C.c_uint64]),
        "trax_filter_phi_decide_host": (C.c_int, [C.c_uint64, P, P, P, P, P, P, P]),
        "trax_physics_eval": (C.c_int, [P, U32, U32, P, P, P, U32, P, P, P, P, C.c_int, C.POINTER(PhysicsRow)]),
        "trax_physics_create": (C.c_int, [C.POINTER(P)]),
        "trax_physics_destroy": (None, [P]),
        "trax_physics_add": (C.c_int, [P, P, P, U32, P, P, P, P, U32, P, P, P, P]),
        "trax_physics_merge": (C.c_int, [P, P]),
        "trax_physics_count": (C.c_int, [P, C.POINTER(U32)]),
        "trax_physics_csv": (C.c_int, [P, C.c_int, U32, C.c_int, P, C.c_uint64, C.POINTER(C.c_uint64)]),
        "trax_hits_truth": (C.c_int, [P, U32, C.POINTER(U32), C.POINTER(U32), P, P, P, P, P, P]),
        "trax_filter_audit_f64": (C.c_int, [P, C.c_uint64, P, C.c_uint64, P, U32, P, P, P, P, C.c_float, C.c_int,
                                            P]),
        "trax_track_eval": (C.c_int, [U32, P, U32, P, P, P, U32, P, P, C.c_uint64, C.POINTER(TrackSummary), P]),
        "trax_math_eval_device": (C.c_int, [P, C.c_int, P, P, P, C.c_uint64]),
        "trax_buf_copy_u32": (C.c_int, [P, P, P, C.c_uint64, P]),
        # stage launchers
        "trax_device_config_create": (C.c_int, [P, P, C.POINTER(P)]),
        "trax_device_config_destroy": (None, [P]),
        "trax_stage_begin": (C.c_int, [P, U32, U32, P]),
        "trax_stage_check": (C.c_int, [P, P]),
        "trax_grid_build_zphi": (C.c_int, [P, U32, U32, U32, U32, U32, P, P, P, HitCols, HitCols, P, P, P]),
        "trax_pair_plan": (C.c_int, [P, P, U32, P, P, P]),
        "trax_pair_count": (C.c_int, [P, P, U32, P, C.POINTER(GridHits), P, U32, P, C.POINTER(Tuning), P]),
        "trax_pair_store": (C.c_int, [P, P, U32, P, C.POINTER(GridHits), P, U32, P, U32, P, P, C.POINTER(Tuning), P]),
        "trax_predict_count": (C.c_int, [P, P, U32, P, C.POINTER(GridHits), U32, P, P, C.POINTER(Tuning), P]),
        "trax_predict_store": (C.c_int, [P, P, U32, P, C.POINTER(GridHits), U32, P, P, U32, P, P,
                                         C.POINTER(Tuning), P]),
        "trax_offsets_fill": (C.c_int, [P, P, U32, U32, P]),
        "trax_filter_count": (C.c_int, [P, P, C.POINTER(GridHits), U32, P, U32, P, P, P, C.POINTER(Tuning), P]),
        "trax_filter_store": (C.c_int, [P, P, U32, C.POINTER(GridHits), U32, P, U32, P, P, P, U32, P, P, P, P,
                                        C.POINTER(Tuning), P]),
        "trax_link_prepare": (C.c_int, [P, P, C.POINTER(GridHits), U32, P, P, P, C.POINTER(LinkBufs),
                                        C.POINTER(Tuning), P]),
        "trax_link_count": (C.c_int, [P, P, U32, C.POINTER(GridHits), U32, P, P, P, P, C.POINTER(LinkBufs),
                                      C.POINTER(Tuning), P]),
        "trax_link_store": (C.c_int, [P, P, U32, C.POINTER(GridHits), U32, P, P, P, P, U32, C.POINTER(LinkBufs),
                                      C.POINTER(Tuning), P]),
        "trax_link_flags": (C.c_int, [P, C.POINTER(GridHits), U32, U32, P, C.POINTER(LinkBufs), P]),
        "trax_link_compact": (C.c_int, [P, U32, U32, U32, C.POINTER(LinkBufs), P]),
        "trax_circle_fit": (C.c_int, [P, C.POINTER(GridHits), U32, P, P, P, P, U32, P, P]),
        "trax_ca_init": (C.c_int, [P, U32, C.POINTER(CABufs), P]),
        "trax_ca_iterate": (C.c_int, [P, U32, U32, P, P, C.POINTER(CABufs), U32, P]),
        "trax_ca_basis_count": (C.c_int, [P, U32, U32, P, P, C.POINTER(CABufs), P]),
        "trax_ca_basis_store": (C.c_int, [P, U32, U32, P, P, P, U32, C.POINTER(CABufs), P]),
        "trax_ca_tracks": (C.c_int, [P, U32, U32, C.POINTER(CABufs), P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise TraxError(rc, lib().trax_last_error().decode())


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------------------
class Config:
    """GridConfig + TripletConfigurations + GeometrySupplement of a workload."""

    def __init__(self, wl: Workload):
        L = lib()
        h = P()
        check(L.trax_config_create(wl.n_z, wl.n_phi, wl.min_pt, C.byref(h)))
        self.h = h
        self.wl = wl
        for t in wl.triplets:
            check(L.trax_config_add_triplet(h, t.layer1, t.layer2, t.layer3, t.dThetaCut, t.sigmaZ, t.dPhiCut,
                                            t.sigmaPhi, t.tipCut, t.pairSpreadZ, t.z0, t.pairSpreadPhi,
                                            t.nCandidates))
        layer, rmin, rmax = geometry_arrays()
        self._geom = (layer, rmin, rmax)
        check(L.trax_config_set_geometry(h, layer.size, ptr(layer), ptr(rmin), ptr(rmax)))
        check(L.trax_config_set_exec(h, wl.n_threads, wl.d_eta_cut))
        self.max_layer = max(t.layer3 for t in wl.triplets)

    @property
    def min_radius_curvature(self) -> float:
        return float(lib().trax_config_min_radius_curvature(self.h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.trax_config_destroy(self.h)
            self.h = None


class HitSet:
    """HitCollection + EventSupplement + LayerSupplement built by addEvent."""

    def __init__(self, cfg: Config, events: List[Event] = ()):
        h = P()
        check(lib().trax_hits_create(cfg.max_layer, C.byref(h)))
        self.h = h
        self.cfg = cfg
        for ev in events:
            self.add(ev)

    def add(self, ev: Event, max_tracks: int = -1, only_tracks: bool = False):
        a = [np.ascontiguousarray(v) for v in (ev.x, ev.y, ev.z, ev.layer, ev.det_id, ev.track_id, ev.track_pt)]
        check(lib().trax_hits_add_event(self.h, self.cfg.h, a[0].size, *[ptr(v) for v in a], max_tracks,
                                        int(only_tracks)))

    def add_pb(self, data: bytes, max_events: int = -1, max_tracks: int = -1, only_tracks: bool = False) -> int:
        """Add the events of a serialized PEventContainer; returns how many."""
        n = U32()
        check(lib().trax_hits_add_pb(self.h, self.cfg.h, data, len(data), max_events, max_tracks, int(only_tracks),
                                     C.byref(n)))
        return int(n.value)

    def add_pb_store(self, data: bytes, skip_events: int = 0, max_events: int = -1, max_tracks: int = -1,
                     only_tracks: bool = False):
        """Add the events of an event store (the length-prefixed stream of
        GenericOutputStore); returns (events added, the store's event count)."""
        n, stored = U32(), U32()
        check(lib().trax_hits_add_pb_store(self.h, self.cfg.h, data, len(data), skip_events, max_events, max_tracks,
                                           int(only_tracks), C.byref(n), C.byref(stored)))
        return int(n.value), int(stored.value)

    @property
    def n_hits(self) -> int:
        return int(lib().trax_hits_count(self.h))

    @property
    def n_events(self) -> int:
        return int(lib().trax_hits_events(self.h))

    def column(self, name: str) -> np.ndarray:
        d, n = P(), C.c_uint64()
        check(lib().trax_hits_column(self.h, name.encode(), C.byref(d), C.byref(n)))
        dt = np.float32 if name in ("x", "y", "z") else np.uint32
        if n.value == 0:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(C.cast(d, C.POINTER(C.c_uint32)), shape=(n.value,)).view(dt).copy()

    def columns(self) -> Dict[str, np.ndarray]:
        return {k: self.column(k) for k in ("x", "y", "z", "layer", "det_id", "hit_id", "event", "event_nhits",
                                            "event_offset", "layer_nhits", "layer_offset")}

    def truth(self, event: Optional[int] = None) -> Dict[str, np.ndarray]:
        """The findable tracks addEvent returned for `event` (None: all events
        in order): track ids, hit offsets per track and the hits' x, y, z,
        layer, each track's hits in layer order."""
        ev = 0xFFFFFFFF if event is None else int(event)
        nt, nh = U32(), U32()
        check(lib().trax_hits_truth(self.h, ev, C.byref(nt), C.byref(nh), None, None, None, None, None, None))
        out = {"track_id": np.zeros(nt.value, np.uint32), "hit_offsets": np.zeros(nt.value + 1, np.uint32),
               "x": np.zeros(nh.value, np.float32), "y": np.zeros(nh.value, np.float32),
               "z": np.zeros(nh.value, np.float32), "layer": np.zeros(nh.value, np.uint32)}
        check(lib().trax_hits_truth(self.h, ev, C.byref(nt), C.byref(nh), *[ptr(out[k]) for k in
                                    ("track_id", "hit_offsets", "x", "y", "z", "layer")]))
        return out

    @property
    def n_findable(self) -> int:
        """Findable tracks of all events: the reference's SIMULATED count."""
        nt, nh = U32(), U32()
        check(lib().trax_hits_truth(self.h, 0xFFFFFFFF, C.byref(nt), C.byref(nh), None, None, None, None, None, None))
        return int(nt.value)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.trax_hits_destroy(self.h)
            self.h = None


def _result_dict(rh) -> Dict[str, np.ndarray]:
    L = lib()
    out = {}
    names = ["grid", "hits.x", "hits.y", "hits.z", "hits.layer", "hits.det_id", "hits.hit_id", "hits.event",
             "pairs", "pair_offsets", "cands", "cand_offsets", "tracklets.h1", "tracklets.h2", "tracklets.h3",
             "tracklet_offsets", "eta", "conn.basis", "conn.followers", "conn.event", "conn.connectable",
             "conn.nonconnectable", "pt", "ca.states", "ca.best_basis", "tracks", "track_offsets",
             "ca.iterations", "diag", "filter.bits"]
    for k in names:
        d, n, f = P(), C.c_uint64(), C.c_int()
        if L.trax_result_get(rh, k.encode(), C.byref(d), C.byref(n), C.byref(f)) != 0:
            continue
        dt = np.float32 if f.value else np.uint32
        if n.value == 0:
            out[k] = np.zeros(0, dt)
        else:
            out[k] = np.ctypeslib.as_array(C.cast(d, C.POINTER(C.c_uint32)), shape=(n.value,)).view(dt).copy()
    for st in ("grid", "pairs", "predict", "filter", "connectivity", "circle", "ca"):
        t = L.trax_result_timing_ms(rh, st.encode())
        if t >= 0:
            out["time." + st] = t
    L.trax_result_destroy(rh)
    return out


class Device:
    """A device context (clever::context)."""

    def __init__(self, device: int = 0):
        h = P()
        check(lib().trax_ctx_create(device, C.byref(h)))
        self.h = h

    def info(self) -> DeviceInfo:
        d = DeviceInfo()
        check(lib().trax_device_info_get(self.h, C.byref(d)))
        return d

    def run_plain(self, cfg: Config, hits: HitSet, stop_after: int = 0) -> Dict[str, np.ndarray]:
        rh = P()
        check(lib().trax_run_plain(self.h, cfg.h, hits.h, stop_after, C.byref(rh)))
        return _result_dict(rh)

    def run_plain_bench(self, cfg: Config, hits: HitSet) -> float:
        """Plain path with only the reference driver's own downloads; host ms."""
        ms = C.c_double()
        check(lib().trax_run_plain_bench(self.h, cfg.h, hits.h, C.byref(ms)))
        return float(ms.value)

    def math_device(self, fn: int, x: np.ndarray, y: Optional[np.ndarray] = None) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = None if y is None else np.ascontiguousarray(y, np.float32)
        out = np.empty_like(x)
        check(lib().trax_math_eval_device(self.h, fn, ptr(x), None if y is None else ptr(y), ptr(out), x.size))
        return out

    def stream(self) -> "Stream":
        return Stream(self)

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.trax_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Stream:
    def __init__(self, dev: Device):
        h = P()
        check(lib().trax_stream_create(dev.h, C.byref(h)))
        self.h = h
        self.dev = dev

    def sync(self):
        check(lib().trax_stream_sync(self.dev.h, self.h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.trax_stream_destroy(self.h)
            self.h = None


class Timer:
    """Device time between start() and stop() on a stream (None = the
    context's default stream), from events recorded on that stream."""

    def __init__(self, dev: Device):
        h = P()
        check(lib().trax_timer_create(dev.h, C.byref(h)))
        self.h = h
        self.dev = dev

    def start(self, stream: Optional[Stream] = None):
        check(lib().trax_timer_start(self.dev.h, self.h, stream.h if stream else None))

    def stop(self, stream: Optional[Stream] = None) -> float:
        ms = C.c_float()
        check(lib().trax_timer_stop(self.dev.h, self.h, stream.h if stream else None, C.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.trax_timer_destroy(self.h)
            self.h = None


def math_host(fn: int, x: np.ndarray, y: Optional[np.ndarray] = None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    y = None if y is None else np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    check(lib().trax_math_eval_host(fn, ptr(x), None if y is None else ptr(y), ptr(out), x.size))
    return out


@dataclass
class Caps:
    max_events: int
    max_hits: int
    max_pairs: int
    max_cands: int
    max_tracklets: int
    max_conn: int
    max_track_entries: int
    use_graph: int = 1  # 0 eager, 1 a graph per phase (phases timed), 2 one graph (only the total timed)
    ca_iterations: int = 0
    per_event: bool = False  # each event its own group (results equal per-event runs)
    fuse_filter: bool = False  # predictor and filter in one pass; the candidates are not stored

    @staticmethod
    def for_hits(max_events: int, max_hits: int, pairs_per_hit: float = 8.0, cands_per_pair: float = 4.0,
                 trk_per_hit: float = 1.0, use_graph: bool = True, per_event: bool = False,
                 fuse_filter: bool = False) -> "Caps":
        p = int(max_hits * pairs_per_hit) + 1024
        q = int(p * cands_per_pair) + 1024
        r = int(max_hits * trk_per_hit) + 1024
        return Caps(max_events, max_hits, p, q, r, 4 * r + 1024, 2 * r + 1024, use_graph, 0, per_event, fuse_filter)

    def grown(self, factor: float = 2.0) -> "Caps":
        g = lambda v: int(v * factor) + 1024
        return Caps(self.max_events, self.max_hits, g(self.max_pairs), g(self.max_cands), g(self.max_tracklets),
                    g(self.max_conn), g(self.max_track_entries), self.use_graph, self.ca_iterations,
                    self.per_event, self.fuse_filter)

    def native(self, use_graph: Optional[bool] = None) -> BatchCaps:
        """The C struct of these capacities."""
        return BatchCaps(self.max_events, self.max_hits, self.max_pairs, self.max_cands, self.max_tracklets,
                         self.max_conn, self.max_track_entries,
                         int(self.use_graph if use_graph is None else use_graph), self.ca_iterations,
                         int(self.per_event), int(self.fuse_filter))


class Batch:
    """Batched pipeline: preallocated workspace, stage graphs replayed per batch."""

    def __init__(self, dev: Device, cfg: Config, caps: Caps, tuning: Optional[Dict[str, int]] = None,
                 stream: Optional[Stream] = None):
        c = caps.native()
        t = Tuning(**(tuning or {}))
        h = P()
        check(lib().trax_batch_create(dev.h, cfg.h, C.byref(c), C.byref(t), None if stream is None else stream.h,
                                      C.byref(h)))
        self.h = h
        self.caps = caps
        self.dev = dev
        self.cfg = cfg
        self.stream = stream

    def submit(self, hits: HitSet):
        check(lib().trax_batch_submit(self.h, hits.h))

    def rerun(self):
        """Run the pipeline again on the inputs already resident on the device."""
        check(lib().trax_batch_rerun(self.h))

    def wait(self) -> int:
        rc = lib().trax_batch_wait(self.h)
        if rc not in (0, TRAX_ERR_CAPACITY):
            check(rc)
        return rc

    def counts(self) -> Dict[str, int]:
        a = (U32 * 8)()
        check(lib().trax_batch_counts(self.h, a))
        keys = ("pairs", "cands", "tracklets", "conn", "track_entries", "ca_iterations", "connectable", "err")
        return dict(zip(keys, list(a)))

    def result(self, names: Optional[List[str]] = None) -> Dict[str, np.ndarray]:
        """Result arrays of the last run (all, or only `names`)."""
        rh = P()
        if names is None:
            check(lib().trax_batch_result(self.h, C.byref(rh)))
        else:
            check(lib().trax_batch_result_select(self.h, ",".join(names).encode(), C.byref(rh)))
        return _result_dict(rh)

    def stage_ms(self, stage: str) -> float:
        v = C.c_float()
        check(lib().trax_batch_stage_ms(self.h, stage.encode(), C.byref(v)))
        return float(v.value)

    @property
    def footprint(self) -> int:
        return int(lib().trax_batch_footprint(self.h))

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.trax_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def run_batch_fit(dev: Device, cfg: Config, hits: HitSet, caps: Caps, tuning=None, tries: int = 6):
    """Run one batch, growing the capacities on overflow; returns (batch, result)."""
    for _ in range(tries):
        b = Batch(dev, cfg, caps, tuning)
        b.submit(hits)
        if b.wait() == 0:
            return b, b.result()
        b.close()
        caps = caps.grown()
    raise TraxError(TRAX_ERR_CAPACITY, "capacities kept overflowing")


def plan_bytes(cfg: Config, caps: Caps):
    """(device bytes, pinned host bytes) a batch of these capacities holds."""
    c = caps.native()
    d, h = C.c_uint64(), C.c_uint64()
    check(lib().trax_batch_plan(cfg.h, C.byref(c), C.byref(d), C.byref(h)))
    return int(d.value), int(h.value)


@dataclass
class OutputBounds:
    """Output counts per input hit of a batch: upper bounds for the capacities."""
    pairs: float
    cands: float
    tracklets: float
    conn: float
    track_entries: float

    @staticmethod
    def measure(counts: List[Dict[str, int]], hits: List[int], headroom: float = 1.25) -> "OutputBounds":
        """The largest ratio seen over batches of a workload, with headroom."""
        keys = ("pairs", "cands", "tracklets", "conn", "track_entries")
        r = {k: max(c[k] / max(h, 1) for c, h in zip(counts, hits)) * headroom for k in keys}
        f = {k: max(v, 1e-3) for k, v in r.items()}  # capacities must be positive
        return OutputBounds(f["pairs"], f["cands"], f["tracklets"], f["conn"], f["track_entries"])


class Server:
    """Continuous-batching serving loop over the batched pipeline."""

    def __init__(self, dev: Device, cfg: Config, bounds: OutputBounds, n_slots: int = 2, max_batch_events: int = 256,
                 max_batch_hits: int = 0, device_budget: int = 0, keep_results: bool = False,
                 use_graph: int = 1, fuse_filter: bool = False, group_by_size: bool = False):
        o = ServeOpts(n_slots, max_batch_events, max_batch_hits, device_budget, bounds.pairs, bounds.cands,
                      bounds.tracklets, bounds.conn, bounds.track_entries, int(keep_results), int(use_graph),
                      int(fuse_filter), int(group_by_size))
        h = P()
        check(lib().trax_server_create(dev.h, cfg.h, C.byref(o), C.byref(h)))
        self.h = h
        self.n_events = 0

    def caps(self) -> Dict[str, int]:
        c = BatchCaps()
        check(lib().trax_server_caps(self.h, C.byref(c)))
        return {n: int(getattr(c, n)) for n, _ in BatchCaps._fields_}

    def replay(self, hits: HitSet, arrival_s: np.ndarray):
        """Replay the events of `hits` arriving at arrival_s (seconds from the
        start); returns (per-event latency in ms, stats dict)."""
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        if arr.size != hits.n_events:
            raise ValueError("one arrival time per event")
        lat = np.zeros(arr.size, np.float64)
        st = ServeStats()
        rc = lib().trax_server_replay(self.h, hits.h, ptr(arr), ptr(lat), C.byref(st))
        if rc not in (0, TRAX_ERR_CAPACITY):
            check(rc)
        self.n_events = arr.size
        stats = {n: (float(getattr(st, n)) if t is C.c_double else int(getattr(st, n))) for n, t in st._fields_}
        return lat, stats

    def event_result(self, event: int, name: str) -> np.ndarray:
        d, n = P(), C.c_uint64()
        check(lib().trax_server_event_result(self.h, event, name.encode(), C.byref(d), C.byref(n)))
        if n.value == 0:
            return np.zeros(0, np.uint32)
        return np.ctypeslib.as_array(C.cast(d, C.POINTER(C.c_uint32)), shape=(n.value,)).copy()

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.trax_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def autotune(dev: Device, cfg: Config, hits: HitSet, caps: Caps, repeats: int = 2):
    """Launch shapes for this workload from the library's autotuner: returns
    (tuning dict for Batch(..., tuning=...), list of trial records)."""
    import json
    c = caps.native(use_graph=False)
    t = Tuning()
    buf = C.create_string_buffer(1 << 16)
    check(lib().trax_autotune(dev.h, cfg.h, hits.h, C.byref(c), repeats, C.byref(t), buf, len(buf)))
    trials = [json.loads(l) for l in buf.value.decode().splitlines() if l.strip()]
    return {n: int(getattr(t, n)) for n, _ in Tuning._fields_}, trials


def filter_phi_decide_host(ux, uy, vx, vy, cut):
    """(surely_outside, passes) of the filter's phi cut, evaluated on the host."""
    a = [np.ascontiguousarray(v, dtype=np.float32) for v in (ux, uy, vx, vy, cut)]
    n = a[0].size
    so, ps = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    check(lib().trax_filter_phi_decide_host(n, *[ptr(v) for v in a], ptr(so), ptr(ps)))
    return so.astype(bool), ps.astype(bool)


def physics_eval(cfg: Config, res: Dict[str, np.ndarray], n_events: int, only_tracks: bool = False):
    """PhysicsRecord per (event, layer triplet) of a result (HIP or oracle):
    list of dicts with found/findable/fakes/clones and the three rates."""
    T = len(cfg.wl.triplets)
    a = {k: np.ascontiguousarray(res[k], dtype=np.uint32) for k in
         ("hits.hit_id", "hits.layer", "hits.event", "tracklets.h1", "tracklets.h2", "tracklets.h3",
          "tracklet_offsets")}
    rows = (PhysicsRow * (n_events * T))()
    check(lib().trax_physics_eval(cfg.h, n_events, a["hits.hit_id"].size, ptr(a["hits.hit_id"]), ptr(a["hits.layer"]),
                                  ptr(a["hits.event"]), a["tracklets.h1"].size, ptr(a["tracklets.h1"]),
                                  ptr(a["tracklets.h2"]), ptr(a["tracklets.h3"]), ptr(a["tracklet_offsets"]),
                                  int(only_tracks), rows))
    return [{n: getattr(r, n) for n, _ in PhysicsRow._fields_} for r in rows]
