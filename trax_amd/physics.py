"""Physics evaluation of a result against the simulation truth, through the
library's host code (trax_amd/csrc/physics.cpp):

- `PhysicsRecords`: PhysicsRecord / PhysicsRecords
  (src/triplet_builder/PhysicsRecord.{h,cpp}): per layer triplet the
  efficiency, fake and clone rates with their spread over events, and eta /
  pT histograms; `csv()` is PhysicsRecords::csvDump's table, `histogram()`
  its per-record histogram files, `write()` the files the reference's driver
  writes (physics.<name>.csv and <name>/eta_<p>.csv, pt_<p>.csv,
  TripletBuilder.cpp:823-834);
- `track_eval`: the track-level evaluation of the CA's chains, printTracks
  (TripletBuilder.cpp:1008-1066), with `summary_line` its stderr line.
"""
import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

from . import api as _api
from .api import Config, HitSet, TrackSummary, check, lib, ptr

TABLE, ETA, PT = 0, 1, 2
LAYOUT_FULL, LAYOUT_PUBLISHED = 0, 1


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class PhysicsRecords:
    """PhysicsRecords: one merged PhysicsRecord per layer triplet."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().trax_physics_create(C.byref(h)))
        self.h = h

    def add(self, cfg: Config, hits: HitSet, res: Dict[str, np.ndarray]) -> "PhysicsRecords":
        """fillData + addRecord for every event of `hits` and layer triplet of
        `cfg` (TripletBuilder.cpp:515-524), from a result's grid-ordered hits
        and tracklets (HIP or oracle)."""
        a = {k: _f32(res[k]) for k in ("hits.x", "hits.y", "hits.z")}
        a.update({k: _u32(res[k]) for k in ("hits.hit_id", "tracklets.h1", "tracklets.h2", "tracklets.h3",
                                            "tracklet_offsets")})
        need = hits.n_events * len(cfg.wl.triplets) + 1
        if a["tracklet_offsets"].size < need:
            raise ValueError(f"tracklet_offsets holds {a['tracklet_offsets'].size} entries, {need} needed")
        check(lib().trax_physics_add(self.h, cfg.h, hits.h, a["hits.x"].size, ptr(a["hits.x"]), ptr(a["hits.y"]),
                                     ptr(a["hits.z"]), ptr(a["hits.hit_id"]), a["tracklets.h1"].size,
                                     ptr(a["tracklets.h1"]), ptr(a["tracklets.h2"]), ptr(a["tracklets.h3"]),
                                     ptr(a["tracklet_offsets"])))
        return self

    def merge(self, other: "PhysicsRecords") -> "PhysicsRecords":
        check(lib().trax_physics_merge(self.h, other.h))
        return self

    def __len__(self) -> int:
        n = C.c_uint32()
        check(lib().trax_physics_count(self.h, C.byref(n)))
        return int(n.value)

    def _text(self, table: int, layer_triplet: int = 0, layout: int = LAYOUT_FULL) -> str:
        need = C.c_uint64()
        check(lib().trax_physics_csv(self.h, table, layer_triplet, layout, None, 0, C.byref(need)))
        buf = C.create_string_buffer(int(need.value))
        check(lib().trax_physics_csv(self.h, table, layer_triplet, layout, buf, len(buf), C.byref(need)))
        return buf.value.decode()

    def csv(self) -> str:
        """PhysicsRecords::csvDump's table: a row per record, then the total."""
        return self._text(TABLE)

    def histogram(self, which: int, layer_triplet: int, layout: int = LAYOUT_FULL) -> str:
        """The eta (ETA) or pT (PT) histogram file of one record
        (layer_triplet = len(self) + 1: the total's)."""
        return self._text(which, layer_triplet, layout)

    def write(self, directory: str, name: str, layout: int = LAYOUT_FULL) -> str:
        """physics.<name>.csv in `directory` and the histogram files under
        <directory>/<name>/, as the reference's driver lays them out; returns
        the table's path."""
        sub = os.path.join(directory, name)
        os.makedirs(sub, exist_ok=True)
        table = self.csv()
        lts = [int(l.split()[0]) for l in table.splitlines()[1:] if l.strip()]
        for lt in lts:
            for which, stem in ((ETA, "eta"), (PT, "pt")):
                with open(os.path.join(sub, f"{stem}_{lt}.csv"), "w") as f:
                    f.write(self.histogram(which, lt, layout))
        path = os.path.join(directory, f"physics.{name}.csv")
        with open(path, "w") as f:
            f.write(table)
        return path

    def __del__(self):
        if getattr(self, "h", None) and _api._lib is not None:
            _api._lib.trax_physics_destroy(self.h)
            self.h = None


def physics_records(cfg: Config, hits: HitSet, res: Dict[str, np.ndarray]) -> PhysicsRecords:
    return PhysicsRecords().add(cfg, hits, res)


def track_eval(res: Dict[str, np.ndarray], simulated: int, kept: bool = False):
    """printTracks' counts over a result's CA chains: dict with total,
    filtered, valids, valid_unique, fakes, simulated and the three float32
    rates; with kept=True also the per-chain flags (bit 0 kept by clone
    removal, bit 1 real)."""
    hid = _u32(res["hits.hit_id"])
    h = [_u32(res[f"tracklets.h{i}"]) for i in (1, 2, 3)]
    off, trk = _u32(res["track_offsets"]), _u32(res["tracks"])
    n_chains = max(off.size - 1, 0)
    if off.size == 0:
        off = np.zeros(1, np.uint32)
    flags = np.zeros(max(n_chains, 1), np.uint8)
    s = TrackSummary()
    check(lib().trax_track_eval(hid.size, ptr(hid), h[0].size, ptr(h[0]), ptr(h[1]), ptr(h[2]), n_chains, ptr(off),
                                ptr(trk), int(simulated), C.byref(s), ptr(flags)))
    out = {n: (float(getattr(s, n)) if t is C.c_float else int(getattr(s, n))) for n, t in TrackSummary._fields_}
    if kept:
        out["kept"] = flags[:n_chains].copy()
    return out


def summary_line(s: Dict, d_eta_cut: float) -> str:
    """printTracks' stderr line (TripletBuilder.cpp:1058-1066): dEtaCut,
    total, filtered, valids, valid unique, fakes, fake rate, clone rate,
    efficiency, simulated; floats as the C++ stream writes them."""
    f32 = lambda v: format(float(np.float32(v)), "g")
    return " ".join([f32(d_eta_cut), str(s["total"]), str(s["filtered"]), str(s["valids"]), str(s["valid_unique"]),
                     str(s["fakes"]), f32(s["fake_rate"]), f32(s["clone_rate"]), f32(s["efficiency"]),
                     str(s["simulated"])])


def findable_count(hits: HitSet, event: Optional[int] = None) -> int:
    return int(hits.truth(event)["track_id"].size)
