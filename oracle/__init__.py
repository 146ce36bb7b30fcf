"""ORACLE -- test infrastructure only (CPU restatement of the reference's
kernels, trax_oracle.cpp, and of its physics evaluation, physics.cpp).  Loaded by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the product never imports it."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "trax_oracle.cpp")
SRC_PHYSICS = os.path.join(HERE, "physics.cpp")
LIB = os.path.join(HERE, "_build", "liboracle.so")

P = C.c_void_p
U32 = C.c_uint32


def build(plant: int = 0) -> str:
    """Build the oracle (plant = 0) or a variant with planted mistake `plant`
    (the mutation hooks of trax_oracle.cpp, for showing that the pins catch
    them); returns the library's path."""
    out = LIB if plant == 0 else os.path.join(HERE, "_build", f"liboracle_plant{plant}.so")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-Wno-unused-function",
           f"-DTRAX_ORACLE_PLANT={int(plant)}", SRC, SRC_PHYSICS, "-o", tmp, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    return out


class Input(C.Structure):
    _fields_ = [
        ("x", P), ("y", P), ("z", P), ("layer", P), ("det_id", P), ("hit_id", P), ("event", P), ("n_hits", U32),
        ("n_events", U32), ("n_layers", U32), ("evt_nhits", P), ("evt_offset", P), ("lay_nhits", P),
        ("lay_offset", P), ("n_z", U32), ("n_phi", U32), ("min_r", P), ("max_r", P), ("n_geom", U32),
        ("n_trip", U32), ("l1", P), ("l2", P), ("l3", P), ("dtheta", P), ("sigma_z", P), ("dphi", P),
        ("sigma_phi", P), ("tip", P), ("z0", P), ("pair_spread_phi", P), ("min_radius_curvature", C.c_float),
        ("n_threads", U32), ("d_eta_cut", C.c_float), ("sparse_pairs", C.c_int), ("stop_after", C.c_int),
        ("n_workers", C.c_int),
    ]


_libs = {}


def lib(path=None):
    """The oracle library (or the variant at `path`)."""
    path = path or LIB
    if path not in _libs:
        if path == LIB and not os.path.exists(LIB):
            build()
        L = C.CDLL(path)
        L.oracle_run.restype = P
        L.oracle_run.argtypes = [C.POINTER(Input)]
        L.oracle_error.restype = C.c_char_p
        L.oracle_error.argtypes = [P]
        L.oracle_get.restype = C.c_longlong
        L.oracle_get.argtypes = [P, C.c_char_p, C.POINTER(P), C.POINTER(C.c_int)]
        L.oracle_timing.restype = C.c_double
        L.oracle_timing.argtypes = [P, C.c_char_p]
        L.oracle_free.argtypes = [P]
        L.oracle_prefix_sum.argtypes = [P, U32, U32]
        L.oracle_run_link.restype = P
        L.oracle_run_link.argtypes = [C.POINTER(Input), P, P, P, U32]
        L.oracle_physics.restype = P
        L.oracle_physics.argtypes = [U32, U32] + [P] * 6 + [U32] + [P] * 4 + [U32] + [P] * 4
        L.oracle_physics_text.restype = C.c_char_p
        L.oracle_physics_text.argtypes = [P, C.c_int, U32, C.c_int]
        L.oracle_physics_fill.argtypes = [P, U32, P, P]
        L.oracle_physics_fill_text.restype = C.c_char_p
        L.oracle_physics_fill_text.argtypes = [P, U32, C.c_int, C.c_int]
        L.oracle_physics_free.argtypes = [P]
        L.oracle_tracks.argtypes = [U32, P, P, P, P, U32, P, P, C.c_ulonglong, P, P, P]
        _libs[path] = L
    return _libs[path]


NAMES = ["grid", "hits.x", "hits.y", "hits.z", "hits.layer", "hits.det_id", "hits.hit_id", "hits.event", "pairs",
         "pair_offsets", "cands", "cand_offsets", "tracklets.h1", "tracklets.h2", "tracklets.h3",
         "tracklet_offsets", "eta", "conn.count_prefix", "conn.basis", "conn.followers", "conn.event",
         "conn.connectable", "conn.nonconnectable", "pt", "ca.states", "ca.iterations", "ca.best_basis", "tracks",
         "track_offsets", "diag"]


def run(cols, wl, min_radius_curvature, stop_after=0, sparse_pairs=True, workers=None, lib_path=None):
    """Run the oracle on HitCollection columns (a dict as produced by
    trax_amd.api.HitSet.columns()) for a Workload.  Returns named arrays.
    lib_path selects a variant built by build(plant)."""
    from trax_amd.config import geometry_arrays, triplet_arrays
    keep = []

    def a(v, dt):
        v = np.ascontiguousarray(v, dtype=dt)
        keep.append(v)
        return C.c_void_p(v.ctypes.data)

    layer, rmin, rmax = geometry_arrays()
    t = triplet_arrays(wl.triplets)
    inp = Input()
    inp.x, inp.y, inp.z = a(cols["x"], np.float32), a(cols["y"], np.float32), a(cols["z"], np.float32)
    inp.layer, inp.det_id = a(cols["layer"], np.uint32), a(cols["det_id"], np.uint32)
    inp.hit_id, inp.event = a(cols["hit_id"], np.uint32), a(cols["event"], np.uint32)
    inp.n_hits = cols["x"].size
    inp.n_events = cols["event_nhits"].size
    inp.n_layers = max(tt.layer3 for tt in wl.triplets)
    inp.evt_nhits, inp.evt_offset = a(cols["event_nhits"], np.uint32), a(cols["event_offset"], np.uint32)
    inp.lay_nhits, inp.lay_offset = a(cols["layer_nhits"], np.uint32), a(cols["layer_offset"], np.uint32)
    inp.n_z, inp.n_phi = wl.n_z, wl.n_phi
    inp.min_r, inp.max_r, inp.n_geom = a(rmin, np.float32), a(rmax, np.float32), rmin.size
    inp.n_trip = len(wl.triplets)
    inp.l1, inp.l2, inp.l3 = a(t["layer1"], np.uint32), a(t["layer2"], np.uint32), a(t["layer3"], np.uint32)
    inp.dtheta, inp.sigma_z = a(t["dThetaCut"], np.float32), a(t["sigmaZ"], np.float32)
    inp.dphi, inp.sigma_phi = a(t["dPhiCut"], np.float32), a(t["sigmaPhi"], np.float32)
    inp.tip, inp.z0 = a(t["tipCut"], np.float32), a(t["z0"], np.float32)
    inp.pair_spread_phi = a(t["pairSpreadPhi"], np.uint32)
    inp.min_radius_curvature = min_radius_curvature
    inp.n_threads = wl.n_threads
    inp.d_eta_cut = wl.d_eta_cut
    inp.sparse_pairs = int(sparse_pairs)
    inp.stop_after = stop_after
    inp.n_workers = workers if workers is not None else min(8, os.cpu_count() or 1)
    L = lib(lib_path)
    h = L.oracle_run(C.byref(inp))
    try:
        err = L.oracle_error(h)
        if err:
            raise RuntimeError("oracle: " + err.decode())
        out = {}
        for k in NAMES:
            d, f = P(), C.c_int()
            n = L.oracle_get(h, k.encode(), C.byref(d), C.byref(f))
            if n < 0:
                continue
            dt = np.float32 if f.value else np.uint32
            out[k] = (np.zeros(0, dt) if n == 0 else
                      np.ctypeslib.as_array(C.cast(d, C.POINTER(C.c_uint32)), shape=(n,)).view(dt).copy())
        for st in ("grid", "pairs", "predict", "filter", "connectivity_ca"):
            v = L.oracle_timing(h, st.encode())
            if v >= 0:
                out["time." + st] = v
        return out
    finally:
        L.oracle_free(h)


def run_link(hits, wl, min_radius_curvature, h1, h2, h3, n_events=1, lib_path=None):
    """Connectivity, circle fit and CA of the oracle on given tracklets.  `hits`
    holds grid-ordered columns x, y, z, layer, event (a hand-built fixture)."""
    keep = []

    def a(v, dt):
        v = np.ascontiguousarray(v, dtype=dt)
        keep.append(v)
        return C.c_void_p(v.ctypes.data)

    inp = Input()
    inp.x, inp.y, inp.z = a(hits["x"], np.float32), a(hits["y"], np.float32), a(hits["z"], np.float32)
    inp.layer, inp.event = a(hits["layer"], np.uint32), a(hits["event"], np.uint32)
    inp.n_hits = len(hits["x"])
    inp.n_events = n_events
    inp.n_layers = max(t.layer3 for t in wl.triplets)
    inp.n_trip = len(wl.triplets)
    inp.n_threads = wl.n_threads
    inp.d_eta_cut = wl.d_eta_cut
    inp.min_radius_curvature = min_radius_curvature
    inp.n_workers = 1
    L = lib(lib_path)
    h = L.oracle_run_link(C.byref(inp), a(h1, np.uint32), a(h2, np.uint32), a(h3, np.uint32), len(h1))
    try:
        out = {}
        for k in NAMES + ["tracklet_offsets"]:
            d, f = P(), C.c_int()
            n = L.oracle_get(h, k.encode(), C.byref(d), C.byref(f))
            if n < 0:
                continue
            dt = np.float32 if f.value else np.uint32
            out[k] = (np.zeros(0, dt) if n == 0 else
                      np.ctypeslib.as_array(C.cast(d, C.POINTER(C.c_uint32)), shape=(n,)).view(dt).copy())
        return out
    finally:
        L.oracle_free(h)


def prefix_sum(values, wg):
    v = np.ascontiguousarray(values, dtype=np.uint32).copy()
    lib().oracle_prefix_sum(C.c_void_p(v.ctypes.data), v.size, wg)
    return v


def findable_tracks(event, wl, max_tracks=-1, only_tracks=False):
    """HitCollection::addEvent's findable tracks (HitCollection.cpp:41-96)
    of one event (an Event of trax_amd.generator): hits of at least minPt
    grouped by simulated track id (id 0 skipped with only_tracks), each
    track's hits ordered by layer (stably), kept when they cover every layer
    the triplets use; at most max_tracks findable tracks (-1: all).  Returns
    {track id: indices of its hits in the event}, ids ascending."""
    used = sorted({l for t in wl.triplets for l in (t.layer1, t.layer2, t.layer3)})
    min_pt = float(np.float32(wl.min_pt))
    tid, layer, pt = (np.asarray(event.track_id), np.asarray(event.layer), np.asarray(event.track_pt))
    groups = {}
    for i in range(tid.size):
        if only_tracks and tid[i] == 0:
            continue
        if float(pt[i]) < min_pt:
            continue
        groups.setdefault(int(tid[i]), []).append(i)
    out, found = {}, 0
    for t in sorted(groups):
        if found == max_tracks:
            break
        if only_tracks and t == 0:
            continue
        hits = sorted(groups[t], key=lambda i: int(layer[i]))
        cov = {int(layer[i]) for i in hits}
        if all(l in cov for l in used):
            out[t] = hits
            found += 1
    return out


def _flat_truth(events, truths):
    ev_off, tids, hoff, xs, ys, zs = [0], [], [0], [], [], []
    for ev, tr in zip(events, truths):
        for t, idx in tr.items():
            tids.append(t)
            xs.extend(np.asarray(ev.x)[idx]), ys.extend(np.asarray(ev.y)[idx]), zs.extend(np.asarray(ev.z)[idx])
            hoff.append(hoff[-1] + len(idx))
        ev_off.append(len(tids))
    return ([np.asarray(v, np.uint32) for v in (ev_off, tids, hoff)] +
            [np.asarray(v, np.float32) for v in (xs, ys, zs)])


def physics(events, truths, res, n_lt, lib_path=None):
    """PhysicsRecords of a result against the findable tracks `truths` (one
    findable_tracks dict per event of `events`).  Returns {"table": csvDump's
    table, ("eta" | "pt", layer triplet, layout): histogram text, "fills":
    [(event, layer triplet, efficiency, fake rate, clone rate)] as filled,
    "fill_hist": [{"eta": text, "pt": text}] of each record as filled}."""
    keep = []

    def a(v, dt):
        v = np.ascontiguousarray(v, dtype=dt)
        keep.append(v)
        return C.c_void_p(v.ctypes.data)

    tr = _flat_truth(events, truths)
    L = lib(lib_path)
    x, y, z = (a(res[k], np.float32) for k in ("hits.x", "hits.y", "hits.z"))
    h = L.oracle_physics(len(events), n_lt, *[a(v, v.dtype) for v in tr], res["hits.x"].size, x, y, z,
                         a(res["hits.hit_id"], np.uint32), res["tracklets.h1"].size,
                         a(res["tracklets.h1"], np.uint32), a(res["tracklets.h2"], np.uint32),
                         a(res["tracklets.h3"], np.uint32), a(res["tracklet_offsets"], np.uint32))
    try:
        out = {"table": L.oracle_physics_text(h, 0, 0, 0).decode()}
        lts = [int(l.split()[0]) for l in out["table"].splitlines()[1:] if l.strip()]
        for lt in lts:
            for which, name in ((1, "eta"), (2, "pt")):
                for layout in (0, 1):
                    out[(name, lt, layout)] = L.oracle_physics_text(h, which, lt, layout).decode()
        fills, ids, rates, i = [], (U32 * 2)(), (C.c_double * 3)(), 0
        hists = []
        while L.oracle_physics_fill(h, i, ids, rates) == 0:
            fills.append((ids[0], ids[1], rates[0], rates[1], rates[2]))
            hists.append({"eta": L.oracle_physics_fill_text(h, i, 1, 0).decode(),
                          "pt": L.oracle_physics_fill_text(h, i, 2, 0).decode()})
            i += 1
        out["fills"] = fills
        out["fill_hist"] = hists
        return out
    finally:
        L.oracle_physics_free(h)


def tracks(res, simulated, lib_path=None):
    """printTracks' counts (TripletBuilder.cpp:1008-1066) over a result's CA
    chains: dict as trax_amd.physics.track_eval returns it, with "kept"."""
    keep = []

    def a(v, dt):
        v = np.ascontiguousarray(v, dtype=dt)
        keep.append(v)
        return C.c_void_p(v.ctypes.data)

    off = np.asarray(res["track_offsets"], np.uint32)
    n_chains = max(off.size - 1, 0)
    counts, rates = np.zeros(6, np.uint64), np.zeros(3, np.float32)
    kept = np.zeros(max(n_chains, 1), np.uint8)
    lib(lib_path).oracle_tracks(res["hits.hit_id"].size, a(res["hits.hit_id"], np.uint32),
                                *[a(res[f"tracklets.h{i}"], np.uint32) for i in (1, 2, 3)], n_chains,
                                a(off if off.size else np.zeros(1), np.uint32), a(res["tracks"], np.uint32),
                                int(simulated), a(counts, np.uint64), a(rates, np.float32), a(kept, np.uint8))
    names = ("total", "filtered", "valids", "valid_unique", "fakes", "simulated")
    out = {n: int(v) for n, v in zip(names, counts)}
    out.update({n: float(v) for n, v in zip(("fake_rate", "clone_rate", "efficiency"), rates)})
    out["kept"] = kept[:n_chains].copy()
    return out
