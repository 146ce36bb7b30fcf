"""Independent pins of the oracle's linking half: TripletConnectivityTight's
eta and dEta cut, the follower ranges and predicate lists, TrackletCircleFitter's
radius and CellularAutomaton's best basis and tracks.

Each check computes its answer without the oracle's formulas:
- eta in float64 from the first to the third hit, links by brute force over
  all tracklet pairs that share two hits, decided in float64 away from the cut;
- the circle radius as the circumradius of the three (x, y) points (side lengths
  and area), not through the paraboloid plane;
- the best basis as the minimum of (pT difference, basis index) over each
  follower's candidates, and the tracks as chains of best bases.
An exact-cut fixture (|delta eta| equal to dEtaCut in float32) pins the strict
comparison, which no random event reaches.  The last test builds the oracle
with each planted mistake of its mutation hooks and shows that the matching
pin fails on it.
"""
import numpy as np
import pytest

import oracle
from helpers import hitset, small_workload
from trax_amd.api import math_host
from trax_amd.config import min_radius_curvature


def _dense_event(tracks=2000, noise=20000, seed=3):
    wl = small_workload(2, tracks=tracks, noise=noise, events=1)
    cfg, hs, _ = hitset(wl, seed=seed)
    return wl, cfg, hs.columns()


@pytest.fixture(scope="module")
def dense_link():
    wl, cfg, cols = _dense_event()
    return wl, cfg, cols, oracle.run(cols, wl, cfg.min_radius_curvature)


# ---- checks (each returns a list of failure reports) -------------------------------
def link_failures(res, wl, n_events=1, margin=1e-5):
    """tripletEtaCalculatorStore / tripletConnectivityTightCount+Store
    (TripletConnectivityTight.h:47-243, .cpp:184-307) against float64 brute
    force, for events whose every (event, layer triplet) group holds tracklets:
    - eta_t = asinh(dz / rho) from the first to the third hit (.h:60-75);
    - tracklet offsets: each group's first tracklet, then R - 1 (the filter's
      last-tracklet entry, TripletThetaPhiFilter.h:250-253);
    - base t (not in the last group, .h:113-115) links follower i (not the last
      tracklet) iff h1_i = h2_t, h2_i = h3_t and |eta_t - eta_i| < dEtaCut;
      pairs within `margin` of the cut in float64 are not compared;
    - the predicate lists split tracklets 0 .. R - 2 by whether they take part
      in a link (.cpp:184-256)."""
    fails = []
    T = len(wl.triplets)
    x, y, z = (res["hits." + k].astype(np.float64) for k in "xyz")
    ev, lay = res["hits.event"].astype(np.int64), res["hits.layer"].astype(np.int64)
    h1, h2, h3 = (res["tracklets.h" + k].astype(np.int64) for k in "123")
    R = h1.size
    eta64 = np.arcsinh((z[h3] - z[h1]) / np.hypot(x[h3] - x[h1], y[h3] - y[h1]))
    eta32 = res["eta"].astype(np.float64)
    bad = np.abs(eta32 - eta64) > 2e-6 * np.maximum(1.0, np.abs(eta64))
    if bad.any():
        fails.append(f"eta: {int(bad.sum())} of {R} differ from float64")
    key = ev[h1] * T + lay[h1] - 1
    G = n_events * T
    assert np.all(np.bincount(key, minlength=G)[:G] > 0), "every group must hold tracklets"
    starts = np.searchsorted(key, np.arange(G))
    want_off = np.append(starts, R - 1)
    if not np.array_equal(res["tracklet_offsets"].astype(np.int64), want_off):
        fails.append("tracklet offsets differ from the group starts + R - 1")
    # every (base, follower) pair sharing two hits
    n_h = x.size
    kf = h1 * n_h + h2
    order = np.argsort(kf, kind="stable")
    kb = h2 * n_h + h3
    lo, hi = np.searchsorted(kf[order], kb, "left"), np.searchsorted(kf[order], kb, "right")
    nb = hi - lo
    base = np.repeat(np.arange(R), nb)
    foll = order[np.concatenate([np.arange(a, b) for a, b in zip(lo, hi)]) if nb.sum() else np.zeros(0, np.int64)]
    keep = (base < starts[G - 1]) & (foll < R - 1)
    base, foll = base[keep], foll[keep]
    de = np.abs(eta64[base] - eta64[foll])
    cut = float(np.float32(wl.d_eta_cut))
    near = np.abs(de - cut) < margin
    want = set(zip(base[(de < cut) & ~near].tolist(), foll[(de < cut) & ~near].tolist()))
    skip = set(zip(base[near].tolist(), foll[near].tolist()))
    got = set(zip(res["conn.basis"].tolist(), res["conn.followers"].tolist()))
    if len(got) != res["conn.basis"].size:
        fails.append("repeated links")
    extra, missing = (got - skip) - want, want - got
    if extra or missing:
        fails.append(f"links: {len(extra)} extra, {len(missing)} missing of {len(want)}")
    if len(skip) > max(5, len(want) // 1000):
        fails.append(f"{len(skip)} pairs within the margin of the cut")
    # predicate lists over tracklets 0 .. R - 2
    part = np.zeros(R, bool)
    part[res["conn.basis"].astype(np.int64)] = True
    part[res["conn.followers"].astype(np.int64)] = True
    idx = np.arange(R - 1)
    if got:
        if not (np.array_equal(res["conn.connectable"], idx[part[:-1]]) and
                np.array_equal(res["conn.nonconnectable"], idx[~part[:-1]])):
            fails.append("predicate lists")
    return fails, len(want)


def circle_failures(res, rel_median=1e-5, rel_p99=5e-4):
    """trackletCircleFitterStore (TrackletCircleFitter.h:47-97) against the
    circumradius of the three hits, |a||b||c| / (2 |cross|) in float64."""
    conn = res["conn.connectable"].astype(np.int64)
    x, y = res["hits.x"].astype(np.float64), res["hits.y"].astype(np.float64)
    t1, t2, t3 = (res["tracklets.h" + k].astype(np.int64)[conn] for k in "123")
    ax, ay, bx, by, cx, cy = x[t1], y[t1], x[t2], y[t2], x[t3], y[t3]
    la, lb, lc = np.hypot(bx - cx, by - cy), np.hypot(ax - cx, ay - cy), np.hypot(ax - bx, ay - by)
    area2 = np.abs((bx - ax) * (cy - ay) - (by - ay) * (cx - ax))
    r64 = la * lb * lc / (2 * area2)
    pt = res["pt"].astype(np.float64)[conn]
    rel = np.abs(pt - r64) / r64
    rel = np.where(np.isfinite(rel), rel, np.inf)
    fails = []
    if np.median(rel) > rel_median or np.quantile(rel, 0.99) > rel_p99:
        fails.append(f"circle radius: median {np.median(rel):.2e}, p99 {np.quantile(rel, 0.99):.2e} relative error")
    return fails, conn.size


def ca_failures(res):
    """followerBestBasisStore, handlersStateStore and trackCollectionStore
    (CellularAutomaton.h:224-323): each follower's best basis is, among the
    connections whose basis state is one below its own, the basis with the
    smallest |pT_b - pT_f| (float32), the lowest basis index on ties (a NaN
    difference at the lowest basis is never replaced, as `d < min` is false for
    it); a track starts at every tracklet that is nobody's best basis, holds
    its state's number of tracklets and follows the best bases down to state 1."""
    fails = []
    basis, foll = res["conn.basis"].astype(np.int64), res["conn.followers"].astype(np.int64)
    st = res["ca.states"].astype(np.int64)
    pt = res["pt"]
    R = st.size
    m = st[basis] + 1 == st[foll]
    b, f = basis[m], foll[m]
    d = np.abs(pt[b] - pt[f])  # float32, as the reference computes it
    want = np.full(R, 0xFFFFFFFF, np.int64)
    o = np.lexsort((b, f))  # by follower, then ascending basis
    b, f, d = b[o], f[o], d[o]
    first = np.ones(f.size, bool)
    first[1:] = f[1:] != f[:-1]
    for s, e in zip(np.nonzero(first)[0], np.append(np.nonzero(first)[0][1:], f.size)):
        dd = d[s:e]
        if np.isnan(dd[0]):
            want[f[s]] = b[s]
            continue
        ok = ~np.isnan(dd)
        k = np.lexsort((b[s:e][ok], dd[ok]))[0]
        want[f[s]] = b[s:e][ok][k]
    if not np.array_equal(res["ca.best_basis"].astype(np.int64), want):
        fails.append(f"best basis: {int((res['ca.best_basis'].astype(np.int64) != want).sum())} of {R} differ")
    best = want
    isbest = np.zeros(R, bool)
    isbest[best[best < R]] = True
    length = np.where(isbest, 0, st)
    off = np.concatenate([[0], np.cumsum(length)])
    if not np.array_equal(res["track_offsets"].astype(np.int64), off):
        fails.append("track offsets")
        return fails
    trk = res["tracks"].astype(np.int64)
    for t in np.nonzero(length)[0][:5000]:
        chain = trk[off[t]:off[t + 1]]
        if chain[0] != t or np.any(chain[1:] != best[chain[:-1]]) or np.any(np.diff(st[chain]) != -1) or \
                st[chain[-1]] != 1:
            fails.append(f"track of tracklet {t}: {chain.tolist()}")
            break
    return fails


# ---- pins -------------------------------------------------------------------------
def test_links_brute_force_float64(dense_link):
    wl, cfg, cols, res = dense_link
    fails, n_links = link_failures(res, wl)
    assert not fails, fails
    assert n_links > 500


def test_circle_radius_is_the_circumradius(dense_link):
    wl, cfg, cols, res = dense_link
    fails, n = circle_failures(res)
    assert not fails, fails
    assert n > 500


def test_ca_best_basis_and_tracks_brute_force(dense_link):
    wl, cfg, cols, res = dense_link
    assert res["ca.iterations"][0] >= 3
    assert not ca_failures(res)


def _edge_fixture():
    """Base A-B-C and two followers B-C-D, B-C-D' whose float32 |delta eta|
    is exactly dEtaCut and just below it; a last tracklet of a third layer
    triplet takes no part.  Every difference vector has rho = 1 exactly and an
    exact dz, so eta = asinh(dz): the base's dz is one ulp of the cut (asinh is
    the identity there), D's gives the float one ulp above the cut, D''s one
    ulp below it."""
    cut = np.float32(0.0256)
    s = np.spacing(cut)
    q0 = np.float32(np.sinh(np.float64(cut)))
    qs = (q0 + np.arange(-400, 401, dtype=np.int64).astype(np.float64) * np.spacing(q0)).astype(np.float32)
    ev = math_host(3, qs)
    up, down = qs[ev == cut + s], qs[ev == cut - s]
    assert up.size and down.size
    x = np.array([1.0, 1.5, 2.0, 2.5, 2.5, 2.0, 2.5, 3.0], np.float32)
    y = np.array([0.1, 0.25, 0.1, 0.25, 0.25, 1.0, 1.2, 1.4], np.float32)
    z = np.array([0.0, 0.0, s, up[0], down[0], 0.0, 0.0, 0.0], np.float32)
    layer = np.array([1, 2, 3, 4, 4, 3, 4, 5], np.uint32)
    hits = dict(x=x, y=y, z=z, layer=layer, event=np.zeros(8, np.uint32))
    h1 = np.array([0, 1, 1, 5], np.uint32)  # A-B-C; B-C-D; B-C-D'; E-F-G
    h2 = np.array([1, 2, 2, 6], np.uint32)
    h3 = np.array([2, 3, 4, 7], np.uint32)
    return hits, h1, h2, h3


def edge_failures(lib_path=None):
    wl = small_workload(2)
    wl.triplets = wl.triplets[:3]
    hits, h1, h2, h3 = _edge_fixture()
    res = oracle.run_link(hits, wl, min_radius_curvature(1.0), h1, h2, h3, lib_path=lib_path)
    fails = []
    cut = np.float32(wl.d_eta_cut)
    e = res["eta"]
    if np.abs(e[0] - e[1]) != cut or not np.abs(e[0] - e[2]) < cut:
        fails.append(f"fixture eta {e.tolist()}")
    links = sorted(zip(res["conn.basis"].tolist(), res["conn.followers"].tolist()))
    if links != [(0, 2)]:
        fails.append(f"links {links} (want only the follower below the cut)")
    return fails


def test_eta_cut_is_strict_at_the_cut():
    assert not edge_failures()


def test_pins_catch_planted_mistakes():
    """Each mutation hook of the oracle (a planted mistake) fails its pin:
    1 `<=` at the eta cut, 2 eta from h1 to h2, 3 the circle offset without the
    paraboloid lift, 4 the best basis by the largest pT difference."""
    wl, cfg, cols = _dense_event(tracks=800, noise=6000, seed=5)
    base = oracle.run(cols, wl, cfg.min_radius_curvature)
    assert not link_failures(base, wl)[0] and not circle_failures(base)[0] and not ca_failures(base)
    assert not edge_failures()
    caught = {}
    for plant in (1, 2, 3, 4):
        path = oracle.build(plant)
        if plant == 1:
            caught[plant] = bool(edge_failures(path))
            continue
        res = oracle.run(cols, wl, cfg.min_radius_curvature, lib_path=path)
        check = {2: lambda r: link_failures(r, wl)[0], 3: lambda r: circle_failures(r)[0], 4: ca_failures}[plant]
        caught[plant] = bool(check(res))
    assert caught == {1: True, 2: True, 3: True, 4: True}, caught
