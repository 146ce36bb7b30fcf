"""Pins of the oracle against what fixes its results independently of it:
the reference's prefix-sum known-answer test, closed forms, GridBuilder's own
verifyGrid invariants, brute force over all hit combinations evaluated in
float64, physics (true helix triplets must be found), and the hand-built
connectivity chains of the reference's test suite."""
import numpy as np
import pytest

import oracle
from helpers import hitset, small_workload
from trax_amd.config import geometry_arrays


# ---- PrefixSum -----------------------------------------------------------------
@pytest.mark.parametrize("wg", [1, 2, 4, 8, 16, 32, 64, 128, 256])
def test_prefix_sum_reference_kat(wg):
    # src/trax_test/trax_test.cpp:447-479: 1 Mi ones, out[N-1] == N-1 for every wg
    n = 1024 * 1024
    out = oracle.prefix_sum(np.ones(n, np.uint32), wg)
    assert out[-1] == n - 1
    assert np.array_equal(out[::4097], np.arange(0, n, 4097, dtype=np.uint32))


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 511, 512, 513, 100_003])
@pytest.mark.parametrize("wg", [64, 256])
def test_prefix_sum_closed_form(n, wg):
    v = np.random.default_rng(n).integers(0, 1 << 16, n, dtype=np.uint32)
    ref = np.concatenate([[0], np.cumsum(v.astype(np.uint64))[:-1]]).astype(np.uint32)
    assert np.array_equal(oracle.prefix_sum(v, wg), ref)


# ---- grid build -----------------------------------------------------------------
@pytest.fixture(scope="module")
def dense():
    wl = small_workload(2)
    wl, cfg, hs, evs = (wl, *hitset(wl, seed=3))
    cols = hs.columns()
    res = oracle.run(cols, wl, cfg.min_radius_curvature)
    return wl, cfg, cols, res, evs


def test_grid_verify_invariants(dense):
    """GridBuilder::verifyGrid (src/algorithms/GridBuilder.cpp:188-225): every hit
    of cell (z, p) lies on the cell's layer and inside its z and phi bounds."""
    wl, cfg, cols, res, _ = dense
    nz, nphi, L = wl.n_z, wl.n_phi, wl.n_layers
    grid = res["grid"].reshape(1, L, nz + 1, nphi + 1)
    x, y, z, layer = res["hits.x"], res["hits.y"], res["hits.z"], res["hits.layer"]
    sz = np.float32(600.0) / np.float32(nz)
    sp = np.float32(2 * np.float32(np.pi)) / np.float32(nphi)
    phi = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    checked = 0
    for l in range(L):
        g = grid[0, l]
        for zz in range(nz):
            for p in range(nphi):
                h = np.arange(g[zz, p], g[zz, p + 1])
                if h.size == 0:
                    continue
                checked += h.size
                assert np.all(layer[h] == l + 1)
                zlo, zhi = np.float32(-300) + np.float32(zz) * sz, np.float32(-300) + np.float32(zz + 1) * sz
                assert np.all((z[h] >= zlo) & (z[h] <= zhi))
                plo = -np.float32(np.pi) + np.float32(p) * sp
                phi_hi = -np.float32(np.pi) + np.float32(p + 1) * sp
                assert np.all((phi[h] >= plo - 1e-6) & (phi[h] <= phi_hi + 1e-6))
    assert checked == x.size  # no hit in the padding row or column


def test_grid_scatter_is_stable_permutation(dense):
    wl, cfg, cols, res, _ = dense
    key_in = np.stack([cols["x"].view(np.uint32), cols["y"].view(np.uint32), cols["z"].view(np.uint32),
                       cols["layer"], cols["det_id"], cols["hit_id"], cols["event"]], 1)
    key_out = np.stack([res["hits.x"].view(np.uint32), res["hits.y"].view(np.uint32),
                        res["hits.z"].view(np.uint32), res["hits.layer"], res["hits.det_id"],
                        res["hits.hit_id"], res["hits.event"]], 1)
    order_in = np.lexsort(key_in.T[::-1])
    order_out = np.lexsort(key_out.T[::-1])
    assert np.array_equal(key_in[order_in], key_out[order_out])
    # map output positions to input positions; inside each cell the input order is kept
    src = np.empty(key_in.shape[0], np.int64)
    src[order_out] = order_in
    g = res["grid"]
    starts = np.unique(g)
    for a, b in zip(starts[:-1], starts[1:]):
        assert np.all(np.diff(src[a:b]) > 0)


def test_grid_histogram_brute_force(dense):
    wl, cfg, cols, res, _ = dense
    nz, nphi, L = wl.n_z, wl.n_phi, wl.n_layers
    x, y, z = cols["x"].astype(np.float64), cols["y"].astype(np.float64), cols["z"].astype(np.float64)
    zs = np.floor((z + 300.0) / (600.0 / nz)).astype(np.int64)
    ps = np.floor((np.arctan2(y, x) + np.pi) / (2 * np.pi / nphi)).astype(np.int64)
    cell = (zs * (nphi + 1) + ps)
    lay = cols["layer"].astype(np.int64) - 1
    hist = np.zeros((L, (nz + 1) * (nphi + 1)), np.int64)
    np.add.at(hist, (lay, cell), 1)
    g = res["grid"].reshape(L, -1).astype(np.int64)
    ora = np.diff(np.concatenate([g.ravel(), [x.size]])).reshape(L, -1)
    # float32 vs float64 cell assignment may differ only for hits on a cell edge
    assert np.abs(ora - hist).sum() <= 4


# ---- pairs: brute force in float64 ---------------------------------------------
def _pair_window64(t, lay1, xo, yo, zo, rmin, rmax, mrc):
    """float64 evaluation of the pair window of PairGeneratorBeamspot.h:108-163
    for outer hits (vectorised): returns zLow, zHigh, phiLow, phiHigh, wrap."""
    z0, tip, sz = float(t.z0), float(t.tipCut), float(t.sigmaZ)
    r1min, r1max = float(rmin[lay1]), float(rmax[lay1])
    s = np.sign(yo)
    r = s * np.hypot(xo, yo)
    cth, ctl = (zo - z0) / r, (zo + z0) / r
    zh = np.maximum(s * r1max * cth + z0, s * r1min * cth + z0) + sz
    zl = np.minimum(s * r1max * ctl - z0, s * r1min * ctl - z0) - sz
    phi = np.arctan2(yo, xo)
    a = np.arccos(r / (2 * mrc))
    dphi = np.maximum(np.abs(a - np.arccos(r1max / (2 * mrc))), np.abs(a - np.arccos(r1min / (2 * mrc))))
    dphi = dphi + np.maximum(np.arctan(tip * (r - r1min)) / (r * r1min), np.arctan(tip * (r - r1max)) / (r * r1max))
    return zl, zh, phi - dphi, phi + dphi


def test_pairs_brute_force_float64():
    wl = small_workload(2, tracks=60, noise=600)
    wl, cfg, hs, _ = (wl, *hitset(wl, seed=21))
    res = oracle.run(hs.columns(), wl, cfg.min_radius_curvature, stop_after=2)
    x, y, z = (res[k].astype(np.float64) for k in ("hits.x", "hits.y", "hits.z"))
    layer = res["hits.layer"]
    _, rmin, rmax = geometry_arrays()
    got = set(map(tuple, res["pairs"].reshape(-1, 2)))
    near, missing, extra = 0, 0, 0
    for p, t in enumerate(wl.triplets):
        inner = np.nonzero(layer == t.layer1)[0]
        outer = np.nonzero(layer == t.layer2)[0]
        zl, zh, pl, ph = _pair_window64(t, t.layer1 - 1, x[outer], y[outer], z[outer], rmin.astype(np.float64),
                                        rmax.astype(np.float64), cfg.min_radius_curvature)
        zi, phii = z[inner][None, :], np.arctan2(y[inner], x[inner])[None, :]
        dphi = (phii - pl[:, None]) % (2 * np.pi)
        width = (ph - pl)[:, None]
        ok = (zi >= zl[:, None]) & (zi <= zh[:, None]) & (dphi <= width)
        margin = np.minimum.reduce([np.abs(zi - zl[:, None]), np.abs(zh[:, None] - zi),
                                    np.abs(dphi), np.abs(width - dphi)])
        for io in range(outer.size):
            for ii in range(inner.size):
                inset = (int(inner[ii]), int(outer[io])) in got
                if inset != bool(ok[io, ii]):
                    if margin[io, ii] < 1e-4:
                        near += 1
                    elif inset:
                        extra += 1
                    else:
                        missing += 1
    assert missing == 0 and extra == 0, (missing, extra, near)
    assert near <= 3


def test_predictor_brute_force_float64():
    """TripletThetaPhiPredictor (.h:77-230): for every pair, the third hits the
    oracle predicts are exactly the layer-3 hits of the same event inside the
    window recomputed in float64 (z from the transverse impact and cot(theta),
    phi from the hit-to-hit and curvature terms), up to hits within 1e-4 of a
    window edge.  Both rOrigin forms coincide in float64."""
    wl = small_workload(2, tracks=60, noise=600)
    wl, cfg, hs, _ = (wl, *hitset(wl, seed=23))
    res = oracle.run(hs.columns(), wl, cfg.min_radius_curvature, stop_after=3)
    x, y, z = (res[k].astype(np.float64) for k in ("hits.x", "hits.y", "hits.z"))
    layer = res["hits.layer"]
    _, rmin, rmax = geometry_arrays()
    mrc = float(cfg.min_radius_curvature)
    pairs = res["pairs"].reshape(-1, 2).astype(np.int64)
    cands = res["cands"].reshape(-1, 2).astype(np.int64)
    ok_c = cands[:, 0] != 0xFFFFFFFF
    got = {}
    for pi, h3 in cands[ok_c]:
        got.setdefault(int(pi), []).append(int(h3))
    phi_all = np.arctan2(y, x)
    near = missing = extra = checked = 0
    for pi, (h1, h2) in enumerate(pairs):
        if h1 == 0xFFFFFFFF:
            continue
        lt = int(layer[h1]) - 1
        t = wl.triplets[lt]
        l3 = t.layer3 - 1
        p1x, p1y, p1z, p2x, p2y, p2z = x[h1], y[h1], z[h1], x[h2], y[h2], z[h2]
        tip2 = (p1y * (p2x - p1x) - p1x * (p2y - p1y)) ** 2 / ((p2x - p1x) ** 2 + (p2y - p1y) ** 2)
        r0 = np.sqrt(p1x * p1x + p1y * p1y - tip2)
        cot = (p2z - p1z) / (np.sqrt(p2x * p2x + p2y * p2y - tip2) - r0)
        za = p1z + (np.sqrt(float(rmax[l3]) ** 2 - tip2) - r0) * cot
        zb = p1z + (np.sqrt(float(rmin[l3]) ** 2 - tip2) - r0) * cot
        zl, zh = min(za, zb) - float(t.sigmaZ), max(za, zb) + float(t.sigmaZ)
        phi = np.arctan2(p2y, p2x)
        dphi = abs((phi - np.arctan2(p1y, p1x) + np.pi) % (2 * np.pi) - np.pi)
        a = np.arccos(np.hypot(p2x - p1x, p2y - p1y) / (2 * mrc))
        m = max(abs(a - np.arccos(float(rmax[l3]) / (2 * mrc))), abs(a - np.arccos(float(rmin[l3]) / (2 * mrc))))
        half = max(dphi, m) + float(t.sigmaPhi)
        same = np.nonzero((layer == t.layer3) & (res["hits.event"] == res["hits.event"][h1]))[0]
        zi, di = z[same], np.abs((phi_all[same] - phi + np.pi) % (2 * np.pi) - np.pi)
        inside = (zi >= zl) & (zi <= zh) & (di <= half)
        margin = np.minimum.reduce([np.abs(zi - zl), np.abs(zh - zi), np.abs(half - di)])
        g = set(got.get(pi, []))
        for k, h in enumerate(same):
            checked += 1
            if (int(h) in g) != bool(inside[k]):
                if margin[k] < 1e-4:
                    near += 1
                elif int(h) in g:
                    extra += 1
                else:
                    missing += 1
    assert checked > 50_000
    assert missing == 0 and extra == 0, (missing, extra, near)
    assert near <= 5


def test_pair_order_is_reference_work_item_order(dense):
    """pairStore (PairGeneratorBeamspot.h:488-521): per (event, layer pair) the
    pairs of work item t (outer hits t, t + nThreads, ...) come first, inner hits
    ascending per outer hit."""
    wl, cfg, cols, res, _ = dense
    pairs = res["pairs"].reshape(-1, 2).astype(np.int64)
    off = res["pair_offsets"]
    grid = res["grid"].reshape(wl.n_layers, -1)
    nt = wl.n_threads
    for p, t in enumerate(wl.triplets):
        seg = pairs[off[p]:off[p + 1]]
        l2start = grid[t.layer2 - 1, 0]
        il = seg[:, 1] - l2start
        key = np.stack([il % nt, il // nt, seg[:, 0]], 1)
        order = np.lexsort(key.T[::-1])
        assert np.array_equal(order, np.arange(seg.shape[0]))
        assert np.all(np.diff(key[:, 0]) >= 0)


# ---- filter: FP64 adjudication ------------------------------------------------------
def _filter64(x, y, z, h1, h2, h3, t, mrc):
    a1 = np.arctan2(np.hypot(x[h2] - x[h1], y[h2] - y[h1]), z[h2] - z[h1])
    a2 = np.arctan2(np.hypot(x[h3] - x[h2], y[h3] - y[h2]), z[h3] - z[h2])
    dth = np.abs(a2 / a1)
    p1, p2 = np.arctan2(y[h2] - y[h1], x[h2] - x[h1]), np.arctan2(y[h3] - y[h2], x[h3] - x[h2])
    dph = (p2 - p1 + np.pi) % (2 * np.pi) - np.pi
    P = [np.stack([x[h], y[h], x[h] ** 2 + y[h] ** 2], 1) for h in (h1, h2, h3)]
    n = np.cross(P[1] - P[0], P[2] - P[0])
    n /= np.linalg.norm(n, axis=1)[:, None]
    c0 = -n[:, :2] / (2 * n[:, 2:3])
    c = -np.sum(n * P[0], 1)
    rad = np.sqrt((1 - n[:, 2] ** 2 - 4 * c * n[:, 2]) / (4 * n[:, 2] ** 2))
    v = -c0 / np.linalg.norm(c0, axis=1)[:, None]
    tip2 = np.sum((c0 + rad[:, None] * v) ** 2, 1)
    ok = (np.abs(dth - 1) <= t.dThetaCut) & (np.abs(dph) <= t.dPhiCut) & (tip2 <= t.tipCut ** 2) & (rad >= mrc)
    margin = np.minimum.reduce([np.abs(np.abs(dth - 1) - t.dThetaCut), np.abs(np.abs(dph) - t.dPhiCut),
                                np.abs(tip2 - t.tipCut ** 2) / max(t.tipCut ** 2, 1e-9), np.abs(rad - mrc) / mrc])
    return ok, margin


def _filter_mca(x, y, z, a, b, c, t, mrc, rng):
    """The filter cuts of one candidate in Monte Carlo arithmetic (Parker,
    Pierce and Eggert 2000): float64 with a random relative error of up to one
    float32 unit roundoff injected after every operation (two for the
    transcendental built-ins), so repeated samples span the results that float32
    rounding anywhere in the formula can produce."""
    u = 2.0 ** -24

    def r(v, k=1):
        return v * (1.0 + rng.uniform(-k, k) * u)

    sub = lambda p, q: r(p - q)
    mul = lambda p, q: r(p * q)
    add = lambda p, q: r(p + q)
    div = lambda p, q: r(p / q)
    sqrt = lambda p: r(np.sqrt(p))
    atan2 = lambda p, q: r(np.arctan2(p, q), 4)
    x1, y1, z1, x2, y2, z2, x3, y3, z3 = x[a], y[a], z[a], x[b], y[b], z[b], x[c], y[c], z[c]
    a1 = atan2(sqrt(add(mul(sub(x2, x1), sub(x2, x1)), mul(sub(y2, y1), sub(y2, y1)))), sub(z2, z1))
    a2 = atan2(sqrt(add(mul(sub(x3, x2), sub(x3, x2)), mul(sub(y3, y2), sub(y3, y2)))), sub(z3, z2))
    dth = abs(div(a2, a1))
    ok = (1 - t.dThetaCut <= dth <= 1 + t.dThetaCut)
    p1, p2 = atan2(sub(y2, y1), sub(x2, x1)), atan2(sub(y3, y2), sub(x3, x2))
    d = sub(p2, p1)
    d = d - 2 * np.pi if d > np.pi else (d + 2 * np.pi if d < -np.pi else d)
    ok &= abs(d) <= t.dPhiCut
    P = [(px, py, add(mul(px, px), mul(py, py))) for px, py in ((x1, y1), (x2, y2), (x3, y3))]
    A = [sub(P[1][i], P[0][i]) for i in range(3)]
    B = [sub(P[2][i], P[0][i]) for i in range(3)]
    n = [sub(mul(A[1], B[2]), mul(A[2], B[1])), sub(mul(A[2], B[0]), mul(A[0], B[2])),
         sub(mul(A[0], B[1]), mul(A[1], B[0]))]
    ln = sqrt(add(add(mul(n[0], n[0]), mul(n[1], n[1])), mul(n[2], n[2])))
    n = [div(v, ln) for v in n]
    cx, cy = div(-n[0], mul(2, n[2])), div(-n[1], mul(2, n[2]))
    cc = -add(add(mul(n[0], P[0][0]), mul(n[1], P[0][1])), mul(n[2], P[0][2]))
    rad = sqrt(div(sub(sub(1, mul(n[2], n[2])), mul(mul(4, cc), n[2])), mul(mul(4, n[2]), n[2])))
    lv = sqrt(add(mul(cx, cx), mul(cy, cy)))
    vx, vy = div(-cx, lv), div(-cy, lv)
    px, py = add(cx, mul(rad, vx)), add(cy, mul(rad, vy))
    tip2 = add(mul(px, px), mul(py, py))
    ok &= tip2 <= t.tipCut * t.tipCut
    ok &= rad >= mrc
    return bool(ok)


def test_filter_cuts_agree_with_float64_up_to_cut_edges(dense):
    wl, cfg, cols, res, _ = dense
    x, y, z = (res[k].astype(np.float64) for k in ("hits.x", "hits.y", "hits.z"))
    pairs = res["pairs"].reshape(-1, 2)
    cands = res["cands"].reshape(-1, 2)
    cands = cands[cands[:, 0] != 0xFFFFFFFF]
    h1, h2, h3 = pairs[cands[:, 0], 0], pairs[cands[:, 0], 1], cands[:, 1]
    lt = res["hits.layer"][h1] - 1
    kept = set(zip(res["tracklets.h1"].tolist(), res["tracklets.h2"].tolist(), res["tracklets.h3"].tolist()))
    in32 = np.array([(a, b, c) in kept for a, b, c in zip(h1.tolist(), h2.tolist(), h3.tolist())])
    # a float32 decision that differs from float64 must be reproduced by some
    # Monte Carlo arithmetic sample at float32 precision: the candidate then sits
    # within float32 rounding of a cut (the circle fit of near-straight tracks is
    # ill-conditioned); a dropped term or a wrong sign is not explained so
    rng = np.random.default_rng(1)
    disagree, unexplained, total = 0, 0, 0
    for k, t in enumerate(wl.triplets):
        m = np.nonzero(lt == k)[0]
        ok64, _ = _filter64(x, y, z, h1[m], h2[m], h3[m], t, cfg.min_radius_curvature)
        d = m[ok64 != in32[m]]
        total += m.size
        disagree += d.size
        for j in d:
            explained = any(_filter_mca(x, y, z, h1[j], h2[j], h3[j], t, cfg.min_radius_curvature, rng)
                            == bool(in32[j]) for _ in range(400))
            unexplained += not explained
    print(f"float32 vs float64 cut decisions: {disagree} of {total} candidates differ, {unexplained} unexplained")
    assert unexplained == 0
    assert disagree <= max(10, total // 100000)


def test_filter_disagreements_are_the_float32_expressions_own(dense):
    """The float64 adjudicator (trax_amd/adjudicate.py) on the oracle's filter
    decisions over a dense event: every candidate whose float64 decision
    differs is decided the same way by a bit-exact float32 host replica of the
    reference's cut expressions (numpy float32 with the library's atan2), so
    the difference is the float32 arithmetic's, and each lies within one unit
    of float32 rounding per operation of its cut's quantity."""
    from trax_amd.adjudicate import adjudicate
    wl, cfg, cols, res, _ = dense
    a = adjudicate(res, cfg, workers=4)
    n = a["float32_float64_disagreements"]
    assert a["candidates"] > 1_000_000 and n <= a["candidates"] // 100_000
    assert a["float32_replica_agrees"] == n
    assert a["within_float32_rounding"] == n
    assert n == sum(a["by_cut"].values())


def test_true_helix_triplets_are_found(dense):
    """Physics: consecutive-layer hit triplets of simulated helix tracks pass the
    reference's cuts (the published efficiencies are 0.79-0.84,
    plotting/data/physics/physics.ttBar.csv:2-6; these tracks are clean)."""
    wl, cfg, cols, res, _ = dense
    hid, layer = res["hits.hit_id"], res["hits.layer"]
    found = set(zip(res["tracklets.h1"].tolist(), res["tracklets.h2"].tolist(), res["tracklets.h3"].tolist()))
    true, hit = 0, 0
    by_track = {}
    for i in np.nonzero(hid > 0)[0]:
        by_track.setdefault(int(hid[i]), {})[int(layer[i])] = int(i)
    for hits in by_track.values():
        for t in wl.triplets:
            if all(l in hits for l in (t.layer1, t.layer2, t.layer3)):
                true += 1
                hit += (hits[t.layer1], hits[t.layer2], hits[t.layer3]) in found
    assert true > 1000
    assert hit / true > 0.95, hit / true


# ---- connectivity and CA on hand-built chains ------------------------------------
def _chain_fixture(triplets, n_layers, cot=0.7, phi0=0.3):
    """Hits numbered as in the fixture; hits of the chain sit on one straight
    line (equal eta), the others are scattered."""
    from trax_amd.config import BARREL_RADII_10
    hit_layer = {}
    for (a, b, c), lt in triplets:
        for h, l in ((a, lt + 1), (b, lt + 2), (c, lt + 3)):
            assert hit_layer.setdefault(h, l) == l
    ids = sorted(hit_layer, key=lambda h: (hit_layer[h], h))
    index = {h: i for i, h in enumerate(ids)}
    rng = np.random.default_rng(0)
    cols = {k: [] for k in ("x", "y", "z", "layer", "event")}
    for h in ids:
        r = BARREL_RADII_10[hit_layer[h] - 1]
        ph = phi0 if h in CHAIN_HITS else rng.uniform(-3, 3)
        ct = cot if h in CHAIN_HITS else rng.uniform(-2, 2)
        cols["x"].append(r * np.cos(ph)); cols["y"].append(r * np.sin(ph)); cols["z"].append(r * ct)
        cols["layer"].append(hit_layer[h]); cols["event"].append(0)
    tr = sorted(((lt, index[a], index[b], index[c]) for (a, b, c), lt in triplets))
    h1, h2, h3 = (np.array([t[k] for t in tr], np.uint32) for k in (1, 2, 3))
    return cols, h1, h2, h3, index


CHAIN_HITS = {1, 4, 6, 16, 30, 36}

# TripletConnectivityTest.h:59-79 (addFourLayerForwardTight); the last triplet's
# first hit 27 is renumbered 40 because one hit has one layer here
TIGHT = [((1, 4, 6), 0), ((1, 3, 5), 0), ((2, 3, 8), 0), ((2, 4, 7), 0),
         ((4, 6, 16), 1), ((9, 12, 17), 1), ((9, 13, 18), 1), ((9, 14, 19), 1),
         ((6, 16, 30), 2), ((15, 22, 29), 2), ((15, 24, 29), 2), ((15, 25, 27), 2),
         ((16, 30, 36), 3), ((28, 32, 37), 3), ((26, 33, 38), 3), ((40, 34, 39), 3)]


def test_connectivity_tight_four_layer_chain():
    wl = small_workload(2)
    wl.triplets = wl.triplets[:4]
    cols, h1, h2, h3, index = _chain_fixture(TIGHT, 6)
    from trax_amd.config import min_radius_curvature
    res = oracle.run_link(cols, wl, min_radius_curvature(1.0), h1, h2, h3)
    tid = {tuple(int(v) for v in t): i for i, t in enumerate(zip(h1, h2, h3))}
    chain = [tid[(index[a], index[b], index[c])] for a, b, c in ((1, 4, 6), (4, 6, 16), (6, 16, 30), (16, 30, 36))]
    links = set(zip(res["conn.basis"].tolist(), res["conn.followers"].tolist()))
    assert links == {(chain[0], chain[1]), (chain[1], chain[2]), (chain[2], chain[3])}
    # the reference test expects connectivity 3 for the last chain triplet
    # (TripletConnectivityTest.h:243-250): three predecessors, CA state 4
    states = res["ca.states"]
    assert [int(states[t]) for t in chain] == [1, 2, 3, 4]
    off = res["track_offsets"]
    lens = np.diff(off)
    assert [int(lens[t]) for t in chain] == [0, 0, 0, 4]
    assert res["tracks"][off[chain[3]]:off[chain[3] + 1]].tolist() == chain[::-1]
    # the last tracklet of the collection is neither base nor follower (.cpp:277)
    last = len(h1) - 1
    assert last not in res["conn.connectable"] and last not in res["conn.nonconnectable"]


def test_ca_states_are_longest_chains(dense):
    """The CA forward phase (CellularAutomaton.h:52-100), iterated until no
    cell lives, leaves each tracklet at 1 + the length of the longest chain of
    connections ending at it, and stops after max + 1 iterations: a closed
    form, computed here over the connections in the layer order of their
    bases (every connection leads to a higher layer triplet)."""
    wl, cfg, cols, res, _ = dense
    basis = res["conn.basis"].astype(np.int64)
    foll = res["conn.followers"].astype(np.int64)
    th1 = res["tracklets.h1"].astype(np.int64)
    lay = res["hits.layer"].astype(np.int64)
    assert basis.size > 1000
    lt_b, lt_f = lay[th1[basis]], lay[th1[foll]]
    assert np.all(lt_f > lt_b)
    st = np.ones(th1.size, np.int64)
    for lt in np.unique(lt_b):
        m = lt_b == lt
        np.maximum.at(st, foll[m], st[basis[m]] + 1)
    assert np.array_equal(st, res["ca.states"].astype(np.int64))
    assert int(res["ca.iterations"][0]) == int(st.max())


def test_ca_best_basis_breaks_pt_ties_by_lowest_index():
    # two bases with the same circle (identical hits) feeding one follower
    wl = small_workload(2)
    wl.triplets = wl.triplets[:3]
    trip = [((1, 4, 6), 0), ((2, 4, 6), 0), ((4, 6, 16), 1), ((6, 16, 30), 2), ((50, 51, 52), 2)]
    cols, h1, h2, h3, index = _chain_fixture(trip, 5)
    # hit 2 duplicates hit 1's position, so both bases have the same pT
    i1, i2 = index[1], index[2]
    for k in ("x", "y", "z"):
        cols[k][i2] = cols[k][i1]
    from trax_amd.config import min_radius_curvature
    res = oracle.run_link(cols, wl, min_radius_curvature(1.0), h1, h2, h3)
    tid = {tuple(int(v) for v in t): i for i, t in enumerate(zip(h1, h2, h3))}
    f = tid[(index[4], index[6], index[16])]
    b1, b2 = tid[(i1, index[4], index[6])], tid[(i2, index[4], index[6])]
    assert res["ca.best_basis"][f] == min(b1, b2)
