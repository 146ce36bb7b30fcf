"""Float64 adjudication of the filter's float32 decisions
(TripletThetaPhiFilter.h:83-150).

Every stored candidate's four cuts are evaluated in float64 from the float32
hit coordinates (the library's threaded host evaluation); a candidate whose
float64 decision differs from the GPU's float32 one (its bit in the filter's
bitmap, or its membership in the tracklets) is a disagreement.  For each
disagreement the cut that decides it is named and its distance to the cut is
reported two ways:

- in float32 ulps of the cut value: |q64 - cut| / ulp(cut), with q the cut's
  quantity (|theta2 / theta1| against 1 -+ dThetaCut, |dphi| against dPhiCut,
  the distance of closest approach against tipCut, the radius against the
  minimum radius of curvature);
- in float32 ulps of the inputs: the smallest k such that moving each of the
  nine hit coordinates by at most k of its own float32 ulps can take q64 to
  the cut, k = |q64 - cut| / sum_i |dq/dx_i| ulp(x_i) (central differences
  of one ulp).  k <= 1 means the decision is within the rounding of the
  inputs themselves.

- in units of the quantity's own float32 rounding: the reference's float32
  expression re-evaluated in Monte Carlo arithmetic (Parker, Pierce and
  Eggert 2000: float64 with a random relative error of up to one float32
  unit roundoff, 2^-24, after every operation, and two after atan2, the
  shared built-in's measured bound), 256 samples; the distance is
  |q64 - cut| over half the samples' spread.  A distance of at most 1 means
  the cut lies within one unit of float32 rounding per operation of the
  quantity: its float32 value can fall on either side, and the disagreement
  is explained by rounding alone.  The circle cuts of near-straight tracks are
  ill-conditioned (the closest approach is the difference of two large
  numbers), so their spread is wide in ulps of the cut.

Which cut decides a disagreement is read from the kernels' own float32
values: a host replica of the cut expressions, bit-exact with the device
(numpy float32, whose IEEE operations round as the device's, and the
library's atan2), gives every cut's float32 quantity and decision; the cut
whose float32 and float64 decisions differ is the one reported, and the
replica's overall decision is checked to be the GPU's.
"""
from typing import Dict, List

import numpy as np

CUTS = ("theta", "phi", "tip", "radius")


def _quantities(p1, p2, p3):
    """The four cut quantities in float64 for hits p1, p2, p3 (arrays of shape
    (n, 3)): theta ratio, |dphi|, distance of closest approach, radius."""
    x1, y1, z1 = p1[:, 0], p1[:, 1], p1[:, 2]
    x2, y2, z2 = p2[:, 0], p2[:, 1], p2[:, 2]
    x3, y3, z3 = p3[:, 0], p3[:, 1], p3[:, 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        th1 = np.arctan2(np.hypot(x2 - x1, y2 - y1), z2 - z1)
        th2 = np.arctan2(np.hypot(x3 - x2, y3 - y2), z3 - z2)
        ratio = np.abs(th2 / th1)
        d = np.arctan2(y3 - y2, x3 - x2) - np.arctan2(y2 - y1, x2 - x1)
        d = np.where(d > np.pi, d - 2 * np.pi, np.where(d < -np.pi, d + 2 * np.pi, d))
        q1 = np.stack([x1, y1, x1 * x1 + y1 * y1], 1)
        q2 = np.stack([x2, y2, x2 * x2 + y2 * y2], 1)
        q3 = np.stack([x3, y3, x3 * x3 + y3 * y3], 1)
        n = np.cross(q2 - q1, q3 - q1)
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        cx, cy = -n[:, 0] / (2 * n[:, 2]), -n[:, 1] / (2 * n[:, 2])
        c = -(n * q1).sum(1)
        radius = np.sqrt((1 - n[:, 2] ** 2 - 4 * c * n[:, 2]) / (4 * n[:, 2] ** 2))
        # closest approach to the origin: |centre| - radius in absolute value
        tip = np.abs(np.hypot(cx, cy) - radius)
    return ratio, np.abs(d), tip, radius


def _mca_quantities(p1, p2, p3, rng, n=256):
    """The four cut quantities of one candidate as the float32 expression of
    the filter computes them (TripletThetaPhiFilter.h:83-150, the kernel's
    order of operations), in Monte Carlo arithmetic over n samples: theta
    ratio, |dphi|, closest approach (sqrt of the squared one) and radius."""
    u = 2.0 ** -24

    def r(v, k=1):
        return v * (1.0 + rng.uniform(-k, k, n) * u)

    sub = lambda a, b: r(a - b)
    mul = lambda a, b: r(a * b)
    add = lambda a, b: r(a + b)
    div = lambda a, b: r(a / b)
    sqrt = lambda a: r(np.sqrt(a))
    atan2 = lambda a, b: r(np.arctan2(a, b), 2)
    (x1, y1, z1), (x2, y2, z2), (x3, y3, z3) = p1, p2, p3
    with np.errstate(divide="ignore", invalid="ignore"):
        a1 = atan2(sqrt(add(mul(sub(x2, x1), sub(x2, x1)), mul(sub(y2, y1), sub(y2, y1)))), sub(z2, z1))
        a2 = atan2(sqrt(add(mul(sub(x3, x2), sub(x3, x2)), mul(sub(y3, y2), sub(y3, y2)))), sub(z3, z2))
        ratio = np.abs(div(a2, a1))
        d = sub(atan2(sub(y3, y2), sub(x3, x2)), atan2(sub(y2, y1), sub(x2, x1)))
        d = np.where(d > np.pi, d - 2 * np.pi, np.where(d < -np.pi, d + 2 * np.pi, d))
        P = [(px, py, add(mul(px, px), mul(py, py))) for px, py in ((x1, y1), (x2, y2), (x3, y3))]
        A = [sub(P[1][i], P[0][i]) for i in range(3)]
        B = [sub(P[2][i], P[0][i]) for i in range(3)]
        nv = [sub(mul(A[1], B[2]), mul(A[2], B[1])), sub(mul(A[2], B[0]), mul(A[0], B[2])),
              sub(mul(A[0], B[1]), mul(A[1], B[0]))]
        rn = div(1.0, sqrt(add(add(mul(nv[0], nv[0]), mul(nv[1], nv[1])), mul(nv[2], nv[2]))))
        nv = [mul(v, rn) for v in nv]
        cx, cy = div(-nv[0], mul(2.0, nv[2])), div(-nv[1], mul(2.0, nv[2]))
        cc = -add(add(mul(nv[0], P[0][0]), mul(nv[1], P[0][1])), mul(nv[2], P[0][2]))
        rad = sqrt(div(sub(sub(1.0, mul(nv[2], nv[2])), mul(mul(4.0, cc), nv[2])), mul(mul(4.0, nv[2]), nv[2])))
        rv = div(1.0, sqrt(add(mul(cx, cx), mul(cy, cy))))
        vx, vy = mul(-cx, rv), mul(-cy, rv)
        px, py = add(cx, mul(rad, vx)), add(cy, mul(rad, vy))
        tip = np.sqrt(add(mul(px, px), mul(py, py)))
    return ratio, np.abs(d), tip, rad


def _float32_quantities(p1, p2, p3):
    """The four cut quantities as the filter kernels compute them, bit for bit:
    the float32 expressions of filter_cuts.h (the reference's order of
    operations, TripletThetaPhiFilter.h:83-150) in numpy float32, whose
    +, -, *, / and sqrt round as the device's do (no contraction), with the
    library's own atan2 (the same built-in the kernels call).  Returns the
    theta ratio, |dphi|, the squared closest approach and the radius."""
    from .api import math_host
    f = np.float32
    x1, y1, z1 = (p1[:, k].astype(f) for k in range(3))
    x2, y2, z2 = (p2[:, k].astype(f) for k in range(3))
    x3, y3, z3 = (p3[:, k].astype(f) for k in range(3))
    atan2 = lambda yy, xx: math_host(0, np.ascontiguousarray(xx, f), np.ascontiguousarray(yy, f))
    PI = f(np.pi)
    with np.errstate(all="ignore"):
        th1 = atan2(np.sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1)), z2 - z1)
        th2 = atan2(np.sqrt((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2)), z3 - z2)
        ratio = np.abs(th2 / th1)
        d = atan2(y3 - y2, x3 - x2) - atan2(y2 - y1, x2 - x1)
        d = d + np.where(d > PI, f(-2) * PI, np.where(d < -PI, f(2) * PI, f(0)))
        p1z, p2z, p3z = x1 * x1 + y1 * y1, x2 * x2 + y2 * y2, x3 * x3 + y3 * y3
        ax, ay, az = x2 - x1, y2 - y1, p2z - p1z
        bx, by, bz = x3 - x1, y3 - y1, p3z - p1z
        nx, ny, nz = ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx
        dd = nx * nx + ny * ny + nz * nz
        rn = f(1) / np.where(dd != 0, np.sqrt(dd), f(1))  # normalize: one reciprocal, three multiplies
        nx, ny, nz = (np.where(dd != 0, v * rn, v) for v in (nx, ny, nz))
        cox, coy = (-nx) / (f(2) * nz), (-ny) / (f(2) * nz)
        cc = -(nx * x1 + ny * y1 + nz * p1z)
        rad = np.sqrt((f(1) - nz * nz - f(4) * cc * nz) / (f(4) * nz * nz))
        vx, vy = -cox, -coy
        dv = vx * vx + vy * vy
        rv = f(1) / np.where(dv != 0, np.sqrt(dv), f(1))
        vx, vy = np.where(dv != 0, vx * rv, vx), np.where(dv != 0, vy * rv, vy)
        px, py = cox + rad * vx, coy + rad * vy
        tip2 = px * px + py * py
    return ratio, np.abs(d), tip2, rad


def _passes32(q, thr32):
    """Per cut, the float32 comparisons of the kernels."""
    ratio, dphi, tip2, rad = q
    lo, hi, phic, tipc, rmin = thr32
    return [(lo <= ratio) & (ratio <= hi), dphi <= phic, tip2 <= tipc * tipc, rad >= rmin]


def _passes(q, thr):
    """Per cut, whether quantity q passes (thr: per-candidate thresholds)."""
    ratio, dphi, tip, radius = q
    lo, hi, phic, tipc, rmin = thr
    return [(lo <= ratio) & (ratio <= hi), dphi <= phic, tip <= tipc, radius >= rmin]


def _distance(q, thr):
    """Per cut, the distance of q from the cut's value (the nearer bound of the theta window)."""
    ratio, dphi, tip, radius = q
    lo, hi, phic, tipc, rmin = thr
    dth = np.minimum(np.abs(ratio - lo), np.abs(ratio - hi))
    return [dth, np.abs(dphi - phic), np.abs(tip - tipc), np.abs(radius - rmin)]


def _cut_value(c, thr, ratio):
    """The value of cut c (the nearer bound of the theta window)."""
    lo, hi, phic, tipc, rmin = thr
    if c == 0:
        return np.where(np.abs(ratio - lo) < np.abs(ratio - hi), lo, hi)
    return (phic, tipc, rmin)[c - 1]


def _key(a, b, c):
    return (a.astype(np.int64) << 42) | (b.astype(np.int64) << 21) | c.astype(np.int64)


def tracklet_keys(res: Dict[str, np.ndarray]) -> np.ndarray:
    """The sorted keys of the tracklets: 21 bits per hit index, so the hits
    must number at most 2^21 (larger results carry "filter.bits")."""
    if res["hits.x"].size > 1 << 21:
        raise ValueError("tracklet keys hold hit indices below 2^21: pass the filter's bits (\"filter.bits\")")
    return np.sort(_key(res["tracklets.h1"], res["tracklets.h2"], res["tracklets.h3"]))


def screen(res: Dict[str, np.ndarray], cfg, workers: int = 16, progress=None):
    """The float64 decision of every stored candidate, from the library's
    threaded host evaluation of the four cuts (trax_filter_audit_f64), and
    the device's decision: the filter's bitmap when the result holds it
    ("filter.bits", batches that store the candidates), else membership of
    (h1, h2, h3) in the tracklets.  Returns (indices of the stored
    candidates, float64 passes, device kept), the last two over those
    indices."""
    import ctypes as C
    from .api import check, lib, ptr
    cands = np.ascontiguousarray(res["cands"], np.uint32)
    pairs = np.ascontiguousarray(res["pairs"], np.uint32)
    cols = [np.ascontiguousarray(res[k], np.float32) for k in ("hits.x", "hits.y", "hits.z")]
    layer = np.ascontiguousarray(res["hits.layer"], np.uint32)
    n = cands.size // 2
    passes = np.zeros(max(n, 1), np.uint8)
    check(lib().trax_filter_audit_f64(cfg.h, n, ptr(cands), pairs.size // 2, ptr(pairs), layer.size,
                                      *[ptr(c) for c in cols], ptr(layer), C.c_float(cfg.min_radius_curvature),
                                      int(workers), ptr(passes)))
    passes = passes[:n]
    if progress:
        progress(f"float64 cuts of {n} candidate slots evaluated")
    idx = np.nonzero(passes != 0x80)[0]
    ok = (passes[idx] & 15) == 15
    if "filter.bits" in res:
        bits = np.unpackbits(np.ascontiguousarray(res["filter.bits"], np.uint32).view(np.uint8), bitorder="little")
        kept = bits[idx].astype(bool)
    else:
        c = cands.reshape(-1, 2)[idx].astype(np.int64)
        pr = pairs.reshape(-1, 2)
        keys = tracklet_keys(res)
        ck = _key(pr[c[:, 0], 0], pr[c[:, 0], 1], c[:, 1])
        pos = np.minimum(np.searchsorted(keys, ck), max(keys.size - 1, 0))
        kept = (keys[pos] == ck) if keys.size else np.zeros(ck.size, bool)
    return idx, ok, kept


def adjudicate(res: Dict[str, np.ndarray], cfg, max_report: int = 64, seed: int = 1, workers: int = 16,
               max_examined: int = 100_000, progress=None) -> Dict:
    """res: a batch or plain result with the candidates stored ("cands", "pairs",
    "hits.*", and "filter.bits" or "tracklets.*"); cfg: its
    trax_amd.api.Config.  Every candidate is screened in float64 (screen);
    the disagreements with the device are examined one by one.  Returns
    counts, the largest distances and up to max_report disagreements in
    detail.  At most max_examined disagreements are examined ("unexamined"
    counts the rest)."""
    triplets, min_radius_curvature = cfg.wl.triplets, cfg.min_radius_curvature
    idx, ok64, kept_all = screen(res, cfg, workers, progress)
    sel = np.nonzero(ok64 != kept_all)[0]
    n_dis = int(sel.size)
    if progress:
        progress(f"{idx.size} candidates, {int(kept_all.sum())} kept, {n_dis} disagreements")
    sel = sel[:max_examined]
    out = {"candidates": int(idx.size), "tracklets": int(kept_all.sum()),
           "float32_float64_disagreements": n_dis, "unexamined": n_dis - int(sel.size),
           "by_cut": {c: 0 for c in CUTS}, "max_ulp_of_cut": None,
           "max_ulp_of_inputs": None, "max_rounding_units": None, "within_1_ulp_of_inputs": 0,
           "within_float32_rounding": 0, "float32_replica_agrees": 0, "details": []}
    # the disagreements alone, examined in float64 with numpy
    x, y, z = (res[k].astype(np.float64) for k in ("hits.x", "hits.y", "hits.z"))
    pairs = res["pairs"].reshape(-1, 2)
    cands = res["cands"].reshape(-1, 2)[idx[sel]].astype(np.int64)
    kept = kept_all[sel]
    h1, h2, h3 = pairs[cands[:, 0], 0], pairs[cands[:, 0], 1], cands[:, 1]
    lt = res["hits.layer"][h1].astype(np.int64) - 1
    f32 = np.float32
    dth = np.array([t.dThetaCut for t in triplets], f32)[lt]
    # the thresholds as the float32 kernel compares against them
    lo = (f32(1) - dth).astype(np.float64)
    hi = (f32(1) + dth).astype(np.float64)
    phic = np.array([t.dPhiCut for t in triplets], f32)[lt].astype(np.float64)
    tipc = np.array([t.tipCut for t in triplets], f32)[lt].astype(np.float64)
    rmin = np.full(h1.size, float(f32(min_radius_curvature)))
    thr = (lo, hi, phic, tipc, rmin)
    P = [np.stack([x[h], y[h], z[h]], 1) for h in (h1, h2, h3)]
    q = _quantities(*P)
    ok_cut = _passes(q, thr)
    bad = np.arange(sel.size)
    if bad.size == 0:
        return out
    rng = np.random.default_rng(seed)
    dist = _distance(q, thr)
    rows: List[Dict] = []
    # the kernels' own float32 values of the disagreeing candidates: which cut
    # float32 decides differently, and by how much its value is off
    thr32 = tuple(np.asarray(t, np.float64)[bad].astype(np.float32) for t in (lo, hi, phic, tipc, rmin))
    q32 = _float32_quantities(*(p[bad] for p in P))
    ok32_cut = _passes32(q32, thr32)
    ok32 = ok32_cut[0] & ok32_cut[1] & ok32_cut[2] & ok32_cut[3]
    # numerical gradient of each cut quantity: one float32 ulp per coordinate
    Pb = [p[bad] for p in P]
    thr_b = tuple(t[bad] for t in thr)
    sens = [np.zeros(bad.size) for _ in CUTS]
    for k in range(3):
        for j in range(3):
            u = np.spacing(np.abs(Pb[k][:, j]).astype(np.float32)).astype(np.float64)
            plus = [p.copy() for p in Pb]
            minus = [p.copy() for p in Pb]
            plus[k][:, j] += u
            minus[k][:, j] -= u
            qp, qm = _quantities(*plus), _quantities(*minus)
            for c in range(4):
                sens[c] += np.abs(qp[c] - qm[c]) / 2
    for r, i in enumerate(bad):
        # the cut whose float32 and float64 decisions differ (the first such)
        flipped = [c for c in range(4) if bool(ok32_cut[c][r]) != bool(ok_cut[c][i])]
        c = flipped[0] if flipped else 3
        name = CUTS[c]
        cutv = float(_cut_value(c, thr_b, q[0][bad])[r])
        ulp_cut = float(dist[c][i] / np.spacing(np.float32(cutv)))
        ulp_in = float(dist[c][i] / sens[c][r]) if sens[c][r] > 0 else float("inf")
        mq = _mca_quantities(*(tuple(p[i]) for p in P), rng)[c]
        half = (np.nanmax(mq) - np.nanmin(mq)) / 2 if np.isfinite(mq).any() else 0.0
        units = float(dist[c][i] / half) if half > 0 else float("inf")
        out["by_cut"][name] += 1
        v32 = float(q32[c][r]) if c != 2 else float(np.sqrt(np.float64(q32[c][r])))
        rows.append({"cut": name, "gpu_kept": bool(kept[i]), "float32_replica_agrees": bool(ok32[r]) == bool(kept[i]),
                     "hits": [int(h1[i]), int(h2[i]), int(h3[i])],
                     "quantity": float([q[0], q[1], q[2], q[3]][c][i]), "quantity_float32": v32, "cut_value": cutv,
                     "ulp_of_cut": ulp_cut, "ulp_of_inputs": ulp_in, "rounding_units": units})
    out["max_ulp_of_cut"] = max(r["ulp_of_cut"] for r in rows)
    out["max_ulp_of_inputs"] = max(r["ulp_of_inputs"] for r in rows)
    out["max_rounding_units"] = max(r["rounding_units"] for r in rows)
    out["within_1_ulp_of_inputs"] = sum(r["ulp_of_inputs"] <= 1.0 for r in rows)
    out["within_float32_rounding"] = sum(r["rounding_units"] <= 1.0 for r in rows)
    out["float32_replica_agrees"] = sum(r["float32_replica_agrees"] for r in rows)
    out["details"] = sorted(rows, key=lambda r: -r["rounding_units"])[:max_report]
    return out
