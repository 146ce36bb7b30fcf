"""Physics and track-level evaluation (trax_amd/physics.py over
trax_amd/csrc/physics.cpp) against the oracle's literal restatement
(oracle/physics.cpp), and the oracle's pins.

The library's PhysicsRecords (PhysicsRecord.{h,cpp}) and printTracks counts
(TripletBuilder.cpp:840-1066) must equal the oracle's text for text, on the
recorded ttbar events and on synthetic events that exercise every class
(found, clone, misc, fake, missed).  The oracle is pinned by what the
mathematics fixes:

- a record's rates merged over events are their plain mean and their sample
  spread (Welford and Chan's merge are exact), the total row over all records;
- a histogram's counts are the classification of each tracklet (numpy, from
  the hit ids and the truth) binned at its float32 eta / pT (numpy float64
  asinh; the circumradius from side lengths and area), up to tracklets
  within rounding of a bin edge; the merged counts are the events' sums, misc
  the first event's (tBinnedData::operator+= leaves it out);
- a merged bin's rates are the truth-weighted efficiency and the
  tracklet-weighted fake and clone rates of the events that hold the bin, one
  more zero weight when the first event lacks it (a default bin of n = 1),
  with the weighted spread;
- clone removal on hand-built tracks whose outcome is worked out by hand.

Each planted mistake of the oracle's hooks 20-23 fails its pin."""
import math
import os
import sys

import numpy as np
import pytest

import oracle
from helpers import hitset, small_workload
from trax_amd.api import HitSet
from trax_amd.physics import ETA, LAYOUT_PUBLISHED, PT, PhysicsRecords, summary_line, track_eval

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
from make_physics_golden import OUT as PHYS_GOLDEN, golden_files, ttbar_inputs  # noqa: E402

PT_PER_CM = 1.602177e-19 * 3.8112 * 1e-2 / 5.344286e-19


# ---- inputs ---------------------------------------------------------------------

class Case:
    def __init__(self, wl, cfg, hs, events, truths, res):
        self.wl, self.cfg, self.hs, self.events, self.truths, self.res = wl, cfg, hs, events, truths, res
        self.T = len(wl.triplets)


def _ttbar():
    wl, cfg, hs, events, truths = ttbar_inputs()
    return Case(wl, cfg, hs, events, truths, oracle.run(hs.columns(), wl, cfg.min_radius_curvature))


def _synthetic(n_events=3, max_tracks=-1):
    # all hits kept (noise and tracks that leave the barrel early stay in the
    # collection), so misc tracklets and the noise's id 0 occur
    wl = small_workload(3, events=n_events, tracks_min=40, tracks_max=300)
    cfg, _, events = hitset(wl, seed=5)
    hs = HitSet(cfg)
    for ev in events:
        hs.add(ev, max_tracks=max_tracks)
    truths = [oracle.findable_tracks(ev, wl, max_tracks=max_tracks) for ev in events]
    return Case(wl, cfg, hs, events, truths, oracle.run(hs.columns(), wl, cfg.min_radius_curvature, workers=4))


@pytest.fixture(scope="module")
def ttbar():
    return _ttbar()


@pytest.fixture(scope="module")
def synthetic():
    return _synthetic()


@pytest.fixture(scope="module", params=["ttbar", "synthetic"])
def case(request, ttbar, synthetic):
    return {"ttbar": ttbar, "synthetic": synthetic}[request.param]


def _table(text):
    lines = text.strip().splitlines()
    assert lines[0].split() == ["layerTriplet", "n", "eff", "effVar", "fr", "frVar", "cr", "crVar"]
    return {int(l.split()[0]): [float(v) for v in l.split()[1:]] for l in lines[1:]}


def _hist(text):
    """(header, {printed bin key: summed values}, keys printed more than once):
    bin keys are written with six significant digits, so two wide-apart
    bins of large pT can print alike."""
    lines = text.strip().splitlines()
    out, dup = {}, set()
    for l in lines[1:]:
        k, *v = l.split()
        if k in out:
            dup.add(k)
            out[k] = [a + b for a, b in zip(out[k], map(float, v))]
        else:
            out[k] = [float(x) for x in v]
    return lines[0].split(), out, dup


def _key(edge):
    return format(float(edge), "g")


def _close(a, b):
    # six significant digits as printed
    return abs(a - b) <= 1e-6 + 2e-5 * max(abs(a), abs(b))


# ---- the library against the oracle ----------------------------------------------

def test_truth_is_addevents_findable_tracks(case):
    for e, ev in enumerate(case.events):
        t = case.hs.truth(e)
        want = case.truths[e]
        assert list(t["track_id"]) == list(want), e
        idx = np.concatenate([np.asarray(v, np.int64) for v in want.values()]) if want else np.zeros(0, np.int64)
        assert np.array_equal(np.diff(t["hit_offsets"]), [len(v) for v in want.values()])
        for k in ("x", "y", "z"):
            assert np.array_equal(t[k].view(np.uint32), np.asarray(getattr(ev, k))[idx].view(np.uint32)), (e, k)
        assert np.array_equal(t["layer"], np.asarray(ev.layer)[idx])
    assert case.hs.n_findable == sum(len(t) for t in case.truths)


def test_truth_with_a_track_limit():
    c = _synthetic(n_events=2, max_tracks=25)
    for e in range(2):
        assert len(c.truths[e]) == 25
        assert list(c.hs.truth(e)["track_id"]) == list(c.truths[e])


def test_library_physics_equals_oracle(case):
    pr = PhysicsRecords().add(case.cfg, case.hs, case.res)
    o = oracle.physics(case.events, case.truths, case.res, case.T)
    assert pr.csv() == o["table"]
    assert len(pr) == case.T
    for (name, lt, layout), text in ((k, v) for k, v in o.items() if isinstance(k, tuple)):
        assert pr.histogram(ETA if name == "eta" else PT, lt, layout) == text, (name, lt, layout)


def test_library_physics_merge_equals_adding_in_order(synthetic):
    """Records of two batches merged (PhysicsRecords::merge) equal the
    records both batches were added to in turn: addRecord merges in the same
    order either way."""
    c = synthetic
    parts = []
    for evs in (c.events[:2], c.events[2:]):
        hs = HitSet(c.cfg)
        for ev in evs:
            hs.add(ev)
        parts.append((hs, oracle.run(hs.columns(), c.wl, c.cfg.min_radius_curvature)))
    one = PhysicsRecords()
    for hs, res in parts:
        one.add(c.cfg, hs, res)
    merged = PhysicsRecords().add(c.cfg, *parts[0])
    merged.merge(PhysicsRecords().add(c.cfg, *parts[1]))
    assert merged.csv() == one.csv()
    for lt in list(range(c.T)) + [c.T + 1]:
        assert merged.histogram(ETA, lt) == one.histogram(ETA, lt)
        assert merged.histogram(PT, lt) == one.histogram(PT, lt)
    assert merged.csv() != PhysicsRecords().add(c.cfg, *parts[0]).csv()


def test_library_writes_the_reference_files(ttbar, tmp_path):
    """The ttbar run's files in the reference's layout (physics.<name>.csv and
    <name>/eta_<p>.csv, pt_<p>.csv) equal the oracle's fixtures, and their
    headers are the published ones (plotting/data/physics/physics.ttBar.csv:1,
    plotting/data/physics/ttbar/eta_0.csv:1 for the five-column layout)."""
    c = ttbar
    pr = PhysicsRecords().add(c.cfg, c.hs, c.res)
    path = pr.write(str(tmp_path), "ttbar")
    for rel in ["physics.ttbar.csv"] + [os.path.join("ttbar", f"{n}_{lt}.csv") for n in ("eta", "pt")
                                        for lt in (0, 1, 2, 3, 5)]:
        with open(os.path.join(tmp_path, rel)) as f, open(os.path.join(PHYS_GOLDEN, rel)) as g:
            assert f.read() == g.read(), rel
    with open(path) as f:
        assert f.readline().strip() == "layerTriplet n eff effVar fr frVar cr crVar"
    assert pr.histogram(ETA, 0, LAYOUT_PUBLISHED).splitlines()[0] == "bin valid fake clones missed"
    s = track_eval(c.res, c.hs.n_findable)
    with open(os.path.join(PHYS_GOLDEN, "tracks.ttbar.txt")) as g:
        assert summary_line(s, c.wl.d_eta_cut) + "\n" == g.read()


def test_fixtures_are_the_oracles(ttbar):
    c = ttbar
    files = golden_files(c.wl, c.events, c.truths, c.res, c.hs.n_findable)
    for rel, text in files.items():
        with open(os.path.join(PHYS_GOLDEN, rel)) as g:
            assert g.read() == text, rel


def test_library_track_eval_equals_oracle(case):
    got = track_eval(case.res, case.hs.n_findable, kept=True)
    want = oracle.tracks(case.res, case.hs.n_findable)
    assert got["total"] > 10
    for k in want:
        if k == "kept":
            assert np.array_equal(got[k], want[k])
        else:
            assert got[k] == want[k], k


# ---- pins of the oracle -------------------------------------------------------------

def merge_failures(o, T):
    """Per layer triplet, the table row against the plain mean and sample
    spread of its records' rates; the total row against all records'."""
    rows = _table(o["table"])
    bad = []
    groups = {lt: [f for f in o["fills"] if f[1] == lt] for lt in range(T)}
    groups[T + 1] = o["fills"]
    for lt, fs in groups.items():
        n, eff, eff_sd, fr, fr_sd, cr, cr_sd = rows[lt]
        v = np.array([f[2:] for f in fs], np.float64)
        want = [len(fs)] + [x for j in range(3) for x in (v[:, j].mean(), v[:, j].std(ddof=1) if len(fs) > 1 else 0)]
        for name, g, w in zip(("n", "eff", "effVar", "fr", "frVar", "cr", "crVar"),
                              [n, eff, eff_sd, fr, fr_sd, cr, cr_sd], want):
            if not _close(g, w):
                bad.append((lt, name, g, w))
    return bad


def test_oracle_record_merge_is_mean_and_sample_spread(case):
    o = oracle.physics(case.events, case.truths, case.res, case.T)
    assert len(o["fills"]) == len(case.events) * case.T
    assert not merge_failures(o, case.T)


def _eta32(dx, dy, dz):
    """eta of float32 coordinate differences in float64 (PhysicsRecord.h:149-160)."""
    dx, dy, dz = (np.asarray(v, np.float32).astype(np.float64) for v in (dx, dy, dz))
    return np.arcsinh(dz / np.sqrt(dx * dx + dy * dy))


def _pt_circumradius(p1, p2, p3):
    """pT of the circle through three points (x, y): R = abc / (4 area), in float64."""
    p1, p2, p3 = (np.asarray(p, np.float64) for p in (p1, p2, p3))
    a = np.hypot(*(p2 - p1).T)
    b = np.hypot(*(p3 - p2).T)
    c = np.hypot(*(p3 - p1).T)
    cross = np.abs((p2 - p1)[:, 0] * (p3 - p1)[:, 1] - (p2 - p1)[:, 1] * (p3 - p1)[:, 0])
    with np.errstate(divide="ignore", invalid="ignore"):
        return a * b * c / (2 * cross) * PT_PER_CM


def _bins(v, rel):
    """Printed key of each value's bin (its lower edge 0.1 * floor(v / 0.1),
    the double the reference forms), and the keys a value within `rel` of an
    edge could take instead (empty when unambiguous)."""
    v = np.asarray(v, np.float64)
    keys, alts = [], []
    with np.errstate(invalid="ignore"):
        k = 0.1 * np.floor(v / 0.1)
        lo = 0.1 * np.floor(v * (1 - rel * np.sign(v)) / 0.1)
        hi = 0.1 * np.floor(v * (1 + rel * np.sign(v)) / 0.1)
    for a, b, c in zip(k, lo, hi):
        keys.append(_key(a))
        alts.append({_key(b), _key(c)} if b != c else set())
    return keys, alts


def _classify(c, e, lt):
    """Class of each tracklet of record (e, lt): 0 found, 1 clone, 2 misc,
    3 fake; and the findable ids not found."""
    res = c.res
    off = res["tracklet_offsets"].astype(np.int64)
    lo, hi = off[e * c.T + lt], off[e * c.T + lt + 1]
    h = [res[f"tracklets.h{i}"][lo:hi].astype(np.int64) for i in (1, 2, 3)]
    ids = [res["hits.hit_id"][x] for x in h]
    valid = (ids[0] == ids[1]) & (ids[0] == ids[2])
    findable = np.isin(ids[0], np.fromiter(c.truths[e].keys(), np.uint32, len(c.truths[e])))
    cls = np.full(hi - lo, 3)
    cls[valid & ~findable] = 2
    vf = np.nonzero(valid & findable)[0]
    _, first = np.unique(ids[0][vf], return_index=True)
    cls[vf] = 1
    cls[vf[first]] = 0
    missed = sorted(set(c.truths[e]) - set(ids[0][vf].tolist()))
    return h, cls, missed


def count_failures(c, o):
    """Merged histogram counts of every layer triplet against the classified,
    binned tracklets and missed tracks of each event."""
    res = c.res
    x, y, z = res["hits.x"], res["hits.y"], res["hits.z"]
    col_of = {0: 0, 3: 1, 1: 2, 2: 4}  # found, fake, clone, misc -> valid, fake, clones, misc
    bad = []
    for lt in range(c.T):
        for name in ("eta", "pt"):
            rel = 1e-6 if name == "eta" else 1e-4
            want, amb, misc_first = {}, {}, {}
            for e in range(len(c.events)):
                h, cls, missed = _classify(c, e, lt)
                if name == "eta":
                    v = _eta32(x[h[2]] - x[h[0]], y[h[2]] - y[h[0]], z[h[2]] - z[h[0]]).astype(np.float32)
                else:
                    v = _pt_circumradius(*[np.stack([x[k], y[k]], 1) for k in h]).astype(np.float32)
                ev = c.events[e]
                mv = []
                for t in missed:
                    idx = c.truths[e][t]
                    i0, im, i1 = idx[0], idx[len(idx) // 2], idx[-1]
                    if name == "eta":
                        mv.append(_eta32(ev.x[i1] - ev.x[i0], ev.y[i1] - ev.y[i0], ev.z[i1] - ev.z[i0]))
                    else:
                        mv.append(_pt_circumradius(*[np.array([[ev.x[i], ev.y[i]]]) for i in (i0, im, i1)])[0])
                keys, alts = _bins(v, rel)
                items = [(k, col_of[int(cl)], a) for k, a, cl in zip(keys, alts, cls)]
                mkeys, malts = _bins(np.asarray(mv, np.float64), rel)
                items += [(k, 3, a) for k, a in zip(mkeys, malts)]
                for k, col, a in items:
                    for alt in a:
                        amb[alt] = amb.get(alt, 0) + 1
                    if col == 4:
                        if e == 0:
                            misc_first[k] = misc_first.get(k, 0) + 1
                        continue
                    want.setdefault(k, [0, 0, 0, 0])[col] += 1
            head, got, _ = _hist(o[(name, lt, 0)])
            assert head[:6] == ["bin", "valid", "fake", "clones", "missed", "misc"]
            for k in set(want) | set(got):
                g = got.get(k, [0] * 11)
                w = want.get(k, [0, 0, 0, 0])
                if sum(abs(g[j] - w[j]) for j in range(4)) > 2 * amb.get(k, 0):
                    bad.append((name, lt, k, g[:5], w))
                if g[4] != misc_first.get(k, 0) and not amb:
                    bad.append((name, lt, k, "misc", g[4], misc_first.get(k, 0)))
    return bad


def test_oracle_histogram_counts_closed_form(case):
    o = oracle.physics(case.events, case.truths, case.res, case.T)
    bad = count_failures(case, o)
    assert not bad, bad[:6]


def rate_failures(c, o):
    """Merged bin rates against the weighted closed forms over the records'
    own bins (each event's record before merging); a record's bins against
    their counts' ratios.  Keys that print alike for two bins are left out."""
    bad = []
    for lt in range(c.T):
        for name in ("eta", "pt"):
            per, dups = [], set()
            for f, hist in zip(o["fills"], o["fill_hist"]):
                if f[1] != lt:
                    continue
                _, h, d = _hist(hist[name])
                per.append(h)
                dups |= d
            for p in per:  # one event: the bin's own ratios, no spread
                for k, (va, fa, cl, mi, mc, eff, effv, fr, frv, cr, crv) in p.items():
                    if k in dups:
                        continue
                    alln = va + fa + cl + mc
                    want = [va / max(va + mi, 1), 0, fa / max(alln, 1), 0, cl / max(alln, 1), 0]
                    if not all(_close(g, w) for g, w in zip([eff, effv, fr, frv, cr, crv], want)):
                        bad.append(("single", name, lt, k, [eff, effv, fr, frv, cr, crv], want))
            _, got, d = _hist(o[(name, lt, 0)])
            for k, g in got.items():
                if k in dups or k in d:
                    continue
                held = [p[k] for p in per if k in p]
                # weights: truth (valid + missed) for the efficiency, all
                # tracklets for fake and clone rates; a default bin (weight 1,
                # rate 0) when the first event lacks the bin
                wt = np.array([v[0] + v[3] for v in held], np.float64)
                wn = np.array([v[0] + v[1] + v[2] + v[4] for v in held], np.float64)
                eff_e = np.array([v[0] / max(v[0] + v[3], 1) for v in held])
                fr_e = np.array([v[1] / max(v[0] + v[1] + v[2] + v[4], 1) for v in held])
                cr_e = np.array([v[2] / max(v[0] + v[1] + v[2] + v[4], 1) for v in held])
                if k not in per[0]:
                    wn, fr_e, cr_e = np.append(wn, 1.0), np.append(fr_e, 0.0), np.append(cr_e, 0.0)
                    wt, eff_e = np.append(wt, 0.0), np.append(eff_e, 0.0)
                n = wn.sum()

                def mean_sd(x, w):
                    m = (w * x).sum() / w.sum() if w.sum() > 0 else float(x[0])
                    m2 = (w * (x - m) ** 2).sum()
                    return m, (math.sqrt(m2 / (n - 1)) if n > 1 else 0.0)

                want = [*mean_sd(eff_e, wt), *mean_sd(fr_e, wn), *mean_sd(cr_e, wn)]
                if not all(_close(a, b) for a, b in zip(g[5:11], want)):
                    bad.append(("merged", name, lt, k, g[5:11], want))
    return bad


def test_oracle_bin_rates_closed_form(ttbar, synthetic):
    for c in (ttbar, synthetic):
        o = oracle.physics(c.events, c.truths, c.res, c.T)
        bad = rate_failures(c, o)
        assert not bad, bad[:4]


def test_synthetic_case_exercises_every_class(synthetic):
    c = synthetic
    seen = set()
    for e in range(len(c.events)):
        for lt in range(c.T):
            _, cls, missed = _classify(c, e, lt)
            seen |= set(cls.tolist())
            if missed:
                seen.add("missed")
    assert seen >= {0, 1, 2, 3, "missed"}


# ---- clone removal on hand-built tracks --------------------------------------------

def _chains(tracks, ids):
    """A result holding the given tracks (hit lists, innermost first) as CA
    chains: each track of k hits is k - 2 tracklets, stored outermost first
    as the CA writes them; `ids` the hits' simulated track ids."""
    h1, h2, h3, trk, off = [], [], [], [], [0]
    for t in tracks:
        chain = []
        for j in range(len(t) - 2):
            h1.append(t[j]), h2.append(t[j + 1]), h3.append(t[j + 2])
            chain.append(len(h1) - 1)
        trk.extend(chain[::-1])
        off.append(len(trk))
    u = lambda v: np.asarray(v, np.uint32)
    return {"hits.hit_id": u(ids), "tracklets.h1": u(h1), "tracklets.h2": u(h2), "tracklets.h3": u(h3),
            "tracks": u(trk), "track_offsets": u(off)}


# hits 0-99 belong to sim track hit // 10 + 1 (hits 90-99: track 10)
IDS = [h // 10 + 1 for h in range(100)]
HAND = [
    # (tracks, expected kept flags, counts: total, filtered, valids, valid_unique, fakes)
    # disjoint tracks are all kept; a 4-hit chain (two tracklets) is no track
    ([[0, 1, 2, 3, 4], [10, 11, 12, 13, 14], [20, 21, 22, 23]], [3, 3, 0], (2, 2, 2, 2, 0)),
    # one shared hit of five (0.2 >= 0.19): similar; equal length keeps both
    ([[0, 1, 2, 3, 4], [4, 15, 16, 17, 18]], [3, 1], (2, 2, 1, 1, 1)),
    # one shared hit of six (0.167 < 0.19): not similar, both kept
    ([[0, 1, 2, 3, 4, 5], [5, 15, 16, 17, 18, 19]], [3, 1], (2, 2, 1, 1, 1)),
    # a longer similar track replaces the kept one; a shorter one is dropped
    ([[0, 1, 2, 3, 4], [0, 1, 2, 3, 4, 5], [1, 2, 3, 6, 7]], [0, 3, 0], (3, 1, 1, 1, 0)),
    # the first similar kept track in list order decides: the new track is
    # shorter than the first (dropped), though longer than the second
    ([[30, 31, 32, 33, 34, 35, 36], [40, 41, 42, 43, 44], [30, 31, 40, 41, 45, 46]], [3, 3, 0], (3, 2, 2, 2, 0)),
    # in the other order the shorter kept track is replaced
    ([[40, 41, 42, 43, 44], [30, 31, 32, 33, 34, 35, 36], [30, 31, 40, 41, 45, 46]], [0, 3, 1], (3, 2, 1, 1, 1)),
    # two real tracks of one sim track: both kept, one unique
    ([[50, 51, 52, 53, 54], [55, 56, 57, 58, 59]], [3, 3], (2, 2, 2, 1, 0)),
]


@pytest.mark.parametrize("i", range(len(HAND)))
def test_clone_removal_hand_built(i):
    tracks, flags, counts = HAND[i]
    res = _chains(tracks, IDS)
    for got in (track_eval(res, 10, kept=True), oracle.tracks(res, 10)):
        assert list(got["kept"]) == flags
        assert (got["total"], got["filtered"], got["valids"], got["valid_unique"], got["fakes"]) == counts
        f = float(counts[1])
        assert got["fake_rate"] == np.float32(counts[4]) / np.float32(f)
        assert got["clone_rate"] == np.float32(counts[2] - counts[3]) / np.float32(f)
        assert got["efficiency"] == np.float32(counts[3]) / np.float32(10)


def test_track_summary_line():
    s = {"total": 155, "filtered": 135, "valids": 111, "valid_unique": 87, "fakes": 24, "simulated": 213,
         "fake_rate": float(np.float32(24) / np.float32(135)), "clone_rate": float(np.float32(24) / np.float32(135)),
         "efficiency": float(np.float32(87) / np.float32(213))}
    assert summary_line(s, np.float32(0.0256)) == "0.0256 155 135 111 87 24 0.177778 0.177778 0.408451 213"


# ---- mutation hooks -------------------------------------------------------------------

def test_physics_pins_catch_planted_mistakes(ttbar, synthetic):
    caught = {}
    for plant in (20, 21, 22, 23):
        path = oracle.build(plant)
        if plant == 22:
            caught[plant] = any(list(oracle.tracks(_chains(t, IDS), 10, lib_path=path)["kept"]) != f
                                for t, f, _ in HAND)
            continue
        fails = False
        for c in (ttbar, synthetic):
            o = oracle.physics(c.events, c.truths, c.res, c.T, lib_path=path)
            if plant == 20:
                fails |= bool(count_failures(c, o))
            elif plant == 21:
                fails |= bool(merge_failures(o, c.T))
            else:
                fails |= bool(rate_failures(c, o))
        caught[plant] = fails
    assert all(caught.values()), caught


# ---- the HIP path's output --------------------------------------------------------

@pytest.mark.gpu
def test_hip_physics_equals_oracle(dev, case, tmp_path):
    """The HIP pipeline's tracklets and tracks (plain and batched) through the
    library's physics give the oracle's files and counts; on ttbar the files
    the reference's driver writes equal the committed fixtures."""
    from trax_amd.api import Caps, run_batch_fit
    want = oracle.physics(case.events, case.truths, case.res, case.T)
    want_tracks = oracle.tracks(case.res, case.hs.n_findable)
    print(f"physics {case.wl.name}: {case.hs.n_events} events, {case.hs.n_hits} hits", flush=True)
    got_plain = dev.run_plain(case.cfg, case.hs)
    b, got_batch = run_batch_fit(dev, case.cfg, case.hs, Caps.for_hits(case.hs.n_events, case.hs.n_hits))
    b.close()
    for got in (got_plain, got_batch):
        pr = PhysicsRecords().add(case.cfg, case.hs, got)
        assert pr.csv() == want["table"]
        for (name, lt, layout), text in ((k, v) for k, v in want.items() if isinstance(k, tuple)):
            assert pr.histogram(ETA if name == "eta" else PT, lt, layout) == text, (name, lt, layout)
        t = track_eval(got, case.hs.n_findable, kept=True)
        assert all(np.array_equal(t[k], want_tracks[k]) if k == "kept" else t[k] == want_tracks[k]
                   for k in want_tracks)
    if case.wl.name.startswith("ttbar"):
        PhysicsRecords().add(case.cfg, case.hs, got_plain).write(str(tmp_path), "ttbar")
        for rel in ["physics.ttbar.csv"] + [os.path.join("ttbar", f"{n}_{lt}.csv") for n in ("eta", "pt")
                                            for lt in (0, 1, 2, 3, 5)]:
            with open(os.path.join(tmp_path, rel)) as f, open(os.path.join(PHYS_GOLDEN, rel)) as g:
                assert f.read() == g.read(), rel
