"""Runtime records in the reference's CSV format (trax_amd.runtime_record)
against the reference's own published tables (tests/golden/runtime/, copied
from plotting/data/runtime/): header, number format and the aggregation
rules (RuntimeRecord.cpp:429-640)."""
import io
import os

import numpy as np
import pytest

from trax_amd import runtime_record as rr

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "runtime")
TABLES = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".csv"))
N_INFO = 7  # events layers layerTriplets threads hits tracks n


def _lines(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return [ln.rstrip("\n") for ln in f if ln.strip()]


def _rows(name):
    """Data lines as (ints, floats) per the table's column types."""
    out = []
    for ln in _lines(name)[1:]:
        tok = ln.split(" ")
        out.append(([int(t) for t in tok[:N_INFO]], [float(t) for t in tok[N_INFO:]]))
    return out


def test_tables_present():
    assert len(TABLES) == 2


@pytest.mark.parametrize("name", TABLES)
def test_four_stage_header_is_the_published_one(name):
    assert rr.header("four_stage") == _lines(name)[0]
    assert len(rr.header("four_stage").split(" ")) == 65


def test_full_header_layout():
    """RuntimeRecords::csvDump (RuntimeRecord.cpp:596-640): the four-stage
    table's columns before the total, then the three later classes, then the
    total; 95 in all, with the connectivity group's odd last name."""
    full = rr.header("full").split(" ")
    four = rr.header("four_stage").split(" ")
    assert len(full) == 95 and len(set(full)) == 95
    assert full[:55] == four[:55] and full[-10:] == four[-10:]
    assert full[55:65] == ["tripletConnectivityTightCount", "tripletConnectivityTightCountVar",
                           "tripletConnectivityTightScan", "tripletConnectivityTightScanVar",
                           "tripletConnectivityTightStore", "tripletConnectivityTightStoreVar",
                           "tripletConnectivityTightWalltime", "tripletConnectivityTightWalltimeVar",
                           "tripletConnectivityTightKernel", "tripletConnectivityTightWalltimeKernelVar"]
    assert full[65] == "trackletCircleFitterCount" and full[75] == "cellularAutomatonCount"
    assert full[84] == "cellularAutomatonKernelVar" and full[85] == "totalCount"


@pytest.mark.parametrize("name", TABLES)
def test_number_format_reproduces_published_text(name):
    """Every published line comes back character for character when its
    values are written the way Utils::csv writes them (utils.h:26-62)."""
    for ln, (ints, floats) in zip(_lines(name)[1:], _rows(name)):
        assert rr.csv(ints + floats) == ln


def _col(name, h):
    return rr.header("four_stage").split(" ").index(h) - N_INFO


@pytest.mark.parametrize("name", TABLES)
def test_published_rows_obey_the_total_and_kernel_rules(name):
    """In the reference's rows a Kernel column is Count + Scan + Store and its
    Var the sum of their Vars; a total column is the sum of the four stages
    and its Var the root of the sum of their squared Vars.  The rules the
    emitter implements, checked on the reference's numbers (six digits)."""
    stages = ["buildGrid", "pairGen", "tripletPredict", "tripletFilter"]
    for _, v in _rows(name):
        c = lambda h: v[_col(name, h)]  # noqa: E731
        for s in stages + ["total"]:
            k = c(f"{s}Count") + c(f"{s}Scan") + c(f"{s}Store")
            assert c(f"{s}Kernel") == pytest.approx(k, rel=2e-5)
            kv = c(f"{s}CountVar") + c(f"{s}ScanVar") + c(f"{s}StoreVar")
            assert c(f"{s}KernelVar") == pytest.approx(kv, rel=2e-5)
        for f in ("Count", "Scan", "Store", "Walltime"):
            assert c(f"total{f}") == pytest.approx(sum(c(s + f) for s in stages), rel=2e-5)
            rss = np.sqrt(sum(c(s + f + "Var") ** 2 for s in stages))
            assert c(f"total{f}Var") == pytest.approx(rss, rel=2e-5)


def _spread(mean, sd, n, rng):
    """n values with exactly this mean and sample standard deviation."""
    if n == 1 or sd == 0.0:
        return np.full(n, mean)
    z = rng.standard_normal(n)
    z = (z - z.mean()) / z.std(ddof=1)
    return mean + sd * z


@pytest.mark.parametrize("name", TABLES)
def test_class_reproduces_published_rows(name):
    """Runs drawn with each published row's means and deviations, gathered by
    RuntimeRecordClass, give back the whole row: mean, sample deviation,
    totals and Kernel columns, in the published column order."""
    rng = np.random.default_rng(7)
    stages = rr.LAYOUTS["four_stage"]
    for ints, v in _rows(name):
        events, layers, lt, threads, hits, tracks, n = ints
        c = lambda h: v[_col(name, h)]  # noqa: E731
        prefixes = dict(rr.STAGES)
        draws = {}
        for d in ("read", "write"):
            for f, h in (("time", "Time"), ("bytes", "Bytes")):
                draws[(d, f)] = _spread(c(f"{d}{h}"), c(f"{d}{h}Var"), n, rng)
        for s in stages:
            for f in ("count", "scan", "store", "walltime"):
                h = prefixes[s] + f.capitalize()
                draws[(s, f)] = _spread(c(h), c(h + "Var"), n, rng)
        recs = rr.RuntimeRecords()
        for i in range(n):
            r = rr.RuntimeRecord(events, layers, lt, threads, hits, tracks)
            r.read = rr.IOInfo(draws[("read", "time")][i], draws[("read", "bytes")][i])
            r.write = rr.IOInfo(draws[("write", "time")][i], draws[("write", "bytes")][i])
            for s in stages:
                r.stages[s] = rr.RuntimeInfo(*(draws[(s, f)][i] for f in ("count", "scan", "store", "walltime")))
            recs.add_record(r)
        lines = recs.csv_dump("four_stage").splitlines()
        assert len(lines) == 2 and lines[0] == rr.header("four_stage")
        tok = lines[1].split(" ")
        assert [int(t) for t in tok[:N_INFO]] == ints
        got = np.array([float(t) for t in tok[N_INFO:]])
        np.testing.assert_allclose(got, np.array(v), rtol=3e-5, atol=1e-9)


def test_mean_and_sample_deviation():
    rng = np.random.default_rng(1)
    x = rng.uniform(0.1, 5.0, (9, 4))
    recs = rr.RuntimeRecords()
    for row in x:
        r = rr.RuntimeRecord(4, 3, 1, 256, 1000, 100)
        r.stages["pairGen"] = rr.RuntimeInfo(*row)
        recs.add_record(r)
    (cls,) = recs.classes
    mean, var = cls.mean("pairGen"), cls.to_var(cls.m2("pairGen"))
    np.testing.assert_allclose([mean.count, mean.scan, mean.store, mean.walltime], x.mean(0), rtol=1e-12)
    np.testing.assert_allclose([var.count, var.scan, var.store, var.walltime], x.std(0, ddof=1), rtol=1e-10)


def test_single_record_has_zero_deviation():
    """n - 1 is clamped to 1 (Utils::clamp, utils.h:65-67): one run gives
    zero Var columns, not a division by zero."""
    r = rr.RuntimeRecord(1, 3, 1, 256, 10, 2).fill_phases({p: 1.5 for p in rr.PHASES})
    recs = rr.RuntimeRecords()
    recs.add_record(r)
    tab = np.genfromtxt(io.StringIO(recs.csv_dump()), delimiter=" ", names=True)
    assert tab["n"] == 1
    assert tab["totalKernelVar"] == 0.0 and tab["buildGridWalltimeVar"] == 0.0


def test_records_group_by_shape_in_first_seen_order():
    recs = rr.RuntimeRecords()
    shapes = [(1, 3, 1, 256, 10, 2), (2, 3, 1, 256, 20, 4), (1, 3, 1, 256, 10, 2), (1, 3, 1, 128, 10, 2)]
    for k, s in enumerate(shapes):
        recs.add_record(rr.RuntimeRecord(*s).fill_phases({p: float(k + 1) for p in rr.PHASES}))
    assert [c.key() for c in recs.classes] == [shapes[0], shapes[1], shapes[3]]
    assert [len(c.records) for c in recs.classes] == [2, 1, 1]
    # a record of another shape is not taken into a class
    recs.classes[1].add_record(rr.RuntimeRecord(*shapes[0]))
    assert len(recs.classes[1].records) == 1


def test_phases_fill_their_classes():
    """Each phase adds to one class's Count, Scan or Store and to that
    class's wall time; every class gets a phase, and the Kernel total is the
    sum of all phase times."""
    ms = {p: float(2 ** k) for k, p in enumerate(rr.PHASES)}  # distinct bits: sums identify phases
    r = rr.RuntimeRecord(1, 3, 1, 256, 10, 2).fill_phases(ms)
    assert set(s for s, _ in rr.PHASE_CLASS.values()) == set(rr.STAGE_NAMES)
    assert r.total().total_kernel() == sum(ms.values()) == r.total().walltime
    assert r.stages["tripletPredict"].store == ms["predict"]
    assert r.stages["pairGen"].scan == ms["pair_scan"]
    assert r.stages["buildGrid"].total_kernel() == ms["grid_count"] + ms["grid_store"]
    assert r.stages["cellularAutomaton"].walltime == ms["ca_forward"] + ms["ca_backward"]
    with pytest.raises(KeyError, match="predict"):
        rr.RuntimeRecord(1, 3, 1, 256, 10, 2).fill_phases({p: 1.0 for p in rr.PHASES if p != "predict"})


@pytest.mark.parametrize("layout,ncol", [("full", 95), ("four_stage", 65)])
def test_dump_reads_back_as_the_plotting_scripts_read_it(layout, ncol):
    """plotRuntimes.py:23-40 reads the tables with genfromtxt(names=True)
    and uses these columns."""
    recs = rr.RuntimeRecords()
    rng = np.random.default_rng(3)
    for e in (1, 2, 4):
        for _ in range(3):
            ms = {p: float(rng.uniform(0.01, 1.0)) * e for p in rr.PHASES}
            recs.add_record(rr.RuntimeRecord(e, 3, 1, 256, 355 * e, 100 * e).fill_phases(ms))
    tab = np.genfromtxt(io.StringIO(recs.csv_dump(layout)), delimiter=" ", names=True)
    assert len(tab.dtype.names) == ncol and tab.shape == (3,)
    assert list(tab["events"]) == [1, 2, 4] and list(tab["n"]) == [3, 3, 3]
    for col in ("totalWalltime", "totalWalltimeVar", "totalKernel", "totalKernelVar", "readTime", "writeTime",
                "buildGridKernel", "pairGenKernel", "tripletPredictKernel", "tripletFilterKernel", "hits", "tracks"):
        assert np.all(np.isfinite(tab[col])), col
    if layout == "full":
        kern = sum(tab[f"{p}Kernel"] for p in ("buildGrid", "pairGen", "tripletPredict", "tripletFilter",
                                              "tripletConnectivityTight", "trackletCircleFitter",
                                              "cellularAutomaton"))
        np.testing.assert_allclose(tab["totalKernel"], kern, rtol=3e-5)


def test_single_record_line():
    """RuntimeRecord::csvDump (RuntimeRecord.cpp:382-407): no n column and no
    Var columns; 6 + 4 + 8 * 5 fields."""
    r = rr.RuntimeRecord(1, 3, 1, 256, 10, 2).fill_phases({p: 0.25 for p in rr.PHASES})
    tok = r.csv_dump().split(" ")
    assert len(tok) == 50 and tok[:6] == ["1", "3", "1", "256", "10", "2"]
    assert tok[-1] == "4" and tok[-2] == "4"  # total kernel and wall time: 16 phases of 0.25 ms
    assert len(r.csv_dump("four_stage").split(" ")) == 35


def test_unknown_layout_rejected():
    with pytest.raises(ValueError):
        rr.RuntimeRecords().csv_dump("five_stage")
