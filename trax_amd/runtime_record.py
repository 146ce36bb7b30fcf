"""Runtime records in the reference's CSV format, so its plotting scripts
(plotting/plotRuntimes.py:23-40 reads them with
``np.genfromtxt(..., delimiter=" ", names=True)``) can put MI355X runs beside
the published ones.

Mirrors src/triplet_builder/RuntimeRecord.{h,cpp}:

* ``RuntimeInfo`` is tRuntimeInfo: the kernel time of one stage split into
  Count, Scan and Store, plus the stage's wall time (RuntimeRecord.h:27-50,
  RuntimeRecord.cpp:4-14).  Times are kept in milliseconds; the reference keeps nanoseconds and
  converts at output (Utils::nsToMs, utils.h:18).
* ``RuntimeRecord`` is one run: the workload shape (events, layers,
  layerTriplets, threads, hits, tracks), transfer time and bytes, and one
  RuntimeInfo per algorithm class (RuntimeRecord.h:74-122).
* ``RuntimeRecordClass`` gathers runs of equal shape (RuntimeRecord.cpp:
  419-427, :523-545) and keeps a running mean and sum of squared deviations
  per field (calculateMeanVar, RuntimeRecord.cpp:492-521).  The ``...Var`` columns hold
  sqrt(M2 / max(n - 1, 1)), the sample standard deviation
  (RuntimeRecordClass::toVar, RuntimeRecord.cpp:468-490); the total's M2 is
  the sum of the stages' M2 and a Kernel column's Var is the sum of its Count,
  Scan and Store Vars (tRuntimeInfo::totalKernel applied to the Var values,
  RuntimeRecord.cpp:53-60 and :429-456).
* ``RuntimeRecords.csv_dump`` writes the header and one line per class
  (RuntimeRecord.cpp:596-640), fields separated by one space (SEP,
  utils.h:13), numbers as C++'s default ostream prints them (``%g``, six
  significant digits).

Kernel time on this side comes from the HIP events a batch records between
its phases (``Batch.stage_ms``).  The reference adds each kernel's profiled
time to Count, Scan or Store when its name holds "Count", "prefixSum" or
"Store" (fillInfo, RuntimeRecord.cpp:123-134); ``PHASE_CLASS`` assigns each
phase the same way.  Phases that fuse a count, a scan and a store into one
pass (the grid build, the predictor) go to Store, the class of the kernel
that writes the stage's output; the CA's forward phase goes to Count (the
reference's forward kernel, named "iteration", matches none of the three
names and so drops out of its Kernel columns; here every phase is counted).
A stage's wall time is its span on the batch's stream (the sum of its
phases), because a replayed batch does not return to the host between
stages; the reference times the stage on the host clock
(tRuntimeInfo::startWalltime/stopWalltime, RuntimeRecord.cpp:62-74).
"""
import math
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

# algorithm classes in output order (RuntimeRecord.cpp:429-456) with the
# column prefix each has in the header (RuntimeRecord.cpp:603-638)
STAGES: Tuple[Tuple[str, str], ...] = (
    ("buildGrid", "buildGrid"),
    ("pairGen", "pairGen"),
    ("tripletPredict", "tripletPredict"),
    ("tripletFilter", "tripletFilter"),
    ("tripletConnectivity", "tripletConnectivityTight"),
    ("trackletCircleFitter", "trackletCircleFitter"),
    ("cellularAutomaton", "cellularAutomaton"),
)
STAGE_NAMES = tuple(s for s, _ in STAGES)

# The reference's data files under plotting/data/runtime/ come from a build
# that timed only the first four classes: 7 + 8 + 5 * 10 = 65 columns, the
# total summed over those four.  The current RuntimeRecords::csvDump writes
# all seven: 7 + 8 + 8 * 10 = 95 columns.
LAYOUTS = {"full": STAGE_NAMES, "four_stage": STAGE_NAMES[:4]}

# batch phase -> (algorithm class, Count / Scan / Store)
PHASE_CLASS: Dict[str, Tuple[str, str]] = {
    "grid_count": ("buildGrid", "store"),  # count, scan and stable scatter in one kernel
    "grid_store": ("buildGrid", "store"),
    "pair_plan": ("pairGen", "count"),  # slots per layer-triplet segment, scanned
    "pair_count": ("pairGen", "count"),
    "pair_scan": ("pairGen", "scan"),
    "pair_store": ("pairGen", "store"),
    "predict": ("tripletPredict", "store"),  # count, scan and store in one pass
    "filter_count": ("tripletFilter", "count"),
    "filter_scan": ("tripletFilter", "scan"),
    "filter_store": ("tripletFilter", "store"),
    "link_prepare": ("tripletConnectivity", "count"),
    "link_count": ("tripletConnectivity", "count"),
    "link_store": ("tripletConnectivity", "store"),
    "circle_fit": ("trackletCircleFitter", "store"),
    "ca_forward": ("cellularAutomaton", "count"),
    "ca_backward": ("cellularAutomaton", "store"),
}
PHASES = tuple(PHASE_CLASS)

_FIELDS = ("count", "scan", "store", "walltime")
_INFO_COLUMNS = ("events", "layers", "layerTriplets", "threads", "hits", "tracks")


def fmt(v) -> str:
    """One field as C++'s default ostream writes it: names and integers as
    they are, floating point with six significant digits (printf's %g)."""
    if isinstance(v, (int, str)):
        return str(v)
    return "%g" % v


def csv(values: Iterable) -> str:
    """Utils::csv (utils.h:26-62): the values joined by one space."""
    return " ".join(fmt(v) for v in values)


@dataclass
class RuntimeInfo:
    """tRuntimeInfo (RuntimeRecord.h:27-50): kernel time by class and wall time, ms."""
    count: float = 0.0
    scan: float = 0.0
    store: float = 0.0
    walltime: float = 0.0

    def total_kernel(self) -> float:
        return self.count + self.scan + self.store

    def __add__(self, o: "RuntimeInfo") -> "RuntimeInfo":
        return RuntimeInfo(self.count + o.count, self.scan + o.scan, self.store + o.store,
                           self.walltime + o.walltime)

    def csv_dump(self, var: Optional["RuntimeInfo"] = None) -> str:
        """Count Scan Store Walltime Kernel, each followed by its Var when
        `var` is given (RuntimeRecord.cpp:44-60)."""
        if var is None:
            return csv([self.count, self.scan, self.store, self.walltime, self.total_kernel()])
        return csv([self.count, var.count, self.scan, var.scan, self.store, var.store, self.walltime, var.walltime,
                    self.total_kernel(), var.total_kernel()])


@dataclass
class IOInfo:
    """tIOInfo: transfer time (ms) and bytes (RuntimeRecord.cpp:76-121)."""
    time: float = 0.0
    bytes: float = 0.0

    def csv_dump(self, var: Optional["IOInfo"] = None) -> str:
        if var is None:
            return csv([self.time, float(self.bytes)])
        return csv([self.time, var.time, float(self.bytes), float(var.bytes)])


def _sum(infos: Iterable[RuntimeInfo]) -> RuntimeInfo:
    total = RuntimeInfo()
    for i in infos:
        total = total + i
    return total


@dataclass
class RuntimeRecord:
    """One run (RuntimeRecord, RuntimeRecord.h:74-122).  `read` is
    device-to-host transfer, `write` host-to-device, as the reference's
    context counts them; both stay zero for a run on inputs already resident
    on the device."""
    events: int
    layers: int
    layer_triplets: int
    threads: int
    hits: int
    tracks: int
    read: IOInfo = field(default_factory=IOInfo)
    write: IOInfo = field(default_factory=IOInfo)
    stages: Dict[str, RuntimeInfo] = field(default_factory=lambda: {s: RuntimeInfo() for s in STAGE_NAMES})

    def key(self) -> Tuple[int, ...]:
        """The fields two records must share to be averaged together
        (RuntimeRecord::operator==, RuntimeRecord.cpp:372-380)."""
        return (self.events, self.layers, self.layer_triplets, self.threads, self.hits, self.tracks)

    def fill_phases(self, phase_ms: Dict[str, float]) -> "RuntimeRecord":
        """fillRuntimes (RuntimeRecord.cpp:136-192) from per-phase times in
        ms: each phase adds to its class's Count, Scan or Store and to the
        stage's wall time.  Every phase of ``PHASES`` must be present."""
        missing = [p for p in PHASES if p not in phase_ms]
        if missing:
            raise KeyError(f"phase times missing: {', '.join(missing)}")
        for s in STAGE_NAMES:
            self.stages[s] = RuntimeInfo()
        for p, (stage, cls) in PHASE_CLASS.items():
            info = self.stages[stage]
            ms = float(phase_ms[p])
            setattr(info, cls, getattr(info, cls) + ms)
            info.walltime += ms
        return self

    def fill_runtimes(self, batch) -> "RuntimeRecord":
        """The phase times of the batch's last completed run."""
        return self.fill_phases({p: batch.stage_ms(p) for p in PHASES})

    def total(self, stages: Sequence[str] = STAGE_NAMES) -> RuntimeInfo:
        return _sum(self.stages[s] for s in stages)

    def csv_dump(self, layout: str = "full") -> str:
        """One line without header (RuntimeRecord::csvDump,
        RuntimeRecord.cpp:382-407): shape, transfers, then Count Scan Store
        Walltime Kernel for each class and for the total."""
        stages = LAYOUTS[layout]
        parts = [csv(self.key()), self.read.csv_dump(), self.write.csv_dump()]
        parts += [self.stages[s].csv_dump() for s in stages]
        parts.append(self.total(stages).csv_dump())
        return " ".join(parts)


class _Moments:
    """Running mean and sum of squared deviations of one record field, updated
    as calculateMeanVar does (RuntimeRecord.cpp:492-521)."""
    __slots__ = ("mean", "m2")

    def __init__(self):
        self.mean = 0.0
        self.m2 = 0.0

    def add(self, x: float, n: int):
        delta = x - self.mean
        self.mean += delta / max(n, 1)
        self.m2 += delta * (x - self.mean)


class RuntimeRecordClass:
    """Runs of one shape with their means and deviations (RuntimeRecordClass,
    RuntimeRecord.h:124-194)."""

    def __init__(self, r: RuntimeRecord):
        self.events, self.layers, self.layer_triplets, self.threads, self.hits, self.tracks = r.key()
        self.records: List[RuntimeRecord] = []
        self._stage = {s: {f: _Moments() for f in _FIELDS} for s in STAGE_NAMES}
        self._io = {d: {f: _Moments() for f in ("time", "bytes")} for d in ("read", "write")}

    def key(self) -> Tuple[int, ...]:
        return (self.events, self.layers, self.layer_triplets, self.threads, self.hits, self.tracks)

    def add_record(self, r: RuntimeRecord):
        """A record of another shape is ignored, as in addRecord
        (RuntimeRecord.cpp:523-545)."""
        if r.key() != self.key():
            return
        self.records.append(r)
        n = len(self.records)
        for s in STAGE_NAMES:
            for f in _FIELDS:
                self._stage[s][f].add(getattr(r.stages[s], f), n)
        for d in ("read", "write"):
            io = getattr(r, d)
            self._io[d]["time"].add(io.time, n)
            self._io[d]["bytes"].add(float(io.bytes), n)

    def _div(self) -> int:
        return max(len(self.records) - 1, 1)  # Utils::clamp(records.size() - 1)

    def mean(self, stage: str) -> RuntimeInfo:
        return RuntimeInfo(*(self._stage[stage][f].mean for f in _FIELDS))

    def m2(self, stage: str) -> RuntimeInfo:
        return RuntimeInfo(*(self._stage[stage][f].m2 for f in _FIELDS))

    def to_var(self, m2: RuntimeInfo) -> RuntimeInfo:
        d = self._div()
        return RuntimeInfo(*(math.sqrt(getattr(m2, f) / d) for f in _FIELDS))

    def io(self, d: str) -> Tuple[IOInfo, IOInfo]:
        k = self._div()
        m = self._io[d]
        return (IOInfo(m["time"].mean, m["bytes"].mean),
                IOInfo(math.sqrt(m["time"].m2 / k), math.sqrt(m["bytes"].m2 / k)))

    def csv_dump(self, layout: str = "full") -> str:
        """One line of the RuntimeRecords table (RuntimeRecordClass::csvDump,
        RuntimeRecord.cpp:429-456)."""
        stages = LAYOUTS[layout]
        parts = [csv(list(self.key()) + [len(self.records)])]
        for d in ("read", "write"):
            mean, var = self.io(d)
            parts.append(mean.csv_dump(var))
        for s in stages:
            parts.append(self.mean(s).csv_dump(self.to_var(self.m2(s))))
        total_mean = _sum(self.mean(s) for s in stages)
        total_m2 = _sum(self.m2(s) for s in stages)
        parts.append(total_mean.csv_dump(self.to_var(total_m2)))
        return " ".join(parts)


def header(layout: str = "full") -> str:
    """The table's header line (RuntimeRecords::csvDump,
    RuntimeRecord.cpp:596-640).  The connectivity group's last column is
    named "tripletConnectivityTightWalltimeKernelVar" there
    (RuntimeRecord.cpp:620), and so here."""
    cols = list(_INFO_COLUMNS) + ["n"]
    for d in ("read", "write"):
        cols += [f"{d}Time", f"{d}TimeVar", f"{d}Bytes", f"{d}BytesVar"]
    prefixes = dict(STAGES)
    for s in LAYOUTS[layout]:
        p = prefixes[s]
        group = [f"{p}{c}{v}" for c in ("Count", "Scan", "Store", "Walltime", "Kernel") for v in ("", "Var")]
        if s == "tripletConnectivity":
            group[-1] = f"{p}WalltimeKernelVar"
        cols += group
    cols += [f"total{c}{v}" for c in ("Count", "Scan", "Store", "Walltime", "Kernel") for v in ("", "Var")]
    return csv(cols)


class RuntimeRecords:
    """Records grouped by shape in order of first appearance
    (RuntimeRecords, RuntimeRecord.h:196-215,
    RuntimeRecord.cpp:556-640)."""

    def __init__(self):
        self.classes: List[RuntimeRecordClass] = []

    def add_record(self, r: RuntimeRecord):
        for c in self.classes:
            if c.key() == r.key():
                c.add_record(r)
                return
        c = RuntimeRecordClass(r)
        c.add_record(r)
        self.classes.append(c)

    def csv_dump(self, layout: str = "full") -> str:
        """Header and one line per class, each ending in a newline."""
        if layout not in LAYOUTS:
            raise ValueError(f"layout must be one of {sorted(LAYOUTS)}")
        return "".join(line + "\n" for line in [header(layout)] + [c.csv_dump(layout) for c in self.classes])
