"""Detector geometry, layer-triplet cuts and grid settings of the five workloads.

Layer radii bracket (min/max radius per layer, GeometrySupplement) are the
values the reference driver derives from configs/detectorRadius.dat and
configs/radiusDictionary.dat (src/triplet_builder/TripletBuilder.cpp:155-184):
the min/max dictionary radius over all detector ids of each layer, layers 1-13.

Layer-triplet cuts follow the reference's configs/layerTriplets.xml:
(1,2,3) uses the values of its first entry (:5-19), (2,3,4) the second, (3,4,5)
the third and every later consecutive triplet (4,5,6) ... (8,9,10) the values
of the fourth entry (:53-67, there written for (4,5,8)).  The fields are the
twelve columns of TripletConfigurations (src/datastructures/TripletConfiguration.h:33),
with pairSpreadZ / pairSpreadPhi / nCandidates as unsigned integers and z0
read as an integer and stored as float (TripletConfiguration.cpp:91).
No random draw is involved: the cut values are fixed constants.

Barrel half-lengths: the three pixel layers use |z| < 26.5 cm; the reference
states no half-length for outer layers, so this project uses 70 cm for the
four layers at r = 25.5 ... 49.8 cm and 110 cm for the three at 60.8 ... 78 cm
(all inside the grid's fixed z range of +-300 cm, src/datastructures/Grid.h:33-34).
"""
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

# GeometrySupplement rows (layer, min radius, max radius), layers 1..13
LAYER_RADIUS_BRACKETS: List[Tuple[int, float, float]] = [
    (1, 4.13, 4.63), (2, 7.04, 7.53), (3, 9.92, 10.4), (4, 23.78, 27.21),
    (5, 32.18, 35.61), (6, 40.26, 43.43), (7, 48.2, 51.38), (8, 58.43, 63.15),
    (9, 66.83, 71.56), (10, 76.06, 79.92), (11, 84.86, 88.73),
    (12, 94.56, 98.42), (13, 106.06, 109.93),
]

BARREL_RADII_10 = [4.4, 7.3, 10.2, 25.5, 33.9, 41.9, 49.8, 60.8, 69.2, 78.0]
BARREL_HALF_LENGTH_10 = [26.5, 26.5, 26.5, 70.0, 70.0, 70.0, 70.0, 110.0, 110.0, 110.0]

# (dThetaCut, sigmaZ, dPhiCut, sigmaPhi, tipCut, pairSpreadZ, z0, pairSpreadPhi, nCandidates)
_XML_ENTRIES = [
    (0.01, 0.037, 0.075, 0.032, 0.5, 2, 10, 2, 4),
    (0.2, 0.207, 0.2, 0.102, 0.5, 2, 10, 2, 16),
    (0.4, 0.207, 0.4, 0.102, 1.0, 3, 15, 2, 64),
    (0.5, 0.207, 0.4, 0.102, 3.5, 4, 30, 2, 128),
]
# the layers of the four entries in configs/layerTriplets.xml
_XML_LAYERS = [(1, 2, 3), (2, 3, 4), (3, 4, 5), (4, 5, 8)]

MIN_PT_GEV = 1.0
B_FIELD_T = 3.8
GRID_MIN_Z, GRID_MAX_Z = -300.0, 300.0


@dataclass
class TripletConfig:
    layer1: int
    layer2: int
    layer3: int
    dThetaCut: float
    sigmaZ: float
    dPhiCut: float
    sigmaPhi: float
    tipCut: float
    pairSpreadZ: int
    z0: float
    pairSpreadPhi: int
    nCandidates: int


def consecutive_triplets(n_triplets: int) -> List[TripletConfig]:
    out = []
    for k in range(n_triplets):
        e = _XML_ENTRIES[min(k, 3)]
        out.append(TripletConfig(k + 1, k + 2, k + 3, e[0], e[1], e[2], e[3], e[4],
                                 e[5], float(e[6]), e[7], e[8]))
    return out


@dataclass
class Workload:
    name: str
    radii: List[float]
    half_lengths: List[float]
    triplets: List[TripletConfig]
    n_z: int
    n_phi: int
    min_pt: float = MIN_PT_GEV
    n_threads: int = 256
    d_eta_cut: float = 0.0256  # TripletBuilder.cpp:793
    extra: dict = field(default_factory=dict)

    @property
    def n_layers(self) -> int:
        return len(self.radii)


def ttbar_workload() -> Workload:
    """The reference's ttbar reconstruction run on recorded simulation
    (configs/ttBar.reco.conf): the four layer triplets of
    configs/layerTriplets.xml with their cuts, CMS barrel layers 1-5 and 8,
    a 50 z x 8 phi grid, minPt 1 GeV, only findable tracks (onlyTracks).  The
    events come from a serialized PEventContainer, not from the generator;
    `radii` only sizes the layer count (8)."""
    trip = [TripletConfig(l1, l2, l3, e[0], e[1], e[2], e[3], e[4], e[5], float(e[6]), e[7], e[8])
            for (l1, l2, l3), e in zip(_XML_LAYERS, _XML_ENTRIES)]
    radii = [0.5 * (lo + hi) for _, lo, hi in LAYER_RADIUS_BRACKETS[:8]]
    return Workload("ttbar-recorded-events", radii, [0.0] * 8, trip, 50, 8,
                    extra=dict(only_tracks=True))


def workload(cfg: int) -> Workload:
    """The five workloads: 1 plain CPU-sized, 2 dense plain, 3 batched,
    4 serving stream, 5 cellular automaton at scale."""
    if cfg == 1:
        return Workload("one-event-100-tracks-3-layers", BARREL_RADII_10[:3],
                        BARREL_HALF_LENGTH_10[:3], consecutive_triplets(1), 8, 32,
                        extra=dict(tracks=100, noise=0, events=1))
    base = dict(radii=BARREL_RADII_10, half_lengths=BARREL_HALF_LENGTH_10,
                triplets=consecutive_triplets(8), n_z=32, n_phi=128)
    if cfg == 2:
        return Workload("dense-event-2000-tracks-20000-noise", **base,
                        extra=dict(tracks=2000, noise=20000, events=1))
    if cfg == 3:
        return Workload("batch-256-events-loguniform-50-5000-tracks", **base,
                        extra=dict(events=256, tracks_min=50, tracks_max=5000,
                                   noise_per_track=5.0))
    if cfg == 4:
        return Workload("stream-10000-events-poisson-arrivals", **base,
                        extra=dict(events=10000, tracks_min=50, tracks_max=5000,
                                   noise_per_track=5.0))
    if cfg == 5:
        return Workload("ca-64-events-5000-tracks-50000-noise", **base,
                        extra=dict(events=64, tracks=5000, noise=50000))
    raise ValueError(f"unknown workload {cfg}")


def geometry_arrays():
    """GeometrySupplement columns: layer (u32), min radius, max radius (f32)."""
    g = np.array(LAYER_RADIUS_BRACKETS, dtype=np.float64)
    return (g[:, 0].astype(np.uint32), g[:, 1].astype(np.float32), g[:, 2].astype(np.float32))


def triplet_arrays(triplets: List[TripletConfig]):
    """TripletConfigurations columns in the reference's order."""
    f32 = lambda k: np.array([getattr(t, k) for t in triplets], dtype=np.float32)
    u32 = lambda k: np.array([getattr(t, k) for t in triplets], dtype=np.uint32)
    return dict(layer1=u32("layer1"), layer2=u32("layer2"), layer3=u32("layer3"),
                dThetaCut=f32("dThetaCut"), sigmaZ=f32("sigmaZ"), dPhiCut=f32("dPhiCut"),
                sigmaPhi=f32("sigmaPhi"), tipCut=f32("tipCut"),
                pairSpreadZ=u32("pairSpreadZ"), z0=f32("z0"),
                pairSpreadPhi=u32("pairSpreadPhi"), nCandidates=u32("nCandidates"))


def min_radius_curvature(min_pt: float, bz: float = 3.8112) -> float:
    """TripletConfiguration.cpp:36-43 and .h:60-64: minPt and Bz are float
    parameters, the product is formed in double and returned as float."""
    q = 1.602177e-19
    gev_c = 5.344286e-19
    bz32 = float(np.float32(bz))
    return float(np.float32(float(np.float32(min_pt)) * gev_c / (bz32 * q) * 1e2))
