"""Seeded synthetic events: helix tracks from the beam line plus uniform noise.

Each event is what the reference reads from a PEvent (src/datastructures/
serialize/Event.proto:1-37): a list of hits with position (x, y, z), layer,
detector id, simulated track id (0 = noise) and simulated track pT.

Tracks: charge +-1, pT log-uniform in [1, 50] GeV/c, eta uniform in
[-2.5, 2.5], phi0 uniform in [0, 2pi), vertex (0, 0, z0) with z0 ~ N(0, 5 cm),
in a solenoid of B = 3.8 T.  The transverse radius of curvature is
R = pT / (0.3 B) m.  At a barrel cylinder of radius r the helix has turned by
asin(r / 2R), so phi_hit = phi0 - q asin(r / 2R) and the transverse path is
s = 2 R asin(r / 2R), giving z = z0 + s sinh(eta).  Hits are smeared by
15 um in r*phi and 20 um in z and kept inside the layer's half-length.
Noise hits sit on a uniformly chosen layer, uniform in phi and z; they carry
track id 0 and a pT above any minimum pT cut so that HitCollection::addEvent
keeps them (src/datastructures/HitCollection.cpp:48-50).

Every event has its own numpy Generator seeded by (seed, event index), so an
event's hits do not depend on which other events share its batch.
"""
from dataclasses import dataclass

import numpy as np

from .config import Workload

NOISE_PT = 1.0e3


@dataclass
class Event:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    layer: np.ndarray
    det_id: np.ndarray
    track_id: np.ndarray
    track_pt: np.ndarray
    n_tracks: int

    @property
    def n_hits(self) -> int:
        return int(self.x.size)


def _rng(seed: int, index: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, index])))


def make_event(wl: Workload, n_tracks: int, n_noise: int, seed: int, index: int) -> Event:
    rng = _rng(seed, index)
    radii = np.asarray(wl.radii, dtype=np.float64)
    half = np.asarray(wl.half_lengths, dtype=np.float64)
    n_layers = radii.size

    q = rng.choice(np.array([-1.0, 1.0]), size=n_tracks)
    pt = np.exp(rng.uniform(np.log(1.0), np.log(50.0), size=n_tracks))
    eta = rng.uniform(-2.5, 2.5, size=n_tracks)
    phi0 = rng.uniform(0.0, 2.0 * np.pi, size=n_tracks)
    z0 = rng.normal(0.0, 5.0, size=n_tracks)
    big_r = pt / (0.3 * 3.8) * 100.0  # cm

    ratio = radii[None, :] / (2.0 * big_r[:, None])
    turn = np.arcsin(ratio)
    phi = phi0[:, None] - q[:, None] * turn
    z = z0[:, None] + 2.0 * big_r[:, None] * turn * np.sinh(eta)[:, None]
    phi = phi + rng.normal(0.0, 1.0, size=phi.shape) * (0.0015 / radii[None, :])
    z = z + rng.normal(0.0, 0.0020, size=z.shape)
    keep = np.abs(z) < half[None, :]

    t_idx, l_idx = np.nonzero(keep)  # track-major, layer-minor order
    tx = radii[l_idx] * np.cos(phi[t_idx, l_idx])
    ty = radii[l_idx] * np.sin(phi[t_idx, l_idx])
    tz = z[t_idx, l_idx]

    nl = rng.integers(0, n_layers, size=n_noise)
    nphi = rng.uniform(0.0, 2.0 * np.pi, size=n_noise)
    nz = rng.uniform(-1.0, 1.0, size=n_noise) * half[nl]
    nx = radii[nl] * np.cos(nphi)
    ny = radii[nl] * np.sin(nphi)

    layer = np.concatenate([l_idx + 1, nl + 1]).astype(np.uint32)
    all_phi = np.concatenate([phi[t_idx, l_idx], nphi])
    all_z = np.concatenate([tz, nz])
    hl = half[layer - 1]
    module = (np.floor(np.mod(all_phi, 2.0 * np.pi) / (2.0 * np.pi) * 64.0) * 32.0
              + np.floor((all_z + hl) / (2.0 * hl) * 32.0)).astype(np.uint32)
    det_id = (300_000_000 + layer.astype(np.uint32) * 100_000 + module).astype(np.uint32)

    return Event(
        x=np.concatenate([tx, nx]).astype(np.float32),
        y=np.concatenate([ty, ny]).astype(np.float32),
        z=all_z.astype(np.float32),
        layer=layer,
        det_id=det_id,
        track_id=np.concatenate([t_idx + 1, np.zeros(n_noise, dtype=np.int64)]).astype(np.uint32),
        track_pt=np.concatenate([pt[t_idx], np.full(n_noise, NOISE_PT)]).astype(np.float32),
        n_tracks=n_tracks,
    )


def event_sizes(wl: Workload, n_events: int, seed: int):
    """(tracks, noise) per event: fixed sizes, or log-uniform tracks with
    Poisson noise of mean noise_per_track * tracks."""
    ex = wl.extra
    if "tracks" in ex:
        return [(int(ex["tracks"]), int(ex["noise"]))] * n_events
    rng = _rng(seed, 1 << 30)
    lo, hi = np.log(ex["tracks_min"]), np.log(ex["tracks_max"])
    tracks = np.rint(np.exp(rng.uniform(lo, hi, size=n_events))).astype(np.int64)
    noise = rng.poisson(ex["noise_per_track"] * tracks)
    return [(int(t), int(n)) for t, n in zip(tracks, noise)]


def make_events(wl: Workload, n_events: int = None, seed: int = 2024, first: int = 0, of: int = None):
    """Events first .. first + n_events - 1 of a list of `of` events (default:
    the first + n_events events); the sizes of a list depend on its length,
    its events on their index and the seed."""
    n_events = wl.extra.get("events", 1) if n_events is None else n_events
    total = first + n_events if of is None else of
    if total < first + n_events:
        raise ValueError("the events asked for lie past the end of the list")
    sizes = event_sizes(wl, total, seed)[first:first + n_events]
    return [make_event(wl, t, n, seed, first + i) for i, (t, n) in enumerate(sizes)]


def _make_chunk(args):
    wl, sizes, seed, first = args
    return [make_event(wl, t, n, seed, first + i) for i, (t, n) in enumerate(sizes)]


def make_events_parallel(wl: Workload, n_events: int, seed: int = 2024, workers: int = 8, chunk: int = 64):
    """make_events over a process pool, chunk by chunk (yields lists of events in
    order).  The events are identical to make_events' (each has its own seed);
    start it before the process touches the GPU."""
    from concurrent.futures import ProcessPoolExecutor
    sizes = event_sizes(wl, n_events, seed)
    jobs = [(wl, sizes[i:i + chunk], seed, i) for i in range(0, n_events, chunk)]
    with ProcessPoolExecutor(max_workers=max(1, workers)) as ex:
        yield from ex.map(_make_chunk, jobs)
