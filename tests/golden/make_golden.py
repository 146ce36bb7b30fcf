"""Writes the golden fixtures of the pipeline: for each seeded case, the
oracle's output arrays (small case: the arrays themselves; larger cases: a
SHA-256 digest of every array's bytes).  Run from the repository root:

    python tests/golden/make_golden.py

The fixtures pin the path's results across changes: the CPU tests check the
oracle still produces them, the GPU tests check the HIP path does."""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from helpers import PIPELINE_KEYS, hitset, small_workload  # noqa: E402

# name -> (workload config, overrides, seed)
CASES = {
    "one-event-100-tracks": (1, {}, 11),
    "dense-2000-tracks-20000-noise": (2, {}, 11),
    "four-uneven-events": (3, dict(events=4, tracks_min=20, tracks_max=400), 11),
    "two-events-ca-density": (5, dict(events=2, tracks=500, noise=5000), 11),
    # one event at the full density of the CA workload (5000 tracks + 50,000
    # noise hits): about 7 minutes of oracle time on 8 threads
    "ca-scale-one-event": (5, dict(events=1, tracks=5000, noise=50000), 11),
    # single events of the 256-event batch of the throughput measurement
    # (log-uniform 50-5000 tracks, seed 2024), each run alone: its largest
    # (4927 tracks), a median one (392) and its smallest (51)
    "batch-256-event-157": (3, dict(first=157, of=256), 2024),
    "batch-256-event-166": (3, dict(first=166, of=256), 2024),
    "batch-256-event-183": (3, dict(first=183, of=256), 2024),
}
FULL = {"one-event-100-tracks"}
SLOW = {"ca-scale-one-event", "batch-256-event-157"}  # not re-run by the CPU tests


def case_workload(name):
    cfg_no, over, seed = CASES[name]
    over = dict(over)
    if "first" in over:
        return small_workload(cfg_no), seed
    if "tracks" in over:
        wl = small_workload(cfg_no, events=over["events"], tracks=over["tracks"], noise=over["noise"])
    elif "events" in over:
        wl = small_workload(cfg_no, events=over.pop("events"), **over)
    else:
        wl = small_workload(cfg_no)
    return wl, seed


def digest(a):
    a = np.ascontiguousarray(a)
    return {"dtype": str(a.dtype), "n": int(a.size), "sha256": hashlib.sha256(a.tobytes()).hexdigest()}


def main(names=None):
    import oracle
    path = os.path.join(HERE, "pipeline_digests.json")
    digests = {}
    if os.path.exists(path):
        with open(path) as f:
            digests = json.load(f)
    for name in names or CASES:
        wl, seed = case_workload(name)
        first = CASES[name][1].get("first")
        cfg, hs, _ = (hitset(wl, seed=seed) if first is None else
                     hitset(wl, n_events=1, seed=seed, first=first, of=CASES[name][1]["of"]))
        res = oracle.run(hs.columns(), wl, cfg.min_radius_curvature, workers=os.cpu_count() or 1)
        keys = [k for k in PIPELINE_KEYS if k in res]
        digests[name] = {k: digest(res[k]) for k in keys}
        if name in FULL:
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **{k: res[k] for k in keys})
        print(name, hs.n_hits, "hits,", res["tracklets.h1"].size, "tracklets")
    with open(path, "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or None)  # optional case names: regenerate only those
