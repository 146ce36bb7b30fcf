"""Writes the physics fixtures of the recorded ttbar events
(tests/golden/ttbar/hits.ttBar.pb through the ttbar configuration, only
findable tracks): the oracle's pipeline, then the oracle's PhysicsRecords and
track-level evaluation of its output, as the reference's driver lays the
files out (TripletBuilder.cpp:823-834):

    tests/golden/ttbar/physics/physics.ttbar.csv       PhysicsRecords::csvDump
    tests/golden/ttbar/physics/ttbar/eta_<p>.csv       histograms, 12 columns
    tests/golden/ttbar/physics/ttbar/pt_<p>.csv
    tests/golden/ttbar/physics/tracks.ttbar.txt        printTracks' stderr line

Run from the repository root:  python tests/golden/make_physics_golden.py"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

OUT = os.path.join(HERE, "ttbar", "physics")
PB = os.path.join(HERE, "ttbar", "hits.ttBar.pb")


def ttbar_inputs():
    """(workload, Config, HitSet, events, findable tracks per event) of the
    ttbar fixture, the events decoded by the tests' own protobuf reader."""
    import oracle
    from test_ttbar import _decode
    from trax_amd.api import Config, HitSet
    from trax_amd.config import ttbar_workload
    wl = ttbar_workload()
    cfg = Config(wl)
    with open(PB, "rb") as f:
        events = _decode(f.read())
    hs = HitSet(cfg)
    for ev in events:
        hs.add(ev, only_tracks=True)
    truths = [oracle.findable_tracks(ev, wl, only_tracks=True) for ev in events]
    return wl, cfg, hs, events, truths


def golden_files(wl, events, truths, res, simulated):
    """{relative path: text} of the fixture files for one oracle result."""
    import oracle
    from trax_amd.physics import summary_line
    o = oracle.physics(events, truths, res, len(wl.triplets))
    files = {"physics.ttbar.csv": o["table"]}
    for (name, lt, layout), text in ((k, v) for k, v in o.items() if isinstance(k, tuple)):
        if layout == 0:
            files[os.path.join("ttbar", f"{name}_{lt}.csv")] = text
    files["tracks.ttbar.txt"] = summary_line(oracle.tracks(res, simulated), wl.d_eta_cut) + "\n"
    return files


def main():
    import oracle
    wl, cfg, hs, events, truths = ttbar_inputs()
    res = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    for rel, text in golden_files(wl, events, truths, res, sum(len(t) for t in truths)).items():
        path = os.path.join(OUT, rel)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            f.write(text)
        print(path)


if __name__ == "__main__":
    main()
