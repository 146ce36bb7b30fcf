"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle):
the oracle must keep producing them (CPU), and the HIP path through the C ABI
must produce them too (GPU), plain and batched."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import oracle
from helpers import hitset
from trax_amd.api import Caps, run_batch_fit

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
from make_golden import CASES, case_workload  # noqa: E402

with open(os.path.join(GOLDEN, "pipeline_digests.json")) as _f:
    DIGESTS = json.load(_f)


def _check(res, name):
    bad = []
    for k, d in DIGESTS[name].items():
        if k not in res:
            bad.append(f"{k}: missing")
            continue
        a = np.ascontiguousarray(res[k])
        if str(a.dtype) != d["dtype"] or a.size != d["n"] or hashlib.sha256(a.tobytes()).hexdigest() != d["sha256"]:
            bad.append(f"{k}: differs from the fixture ({a.dtype}, {a.size} vs {d['dtype']}, {d['n']})")
    return bad


def _inputs(name):
    wl, seed = case_workload(name)
    first = CASES[name][1].get("first")
    cfg, hs, _ = (hitset(wl, seed=seed) if first is None else
                 hitset(wl, n_events=1, seed=seed, first=first, of=CASES[name][1]["of"]))
    return wl, cfg, hs


def test_full_fixture_arrays():
    wl, cfg, hs = _inputs("one-event-100-tracks")
    res = oracle.run(hs.columns(), wl, cfg.min_radius_curvature)
    with np.load(os.path.join(GOLDEN, "one-event-100-tracks.npz")) as z:
        for k in z.files:
            assert np.array_equal(res[k], z[k]), k


@pytest.mark.parametrize("name", ["one-event-100-tracks", "four-uneven-events", "two-events-ca-density",
                                  "batch-256-event-166", "batch-256-event-183"])
def test_oracle_reproduces_fixtures(name):
    wl, cfg, hs = _inputs(name)
    res = oracle.run(hs.columns(), wl, cfg.min_radius_curvature, workers=4)
    bad = _check(res, name)
    assert not bad, "\n".join(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_hip_plain_path_matches_fixtures(dev, name):
    wl, cfg, hs = _inputs(name)
    print(f"golden {name}: {hs.n_events} events, {hs.n_hits} hits", flush=True)
    bad = _check(dev.run_plain(cfg, hs), name)
    assert not bad, "\n".join(bad)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_hip_batched_path_matches_fixtures(dev, name):
    wl, cfg, hs = _inputs(name)
    print(f"golden batched {name}: {hs.n_events} events, {hs.n_hits} hits", flush=True)
    b, res = run_batch_fit(dev, cfg, hs, Caps.for_hits(hs.n_events, hs.n_hits))
    b.close()
    bad = _check(res, name)
    assert not bad, "\n".join(bad)
