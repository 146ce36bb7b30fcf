#!/usr/bin/env python3
"""Benchmark of the triplet-seeding path on one MI355X.

Workload: 256 events with log-uniform track counts in [50, 5000] and Poisson
noise (mean 5 hits per track), 10 barrel layers, 8 consecutive layer triplets,
32 z x 128 phi grid.  A step is one pass of the whole path (grid build, pairs,
prediction, filter, connectivity, circle fit, cellular automaton) over the
256-event batch through the batched pipeline captured as one graph, with the
hits already resident in device memory (`value`); a second run with a graph
per phase times every phase with HIP events, and the rate from the first H2D
of the hits to the last D2H of the tracklets and tracks (the serving loop's two
slots, every event arriving at once) is reported beside it.
The plain per-event path (the reference's call sequence: one event per launch
sequence, blocking count read-backs, exact allocations, its own transfers) is
timed on a sample of the same events for comparison, the filter's float32
decisions on every candidate of the batch are adjudicated against float64, and
the oracle (the CPU restatement of the reference's kernels) runs on a sample
of the batch that matches its mix of sizes as the CPU baseline.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_EVENTS = 256
SEED = 2024

# HBM peak of MI355X (bytes/s)
HBM_PEAK = 8.0e12


def phase_bytes(n_hits, seg_hits, trip_hits, counts, cells, n_threads, fused):
    """Algorithmic (compulsory) HBM bytes of each phase: every input stream read
    once, every output written once; the per-unit figures of the measurement
    plan (grid 68 B/hit + 16 B/cell, scan 8 B/element, ...).  With the filter
    fused into the predictor, the predict phase reads the pairs and the hits
    of the three layers and writes the tracklets (8 B/pair + 16 B/hit + 12 B
    per tracklet); the filter phases only place the tracklet offsets."""
    P, Q, R, K = counts["pairs"], counts["cands"], counts["tracklets"], counts["conn"]
    slots = counts.get("slots", 0)
    words = (Q + 63) // 64
    if fused:
        return {
            # the grid build is one kernel (count, scan, scatter), run in the grid_count phase
            "grid_count": 68 * n_hits + 16 * cells,
            "pair_count": 12 * seg_hits + 4 * slots,
            "pair_scan": 8 * slots,
            "pair_store": 12 * seg_hits + 4 * slots + 8 * P,
            "predict": 8 * P + 16 * trip_hits + 12 * R,
            # the pass's tracklets copied from the waves' ranges into pair order, chunk counts scanned
            "filter_count": 24 * R + 16 * ((P + 63) // 64),
            "filter_store": 12 * R,
            "link_count": 12 * R + 4 * R + 8 * R,
            "link_store": 12 * R + 4 * R + 12 * K,
        }
    return {
        "grid_count": 68 * n_hits + 16 * cells,       # the one-kernel grid build: 12 B/hit count + 56 B/hit scatter
        "pair_count": 12 * seg_hits + 4 * slots,
        "pair_scan": 8 * slots,
        "pair_store": 12 * seg_hits + 4 * slots + 8 * P,
        "predict": 8 * P + 16 * trip_hits + 4 * P + 8 * Q,   # pairs, hit x/y/z/phi, offsets, candidates
        "filter_count": 8 * Q + 8 * Q + 12 * words,
        "filter_scan": 8 * words,
        "filter_store": 12 * words + 8 * R + 12 * R,
        "link_count": 12 * R + 4 * R + 8 * R,
        "link_store": 12 * R + 4 * R + 12 * K,
    }


def scan_hbm(dev, n=1 << 28, reps=10):
    """HBM rate of the prefix-sum primitive alone (trax_prefix_sum, the
    reference's PrefixSum::run) at a size where HBM, not launch latency,
    bounds it: n u32 values read and written once (8 B per element), timed by
    device events on the stream the scan runs on.  The scans inside the
    pipeline cover a few MB each and are bound by launch latency instead."""
    import ctypes as C
    import numpy as np
    from trax_amd.api import Timer, check, lib
    L = lib()
    buf, tot = C.c_void_p(), C.c_void_p()
    check(L.trax_buf_alloc(dev.h, 4 * n, C.byref(buf)))
    check(L.trax_buf_alloc(dev.h, 4, C.byref(tot)))
    try:
        check(L.trax_buf_fill_u32(dev.h, buf, 1, n, None))
        check(L.trax_prefix_sum(dev.h, buf, n, tot, None))  # also sizes the scan's scratch
        total = np.zeros(1, np.uint32)
        check(L.trax_d2h(dev.h, total.ctypes.data, tot, 4, None))
        check(L.trax_stream_sync(dev.h, None))
        if int(total[0]) != n:
            raise RuntimeError(f"scan total {int(total[0])} != {n}")
        timer = Timer(dev)
        timer.start()
        for _ in range(reps):
            check(L.trax_prefix_sum(dev.h, buf, n, tot, None))
        ms = timer.stop() / reps
    finally:
        L.trax_buf_free(dev.h, buf)
        L.trax_buf_free(dev.h, tot)
    gbs = 8 * n / (ms * 1e-3) / 1e9
    return {"elements": n, "ms": round(ms, 4), "gbs": round(gbs, 1), "frac": round(gbs * 1e9 / HBM_PEAK, 4)}


def fp64_check(dev, wl, cfg, hs, caps):
    """The filter's float32 decisions on every candidate of the batch against a
    float64 recomputation (trax_amd/adjudicate.py): the batch runs once more
    through the unfused path, which stores the candidates, and every
    candidate's four cuts (TripletThetaPhiFilter.h:83-150) are evaluated in
    float64 on the host.  Each disagreement is reported with its cut and its
    distance to the cut in float32 ulps of the cut value, in float32 ulps of
    the inputs, and in units of the float32 rounding of the cut's quantity
    (Monte Carlo arithmetic)."""
    from trax_amd.adjudicate import adjudicate
    from trax_amd.api import Batch
    t0 = time.perf_counter()
    print("float64 check: the batch through the unfused path", file=sys.stderr, flush=True)
    b = Batch(dev, cfg, caps)
    b.submit(hs)
    if b.wait() != 0:
        raise RuntimeError("capacity overflow in the float64 check")
    res = b.result(["cands", "pairs", "hits.x", "hits.y", "hits.z", "hits.layer", "filter.bits"])
    b.close()
    a = adjudicate(res, cfg, workers=min(16, os.cpu_count() or 1),
                   progress=lambda m: print(f"float64 check: {m}", file=sys.stderr, flush=True))
    a["events"] = hs.n_events
    a["details"] = a["details"][:4]
    a["host_s"] = round(time.perf_counter() - t0, 1)
    return a


def serving_stream(wl, n_events, seed, workers):
    """The 10,000-event stream of the serving measurement, generated on a
    process pool (before the process touches the GPU) into one HitSet."""
    from trax_amd.api import Config, HitSet
    from trax_amd.generator import make_events_parallel
    cfg = Config(wl)
    hs = HitSet(cfg)
    for chunk in make_events_parallel(wl, n_events, seed=seed, workers=workers):
        for ev in chunk:
            hs.add(ev)
    return cfg, hs


def serve(dev, cfg, hs, bounds, rates_hz, seed, n_slots, max_batch_events, slot_hits, fused, group):
    """Replay the stream with Poisson arrivals at each rate of rates_hz through
    one continuous-batching server (events in arrival order, or grouped by hit
    count); latency = arrival to results on the host.  One record per rate."""
    import numpy as np
    from trax_amd.api import Server
    sv = Server(dev, cfg, bounds, n_slots=n_slots, max_batch_events=max_batch_events, max_batch_hits=slot_hits,
                device_budget=64 << 30, fuse_filter=fused, group_by_size=group)
    caps = sv.caps()
    out = []
    for rate_hz in rates_hz:
        rng = np.random.default_rng(seed)
        arrival = np.cumsum(rng.exponential(1.0 / rate_hz, hs.n_events))
        arrival -= arrival[0]
        _, st = sv.replay(hs, arrival)
        out.append({"events": st["events"], "arrival": "poisson", "rate_hz": rate_hz, "seed": seed,
                    "admission": "grouped-by-hit-count" if group else "arrival-order",
                    "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
                    "max_ms": round(st["max_ms"], 3), "events_per_s": round(st["events_per_s"], 1),
                    "tracklets_per_s": round(st["tracklets_per_s"], 1), "batches": st["batches"],
                    "mean_batch_events": round(st["mean_batch_events"], 2),
                    "max_batch_events": st["max_batch_events"], "resubmits": st["resubmits"],
                    "failed": st["failed"], "slots": n_slots, "slot_hit_capacity": caps["max_hits"],
                    "device_bytes": st["device_bytes"],
                    "mean_ms": {k: round(st[k + "_ms"], 3) for k in ("queue", "pipeline", "copy", "split")}})
        print(f"serving {out[-1]['admission']} at {rate_hz:.0f}/s: p50 {out[-1]['p50_ms']} ms, "
              f"p99 {out[-1]['p99_ms']} ms, {out[-1]['events_per_s']} events/s", file=sys.stderr, flush=True)
    sv.close()
    return out


# the kernel each phase's time belongs to (the phase also holds small fills)
PHASE_KERNEL = {"grid_count": "k_grid_build", "grid_store": "k_grid_build", "pair_count": "k_pair_count",
                "pair_store": "k_pair_store", "predict": "k_predict_fused", "filter_count": "k_filter_count",
                "filter_store": "k_filter_store", "link_count": "k_link_count", "link_store": "k_link_store"}
PHASE_KERNEL_FUSED = dict(PHASE_KERNEL, predict="k_predict_filter_wave", filter_count="k_chunk_copy3",
                          filter_store="k_filter_offsets")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, calibrated)
    from the rocprofv3 counter summary committed under profiles/, or None."""
    path = os.path.join(ROOT, "profiles", "hbm-traffic-256-events.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            p = line.split()
            if p and not line.startswith("#") and p[0] == kernel and len(p) == 4:
                return (int(p[2]) + int(p[3])) // max(int(p[1]), 1)
    return None


def measured_valu_busy(kernel):
    """Fraction of the SIMDs' VALU issue the kernel used (SQ_ACTIVE_INST_VALU
    over the kernel's quad-cycles) from the rocprofv3 summary committed under
    profiles/, or None: the bound of a kernel that HBM does not bound."""
    path = os.path.join(ROOT, "profiles", "valu-issue-256-events.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            p = line.split()
            if p and not line.startswith("#") and p[0] == kernel and len(p) == 6:
                return float(p[5])
    return None


def end_to_end(dev, cfg, events, bounds, batch_hits, fused, reps=4):
    """Events/s from the first H2D to the last D2H: `reps` copies of the
    step's events, all arriving at once, through the serving loop with two
    slots that each hold a whole step (256 events); every event's hits go up
    from pinned host memory and its tracklets and CA tracks come back."""
    import numpy as np
    from trax_amd.api import HitSet, Server
    hs = HitSet(cfg, [ev for _ in range(reps) for ev in events])
    sv = Server(dev, cfg, bounds, n_slots=2, max_batch_events=len(events), max_batch_hits=batch_hits + 4096,
                device_budget=64 << 30, fuse_filter=fused, group_by_size=False)
    sv.replay(hs, np.zeros(hs.n_events))  # warm-up: graph capture of both slots
    _, st = sv.replay(hs, np.zeros(hs.n_events))
    sv.close()
    return {"events": hs.n_events, "events_per_s": round(st["events_per_s"], 1), "batches": st["batches"],
            "mean_batch_events": round(st["mean_batch_events"], 2), "slots": 2,
            "mean_ms": {k: round(st[k + "_ms"], 3) for k in ("pipeline", "copy", "split")}}


def cpu_baseline(wl, events, cfg_cls, hitset_cls, n_sample=32):
    """The oracle (the CPU restatement of the reference's kernels) on a sample
    of the step's events that matches its mix of sizes: every eighth event in
    order of hit count, run as one event group on the host's threads."""
    import oracle
    cores = min(16, os.cpu_count() or 1)
    order = sorted(range(len(events)), key=lambda e: events[e].n_hits)
    step = max(1, len(events) // n_sample)
    sample = [events[e] for e in order[step // 2::step]][:n_sample]
    cfg = cfg_cls(wl)
    hs = hitset_cls(cfg, sample)
    t0 = time.perf_counter()
    oracle.run(hs.columns(), wl, cfg.min_radius_curvature, workers=cores)
    t = time.perf_counter() - t0
    mean_all = sum(ev.n_hits for ev in events) / len(events)
    return {"value": len(sample) / t, "unit": "events/s", "cores": cores, "kind": "port",
            "hits_per_s": round(hs.n_hits / t, 1),
            "sample": f"{len(sample)} of the {len(events)} batch events (every {step}th by hit count: "
                      f"{hs.n_hits / len(sample):.0f} hits per event against {mean_all:.0f} in the batch) as one "
                      f"event group, oracle (C++ restatement of the reference kernels) on {cores} host threads, "
                      f"{t:.1f} s"}


def write_runtime_csv(path, wl, hs, events, step_phase_ms):
    """The timed steps as one row of the reference's runtime table (mean and
    deviation over the steps), inputs resident so no transfer columns."""
    from trax_amd import runtime_record as rr
    recs = rr.RuntimeRecords()
    shape = (len(events), wl.n_layers, len(wl.triplets), wl.n_threads, hs.n_hits, sum(e.n_tracks for e in events))
    for ms in step_phase_ms:
        recs.add_record(rr.RuntimeRecord(*shape).fill_phases(ms))
    with open(path, "w") as f:
        f.write(recs.csv_dump())


def runtime_sweep(dev, path, reps=19, max_events=2048):
    """The shape of the reference's published event sweep
    (plotting/data/runtime/runtime.bigOne.local.conf.events.100.test.gpu.csv:
    1, 2, 4, ... 2048 events of 100 tracks, one layer triplet, 19 runs each),
    here on the one-event-100-tracks geometry (3 layers, 8 z x 32 phi grid),
    batched with the filter fused: one row per event count in the reference's
    65-column layout (grid, pair, predict and filter stages)."""
    from trax_amd import make_events, workload
    from trax_amd import runtime_record as rr
    from trax_amd.api import Caps, Config, HitSet, run_batch_fit
    wl = workload(1)
    recs = rr.RuntimeRecords()
    n = 1
    while n <= max_events:
        events = make_events(wl, n, seed=SEED + 3)
        cfg = Config(wl)
        hs = HitSet(cfg, events)
        b, _ = run_batch_fit(dev, cfg, hs, Caps.for_hits(n, hs.n_hits, per_event=True, fuse_filter=True))
        shape = (n, wl.n_layers, len(wl.triplets), wl.n_threads, hs.n_hits, sum(e.n_tracks for e in events))
        for _ in range(reps):
            b.rerun()
            if b.wait() != 0:
                raise RuntimeError("capacity overflow in the event sweep")
            recs.add_record(rr.RuntimeRecord(*shape).fill_runtimes(b))
        b.close()
        print(f"event sweep: {n} events", file=sys.stderr, flush=True)
        n *= 2
    with open(path, "w") as f:
        f.write(recs.csv_dump("four_stage"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-events", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shapes instead of the autotuner's")
    ap.add_argument("--no-fuse", action="store_true",
                    help="predictor and filter as separate passes (the candidates stored) instead of one")
    ap.add_argument("--tuning", type=str, default=None,
                    help="fixed launch shapes as JSON (e.g. the autotuner's choice), skipping the autotuner")
    ap.add_argument("--e2e-reps", type=int, default=8,
                    help="copies of the step's events in the end-to-end (first H2D to last D2H) run (0 = skip); "
                         "the first upload and the last download are not overlapped, so more copies weigh them less")
    ap.add_argument("--no-fp64-check", action="store_true",
                    help="skip the float64 adjudication of the filter's decisions (a run for the profiler)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device copy (traffic-counter calibration)")
    ap.add_argument("--serve-events", type=int, default=10000, help="events of the serving stream (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=4000.0, help="Poisson arrival rate of the stream (events/s)")
    ap.add_argument("--serve-sweep", type=str, default="8000,16000,32000,64000",
                    help="further arrival rates (events/s) of the load sweep, comma-separated ('' = none)")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-max-batch", type=int, default=256)
    ap.add_argument("--serve-slot-hits", type=int, default=2_000_000,
                    help="hit capacity of each serving slot (its output capacities scale with it)")
    ap.add_argument("--runtime-csv", type=str, default=None,
                    help="also write the timed steps as the reference's runtime table (RuntimeRecords::csvDump)")
    ap.add_argument("--runtime-sweep", type=str, default=None,
                    help="also run the reference's published event sweep shape and write its runtime table")
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="processes generating the serving stream")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: events shard across GPUs with no communication, "
              f"so --gpus {args.gpus} is not measured here", file=sys.stderr)
        sys.exit(2)

    from trax_amd import make_events, workload
    from trax_amd.api import Batch, Caps, Config, Device, HitSet, OutputBounds, autotune, run_batch_fit

    wl = workload(3)
    t_gen = time.perf_counter()
    stream = serving_stream(wl, args.serve_events, SEED + 1, args.workers) if args.serve_events else None
    events = make_events(wl, N_EVENTS, seed=SEED)
    cfg = Config(wl)
    hs = HitSet(cfg, events)
    print(f"inputs generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    dev = Device(0)

    # capacity plan from a first run (grown on overflow), then fixed with headroom
    probe, _ = run_batch_fit(dev, cfg, hs, Caps.for_hits(N_EVENTS, hs.n_hits, use_graph=False))
    c = probe.counts()
    probe.close()
    # output bounds per hit for the server: the whole mix and its 16 densest events
    dense = HitSet(cfg, sorted(events, key=lambda ev: ev.n_hits)[-16:])
    probe, _ = run_batch_fit(dev, cfg, dense, Caps.for_hits(16, dense.n_hits, use_graph=False))
    bounds = OutputBounds.measure([c, probe.counts()], [hs.n_hits, dense.n_hits], headroom=1.25)
    probe.close()
    head = lambda v: int(v * 1.15) + 4096
    # the timed steps replay the whole sequence as one graph (use_graph 2); the
    # per-phase times come from a second run with a graph per phase (use_graph 1)
    caps = Caps(N_EVENTS, hs.n_hits, head(c["pairs"]), head(c["cands"]), head(c["tracklets"]), head(c["conn"]),
                head(c["track_entries"]), use_graph=0 if args.no_graph else 2, per_event=True,
                fuse_filter=not args.no_fuse)
    # launch shapes for this workload from the autotuner (results do not depend on them)
    t_tune = time.perf_counter()
    if args.tuning:
        tuning, trials = json.loads(args.tuning), []
    elif args.no_autotune:
        tuning, trials = {}, []
    else:
        tuning, trials = autotune(dev, cfg, hs, caps, repeats=2)
    tune_s = time.perf_counter() - t_tune
    print(f"autotuned in {tune_s:.1f} s: {tuning}", file=sys.stderr, flush=True)
    b = Batch(dev, cfg, caps, tuning)
    b.submit(hs)  # upload once; later steps run on the resident hits
    if b.wait() != 0:
        raise RuntimeError("capacity overflow after planning")
    for _ in range(args.warmup):
        b.rerun()
        b.wait()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.rerun()
        if b.wait() != 0:
            raise RuntimeError("capacity overflow")
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    counts = b.counts()
    ev_s = N_EVENTS / (ms_step / 1e3)
    trip_s = counts["tracklets"] / (ms_step / 1e3)
    b.close()

    # the phases of the same steps, each timed by HIP events on the batch's
    # stream (a graph per phase, or eager launches with --no-graph)
    phases = ["grid_count", "grid_store", "pair_plan", "pair_count", "pair_scan", "pair_store", "predict",
              "filter_count", "filter_scan", "filter_store", "link_prepare", "link_count", "link_store",
              "circle_fit", "ca_forward", "ca_backward"]
    caps_phase = Caps(**{**caps.__dict__, "use_graph": 0 if args.no_graph else 1})
    b = Batch(dev, cfg, caps_phase, tuning)
    b.submit(hs)
    if b.wait() != 0:
        raise RuntimeError("capacity overflow after planning")
    for _ in range(args.warmup):
        b.rerun()
        b.wait()
    acc = dict.fromkeys(phases, 0.0)
    step_phase_ms = []
    t2 = time.perf_counter()
    for _ in range(args.steps):
        b.rerun()
        if b.wait() != 0:
            raise RuntimeError("capacity overflow")
        step_phase_ms.append({p: b.stage_ms(p) for p in phases})
    ms_step_phase_graphs = (time.perf_counter() - t2) * 1e3 / args.steps
    for step in step_phase_ms:
        for p in phases:
            acc[p] += step[p]
    if args.runtime_csv:
        write_runtime_csv(args.runtime_csv, wl, hs, events, step_phase_ms)
    phase_ms = {p: acc[p] / args.steps for p in phases}

    b.close()  # the servers below hold their own workspaces
    serving, serving_sweep = None, []
    if stream is not None:
        print("serving stream", file=sys.stderr, flush=True)
        rates = [args.serve_rate] + [float(r) for r in args.serve_sweep.split(",") if r.strip()]
        for group in (True, False):
            serving_sweep += serve(dev, stream[0], stream[1], bounds, rates, SEED + 2, args.serve_slots,
                                   args.serve_max_batch, args.serve_slot_hits, not args.no_fuse, group)
        serving = serving_sweep[0]  # grouped by hit count, at the stated rate

    # from the first H2D to the last D2H: the step's events, four times over,
    # all waiting at the start, through the continuous-batching server (hits
    # uploaded from pinned host memory, tracklets and CA tracks downloaded and
    # split per event, two batches in flight so transfers overlap compute)
    e2e, e2e_ev_s = None, None
    if args.e2e_reps > 0:
        e2e = end_to_end(dev, cfg, events, bounds, hs.n_hits, not args.no_fuse, reps=args.e2e_reps)
        e2e_ev_s = e2e["events_per_s"]
        print(f"end to end: {e2e}", file=sys.stderr, flush=True)

    # plain per-event path on a sample of the same events
    n_plain = min(args.plain_events, N_EVENTS)
    plain_ev_s, batch_ev_s_same_mix = None, None
    if n_plain > 0:
        singles = [HitSet(cfg, [ev]) for ev in events[:n_plain]]
        dev.run_plain_bench(cfg, singles[0])  # warm-up
        plain_ms = sum(dev.run_plain_bench(cfg, s) for s in singles)
        sample_hits = sum(s.n_hits for s in singles)
        plain_ev_s = n_plain / (plain_ms / 1e3)
        # the batch rates on the same mix of event sizes (the plain path's time
        # includes its uploads and downloads, so the end-to-end rate is the
        # like-for-like one)
        mix = (hs.n_hits / N_EVENTS) / (sample_hits / n_plain)
        batch_ev_s_same_mix = ev_s * mix
        batch_e2e_ev_s_same_mix = e2e_ev_s * mix if e2e_ev_s else None

    if args.copy_calibration:
        # a plain copy of known size: HBM traffic counters are scaled by what
        # they read for it (4 B read + 4 B written per element)
        import ctypes as C
        from trax_amd.api import check, lib
        n = 1 << 26
        a, bb = C.c_void_p(), C.c_void_p()
        check(lib().trax_buf_alloc(dev.h, 4 * n, C.byref(a)))
        check(lib().trax_buf_alloc(dev.h, 4 * n, C.byref(bb)))
        check(lib().trax_buf_fill_u32(dev.h, a, 7, n, None))
        for _ in range(3):
            check(lib().trax_buf_copy_u32(dev.h, bb, a, n, None))
        check(lib().trax_stream_sync(dev.h, None))
        lib().trax_buf_free(dev.h, a)
        lib().trax_buf_free(dev.h, bb)

    scan = scan_hbm(dev)
    print(f"scan alone: {scan}", file=sys.stderr, flush=True)
    caps64 = Caps(N_EVENTS, hs.n_hits, caps.max_pairs, head(c["cands"]), caps.max_tracklets, caps.max_conn,
                  caps.max_track_entries, use_graph=False, per_event=True, fuse_filter=False)
    fp64 = None if args.no_fp64_check else fp64_check(dev, wl, cfg, hs, caps64)
    print(f"float64 check: {fp64}", file=sys.stderr, flush=True)

    # roofline of the dominant phase, HBM-bound
    cols = hs.columns()
    lay = cols["layer_nhits"].reshape(N_EVENTS, wl.n_layers).astype(int)
    seg_hits = sum(int(lay[:, t.layer1 - 1].sum() + lay[:, t.layer2 - 1].sum()) for t in wl.triplets)
    trip_hits = sum(int(lay[:, t.layer1 - 1].sum() + lay[:, t.layer2 - 1].sum() + lay[:, t.layer3 - 1].sum())
                    for t in wl.triplets)
    cells = N_EVENTS * wl.n_layers * (wl.n_z + 1) * (wl.n_phi + 1)
    counts["slots"] = hs.n_hits + N_EVENTS * len(wl.triplets) * wl.n_threads
    pb = phase_bytes(hs.n_hits, seg_hits, trip_hits, counts, cells, wl.n_threads, not args.no_fuse)
    kernel_of = PHASE_KERNEL if args.no_fuse else PHASE_KERNEL_FUSED
    gbs = {p: pb[p] / (phase_ms[p] * 1e-3) / 1e9 for p in pb if phase_ms[p] > 0}
    dom = max(pb, key=lambda p: phase_ms[p])
    achieved = pb[dom] / (phase_ms[dom] * 1e-3)
    grid_ms = phase_ms["grid_count"] + phase_ms["grid_store"]
    grid_gbs = pb["grid_count"] / (grid_ms * 1e-3) / 1e9

    line = {
        "metric": "events/s & triplets/s, batched vs plain per-event HIP; p50/p99 event latency",
        "value": round(ev_s, 1),
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "ms_per_step_phase_graphs": round(ms_step_phase_graphs, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "events_per_step": N_EVENTS, "hits_per_step": hs.n_hits,
                   "layers": wl.n_layers, "layer_triplets": len(wl.triplets), "grid": [wl.n_z, wl.n_phi],
                   "seed": SEED, "graph": "eager" if args.no_graph else "one graph per step", "event_groups": "per-event",
                   "predictor_filter": "separate" if args.no_fuse else "fused",
                   "tuning": tuning, "autotune_s": round(tune_s, 2)},
        "triplets_per_s": round(trip_s, 1),
        "triplets_per_step": counts["tracklets"],
        "pairs_per_step": counts["pairs"],
        "candidates_per_step": counts["cands"],
        "plain_events_per_s": round(plain_ev_s, 1) if plain_ev_s else None,
        "plain_sample_events": n_plain,
        "batched_vs_plain_same_mix": round(batch_ev_s_same_mix / plain_ev_s, 2) if plain_ev_s else None,
        "events_per_s_h2d_to_d2h": e2e_ev_s,
        "h2d_to_d2h": e2e,
        "batched_h2d_to_d2h_vs_plain_same_mix":
            round(batch_e2e_ev_s_same_mix / plain_ev_s, 2) if plain_ev_s and e2e_ev_s else None,
        "grid_build_hbm_gbs": round(grid_gbs, 1),
        "grid_build_hbm_frac": round(grid_gbs * 1e9 / HBM_PEAK, 4),
        "scan_hbm": scan,
        "fp64_check": fp64,
        "phase_ms": {k: round(v, 4) for k, v in phase_ms.items()},
        "phase_gbs": {k: round(v, 1) for k, v in gbs.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": measured_traffic(kernel_of.get(dom, dom)),
                     "kernel": kernel_of.get(dom, dom), "algorithmic_bytes": int(pb[dom]),
                     "valu_busy": measured_valu_busy(kernel_of.get(dom, dom))},
        "latency_ms": {"p50": serving["p50_ms"] if serving else None,
                       "p99": serving["p99_ms"] if serving else None},
        "serving": serving,
        "serving_sweep": [{k: r[k] for k in ("admission", "rate_hz", "events_per_s", "p50_ms", "p99_ms",
                                             "mean_batch_events", "batches")} for r in serving_sweep],
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(wl, events, Config, HitSet)
    if args.runtime_sweep:
        runtime_sweep(dev, args.runtime_sweep)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
