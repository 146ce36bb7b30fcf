// Serving loop: events arrive one by one, are grouped into batches of whole
// events and run through the batched pipeline on several in-flight slots
// (one stream and one workspace each), with each event's results copied back
// to the host as soon as its batch completes.
//
// The reference processes a fixed event group per call with one in-order
// queue (src/triplet_builder/TripletBuilder.cpp:297-552); this loop keeps the
// reference's per-event results (tracklets of TripletThetaPhiFilter, tracks of
// CellularAutomaton) and adds continuous batching in front of it:
//   - admission: a batch takes waiting events while its hit count fits the
//     slot's hit capacity, which the device-memory budget sets; output
//     capacities are bounds per input hit measured on the workload.  In
//     arrival order (consecutive events of the stream), or grouped by hit
//     count: the oldest waiting event's size class (hit count by powers of
//     four) first, in arrival order, then the other waiting events to fill
//     the batch, so that events of like size share batches;
//   - an output that still overflows re-queues the batch's events at the
//     front, and the next batches take half as many events;
//   - completion: the counts are read when the slot's stream has drained, the
//     results of exactly that size are copied back, and the events complete
//     when that copy has landed;
//   - overlap: a slot's uploads go out as soon as it takes a batch, but its
//     pipeline waits for the previously submitted one (an event wait on the
//     device), so pipelines run one at a time and every slot's uploads and
//     downloads overlap another slot's pipeline.  Two pipelines side by side
//     take as long as one after the other, and would finish together, with
//     both downloads left to follow.  Each slot's batch keeps all its work on
//     the slot's one stream (no second stream for the fine walk structure):
//     the device has few hardware queues (four by default), streams beyond
//     them share one, and a slot whose stream shares a queue with another
//     slot's pipeline sees its downloads' completion only after that
//     pipeline.  The counts and the CA's flags are copied back at the end of
//     the pipeline itself for the same reason.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <thread>

#include "capi_internal.h"

using namespace trax;
using namespace trax_capi;

namespace {

using Clock = std::chrono::steady_clock;

struct EventOut {
    std::vector<uint32_t> h1, h2, h3, track_offsets, tracks, track_hits, track_hit_offsets;
};

struct Slot {
    trax_batch* b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t ev_pipe = nullptr, ev_dl = nullptr;
    int state = 0;  // 0 free, 1 pipeline running, 2 results downloading
    std::vector<uint32_t> events;    // stream indices of the batch's events
    std::vector<uint32_t> hit_base;  // batch-relative first hit of each event, plus the total
    uint32_t R = 0, TE = 0, K = 0;
    double t_submit = 0, t_pipe = 0;  // seconds from the replay's start
    uint32_t *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *handler = nullptr, *tracks = nullptr;  // pinned
};

template <class T>
T* pinned(size_t n) {
    void* p = nullptr;
    hip_check(hipHostMalloc(&p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    return static_cast<T*>(p);
}

}  // namespace

struct trax_server {
    trax_ctx* ctx = nullptr;
    trax_config cfg;
    trax_serve_opts opts{};
    trax_batch_caps caps{};
    uint64_t device_bytes = 0;
    std::vector<Slot> slots;
    // the replayed stream, in pinned memory
    float *x = nullptr, *y = nullptr, *z = nullptr;
    uint32_t *layer = nullptr, *det = nullptr, *hid = nullptr;
    uint64_t store_hits = 0;
    std::vector<uint32_t> evt_off, evt_n, lay_n, lay_off;
    uint32_t L = 0, n_events = 0;
    // per event
    std::vector<double> arrival, done;
    std::vector<uint32_t> n_tracklets, n_track_entries;
    std::vector<EventOut> out;
    std::vector<int> failed;
    // counters
    uint32_t batches = 0, resubmits = 0, max_batch = 0;
    uint64_t batch_event_sum = 0;
    double sum_queue = 0, sum_pipe = 0, sum_copy = 0, sum_split = 0;
    // the completion event of the most recently submitted batch's pipeline:
    // each pipeline starts after the previous one, so the slots' pipelines
    // run one at a time and a slot's copies overlap another slot's pipeline
    hipEvent_t last_pipe = nullptr;

    ~trax_server() { release(); }

    void free_store() {
        for (void* p : {(void*)x, (void*)y, (void*)z, (void*)layer, (void*)det, (void*)hid})
            if (p) (void)hipHostFree(p);
        x = y = z = nullptr;
        layer = det = hid = nullptr;
    }

    void release() {
        for (Slot& s : slots) {
            if (s.st) (void)hipStreamSynchronize(s.st);
            if (s.b) trax_batch_destroy(s.b);
            for (void* p : {(void*)s.h1, (void*)s.h2, (void*)s.h3, (void*)s.handler, (void*)s.tracks})
                if (p) (void)hipHostFree(p);
            if (s.ev_pipe) (void)hipEventDestroy(s.ev_pipe);
            if (s.ev_dl) (void)hipEventDestroy(s.ev_dl);
            if (s.st) (void)hipStreamDestroy(s.st);
        }
        slots.clear();
        free_store();
    }

    void load(const trax_hits* h) {
        free_store();
        const HitCollection& c = h->hits;
        uint64_t running = 0;
        for (uint32_t e = 0; e < h->n_events; ++e) {
            if (h->evt_offset[e] != running) throw InputError("event offsets are not running sums");
            running += h->evt_nhits[e];
        }
        if (running != c.size()) throw InputError("event sizes do not add up to the hit count");
        store_hits = c.size();
        x = pinned<float>(store_hits); y = pinned<float>(store_hits); z = pinned<float>(store_hits);
        layer = pinned<uint32_t>(store_hits); det = pinned<uint32_t>(store_hits); hid = pinned<uint32_t>(store_hits);
        std::memcpy(x, c.x.data(), store_hits * 4); std::memcpy(y, c.y.data(), store_hits * 4);
        std::memcpy(z, c.z.data(), store_hits * 4); std::memcpy(layer, c.layer.data(), store_hits * 4);
        std::memcpy(det, c.detId.data(), store_hits * 4); std::memcpy(hid, c.hitId.data(), store_hits * 4);
        n_events = h->n_events;
        L = h->max_layer;
        evt_off = h->evt_offset; evt_n = h->evt_nhits;
        lay_n = h->lay_nhits; lay_off = h->lay_offset;
    }

    // size class of an event for grouping: its hit count by powers of four
    uint32_t size_class(uint32_t e) const {
        uint32_t c = 0;
        for (uint64_t n = evt_n[e]; n >= 4; n >>= 2) ++c;
        return c;
    }

    // stage the events into the slot's workspace and enqueue the pipeline
    void submit(Slot& s, std::vector<uint32_t> evs, double now) {
        trax_batch* b = s.b;
        s.t_submit = now;
        sum_queue += now - arrival[evs[0]];
        const uint32_t E = (uint32_t)evs.size(), Ecap = b->E;
        s.events = std::move(evs);
        s.hit_base.assign(E + 1, 0);
        uint32_t* sp = b->staging + 7ull * b->caps.max_hits;  // supplements of the staging area
        for (uint32_t k = 0; k < E; ++k) {
            const uint32_t e = s.events[k];
            sp[k] = s.hit_base[k];
            std::memcpy(sp + Ecap + (size_t)k * L, &lay_n[(size_t)e * L], L * 4ull);
            std::memcpy(sp + Ecap + (size_t)Ecap * L + (size_t)k * L, &lay_off[(size_t)e * L], L * 4ull);
            s.hit_base[k + 1] = s.hit_base[k] + evt_n[e];
        }
        const uint32_t N = s.hit_base[E];
        uint32_t* shape = sp + Ecap + 2ull * Ecap * L;
        shape[0] = E;
        shape[1] = N;
        // consecutive events of the stream are one contiguous range of the
        // pinned store and go up directly; others are gathered into the
        // slot's pinned staging area first (the slot's stream has drained, so
        // the area is free), one copy per run of consecutive events
        bool consecutive = true;
        for (uint32_t k = 1; k < E; ++k) consecutive = consecutive && s.events[k] == s.events[k - 1] + 1;
        const uint32_t* src[6] = {reinterpret_cast<const uint32_t*>(x), reinterpret_cast<const uint32_t*>(y),
                                  reinterpret_cast<const uint32_t*>(z), layer, det, hid};
        uint32_t* dst[6] = {reinterpret_cast<uint32_t*>(b->hin.x), reinterpret_cast<uint32_t*>(b->hin.y),
                            reinterpret_cast<uint32_t*>(b->hin.z), b->hin.layer, b->hin.det, b->hin.id};
        for (int c = 0; c < 6; ++c) {
            const uint32_t* from = src[c] + evt_off[s.events[0]];
            if (!consecutive) {
                uint32_t* stg = b->staging + (size_t)c * b->caps.max_hits;
                for (uint32_t k = 0; k < E;) {
                    uint32_t k1 = k + 1;
                    while (k1 < E && s.events[k1] == s.events[k1 - 1] + 1) ++k1;
                    const uint64_t n_run = s.hit_base[k1] - s.hit_base[k];
                    std::memcpy(stg + s.hit_base[k], src[c] + evt_off[s.events[k]], n_run * 4);
                    k = k1;
                }
                from = stg;
            }
            hip_check(hipMemcpyAsync(dst[c], from, N * 4ull, hipMemcpyHostToDevice, s.st), "h2d");
        }
        hip_check(hipMemcpyAsync(b->sup, sp, (Ecap + 2ull * Ecap * L) * 4, hipMemcpyHostToDevice, s.st), "h2d");
        hip_check(hipMemcpyAsync(b->cnt, shape, 8, hipMemcpyHostToDevice, s.st), "h2d");
        hip_check(launch_event_ids(b->hin.evt, b->sup, b->cnt, Ecap, s.st), "event ids");
        b->n_events = E;
        b->n_hits = N;
        // the uploads above go out at once; the pipeline waits for the one before it
        if (last_pipe && last_pipe != s.ev_pipe) hip_check(hipStreamWaitEvent(s.st, last_pipe, 0), "wait");
        b->launch();
        hip_check(hipEventRecord(s.ev_pipe, s.st), "event");
        last_pipe = s.ev_pipe;
        s.state = 1;
        ++batches;
        batch_event_sum += E;
        max_batch = std::max(max_batch, E);
    }

    // the pipeline has drained: copy back exactly the results it produced
    // returns false if the batch overflowed a capacity (events re-queued)
    bool collect(Slot& s) {
        trax_batch* b = s.b;
        const int rc = b->finish();
        if (rc == TRAX_ERR_CAPACITY) return false;
        if (rc != TRAX_OK) throw InputError(g_last_error);
        const DevCounts& c = *b->h_counts;
        s.R = c.n_tracklets;
        s.K = c.n_conn;
        s.TE = s.K ? c.n_track_entries : 0;
        auto d2h = [&](uint32_t* dst, const uint32_t* src, uint64_t n) {
            if (n) hip_check(hipMemcpyAsync(dst, src, n * 4, hipMemcpyDeviceToHost, s.st), "d2h");
        };
        d2h(s.h1, b->th1, s.R);
        d2h(s.h2, b->th2, s.R);
        d2h(s.h3, b->th3, s.R);
        if (s.K) {
            d2h(s.handler, b->cb.handler, s.R + 1ull);
            d2h(s.tracks, b->cb.tracks, s.TE);
        }
        hip_check(hipEventRecord(s.ev_dl, s.st), "event");
        s.state = 2;
        return true;
    }

    // results are on the host: split them by event
    void complete(Slot& s, double now) {
        const uint32_t E = (uint32_t)s.events.size();
        uint32_t t_lo = 0;
        for (uint32_t k = 0; k < E; ++k) {
            const uint32_t e = s.events[k];
            // tracklets are ordered by event; their first hit says which
            const uint32_t hi_hit = s.hit_base[k + 1];
            const uint32_t t_hi = (uint32_t)(std::lower_bound(s.h1 + t_lo, s.h1 + s.R, hi_hit) - s.h1);
            done[e] = now;
            n_tracklets[e] = t_hi - t_lo;
            // the event has a connection of its own iff one of its chains is
            // longer than one tracklet; without one a run of the event alone
            // has no CA outputs (the driver skips the CA, TripletBuilder.cpp:414-423), and so
            // neither has its result here
            bool has_conn = false;
            if (s.K)
                for (uint32_t t = t_lo; t < t_hi && !has_conn; ++t) has_conn = s.handler[t + 1] - s.handler[t] > 1;
            const uint32_t tr_lo = has_conn ? s.handler[t_lo] : 0, tr_hi = has_conn ? s.handler[t_hi] : 0;
            n_track_entries[e] = tr_hi - tr_lo;
            if (opts.keep_results) {
                EventOut& o = out[e];
                const uint32_t hb = s.hit_base[k];
                o.h1.resize(t_hi - t_lo); o.h2.resize(t_hi - t_lo); o.h3.resize(t_hi - t_lo);
                for (uint32_t t = t_lo; t < t_hi; ++t) {
                    o.h1[t - t_lo] = s.h1[t] - hb;
                    o.h2[t - t_lo] = s.h2[t] - hb;
                    o.h3[t - t_lo] = s.h3[t] - hb;
                }
                o.track_offsets.clear();
                o.tracks.clear();
                o.track_hits.clear();
                o.track_hit_offsets.assign(1, 0);
                if (has_conn) {
                    o.track_offsets.resize(t_hi - t_lo + 1);
                    for (uint32_t t = t_lo; t <= t_hi; ++t) o.track_offsets[t - t_lo] = s.handler[t] - tr_lo;
                    o.tracks.resize(tr_hi - tr_lo);
                    for (uint32_t i = tr_lo; i < tr_hi; ++i) o.tracks[i - tr_lo] = s.tracks[i] - t_lo;
                    // extractTracks (TripletBuilder.cpp:955-988): chains of at
                    // least three tracklets, innermost first, as hit lists
                    for (uint32_t t = t_lo; t < t_hi; ++t) {
                        const uint32_t off = s.handler[t], len = s.handler[t + 1] - off;
                        if (len < 3) continue;
                        for (uint32_t j = 0; j < len; ++j) {
                            const uint32_t tr = s.tracks[off + (len - 1 - j)];
                            if (j == 0) {
                                o.track_hits.push_back(s.h1[tr] - hb);
                                o.track_hits.push_back(s.h2[tr] - hb);
                            }
                            o.track_hits.push_back(s.h3[tr] - hb);
                        }
                        o.track_hit_offsets.push_back((uint32_t)o.track_hits.size());
                    }
                }
            }
            t_lo = t_hi;
        }
        s.state = 0;
        s.events.clear();
    }
};

extern "C" int trax_server_create(trax_ctx* ctx, const trax_config* cfg, const trax_serve_opts* opts,
                                  trax_server** out) {
    return guarded([&] {
        if (!ctx || !cfg || !opts || !out) return fail(TRAX_ERR_ARG, "null argument");
        validate_config(cfg);
        hip_check(hipSetDevice(ctx->c->device()), "hipSetDevice");
        std::unique_ptr<trax_server> sv(new trax_server);
        sv->ctx = ctx;
        sv->cfg = *cfg;
        sv->opts = *opts;
        trax_serve_opts& o = sv->opts;
        if (!o.n_slots) o.n_slots = 2;
        if (!o.max_batch_events) o.max_batch_events = 256;
        if (!o.device_budget) o.device_budget = 32ull << 30;
        if (!(o.pairs_per_hit > 0) || !(o.cands_per_hit > 0) || !(o.tracklets_per_hit > 0) || !(o.conn_per_hit > 0) ||
            !(o.track_entries_per_hit > 0))
            return fail(TRAX_ERR_ARG, "output bounds per hit must be positive");
        auto caps_for = [&](uint64_t hits) {
            trax_batch_caps c{};
            auto cap = [&](float per_hit) {
                const double v = std::ceil((double)per_hit * (double)hits) + 4096.0;
                return (uint32_t)std::min(v, 4.0e9);
            };
            c.max_events = o.max_batch_events;
            c.max_hits = (uint32_t)hits;
            c.max_pairs = cap(o.pairs_per_hit);
            c.max_cands = cap(o.cands_per_hit);
            c.max_tracklets = cap(o.tracklets_per_hit);
            c.max_conn = cap(o.conn_per_hit);
            c.max_track_entries = cap(o.track_entries_per_hit);
            c.use_graph = o.use_graph;
            c.fuse_filter = o.fuse_filter;
            c.per_event = 1;  // an event's results do not depend on the batch it shares
            return c;
        };
        auto bytes_for = [&](uint64_t hits) -> uint64_t {
            trax_batch_caps c = caps_for(hits);
            uint64_t dev = 0;
            if (trax_batch_plan(cfg, &c, &dev, nullptr) != TRAX_OK) return ~0ull;
            return dev;
        };
        const uint64_t per_slot = o.device_budget / o.n_slots;
        uint64_t hits = o.max_batch_hits;
        if (!hits) {  // the largest hit capacity whose workspace fits the budget
            uint64_t lo = 0, hi = 1ull << 28;
            while (lo < hi) {
                const uint64_t mid = (lo + hi + 1) / 2;
                if (bytes_for(mid) <= per_slot) lo = mid;
                else hi = mid - 1;
            }
            hits = lo;
        }
        if (hits == 0 || bytes_for(hits) > per_slot)
            return fail(TRAX_ERR_CAPACITY, "the device budget does not hold one slot of the requested size");
        o.max_batch_hits = (uint32_t)hits;
        sv->caps = caps_for(hits);
        sv->slots.resize(o.n_slots);
        for (Slot& s : sv->slots) {
            hip_check(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreateWithFlags(&s.ev_pipe, hipEventDisableTiming), "hipEventCreate");
            hip_check(hipEventCreateWithFlags(&s.ev_dl, hipEventDisableTiming), "hipEventCreate");
            trax_stream ts;
            ts.s = s.st;
            const int rc = trax_batch_create(ctx, cfg, &sv->caps, nullptr, &ts, &s.b);
            if (rc != TRAX_OK) return rc;
            s.b->single_stream();  // one stream per slot (see the overlap note above)
            sv->device_bytes += trax_batch_footprint(s.b);
            const uint64_t R = sv->caps.max_tracklets;
            s.h1 = pinned<uint32_t>(R); s.h2 = pinned<uint32_t>(R); s.h3 = pinned<uint32_t>(R);
            s.handler = pinned<uint32_t>(R + 1);
            s.tracks = pinned<uint32_t>(sv->caps.max_track_entries);
        }
        *out = sv.release();
        return TRAX_OK;
    });
}

extern "C" void trax_server_destroy(trax_server* s) { delete s; }

extern "C" int trax_server_caps(const trax_server* s, trax_batch_caps* caps) {
    if (!s || !caps) return fail(TRAX_ERR_ARG, "null argument");
    *caps = s->caps;
    return TRAX_OK;
}

extern "C" int trax_server_replay(trax_server* sv, const trax_hits* events, const double* arrival_s,
                                  double* latency_ms, trax_serve_stats* stats) {
    return guarded([&] {
        if (!sv || !events || !arrival_s) return fail(TRAX_ERR_ARG, "null argument");
        if (events->max_layer != sv->cfg.trip.maxLayer())
            return fail(TRAX_ERR_INPUT, "hit collection layers differ from the configuration");
        const uint32_t n = events->n_events;
        for (uint32_t i = 1; i < n; ++i)
            if (!(arrival_s[i] >= arrival_s[i - 1])) return fail(TRAX_ERR_ARG, "arrival times must be non-decreasing");
        for (uint32_t i = 0; i < n; ++i)
            if (events->evt_nhits[i] > sv->caps.max_hits)
                return fail(TRAX_ERR_CAPACITY, "an event has more hits than a batch holds");
        hip_check(hipSetDevice(sv->ctx->c->device()), "hipSetDevice");
        sv->load(events);
        sv->arrival.assign(arrival_s, arrival_s + n);
        sv->done.assign(n, -1.0);
        sv->n_tracklets.assign(n, 0);
        sv->n_track_entries.assign(n, 0);
        sv->failed.assign(n, 0);
        sv->out.assign(sv->opts.keep_results ? n : 0, EventOut{});
        sv->batches = sv->resubmits = sv->max_batch = 0;
        sv->batch_event_sum = 0;
        sv->sum_queue = sv->sum_pipe = sv->sum_copy = sv->sum_split = 0;
        sv->last_pipe = nullptr;

        std::deque<uint32_t> waiting;
        uint32_t next = 0, finished = 0, limit = sv->opts.max_batch_events;
        const Clock::time_point t0 = Clock::now();
        auto now_s = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
        while (finished < n) {
            bool progressed = false;
            double now = now_s();
            while (next < n && sv->arrival[next] <= now) waiting.push_back(next++);
            for (Slot& s : sv->slots) {
                if (s.state == 1 && hipEventQuery(s.ev_pipe) == hipSuccess) {
                    progressed = true;
                    s.t_pipe = now_s();
                    sv->sum_pipe += s.t_pipe - s.t_submit;
                    if (!sv->collect(s)) {  // overflow: re-queue in order, smaller batches
                        ++sv->resubmits;
                        if (s.events.size() == 1) {
                            sv->failed[s.events[0]] = 1;
                            sv->done[s.events[0]] = now_s();
                            ++finished;
                        } else {
                            for (auto it = s.events.rbegin(); it != s.events.rend(); ++it) waiting.push_front(*it);
                            limit = std::max<uint32_t>(1, (uint32_t)s.events.size() / 2);
                        }
                        s.events.clear();
                        s.state = 0;
                    }
                } else if (s.state == 2 && hipEventQuery(s.ev_dl) == hipSuccess) {
                    progressed = true;
                    finished += (uint32_t)s.events.size();
                    const double t_dl = now_s();
                    sv->sum_copy += t_dl - s.t_pipe;
                    sv->complete(s, t_dl);
                    sv->sum_split += now_s() - t_dl;
                    limit = std::min(sv->opts.max_batch_events, limit * 2);
                }
            }
            for (Slot& s : sv->slots) {
                if (s.state != 0 || waiting.empty()) continue;
                std::vector<uint32_t> evs;
                uint64_t hits = 0;
                if (!sv->opts.group_by_size) {
                    // events in arrival order, all consecutive, within the hit capacity
                    while (!waiting.empty() && evs.size() < limit) {
                        const uint32_t e = waiting.front();
                        if (!evs.empty() && (e != evs.back() + 1 || hits + sv->evt_n[e] > sv->caps.max_hits)) break;
                        hits += sv->evt_n[e];
                        evs.push_back(e);
                        waiting.pop_front();
                    }
                } else {
                    // the oldest event's size class first, then the rest, each in arrival order
                    const uint32_t cls = sv->size_class(waiting.front());
                    std::vector<char> taken(waiting.size(), 0);
                    for (int pass = 0; pass < 2; ++pass)
                        for (size_t i = 0; i < waiting.size() && evs.size() < limit; ++i) {
                            const uint32_t e = waiting[i];
                            if (taken[i] || (pass == 0 && sv->size_class(e) != cls)) continue;
                            if (hits + sv->evt_n[e] > sv->caps.max_hits) {
                                if (evs.empty()) continue;  // cannot happen: every event fits a batch alone
                                break;
                            }
                            hits += sv->evt_n[e];
                            evs.push_back(e);
                            taken[i] = 1;
                        }
                    std::deque<uint32_t> rest;
                    for (size_t i = 0; i < waiting.size(); ++i)
                        if (!taken[i]) rest.push_back(waiting[i]);
                    waiting.swap(rest);
                    std::sort(evs.begin(), evs.end());  // stream order inside the batch: runs copy in one piece
                }
                sv->submit(s, std::move(evs), now_s());
                progressed = true;
            }
            if (!progressed) {
                // idle: sleep until the next arrival, polling the slots meanwhile
                const double wait = next < n ? sv->arrival[next] - now_s() : 1e-4;
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 2e-5)));
            }
        }
        const double wall = now_s();
        std::vector<double> lat;
        lat.reserve(n);
        uint64_t trk = 0, te = 0;
        uint32_t n_failed = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const double l = (sv->done[i] - sv->arrival[i]) * 1e3;
            if (latency_ms) latency_ms[i] = sv->failed[i] ? -1.0 : l;
            if (sv->failed[i]) { ++n_failed; continue; }
            lat.push_back(l);
            trk += sv->n_tracklets[i];
            te += sv->n_track_entries[i];
        }
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->events = n;
            stats->failed = n_failed;
            stats->batches = sv->batches;
            stats->resubmits = sv->resubmits;
            stats->max_batch_events = sv->max_batch;
            stats->mean_batch_events = sv->batches ? (double)sv->batch_event_sum / sv->batches : 0.0;
            stats->slot_hits = sv->caps.max_hits;
            stats->device_bytes = sv->device_bytes;
            stats->tracklets = trk;
            stats->track_entries = te;
            stats->wall_s = wall;
            stats->events_per_s = wall > 0 ? n / wall : 0.0;
            stats->tracklets_per_s = wall > 0 ? trk / wall : 0.0;
            const double nb = sv->batches ? (double)sv->batches : 1.0;
            stats->queue_ms = sv->sum_queue / nb * 1e3;
            stats->pipeline_ms = sv->sum_pipe / nb * 1e3;
            stats->copy_ms = sv->sum_copy / nb * 1e3;
            stats->split_ms = sv->sum_split / nb * 1e3;
            if (!lat.empty()) {
                std::sort(lat.begin(), lat.end());
                auto q = [&](double p) {  // nearest-rank percentile
                    size_t k = (size_t)std::ceil(p * lat.size());
                    return lat[std::min(lat.size() - 1, k ? k - 1 : 0)];
                };
                stats->p50_ms = q(0.50);
                stats->p99_ms = q(0.99);
                stats->max_ms = lat.back();
            }
        }
        sv->free_store();
        return n_failed ? fail(TRAX_ERR_CAPACITY, "some events overflowed a capacity on their own") : TRAX_OK;
    });
}

extern "C" int trax_server_event_result(const trax_server* sv, uint32_t event, const char* name,
                                        const uint32_t** data, uint64_t* count) {
    if (!sv || !name || !data || !count) return fail(TRAX_ERR_ARG, "null argument");
    if (event >= sv->n_tracklets.size()) return fail(TRAX_ERR_ARG, "no such event");
    if (std::strcmp(name, "n_tracklets") == 0) {
        *data = &sv->n_tracklets[event];
        *count = 1;
        return TRAX_OK;
    }
    if (std::strcmp(name, "n_track_entries") == 0) {
        *data = &sv->n_track_entries[event];
        *count = 1;
        return TRAX_OK;
    }
    if (!sv->opts.keep_results) return fail(TRAX_ERR_STATE, "the server does not keep results");
    const EventOut& o = sv->out[event];
    const std::vector<uint32_t>* v = nullptr;
    if (std::strcmp(name, "tracklets.h1") == 0) v = &o.h1;
    else if (std::strcmp(name, "tracklets.h2") == 0) v = &o.h2;
    else if (std::strcmp(name, "tracklets.h3") == 0) v = &o.h3;
    else if (std::strcmp(name, "track_offsets") == 0) v = &o.track_offsets;
    else if (std::strcmp(name, "tracks") == 0) v = &o.tracks;
    else if (std::strcmp(name, "track_hits") == 0) v = &o.track_hits;
    else if (std::strcmp(name, "track_hit_offsets") == 0) v = &o.track_hit_offsets;
    if (!v) return fail(TRAX_ERR_ARG, std::string("no result array ") + name);
    *data = v->data();
    *count = v->size();
    return TRAX_OK;
}
