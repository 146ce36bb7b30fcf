// Stage launchers of the C ABI (include/trax_hip.h, section 2b): one call per
// kernel group of the reference's algorithm classes, on caller-owned device
// buffers, with the sizes of every stage passed from the host.  Each launcher
// writes the sizes its kernels read into the context's device-side record
// (by a kernel on the caller's stream, so a chain stays asynchronous and in
// order), then enqueues the same launchers the host classes and the batched
// pipeline use.
#include "capi_internal.h"

#include <cstring>

using namespace trax;
using namespace trax_capi;

struct trax_device_config {
    std::unique_ptr<ConfigMirror> cm;
    GridGeom g;
    uint32_t n_trip = 0;
};

namespace {

DevCounts* rec(trax_ctx* ctx) { return ctx->c->counts(); }

HitColumns cols_of(const trax_grid_hits* h) {
    HitColumns c;
    c.x = const_cast<float*>(h->x);
    c.y = const_cast<float*>(h->y);
    c.z = const_cast<float*>(h->z);
    c.layer = const_cast<uint32_t*>(h->layer);
    c.evt = const_cast<uint32_t*>(h->event);
    c.zphi = reinterpret_cast<float2*>(const_cast<float*>(h->zphi));
    return c;
}

void need(bool ok, const char* what) {
    if (!ok) throw std::invalid_argument(what);
}

void check_hits(const trax_grid_hits* h) {
    need(h && h->x && h->y && h->z && h->layer && h->event && h->zphi, "grid hits: null column");
    need((reinterpret_cast<uintptr_t>(h->zphi) & 15u) == 0, "grid hits: the (z, phi) column must be 16-byte aligned");
}

// sizes a stage's kernels read, written into the context's record on the stream
struct Sizes {
    trax_ctx* ctx;
    hipStream_t st;
    CountWrites w;
    Sizes(trax_ctx* c, trax_stream* s) : ctx(c), st(pick_stream(c, s)) {}
    Sizes& put(uint32_t DevCounts::*field, uint32_t v) {
        DevCounts* d = rec(ctx);
        w.put(&(d->*field), d, v);
        return *this;
    }
    void write() { hip_check(launch_set_counts(rec(ctx), w, st), "stage sizes"); }
};

LinkBufs link_bufs(const trax_link_bufs* b) {
    need(b != nullptr, "null link buffers");
    LinkBufs l;
    l.eta = b->eta;
    l.hit_count = b->hit_count;
    l.hit_cursor = b->hit_cursor;
    l.csr = b->csr;
    l.conn = b->conn;
    l.counts = b->counts;
    l.basis = b->basis;
    l.foll = b->followers;
    l.pev = b->event;
    l.flag_a = b->flag_a;
    l.flag_b = b->flag_b;
    l.connectable = b->connectable;
    l.nonconnectable = b->nonconnectable;
    l.runs = 0;       // the per-first-hit buckets, as the plain path builds them
    l.per_event = 0;  // the chain is one event group
    return l;
}

CABufs ca_bufs(const trax_ca_bufs* b) {
    need(b != nullptr, "null CA buffers");
    CABufs c;
    c.st_a = b->states;
    c.st_b = b->states_next;
    c.flags = b->flags;
    c.fbc = b->basis_count;
    c.cursor = b->cursor;
    c.bidx = b->basis_index;
    c.bdiff = b->basis_diff;
    c.best = b->best_basis;
    c.isbest = b->is_best;
    c.handler = b->track_offsets;
    c.tracks = b->tracks;
    return c;
}

}  // namespace

extern "C" int trax_device_config_create(trax_ctx* ctx, const trax_config* cfg, trax_device_config** out) {
    return guarded([&] {
        need(ctx && cfg && out, "null argument");
        validate_config(cfg);
        std::unique_ptr<trax_device_config> dc(new trax_device_config);
        dc->cm.reset(new ConfigMirror(*ctx->c, cfg->trip, cfg->geom, cfg->n_threads, cfg->d_eta_cut));
        GridConfig gc = cfg->grid;
        gc.nLayers = cfg->trip.maxLayer();
        dc->g = gc.geom();
        dc->n_trip = cfg->trip.size();
        *out = dc.release();
        return TRAX_OK;
    });
}

extern "C" void trax_device_config_destroy(trax_device_config* dc) { delete dc; }

extern "C" int trax_stage_begin(trax_ctx* ctx, uint32_t n_events, uint32_t n_hits, trax_stream* s) {
    return guarded([&] {
        need(ctx != nullptr, "null context");
        Sizes z(ctx, s);
        hip_check(launch_fill(reinterpret_cast<uint32_t*>(rec(ctx)) + kCountsShapeWords, 0u,
                              sizeof(DevCounts) / 4 - kCountsShapeWords, z.st),
                  "clear sizes");
        z.put(&DevCounts::n_events, n_events).put(&DevCounts::n_hits, n_hits).write();
        return TRAX_OK;
    });
}

extern "C" int trax_stage_check(trax_ctx* ctx, trax_stream* s) {
    return guarded([&] {
        need(ctx != nullptr, "null context");
        const hipStream_t st = pick_stream(ctx, s);
        uint32_t err = 0;
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        hip_check(hipMemcpy(&err, &rec(ctx)->err, sizeof(err), hipMemcpyDeviceToHost), "d2h errors");
        if (err & kErrCapacity) return fail(TRAX_ERR_CAPACITY, "an output exceeded its size");
        if (err & kErrCellRange) return fail(TRAX_ERR_INPUT, "a hit lies outside the z-phi grid");
        if (err & kErrIndexRange) return fail(TRAX_ERR_INPUT, "an index read from device memory is out of range");
        if (err) return fail(TRAX_ERR_INPUT, "device reported an input error");
        return TRAX_OK;
    });
}

extern "C" int trax_grid_build_zphi(trax_ctx* ctx, uint32_t n_events, uint32_t n_layers, uint32_t n_hits, uint32_t n_z,
                                    uint32_t n_phi, const uint32_t* d_event_offsets, const uint32_t* d_layer_nhits,
                                    const uint32_t* d_layer_offsets, trax_hit_cols in, trax_hit_cols_out out,
                                    float* d_zphi, uint32_t* d_grid, trax_stream* s) {
    return guarded([&] {
        need(ctx && d_zphi && d_grid, "null argument");
        need((reinterpret_cast<uintptr_t>(d_zphi) & 15u) == 0, "the (z, phi) column must be 16-byte aligned");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_events, n_events).put(&DevCounts::n_hits, n_hits).write();
        GridConfig gc;
        gc.nEvents = n_events; gc.nLayers = n_layers; gc.nSectorsZ = n_z; gc.nSectorsPhi = n_phi;
        const GridGeom g = gc.geom();
        HitColumns hin, hout;
        hin.x = const_cast<float*>(in.x); hin.y = const_cast<float*>(in.y); hin.z = const_cast<float*>(in.z);
        hin.layer = const_cast<uint32_t*>(in.layer); hin.det = const_cast<uint32_t*>(in.det_id);
        hin.id = const_cast<uint32_t*>(in.hit_id); hin.evt = const_cast<uint32_t*>(in.event);
        hout.x = out.x; hout.y = out.y; hout.z = out.z; hout.layer = out.layer; hout.det = out.det_id;
        hout.id = out.hit_id; hout.evt = out.event;
        hout.zphi = reinterpret_cast<float2*>(d_zphi);
        const LaunchCfg lc;
        hip_check(launch_grid_count(g, n_events * n_layers, d_event_offsets, d_layer_nhits, d_layer_offsets, in.x, in.y,
                                    in.z, d_grid, rec(ctx), lc, z.st),
                  "grid count");
        hip_check(launch_grid_store(g, n_events * n_layers, d_event_offsets, d_layer_nhits, d_layer_offsets, d_grid, hin,
                                    hout, n_hits, rec(ctx), lc, z.st),
                  "grid store");
        return TRAX_OK;
    });
}

// ---- pairs -----------------------------------------------------------------------

extern "C" int trax_pair_plan(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const uint32_t* d_grid,
                              uint32_t* d_seg_base, trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_grid && d_seg_base, "null argument");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_events, n_events).write();
        GridGeom g = dc->g;
        hip_check(launch_pair_plan(g, dc->cm->cols(), n_events * dc->n_trip, d_grid, d_seg_base, rec(ctx), z.st),
                  "pair plan");
        return TRAX_OK;
    });
}

extern "C" int trax_pair_count(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const uint32_t* d_grid,
                               const trax_grid_hits* hits, const uint32_t* d_seg_base, uint32_t n_slots,
                               uint32_t* d_slot_counts, const trax_tuning* tuning, trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_grid && d_seg_base && d_slot_counts, "null argument");
        check_hits(hits);
        Sizes z(ctx, s);
        const uint32_t nseg = n_events * dc->n_trip;
        z.put(&DevCounts::n_events, n_events).put(&DevCounts::n_segments, nseg).put(&DevCounts::n_pair_slots, n_slots)
            .write();
        const HitColumns hc = cols_of(hits);
        hip_check(launch_pair_count(dc->g, dc->cm->cols(), nseg, n_slots, d_grid, hc.x, hc.y, hc.z, hc.zphi, d_seg_base,
                                    d_slot_counts, rec(ctx), launch_cfg(tuning), z.st),
                  "pair count");
        return TRAX_OK;
    });
}

extern "C" int trax_pair_store(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const uint32_t* d_grid,
                               const trax_grid_hits* hits, const uint32_t* d_seg_base, uint32_t n_slots,
                               const uint32_t* d_slot_counts, uint32_t n_pairs, uint32_t* d_pairs,
                               uint32_t* d_pair_offsets, const trax_tuning* tuning, trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_grid && d_seg_base && d_slot_counts && d_pairs && d_pair_offsets, "null argument");
        check_hits(hits);
        Sizes z(ctx, s);
        const uint32_t nseg = n_events * dc->n_trip;
        z.put(&DevCounts::n_events, n_events).put(&DevCounts::n_segments, nseg).put(&DevCounts::n_pair_slots, n_slots)
            .put(&DevCounts::n_pairs, n_pairs).write();
        const HitColumns hc = cols_of(hits);
        hip_check(launch_pair_store(dc->g, dc->cm->cols(), nseg, n_slots, n_pairs, d_grid, hc.x, hc.y, hc.z, hc.zphi,
                                    d_seg_base, d_slot_counts, d_pairs, d_pair_offsets, rec(ctx), launch_cfg(tuning),
                                    z.st),
                  "pair store");
        return TRAX_OK;
    });
}

// ---- predictor ---------------------------------------------------------------------

extern "C" int trax_predict_count(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events,
                                  const uint32_t* d_grid, const trax_grid_hits* hits, uint32_t n_pairs,
                                  const uint32_t* d_pairs, uint32_t* d_counts, const trax_tuning* tuning,
                                  trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_grid && d_pairs && d_counts, "null argument");
        check_hits(hits);
        Sizes z(ctx, s);
        z.put(&DevCounts::n_events, n_events).put(&DevCounts::n_hits, hits->n_hits).put(&DevCounts::n_pairs, n_pairs)
            .write();
        const HitColumns hc = cols_of(hits);
        hip_check(launch_predict_count(dc->g, dc->cm->cols(), n_pairs, d_grid, hc.x, hc.y, hc.z, hc.zphi, hc.layer,
                                       hc.evt, d_pairs, d_counts, rec(ctx), launch_cfg(tuning), z.st),
                  "predict count");
        return TRAX_OK;
    });
}

extern "C" int trax_predict_store(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events,
                                  const uint32_t* d_grid, const trax_grid_hits* hits, uint32_t n_pairs,
                                  const uint32_t* d_pairs, const uint32_t* d_counts, uint32_t n_cands,
                                  uint32_t* d_cands, uint32_t* d_cand_offsets, const trax_tuning* tuning,
                                  trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_grid && d_pairs && d_counts && d_cands && d_cand_offsets, "null argument");
        check_hits(hits);
        Sizes z(ctx, s);
        z.put(&DevCounts::n_events, n_events).put(&DevCounts::n_hits, hits->n_hits).put(&DevCounts::n_pairs, n_pairs)
            .put(&DevCounts::n_cands, n_cands).write();
        // the offsets start zeroed (the Pairing's offsets, Pairings.h:9-59)
        hip_check(launch_fill(d_cand_offsets, 0u, (uint64_t)n_events * dc->n_trip + 1, z.st), "fill");
        const HitColumns hc = cols_of(hits);
        hip_check(launch_predict_store(dc->g, dc->cm->cols(), n_pairs, n_cands, d_grid, hc.x, hc.y, hc.z, hc.zphi,
                                       hc.layer, hc.evt, d_pairs, d_counts, d_cands, d_cand_offsets, rec(ctx),
                                       launch_cfg(tuning), z.st),
                  "predict store");
        return TRAX_OK;
    });
}

extern "C" int trax_offsets_fill(trax_ctx* ctx, uint32_t* d_offsets, uint32_t n_events, uint32_t n_triplets,
                                 trax_stream* s) {
    return guarded([&] {
        need(ctx && d_offsets, "null argument");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_events, n_events).write();
        hip_check(launch_forward_fill(d_offsets, rec(ctx), n_triplets, z.st), "offsets fill");
        return TRAX_OK;
    });
}

// ---- filter ------------------------------------------------------------------------

extern "C" int trax_filter_count(trax_ctx* ctx, const trax_device_config* dc, const trax_grid_hits* hits,
                                 uint32_t n_pairs, const uint32_t* d_pairs, uint32_t n_cands, const uint32_t* d_cands,
                                 uint64_t* d_bitmap, uint32_t* d_word_counts, const trax_tuning* tuning,
                                 trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_pairs && d_cands && d_bitmap && d_word_counts, "null argument");
        check_hits(hits);
        Sizes z(ctx, s);
        z.put(&DevCounts::n_hits, hits->n_hits).put(&DevCounts::n_pairs, n_pairs).put(&DevCounts::n_cands, n_cands)
            .write();
        const HitColumns hc = cols_of(hits);
        hip_check(launch_filter_count(dc->cm->cols(), n_cands, hc.x, hc.y, hc.z, hc.layer, d_pairs, d_cands, d_bitmap,
                                      d_word_counts, rec(ctx), launch_cfg(tuning), z.st),
                  "filter count");
        return TRAX_OK;
    });
}

extern "C" int trax_filter_store(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events,
                                 const trax_grid_hits* hits, uint32_t n_pairs, const uint32_t* d_pairs,
                                 uint32_t n_cands, const uint32_t* d_cands, const uint64_t* d_bitmap,
                                 const uint32_t* d_word_counts, uint32_t n_tracklets, uint32_t* d_h1, uint32_t* d_h2,
                                 uint32_t* d_h3, uint32_t* d_tracklet_offsets, const trax_tuning* tuning,
                                 trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_pairs && d_cands && d_bitmap && d_word_counts && d_h1 && d_h2 && d_h3 && d_tracklet_offsets,
             "null argument");
        check_hits(hits);
        Sizes z(ctx, s);
        z.put(&DevCounts::n_events, n_events).put(&DevCounts::n_hits, hits->n_hits).put(&DevCounts::n_pairs, n_pairs)
            .put(&DevCounts::n_cands, n_cands).put(&DevCounts::n_filter_words, (n_cands + 63) / 64)
            .put(&DevCounts::n_tracklets, n_tracklets).write();
        hip_check(launch_fill(d_tracklet_offsets, 0u, (uint64_t)n_events * dc->n_trip + 1, z.st), "fill");
        const TripletCols tc = dc->cm->cols();
        const HitColumns hc = cols_of(hits);
        hip_check(launch_filter_store(tc, n_cands, n_tracklets, hc.layer, hc.evt, d_pairs, d_cands, d_bitmap,
                                      d_word_counts, d_h1, d_h2, d_h3, d_tracklet_offsets, rec(ctx), launch_cfg(tuning),
                                      z.st),
                  "filter store");
        hip_check(launch_filter_offset_store(tc, n_tracklets, hc.layer, hc.evt, d_h1, d_tracklet_offsets, rec(ctx), z.st),
                  "filter offsets");
        return TRAX_OK;
    });
}

// ---- connectivity and circle fit ---------------------------------------------------------

extern "C" int trax_link_prepare(trax_ctx* ctx, const trax_device_config* dc, const trax_grid_hits* hits,
                                 uint32_t n_tracklets, const uint32_t* d_h1, const uint32_t* d_h2,
                                 const uint32_t* d_h3, const trax_link_bufs* b, const trax_tuning* tuning,
                                 trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_h1 && d_h2 && d_h3, "null argument");
        check_hits(hits);
        const LinkBufs lb = link_bufs(b);
        need(lb.eta && lb.hit_count && lb.hit_cursor && lb.csr && lb.conn, "link buffers: null eta, bucket or flag");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_hits, hits->n_hits).put(&DevCounts::n_tracklets, n_tracklets).write();
        const HitColumns hc = cols_of(hits);
        const ScanScratch& ss = ctx->c->scan_scratch(hits->n_hits + 1ull);
        hip_check(launch_link_prepare(dc->cm->cols(), hits->n_hits, n_tracklets, hc.x, hc.y, hc.z, d_h1, d_h2, d_h3, lb,
                                      rec(ctx), ss, launch_cfg(tuning), z.st),
                  "link prepare");
        return TRAX_OK;
    });
}

extern "C" int trax_link_count(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events,
                               const trax_grid_hits* hits, uint32_t n_tracklets, const uint32_t* d_h1,
                               const uint32_t* d_h2, const uint32_t* d_h3, const uint32_t* d_tracklet_offsets,
                               const trax_link_bufs* b, const trax_tuning* tuning, trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_h1 && d_h2 && d_h3 && d_tracklet_offsets, "null argument");
        check_hits(hits);
        const LinkBufs lb = link_bufs(b);
        need(lb.eta && lb.hit_count && lb.csr && lb.conn && lb.counts, "link buffers: null eta, bucket or count");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_events, n_events).put(&DevCounts::n_hits, hits->n_hits)
            .put(&DevCounts::n_tracklets, n_tracklets).write();
        const HitColumns hc = cols_of(hits);
        hip_check(launch_link_count(dc->cm->cols(), n_tracklets, hc.layer, hc.evt, d_h1, d_h2, d_h3, d_tracklet_offsets,
                                    lb, rec(ctx), launch_cfg(tuning), z.st),
                  "link count");
        return TRAX_OK;
    });
}

extern "C" int trax_link_store(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events,
                               const trax_grid_hits* hits, uint32_t n_tracklets, const uint32_t* d_h1,
                               const uint32_t* d_h2, const uint32_t* d_h3, const uint32_t* d_tracklet_offsets,
                               uint32_t n_conn, const trax_link_bufs* b, const trax_tuning* tuning, trax_stream* s) {
    return guarded([&] {
        need(ctx && dc && d_h1 && d_h2 && d_h3 && d_tracklet_offsets, "null argument");
        check_hits(hits);
        const LinkBufs lb = link_bufs(b);
        need(lb.eta && lb.hit_count && lb.csr && lb.counts && lb.basis && lb.foll && lb.pev,
             "link buffers: null eta, bucket, count or connection list");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_events, n_events).put(&DevCounts::n_hits, hits->n_hits)
            .put(&DevCounts::n_tracklets, n_tracklets).put(&DevCounts::n_conn, n_conn).write();
        const HitColumns hc = cols_of(hits);
        hip_check(launch_link_store(dc->cm->cols(), n_tracklets, n_conn, hc.layer, hc.evt, d_h1, d_h2, d_h3,
                                    d_tracklet_offsets, lb, rec(ctx), launch_cfg(tuning), z.st),
                  "link store");
        return TRAX_OK;
    });
}

extern "C" int trax_link_flags(trax_ctx* ctx, const trax_grid_hits* hits, uint32_t n_tracklets, uint32_t n_conn,
                               const uint32_t* d_h1, const trax_link_bufs* b, trax_stream* s) {
    return guarded([&] {
        need(ctx && d_h1, "null argument");
        check_hits(hits);
        const LinkBufs lb = link_bufs(b);
        need(lb.conn && lb.flag_a && lb.flag_b, "link buffers: null flag");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_hits, hits->n_hits).put(&DevCounts::n_tracklets, n_tracklets).put(&DevCounts::n_conn, n_conn)
            .write();
        hip_check(launch_link_flags(n_tracklets, hits->event, d_h1, lb, rec(ctx), LaunchCfg(), z.st), "link flags");
        return TRAX_OK;
    });
}

extern "C" int trax_link_compact(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_connectable,
                                 uint32_t n_nonconnectable, const trax_link_bufs* b, trax_stream* s) {
    return guarded([&] {
        need(ctx != nullptr, "null context");
        const LinkBufs lb = link_bufs(b);
        need(lb.flag_a && lb.flag_b && lb.connectable && lb.nonconnectable, "link buffers: null flag or list");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_tracklets, n_tracklets).put(&DevCounts::n_connectable, n_connectable)
            .put(&DevCounts::n_nonconnectable, n_nonconnectable).write();
        hip_check(launch_link_compact(n_tracklets, lb, rec(ctx), LaunchCfg(), z.st), "link compact");
        return TRAX_OK;
    });
}

extern "C" int trax_circle_fit(trax_ctx* ctx, const trax_grid_hits* hits, uint32_t n_tracklets, const uint32_t* d_h1,
                               const uint32_t* d_h2, const uint32_t* d_h3, const uint32_t* d_connectable,
                               uint32_t n_connectable, float* d_pt, trax_stream* s) {
    return guarded([&] {
        need(ctx && d_h1 && d_h2 && d_h3 && d_connectable && d_pt, "null argument");
        check_hits(hits);
        Sizes z(ctx, s);
        z.put(&DevCounts::n_hits, hits->n_hits).put(&DevCounts::n_tracklets, n_tracklets)
            .put(&DevCounts::n_connectable, n_connectable).write();
        LinkBufs lb;
        lb.connectable = const_cast<uint32_t*>(d_connectable);
        lb.pt = d_pt;
        hip_check(launch_circle_fit(n_tracklets, hits->x, hits->y, d_h1, d_h2, d_h3, lb, rec(ctx), LaunchCfg(), z.st),
                  "circle fit");
        return TRAX_OK;
    });
}

// ---- cellular automaton ----------------------------------------------------------------

extern "C" int trax_ca_init(trax_ctx* ctx, uint32_t n_tracklets, const trax_ca_bufs* b, trax_stream* s) {
    return guarded([&] {
        need(ctx != nullptr, "null context");
        const CABufs cb = ca_bufs(b);
        need(cb.st_a && cb.st_b && cb.flags, "CA buffers: null state or flag");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_tracklets, n_tracklets).write();
        hip_check(launch_ca_init(n_tracklets, cb, rec(ctx), z.st), "ca init");
        return TRAX_OK;
    });
}

extern "C" int trax_ca_iterate(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_conn, const uint32_t* d_basis,
                               const uint32_t* d_followers, const trax_ca_bufs* b, uint32_t iteration,
                               trax_stream* s) {
    return guarded([&] {
        need(ctx && d_basis && d_followers, "null argument");
        const CABufs cb = ca_bufs(b);
        need(cb.st_a && cb.st_b && cb.flags, "CA buffers: null state or flag");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_tracklets, n_tracklets).put(&DevCounts::n_conn, n_conn).write();
        hip_check(launch_ca_iterations(n_conn, d_basis, d_followers, cb, rec(ctx), iteration, 1, LaunchCfg(), z.st),
                  "ca iteration");
        return TRAX_OK;
    });
}

extern "C" int trax_ca_basis_count(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_conn, const uint32_t* d_basis,
                                   const uint32_t* d_followers, const trax_ca_bufs* b, trax_stream* s) {
    return guarded([&] {
        need(ctx && d_basis && d_followers, "null argument");
        const CABufs cb = ca_bufs(b);
        need(cb.st_a && cb.fbc && cb.cursor, "CA buffers: null state, count or cursor");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_tracklets, n_tracklets).put(&DevCounts::n_conn, n_conn).write();
        hip_check(launch_ca_basis_count(n_tracklets, n_conn, d_basis, d_followers, cb, rec(ctx), LaunchCfg(), z.st),
                  "ca basis count");
        return TRAX_OK;
    });
}

extern "C" int trax_ca_basis_store(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_conn, const uint32_t* d_basis,
                                   const uint32_t* d_followers, const float* d_pt, uint32_t n_basis,
                                   const trax_ca_bufs* b, trax_stream* s) {
    return guarded([&] {
        need(ctx && d_basis && d_followers && d_pt, "null argument");
        const CABufs cb = ca_bufs(b);
        need(cb.st_a && cb.fbc && cb.cursor && cb.bidx && cb.bdiff && cb.best && cb.isbest && cb.handler,
             "CA buffers: null basis list, best basis or track offsets");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_tracklets, n_tracklets).put(&DevCounts::n_conn, n_conn).put(&DevCounts::n_basis, n_basis)
            .write();
        hip_check(launch_ca_basis_store(n_tracklets, n_conn, n_basis, d_basis, d_followers, d_pt, cb, rec(ctx),
                                        LaunchCfg(), z.st),
                  "ca basis store");
        return TRAX_OK;
    });
}

extern "C" int trax_ca_tracks(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_track_entries, const trax_ca_bufs* b,
                              trax_stream* s) {
    return guarded([&] {
        need(ctx != nullptr, "null context");
        const CABufs cb = ca_bufs(b);
        need(cb.best && cb.handler && cb.tracks, "CA buffers: null best basis, track offsets or tracks");
        Sizes z(ctx, s);
        z.put(&DevCounts::n_tracklets, n_tracklets).put(&DevCounts::n_track_entries, n_track_entries).write();
        hip_check(launch_ca_tracks(n_tracklets, n_track_entries, cb, rec(ctx), ScanScratch(), LaunchCfg(), z.st, 0),
                  "ca tracks");
        return TRAX_OK;
    });
}
