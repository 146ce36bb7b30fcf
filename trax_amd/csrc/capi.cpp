// C ABI of the gfx950 triplet-seeding library (declared in include/trax_hip.h).
#include "capi_internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <set>
#include <stdexcept>

using namespace trax;
using namespace trax_capi;

extern "C" const char* trax_last_error(void) { return g_last_error.c_str(); }

// ---- scratch of the no-allocation prefix sum --------------------------------
// trax_prefix_sum_noalloc keeps the scan's status words (the tile sums of the
// reduce-then-scan form, or the tile status and tile counter of the look-back
// form) in this file-scope device array instead of memory it would allocate.
// Every device holds its own copy until the process ends.  In host code the
// array's name is not a device address: the address of the current device's
// copy comes from hipGetSymbolAddress, taken once per device in
// trax_ctx_create, and only that address is passed to kernels.
static constexpr uint32_t kNoAllocScanMax = 1u << 28;  // elements per call, at most (1 GiB of u32)
static constexpr uint32_t kNoAllocScanWords = (uint32_t)(kNoAllocScanMax / kScanTileElements) + 3;  // 256 KiB
static __device__ __align__(256) uint64_t scratch[kNoAllocScanWords];
namespace {
constexpr int kMaxDevices = 64;
std::atomic<uint64_t*> g_scratch_addr[kMaxDevices];

// the address of the device's copy of `scratch` (device current); once per device
int noalloc_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(TRAX_ERR_ARG, "device index above the library's table");
    if (g_scratch_addr[device].load()) return TRAX_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(TRAX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_addr[device].store(static_cast<uint64_t*>(p));
    return TRAX_OK;
}
}  // namespace

// ---- context, streams, memory ---------------------------------------------

extern "C" int trax_ctx_create(int device, trax_ctx** out) {
    return guarded([&] {
        if (!out) return fail(TRAX_ERR_ARG, "null output");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0)
            return fail(TRAX_ERR_HIP, "no such HIP device");
        std::unique_ptr<trax_ctx> c(new trax_ctx);
        c->c.reset(new Context(device));  // makes the device current
        const int rc = noalloc_scratch_setup(device);
        if (rc != TRAX_OK) return rc;
        *out = c.release();
        return TRAX_OK;
    });
}

extern "C" void trax_ctx_destroy(trax_ctx* ctx) { delete ctx; }

extern "C" int trax_device_info_get(trax_ctx* ctx, trax_device_info* out) {
    return guarded([&] {
        hipDeviceProp_t p;
        hip_check(hipGetDeviceProperties(&p, ctx->c->device()), "hipGetDeviceProperties");
        std::memset(out, 0, sizeof(*out));
        std::strncpy(out->name, p.gcnArchName, sizeof(out->name) - 1);
        out->cu_count = (uint32_t)p.multiProcessorCount;
        out->lds_bytes_per_block = (uint32_t)p.sharedMemPerBlock;
        out->max_block = (uint32_t)p.maxThreadsPerBlock;
        out->wave_size = (uint32_t)p.warpSize;
        out->global_mem_bytes = (uint64_t)p.totalGlobalMem;
        return TRAX_OK;
    });
}

extern "C" int trax_stream_create(trax_ctx* ctx, trax_stream** out) {
    return guarded([&] {
        hip_check(hipSetDevice(ctx->c->device()), "hipSetDevice");
        trax_stream* s = new trax_stream;
        hip_check(hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking), "hipStreamCreate");
        *out = s;
        return TRAX_OK;
    });
}

extern "C" int trax_stream_destroy(trax_stream* s) {
    return guarded([&] {
        if (s) {
            hip_check(hipStreamDestroy(s->s), "hipStreamDestroy");
            delete s;
        }
        return TRAX_OK;
    });
}

static hipStream_t pick(trax_ctx* ctx, trax_stream* s) { return pick_stream(ctx, s); }

extern "C" int trax_stream_sync(trax_ctx* ctx, trax_stream* s) {
    return guarded([&] {
        hip_check(hipStreamSynchronize(pick(ctx, s)), "hipStreamSynchronize");
        return TRAX_OK;
    });
}

extern "C" int trax_timer_create(trax_ctx* ctx, trax_timer** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("trax_timer_create: out is NULL");
        hip_check(hipSetDevice(ctx->c->device()), "hipSetDevice");
        std::unique_ptr<trax_timer> t(new trax_timer);
        hip_check(hipEventCreate(&t->start), "hipEventCreate");
        const hipError_t e = hipEventCreate(&t->stop);
        if (e != hipSuccess) {
            (void)hipEventDestroy(t->start);
            hip_check(e, "hipEventCreate");
        }
        *out = t.release();
        return TRAX_OK;
    });
}

extern "C" int trax_timer_destroy(trax_timer* t) {
    return guarded([&] {
        if (t) {
            hip_check(hipEventDestroy(t->start), "hipEventDestroy");
            hip_check(hipEventDestroy(t->stop), "hipEventDestroy");
            delete t;
        }
        return TRAX_OK;
    });
}

extern "C" int trax_timer_start(trax_ctx* ctx, trax_timer* t, trax_stream* s) {
    return guarded([&] {
        if (!t) throw std::invalid_argument("trax_timer_start: timer is NULL");
        hip_check(hipEventRecord(t->start, pick(ctx, s)), "hipEventRecord");
        return TRAX_OK;
    });
}

extern "C" int trax_timer_stop(trax_ctx* ctx, trax_timer* t, trax_stream* s, float* ms) {
    return guarded([&] {
        if (!t || !ms) throw std::invalid_argument("trax_timer_stop: timer or ms is NULL");
        hip_check(hipEventRecord(t->stop, pick(ctx, s)), "hipEventRecord");
        hip_check(hipEventSynchronize(t->stop), "hipEventSynchronize");
        hip_check(hipEventElapsedTime(ms, t->start, t->stop), "hipEventElapsedTime");
        return TRAX_OK;
    });
}

extern "C" int trax_buf_alloc(trax_ctx* ctx, uint64_t bytes, void** dptr) {
    return guarded([&] {
        hip_check(hipSetDevice(ctx->c->device()), "hipSetDevice");
        hip_check(hipMalloc(dptr, bytes ? bytes : 4), "hipMalloc");
        return TRAX_OK;
    });
}

extern "C" int trax_buf_free(trax_ctx* ctx, void* dptr) {
    return guarded([&] {
        (void)ctx;
        hip_check(hipFree(dptr), "hipFree");
        return TRAX_OK;
    });
}

extern "C" int trax_buf_fill_u32(trax_ctx* ctx, void* dptr, uint32_t value, uint64_t count, trax_stream* s) {
    return guarded([&] {
        hip_check(launch_fill(static_cast<uint32_t*>(dptr), value, count, pick(ctx, s)), "fill");
        return TRAX_OK;
    });
}

extern "C" int trax_buf_copy_u32(trax_ctx* ctx, void* dst, const void* src, uint64_t count, trax_stream* s) {
    return guarded([&] {
        hip_check(launch_copy(static_cast<uint32_t*>(dst), static_cast<const uint32_t*>(src), count, pick(ctx, s)),
                  "copy");
        return TRAX_OK;
    });
}

extern "C" int trax_h2d(trax_ctx* ctx, void* dst, const void* src, uint64_t bytes, trax_stream* s) {
    return guarded([&] {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, pick(ctx, s)), "h2d");
        return TRAX_OK;
    });
}

extern "C" int trax_d2h(trax_ctx* ctx, void* dst, const void* src, uint64_t bytes, trax_stream* s) {
    return guarded([&] {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, pick(ctx, s)), "d2h");
        return TRAX_OK;
    });
}

// ---- kernel launchers -------------------------------------------------------

extern "C" int trax_prefix_sum(trax_ctx* ctx, uint32_t* d_data, uint32_t n, uint32_t* d_total, trax_stream* s) {
    return guarded([&] {
        hip_check(launch_scan(d_data, nullptr, n, n, d_total, ctx->c->scan_scratch(n), pick(ctx, s)), "scan");
        return TRAX_OK;
    });
}

extern "C" int trax_prefix_sum_noalloc(trax_ctx* ctx, uint32_t* d_data, uint32_t n, uint32_t* d_total,
                                       trax_stream* s) {
    if (!ctx) return fail(TRAX_ERR_ARG, "null context");
    if (n > kNoAllocScanMax)
        return fail(TRAX_ERR_ARG, "trax_prefix_sum_noalloc takes at most 2^28 elements (its scratch holds " +
                                      std::to_string(kNoAllocScanWords) + " status words)");
    const int device = ctx->c->device();
    uint64_t* addr = (device >= 0 && device < kMaxDevices) ? g_scratch_addr[device].load() : nullptr;
    if (!addr) return fail(TRAX_ERR_STATE, "no scratch address for the context's device");
    ScanScratch ss;
    ss.status = addr;
    ss.max_tiles = kNoAllocScanWords - 1;  // one more word past the tiles, as the context's scratch has
    const hipError_t e = launch_scan(d_data, nullptr, n, n, d_total, ss, pick(ctx, s));
    if (e != hipSuccess) return fail(TRAX_ERR_HIP, std::string("scan: ") + hipGetErrorString(e));
    return TRAX_OK;
}

extern "C" int trax_grid_build(trax_ctx* ctx, uint32_t n_events, uint32_t n_layers, uint32_t n_hits, uint32_t n_z,
                               uint32_t n_phi, const uint32_t* d_event_offsets, const uint32_t* d_layer_nhits,
                               const uint32_t* d_layer_offsets, trax_hit_cols in, trax_hit_cols_out out,
                               uint32_t* d_grid, trax_stream* s) {
    return guarded([&] {
        Context& c = *ctx->c;
        hipStream_t prev = c.stream();
        c.set_stream(pick(ctx, s));
        GridConfig gc;
        gc.nEvents = n_events; gc.nLayers = n_layers; gc.nSectorsZ = n_z; gc.nSectorsPhi = n_phi;
        c.reset_counts();
        c.set_batch_shape(n_events, n_hits);
        HitColumns hin, hout;
        hin.x = const_cast<float*>(in.x); hin.y = const_cast<float*>(in.y); hin.z = const_cast<float*>(in.z);
        hin.layer = const_cast<uint32_t*>(in.layer); hin.det = const_cast<uint32_t*>(in.det_id);
        hin.id = const_cast<uint32_t*>(in.hit_id); hin.evt = const_cast<uint32_t*>(in.event);
        hout.x = out.x; hout.y = out.y; hout.z = out.z; hout.layer = out.layer; hout.det = out.det_id;
        hout.id = out.hit_id; hout.evt = out.event;
        const GridGeom g = gc.geom();
        hip_check(launch_grid_count(g, n_events * n_layers, d_event_offsets, d_layer_nhits, d_layer_offsets, in.x, in.y,
                                    in.z, d_grid, c.counts(), c.launch(), c.stream()),
                  "grid count");
        hip_check(launch_grid_store(g, n_events * n_layers, d_event_offsets, d_layer_nhits, d_layer_offsets, d_grid, hin,
                                    hout, n_hits, c.counts(), c.launch(), c.stream()),
                  "grid store");
        c.finish_default_queue();
        c.check_errors();
        c.set_stream(prev);
        return TRAX_OK;
    });
}

// ---- hits and configuration -----------------------------------------------------

extern "C" int trax_hits_create(uint32_t max_layer, trax_hits** out) {
    return guarded([&] {
        if (max_layer == 0) return fail(TRAX_ERR_ARG, "max_layer must be positive");
        trax_hits* h = new trax_hits;
        h->max_layer = max_layer;
        *out = h;
        return TRAX_OK;
    });
}

extern "C" void trax_hits_destroy(trax_hits* h) { delete h; }

static void add_event(trax_hits* h, const trax_config* cfg, const std::vector<PHit>& ev, int max_tracks,
                      int only_tracks);

extern "C" int trax_hits_add_event(trax_hits* h, const trax_config* cfg, uint32_t n, const float* x, const float* y,
                                   const float* z, const uint32_t* layer, const uint32_t* det_id,
                                   const uint32_t* track_id, const float* track_pt, int max_tracks, int only_tracks) {
    return guarded([&] {
        if (cfg->trip.maxLayer() != h->max_layer)
            return fail(TRAX_ERR_ARG, "configuration's highest layer differs from the collection's");
        std::vector<PHit> ev(n);
        for (uint32_t i = 0; i < n; ++i) ev[i] = PHit{x[i], y[i], z[i], layer[i], det_id[i], track_id[i], track_pt[i]};
        add_event(h, cfg, ev, max_tracks, only_tracks);
        return TRAX_OK;
    });
}

// HitCollection::addEvent on the collection's growing supplements
static void add_event(trax_hits* h, const trax_config* cfg, const std::vector<PHit>& ev, int max_tracks,
                      int only_tracks) {
    {
        const uint32_t e = h->n_events, L = h->max_layer;
        // the supplements grow by one event in place (moved in and out, not copied)
        EventSupplement es(0);
        LayerSupplement ls(L, 0);
        es.nEvents = e + 1;
        es.nHits = std::move(h->evt_nhits);
        es.offset = std::move(h->evt_offset);
        ls.nHits = std::move(h->lay_nhits);
        ls.offset = std::move(h->lay_offset);
        es.nHits.resize(e + 1, 0);
        es.offset.resize(e + 1, 0);
        ls.nHits.resize((size_t)(e + 1) * L, 0);
        ls.offset.resize((size_t)(e + 1) * L, 0);
        try {
            h->truth.push_back(h->hits.addEvent(ev, es, e, ls, cfg->trip, max_tracks, only_tracks != 0));
        } catch (...) {
            es.nHits.resize(e); es.offset.resize(e);
            ls.nHits.resize((size_t)e * L); ls.offset.resize((size_t)e * L);
            h->evt_nhits = std::move(es.nHits); h->evt_offset = std::move(es.offset);
            h->lay_nhits = std::move(ls.nHits); h->lay_offset = std::move(ls.offset);
            throw;
        }
        h->evt_nhits = std::move(es.nHits); h->evt_offset = std::move(es.offset);
        h->lay_nhits = std::move(ls.nHits); h->lay_offset = std::move(ls.offset);
        h->n_events = e + 1;
    }
}

extern "C" int trax_hits_add_pb(trax_hits* h, const trax_config* cfg, const void* data, uint64_t size, int max_events,
                                int max_tracks, int only_tracks, uint32_t* n_added) {
    return guarded([&] {
        if (!h || !cfg || (!data && size)) return fail(TRAX_ERR_ARG, "null argument");
        if (cfg->trip.maxLayer() != h->max_layer)
            return fail(TRAX_ERR_ARG, "configuration's highest layer differs from the collection's");
        const std::vector<PEventData> evs = decodeEventContainer(static_cast<const uint8_t*>(data), size);
        uint32_t added = 0;
        for (const PEventData& ev : evs) {
            if (max_events >= 0 && added >= (uint32_t)max_events) break;
            add_event(h, cfg, ev.hits, max_tracks, only_tracks);
            ++added;
        }
        if (n_added) *n_added = added;
        return TRAX_OK;
    });
}

extern "C" int trax_hits_add_pb_store(trax_hits* h, const trax_config* cfg, const void* data, uint64_t size,
                                      int skip_events, int max_events, int max_tracks, int only_tracks,
                                      uint32_t* n_added, uint32_t* n_stored) {
    return guarded([&] {
        if (!h || !cfg || (!data && size)) return fail(TRAX_ERR_ARG, "null argument");
        if (cfg->trip.maxLayer() != h->max_layer)
            return fail(TRAX_ERR_ARG, "configuration's highest layer differs from the collection's");
        const std::vector<PEventData> evs =
            decodeEventStore(static_cast<const uint8_t*>(data), size, skip_events > 0 ? (uint32_t)skip_events : 0u,
                             max_events >= 0 ? (uint32_t)max_events : 0xffffffffu, n_stored);
        for (const PEventData& ev : evs) add_event(h, cfg, ev.hits, max_tracks, only_tracks);
        if (n_added) *n_added = (uint32_t)evs.size();
        return TRAX_OK;
    });
}

extern "C" uint32_t trax_hits_count(const trax_hits* h) { return h->hits.size(); }
extern "C" uint32_t trax_hits_events(const trax_hits* h) { return h->n_events; }

extern "C" int trax_hits_column(const trax_hits* h, const char* name, const void** data, uint64_t* count) {
    const std::string k(name);
    const HitCollection& c = h->hits;
    auto ret = [&](const void* p, size_t n) { *data = p; *count = n; return TRAX_OK; };
    if (k == "x") return ret(c.x.data(), c.x.size());
    if (k == "y") return ret(c.y.data(), c.y.size());
    if (k == "z") return ret(c.z.data(), c.z.size());
    if (k == "layer") return ret(c.layer.data(), c.layer.size());
    if (k == "det_id") return ret(c.detId.data(), c.detId.size());
    if (k == "hit_id") return ret(c.hitId.data(), c.hitId.size());
    if (k == "event") return ret(c.event.data(), c.event.size());
    if (k == "event_nhits") return ret(h->evt_nhits.data(), h->evt_nhits.size());
    if (k == "event_offset") return ret(h->evt_offset.data(), h->evt_offset.size());
    if (k == "layer_nhits") return ret(h->lay_nhits.data(), h->lay_nhits.size());
    if (k == "layer_offset") return ret(h->lay_offset.data(), h->lay_offset.size());
    return fail(TRAX_ERR_ARG, "unknown column " + k);
}

extern "C" int trax_config_create(uint32_t n_z, uint32_t n_phi, float min_pt, trax_config** out) {
    return guarded([&] {
        if (!n_z || !n_phi) return fail(TRAX_ERR_ARG, "grid needs at least one cell");
        trax_config* c = new trax_config;
        c->grid.nSectorsZ = n_z;
        c->grid.nSectorsPhi = n_phi;
        c->trip.minPt_ = min_pt;
        *out = c;
        return TRAX_OK;
    });
}

extern "C" void trax_config_destroy(trax_config* c) { delete c; }

extern "C" int trax_config_add_triplet(trax_config* c, uint32_t layer1, uint32_t layer2, uint32_t layer3,
                                       float d_theta_cut, float sigma_z, float d_phi_cut, float sigma_phi,
                                       float tip_cut, uint32_t pair_spread_z, float z0, uint32_t pair_spread_phi,
                                       uint32_t n_candidates) {
    if (!layer1 || !layer2 || !layer3) return fail(TRAX_ERR_ARG, "layers are numbered from 1");
    c->trip.addWithValue(layer1, layer2, layer3, d_theta_cut, sigma_z, d_phi_cut, sigma_phi, tip_cut, pair_spread_z,
                         z0, pair_spread_phi, n_candidates);
    return TRAX_OK;
}

extern "C" int trax_config_set_geometry(trax_config* c, uint32_t n, const uint32_t* layer, const float* min_radius,
                                        const float* max_radius) {
    c->geom = GeometrySupplement();
    for (uint32_t i = 0; i < n; ++i) c->geom.addWithValue(layer[i], min_radius[i], max_radius[i]);
    return TRAX_OK;
}

extern "C" int trax_config_set_exec(trax_config* c, uint32_t n_threads, float d_eta_cut) {
    if (n_threads == 0) return fail(TRAX_ERR_ARG, "n_threads must be positive");
    c->n_threads = n_threads;
    c->d_eta_cut = d_eta_cut;
    return TRAX_OK;
}

extern "C" float trax_config_min_radius_curvature(const trax_config* c) { return c->trip.minRadiusCurvature(); }

extern "C" int trax_config_validate(const trax_config* c) {
    if (!c) return fail(TRAX_ERR_ARG, "null configuration");
    return guarded([&] {
        validate_config(c);
        return TRAX_OK;
    });
}

void trax_capi::validate_config(const trax_config* cfg) {
    const uint32_t L = cfg->trip.maxLayer();
    if (cfg->trip.size() == 0) throw InputError("no layer triplet configured");
    // the fused predictor keeps a pair's layer triplet in 8 bits
    if (cfg->trip.size() > 255) throw InputError("more than 255 layer triplets");
    if (cfg->geom.size() < L) throw InputError("geometry supplement has fewer layers than the triplets use");
    for (uint32_t i = 0; i < cfg->trip.size(); ++i) {
        if (cfg->trip.layer1[i] == 0 || cfg->trip.layer2[i] == 0 || cfg->trip.layer3[i] == 0)
            throw InputError("layers are numbered from 1");
        if (cfg->trip.layer1[i] > L || cfg->trip.layer2[i] > L) throw InputError("bad layer triplet");
    }
}

// ---- plain per-event-group path ---------------------------------------------------

// collect = false keeps only the downloads the reference driver itself makes
// (grid and hits after the build, tracklets after the filter, the CA outputs)
static int run_plain_impl(trax_ctx* ctx, const trax_config* cfg, const trax_hits* h, int stop_after, bool collect,
                          trax_result** out) {
    return guarded([&] {
        validate_config(cfg);
        Context& c = *ctx->c;
        hip_check(hipSetDevice(c.device()), "hipSetDevice");
        auto* r = new trax_result;
        std::unique_ptr<trax_result> guard(r);
        auto D = [&](const UVec& v) { return collect ? v.download(c) : std::vector<uint32_t>(); };
        const uint32_t E = h->n_events, L = h->max_layer, T = cfg->trip.size(), N = h->hits.size();
        const uint32_t nt = cfg->n_threads;
        auto t0 = std::chrono::steady_clock::now();
        auto lap = [&](const char* k) {
            auto t1 = std::chrono::steady_clock::now();
            r->timings[k] = std::chrono::duration<double, std::milli>(t1 - t0).count();
            t0 = t1;
        };
        auto want = [&](int s) { return stop_after == 0 || stop_after >= s; };
        c.reset_counts();
        GridConfig gc = cfg->grid;
        gc.nEvents = E;
        gc.nLayers = L;
        Grid grid(c, gc);
        HitCollection hits = h->hits;
        EventSupplement es(E);
        es.nHits = h->evt_nhits; es.offset = h->evt_offset;
        LayerSupplement ls(L, E);
        ls.nHits = h->lay_nhits; ls.offset = h->lay_offset;
        hits.transfer.initBuffers(c, hits);  // hits.transfer.initBuffers / toDevice (TripletBuilder.cpp:297-298)
        hits.transfer.toDevice(c, hits);
        const GeometrySupplement& geom = cfg->geom;
        const TripletConfigurations& trip = cfg->trip;
        GridBuilder(c).run(hits, nt, es, ls, grid);
        lap("grid");
        r->put("grid", grid.host);
        r->putf("hits.x", hits.x); r->putf("hits.y", hits.y); r->putf("hits.z", hits.z);
        r->put("hits.layer", hits.layer); r->put("hits.det_id", hits.detId);
        r->put("hits.hit_id", hits.hitId); r->put("hits.event", hits.event);
        auto diag = [&] {
            std::vector<uint32_t> d(4);
            d[0] = c.download_count(&c.counts()->diag_pair_short);
            d[1] = c.download_count(&c.counts()->diag_pred_short);
            d[2] = c.download_count(&c.counts()->diag_pred_trunc);
            d[3] = c.download_count(&c.counts()->diag_wrap_overlap);
            r->put("diag", d);
        };
        if (!want(2)) { diag(); *out = guard.release(); return TRAX_OK; }
        std::unique_ptr<Pairing> pairs(PairGeneratorBeamspot(c).run(hits, geom, nt, trip, grid));
        r->put("pairs", D(pairs->pairing));
        r->put("pair_offsets", D(pairs->pairingOffsets));
        lap("pairs");
        if (!want(3)) { diag(); *out = guard.release(); return TRAX_OK; }
        std::unique_ptr<Pairing> cands(TripletThetaPhiPredictor(c).run(hits, DetectorGeometry(), geom, Dictionary(),
                                                                      (int)nt, trip, grid, *pairs));
        r->put("cands", D(cands->pairing));
        r->put("cand_offsets", D(cands->pairingOffsets));
        lap("predict");
        if (!want(4)) { diag(); *out = guard.release(); return TRAX_OK; }
        std::unique_ptr<TrackletCollection> tr(TripletThetaPhiFilter(c).run(hits, grid, *pairs, *cands, (int)nt, trip));
        pairs.reset();  // TripletBuilder.cpp:357-358
        cands.reset();
        r->put("tracklets.h1", tr->h1); r->put("tracklets.h2", tr->h2); r->put("tracklets.h3", tr->h3);
        r->put("tracklet_offsets", D(tr->trackletOffsets));
        lap("filter");
        if (collect) diag();
        if (!want(5)) { *out = guard.release(); return TRAX_OK; }
        // TripletBuilder.cpp:395-469: connectivity, circle fit, CA; the caller owns what run returns
        ConnectivityDiag cdiag;
        UVec *basis_p, *foll_p, *connectable_p, *nonconnectable_p;
        std::tie(basis_p, foll_p, connectable_p, nonconnectable_p) =
            TripletConnectivityTight(c).run(hits, *tr, trip, cfg->d_eta_cut, nt, false, &cdiag);
        std::unique_ptr<UVec> basis(basis_p), foll(foll_p), connectable(connectable_p),
            nonconnectable(nonconnectable_p);
        if (cdiag.eta && collect) r->putf("eta", cdiag.eta->download(c));
        else r->putf("eta", {});
        auto dl = [&](const UVec* v, bool always = false) {
            return (v && (collect || always)) ? v->download(c) : std::vector<uint32_t>();
        };
        r->put("conn.basis", dl(basis.get()));
        r->put("conn.followers", dl(foll.get()));
        r->put("conn.event", dl(cdiag.event.get()));
        r->put("conn.connectable", dl(connectable.get()));
        r->put("conn.nonconnectable", dl(nonconnectable.get()));
        lap("connectivity");
        if (!basis) { *out = guard.release(); return TRAX_OK; }  // TripletBuilder.cpp:414-422
        std::unique_ptr<FVec> pt(TrackletCircleFitter(c).run(hits, *tr, *connectable, nt));
        lap("circle");
        CADiag adiag;
        const UVec *tracks_p, *track_off_p;
        std::tie(tracks_p, track_off_p) = CellularAutomaton(c).run(*basis, *foll, *pt, *connectable, nt, false, &adiag);
        std::unique_ptr<const UVec> tracks(tracks_p), track_off(track_off_p);
        r->putf("pt", pt->download(c));  // TripletBuilder.cpp:474-481
        if (collect) r->put("ca.states", adiag.states->download(c));
        if (collect) r->put("ca.best_basis", adiag.bestBasis->download(c));
        r->put("tracks", tracks->download(c));
        r->put("track_offsets", track_off->download(c));
        r->put("ca.iterations", {adiag.iterations});
        lap("ca");
        *out = guard.release();
        return TRAX_OK;
    });
}

extern "C" int trax_run_plain(trax_ctx* ctx, const trax_config* cfg, const trax_hits* h, int stop_after,
                              trax_result** out) {
    return run_plain_impl(ctx, cfg, h, stop_after, true, out);
}

extern "C" int trax_run_plain_bench(trax_ctx* ctx, const trax_config* cfg, const trax_hits* h, double* ms) {
    trax_result* r = nullptr;
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = run_plain_impl(ctx, cfg, h, 0, false, &r);
    const auto t1 = std::chrono::steady_clock::now();
    if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    delete r;
    return rc;
}

// ---- batched pipeline -------------------------------------------------------------

const char* const trax_capi::kPhases[kNumStages] = {
    "grid_count", "grid_store",   "pair_plan",   "pair_count",   "pair_scan",    "pair_store", "predict",    "filter_count",
    "filter_scan", "filter_store", "link_prepare", "link_count", "link_store", "circle_fit", "ca_forward", "ca_backward"};
const trax_capi::StageRange trax_capi::kStageRanges[6] = {{"grid", 0, 1},   {"pairs", 2, 5},
                                                          {"predict", 6, 6}, {"filter", 7, 9},
                                                          {"connectivity", 10, 13}, {"ca", 14, 15}};

void trax_batch::enqueue_fine(hipStream_t s) {
    FineGeom fg;
    hip_check(launch_fine_grid(cfg.grid.geom(), E * L, caps.max_hits, grid, hout.zphi, cnt, fine, lc, fg, s),
              "fine grid");
}

void trax_batch::enqueue_all(bool phase_events) {
    for (int k = 0; k < kNumStages; ++k) {
        if (phase_events) hip_check(hipEventRecord(ev[k], st), "event");
        if (k == kPredictStage) hip_check(hipStreamWaitEvent(st, ev_fine, 0), "wait fine grid");
        enqueue_stage(k);
        if (k == kGridStage) {  // fork: the fine walk structure beside the pair stages
            hip_check(hipEventRecord(ev_grid, st), "event");
            hip_check(hipStreamWaitEvent(side(), ev_grid, 0), "wait grid");
            enqueue_fine(side());
            hip_check(hipEventRecord(ev_fine, side()), "event");
        }
    }
    if (phase_events) hip_check(hipEventRecord(ev[kNumStages], st), "event");
}

bool trax_batch::capture_whole() {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
    bool ok = true;
    try {
        enqueue_all(false);  // the runtime does not time events recorded inside a graph
    } catch (const std::exception&) {
        ok = false;
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (ok && e == hipSuccess && g && hipGraphInstantiate(&whole_exec, g, nullptr, nullptr, 0) == hipSuccess) {
        whole_graph = g;
        return true;
    }
    if (g) (void)hipGraphDestroy(g);
    whole_exec = nullptr;
    (void)hipGetLastError();  // a failed capture leaves no error behind for later calls
    (void)hipStreamSynchronize(st);
    (void)hipStreamSynchronize(side());
    return false;
}

void trax_batch::single_stream() {
    if (!st2 || captured) return;
    (void)hipStreamSynchronize(st2);
    (void)hipStreamDestroy(st2);
    st2 = nullptr;
}

void trax_batch::launch() {
    if (caps.use_graph && !captured) {
        whole = caps.use_graph == 2 && capture_whole();
        if (!whole) {
            for (int k = 0; k < kNumStages; ++k) {
                hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
                enqueue_stage(k);
                hip_check(hipStreamEndCapture(st, &graphs[k]), "end capture");
                hip_check(hipGraphInstantiate(&execs[k], graphs[k], nullptr, nullptr, 0), "instantiate");
            }
            hip_check(hipStreamBeginCapture(side(), hipStreamCaptureModeThreadLocal), "begin capture");
            enqueue_fine(side());
            hip_check(hipStreamEndCapture(side(), &fine_graph), "end capture");
            hip_check(hipGraphInstantiate(&fine_exec, fine_graph, nullptr, nullptr, 0), "instantiate");
        }
        captured = true;
    }
    if (!caps.use_graph) {
        enqueue_all(true);
    } else if (whole) {
        hip_check(hipEventRecord(ev[0], st), "event");
        hip_check(hipGraphLaunch(whole_exec, st), "graph launch");
        hip_check(hipEventRecord(ev[kNumStages], st), "event");
    } else {
        for (int k = 0; k < kNumStages; ++k) {
            hip_check(hipEventRecord(ev[k], st), "event");
            if (k == kPredictStage) hip_check(hipStreamWaitEvent(st, ev_fine, 0), "wait fine grid");
            hip_check(hipGraphLaunch(execs[k], st), "graph launch");
            if (k == kGridStage) {  // fork: the fine walk structure beside the pair stages
                hip_check(hipEventRecord(ev_grid, st), "event");
                hip_check(hipStreamWaitEvent(side(), ev_grid, 0), "wait grid");
                hip_check(hipGraphLaunch(fine_exec, side()), "graph launch");
                hip_check(hipEventRecord(ev_fine, side()), "event");
            }
        }
        hip_check(hipEventRecord(ev[kNumStages], st), "event");
    }
    hip_check(hipMemcpyAsync(h_counts, cnt, sizeof(DevCounts), hipMemcpyDeviceToHost, st), "d2h");
    // the CA's last iteration flags come back with the counts, queued behind
    // this pipeline only: a read queued later on this stream could wait
    // behind work that other streams queued in the meantime
    hip_check(hipMemcpyAsync(h_flags, cb.flags, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "d2h");
    submitted = true;
    have_counts = false;
}

int trax_batch::finish() {
    // the CA forward phase ran a fixed number of iterations; continue on the
    // host if the last one still found living cells
    uint32_t* flags = h_flags;  // copied back with the counts at the end of launch()
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    uint32_t k = ca_iter;
    if (h_counts->n_conn && !(h_counts->err & kErrVoid) && flags[(k - 1) % 3]) {
        while (true) {
            hip_check(launch_ca_iterations(caps.max_conn, lb.basis, lb.foll, cb, cnt, k, 1, lc, st), "ca iteration");
            hip_check(hipMemcpyAsync(flags, cb.flags, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "d2h");
            hip_check(hipStreamSynchronize(st), "sync");
            if (!flags[k % 3]) break;
            ++k;
        }
        enqueue_ca_tail();
        hip_check(hipMemcpyAsync(h_counts, cnt, sizeof(DevCounts), hipMemcpyDeviceToHost, st), "d2h");
        hip_check(hipStreamSynchronize(st), "sync");
    }
    have_counts = true;
    const uint32_t err = h_counts->err;
    if (err & kErrCapacity) return fail(TRAX_ERR_CAPACITY, "an output exceeded its capacity");
    if (err & kErrCellRange) return fail(TRAX_ERR_INPUT, "a hit lies outside the z-phi grid");
    if (err & kErrIndexRange) return fail(TRAX_ERR_INPUT, "an index read from device memory is out of range");
    if (err) return fail(TRAX_ERR_INPUT, "device reported an input error");
    return TRAX_OK;
}

void trax_batch::enqueue_stage(int s) {
    const GridGeom g = cfg.grid.geom();
    const TripletCols tc = cm->cols();
    const uint32_t* eo = sup;
    const uint32_t* ln = sup + E;
    const uint32_t* lo = sup + E + (size_t)E * L;
    const uint32_t nseg = E * T;
    const uint32_t P = caps.max_pairs, Q = caps.max_cands, R = caps.max_tracklets, K = caps.max_conn,
                   N = caps.max_hits;
    auto ok = [](hipError_t e, const char* w) { hip_check(e, w); };
    switch (s) {
        case 0:  // grid_count: the whole grid build (count, scan, stable scatter) in one kernel
            ok(launch_grid_build(g, E * L, eo, ln, lo, grid_rec, hin, hout, grid, grid_perm, N, cnt, lc, st, true),
               "grid build");  // the run's counts are cleared first, in the same launches
            break;
        case 1:  // grid_store: part of the grid build above
            break;
        case 2:  // pair_plan
            ok(launch_pair_plan(g, tc, nseg, grid, seg_base, cnt, st), "pair plan");
            ok(launch_scan(seg_base, &cnt->n_segments, 1u, nseg + 1ull, &cnt->n_pair_slots, scan, st), "scan");
            ok(launch_slot_segments(seg_base, slot_cap, slot_seg, cnt, st), "slot segments");
            break;
        case 3:  // pair_count
            ok(launch_pair_count(g, tc, nseg, slot_cap, grid, hout.x, hout.y, hout.z, hout.zphi, seg_base, slot_counts,
                                 cnt, lc, st, slot_seg),
               "pair count");
            break;
        case 4:  // pair_scan
            ok(launch_scan(slot_counts, &cnt->n_pair_slots, 1u, slot_cap + 1ull, &cnt->n_pairs, scan, st), "scan");
            break;
        case 5:  // pair_store
            ok(launch_pair_store(g, tc, nseg, slot_cap, P, grid, hout.x, hout.y, hout.z, hout.zphi, seg_base,
                                 slot_counts, pairs, pair_off, cnt, lc, st, slot_seg),
               "pair store");
            break;
        case 6:  // predict: count, scan and store in one pass (with the filter when fused, its outputs placed in
                 // the filter_count phase)
            ok(launch_fill(cand_off, 0u, nseg + 1ull, st), "fill");
            if (caps.fuse_filter)
                ok(launch_predict_filter(g, tc, E * L, N, P, R, grid, hout.x, hout.y, hout.z, hout.zphi, hout.layer,
                                         hout.evt, pairs, cand_off, th1, th2, th3, cnt, fine, scan, lc, st),
                   "predict + filter");
            else
                ok(launch_predict_fused(g, tc, E * L, N, P, Q, grid, hout.x, hout.y, hout.z, hout.zphi, hout.layer,
                                        hout.evt, pairs, cands, cand_off, cnt, fine, scan, lc, st),
                   "predict");
            break;
        case 7:  // filter_count (fused: the predict pass's tracklets and candidate groups placed in pair order)
            if (caps.fuse_filter) {
                ok(launch_predict_filter_place(tc, P, R, cand_off, th1, th2, th3, cnt, fine, scan, lc, st),
                   "predict + filter placement");
                break;
            }
            ok(launch_filter_count(tc, Q, hout.x, hout.y, hout.z, hout.layer, pairs, cands, bitmap, word_counts, cnt,
                                   lc, st),
               "filter count");
            break;
        case 8:  // filter_scan
            if (!caps.fuse_filter)
                ok(launch_scan(word_counts, &cnt->n_filter_words, 1u, Q / 64 + 2ull, &cnt->n_tracklets, scan, st),
                   "scan");
            ok(launch_fill(t_off, 0u, nseg + 1ull, st), "fill");
            break;
        case 9:  // filter_store
            if (caps.fuse_filter)
                ok(launch_filter_offsets(tc, R, hout.layer, hout.evt, th1, t_off, cnt, lc, st), "filter offsets");
            else {
                ok(launch_filter_store(tc, Q, R, hout.layer, hout.evt, pairs, cands, bitmap, word_counts, th1, th2,
                                       th3, t_off, cnt, lc, st),
                   "filter store");
                ok(launch_filter_offsets(tc, R, hout.layer, hout.evt, th1, t_off, cnt, lc, st), "filter offsets");
            }
            break;
        case 10:  // link_prepare
            ok(launch_link_prepare(tc, N, R, hout.x, hout.y, hout.z, th1, th2, th3, lb, cnt, scan, lc, st),
               "link prepare");
            ok(launch_link_events(tc, E, R, hout.layer, hout.evt, th1, t_off, lb, cnt, st), "link events");
            break;
        case 11:  // link_count: count and store in one pass, the chunk counts scanned
            ok(launch_link_fused(tc, R, hout.layer, hout.evt, th1, th2, th3, t_off, lb, cnt, scan, lc, st), "link");
            break;
        case 12:  // link_store: the connections placed in base order
            ok(launch_link_fused_place(R, K, lb, cnt, lc, st), "link place");
            ok(launch_link_flags(R, hout.evt, th1, lb, cnt, lc, st), "link flags");
            ok(launch_scan(lb.flag_a, &cnt->n_tracklets, 1u, R + 1ull, &cnt->n_connectable, scan, st), "scan");
            ok(launch_scan(lb.flag_b, &cnt->n_tracklets, 1u, R + 1ull, &cnt->n_nonconnectable, scan, st), "scan");
            ok(launch_link_compact(R, lb, cnt, lc, st), "link compact");
            break;
        case 13:  // circle_fit
            ok(launch_circle_fit(R, hout.x, hout.y, th1, th2, th3, lb, cnt, lc, st), "circle");
            break;
        case 14:  // ca_forward
            ok(launch_ca_init(R, cb, cnt, st), "ca init");
            if (ca_layered)
                ok(launch_ca_layers(K, T, lb.basis, lb.foll, th1, hout.layer, cb, cnt, scan, lc, st), "ca layers");
            else
                ok(launch_ca_iterations(K, lb.basis, lb.foll, cb, cnt, 0, ca_iter, lc, st), "ca iterations");
            break;
        case 15:  // ca_backward
            enqueue_ca_tail();
            break;
    }
}

void trax_batch::enqueue_ca_tail() {
    const uint32_t R = caps.max_tracklets, K = caps.max_conn;
    hip_check(launch_ca_basis_count(R, K, lb.basis, lb.foll, cb, cnt, lc, st), "ca basis count");
    hip_check(launch_scan(cb.fbc, &cnt->n_tracklets, 1u, R + 1ull, &cnt->n_basis, scan, st), "scan");
    hip_check(launch_ca_basis_store(R, K, K, lb.basis, lb.foll, lb.pt, cb, cnt, lc, st), "ca basis store");
    hip_check(launch_scan(cb.handler, &cnt->n_tracklets, 1u, R + 1ull, &cnt->n_track_entries, scan, st), "scan");
    hip_check(launch_ca_tracks(R, caps.max_track_entries, cb, cnt, scan, lc, st, 0), "ca tracks");
}

namespace {

// fix the batch's shape from the configuration and capacities
void batch_shape(trax_batch& B, const trax_config* cfg, const trax_batch_caps* caps) {
    validate_config(cfg);
    if (!caps->max_events || !caps->max_hits) throw InputError("capacities need at least one event and hit");
    if (caps->use_graph > 2) throw InputError("use_graph is 0 (eager), 1 (a graph per phase) or 2 (one graph)");
    B.cfg = *cfg;
    B.caps = *caps;
    B.E = caps->max_events;
    B.L = cfg->trip.maxLayer();
    B.T = cfg->trip.size();
    B.cfg.grid.nEvents = B.E;
    B.cfg.grid.nLayers = B.L;
    B.C = B.cfg.grid.cells();
    B.ca_iter = caps->ca_iterations ? caps->ca_iterations : B.T + 2;
    // every connection leads to a higher layer triplet when each triplet's
    // second layer is above the layer that indexes it (its first)
    B.ca_layered = caps->ca_iterations == 0 && B.T <= 256;
    for (uint32_t i = 0; i < B.T; ++i) B.ca_layered = B.ca_layered && cfg->trip.layer2[i] >= i + 2;
    const uint64_t slot_cap = (uint64_t)caps->max_hits + (uint64_t)B.E * B.T * cfg->n_threads;
    if (slot_cap >= 0xffffffffull || (uint64_t)B.E * B.L * B.C >= 0xffffffffull)
        throw InputError("capacities too large");
    B.slot_cap = (uint32_t)slot_cap;
    const uint64_t P = caps->max_pairs, Q = caps->max_cands, R = caps->max_tracklets;
    B.scan_cap = (uint32_t)std::max<uint64_t>({caps->max_hits + 1ull, slot_cap + 1, P + 1, Q / 64 + 2, R + 1,
                                               (uint64_t)B.E * B.T + 1, predict_fused_chunks(P) + 1,
                                               link_chunks(R) + 1, (uint64_t)B.T * kCaTableBlocks});
}

uint32_t batch_scan_tiles(const trax_batch& B) {
    // the status words also hold the fused predictor's work counters
    return std::max({scan_tiles(B.scan_cap), predict_fused_tiles(B.caps.max_pairs), kPredCounterWords / 2});
}

}  // namespace

// one arena for every buffer of the pipeline; run once to measure the arena,
// once to carve it
void trax_batch::layout(Arena& a) {
    const uint64_t E = this->E, L = this->L, T = this->T, C = this->C, N = caps.max_hits, P = caps.max_pairs,
                   Q = caps.max_cands, R = caps.max_tracklets, K = caps.max_conn, TE = caps.max_track_entries;
    trax_batch& B = *this;
        auto cols = [&](HitColumns& c) {
            c.x = a.take<float>(N); c.y = a.take<float>(N); c.z = a.take<float>(N);
            c.layer = a.take<uint32_t>(N); c.det = a.take<uint32_t>(N); c.id = a.take<uint32_t>(N);
            c.evt = a.take<uint32_t>(N);
        };
        cols(B.hin);
        cols(B.hout);
        B.hout.zphi = a.take<float2>((uint64_t)N + kZphiSlack);
        B.fine.grid = a.take<uint32_t>(E * L * fine_geom(B.cfg.grid.geom(), fine_max(B.cfg.grid.geom())).cells);
        B.fine.hits = a.take<float4>((uint64_t)N + kFineSlack);
        if (caps.fuse_filter) {
            B.hout.xy = a.take<float2>(N);
            B.fine.xy = B.hout.xy;
            B.fine.pool_cap = kFusedPoolCap;
            B.fine.pool_blocks = kFusedPoolBlocks;
            B.fine.pool = a.take<uint2>((uint64_t)kFusedPoolCap * kFusedPoolBlocks);
            const uint64_t chunks = predict_fused_chunks(P) + 1;
            B.fine.chunk_cands = a.take<uint32_t>(chunks);
            B.fine.chunk_trk = a.take<uint32_t>(chunks);
            B.fine.chunk_at = a.take<uint32_t>(chunks);
            B.fine.group_chunk = a.take<uint32_t>(E * T + 1);
            B.fine.scratch_cap = (uint32_t)fused_scratch_cap(R);
            B.fine.scratch = a.take<uint32_t>(3ull * B.fine.scratch_cap);
        }
        B.sup = a.take<uint32_t>(E + 2 * E * L);
        B.grid = a.take<uint32_t>(E * L * C);
        B.grid_perm = a.take<uint32_t>(N + (uint64_t)E * L);  // scatter scratch, then the item order
        B.grid_rec = a.take<float4>(2 * (uint64_t)N);
        B.seg_base = a.take<uint32_t>(E * T + 1);
        B.slot_counts = a.take<uint32_t>(slot_cap + 1);
        B.slot_seg = a.take<uint32_t>(slot_seg_entries(slot_cap));
        B.pairs = a.take<uint32_t>(2 * P);
        B.pair_off = a.take<uint32_t>(E * T + 1);
        B.pred_counts = a.take<uint32_t>(P + 1);
        const uint64_t Qs = caps.fuse_filter ? 0 : Q;  // candidates are stored only when not fused
        B.cands = a.take<uint32_t>(2 * Qs);
        B.cand_off = a.take<uint32_t>(E * T + 1);
        B.bitmap = a.take<uint64_t>(Qs / 64 + 2);
        B.word_counts = a.take<uint32_t>(Qs / 64 + 2);
        B.th1 = a.take<uint32_t>(R); B.th2 = a.take<uint32_t>(R); B.th3 = a.take<uint32_t>(R);
        B.t_off = a.take<uint32_t>(E * T + 1);
        LinkBufs& lb = B.lb;
        lb.eta = a.take<float>(R);
        lb.hit_count = a.take<uint32_t>(N + 1);
        lb.hit_cursor = a.take<uint32_t>(N);
        lb.csr = a.take<uint32_t>(R);
        lb.conn = a.take<uint32_t>(R + 1);
        lb.counts = a.take<uint32_t>(R + 1);
        lb.chunk_n = a.take<uint32_t>(link_chunks(R) + 1);
        lb.chunk_at = a.take<uint32_t>(link_chunks(R) + 1);
        lb.scratch_cap = (uint32_t)link_scratch_cap(K);
        lb.scratch = a.take<uint32_t>(3ull * lb.scratch_cap);
        lb.basis = a.take<uint32_t>(K); lb.foll = a.take<uint32_t>(K); lb.pev = a.take<uint32_t>(K);
        lb.flag_a = a.take<uint32_t>(R + 1); lb.flag_b = a.take<uint32_t>(R + 1);
        lb.connectable = a.take<uint32_t>(R); lb.nonconnectable = a.take<uint32_t>(R);
        lb.pt = a.take<float>(R);
        lb.per_event = caps.per_event ? 1u : 0u;
        {
            // one run of tracklets per middle hit when no layer is the middle layer of two triplets
            std::set<uint32_t> mid;
            for (uint32_t i = 0; i < cfg.trip.size(); ++i) mid.insert(cfg.trip.layer2[i]);
            lb.runs = mid.size() == cfg.trip.size() ? 1u : 0u;
        }
        if (lb.per_event) {
            lb.ev_first = a.take<uint32_t>(E); lb.ev_last = a.take<uint32_t>(E); lb.ev_has = a.take<uint32_t>(E);
            lb.ev_off = a.take<uint32_t>(E * (T + 1));
        }
        CABufs& cb = B.cb;
        cb.st_a = a.take<uint32_t>(R); cb.st_b = a.take<uint32_t>(R);
        cb.fbc = a.take<uint32_t>(R + 1); cb.cursor = a.take<uint32_t>(R);
        cb.bidx = a.take<uint32_t>(K); cb.bdiff = a.take<float>(K);
        cb.best = a.take<uint32_t>(R); cb.isbest = a.take<uint32_t>(R);
        cb.handler = a.take<uint32_t>(R + 1);
        cb.tracks = a.take<uint32_t>(TE);
        cb.flags = a.take<uint32_t>(3);
        cb.clt = B.ca_layered ? a.take<uint8_t>(K) : nullptr;
        cb.lt_conn = B.ca_layered ? a.take<uint2>(K) : nullptr;
        cb.lt_tab = B.ca_layered ? a.take<uint32_t>((uint64_t)B.T * kCaTableBlocks + 1) : nullptr;
}

extern "C" int trax_batch_plan(const trax_config* cfg, const trax_batch_caps* caps, uint64_t* device_bytes,
                               uint64_t* pinned_bytes) {
    return guarded([&] {
        trax_batch B;
        batch_shape(B, cfg, caps);
        Arena measure;
        measure.cap = ~(size_t)0;
        B.layout(measure);
        if (device_bytes)
            *device_bytes = measure.used + (uint64_t)(batch_scan_tiles(B) + 1) * 8 + sizeof(DevCounts);
        if (pinned_bytes)
            *pinned_bytes = (7ull * caps->max_hits + B.E + 2ull * B.E * B.L + 2) * 4 + sizeof(DevCounts);
        return TRAX_OK;
    });
}

extern "C" int trax_batch_create(trax_ctx* ctx, const trax_config* cfg, const trax_batch_caps* caps,
                                 const trax_tuning* tuning, trax_stream* s, trax_batch** out) {
    return guarded([&] {
        hip_check(hipSetDevice(ctx->c->device()), "hipSetDevice");
        std::unique_ptr<trax_batch> b(new trax_batch);
        b->owner = ctx;
        b->st = pick(ctx, s);
        b->lc = launch_cfg(tuning);
        if (!predict_variant_supported(b->lc.predict_block, b->lc.predict_list, b->lc.predict_sort))
            throw InputError("unsupported fused predictor variant (block, list, sort)");
        if (b->lc.grid_stage_kib > kGridLdsMax / 1024 || b->lc.grid_order > 2 ||
            (b->lc.predict_wave > 64 && b->lc.predict_wave != kPredTiles))
            throw InputError("grid build: LDS per block at most 160 KiB, item order 0, 1 or 2; predictor "
                             "chunks per claim 0 to 64, or 128 for tiles");
        for (uint32_t v : {b->lc.grid_count_block, b->lc.grid_store_block, b->lc.pair_block, b->lc.filter_block,
                           b->lc.link_block})
            if (v % 64 != 0 || v > 1024) throw InputError("block sizes must be multiples of 64 up to 1024");
        predict_fused_prepare(b->lc);  // before any capture
        grid_build_prepare();
        auto& B = *b;
        batch_shape(B, cfg, caps);
        {
            const uint32_t F = b->lc.predict_fine ? b->lc.predict_fine : 8, Fmax = fine_max(B.cfg.grid.geom());
            if ((F & (F - 1)) != 0 || F > Fmax)
                throw InputError("predict_fine must be a power of two up to " + std::to_string(Fmax) +
                                 " for this grid");
        }
        const uint64_t E = B.E, L = B.L, N = caps->max_hits, P = caps->max_pairs, Q = caps->max_cands,
                       R = caps->max_tracklets, K = caps->max_conn, TE = caps->max_track_entries;
        B.cm.reset(new ConfigMirror(*ctx->c, cfg->trip, cfg->geom, cfg->n_threads, cfg->d_eta_cut));
        Arena measure;
        measure.cap = ~(size_t)0;
        B.layout(measure);
        B.arena.cap = measure.used;
        hip_check(hipMalloc(&B.arena_mem, B.arena.cap), "hipMalloc(arena)");
        B.arena.base = B.arena_mem;
        B.layout(B.arena);
        const uint32_t tiles = batch_scan_tiles(B);
        hip_check(hipMalloc(&B.scan.status, (size_t)(tiles + 1) * sizeof(uint64_t)), "hipMalloc(scan)");
        B.scan.max_tiles = tiles;
        hip_check(hipMalloc(&B.cnt, sizeof(DevCounts)), "hipMalloc(counts)");
        DevCounts init{};
        init.cap_slots = B.slot_cap; init.cap_pairs = (uint32_t)P; init.cap_cands = (uint32_t)Q;
        init.cap_tracklets = (uint32_t)R; init.cap_conn = (uint32_t)K; init.cap_basis = (uint32_t)K;
        init.cap_tracks = (uint32_t)TE;
        hip_check(hipMemcpy(B.cnt, &init, sizeof(DevCounts), hipMemcpyHostToDevice), "h2d counts");
        B.staging_words = 7 * N + E + 2 * E * L + 2;
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&B.staging), B.staging_words * 4, hipHostMallocDefault),
                  "hipHostMalloc");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&B.h_counts), sizeof(DevCounts), hipHostMallocDefault),
                  "hipHostMalloc");
        // pinned, so that reading them back is an ordinary asynchronous copy on the batch's stream
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&B.h_flags), 4 * sizeof(uint32_t), hipHostMallocDefault),
                  "hipHostMalloc");
        for (auto& e : B.ev) hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipStreamCreateWithFlags(&B.st2, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreateWithFlags(&B.ev_grid, hipEventDisableTiming), "hipEventCreate");
        hip_check(hipEventCreateWithFlags(&B.ev_fine, hipEventDisableTiming), "hipEventCreate");
        B.fine.prebuilt = true;
        (void)stride_blocks(1, 256, B.lc);  // resolve the CU count before any capture
        *out = b.release();
        return TRAX_OK;
    });
}

extern "C" void trax_batch_destroy(trax_batch* b) {
    if (!b) return;
    (void)hipStreamSynchronize(b->st);
    if (b->st2) (void)hipStreamSynchronize(b->st2);
    for (int i = 0; i < kNumStages; ++i) {
        if (b->execs[i]) (void)hipGraphExecDestroy(b->execs[i]);
        if (b->graphs[i]) (void)hipGraphDestroy(b->graphs[i]);
    }
    if (b->fine_exec) (void)hipGraphExecDestroy(b->fine_exec);
    if (b->whole_exec) (void)hipGraphExecDestroy(b->whole_exec);
    if (b->whole_graph) (void)hipGraphDestroy(b->whole_graph);
    if (b->fine_graph) (void)hipGraphDestroy(b->fine_graph);
    if (b->ev_grid) (void)hipEventDestroy(b->ev_grid);
    if (b->ev_fine) (void)hipEventDestroy(b->ev_fine);
    if (b->st2) (void)hipStreamDestroy(b->st2);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->arena_mem) (void)hipFree(b->arena_mem);
    if (b->scan.status) (void)hipFree(b->scan.status);
    if (b->cnt) (void)hipFree(b->cnt);
    if (b->staging) (void)hipHostFree(b->staging);
    if (b->h_counts) (void)hipHostFree(b->h_counts);
    if (b->h_flags) (void)hipHostFree(b->h_flags);
    delete b;
}

extern "C" uint64_t trax_batch_footprint(const trax_batch* b) {
    return b->arena.cap + (uint64_t)(b->scan.max_tiles + 1) * 8 + sizeof(DevCounts);
}

extern "C" int trax_batch_submit(trax_batch* b, const trax_hits* h) {
    return guarded([&] {
        const uint32_t E = h->n_events, N = h->hits.size(), L = b->L;
        if (E == 0) throw InputError("empty hit collection");
        if (E > b->E || N > b->caps.max_hits) throw CapacityError("batch exceeds the event or hit capacity");
        if (h->max_layer != L) throw InputError("hit collection layers differ from the configuration");
        uint32_t running = 0;
        for (uint32_t e = 0; e < E; ++e) {
            if (h->evt_offset[e] != running) throw InputError("event offsets are not running sums");
            running += h->evt_nhits[e];
        }
        hip_check(hipSetDevice(b->owner->c->device()), "hipSetDevice");
        hip_check(hipStreamSynchronize(b->st), "hipStreamSynchronize");  // the staging buffer is reused
        const uint32_t Ecap = b->E, Ncap = b->caps.max_hits;
        uint32_t* s = b->staging;
        const HitCollection& c = h->hits;
        std::memcpy(s + 0 * (size_t)Ncap, c.x.data(), N * 4ull);
        std::memcpy(s + 1 * (size_t)Ncap, c.y.data(), N * 4ull);
        std::memcpy(s + 2 * (size_t)Ncap, c.z.data(), N * 4ull);
        std::memcpy(s + 3 * (size_t)Ncap, c.layer.data(), N * 4ull);
        std::memcpy(s + 4 * (size_t)Ncap, c.detId.data(), N * 4ull);
        std::memcpy(s + 5 * (size_t)Ncap, c.hitId.data(), N * 4ull);
        std::memcpy(s + 6 * (size_t)Ncap, c.event.data(), N * 4ull);
        uint32_t* sp = s + 7 * (size_t)Ncap;
        std::memcpy(sp, h->evt_offset.data(), E * 4ull);
        std::memcpy(sp + Ecap, h->lay_nhits.data(), (size_t)E * L * 4);
        std::memcpy(sp + Ecap + (size_t)Ecap * L, h->lay_offset.data(), (size_t)E * L * 4);
        uint32_t* shape = sp + Ecap + 2 * (size_t)Ecap * L;
        shape[0] = E;
        shape[1] = N;
        hipStream_t st = b->st;
        float* dcols[3] = {b->hin.x, b->hin.y, b->hin.z};
        uint32_t* ucols[4] = {b->hin.layer, b->hin.det, b->hin.id, b->hin.evt};
        for (int k = 0; k < 3; ++k)
            hip_check(hipMemcpyAsync(dcols[k], s + k * (size_t)Ncap, N * 4ull, hipMemcpyHostToDevice, st), "h2d");
        for (int k = 0; k < 4; ++k)
            hip_check(hipMemcpyAsync(ucols[k], s + (3 + k) * (size_t)Ncap, N * 4ull, hipMemcpyHostToDevice, st), "h2d");
        hip_check(hipMemcpyAsync(b->sup, sp, (Ecap + 2 * (size_t)Ecap * L) * 4, hipMemcpyHostToDevice, st), "h2d");
        hip_check(hipMemcpyAsync(b->cnt, shape, 8, hipMemcpyHostToDevice, st), "h2d");
        b->n_events = E;
        b->n_hits = N;
        b->launch();
        b->have_counts = false;
        return TRAX_OK;
    });
}

extern "C" int trax_batch_rerun(trax_batch* b) {
    return guarded([&] {
        if (!b->n_events) return fail(TRAX_ERR_STATE, "no batch uploaded");
        hip_check(hipSetDevice(b->owner->c->device()), "hipSetDevice");
        b->launch();
        return TRAX_OK;
    });
}

extern "C" int trax_batch_wait(trax_batch* b) {
    return guarded([&] {
        if (!b->submitted) return fail(TRAX_ERR_STATE, "nothing submitted");
        hip_check(hipSetDevice(b->owner->c->device()), "hipSetDevice");
        return b->finish();
    });
}

extern "C" int trax_batch_counts(trax_batch* b, uint32_t out[8]) {
    if (!b->have_counts) return fail(TRAX_ERR_STATE, "no completed run");
    const DevCounts& c = *b->h_counts;
    out[0] = c.n_pairs; out[1] = c.n_cands; out[2] = c.n_tracklets; out[3] = c.n_conn;
    out[4] = c.n_track_entries; out[5] = c.ca_iterations; out[6] = c.n_connectable; out[7] = c.err;
    return TRAX_OK;
}

extern "C" int trax_batch_stage_ms(trax_batch* b, const char* stage, float* ms) {
    return guarded([&] {
        if (b->whole && std::strcmp(stage, "total") != 0)
            return fail(TRAX_ERR_STATE, "phase times need per-phase graphs (use_graph = 1) or eager launches");
        for (int k = 0; k < kNumStages; ++k)
            if (std::strcmp(stage, kPhases[k]) == 0) {
                hip_check(hipEventElapsedTime(ms, b->ev[k], b->ev[k + 1]), "hipEventElapsedTime");
                return TRAX_OK;
            }
        for (const StageRange& r : kStageRanges)
            if (std::strcmp(stage, r.name) == 0) {
                hip_check(hipEventElapsedTime(ms, b->ev[r.first], b->ev[r.last + 1]), "hipEventElapsedTime");
                return TRAX_OK;
            }
        if (std::strcmp(stage, "total") == 0) {
            hip_check(hipEventElapsedTime(ms, b->ev[0], b->ev[kNumStages]), "hipEventElapsedTime");
            return TRAX_OK;
        }
        return fail(TRAX_ERR_ARG, "unknown stage");
    });
}

// names: comma-separated result arrays to download (null: all of them)
static int batch_result_impl(trax_batch* b, const char* names, trax_result** out) {
    return guarded([&] {
        if (!b->have_counts) return fail(TRAX_ERR_STATE, "no completed run");
        const DevCounts& c = *b->h_counts;
        auto* r = new trax_result;
        std::unique_ptr<trax_result> guard(r);
        const uint32_t E = b->n_events, N = b->n_hits, T = b->T;
        const std::string sel = names ? std::string(",") + names + "," : std::string();
        auto want = [&](const char* k) { return !names || sel.find(std::string(",") + k + ",") != std::string::npos; };
        auto dl = [&](const void* src, uint64_t n) {
            std::vector<uint32_t> v(n);
            if (n) hip_check(hipMemcpy(v.data(), src, n * 4, hipMemcpyDeviceToHost), "d2h");
            return v;
        };
        auto dlf = [&](const void* src, uint64_t n) {
            std::vector<float> v(n);
            if (n) hip_check(hipMemcpy(v.data(), src, n * 4, hipMemcpyDeviceToHost), "d2h");
            return v;
        };
        const uint32_t P = std::min(c.n_pairs, b->caps.max_pairs), Q = std::min(c.n_cands, b->caps.max_cands),
                       R = std::min(c.n_tracklets, b->caps.max_tracklets), K = std::min(c.n_conn, b->caps.max_conn);
        if (want("grid")) r->put("grid", dl(b->grid, (uint64_t)E * b->L * b->C));
        if (want("hits.x")) r->putf("hits.x", dlf(b->hout.x, N)); if (want("hits.y")) r->putf("hits.y", dlf(b->hout.y, N)); if (want("hits.z")) r->putf("hits.z", dlf(b->hout.z, N));
        if (want("hits.layer")) r->put("hits.layer", dl(b->hout.layer, N)); if (want("hits.det_id")) r->put("hits.det_id", dl(b->hout.det, N));
        if (want("hits.hit_id")) r->put("hits.hit_id", dl(b->hout.id, N)); if (want("hits.event")) r->put("hits.event", dl(b->hout.evt, N));
        if (want("pairs")) r->put("pairs", dl(b->pairs, 2ull * P));
        if (want("pair_offsets")) r->put("pair_offsets", dl(b->pair_off, (uint64_t)E * T + 1));
        if (want("cands") && !b->caps.fuse_filter) r->put("cands", dl(b->cands, 2ull * Q));
        if (want("cand_offsets")) r->put("cand_offsets", dl(b->cand_off, (uint64_t)E * T + 1));
        // the filter's decision per stored candidate: bit i % 64 of u64 word i / 64, as u32 halves
        if (want("filter.bits") && !b->caps.fuse_filter)
            r->put("filter.bits", dl(reinterpret_cast<const uint32_t*>(b->bitmap), 2ull * ((Q + 63) / 64)));
        if (want("tracklets.h1")) r->put("tracklets.h1", dl(b->th1, R)); if (want("tracklets.h2")) r->put("tracklets.h2", dl(b->th2, R)); if (want("tracklets.h3")) r->put("tracklets.h3", dl(b->th3, R));
        if (want("tracklet_offsets")) r->put("tracklet_offsets", dl(b->t_off, (uint64_t)E * T + 1));
        if (want("eta")) r->putf("eta", dlf(b->lb.eta, R));
        if (want("conn.basis")) r->put("conn.basis", dl(b->lb.basis, K));
        if (want("conn.followers")) r->put("conn.followers", dl(b->lb.foll, K));
        if (want("conn.event")) r->put("conn.event", dl(b->lb.pev, K));
        if (want("conn.connectable")) r->put("conn.connectable", dl(b->lb.connectable, K ? std::min(c.n_connectable, R) : 0));
        if (want("conn.nonconnectable")) r->put("conn.nonconnectable", dl(b->lb.nonconnectable, K ? std::min(c.n_nonconnectable, R) : 0));
        if (K) {
            if (want("pt")) r->putf("pt", dlf(b->lb.pt, R));
            if (want("ca.states")) r->put("ca.states", dl(b->cb.st_a, R));
            if (want("ca.best_basis")) r->put("ca.best_basis", dl(b->cb.best, R));
            if (want("tracks")) r->put("tracks", dl(b->cb.tracks, std::min(c.n_track_entries, b->caps.max_track_entries)));
            if (want("track_offsets")) r->put("track_offsets", dl(b->cb.handler, R + 1ull));
            if (want("ca.iterations")) r->put("ca.iterations", {c.ca_iterations});
        }
        if (want("diag")) r->put("diag", {c.diag_pair_short, c.diag_pred_short, c.diag_pred_trunc, c.diag_wrap_overlap});
        *out = guard.release();
        return TRAX_OK;
    });
}

extern "C" int trax_batch_result(trax_batch* b, trax_result** out) { return batch_result_impl(b, nullptr, out); }

extern "C" int trax_batch_result_select(trax_batch* b, const char* names, trax_result** out) {
    if (!names) return fail(TRAX_ERR_ARG, "null names");
    return batch_result_impl(b, names, out);
}

// ---- results --------------------------------------------------------------------

extern "C" int trax_result_get(const trax_result* r, const char* name, const void** data, uint64_t* count,
                               int* is_float) {
    auto it = r->arrays.find(name);
    if (it == r->arrays.end()) return fail(TRAX_ERR_ARG, std::string("no result array ") + name);
    *is_float = it->second.is_float ? 1 : 0;
    if (it->second.is_float) {
        *data = it->second.f.data();
        *count = it->second.f.size();
    } else {
        *data = it->second.u.data();
        *count = it->second.u.size();
    }
    return TRAX_OK;
}

extern "C" double trax_result_timing_ms(const trax_result* r, const char* stage) {
    auto it = r->timings.find(stage);
    return it == r->timings.end() ? -1.0 : it->second;
}

extern "C" void trax_result_destroy(trax_result* r) { delete r; }

// ---- math probes ----------------------------------------------------------------

namespace {
__host__ __device__ inline float math_fn(int fn, float x, float y) {
    switch (fn) {
        case 0: return tx_atan2f(y, x);
        case 1: return tx_acosf(x);
        case 2: return tx_atanf(x);
        default: return tx_asinhf(x);
    }
}
__global__ void k_math(int fn, const float* x, const float* y, float* out, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = math_fn(fn, x[i], y ? y[i] : 0.0f);
}
}  // namespace

extern "C" int trax_math_eval_host(int fn, const float* x, const float* y, float* out, uint64_t n) {
    if (fn < 0 || fn > 3) return fail(TRAX_ERR_ARG, "unknown function");
    for (uint64_t i = 0; i < n; ++i) out[i] = math_fn(fn, x[i], y ? y[i] : 0.0f);
    return TRAX_OK;
}

extern "C" int trax_filter_phi_decide_host(uint64_t n, const float* ux, const float* uy, const float* vx,
                                           const float* vy, const float* cut, uint8_t* surely_outside,
                                           uint8_t* passes) {
    if (!ux || !uy || !vx || !vy || !cut || !surely_outside || !passes) return fail(TRAX_ERR_ARG, "null argument");
    const float PI = TX_PI_F;
    for (uint64_t i = 0; i < n; ++i) {
        surely_outside[i] = tx_phi_surely_outside(ux[i], uy[i], vx[i], vy[i], tx_phi_prefilter_tan(cut[i])) ? 1 : 0;
        const float angle1 = tx_atan2f(uy[i], ux[i]);  // TripletThetaPhiFilter.h:95-103
        const float angle2 = tx_atan2f(vy[i], vx[i]);
        float delta = angle2 - angle1;
        delta += (delta > PI) ? -2 * PI : (delta < -PI) ? 2 * PI : 0;
        passes[i] = fabsf(delta) <= cut[i] ? 1 : 0;
    }
    return TRAX_OK;
}

extern "C" int trax_math_eval_device(trax_ctx* ctx, int fn, const float* x, const float* y, float* out, uint64_t n) {
    return guarded([&] {
        if (fn < 0 || fn > 3) return fail(TRAX_ERR_ARG, "unknown function");
        Context& c = *ctx->c;
        FVec dx(c, n), dy(c, n), dout(c, n);
        hip_check(hipMemcpy(dx.get_mem(), x, n * 4, hipMemcpyHostToDevice), "h2d");
        if (y) hip_check(hipMemcpy(dy.get_mem(), y, n * 4, hipMemcpyHostToDevice), "h2d");
        uint64_t blocks = (n + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        if (blocks == 0) blocks = 1;
        hipLaunchKernelGGL(k_math, dim3((uint32_t)blocks), dim3(256), 0, c.stream(), fn, dx.get_mem(),
                           y ? dy.get_mem() : nullptr, dout.get_mem(), n);
        hip_check(hipGetLastError(), "launch");
        c.finish_default_queue();
        hip_check(hipMemcpy(out, dout.get_mem(), n * 4, hipMemcpyDeviceToHost), "d2h");
        return TRAX_OK;
    });
}
