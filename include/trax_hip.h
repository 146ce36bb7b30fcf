/* trax_hip.h -- C ABI of the MI355X (gfx950) triplet-seeding library.
 *
 * This is the device boundary that replaces the `clever` OpenCL layer the
 * reference's algorithm classes sit on (src/datastructures/KernelWrapper.hpp:
 * 12-35 and the clever::context / clever::vector / KERNEL_CLASS call sites,
 * e.g. src/algorithms/GridBuilder.cpp:15-161).  Plain C, opaque handles, int
 * status codes (0 = ok, negative = error; trax_last_error() gives the text),
 * no exceptions across the ABI.  Device memory is passed as plain pointers.
 *
 * Three levels:
 *   1. context, streams, device memory (clever::context, clever::vector,
 *      transfer::download / downloadScalar);
 *   2. one launcher per kernel group of the reference (KERNEL_CLASS objects of
 *      GridBuilder, PrefixSum, PairGeneratorBeamspot, TripletThetaPhiPredictor,
 *      TripletThetaPhiFilter, TripletConnectivityTight, TrackletCircleFitter,
 *      CellularAutomaton), taking device pointers;
 *   3. whole-path entry points: the per-event-group call sequence of
 *      buildTriplets (src/triplet_builder/TripletBuilder.cpp:297-510) and the
 *      batched graph-captured pipeline, with results as named arrays.
 */
#ifndef TRAX_HIP_H
#define TRAX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRAX_OK 0
#define TRAX_ERR_ARG (-1)      /* invalid argument */
#define TRAX_ERR_HIP (-2)      /* HIP runtime error */
#define TRAX_ERR_INPUT (-3)    /* input violates a precondition (e.g. a hit outside the grid) */
#define TRAX_ERR_CAPACITY (-4) /* an output exceeded the preallocated capacity */
#define TRAX_ERR_STATE (-5)    /* call out of sequence */

typedef struct trax_ctx trax_ctx;       /* clever::context: device, default queue, scratch */
typedef struct trax_stream trax_stream; /* an in-order device queue */
typedef struct trax_hits trax_hits;     /* HitCollection + EventSupplement + LayerSupplement (host) */
typedef struct trax_config trax_config; /* GridConfig + TripletConfigurations + GeometrySupplement */
typedef struct trax_result trax_result; /* named output arrays of one run */
typedef struct trax_batch trax_batch;   /* batched pipeline: capacities, workspace, graph */

/* text of the last error of the calling thread (static storage) */
const char* trax_last_error(void);

/* ---- 1. context, streams, memory --------------------------------------- */

typedef struct {
    char name[64];
    uint32_t cu_count;
    uint32_t lds_bytes_per_block; /* clever::context::getLocalMemSize */
    uint32_t max_block;
    uint32_t wave_size;
    uint64_t global_mem_bytes;
} trax_device_info;

/* TripletBuilder.cpp:66 createContext */
int trax_ctx_create(int device, trax_ctx** out);
void trax_ctx_destroy(trax_ctx* ctx);
int trax_device_info_get(trax_ctx* ctx, trax_device_info* out);

int trax_stream_create(trax_ctx* ctx, trax_stream** out);
int trax_stream_destroy(trax_stream* s);
/* NULL = the context's default stream (clever::context::finish_default_queue) */
int trax_stream_sync(trax_ctx* ctx, trax_stream* s);

/* Device timer: events recorded on a stream around launches, read back as
 * elapsed milliseconds.  Replaces the cl_events that the kernel classes
 * collect and clever::context::getKernelPerf reads
 * (TripletBuilder.cpp:491-507, RuntimeRecord.cpp:151-160).  stop records the
 * end event, waits for it and writes the time since start to *ms. */
typedef struct trax_timer trax_timer;
int trax_timer_create(trax_ctx* ctx, trax_timer** out);
int trax_timer_destroy(trax_timer* t);
int trax_timer_start(trax_ctx* ctx, trax_timer* t, trax_stream* s);
int trax_timer_stop(trax_ctx* ctx, trax_timer* t, trax_stream* s, float* ms);

/* clever::vector<T,1>(count, ctx) / release_buffer */
int trax_buf_alloc(trax_ctx* ctx, uint64_t bytes, void** dptr);
int trax_buf_free(trax_ctx* ctx, void* dptr);
/* fill with a 32-bit value by a kernel (safe inside captured graphs) */
int trax_buf_fill_u32(trax_ctx* ctx, void* dptr, uint32_t value, uint64_t count, trax_stream* s);
/* device-to-device copy of count u32 values by a kernel (the CA's memcpy
 * kernel, src/algorithms/CellularAutomaton.h:120-139); also the known-size
 * copy used to calibrate HBM traffic counters */
int trax_buf_copy_u32(trax_ctx* ctx, void* dst, const void* src, uint64_t count, trax_stream* s);
/* clever transfer::toDevice / download, asynchronous on the stream */
int trax_h2d(trax_ctx* ctx, void* dst, const void* src, uint64_t bytes, trax_stream* s);
int trax_d2h(trax_ctx* ctx, void* dst, const void* src, uint64_t bytes, trax_stream* s);

/* ---- 2. kernel launchers ------------------------------------------------ */

/* PrefixSum::run (src/algorithms/PrefixSum.h:43-44, PrefixSum.cpp:8-36):
 * exclusive in-place scan of n u32 values; the sum of all n values is written
 * to *d_total when d_total is not NULL (the element the reference downloads
 * with downloadScalar after scanning count+1 entries). */
int trax_prefix_sum(trax_ctx* ctx, uint32_t* d_data, uint32_t n, uint32_t* d_total, trax_stream* s);

/* The same scan without allocating: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The scan's status
 * words live in a file-scope device array of 32771 u64 (256 KiB) that every
 * device holds until the process ends, so n is at most 2^28 (TRAX_ERR_ARG
 * above); no hipMalloc, hipFree or allocating library call is made.  All calls
 * on one device share that array: calls that overlap on one device (on
 * different streams, or from several threads) race and must be serialised by
 * the caller.  Results are those of trax_prefix_sum. */
int trax_prefix_sum_noalloc(trax_ctx* ctx, uint32_t* d_data, uint32_t n, uint32_t* d_total, trax_stream* s);

/* GridBuilder::run (src/algorithms/GridBuilder.h:42): gridCount + PrefixSum +
 * gridStore.  Hits in the seven SoA columns are scattered into the z-phi grid
 * (stable within a cell).  d_grid receives E*L*(nZ+1)*(nPhi+1) cells; the
 * *_out columns receive the reordered hits.  The supplements must be the
 * running sums HitCollection::addEvent produces. */
typedef struct {
    const float* x; const float* y; const float* z;
    const uint32_t* layer; const uint32_t* det_id; const uint32_t* hit_id; const uint32_t* event;
} trax_hit_cols;
typedef struct {
    float* x; float* y; float* z;
    uint32_t* layer; uint32_t* det_id; uint32_t* hit_id; uint32_t* event;
} trax_hit_cols_out;
int trax_grid_build(trax_ctx* ctx, uint32_t n_events, uint32_t n_layers, uint32_t n_hits, uint32_t n_z,
                    uint32_t n_phi, const uint32_t* d_event_offsets, const uint32_t* d_layer_nhits,
                    const uint32_t* d_layer_offsets, trax_hit_cols in, trax_hit_cols_out out, uint32_t* d_grid,
                    trax_stream* s);

/* ---- 3. whole-path entry points ---------------------------------------- */

/* HitCollection / EventSupplement / LayerSupplement (src/datastructures/
 * HitCollection.h:28, EventSupplement.h:14, LayerSupplement.h:13) */
int trax_hits_create(uint32_t max_layer, trax_hits** out);
void trax_hits_destroy(trax_hits* h);
/* HitCollection::addEvent (src/datastructures/HitCollection.cpp:21-180):
 * one event given as its raw hit list (PEvent hits: position, layer,
 * detector id, simulated track id, simulated track pT).  Uses the layers of
 * the configuration's layer triplets and its minPt; max_tracks = -1 for all. */
int trax_hits_add_event(trax_hits* h, const trax_config* cfg, uint32_t n, const float* x, const float* y,
                        const float* z, const uint32_t* layer, const uint32_t* det_id,
                        const uint32_t* track_id, const float* track_pt, int max_tracks, int only_tracks);
/* EventLoader (src/triplet_builder/EventLoader.h:15-67) + addEvent: the
 * events of a serialized PB_Event::PEventContainer (Event.proto:1-37),
 * decoded from the wire format without libprotobuf, each added as by
 * trax_hits_add_event; max_events = -1 for all.  *n_added: events added. */
int trax_hits_add_pb(trax_hits* h, const trax_config* cfg, const void* data, uint64_t size, int max_events,
                     int max_tracks, int only_tracks, uint32_t* n_added);
/* The same from an event store (GenericOutputStore / GenericInputStore,
 * src/datastructures/serialize/PEventStore.h:29-248, read as EventStoreLoader
 * does, src/triplet_builder/EventLoader.h:171-206): a little-endian u32 event
 * count, then per event an entry type varint (1), a size varint and the
 * serialized PB_Event::PEvent.  Reading stops at the count, at another entry
 * type, at a zero size or at a truncated entry; the first skip_events are
 * dropped and at most max_events (-1 = all) added.  *n_added: events added;
 * *n_stored: the store's event count (either may be NULL).  A malformed event
 * gives TRAX_ERR_INPUT. */
int trax_hits_add_pb_store(trax_hits* h, const trax_config* cfg, const void* data, uint64_t size, int skip_events,
                           int max_events, int max_tracks, int only_tracks, uint32_t* n_added, uint32_t* n_stored);
uint32_t trax_hits_count(const trax_hits* h);
uint32_t trax_hits_events(const trax_hits* h);
/* column access: "x","y","z","layer","det_id","hit_id","event",
 * "event_nhits","event_offset","layer_nhits","layer_offset" */
int trax_hits_column(const trax_hits* h, const char* name, const void** data, uint64_t* count);

/* configuration: grid, layer triplets (TripletConfiguration.h:33 column
 * order), per-layer radius brackets (GeometrySupplement.h:21), minPt */
int trax_config_create(uint32_t n_z, uint32_t n_phi, float min_pt, trax_config** out);
void trax_config_destroy(trax_config* c);
int trax_config_add_triplet(trax_config* c, uint32_t layer1, uint32_t layer2, uint32_t layer3, float d_theta_cut,
                            float sigma_z, float d_phi_cut, float sigma_phi, float tip_cut, uint32_t pair_spread_z,
                            float z0, uint32_t pair_spread_phi, uint32_t n_candidates);
int trax_config_set_geometry(trax_config* c, uint32_t n, const uint32_t* layer, const float* min_radius,
                             const float* max_radius);
/* work-group size of the reference (fixes the pair order) and dEtaCut */
int trax_config_set_exec(trax_config* c, uint32_t n_threads, float d_eta_cut);
float trax_config_min_radius_curvature(const trax_config* c);
/* Checks the preconditions every path relies on before it touches a device:
 * at least one and at most 255 layer triplets, layers numbered from 1, the
 * first two layers of each triplet not above the largest third layer, and a
 * geometry row for every layer the triplets use.  The reference assumes them
 * without checking (the triplet index is hitLayer - 1,
 * src/algorithms/TripletThetaPhiPredictor.h:88; radii are indexed by layer - 1,
 * GeometrySupplement.h:21-81).  Every run entry point applies the same check. */
int trax_config_validate(const trax_config* c);

/* The plain per-event-group path: buildTriplets' call sequence over the host
 * algorithm classes, one blocking count download per stage, exact-size
 * allocations, CA convergence checked on the host every iteration
 * (TripletBuilder.cpp:297-510).  stop_after: 0 = all stages, else the last
 * stage to run (1 grid, 2 pairs, 3 predictor, 4 filter, 5 connectivity+CA). */
int trax_run_plain(trax_ctx* ctx, const trax_config* cfg, const trax_hits* hits, int stop_after,
                   trax_result** out);
/* The same sequence with only the downloads the reference driver makes
 * (grid and hits after GridBuilder, tracklets after the filter, CA outputs);
 * *ms receives the host wall time from upload to the last download. */
int trax_run_plain_bench(trax_ctx* ctx, const trax_config* cfg, const trax_hits* hits, double* ms);

/* Batched pipeline: every stage of many events in one launch sequence with
 * device-resident sizes, preallocated capacities, captured in a hipGraph and
 * replayed per batch.  Capacities are upper bounds per event group. */
typedef struct {
    uint32_t max_events;
    uint32_t max_hits;
    uint32_t max_pairs;
    uint32_t max_cands;
    uint32_t max_tracklets;
    uint32_t max_conn;
    uint32_t max_track_entries;
    uint32_t use_graph;    /* capture the sequence once, replay per batch */
    uint32_t ca_iterations; /* CA iterations in the sequence; 0 = one pass per layer triplet when every
                               triplet's second layer is above its first (same states and iteration
                               count), else layer triplets + 2 iterations */
    /* 0: the batch is one event group, as buildTriplets runs a group of events
     *    (the connectivity stage leaves out the group's last tracklet,
     *    TripletConnectivityTight.cpp:277 and TripletThetaPhiFilter.h:250-253);
     * 1: every event is its own group, so each event's results equal a run of
     *    that event alone, whatever batch it shares */
    uint32_t per_event;
    /* 1: the predictor and the filter run as one pass: the tracklets come
     *    straight from the window walks and the candidates (TripletThetaPhi-
     *    Predictor's output) are counted, with their group offsets, but not
     *    stored, so max_cands does not size any buffer and the "cands" result
     *    is absent; every other result is the same */
    uint32_t fuse_filter;
} trax_batch_caps;
/* launch shapes (0 = default): block sizes per kernel family, blocks per CU
 * of the grid-stride kernels, and the fused predictor's variant: (block,
 * LDS list entries per pair, walk order) one of (128, 32|64), (256,
 * 24|32|48|64), (512, 24|32) in pair order or (256, 32) longest walk first;
 * 0 = (256, 24, pair order); and the z sub-rows per grid row of the fused
 * predictor's walk structure (1, 2, 4, 8 or 16; 0 = 8).  Results do not
 * depend on them. */
typedef struct {
    uint32_t grid_count_block, grid_store_block, pair_block, predict_block, filter_block, link_block;
    uint32_t blocks_per_cu;
    uint32_t predict_list;
    uint32_t predict_sort;
    uint32_t predict_fine;
} trax_tuning;
/* device and pinned host bytes a batch of these capacities would hold, without
 * allocating (the planning step of buildTriplets' per-group buffer sizing,
 * TripletBuilder.cpp:297-510, done once for all groups) */
int trax_batch_plan(const trax_config* cfg, const trax_batch_caps* caps, uint64_t* device_bytes,
                    uint64_t* pinned_bytes);
int trax_batch_create(trax_ctx* ctx, const trax_config* cfg, const trax_batch_caps* caps,
                      const trax_tuning* tuning, trax_stream* s, trax_batch** out);
void trax_batch_destroy(trax_batch* b);
/* Autotuner: picks every launch shape of trax_tuning for this workload by
 * coordinate descent over the supported values, timing the phases each one
 * drives with HIP events on a batch of the sample events (eager launches,
 * best of `repeats` runs per candidate) and checking that every candidate
 * gives the same tracklets.  *best receives the fastest shapes; report (if
 * not NULL) receives one JSON line per trial.  Returns TRAX_ERR_STATE if a
 * candidate changed the results. */
int trax_autotune(trax_ctx* ctx, const trax_config* cfg, const trax_hits* sample, const trax_batch_caps* caps,
                  uint32_t repeats, trax_tuning* best, char* report, uint64_t report_bytes);
/* upload a HitCollection and enqueue the whole pipeline; asynchronous */
int trax_batch_submit(trax_batch* b, const trax_hits* hits);
/* enqueue the pipeline again on the inputs already resident on the device */
int trax_batch_rerun(trax_batch* b);
/* wait for the submitted run; TRAX_ERR_CAPACITY if an output overflowed */
int trax_batch_wait(trax_batch* b);
/* results of the last completed run (downloads them) */
int trax_batch_result(trax_batch* b, trax_result** out);
/* the same for the named arrays only (comma-separated, names as in
 * trax_result_get): large batches without copying every intermediate */
int trax_batch_result_select(trax_batch* b, const char* names, trax_result** out);
/* counts of the last completed run: pairs, candidates, tracklets,
 * connections, track entries, CA iterations */
int trax_batch_counts(trax_batch* b, uint32_t out[8]);
/* device bytes held by the batch workspace */
uint64_t trax_batch_footprint(const trax_batch* b);

/* timing of the last batched run from HIP events recorded between the phases
 * on the batch's stream, in milliseconds.  Phases: "grid_count", "grid_store",
 * "pair_plan", "pair_count", "pair_scan", "pair_store", "predict" (count,
 * scan and store of the predictor in one pass), "filter_count",
 * "filter_scan", "filter_store", "link_prepare", "link_count", "link_store",
 * "circle_fit", "ca_forward", "ca_backward"; stages (the reference's
 * algorithm classes): "grid", "pairs", "predict", "filter", "connectivity",
 * "ca"; and "total". */
int trax_batch_stage_ms(trax_batch* b, const char* stage, float* ms);

/* Serving loop: events arrive one at a time and are grouped into batches of
 * whole events (continuous batching), run on n_slots batches in flight, each
 * with its own stream and workspace; each event's tracklets and tracks are
 * copied back as soon as its batch completes.  No counterpart in the
 * reference, which runs fixed event groups on one queue
 * (TripletBuilder.cpp:297-552); the per-event results are the same. */
typedef struct trax_server trax_server;
typedef struct {
    uint32_t n_slots;          /* batches in flight (default 2) */
    uint32_t max_batch_events; /* events per batch (default 256) */
    uint32_t max_batch_hits;   /* hits per batch; 0 = the most the device budget holds */
    uint64_t device_budget;    /* device bytes for all slots (0 = 32 GiB) */
    /* output capacities per input hit of a batch: upper bounds measured on
     * the workload; a batch that still overflows is re-run as smaller batches */
    float pairs_per_hit, cands_per_hit, tracklets_per_hit, conn_per_hit, track_entries_per_hit;
    uint32_t keep_results; /* keep each event's tracklets and tracks (trax_server_event_result) */
    uint32_t use_graph;
    uint32_t fuse_filter;  /* the slots' batches run the predictor and the filter as one pass */
} trax_serve_opts;
typedef struct {
    uint32_t events, failed, batches, resubmits, max_batch_events, slot_hits;
    uint64_t tracklets, track_entries, device_bytes;
    double wall_s, events_per_s, tracklets_per_s, mean_batch_events;
    double p50_ms, p99_ms, max_ms; /* latency: arrival to results on the host */
    /* means over batches: arrival of the first event to submission, submission
     * to the end of the pipeline, the result copy, the host-side split */
    double queue_ms, pipeline_ms, copy_ms, split_ms;
} trax_serve_stats;
int trax_server_create(trax_ctx* ctx, const trax_config* cfg, const trax_serve_opts* opts, trax_server** out);
void trax_server_destroy(trax_server* s);
/* capacities of each slot (hit capacity from the device budget) */
int trax_server_caps(const trax_server* s, trax_batch_caps* caps);
/* Replay a stream in real time: event i of `events` arrives arrival_s[i]
 * seconds (non-decreasing) after the call starts; returns when every event's
 * results are on the host.  latency_ms[i] (optional) receives completion minus
 * arrival (-1 for an event that overflowed a capacity on its own). */
int trax_server_replay(trax_server* s, const trax_hits* events, const double* arrival_s, double* latency_ms,
                       trax_serve_stats* stats);
/* per-event results of the last replay, hit indices relative to the event's
 * grid-ordered hits and tracklet indices relative to the event's tracklets:
 * "tracklets.h1", "tracklets.h2", "tracklets.h3", "track_offsets", "tracks"
 * (keep_results only), "n_tracklets", "n_track_entries" */
int trax_server_event_result(const trax_server* s, uint32_t event, const char* name, const uint32_t** data,
                             uint64_t* count);

/* named result arrays: "grid", "hits.x", ..., "pairs", "pair_offsets",
 * "cands", "cand_offsets", "tracklets.h1/h2/h3", "tracklet_offsets", "eta",
 * "conn.basis", "conn.followers", "conn.event", "conn.connectable",
 * "conn.nonconnectable", "pt", "ca.states", "ca.best_basis", "tracks",
 * "track_offsets", "ca.iterations", "diag" */
int trax_result_get(const trax_result* r, const char* name, const void** data, uint64_t* count, int* is_float);
double trax_result_timing_ms(const trax_result* r, const char* stage);
void trax_result_destroy(trax_result* r);

/* PhysicsRecord::fillData (src/triplet_builder/PhysicsRecord.cpp:86-255):
 * tracklets against the simulation truth per (event, layer triplet).  Inputs
 * are the grid-ordered hits' simulated track ids, layers and events, the
 * tracklets and their (reference) offsets.  Findable tracks: ids whose hits
 * cover every layer the configuration's triplets use (not id 0 with
 * only_tracks).  rows: n_events * layer triplets entries, event-major. */
typedef struct {
    uint32_t event, layer_triplet, n_tracklets, n_findable, found, fakes, clones, misc;
    double efficiency, fake_rate, clone_rate;
} trax_physics_row;
int trax_physics_eval(const trax_config* cfg, uint32_t n_events, uint32_t n_hits, const uint32_t* hit_id,
                      const uint32_t* hit_layer, const uint32_t* hit_event, uint32_t n_tracklets, const uint32_t* h1,
                      const uint32_t* h2, const uint32_t* h3, const uint32_t* tracklet_offsets, int only_tracks,
                      trax_physics_row* rows);

/* host and device evaluation of the shared float built-ins (atan2, acos,
 * atan, asinh), for accuracy and host/device agreement tests.
 * fn: 0 atan2(y, x), 1 acos(x), 2 atan(x), 3 asinh(x) */
int trax_math_eval_host(int fn, const float* x, const float* y, float* out, uint64_t n);
int trax_math_eval_device(trax_ctx* ctx, int fn, const float* x, const float* y, float* out, uint64_t n);
/* host evaluation of the filter's phi cut (TripletThetaPhiFilter.h:95-103)
 * on hit differences u = p2 - p1, v = p3 - p2: passes[i] = the reference's
 * decision, surely_outside[i] = the prefilter's proof of rejection (which
 * must imply !passes[i]); for tests */
int trax_filter_phi_decide_host(uint64_t n, const float* ux, const float* uy, const float* vx, const float* vy,
                                const float* cut, uint8_t* surely_outside, uint8_t* passes);

#ifdef __cplusplus
}
#endif

#endif /* TRAX_HIP_H */
