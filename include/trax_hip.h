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

/* launch shapes (0 = default): block sizes per kernel family, blocks per CU
 * of the grid-stride kernels, and the fused predictor's variant: (block,
 * LDS list entries per pair, walk order) one of (128, 32|64), (256,
 * 24|32|48|64), (512, 24|32) in pair order or (256, 32) longest walk first;
 * 0 = (256, 24, pair order); the z sub-rows per grid row of the fused
 * predictor's walk structure (1, 2, 4, 8 or 16; 0 = 8); and the batched grid
 * build's block (grid_count_block; 0 = 1024), item order (0 or 1 = largest
 * (event, layer) items first, 2 = item order) and LDS per block in KiB (up to
 * 160; above 64 the items that fit are staged whole in LDS; 0 = 160); and
 * the fused predictor + filter's scheduling (every wave on its own 64-pair
 * chunks, k = 1..64 consecutive chunks a claim, from the region of the pairs
 * of its XCD first; 0 = 2; 128 = tiles of the block's pairs); and the pair
 * kernels' block order (pair_xcd XCDs given contiguous slices of every
 * stride; 0 = 8, 1 = the plain grid-stride order).  Results do not depend on
 * them. */
typedef struct {
    uint32_t grid_count_block, grid_store_block, pair_block, predict_block, filter_block, link_block;
    uint32_t blocks_per_cu;
    uint32_t predict_list;
    uint32_t predict_sort;
    uint32_t predict_fine;
    uint32_t grid_order, grid_stage_kib;
    uint32_t predict_wave;
    uint32_t pair_xcd;
} trax_tuning;

/* ---- 2b. stage launchers ------------------------------------------------
 *
 * One launcher per kernel group of the reference's algorithm classes, on
 * device buffers the caller allocates, so that a driver can keep the
 * reference's call sequence (TripletBuilder.cpp:297-510) stage by stage.  As
 * there, sizes come from the host: every count a stage needs is the total of
 * a trax_prefix_sum over the previous stage's counts, downloaded (the
 * reference's downloadScalar) and passed on.  Launchers enqueue on the stream
 * and return; the stages of one chain share the context's device-side record
 * of sizes and error bits, so they run on one stream, between
 * trax_stage_begin and trax_stage_check.  A device-side error (a hit outside
 * the grid, an index out of range, an output past its size) stops the later
 * stages' kernels and is returned by trax_stage_check.  tuning may be NULL
 * (default launch shapes; results do not depend on them).
 *
 * The layer triplets (TripletConfigurations) and the geometry supplement go
 * to the device once, as a trax_device_config; the work-group size of the
 * configuration (trax_config_set_exec) fixes the pair order as the
 * reference's nThreads does, and its dEtaCut is the connectivity cut. */
typedef struct trax_device_config trax_device_config;
int trax_device_config_create(trax_ctx* ctx, const trax_config* cfg, trax_device_config** out);
void trax_device_config_destroy(trax_device_config* dc);

/* the hit columns after the grid build (grid order), as every later stage
 * reads them: the HitCollection columns the kernels use and the (z, phi)
 * column the window walks read (2 floats per hit, (z, atan2(y, x)), and room
 * for 8 more floats past the last hit) */
typedef struct {
    const float* x; const float* y; const float* z;
    const uint32_t* layer; const uint32_t* event;
    const float* zphi;
    uint32_t n_hits;
} trax_grid_hits;

/* clears the context's size record and error bits and sets the shape of the
 * event group (n_events events, n_hits hits) */
int trax_stage_begin(trax_ctx* ctx, uint32_t n_events, uint32_t n_hits, trax_stream* s);
/* waits for the stream and returns the chain's device-side errors as status */
int trax_stage_check(trax_ctx* ctx, trax_stream* s);

/* GridBuilder::run (GridBuilder.h:42) writing the (z, phi) column as well:
 * d_zphi holds 2 * n_hits + 8 floats */
int trax_grid_build_zphi(trax_ctx* ctx, uint32_t n_events, uint32_t n_layers, uint32_t n_hits, uint32_t n_z,
                         uint32_t n_phi, const uint32_t* d_event_offsets, const uint32_t* d_layer_nhits,
                         const uint32_t* d_layer_offsets, trax_hit_cols in, trax_hit_cols_out out, float* d_zphi,
                         uint32_t* d_grid, trax_stream* s);

/* PairGeneratorBeamspot (PairGeneratorBeamspot.cpp:5-183; pairCount .h:43-246,
 * pairStore .h:448-525).  plan: d_seg_base[n_events * T + 1] receives the
 * count slots of every (event, layer pair), nThreads * ceil(N_outer /
 * nThreads), and 0 last; its scan gives n_slots.  count: d_slot_counts[n_slots
 * + 1] receives the pairs of every slot in the reference's work-item order;
 * its scan gives n_pairs.  store: d_pairs[2 * n_pairs] (inner, outer) and
 * d_pair_offsets[n_events * T + 1]. */
int trax_pair_plan(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const uint32_t* d_grid,
                   uint32_t* d_seg_base, trax_stream* s);
int trax_pair_count(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const uint32_t* d_grid,
                    const trax_grid_hits* hits, const uint32_t* d_seg_base, uint32_t n_slots, uint32_t* d_slot_counts,
                    const trax_tuning* tuning, trax_stream* s);
int trax_pair_store(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const uint32_t* d_grid,
                    const trax_grid_hits* hits, const uint32_t* d_seg_base, uint32_t n_slots,
                    const uint32_t* d_slot_counts, uint32_t n_pairs, uint32_t* d_pairs, uint32_t* d_pair_offsets,
                    const trax_tuning* tuning, trax_stream* s);

/* TripletThetaPhiPredictor (TripletThetaPhiPredictor.cpp:10-170; predictCount
 * .h:49-230, predictStore .h:232-430).  count: d_counts[n_pairs + 1] receives
 * the candidates of every pair; its scan gives n_cands.  store: d_cands[2 *
 * n_cands] (pair, third hit) and the group ends in d_cand_offsets[n_events *
 * T + 1] (zeroed first); trax_offsets_fill then fills the empty groups
 * (predictOffsetMonotonizeStore, .h:432-447). */
int trax_predict_count(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const uint32_t* d_grid,
                       const trax_grid_hits* hits, uint32_t n_pairs, const uint32_t* d_pairs, uint32_t* d_counts,
                       const trax_tuning* tuning, trax_stream* s);
int trax_predict_store(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const uint32_t* d_grid,
                       const trax_grid_hits* hits, uint32_t n_pairs, const uint32_t* d_pairs, const uint32_t* d_counts,
                       uint32_t n_cands, uint32_t* d_cands, uint32_t* d_cand_offsets, const trax_tuning* tuning,
                       trax_stream* s);
/* the offsets "monotonize" (TripletThetaPhiPredictor.h:432-447,
 * TripletThetaPhiFilter.h:267-280) as the sequential forward fill of the zero
 * entries of d_offsets[n_events * n_triplets + 1] */
int trax_offsets_fill(trax_ctx* ctx, uint32_t* d_offsets, uint32_t n_events, uint32_t n_triplets, trax_stream* s);

/* TripletThetaPhiFilter (TripletThetaPhiFilter.cpp:10-208).  count
 * (filterCount + filterPopCount, .h:50-179): one validity bit per candidate in
 * d_bitmap[ceil(n_cands / 64)] (u64 words) and their counts in
 * d_word_counts[words + 1]; its scan gives n_tracklets.  store (filterStore +
 * filterOffsetStore, .h:181-261): the tracklets' hits d_h1, d_h2, d_h3
 * [n_tracklets] and d_tracklet_offsets[n_events * T + 1] (zeroed first, the
 * last tracklet's entry R - 1 as in the reference); trax_offsets_fill then
 * fills the empty groups. */
int trax_filter_count(trax_ctx* ctx, const trax_device_config* dc, const trax_grid_hits* hits, uint32_t n_pairs,
                      const uint32_t* d_pairs, uint32_t n_cands, const uint32_t* d_cands, uint64_t* d_bitmap,
                      uint32_t* d_word_counts, const trax_tuning* tuning, trax_stream* s);
int trax_filter_store(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const trax_grid_hits* hits,
                      uint32_t n_pairs, const uint32_t* d_pairs, uint32_t n_cands, const uint32_t* d_cands,
                      const uint64_t* d_bitmap, const uint32_t* d_word_counts, uint32_t n_tracklets, uint32_t* d_h1,
                      uint32_t* d_h2, uint32_t* d_h3, uint32_t* d_tracklet_offsets, const trax_tuning* tuning,
                      trax_stream* s);

/* TripletConnectivityTight (TripletConnectivityTight.cpp:5-474) and
 * TrackletCircleFitter (TrackletCircleFitter.cpp:4-48) buffers, sized by the
 * caller as noted */
typedef struct {
    float* eta;              /* n_tracklets */
    uint32_t* hit_count;     /* n_hits + 1: the tracklets of every first hit, scanned (CSR starts) */
    uint32_t* hit_cursor;    /* n_hits */
    uint32_t* csr;           /* n_tracklets */
    uint32_t* conn;          /* n_tracklets + 1: takes part in a connection */
    uint32_t* counts;        /* n_tracklets + 1: followers per base; its scan gives n_conn */
    uint32_t* basis;         /* n_conn */
    uint32_t* followers;     /* n_conn */
    uint32_t* event;         /* n_conn */
    uint32_t* flag_a;        /* n_tracklets + 1: connectable predicate; its scan gives n_connectable */
    uint32_t* flag_b;        /* n_tracklets + 1: inverted predicate; its scan gives n_nonconnectable */
    uint32_t* connectable;   /* n_connectable */
    uint32_t* nonconnectable;/* n_nonconnectable */
} trax_link_bufs;
/* tripletEtaCalculatorStore (.h:47-79) and the per-first-hit buckets of the tracklets */
int trax_link_prepare(trax_ctx* ctx, const trax_device_config* dc, const trax_grid_hits* hits, uint32_t n_tracklets,
                      const uint32_t* d_h1, const uint32_t* d_h2, const uint32_t* d_h3, const trax_link_bufs* b,
                      const trax_tuning* tuning, trax_stream* s);
/* tripletConnectivityTightCount (.h:83-167) */
int trax_link_count(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const trax_grid_hits* hits,
                    uint32_t n_tracklets, const uint32_t* d_h1, const uint32_t* d_h2, const uint32_t* d_h3,
                    const uint32_t* d_tracklet_offsets, const trax_link_bufs* b, const trax_tuning* tuning,
                    trax_stream* s);
/* tripletConnectivityTightStore (.h:169-243), counts scanned */
int trax_link_store(trax_ctx* ctx, const trax_device_config* dc, uint32_t n_events, const trax_grid_hits* hits,
                    uint32_t n_tracklets, const uint32_t* d_h1, const uint32_t* d_h2, const uint32_t* d_h3,
                    const uint32_t* d_tracklet_offsets, uint32_t n_conn, const trax_link_bufs* b,
                    const trax_tuning* tuning, trax_stream* s);
/* predicateInversionStore (.h:327-346) over the first n_tracklets - 1
 * tracklets (.cpp:277): flag_a and flag_b */
int trax_link_flags(trax_ctx* ctx, const trax_grid_hits* hits, uint32_t n_tracklets, uint32_t n_conn,
                    const uint32_t* d_h1, const trax_link_bufs* b, trax_stream* s);
/* streamCompactionGetValidIndexesStore (.h:251-273), flags scanned */
int trax_link_compact(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_connectable, uint32_t n_nonconnectable,
                      const trax_link_bufs* b, trax_stream* s);
/* trackletCircleFitterStore (TrackletCircleFitter.h:32-101): d_pt[n_tracklets]
 * receives the radius of every listed tracklet, 0 for the others */
int trax_circle_fit(trax_ctx* ctx, const trax_grid_hits* hits, uint32_t n_tracklets, const uint32_t* d_h1,
                    const uint32_t* d_h2, const uint32_t* d_h3, const uint32_t* d_connectable, uint32_t n_connectable,
                    float* d_pt, trax_stream* s);

/* CellularAutomaton (CellularAutomaton.cpp:7-464) buffers */
typedef struct {
    uint32_t* states;        /* n_tracklets: the states once the forward phase ends */
    uint32_t* states_next;   /* n_tracklets */
    uint32_t* flags;         /* 3 */
    uint32_t* basis_count;   /* n_tracklets + 1: bases per follower; its scan gives n_basis */
    uint32_t* cursor;        /* n_tracklets */
    uint32_t* basis_index;   /* n_basis */
    float* basis_diff;       /* n_basis */
    uint32_t* best_basis;    /* n_tracklets */
    uint32_t* is_best;       /* n_tracklets */
    uint32_t* track_offsets; /* n_tracklets + 1: track lengths; its scan gives n_track_entries */
    uint32_t* tracks;        /* n_track_entries */
} trax_ca_bufs;
/* states and next states 1, flags cleared */
int trax_ca_init(trax_ctx* ctx, uint32_t n_tracklets, const trax_ca_bufs* b, trax_stream* s);
/* forward iteration number `iteration` (0, 1, ...) of the CA (iteration,
 * .h:52-100); afterwards flags[iteration % 3] is non-zero while cells live
 * (the reference's living-cell count, CellularAutomaton.cpp:40-76) */
int trax_ca_iterate(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_conn, const uint32_t* d_basis,
                    const uint32_t* d_followers, const trax_ca_bufs* b, uint32_t iteration, trax_stream* s);
/* followerBasisCount (.h:142-168) */
int trax_ca_basis_count(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_conn, const uint32_t* d_basis,
                        const uint32_t* d_followers, const trax_ca_bufs* b, trax_stream* s);
/* followerBasisStore, followerBestBasisStore and handlersStateStore
 * (.h:170-288), basis counts scanned */
int trax_ca_basis_store(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_conn, const uint32_t* d_basis,
                        const uint32_t* d_followers, const float* d_pt, uint32_t n_basis, const trax_ca_bufs* b,
                        trax_stream* s);
/* trackCollectionStore (.h:290-323), track lengths scanned */
int trax_ca_tracks(trax_ctx* ctx, uint32_t n_tracklets, uint32_t n_track_entries, const trax_ca_bufs* b,
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
    /* 0: launch the sequence eagerly every run; 1: capture it once, one
     * graph per phase, replayed per run (every phase timed by
     * trax_batch_stage_ms); 2: capture it once as one graph (one launch per
     * run, no gaps between phases; only "total" is timed) */
    uint32_t use_graph;
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
    uint32_t use_graph;    /* the slots' batches: 0 eager, 1 a graph per phase, 2 one graph (as trax_batch_caps) */
    uint32_t fuse_filter;  /* the slots' batches run the predictor and the filter as one pass */
    /* 0: a batch takes consecutive waiting events in arrival order; 1: events
     * are grouped by hit count: the oldest waiting event's size class (hit
     * count by powers of four) first, in arrival order, then any other
     * waiting events that fit */
    uint32_t group_by_size;
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
 * "track_offsets", "ca.iterations", "diag"; batches that store the
 * candidates also give "filter.bits", the filter's decision per candidate
 * (bit i % 64 of the u64 word i / 64, as two u32 halves per word) */
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

/* PhysicsRecords (src/triplet_builder/PhysicsRecord.h:205-214): one merged
 * PhysicsRecord per layer triplet, with its eta and pT histograms (bins of
 * 0.1, counts valid / fake / clones / missed / misc and the per-bin rates). */
typedef struct trax_physics trax_physics;
int trax_physics_create(trax_physics** out);
void trax_physics_destroy(trax_physics* p);
/* PhysicsRecord(e, events, p, triplets).fillData + PhysicsRecords::addRecord
 * for every event e of the collection and layer triplet p, in that order
 * (TripletBuilder.cpp:515-524; PhysicsRecord.cpp:86-255, 349-360): the
 * result's grid-ordered hit coordinates and simulated track ids, its
 * tracklets and (reference) offsets, against the findable tracks
 * HitCollection::addEvent returned for each event of `hits`.  On error the
 * records are unchanged. */
int trax_physics_add(trax_physics* p, const trax_config* cfg, const trax_hits* hits, uint32_t n_hits, const float* x,
                     const float* y, const float* z, const uint32_t* hit_id, uint32_t n_tracklets, const uint32_t* h1,
                     const uint32_t* h2, const uint32_t* h3, const uint32_t* tracklet_offsets);
/* PhysicsRecords::merge (PhysicsRecord.cpp:362-367) */
int trax_physics_merge(trax_physics* into, const trax_physics* from);
int trax_physics_count(const trax_physics* p, uint32_t* n);
/* PhysicsRecords::csvDump (PhysicsRecord.cpp:369-392) as text: table 0 the
 * records' table ("layerTriplet n eff effVar fr frVar cr crVar", one row per
 * record, then the total row numbered records + 1); table 1 / 2 the eta / pT
 * histogram file of the record of `layer_triplet` (the total's for records
 * + 1), layout 0 as tHistogram::csvDump writes it (12 columns), layout 1 the
 * five columns "bin valid fake clones missed" of the published histogram
 * files.  Writes at most `size` bytes with a terminating NUL; *needed is the
 * full length + 1 (TRAX_ERR_CAPACITY when it exceeds size). */
int trax_physics_csv(const trax_physics* p, int table, uint32_t layer_triplet, int layout, char* buf, uint64_t size,
                     uint64_t* needed);
/* the findable tracks HitCollection::addEvent returned for `event`
 * (UINT32_MAX: all events in order), track ids ascending, each track's hits
 * in layer order.  With every output null, *n_tracks / *n_hits receive the
 * counts; otherwise they give the arrays' capacities and receive the counts;
 * hit_offsets has n_tracks + 1 entries. */
int trax_hits_truth(const trax_hits* h, uint32_t event, uint32_t* n_tracks, uint32_t* n_hits, uint32_t* track_id,
                    uint32_t* hit_offsets, float* x, float* y, float* z, uint32_t* layer);

/* Track-level evaluation of the CA's chains, printTracks
 * (TripletBuilder.cpp:1008-1066): extractTracks (:955-988, chains of at
 * least three tracklets as hit lists), cloneRemoval (:880-913, a track whose
 * hits are shared at 0.19 of the shorter's length or more with a kept track
 * replaces it when longer, is kept beside it when as long, is dropped when
 * shorter), splitRealFakeTracks (:915-929) and findUniqueRealTracks
 * (:990-1006).  simulated: the findable tracks of the events (the
 * reference's SIMULATED).  kept (optional, one per chain): bit 0 the chain's
 * track survives clone removal, bit 1 it is real. */
typedef struct {
    uint64_t total, filtered, valids, valid_unique, fakes, simulated;
    float fake_rate, clone_rate, efficiency;
} trax_track_summary;
int trax_track_eval(uint32_t n_hits, const uint32_t* hit_id, uint32_t n_tracklets, const uint32_t* h1,
                    const uint32_t* h2, const uint32_t* h3, uint32_t n_chains, const uint32_t* track_offsets,
                    const uint32_t* tracks, uint64_t simulated, trax_track_summary* out, uint8_t* kept);

/* Float64 audit of the filter's decisions (TripletThetaPhiFilter.h:83-150)
 * on the host's threads: for each of n_cands stored candidates ((pair,
 * third hit) interleaved, pairs (h1, h2) interleaved, hit columns
 * grid-ordered), pass[i] gets bit 0 theta window, 1 phi, 2 closest
 * approach, 3 radius when that cut passes in float64 from the float32
 * coordinates, against the float32 thresholds the kernels compare with;
 * 0x80 for an empty slot.  No counterpart in the reference: it measures how
 * the float32 decisions differ from float64 ones. */
int trax_filter_audit_f64(const trax_config* cfg, uint64_t n_cands, const uint32_t* cands, uint64_t n_pairs,
                          const uint32_t* pairs, uint32_t n_hits, const float* x, const float* y, const float* z,
                          const uint32_t* layer, float min_radius_curvature, int n_threads, uint8_t* pass);

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
