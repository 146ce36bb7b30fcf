// Host-side launchers of the gfx950 kernels (internal C++ interface used by
// the C ABI and the host algorithm classes).  Every launcher enqueues on the
// given stream and returns the HIP status of the launch.  Work sizes that
// depend on data (pairs, candidates, tracklets, connections) are read by the
// kernels from the device-resident DevCounts record; the *_cap arguments are
// host-side upper bounds used to size grids and to detect overflow.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace trax {

struct ScanScratch {
    uint64_t* status = nullptr;  // max_tiles + 1 words; the last one holds the tile counter
    uint32_t max_tiles = 0;
};

struct HitColumns {
    float* x = nullptr; float* y = nullptr; float* z = nullptr;
    uint32_t* layer = nullptr; uint32_t* det = nullptr; uint32_t* id = nullptr; uint32_t* evt = nullptr;
    // device-only column written by the grid scatter: (z, phi) of every
    // reordered hit with phi = atan2(y, x), the value every window test of
    // the reference recomputes (bit-identical: the same built-in on the same
    // inputs); one 8-byte load per visited hit in the window walks
    float2* zphi = nullptr;
    // device-only column written by the grid scatter: (x, y) of every
    // reordered hit, one 8-byte load for the fused filter's phi prefilter
    float2* xy = nullptr;
};

// per-kernel block sizes (0 = default); chosen by the autotuner
struct LaunchCfg {
    uint32_t grid_count_block = 0;
    uint32_t grid_store_block = 0;
    uint32_t pair_block = 0;
    uint32_t predict_block = 0;
    uint32_t filter_block = 0;
    uint32_t link_block = 0;
    uint32_t blocks_per_cu = 0;  // grid-stride kernels: blocks per CU (0 = 8)
    uint32_t predict_list = 0;   // fused predictor: LDS list entries per pair (0 = 24)
    uint32_t predict_sort = 0;   // fused predictor: walk a tile's pairs longest first
    uint32_t predict_fine = 0;   // fused predictor: z sub-rows per grid row (power of two; 0 = 8)
    uint32_t pair_xcd = 0;       // pair kernels: stride slices per XCD (0 = 8, one per XCD; 1 = plain grid-stride)
    uint32_t predict_wave = 0;   // fused predictor + filter: every wave on its own 64-pair chunks, k (1..64; 0 = 2)
                                 // a claim; kPredTiles: tiles of the block's pairs
    uint32_t grid_order = 0;     // grid build: 0 or 1 = the largest (event, layer) items first, 2 = item order
    uint32_t grid_stage_kib = 0; // grid build: LDS per block in KiB (above 64: items staged whole; 0 = 160)
};

// The fused predictor's own walk structure: every z row of every (event,
// layer) grid split into F sub-rows by z within the row, the hits of a row
// reordered (stable) by (sub-row, phi cell), each with its grid index.  Sub-row
// s of row r holds the hits whose fine_bin(z) is s; the prefix entries are laid
// out like the grid's, (n_z * F) rows of n_phi + 1 cell starts and one final
// entry, and use the grid's hit positions (the reordering stays inside a row).
struct FineGeom {
    uint32_t F = 1, shift = 0;  // sub-rows per row, log2 F
    float inv_sub = 0.0f;       // F / sector_z
    uint32_t cells = 0;         // n_z * F * (n_phi + 1) + 1 prefix entries per (event, layer)
};
FineGeom fine_geom(const GridGeom& g, uint32_t F);
// entries of room past the last hit of every (z, phi) column: the pair walks
// read it in aligned 16-byte blocks of four entries
constexpr uint32_t kZphiSlack = 4;
// entries of room past the last fine hit: the fine walk reads four entries
// at a time from a segment and masks the ones past its end
constexpr uint32_t kFineSlack = 4;
constexpr uint32_t kFineMax = 16;
// largest sub-row count the launch supports for this grid (power of two <= kFineMax)
uint32_t fine_max(const GridGeom& g);
struct FineBufs {
    uint32_t* grid = nullptr;  // E * L * fine_geom(g, kFineMax).cells prefix entries
    float4* hits = nullptr;    // (z, phi, grid index as bits, 0) at fine positions; N + kFineSlack entries
    // the fused predictor + filter: per resident block, a pool of pool_cap
    // store-window matches beyond a pair's LDS list
    uint2* pool = nullptr;
    uint32_t pool_cap = 0, pool_blocks = 0;
    // ... per chunk of 64 pairs: candidates, tracklets, and where its
    // tracklets went in the scratch (predict_fused_chunks(n_pair_cap) + 1 entries each)
    uint32_t* chunk_cands = nullptr;
    uint32_t* chunk_trk = nullptr;
    uint32_t* chunk_at = nullptr;
    uint32_t* group_chunk = nullptr;  // E * T + 1: the chunk of a candidate group's end
    uint32_t* scratch = nullptr;     // 3 * scratch_cap: the chunks' tracklets in arrival order
    uint32_t scratch_cap = 0;        // fused_scratch_cap(n_trk_cap)
    const float2* xy = nullptr;      // (x, y) of every grid-ordered hit (the grid scatter's column)
    bool prebuilt = false;           // grid and hits built by launch_fine_grid before the predictor's launch
};
constexpr uint32_t kFusedPoolCap = 8192, kFusedPoolBlocks = 2048;
// the fused predictor's work counters: kPredRegions of them, a cache line
// (kPredCounterStride u32) apart, at the start of the scan status words
constexpr uint32_t kPredRegions = 8, kPredCounterStride = 16, kPredCounterWords = kPredRegions * kPredCounterStride;
constexpr uint32_t kPredTiles = 128;  // LaunchCfg::predict_wave: the block-tiled pass
// tracklet scratch of the fused predictor + filter for n_trk_cap tracklets
uint64_t fused_scratch_cap(uint64_t n_trk_cap);

// layer-triplet configuration columns on the device (TripletConfigurations,
// src/datastructures/TripletConfiguration.h:33) and GeometrySupplement radii
struct TripletCols {
    const uint32_t* l1; const uint32_t* l2; const uint32_t* l3;
    const float* dtheta; const float* sigma_z; const float* dphi; const float* sigma_phi;
    const float* tip; const float* z0; const uint32_t* pair_spread_phi;
    const float* min_r; const float* max_r;  // index layer - 1
    // acos(radius / (2 minRadiusCurvature)) of every layer's min and max
    // radius: the per-layer terms of the pair and predictor phi windows,
    // computed once on the host with the same built-in
    const float* acos_min_r; const float* acos_max_r;
    // per layer triplet: tan(dPhiCut + 1e-5) of the filter's phi-cut prefilter
    // (tx_phi_surely_outside), non-positive when disabled
    const double* tan_dphi;
    uint32_t n_trip;
    uint32_t n_geom;
    float min_radius_curvature;
    uint32_t n_threads;  // the reference's work-group size; fixes the pair order
    float d_eta_cut;
};

// status words (u64) a scan of n_cap elements needs: one per tile of
// kScanTileElements elements, and two more
constexpr uint64_t kScanTileElements = 8192;
uint32_t scan_tiles(uint64_t n_cap);
uint32_t stride_blocks(uint64_t work_cap, uint32_t block, const LaunchCfg& lc);
hipError_t launch_fill(uint32_t* p, uint32_t v, uint64_t n, hipStream_t st);
// writes of DevCounts fields (word index, value) by a kernel on the stream, in order
struct CountWrites {
    uint32_t n = 0;
    uint32_t idx[12];
    uint32_t val[12];
    void put(const uint32_t* field, const DevCounts* base, uint32_t v) {
        idx[n] = (uint32_t)(field - reinterpret_cast<const uint32_t*>(base));
        val[n++] = v;
    }
};
hipError_t launch_set_counts(DevCounts* cnt, const CountWrites& w, hipStream_t st);
hipError_t launch_copy(uint32_t* dst, const uint32_t* src, uint64_t n, hipStream_t st);
// event column of a batch of whole events from its event offsets (sizes on device)
hipError_t launch_event_ids(uint32_t* evt, const uint32_t* evt_offset, const DevCounts* cnt, uint32_t n_events_cap,
                            hipStream_t st);
// fill *d_n + n_add entries (at most n_cap)
hipError_t launch_fill_dn(uint32_t* p, uint32_t v, const uint32_t* d_n, uint32_t n_add, uint64_t n_cap,
                          hipStream_t st);
// exclusive scan in place of (*d_n + n_add) elements (d_n may be null), at most
// n_cap; writes the sum of all elements to *d_total when d_total is non-null
hipError_t launch_scan(uint32_t* data, const uint32_t* d_n, uint32_t n_add, uint64_t n_cap,
                       uint32_t* d_total, const ScanScratch& s, hipStream_t st);
// Output produced in chunks of 64 work items that finish in any order: each
// chunk holds a range of a scratch taken kChunkRange entries per atomic, and
// records its count and where its entries went; once the counts are scanned,
// this copies every chunk's three columns to their place in chunk order
// (*d_chunks chunks, at most max_chunks)
constexpr uint32_t kChunkRange = 256;
hipError_t launch_chunk_copy3(const uint32_t* d_chunks, uint64_t max_chunks, const uint32_t* chunk_n,
                              const uint32_t* chunk_at, const uint32_t* scratch, uint32_t scratch_cap, uint32_t* o1,
                              uint32_t* o2, uint32_t* o3, uint32_t out_cap, DevCounts* cnt, const LaunchCfg& lc,
                              hipStream_t st);
// forward fill of zero entries of an offsets array (E*T + 1 entries, E on device)
hipError_t launch_forward_fill(uint32_t* off, const DevCounts* cnt, uint32_t n_trip, hipStream_t st);

// GridBuilder
// count, scan and stable scatter in one kernel (batched path); perm: N u32 scratch;
// rec: 2 N float4 scratch (null: none) where the items that are not staged in
// LDS leave each hit's (x, y, z, phi) and (layer, detId, hitId, event) in the
// count pass, for the scatter to read as one 32-byte record per hit
// hits of one (event, layer) that the one-kernel grid build keeps in LDS at a
// block size (0 when the grid's cells do not fit a u16 or LDS); larger layers
// take its global-memory path
constexpr uint32_t kGridLdsHitsMax = 8192;
constexpr size_t kGridLdsBase = 64 * 1024;   // LDS per block of the unstaged build
constexpr size_t kGridLdsMax = 160 * 1024;   // gfx950's LDS per workgroup
hipError_t launch_grid_build(const GridGeom& g, uint32_t n_items_cap, const uint32_t* evt_offset,
                             const uint32_t* lay_nhits, const uint32_t* lay_offset, float4* rec,
                             const HitColumns& in, const HitColumns& out, uint32_t* grid, uint32_t* perm,
                             uint32_t n_hits_cap, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st,
                             bool reset_counts = false);  // also clear the run's counts (after the shape) first
void grid_build_prepare();
hipError_t launch_grid_count(const GridGeom& g, uint32_t n_items_cap, const uint32_t* evt_offset,
                             const uint32_t* lay_nhits, const uint32_t* lay_offset, const float* hx,
                             const float* hy, const float* hz, uint32_t* grid, DevCounts* cnt,
                             const LaunchCfg& lc, hipStream_t st);
hipError_t launch_grid_store(const GridGeom& g, uint32_t n_items_cap, const uint32_t* evt_offset,
                             const uint32_t* lay_nhits, const uint32_t* lay_offset, const uint32_t* grid,
                             const HitColumns& in, const HitColumns& out, uint32_t n_hits_cap,
                             DevCounts* cnt, const LaunchCfg& lc, hipStream_t st);

// PairGeneratorBeamspot: plan -> scan(seg_base) -> count -> scan(slot_counts) -> store
// segment of every 2^kSlotSegShift-th pair slot, so that a slot's segment is
// searched among the few segments of its block of slots only
constexpr uint32_t kSlotSegShift = 10;
inline uint64_t slot_seg_entries(uint64_t n_slot_cap) { return (n_slot_cap >> kSlotSegShift) + 2; }
hipError_t launch_slot_segments(const uint32_t* seg_base, uint64_t n_slot_cap, uint32_t* slot_seg,
                                const DevCounts* cnt, hipStream_t st);
hipError_t launch_pair_plan(const GridGeom& g, const TripletCols& tc, uint32_t n_seg_cap,
                            const uint32_t* grid, uint32_t* seg_base, DevCounts* cnt, hipStream_t st);
hipError_t launch_pair_count(const GridGeom& g, const TripletCols& tc, uint32_t n_seg_cap,
                             uint32_t n_slot_cap, const uint32_t* grid, const float* hx, const float* hy,
                             const float* hz, const float2* hzp, const uint32_t* seg_base, uint32_t* slot_counts,
                             DevCounts* cnt, const LaunchCfg& lc, hipStream_t st,
                             const uint32_t* slot_seg = nullptr);  // null: search every segment
hipError_t launch_pair_store(const GridGeom& g, const TripletCols& tc, uint32_t n_seg_cap,
                             uint32_t n_slot_cap, uint32_t n_pair_cap, const uint32_t* grid,
                             const float* hx, const float* hy, const float* hz, const float2* hzp,
                             const uint32_t* seg_base, const uint32_t* slot_counts, uint32_t* pairs,
                             uint32_t* pair_offsets, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st,
                             const uint32_t* slot_seg = nullptr);

// TripletThetaPhiPredictor: count -> scan(counts) -> store (+ group offsets) -> launch_forward_fill
hipError_t launch_predict_count(const GridGeom& g, const TripletCols& tc, uint32_t n_pair_cap,
                                const uint32_t* grid, const float* hx, const float* hy, const float* hz,
                                const float2* hzp, const uint32_t* hlayer, const uint32_t* hevt,
                                const uint32_t* pairs, uint32_t* counts, DevCounts* cnt, const LaunchCfg& lc,
                                hipStream_t st);
hipError_t launch_predict_store(const GridGeom& g, const TripletCols& tc, uint32_t n_pair_cap,
                                uint32_t n_cand_cap, const uint32_t* grid, const float* hx, const float* hy,
                                const float* hz, const float2* hzp, const uint32_t* hlayer, const uint32_t* hevt,
                                const uint32_t* pairs, const uint32_t* counts, uint32_t* cands,
                                uint32_t* cand_offsets, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st);
// count + scan + store in one pass (the offsets array must be zeroed before);
// uses predict_fused_tiles(n_pair_cap) look-back words of the scan scratch
uint32_t predict_fused_tiles(uint64_t n_pair_cap);
// chunks of 64 pairs of the fused predictor + filter
uint64_t predict_fused_chunks(uint64_t n_pair_cap);
// fused predictor variants (block size, LDS list entries per pair)
bool predict_variant_supported(uint32_t block, uint32_t list, uint32_t sort);
// resolves the variant's resident-block count (call outside graph capture)
void predict_fused_prepare(const LaunchCfg& lc);
// builds the fine walk structure of every (event, layer) first, then walks it
hipError_t launch_predict_fused(const GridGeom& g, const TripletCols& tc, uint32_t n_el_cap, uint32_t n_hit_cap,
                                uint32_t n_pair_cap, uint32_t n_cand_cap, const uint32_t* grid, const float* hx,
                                const float* hy, const float* hz, const float2* hzp, const uint32_t* hlayer,
                                const uint32_t* hevt, const uint32_t* pairs, uint32_t* cands, uint32_t* cand_offsets,
                                DevCounts* cnt, const FineBufs& fb, const ScanScratch& ss, const LaunchCfg& lc,
                                hipStream_t st);

// predictor and filter fused (batched path): tracklets straight from the
// walk, candidates counted (and their group offsets) but not stored; the
// tracklet offsets come from launch_filter_offsets afterwards
// the fine walk structure (fb.grid, fb.hits) of the grid-ordered hits for the
// launch's sub-row count; the predictor launches build it themselves unless
// fb.prebuilt (a caller may build it on another stream, after the grid build)
hipError_t launch_fine_grid(const GridGeom& g, uint32_t n_el_cap, uint32_t n_hit_cap, const uint32_t* grid,
                            const float2* hzp, DevCounts* cnt, const FineBufs& fb, const LaunchCfg& lc, FineGeom& fg,
                            hipStream_t st);
hipError_t launch_predict_filter(const GridGeom& g, const TripletCols& tc, uint32_t n_el_cap, uint32_t n_hit_cap,
                                 uint32_t n_pair_cap, uint32_t n_trk_cap, const uint32_t* grid, const float* hx,
                                 const float* hy, const float* hz, const float2* hzp, const uint32_t* hlayer,
                                 const uint32_t* hevt, const uint32_t* pairs, uint32_t* cand_offsets, uint32_t* th1,
                                 uint32_t* th2, uint32_t* th3, DevCounts* cnt, const FineBufs& fb,
                                 const ScanScratch& ss, const LaunchCfg& lc, hipStream_t st);
// the fused pass's outputs placed in pair order (after launch_predict_filter on
// the same buffers): the chunk counts scanned, the tracklets copied from the
// waves' scratch ranges, the candidate-group offsets filled in
hipError_t launch_predict_filter_place(const TripletCols& tc, uint32_t n_pair_cap, uint32_t n_trk_cap,
                                       uint32_t* cand_offsets, uint32_t* th1, uint32_t* th2, uint32_t* th3,
                                       DevCounts* cnt, const FineBufs& fb, const ScanScratch& ss, const LaunchCfg& lc,
                                       hipStream_t st);

// TripletThetaPhiFilter: count (bitmap + per-word counts) -> scan -> store -> launch_filter_offsets
hipError_t launch_filter_count(const TripletCols& tc, uint32_t n_cand_cap, const float* hx,
                               const float* hy, const float* hz, const uint32_t* hlayer,
                               const uint32_t* pairs, const uint32_t* cands, uint64_t* bitmap,
                               uint32_t* word_counts, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st);
hipError_t launch_filter_store(const TripletCols& tc, uint32_t n_cand_cap, uint32_t n_trk_cap,
                               const uint32_t* hlayer, const uint32_t* hevt, const uint32_t* pairs,
                               const uint32_t* cands, const uint64_t* bitmap, const uint32_t* word_counts,
                               uint32_t* th1, uint32_t* th2, uint32_t* th3, uint32_t* t_off,
                               DevCounts* cnt, const LaunchCfg& lc, hipStream_t st);
// filterOffsetStore and its monotonize (the forward fill), over tracklets already stored
// (the offsets array zeroed before)
// filterOffsetStore alone (the monotonize is launch_forward_fill)
hipError_t launch_filter_offset_store(const TripletCols& tc, uint32_t n_trk_cap, const uint32_t* hlayer,
                                      const uint32_t* hevt, const uint32_t* th1, uint32_t* t_off, DevCounts* cnt,
                                      hipStream_t st);
hipError_t launch_filter_offsets(const TripletCols& tc, uint32_t n_trk_cap, const uint32_t* hlayer,
                                 const uint32_t* hevt, const uint32_t* th1, uint32_t* t_off, DevCounts* cnt,
                                 const LaunchCfg& lc, hipStream_t st);

// TripletConnectivityTight
struct LinkBufs {
    float* eta = nullptr;            // R
    uint32_t* hit_count = nullptr;   // n_hits + 1: tracklets per inner hit, scanned to CSR starts
    uint32_t* hit_cursor = nullptr;  // n_hits
    uint32_t* csr = nullptr;         // R
    uint32_t* conn = nullptr;        // R: connectivity predicate
    uint32_t* counts = nullptr;      // R + 1: followers per base, scanned
    // the batched path's single connectivity pass (launch_link_fused): per
    // 64-tracklet chunk its connections and where they went in the scratch
    // the tracklets come in one run per middle hit (each layer is the middle
    // layer of at most one layer triplet): run_start / run_end per hit are
    // hit_count / hit_cursor, and a base's followers are the run of its third
    // hit (in ascending index) instead of a CSR bucket
    uint32_t runs = 0;
    uint32_t* chunk_n = nullptr;     // link_chunks(R) + 1
    uint32_t* chunk_at = nullptr;    // link_chunks(R) + 1
    uint32_t* scratch = nullptr;     // 3 * scratch_cap: basis, follower, event
    uint32_t scratch_cap = 0;        // link_scratch_cap(K)
    uint32_t* basis = nullptr; uint32_t* foll = nullptr; uint32_t* pev = nullptr;  // K
    uint32_t* flag_a = nullptr; uint32_t* flag_b = nullptr;  // R + 1
    uint32_t* connectable = nullptr; uint32_t* nonconnectable = nullptr;  // R
    float* pt = nullptr;             // R
    // each event its own group (per_event != 0): the connectivity stage sees
    // every event as a run of that event alone would (its own tracklet
    // offsets, its own last tracklet left out)
    uint32_t per_event = 0;
    uint32_t* ev_first = nullptr;  // E: first tracklet of each event (sentinel: none)
    uint32_t* ev_last = nullptr;   // E: last tracklet of each event
    uint32_t* ev_has = nullptr;    // E: the event has a connection
    uint32_t* ev_off = nullptr;    // E * (T + 1): the event's own tracklet offsets
};
hipError_t launch_link_prepare(const TripletCols& tc, uint32_t n_hits_cap, uint32_t n_trk_cap,
                               const float* hx, const float* hy, const float* hz, const uint32_t* th1,
                               const uint32_t* th2, const uint32_t* th3, const LinkBufs& b, DevCounts* cnt,
                               const ScanScratch& ss, const LaunchCfg& lc, hipStream_t st);
hipError_t launch_link_count(const TripletCols& tc, uint32_t n_trk_cap, const uint32_t* hlayer,
                             const uint32_t* hevt, const uint32_t* th1, const uint32_t* th2,
                             const uint32_t* th3, const uint32_t* t_off, const LinkBufs& b, DevCounts* cnt,
                             const LaunchCfg& lc, hipStream_t st);
// count and store in one pass (batched path): connections per 64-tracklet
// chunk into the scratch, the chunk counts scanned (n_conn), then
// launch_link_fused_place copies them to basis / followers / event
uint64_t link_chunks(uint64_t n_trk_cap);
uint64_t link_scratch_cap(uint64_t n_conn_cap);
hipError_t launch_link_fused(const TripletCols& tc, uint32_t n_trk_cap, const uint32_t* hlayer, const uint32_t* hevt,
                             const uint32_t* th1, const uint32_t* th2, const uint32_t* th3, const uint32_t* t_off,
                             const LinkBufs& b, DevCounts* cnt, const ScanScratch& ss, const LaunchCfg& lc,
                             hipStream_t st);
hipError_t launch_link_fused_place(uint32_t n_trk_cap, uint32_t n_conn_cap, const LinkBufs& b, DevCounts* cnt,
                                   const LaunchCfg& lc, hipStream_t st);
hipError_t launch_link_store(const TripletCols& tc, uint32_t n_trk_cap, uint32_t n_conn_cap,
                             const uint32_t* hlayer, const uint32_t* hevt, const uint32_t* th1,
                             const uint32_t* th2, const uint32_t* th3, const uint32_t* t_off,
                             const LinkBufs& b, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st);
// each event its own group: first/last tracklet and own tracklet offsets of
// every event (only when b.per_event)
hipError_t launch_link_events(const TripletCols& tc, uint32_t n_events_cap, uint32_t n_trk_cap,
                              const uint32_t* hlayer, const uint32_t* hevt, const uint32_t* th1,
                              const uint32_t* t_off, const LinkBufs& b, DevCounts* cnt, hipStream_t st);
hipError_t launch_link_flags(uint32_t n_trk_cap, const uint32_t* hevt, const uint32_t* th1, const LinkBufs& b,
                             DevCounts* cnt, const LaunchCfg& lc, hipStream_t st);
hipError_t launch_link_compact(uint32_t n_trk_cap, const LinkBufs& b, DevCounts* cnt, const LaunchCfg& lc,
                               hipStream_t st);
hipError_t launch_circle_fit(uint32_t n_trk_cap, const float* hx, const float* hy, const uint32_t* th1,
                             const uint32_t* th2, const uint32_t* th3, const LinkBufs& b, DevCounts* cnt,
                             const LaunchCfg& lc, hipStream_t st);

// CellularAutomaton
struct CABufs {
    uint32_t* st_a = nullptr; uint32_t* st_b = nullptr;  // R
    uint32_t* fbc = nullptr;       // R + 1
    uint32_t* cursor = nullptr;    // R
    uint32_t* bidx = nullptr;      // basis capacity
    float* bdiff = nullptr;        // basis capacity
    uint32_t* best = nullptr;      // R
    uint32_t* isbest = nullptr;    // R
    uint32_t* handler = nullptr;   // R + 1
    uint32_t* tracks = nullptr;    // track-entry capacity
    uint32_t* flags = nullptr;     // 3 rotating "alive" flags
    uint8_t* clt = nullptr;        // connection capacity: the layer triplet of each connection's basis
    uint2* lt_conn = nullptr;      // connection capacity: (basis, follower) bucketed by that layer triplet
    uint32_t* lt_tab = nullptr;    // layer triplets * kCaTableBlocks + 1: per-block bucket counts, then starts
};
constexpr uint32_t kCaTableBlocks = 1024;
hipError_t launch_ca_init(uint32_t n_trk_cap, const CABufs& b, DevCounts* cnt, hipStream_t st);
// iterations first_iter .. first_iter + n_iter - 1; an iteration after the
// one that found no living cell returns at once
hipError_t launch_ca_iterations(uint32_t n_conn_cap, const uint32_t* basis, const uint32_t* foll,
                                const CABufs& b, DevCounts* cnt, uint32_t first_iter, uint32_t n_iter,
                                const LaunchCfg& lc, hipStream_t st);
// The forward phase in one pass per layer triplet instead of iterations, for
// configurations where every connection leads to a higher layer triplet
// (each triplet's second layer above its first): the same states and the
// same iteration count (the iteration that finds no living cell included).
// th1 and hlayer give each basis's layer triplet; n_trip passes.
hipError_t launch_ca_layers(uint32_t n_conn_cap, uint32_t n_trip, const uint32_t* basis, const uint32_t* foll,
                            const uint32_t* th1, const uint32_t* hlayer, const CABufs& b, DevCounts* cnt,
                            const ScanScratch& ss, const LaunchCfg& lc, hipStream_t st);
hipError_t launch_ca_basis_count(uint32_t n_trk_cap, uint32_t n_conn_cap, const uint32_t* basis,
                                 const uint32_t* foll, const CABufs& b, DevCounts* cnt, const LaunchCfg& lc,
                                 hipStream_t st);
hipError_t launch_ca_basis_store(uint32_t n_trk_cap, uint32_t n_conn_cap, uint32_t n_basis_cap,
                                 const uint32_t* basis, const uint32_t* foll, const float* pt,
                                 const CABufs& b, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st);
hipError_t launch_ca_tracks(uint32_t n_trk_cap, uint32_t n_track_cap, const CABufs& b, DevCounts* cnt,
                            const ScanScratch& ss, const LaunchCfg& lc, hipStream_t st, int part);

}  // namespace trax
