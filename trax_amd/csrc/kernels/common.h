// Device-side helpers shared by the gfx950 kernels (wave64 primitives, LDS
// block scans) and the device-resident size record every stage reads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../trax_math.h"

namespace trax {

constexpr uint32_t kSentinel = 0xffffffffu;
constexpr int kWave = 64;

// error bits set on the device (DevCounts::err)
constexpr uint32_t kErrCellRange = 1u;     // a hit maps outside the z-phi grid
constexpr uint32_t kErrCapacity = 2u;      // an output exceeded its preallocated capacity
constexpr uint32_t kErrIndexRange = 4u;    // an index read from an input is out of range

// Device-resident sizes of one pipeline run.  Stage kernels read their work
// size from here, so that one launch sequence (and one captured graph) serves
// any batch within the preallocated capacities.
struct DevCounts {
    uint32_t n_events;       // batch shape, written by the host before a run
    uint32_t n_hits;
    // capacities of the output buffers (0xffffffff: sized exactly by the host)
    uint32_t cap_slots, cap_pairs, cap_cands, cap_tracklets, cap_conn, cap_basis, cap_tracks;
    // sizes and flags written by the kernels, cleared before every run
    uint32_t n_segments;     // events * layer triplets
    uint32_t n_filter_words; // ceil(candidates / 64)
    uint32_t n_pair_slots;   // pair count slots (nThreads-major layout)
    uint32_t n_pairs;        // P
    uint32_t n_cands;        // Q
    uint32_t n_tracklets;    // R
    uint32_t n_conn;         // K, (basis, follower) connections
    uint32_t n_connectable;
    uint32_t n_nonconnectable;
    uint32_t n_basis;        // CA backward: basis entries
    uint32_t n_track_entries;
    uint32_t alive;          // CA living cells of the last iteration
    uint32_t ca_iterations;
    uint32_t err;
    uint32_t diag_pair_short;   // pair slots counted but not written
    uint32_t diag_pred_short;   // candidate slots counted but not written
    uint32_t diag_pred_trunc;   // pairs whose store found more than counted
    uint32_t diag_wrap_overlap; // outer hits whose wrapped phi window overlaps itself
    uint32_t trk_alloc;         // fused predictor + filter: tracklet scratch handed out to tiles
    uint32_t n_pred_chunks;     // fused predictor + filter: 64-pair chunks of the run
    uint32_t link_alloc;        // single-pass connectivity: connection scratch handed out to chunks
    uint32_t n_link_chunks;     // single-pass connectivity: 64-tracklet chunks of the run
    uint32_t pad[1];
};
constexpr uint32_t kCountsShapeWords = 9;  // n_events, n_hits and the seven capacities

// a device-resident size bounded by its buffer's capacity; an excess is
// reported through the error bits and the work is cut to the capacity
__device__ inline uint32_t bounded(uint32_t v, uint32_t cap, DevCounts* cnt) {
    if (v > cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&cnt->err, kErrCapacity);
        return cap;
    }
    return v;
}

// Error bits that void a run: an output cut at its capacity, a hit outside
// the grid or a bad index leaves entries that later stages must not read as
// addresses.  Every stage kernel that reads an earlier stage's output tests
// this at entry and does no work once a kernel before it has set one of them
// (kernels on one stream run in order, so the bits an earlier kernel set are
// visible).  The host reports the error and discards the run.
constexpr uint32_t kErrVoid = kErrCellRange | kErrCapacity | kErrIndexRange;
__device__ inline uint32_t load_err(const DevCounts* cnt) {
    return __hip_atomic_load(&cnt->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wave-uniform; for kernels without block barriers
__device__ inline bool run_void(const DevCounts* cnt) {
    return (__builtin_amdgcn_readfirstlane(load_err(cnt)) & kErrVoid) != 0u;
}
// block-uniform (one read by thread 0, shared through LDS); contains __syncthreads()
__device__ inline bool block_run_void(const DevCounts* cnt) {
    __shared__ uint32_t s_void;
    if (threadIdx.x == 0) s_void = load_err(cnt) & kErrVoid;
    __syncthreads();
    const bool v = s_void != 0u;
    __syncthreads();
    return v;
}

// An index read from device memory, checked against the size of the array it
// addresses before it is used as an address: false (and kErrIndexRange) when
// it is out of range, and the work item is skipped.
__device__ inline bool index_ok(uint32_t i, uint32_t n, DevCounts* cnt) {
    if (i < n) return true;
    atomicOr(&cnt->err, kErrIndexRange);
    return false;
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

// makes a wave's LDS writes visible to its other lanes (for waves that run
// their own loops, without a block barrier)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A block index for grid-stride loops that gives each XCD a contiguous slice
// of every stride: blocks are dispatched to the xcds XCDs round-robin
// (block b to XCD b % xcds), and block b takes the slice position
// (b % xcds) * (G / xcds) + b / xcds of the G blocks, so work items next to
// each other, which read the same data, run on one XCD and share its L2.  A
// bijection of [0, G) (the last G % xcds blocks keep their index); xcds <= 1
// keeps every index.
__device__ inline uint32_t xcd_block(uint32_t xcds) {
    const uint32_t b = blockIdx.x, G = gridDim.x;
    if (xcds <= 1) return b;
    const uint32_t full = G - G % xcds;
    if (b >= full) return b;
    return (b % xcds) * (full / xcds) + b / xcds;
}

__device__ inline uint64_t lanemask_lt() {
    const uint32_t l = lane_id();
    return (l == 0) ? 0ull : ((~0ull) >> (64 - l));
}

// inclusive wave scan of u32
__device__ inline uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, kWave);
        if (l >= (uint32_t)d) v += o;
    }
    return v;
}

__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Exclusive scan of one u32 per thread over the whole block.  `wsum` is LDS
// scratch of blockDim/64 + 1 entries.  Returns the exclusive prefix; *total
// receives the block sum.  Contains __syncthreads().
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const uint32_t tid = threadIdx.x, w = tid / kWave, nw = blockDim.x / kWave;
    const uint32_t inc = wave_incl_scan(v);
    if (lane_id() == kWave - 1) wsum[w] = inc;
    __syncthreads();
    if (tid < kWave) {  // the first wave scans the block's wave sums (at most 16)
        const uint32_t x = tid < nw ? wsum[tid] : 0u;
        const uint32_t xi = wave_incl_scan(x);
        if (tid < nw) wsum[tid] = xi - x;
        if (tid + 1 == nw) wsum[nw] = xi;
    }
    __syncthreads();
    const uint32_t ex = wsum[w] + inc - v;
    *total = wsum[nw];
    __syncthreads();
    return ex;
}

// Exclusive scan of n entries held in LDS, in place, by the whole block.
// Returns the total.  Contains __syncthreads().
__device__ inline uint32_t lds_excl_scan(uint32_t* a, uint32_t n, uint32_t* wsum) {
    const uint32_t tid = threadIdx.x, bd = blockDim.x;
    const uint32_t per = (n + bd - 1) / bd;
    const uint32_t lo = tid * per < n ? tid * per : n;
    const uint32_t hi = lo + per < n ? lo + per : n;
    uint32_t s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += a[i];
    uint32_t total;
    uint32_t run = block_excl_scan(s, wsum, &total);
    for (uint32_t i = lo; i < hi; ++i) {
        const uint32_t t = a[i];
        a[i] = run;
        run += t;
    }
    __syncthreads();
    return total;
}

__device__ inline uint64_t atomic_load_u64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void atomic_store_u64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back over tiles processed in increasing order (tile ids come
// from an atomic counter, so every predecessor of a running tile has been
// started).  Called by all 64 lanes of one wave; publishes the tile's
// aggregate, returns the exclusive prefix of the tile and publishes its
// inclusive prefix.  status words: (flag << 32) | value, flag 1 = aggregate,
// 2 = inclusive prefix.  Each step reads the 64 nearest unread predecessors
// and waits only for those up to the nearest inclusive prefix.
__device__ inline uint32_t tile_lookback(uint64_t* status, uint32_t tile, uint32_t total) {
    const uint32_t lane = lane_id();
    if (tile == 0) {
        if (lane == 0) atomic_store_u64(&status[0], (2ull << 32) | total);
        return 0u;
    }
    if (lane == 0) atomic_store_u64(&status[tile], (1ull << 32) | total);
    uint32_t excl = 0;
    int64_t j = (int64_t)tile - 1 - (int64_t)lane;
    while (true) {
        uint64_t s, incl;
        uint32_t flag;
        while (true) {
            s = (j >= 0) ? atomic_load_u64(&status[j]) : (2ull << 32);
            flag = (uint32_t)(s >> 32);
            incl = __ballot(flag == 2u);
            // only the tiles up to the nearest inclusive prefix are needed
            const uint64_t need = incl ? ((incl & (~incl + 1ull)) << 1) - 1ull : ~0ull;
            if ((__ballot(flag == 0u) & need) == 0ull) break;
            __builtin_amdgcn_s_sleep(1);
        }
        const uint32_t stop = incl ? (uint32_t)(__ffsll((unsigned long long)incl) - 1) : 63u;
        excl += wave_sum(lane <= stop ? (uint32_t)s : 0u);
        if (incl) break;
        j -= kWave;
    }
    if (lane == 0) atomic_store_u64(&status[tile], (2ull << 32) | (uint32_t)(excl + total));
    return excl;
}

// upper_bound - 1 over a sorted u32 array: last index s with a[s] <= key
__device__ inline uint32_t find_segment_in(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t key) {
    while (hi - lo > 1) {  // search in [lo, hi), a[lo] <= key
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] <= key) lo = mid; else hi = mid;
    }
    return lo;
}
__device__ inline uint32_t find_segment(const uint32_t* a, uint32_t n_seg, uint32_t key) {
    return find_segment_in(a, 0, n_seg, key);
}

// grid geometry (GridConfig, src/datastructures/Grid.h:5-39)
struct GridGeom {
    float min_z, sector_z;
    uint32_t n_z;
    float min_phi, sector_phi;
    uint32_t n_phi;
    uint32_t n_layers;
    uint32_t cells;  // (n_z + 1) * (n_phi + 1)
};

// cell of a hit, gridCount (src/algorithms/GridBuilder.h:80-84); kSentinel when
// the z or phi sector falls outside the (n_z + 1) x (n_phi + 1) grid, where the
// reference would index its local histogram out of bounds
__device__ inline uint32_t hit_cell_phi(const GridGeom& g, float z, float phi) {
    const uint32_t zs = tx_f2u_sat(floorf((z - g.min_z) / g.sector_z));
    const uint32_t ps = tx_f2u_sat(floorf((phi - g.min_phi) / g.sector_phi));
    if (zs > g.n_z || ps > g.n_phi) return kSentinel;
    return zs * (g.n_phi + 1) + ps;
}
__device__ inline uint32_t hit_cell(const GridGeom& g, float x, float y, float z) {
    return hit_cell_phi(g, z, tx_atan2f(y, x));
}

}  // namespace trax
