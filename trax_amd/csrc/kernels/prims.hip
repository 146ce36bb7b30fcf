// Generic primitives: exclusive u32 scan (the PrefixSum of the reference,
// src/algorithms/PrefixSum.h:58-248) and u32 fill.
//
// The reference scans with a recursive Blelloch tree (2*wg elements per
// work-group, one launch per level plus a uniform-add pass per level).  Here
// an array whose capacity is at most kRtsMaxN elements is scanned in three
// launches (reduce, then scan): the sums of 8192-element tiles, one block
// scanning those sums, then every tile scanned again from its tile's prefix
// (12 bytes per element of HBM traffic, no tile waits on another).  A larger
// one takes a single-pass decoupled look-back scan: 1024 persistent blocks of
// 512 threads take 16384-element tiles in increasing order from a counter,
// publish each tile's aggregate and look back over their predecessors (8
// bytes per element).  In both, each wave owns consecutive elements and moves
// them in rows of 64 16-byte vectors (1 KB contiguous per instruction); a row
// is scanned across the wave's lanes.  Measured on MI355X: reduce-then-scan
// is the faster of the two up to about 32 M elements, look-back beyond.  The
// result is the same exclusive running sum modulo 2^32 either way.
#include "kernels.h"

namespace trax {
namespace {

constexpr int kScanItems = 32;               // elements per thread in both forms
constexpr int kRtsBlock = 256;
constexpr uint32_t kRtsTile = kRtsBlock * kScanItems;
constexpr int kLbBlock = 512;
constexpr uint32_t kLbTile = kLbBlock * kScanItems;
constexpr uint32_t kLbBlocks = 1024;         // persistent look-back blocks
constexpr uint64_t kRtsMaxN = 1ull << 25;

__device__ inline uint32_t scan_size(const uint32_t* d_n, uint32_t n_add, uint32_t n_cap) {
    const uint32_t n = (d_n ? *d_n : 0u) + n_add;
    return n > n_cap ? n_cap : n;  // a size past its buffer is reported by the stage kernels
}

// One wave's segment of kWave*kScanItems elements from seg: row q, lane l is
// the 16-byte vector q*64+l (vector loads when vec; element loads, bounded by
// n, otherwise)
template <int R>
__device__ inline void seg_load(const uint32_t* data, uint64_t seg, uint32_t n, bool vec, uint4 (&a)[R]) {
    const uint32_t lane = lane_id();
    if (vec) {
        const uint4* p = reinterpret_cast<const uint4*>(data + seg);
#pragma unroll
        for (int q = 0; q < R; ++q) a[q] = p[q * kWave + lane];
    } else {
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const uint64_t e = seg + 4ull * (q * kWave + lane);
            a[q].x = e < n ? data[e] : 0u;
            a[q].y = e + 1 < n ? data[e + 1] : 0u;
            a[q].z = e + 2 < n ? data[e + 2] : 0u;
            a[q].w = e + 3 < n ? data[e + 3] : 0u;
        }
    }
}
template <int R>
__device__ inline void seg_store(uint32_t* data, uint64_t seg, uint32_t n, bool vec, const uint4 (&a)[R]) {
    const uint32_t lane = lane_id();
    if (vec) {
        uint4* p = reinterpret_cast<uint4*>(data + seg);
#pragma unroll
        for (int q = 0; q < R; ++q) p[q * kWave + lane] = a[q];
    } else {
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const uint64_t e = seg + 4ull * (q * kWave + lane);
            if (e < n) data[e] = a[q].x;
            if (e + 1 < n) data[e + 1] = a[q].y;
            if (e + 2 < n) data[e + 2] = a[q].z;
            if (e + 3 < n) data[e + 3] = a[q].w;
        }
    }
}
// exclusive prefix of each row within the wave's segment; returns the segment sum
template <int R>
__device__ inline uint32_t seg_scan(const uint4 (&a)[R], uint32_t (&rex)[R]) {
    uint32_t run = 0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const uint32_t v = a[q].x + a[q].y + a[q].z + a[q].w;
        const uint32_t inc = wave_incl_scan(v);
        rex[q] = run + inc - v;
        run += __shfl(inc, kWave - 1, kWave);
    }
    return run;
}
// the segment's elements replaced by their exclusive running sums from p
template <int R>
__device__ inline void seg_apply(uint4 (&a)[R], const uint32_t (&rex)[R], uint32_t p) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
        uint32_t r = p + rex[q];
        uint4 o;
        o.x = r; r += a[q].x;
        o.y = r; r += a[q].y;
        o.z = r; r += a[q].z;
        o.w = r;
        a[q] = o;
    }
}

// look-back form: persistent blocks, tiles in increasing order from counter
template <bool kVec>
__global__ __launch_bounds__(kLbBlock) void k_scan_lookback(uint32_t* __restrict__ data,
                                                            const uint32_t* __restrict__ d_n, uint32_t n_add,
                                                            uint32_t* __restrict__ d_total,
                                                            uint64_t* __restrict__ status,
                                                            uint32_t* __restrict__ counter, uint32_t n_cap) {
    constexpr int R = kScanItems / 4;
    constexpr int kWaves = kLbBlock / kWave;
    __shared__ uint32_t s_wp[kWaves];
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x, w = tid / kWave;
    const uint32_t n = scan_size(d_n, n_add, n_cap);
    while (true) {
        __syncthreads();  // s_tile and s_wp of the previous tile are consumed
        if (tid == 0) s_tile = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        const uint64_t base = (uint64_t)tile * kLbTile;
        if (base >= n) {
            if (tile == 0 && tid == 0 && d_total) *d_total = 0;
            return;
        }
        const uint64_t seg = base + (uint64_t)w * (kWave * kScanItems);
        const bool vec = kVec && base + kLbTile <= n;
        uint4 a[R];
        uint32_t rex[R];
        seg_load(data, seg, n, vec, a);
        const uint32_t wsum = seg_scan(a, rex);
        if (lane_id() == kWave - 1) s_wp[w] = wsum;
        __syncthreads();
        if (tid < kWave) {
            const uint32_t wv = tid < kWaves ? s_wp[tid] : 0u;
            const uint32_t winc = wave_incl_scan(wv);
            const uint32_t total = __shfl(winc, kWaves - 1, kWave);
            const uint32_t excl = tile_lookback(status, tile, total);
            if (tid < kWaves) s_wp[tid] = excl + winc - wv;
            if (d_total && tid == 0 && base + kLbTile >= n) *d_total = excl + total;
        }
        __syncthreads();
        seg_apply(a, rex, s_wp[w]);
        seg_store(data, seg, n, vec, a);
    }
}

// reduce-then-scan form: 1. tile sums
template <bool kVec>
__global__ __launch_bounds__(kRtsBlock) void k_scan_tile_sums(const uint32_t* __restrict__ data,
                                                              const uint32_t* __restrict__ d_n, uint32_t n_add,
                                                              uint32_t* __restrict__ sums, uint32_t n_cap) {
    constexpr int R = kScanItems / 4;
    __shared__ uint32_t s_wsum[kRtsBlock / kWave + 1];
    const uint32_t n = scan_size(d_n, n_add, n_cap);
    const uint64_t base = (uint64_t)blockIdx.x * kRtsTile;
    if (base >= n) return;  // block-uniform
    uint32_t sum = 0;
    if (kVec && base + kRtsTile <= n) {
        const uint4* p = reinterpret_cast<const uint4*>(data + base);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const uint4 a = p[q * kRtsBlock + threadIdx.x];
            sum += a.x + a.y + a.z + a.w;
        }
    } else {
        const uint64_t end = base + kRtsTile < n ? base + kRtsTile : n;
        for (uint64_t e = base + threadIdx.x; e < end; e += kRtsBlock) sum += data[e];
    }
    uint32_t total;
    block_excl_scan(sum, s_wsum, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// 2. exclusive scan of the tile sums in one block; the total to d_total
__global__ __launch_bounds__(1024) void k_scan_sums(uint32_t* __restrict__ sums, const uint32_t* __restrict__ d_n,
                                                    uint32_t n_add, uint32_t* __restrict__ d_total, uint32_t n_cap) {
    __shared__ uint32_t s_wsum[1024 / kWave + 1];
    const uint32_t n = scan_size(d_n, n_add, n_cap);
    const uint32_t m = (uint32_t)(((uint64_t)n + kRtsTile - 1) / kRtsTile);
    uint32_t carry = 0;
    for (uint32_t b = 0; b < m; b += blockDim.x) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < m ? sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_excl_scan(v, s_wsum, &total);
        if (i < m) sums[i] = carry + ex;
        carry += total;
        __syncthreads();  // s_wsum of this round is consumed
    }
    if (d_total && threadIdx.x == 0) *d_total = carry;
}

// 3. every tile scanned from its prefix
template <bool kVec>
__global__ __launch_bounds__(kRtsBlock) void k_scan_down(uint32_t* __restrict__ data,
                                                         const uint32_t* __restrict__ d_n, uint32_t n_add,
                                                         const uint32_t* __restrict__ sums, uint32_t n_cap) {
    constexpr int R = kScanItems / 4;
    constexpr int kWaves = kRtsBlock / kWave;
    __shared__ uint32_t s_wp[kWaves];
    const uint32_t tid = threadIdx.x, w = tid / kWave;
    const uint32_t n = scan_size(d_n, n_add, n_cap);
    const uint64_t base = (uint64_t)blockIdx.x * kRtsTile;
    if (base >= n) return;  // block-uniform
    const uint64_t seg = base + (uint64_t)w * (kWave * kScanItems);
    const bool vec = kVec && base + kRtsTile <= n;
    uint4 a[R];
    uint32_t rex[R];
    seg_load(data, seg, n, vec, a);
    const uint32_t wsum = seg_scan(a, rex);
    if (lane_id() == kWave - 1) s_wp[w] = wsum;
    __syncthreads();
    if (tid < kWave) {
        const uint32_t wv = tid < kWaves ? s_wp[tid] : 0u;
        const uint32_t winc = wave_incl_scan(wv);
        if (tid < kWaves) s_wp[tid] = sums[blockIdx.x] + winc - wv;
    }
    __syncthreads();
    seg_apply(a, rex, s_wp[w]);
    seg_store(data, seg, n, vec, a);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_copy_u32(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// event column of a batch of whole events: hit i of event e gets e, for the
// n_events (device) events whose hits start at evt_offset[e] (running sums)
__global__ void k_event_ids(uint32_t* __restrict__ evt, const uint32_t* __restrict__ evt_offset,
                            const DevCounts* __restrict__ cnt) {
    const uint32_t E = cnt->n_events, N = cnt->n_hits;
    for (uint32_t e = blockIdx.x; e < E; e += gridDim.x) {
        const uint32_t lo = evt_offset[e];
        const uint32_t hi = e + 1 < E ? evt_offset[e + 1] : N;
        for (uint32_t i = lo + threadIdx.x; i < hi && i < N; i += blockDim.x) evt[i] = e;
    }
}

__global__ void k_fill_u32_dn(uint32_t* __restrict__ p, uint32_t v, const uint32_t* __restrict__ d_n,
                              uint32_t n_add, uint64_t n_cap) {
    uint64_t n = (uint64_t)*d_n + n_add;
    if (n > n_cap) n = n_cap;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Sequential forward fill of zero entries ("monotonize", TripletThetaPhiPredictor.h:
// 432-447, TripletThetaPhiFilter.h:267-280, read in program order): one block
// scans with the operator (a, b) -> b != 0 ? b : a.
__global__ __launch_bounds__(256) void k_forward_fill(uint32_t* __restrict__ off,
                                                     const DevCounts* __restrict__ cnt, uint32_t n_trip) {
    __shared__ uint32_t s_last[256 / kWave];
    __shared__ uint32_t s_carry;
    const uint32_t n = cnt->n_events * n_trip + 1;
    const uint32_t tid = threadIdx.x, w = tid / kWave, l = lane_id();
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += blockDim.x) {
        const uint32_t i = base + tid;
        uint32_t v = i < n ? off[i] : 0u;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const uint32_t o = __shfl_up(v, d, kWave);
            if (l >= (uint32_t)d && v == 0u) v = o;
        }
        if (l == kWave - 1) s_last[w] = v;
        __syncthreads();
        uint32_t carry = s_carry;
        for (uint32_t k = 0; k < w; ++k)
            if (s_last[k] != 0u) carry = s_last[k];
        if (v == 0u) v = carry;
        if (i < n) off[i] = v;
        __syncthreads();
        if (tid == blockDim.x - 1) s_carry = v;
        __syncthreads();
    }
}

__global__ void k_set_counts(DevCounts* __restrict__ cnt, CountWrites w) {
    uint32_t* c = reinterpret_cast<uint32_t*>(cnt);
    if (threadIdx.x == 0)
        for (uint32_t i = 0; i < w.n; ++i) c[w.idx[i]] = w.val[i];
}

}  // namespace

hipError_t launch_set_counts(DevCounts* cnt, const CountWrites& w, hipStream_t st) {
    if (w.n > 12) return hipErrorInvalidValue;
    for (uint32_t i = 0; i < w.n; ++i)
        if (w.idx[i] >= sizeof(DevCounts) / 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_set_counts, dim3(1), dim3(64), 0, st, cnt, w);
    return hipGetLastError();
}

uint32_t stride_blocks(uint64_t work_cap, uint32_t block, const LaunchCfg& lc) {
    static uint32_t n_cu = 0;
    if (n_cu == 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n_cu = (uint32_t)v;
        else
            n_cu = 256;
    }
    const uint64_t per_cu = lc.blocks_per_cu ? lc.blocks_per_cu : 8;
    uint64_t want = (work_cap + block - 1) / block;
    const uint64_t mx = per_cu * n_cu;
    if (want > mx) want = mx;
    if (want < 1) want = 1;
    return (uint32_t)want;
}

namespace {

// each chunk's three output columns from a scratch to their place in chunk
// order (chunk_n scanned: a chunk's count is the next entry minus its own;
// chunk_at: where its entries are in the scratch).  Chunks hold a dozen or so
// entries, so a wave takes kCopyGroup consecutive chunks at a time, whose
// outputs are one contiguous range: its lanes write that range in order, each
// finding its entry's chunk among the group's (one chunk's first output per
// lane) by a binary search over lanes.  Writes are coalesced, and so are the
// reads within a chunk; groups of 16 keep enough waves busy when there are
// few chunks.
constexpr uint32_t kCopyGroup = 16;
__global__ void k_chunk_copy3(const uint32_t* __restrict__ d_chunks, const uint32_t* __restrict__ chunk_n,
                              const uint32_t* __restrict__ chunk_at, const uint32_t* __restrict__ scratch,
                              uint32_t scratch_cap, uint32_t* __restrict__ o1, uint32_t* __restrict__ o2,
                              uint32_t* __restrict__ o3, uint32_t out_cap, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;  // the chunks' ranges may not all be written
    const uint32_t n_chunks = *d_chunks;
    const uint32_t lane = lane_id();
    const uint32_t n_waves = gridDim.x * blockDim.x / kWave;
    bool over = false;
    for (uint32_t g0 = (blockIdx.x * blockDim.x + threadIdx.x) / kWave * kCopyGroup; g0 < n_chunks;
         g0 += n_waves * kCopyGroup) {  // wave-uniform
        const uint32_t nc = n_chunks - g0 < kCopyGroup ? n_chunks - g0 : kCopyGroup;
        const uint32_t t = g0 + (lane < nc ? lane : 0u);
        const uint32_t my_dst = chunk_n[t], my_src = chunk_at[t];
        const uint32_t d0 = __shfl(my_dst, 0, kWave), d1 = chunk_n[g0 + nc];
        for (uint32_t base = d0; base < d1; base += kWave) {  // wave-uniform
            const uint32_t d = base + lane;
            // the last of the group's chunks whose first output is at most d
            uint32_t c = 0;
#pragma unroll
            for (uint32_t step = kCopyGroup / 2; step; step >>= 1) {
                const uint32_t cand = c + step;
                const uint32_t v = __shfl(my_dst, cand < nc ? cand : 0u, kWave);
                if (cand < nc && v <= d) c = cand;
            }
            const uint32_t c_dst = __shfl(my_dst, c, kWave), c_src = __shfl(my_src, c, kWave);
            if (d >= d1) continue;
            const uint32_t k = d - c_dst;
            if ((uint64_t)d >= out_cap) {
                over = true;
            } else if ((uint64_t)c_src + k < scratch_cap) {
                o1[d] = scratch[c_src + k];
                o2[d] = scratch[(size_t)scratch_cap + c_src + k];
                o3[d] = scratch[2 * (size_t)scratch_cap + c_src + k];
            } else {  // never left unwritten below the output's capacity
                o1[d] = o2[d] = o3[d] = kSentinel;
                over = true;
            }
        }
    }
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

}  // namespace

hipError_t launch_chunk_copy3(const uint32_t* d_chunks, uint64_t max_chunks, const uint32_t* chunk_n,
                              const uint32_t* chunk_at, const uint32_t* scratch, uint32_t scratch_cap, uint32_t* o1,
                              uint32_t* o2, uint32_t* o3, uint32_t out_cap, DevCounts* cnt, const LaunchCfg& lc,
                              hipStream_t st) {
    hipLaunchKernelGGL(k_chunk_copy3, dim3(stride_blocks(max_chunks * (kWave / kCopyGroup), 256, lc)), dim3(256), 0,
                       st, d_chunks,
                       chunk_n, chunk_at, scratch, scratch_cap, o1, o2, o3, out_cap, cnt);
    return hipGetLastError();
}

hipError_t launch_forward_fill(uint32_t* off, const DevCounts* cnt, uint32_t n_trip, hipStream_t st) {
    hipLaunchKernelGGL(k_forward_fill, dim3(1), dim3(256), 0, st, off, cnt, n_trip);
    return hipGetLastError();
}

// status words (u64) the scan needs at a capacity: tile sums of the
// reduce-then-scan form, or the look-back form's tile status and counter
static_assert(kRtsTile == kScanTileElements, "the scan's status words are counted in reduce-then-scan tiles");
uint32_t scan_tiles(uint64_t n_cap) { return (uint32_t)((n_cap + kRtsTile - 1) / kRtsTile) + 2; }

hipError_t launch_fill(uint32_t* p, uint32_t v, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)blocks), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_copy(uint32_t* dst, const uint32_t* src, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_copy_u32, dim3((uint32_t)blocks), dim3(256), 0, st, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_event_ids(uint32_t* evt, const uint32_t* evt_offset, const DevCounts* cnt, uint32_t n_events_cap,
                            hipStream_t st) {
    const uint32_t blocks = n_events_cap < 1024 ? (n_events_cap ? n_events_cap : 1) : 1024;
    hipLaunchKernelGGL(k_event_ids, dim3(blocks), dim3(256), 0, st, evt, evt_offset, cnt);
    return hipGetLastError();
}

hipError_t launch_fill_dn(uint32_t* p, uint32_t v, const uint32_t* d_n, uint32_t n_add, uint64_t n_cap,
                          hipStream_t st) {
    uint64_t blocks = (n_cap + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32_dn, dim3((uint32_t)blocks), dim3(256), 0, st, p, v, d_n, n_add, n_cap);
    return hipGetLastError();
}

hipError_t launch_scan(uint32_t* data, const uint32_t* d_n, uint32_t n_add, uint64_t n_cap,
                       uint32_t* d_total, const ScanScratch& s, hipStream_t st) {
    const uint32_t tiles = scan_tiles(n_cap);
    if (tiles > s.max_tiles) return hipErrorInvalidValue;
    const uint32_t cap = (uint32_t)(n_cap < 0xffffffffull ? n_cap : 0xffffffffull);
    const bool vec = (reinterpret_cast<uintptr_t>(data) & 15u) == 0;
    if (n_cap <= kRtsMaxN) {
        const uint32_t blocks = std::max<uint32_t>(1u, (uint32_t)((n_cap + kRtsTile - 1) / kRtsTile));
        uint32_t* sums = reinterpret_cast<uint32_t*>(s.status);
        if (vec)
            hipLaunchKernelGGL(k_scan_tile_sums<true>, dim3(blocks), dim3(kRtsBlock), 0, st, data, d_n, n_add, sums, cap);
        else
            hipLaunchKernelGGL(k_scan_tile_sums<false>, dim3(blocks), dim3(kRtsBlock), 0, st, data, d_n, n_add, sums,
                               cap);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, st, sums, d_n, n_add, d_total, cap);
        if (vec)
            hipLaunchKernelGGL(k_scan_down<true>, dim3(blocks), dim3(kRtsBlock), 0, st, data, d_n, n_add, sums, cap);
        else
            hipLaunchKernelGGL(k_scan_down<false>, dim3(blocks), dim3(kRtsBlock), 0, st, data, d_n, n_add, sums, cap);
        return hipGetLastError();
    }
    // status words and the tile counter are cleared by a kernel, not a
    // memset node, so that the sequence can be captured into a graph
    const uint32_t lb_tiles = (uint32_t)((n_cap + kLbTile - 1) / kLbTile) + 1;
    hipError_t e = launch_fill(reinterpret_cast<uint32_t*>(s.status), 0u, 2ull * lb_tiles + 2, st);
    if (e != hipSuccess) return e;
    const uint32_t blocks = std::min<uint32_t>(lb_tiles, kLbBlocks);
    uint32_t* counter = reinterpret_cast<uint32_t*>(s.status + lb_tiles);
    if (vec)
        hipLaunchKernelGGL(k_scan_lookback<true>, dim3(blocks), dim3(kLbBlock), 0, st, data, d_n, n_add, d_total,
                           s.status, counter, cap);
    else
        hipLaunchKernelGGL(k_scan_lookback<false>, dim3(blocks), dim3(kLbBlock), 0, st, data, d_n, n_add, d_total,
                           s.status, counter, cap);
    return hipGetLastError();
}

}  // namespace trax
