// Generic primitives: exclusive u32 scan (the PrefixSum of the reference,
// src/algorithms/PrefixSum.h:58-248) and u32 fill.
//
// The reference scans with a recursive Blelloch tree (2*wg elements per
// work-group, one launch per level plus a uniform-add pass per level).  Here
// the scan is a single-pass decoupled look-back scan: each 256-thread block
// owns a tile of 2048 elements, publishes its aggregate, and wave 0 looks back
// over up to 64 predecessor tiles at a time.  Every element is read once and
// written once (8 bytes per element of HBM traffic).  The result is the same
// exclusive running sum modulo 2^32.
#include "kernels.h"

namespace trax {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ inline uint32_t pad_idx(uint32_t i) { return i + (i >> 5); }

__global__ __launch_bounds__(kScanBlock) void k_scan_excl(uint32_t* __restrict__ data,
                                                        const uint32_t* __restrict__ d_n, uint32_t n_add,
                                                        uint32_t* __restrict__ d_total,
                                                        uint64_t* __restrict__ status,
                                                        uint32_t* __restrict__ counter, uint32_t n_cap) {
    __shared__ uint32_t s_items[kScanTile + kScanTile / 32];
    __shared__ uint32_t s_wsum[kScanBlock / kWave + 1];
    __shared__ uint32_t s_tile, s_prefix;
    const uint32_t tid = threadIdx.x;
    uint32_t n = (d_n ? *d_n : 0u) + n_add;
    if (n > n_cap) n = n_cap;  // a size past its buffer is reported by the stage kernels
    // persistent: tiles are taken in increasing order from the counter, so a
    // tile's look-back only waits on tiles held by running blocks
    while (true) {
    __syncthreads();  // s_items and s_prefix of the previous tile are consumed
    if (tid == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t base = (uint64_t)tile * kScanTile;
    if (base >= n) {
        if (tile == 0 && tid == 0 && d_total) *d_total = 0;
        return;
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t li = k * kScanBlock + tid;
        const uint64_t i = base + li;
        s_items[pad_idx(li)] = i < n ? data[i] : 0u;
    }
    __syncthreads();
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = s_items[pad_idx(tid * kScanItems + k)];
        sum += v[k];
    }
    uint32_t total;
    const uint32_t ex = block_excl_scan(sum, s_wsum, &total);
    if (tid < kWave) {
        uint32_t excl = 0;
        if (tile == 0) {
            if (tid == 0) atomic_store_u64(&status[0], (2ull << 32) | total);
        } else {
            if (tid == 0) atomic_store_u64(&status[tile], (1ull << 32) | total);
            int64_t j = (int64_t)tile - 1 - (int64_t)tid;
            while (true) {
                uint64_t s;
                uint32_t flag;
                while (true) {
                    s = (j >= 0) ? atomic_load_u64(&status[j]) : (2ull << 32);
                    flag = (uint32_t)(s >> 32);
                    if (__ballot(flag == 0u) == 0ull) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                const uint64_t incl = __ballot(flag == 2u);
                const uint32_t stop = incl ? (uint32_t)(__ffsll((unsigned long long)incl) - 1) : 63u;
                excl += wave_sum(tid <= stop ? (uint32_t)s : 0u);
                if (incl) break;
                j -= kWave;
            }
            if (tid == 0) atomic_store_u64(&status[tile], (2ull << 32) | (uint32_t)(excl + total));
        }
        if (tid == 0) s_prefix = excl;
    }
    __syncthreads();
    uint32_t run = s_prefix + ex;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        s_items[pad_idx(tid * kScanItems + k)] = run;
        run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t li = k * kScanBlock + tid;
        const uint64_t i = base + li;
        if (i < n) data[i] = s_items[pad_idx(li)];
    }
    if (d_total && tid == 0 && base + kScanTile >= n) *d_total = s_prefix + total;
    }
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

}  // namespace

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
// order: one wave per chunk (chunk_n scanned: a chunk's count is the next
// entry minus its own; chunk_at: where its entries are in the scratch)
__global__ void k_chunk_copy3(const uint32_t* __restrict__ d_chunks, const uint32_t* __restrict__ chunk_n,
                              const uint32_t* __restrict__ chunk_at, const uint32_t* __restrict__ scratch,
                              uint32_t scratch_cap, uint32_t* __restrict__ o1, uint32_t* __restrict__ o2,
                              uint32_t* __restrict__ o3, uint32_t out_cap, DevCounts* __restrict__ cnt) {
    const uint32_t n_chunks = *d_chunks;
    const uint32_t lane = lane_id();
    bool over = false;
    for (uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) / kWave; t < n_chunks;
         t += gridDim.x * blockDim.x / kWave) {
        const uint32_t dst = chunk_n[t], m = chunk_n[t + 1] - dst, src = chunk_at[t];
        for (uint32_t k = lane; k < m; k += kWave) {
            if ((uint64_t)src + k < scratch_cap && (uint64_t)dst + k < out_cap) {
                o1[dst + k] = scratch[src + k];
                o2[dst + k] = scratch[(size_t)scratch_cap + src + k];
                o3[dst + k] = scratch[2 * (size_t)scratch_cap + src + k];
            } else {
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
    hipLaunchKernelGGL(k_chunk_copy3, dim3(stride_blocks(max_chunks * kWave, 256, lc)), dim3(256), 0, st, d_chunks,
                       chunk_n, chunk_at, scratch, scratch_cap, o1, o2, o3, out_cap, cnt);
    return hipGetLastError();
}

hipError_t launch_forward_fill(uint32_t* off, const DevCounts* cnt, uint32_t n_trip, hipStream_t st) {
    hipLaunchKernelGGL(k_forward_fill, dim3(1), dim3(256), 0, st, off, cnt, n_trip);
    return hipGetLastError();
}

uint32_t scan_tiles(uint64_t n_cap) { return (uint32_t)((n_cap + kScanTile - 1) / kScanTile) + 1; }

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
    // status words and the tile counter are cleared by a kernel, not a
    // memset node, so that the sequence can be captured into a graph
    hipError_t e = launch_fill(reinterpret_cast<uint32_t*>(s.status), 0u, 2ull * tiles + 2, st);
    if (e != hipSuccess) return e;
    LaunchCfg lc;
    const uint32_t blocks = std::min<uint32_t>(tiles, stride_blocks((uint64_t)tiles * kScanBlock, kScanBlock, lc));
    hipLaunchKernelGGL(k_scan_excl, dim3(blocks), dim3(kScanBlock), 0, st, data, d_n, n_add, d_total,
                       s.status, reinterpret_cast<uint32_t*>(s.status + tiles),
                       (uint32_t)(n_cap < 0xffffffffull ? n_cap : 0xffffffffull));
    return hipGetLastError();
}

}  // namespace trax
