// GridBuilder kernels (src/algorithms/GridBuilder.h:48-427, GridBuilder.cpp:6-161).
//
// Layout in HBM: hits are seven SoA columns (x, y, z: f32; layer, detId, hitId,
// event: u32), ordered event-major then layer-major; the grid is
// E * L * (nZ+1)(nPhi+1) u32 cells, cell (e, l, z, p) at
// ((e*L + l) * (nZ+1) + z) * (nPhi+1) + p, holding after the build the global
// index of the first hit of that cell.
//
// grid count: one block per (event, layer); an LDS histogram of the cells,
//   then an LDS exclusive scan of it, written out already offset by the
//   (event, layer)'s first hit.  This fuses the reference's gridCount,
//   the global PrefixSum over the grid and its download into one pass:
//   reads 12 B per hit, writes 4 B per cell.
// grid store: one block per (event, layer); the scanned cells are staged in
//   LDS as write cursors and the hits are scattered in input order within
//   each cell (a stable scatter: one legal outcome of the reference's
//   atomic_inc-ordered gridStore).  Each 64-hit wave ranks its lanes among
//   equal cells with ballots; waves of a block take turns advancing the
//   cursors so that the order is the input order.  Reads and writes 28 B per
//   hit, reads 4 B per cell.
#include "kernels.h"

namespace trax {
namespace {

__global__ __launch_bounds__(1024) void k_grid_count(GridGeom g, const uint32_t* __restrict__ d_nevents,
                                                   const uint32_t* __restrict__ evt_offset,
                                                   const uint32_t* __restrict__ lay_nhits,
                                                   const uint32_t* __restrict__ lay_offset,
                                                   const float* __restrict__ hx,
                                                   const float* __restrict__ hy,
                                                   const float* __restrict__ hz,
                                                   uint32_t* __restrict__ grid, DevCounts* __restrict__ cnt) {
    extern __shared__ uint32_t lds[];
    uint32_t* hist = lds;
    uint32_t* wsum = lds + g.cells;
    const uint32_t n_items = *d_nevents * g.n_layers;
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const uint32_t e = item / g.n_layers;
        for (uint32_t i = threadIdx.x; i < g.cells; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        const uint32_t n = lay_nhits[item];
        const uint32_t off = evt_offset[e] + lay_offset[item];
        bool bad = false;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const uint32_t c = hit_cell(g, hx[off + i], hy[off + i], hz[off + i]);
            if (c < g.cells) atomicAdd(&hist[c], 1u);
            else bad = true;
        }
        if (bad) atomicOr(&cnt->err, kErrCellRange);
        __syncthreads();
        lds_excl_scan(hist, g.cells, wsum);
        uint32_t* out = grid + (size_t)item * g.cells;
        for (uint32_t i = threadIdx.x; i < g.cells; i += blockDim.x) out[i] = off + hist[i];
        __syncthreads();
    }
}

struct HitCols {
    const float* x; const float* y; const float* z;
    const uint32_t* layer; const uint32_t* det; const uint32_t* id; const uint32_t* evt;
};
struct HitColsOut {
    float* x; float* y; float* z;
    uint32_t* layer; uint32_t* det; uint32_t* id; uint32_t* evt;
    float2* zphi;
    float2* xy;
};

__global__ __launch_bounds__(1024) void k_grid_store(GridGeom g, const uint32_t* __restrict__ d_nevents,
                                                   const uint32_t* __restrict__ evt_offset,
                                                   const uint32_t* __restrict__ lay_nhits,
                                                   const uint32_t* __restrict__ lay_offset,
                                                   const uint32_t* __restrict__ grid, HitCols in,
                                                   HitColsOut out, uint32_t n_hits_cap) {
    extern __shared__ uint32_t wp[];
    const uint32_t n_items = *d_nevents * g.n_layers;
    const uint32_t wave = threadIdx.x / kWave, n_waves = blockDim.x / kWave;
    const uint32_t lane = lane_id();
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const uint32_t e = item / g.n_layers;
        const uint32_t* gsl = grid + (size_t)item * g.cells;
        for (uint32_t i = threadIdx.x; i < g.cells; i += blockDim.x) wp[i] = gsl[i];
        __syncthreads();
        const uint32_t n = lay_nhits[item];
        const uint32_t off = evt_offset[e] + lay_offset[item];
        for (uint32_t chunk = 0; chunk < n; chunk += blockDim.x) {
            const uint32_t i = chunk + threadIdx.x;
            bool valid = i < n;
            float x = 0.f, y = 0.f, z = 0.f, phi = 0.f;
            uint32_t c = kSentinel;
            if (valid) {
                x = in.x[off + i]; y = in.y[off + i]; z = in.z[off + i];
                phi = tx_atan2f(y, x);
                c = hit_cell_phi(g, z, phi);
                valid = c < g.cells;
            }
            // rank among lanes of this wave with the same cell
            uint64_t todo = __ballot(valid);
            uint32_t rank = 0, num = 0;
            bool leader = false;
            while (todo) {
                const int src = __ffsll((unsigned long long)todo) - 1;
                const uint32_t cs = __builtin_amdgcn_readlane(c, src);
                const bool mine = valid && c == cs;
                const uint64_t m = __ballot(mine);
                if (mine) {
                    rank = __popcll(m & lanemask_lt());
                    num = __popcll(m);
                    leader = (lane == (uint32_t)src);
                }
                todo &= ~m;
            }
            uint32_t pos = 0;
            for (uint32_t w = 0; w < n_waves; ++w) {
                if (w == wave && valid) {
                    const uint32_t b = wp[c];
                    pos = b + rank;
                    __builtin_amdgcn_wave_barrier();
                    if (leader) wp[c] = b + num;
                }
                __syncthreads();
            }
            if (valid && pos < n_hits_cap) {
                out.x[pos] = x; out.y[pos] = y; out.z[pos] = z;
                out.layer[pos] = in.layer[off + i];
                out.det[pos] = in.det[off + i];
                out.id[pos] = in.id[off + i];
                out.evt[pos] = in.evt[off + i];
                if (out.zphi) out.zphi[pos] = make_float2(z, phi);
                if (out.xy) out.xy[pos] = make_float2(x, y);
            }
        }
        __syncthreads();
    }
}

// The whole grid build of one (event, layer) in one block (batched path):
// the LDS histogram of the cells and its scan as in k_grid_count (the grid
// written out), then the scatter in three steps instead of k_grid_store's
// wave-ordered cursors: every hit takes a position in its cell with an LDS
// atomic and leaves its local index there (perm), each cell's few indices are
// sorted ascending (so a cell holds its hits in input order: the same stable
// outcome), and the outputs are gathered in output order, coalesced.  A layer
// of at most lds_hits hits keeps its cells and perm in LDS (2 B per hit each):
// x, y, z are read from HBM once and the sort never leaves LDS.  A larger
// layer re-reads x, y, z for the scatter and keeps perm in global memory at
// the output positions.  Reads 12 B per hit for the count, 28 B per hit for
// the gather; writes 4 B per cell and 44 B per hit of output (52 B with the
// (x, y) column).

// insertion sort of each cell's entries of perm; ends[c] is the end of cell c
template <typename T>
__device__ inline void sort_cells(T* perm, const uint32_t* ends, uint32_t cells) {
    for (uint32_t c = threadIdx.x; c < cells; c += blockDim.x) {
        const uint32_t lo = c ? ends[c - 1] : 0u, hi = ends[c];
        for (uint32_t k = lo + 1; k < hi; ++k) {
            const T v = perm[k];
            uint32_t j = k;
            while (j > lo && perm[j - 1] > v) {
                perm[j] = perm[j - 1];
                --j;
            }
            perm[j] = v;
        }
    }
}

// The stable position of the hit at unsorted position p of its cell [lo, hi)
// (placed there by an atomic): lo + the number of the cell's hits with a
// smaller input index.
template <typename T>
__device__ inline uint32_t stable_pos(const T* perm, uint32_t p, uint32_t lo, uint32_t hi) {
    if (hi - lo < 2) return p;
    const T v = perm[p];
    uint32_t r = lo;
    for (uint32_t k = lo; k < hi; ++k) r += perm[k] < v ? 1u : 0u;
    return r;
}

// Items (event, layer) ordered by size, largest first (by the bit length of
// the hit count: a coarse order, enough to start the long items first); one
// block.  order[t] is the item the build's t-th block takes.
__global__ __launch_bounds__(1024) void k_grid_order(const uint32_t* __restrict__ d_nevents, uint32_t n_layers,
                                                     const uint32_t* __restrict__ lay_nhits,
                                                     uint32_t* __restrict__ order, uint32_t n_items_cap) {
    __shared__ uint32_t bucket[33];
    const uint32_t n_items = min(*d_nevents * n_layers, n_items_cap);
    if (threadIdx.x < 33) bucket[threadIdx.x] = 0;
    __syncthreads();
    // leading zeros of the hit count: 0 for the largest items, 32 for empty ones
    auto key = [&](uint32_t item) { return (uint32_t)__clz((int)lay_nhits[item]); };
    for (uint32_t i = threadIdx.x; i < n_items; i += blockDim.x) atomicAdd(&bucket[key(i)], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t b = 0; b < 33; ++b) {
            const uint32_t c = bucket[b];
            bucket[b] = run;
            run += c;
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n_items; i += blockDim.x) order[atomicAdd(&bucket[key(i)], 1u)] = i;
}

constexpr uint32_t kGridUnroll = 4;       // hits per thread whose loads are issued together
constexpr uint16_t kNoCell = 0xffffu;     // a hit outside the grid, in the LDS cell list

__global__ __launch_bounds__(1024) void k_grid_build(GridGeom g, const uint32_t* __restrict__ d_nevents,
                                                   const uint32_t* __restrict__ evt_offset,
                                                   const uint32_t* __restrict__ lay_nhits,
                                                   const uint32_t* __restrict__ lay_offset, HitCols in,
                                                   HitColsOut out, uint32_t* __restrict__ grid,
                                                   uint32_t* __restrict__ perm,
                                                   const uint32_t* __restrict__ order, uint4* __restrict__ rec,
                                                   uint32_t n_hits_cap,
                                                   uint32_t lds_hits, uint32_t rank_mode,
                                                   DevCounts* __restrict__ cnt) {
    extern __shared__ uint32_t lds[];
    uint32_t* hist = lds;
    uint32_t* wsum = lds + g.cells;
    uint16_t* lcell = reinterpret_cast<uint16_t*>(wsum + blockDim.x / kWave + 1);
    uint16_t* lperm = lcell + lds_hits;
    const uint32_t n_items = *d_nevents * g.n_layers;
    const uint32_t bs = blockDim.x;
    for (uint32_t t = blockIdx.x; t < n_items; t += gridDim.x) {
        const uint32_t item = order ? order[t] : t;
        const uint32_t e = item / g.n_layers;
        for (uint32_t i = threadIdx.x; i < g.cells; i += bs) hist[i] = 0;
        __syncthreads();
        const uint32_t n = lay_nhits[item];
        const uint32_t off = evt_offset[e] + lay_offset[item];
        const bool fits = (uint64_t)off + n <= n_hits_cap;
        const bool local = n <= lds_hits;  // block-uniform
        bool bad = false;
        for (uint32_t i0 = threadIdx.x; i0 < n; i0 += kGridUnroll * bs) {
            float x[kGridUnroll], y[kGridUnroll], z[kGridUnroll];
#pragma unroll
            for (uint32_t k = 0; k < kGridUnroll; ++k) {
                const uint32_t i = i0 + k * bs;
                if (i < n) { x[k] = in.x[off + i]; y[k] = in.y[off + i]; z[k] = in.z[off + i]; }
            }
            if (rec && fits) {
                // each hit's seven columns as one 32-byte record, written in
                // input order (coalesced) for the gather below
#pragma unroll
                for (uint32_t k = 0; k < kGridUnroll; ++k) {
                    const uint32_t i = i0 + k * bs;
                    if (i < n) {
                        const size_t r = 2 * ((size_t)off + i);
                        rec[r] = make_uint4(__float_as_uint(x[k]), __float_as_uint(y[k]), __float_as_uint(z[k]),
                                            in.layer[off + i]);
                        rec[r + 1] = make_uint4(in.det[off + i], in.id[off + i], in.evt[off + i], 0u);
                    }
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < kGridUnroll; ++k) {
                const uint32_t i = i0 + k * bs;
                if (i < n) {
                    const uint32_t c = hit_cell(g, x[k], y[k], z[k]);
                    if (c < g.cells) atomicAdd(&hist[c], 1u);
                    else bad = true;
                    if (local) lcell[i] = c < g.cells ? (uint16_t)c : kNoCell;
                }
            }
        }
        if (bad) atomicOr(&cnt->err, kErrCellRange);
        __syncthreads();
        lds_excl_scan(hist, g.cells, wsum);
        uint32_t* gout = grid + (size_t)item * g.cells;
        for (uint32_t i = threadIdx.x; i < g.cells; i += bs) gout[i] = off + hist[i];
        __syncthreads();
        if (fits) {
            // positions: hist becomes each cell's end
            if (local) {
                for (uint32_t i = threadIdx.x; i < n; i += bs) {
                    const uint16_t c = lcell[i];
                    if (c != kNoCell) lperm[atomicAdd(&hist[c], 1u)] = (uint16_t)i;
                }
                __syncthreads();
                if (rank_mode == 0) sort_cells(lperm, hist, g.cells);
            } else {
                for (uint32_t i = threadIdx.x; i < n; i += bs) {
                    const uint32_t c = hit_cell(g, in.x[off + i], in.y[off + i], in.z[off + i]);
                    if (c < g.cells) perm[off + atomicAdd(&hist[c], 1u)] = i;
                }
                __syncthreads();
                if (rank_mode == 0) sort_cells(perm + off, hist, g.cells);
            }
            if (rank_mode == 0) __syncthreads();
            const uint32_t placed = g.cells ? hist[g.cells - 1] : 0u;  // hits with a cell
            for (uint32_t p0 = threadIdx.x; p0 < placed; p0 += kGridUnroll * bs) {
                uint32_t src[kGridUnroll], dst[kGridUnroll];
                float x[kGridUnroll], y[kGridUnroll], z[kGridUnroll];
                uint32_t la[kGridUnroll], de[kGridUnroll], id[kGridUnroll], ev[kGridUnroll];
#pragma unroll
                for (uint32_t k = 0; k < kGridUnroll; ++k) {
                    const uint32_t p = p0 + k * bs;
                    src[k] = off;
                    dst[k] = p;
                    if (p < placed) {
                        const uint32_t i = local ? (uint32_t)lperm[p] : perm[off + p];
                        src[k] = off + i;
                        if (rank_mode != 0) {
                            // the cell of position p: its hit's cell (local) or recomputed
                            const uint32_t c = local ? (uint32_t)lcell[i]
                                                     : hit_cell(g, in.x[off + i], in.y[off + i], in.z[off + i]);
                            const uint32_t lo = c ? hist[c - 1] : 0u, hi = hist[c];
                            dst[k] = local ? stable_pos(lperm, p, lo, hi) : stable_pos(perm + off, p, lo, hi);
                        }
                    }
                }
#pragma unroll
                for (uint32_t k = 0; k < kGridUnroll; ++k) {
                    if (p0 + k * bs < placed) {
                        const uint32_t i = src[k];
                        if (rec) {
                            const uint4 a = rec[2 * (size_t)i], b = rec[2 * (size_t)i + 1];
                            x[k] = __uint_as_float(a.x); y[k] = __uint_as_float(a.y); z[k] = __uint_as_float(a.z);
                            la[k] = a.w; de[k] = b.x; id[k] = b.y; ev[k] = b.z;
                        } else {
                            x[k] = in.x[i]; y[k] = in.y[i]; z[k] = in.z[i];
                            la[k] = in.layer[i]; de[k] = in.det[i]; id[k] = in.id[i]; ev[k] = in.evt[i];
                        }
                    }
                }
#pragma unroll
                for (uint32_t k = 0; k < kGridUnroll; ++k) {
                    if (p0 + k * bs < placed) {
                        const uint32_t pos = off + dst[k];
                        out.x[pos] = x[k]; out.y[pos] = y[k]; out.z[pos] = z[k];
                        out.layer[pos] = la[k];
                        out.det[pos] = de[k];
                        out.id[pos] = id[k];
                        out.evt[pos] = ev[k];
                        if (out.zphi) out.zphi[pos] = make_float2(z[k], tx_atan2f(y[k], x[k]));
                        if (out.xy) out.xy[pos] = make_float2(x[k], y[k]);
                    }
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_grid_build(const GridGeom& g, uint32_t n_items_cap, const uint32_t* evt_offset,
                             const uint32_t* lay_nhits, const uint32_t* lay_offset, const uint32_t* grid_unused,
                             const HitColumns& in, const HitColumns& out, uint32_t* grid, uint32_t* perm,
                             uint32_t n_hits_cap, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st,
                             uint4* rec) {
    (void)grid_unused;
    const uint32_t bs = lc.grid_count_block ? lc.grid_count_block : 256;
    const uint32_t blocks = n_items_cap ? n_items_cap : 1;
    const size_t base = (size_t)(g.cells + bs / kWave + 1) * sizeof(uint32_t);
    uint32_t lds_hits = grid_lds_hits(g, bs);
    if (lc.grid_lds_hits) lds_hits = std::min(lds_hits, lc.grid_lds_hits & ~63u);
    const size_t lds = base + (size_t)lds_hits * 2 * sizeof(uint16_t);
    // perm holds n_hits_cap scatter entries, then the item order (n_items_cap)
    uint32_t* order = nullptr;
    if (lc.grid_order) {
        order = perm + n_hits_cap;
        hipLaunchKernelGGL(k_grid_order, dim3(1), dim3(1024), 0, st, &cnt->n_events, g.n_layers, lay_nhits, order,
                           n_items_cap);
    }
    HitCols ic{in.x, in.y, in.z, in.layer, in.det, in.id, in.evt};
    HitColsOut oc{out.x, out.y, out.z, out.layer, out.det, out.id, out.evt, out.zphi, out.xy};
    hipLaunchKernelGGL(k_grid_build, dim3(blocks), dim3(bs), lds, st, g, &cnt->n_events, evt_offset, lay_nhits,
                       lay_offset, ic, oc, grid, perm, order, lc.grid_gather ? rec : nullptr, n_hits_cap, lds_hits,
                       lc.grid_rank, cnt);
    return hipGetLastError();
}

uint32_t grid_lds_hits(const GridGeom& g, uint32_t block) {
    // LDS per block stays within 64 KiB; cells are kept as u16 with 0xffff free
    constexpr size_t kBudget = 64 * 1024;
    const size_t base = (size_t)(g.cells + block / kWave + 1) * sizeof(uint32_t);
    if (g.cells >= kNoCell || base >= kBudget) return 0;
    const size_t fit = (kBudget - base) / (2 * sizeof(uint16_t));
    return (uint32_t)std::min<size_t>(fit & ~(size_t)63, kGridLdsHitsMax);
}

hipError_t launch_grid_count(const GridGeom& g, uint32_t n_items_cap, const uint32_t* evt_offset,
                             const uint32_t* lay_nhits, const uint32_t* lay_offset, const float* hx,
                             const float* hy, const float* hz, uint32_t* grid, DevCounts* cnt,
                             const LaunchCfg& lc, hipStream_t st) {
    const uint32_t bc = lc.grid_count_block ? lc.grid_count_block : 256;
    const uint32_t blocks = n_items_cap ? n_items_cap : 1;
    const size_t lds = (size_t)(g.cells + bc / kWave + 1) * sizeof(uint32_t);
    hipLaunchKernelGGL(k_grid_count, dim3(blocks), dim3(bc), lds, st, g, &cnt->n_events, evt_offset,
                       lay_nhits, lay_offset, hx, hy, hz, grid, cnt);
    return hipGetLastError();
}

hipError_t launch_grid_store(const GridGeom& g, uint32_t n_items_cap, const uint32_t* evt_offset,
                             const uint32_t* lay_nhits, const uint32_t* lay_offset, const uint32_t* grid,
                             const HitColumns& in, const HitColumns& out, uint32_t n_hits_cap,
                             DevCounts* cnt, const LaunchCfg& lc, hipStream_t st) {
    const uint32_t bs = lc.grid_store_block ? lc.grid_store_block : 256;
    const uint32_t blocks = n_items_cap ? n_items_cap : 1;
    const size_t lds = (size_t)g.cells * sizeof(uint32_t);
    HitCols ic{in.x, in.y, in.z, in.layer, in.det, in.id, in.evt};
    HitColsOut oc{out.x, out.y, out.z, out.layer, out.det, out.id, out.evt, out.zphi, out.xy};
    hipLaunchKernelGGL(k_grid_store, dim3(blocks), dim3(bs), lds, st, g, &cnt->n_events, evt_offset,
                       lay_nhits, lay_offset, grid, ic, oc, n_hits_cap);
    return hipGetLastError();
}

}  // namespace trax
