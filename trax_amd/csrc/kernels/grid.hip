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
// outcome), and the outputs are written in output order, coalesced.
//
// Where the outputs' sources come from depends on the item's size:
//  - staged (n <= stage_cap): the count pass reads the item's x, y, z, detId
//    and hitId once, coalesced, into LDS (20 B per hit), and layer and event,
//    which are one value each for the whole item; the outputs are read from
//    LDS in output order.  Every input byte is read from HBM once, coalesced,
//    and every output byte written once.  An item whose layer or event column
//    is not uniform reads those two columns at the sources instead.
//  - local (n <= lds_hits): cells and perm in LDS (2 B per hit each); with a
//    record buffer, the count pass also writes each hit's (x, y, z, phi) and
//    (layer, detId, hitId, event) as one 32-byte record at the hit's source
//    position, and the outputs gather one record per hit (one cache line, and
//    no second atan2); without one, the seven columns are gathered.
//  - global: perm in global memory at the output positions, x, y, z (or the
//    record) re-read for the scatter.
// Reads 28 B per hit (staged) or 28 + 32 B per hit (records; 12 + 28 B
// without); writes 4 B per cell and 44 B per hit of output (52 B with the
// (x, y) column), and 32 B per hit of records.

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

// Items (event, layer) ordered by size, largest first (by the bit length of
// the hit count: a coarse order, enough to start the long items first); one
// block.  order[t] is the item the build's t-th block takes.
__global__ __launch_bounds__(1024) void k_grid_order(const uint32_t* __restrict__ d_nevents, uint32_t n_layers,
                                                     const uint32_t* __restrict__ lay_nhits,
                                                     uint32_t* __restrict__ order, uint32_t n_items_cap,
                                                     uint32_t* __restrict__ reset, uint32_t reset_words) {
    __shared__ uint32_t bucket[33];
    // words to clear before the build (the run's counts after the batch's
    // shape), in the same launch rather than a fill of their own
    for (uint32_t i = threadIdx.x; i < reset_words; i += blockDim.x) reset[i] = 0u;
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

// LDS of one block: hist (cells) | wsum (waves + 1) | mixed flag | the staged
// columns x, y, z, detId, hitId (stage_cap each) | cell and perm lists
// (lds_hits u16 each)
struct GridLds {
    uint32_t *hist, *wsum, *mixed, *sx, *sy, *sz, *sdet, *sid;
    uint16_t *lcell, *lperm;
};
__device__ inline GridLds grid_lds(uint32_t* lds, uint32_t cells, uint32_t stage_cap, uint32_t lds_hits) {
    GridLds L;
    L.hist = lds;
    L.wsum = L.hist + cells;
    L.mixed = L.wsum + blockDim.x / kWave + 1;
    L.sx = L.mixed + 1;
    L.sy = L.sx + stage_cap;
    L.sz = L.sy + stage_cap;
    L.sdet = L.sz + stage_cap;
    L.sid = L.sdet + stage_cap;
    L.lcell = reinterpret_cast<uint16_t*>(L.sid + stage_cap);
    L.lperm = L.lcell + lds_hits;
    return L;
}
size_t grid_lds_bytes(uint32_t cells, uint32_t block, uint32_t stage_cap, uint32_t lds_hits) {
    return (size_t)(cells + block / kWave + 2 + 5ull * stage_cap) * 4 + (size_t)lds_hits * 4;
}

__global__ __launch_bounds__(1024) void k_grid_build(GridGeom g, const uint32_t* __restrict__ d_nevents,
                                                   const uint32_t* __restrict__ evt_offset,
                                                   const uint32_t* __restrict__ lay_nhits,
                                                   const uint32_t* __restrict__ lay_offset, HitCols in,
                                                   HitColsOut out, uint32_t* __restrict__ grid,
                                                   uint32_t* __restrict__ perm,
                                                   const uint32_t* __restrict__ order, uint32_t n_hits_cap,
                                                   uint32_t lds_hits, uint32_t stage_cap, float4* __restrict__ rec,
                                                   DevCounts* __restrict__ cnt) {
    extern __shared__ uint32_t lds[];
    const GridLds L = grid_lds(lds, g.cells, stage_cap, lds_hits);
    uint32_t* hist = L.hist;
    const uint32_t n_items = *d_nevents * g.n_layers;
    const uint32_t bs = blockDim.x;
    // the block's next item (its index, size and first hit), loaded one item
    // ahead so that those loads overlap the current item's work
    uint32_t nx_item = 0, nx_n = 0, nx_off = 0;
    auto fetch = [&](uint32_t tt) {
        nx_item = order ? order[tt] : tt;
        nx_n = lay_nhits[nx_item];
        nx_off = evt_offset[nx_item / g.n_layers] + lay_offset[nx_item];
    };
    if (blockIdx.x < n_items) fetch(blockIdx.x);
    for (uint32_t t = blockIdx.x; t < n_items; t += gridDim.x) {
        const uint32_t item = nx_item, n = nx_n, off = nx_off;
        if (t + gridDim.x < n_items) fetch(t + gridDim.x);
        for (uint32_t i = threadIdx.x; i < g.cells; i += bs) hist[i] = 0;
        if (threadIdx.x == 0) *L.mixed = 0;
        __syncthreads();
        const bool fits = (uint64_t)off + n <= n_hits_cap;
        const bool staged = fits && n <= stage_cap;  // block-uniform
        const bool local = n <= lds_hits;            // block-uniform
        const bool use_rec = rec && fits && !staged;  // block-uniform: records at the sources
        // the item's layer and event: one value each for a uniform item
        const uint32_t la0 = (staged && n) ? in.layer[off] : 0u, ev0 = (staged && n) ? in.evt[off] : 0u;
        bool bad = false, mixed = false;
        for (uint32_t i0 = threadIdx.x; i0 < n; i0 += kGridUnroll * bs) {
            float x[kGridUnroll], y[kGridUnroll], z[kGridUnroll];
            uint32_t de[kGridUnroll], id[kGridUnroll], la[kGridUnroll], ev[kGridUnroll];
#pragma unroll
            for (uint32_t k = 0; k < kGridUnroll; ++k) {
                const uint32_t i = i0 + k * bs;
                if (i < n) {
                    x[k] = in.x[off + i]; y[k] = in.y[off + i]; z[k] = in.z[off + i];
                    if (staged || use_rec) {
                        de[k] = in.det[off + i]; id[k] = in.id[off + i];
                        la[k] = in.layer[off + i]; ev[k] = in.evt[off + i];
                    }
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < kGridUnroll; ++k) {
                const uint32_t i = i0 + k * bs;
                if (i < n) {
                    if (staged) {
                        L.sx[i] = __float_as_uint(x[k]); L.sy[i] = __float_as_uint(y[k]);
                        L.sz[i] = __float_as_uint(z[k]); L.sdet[i] = de[k]; L.sid[i] = id[k];
                        mixed |= la[k] != la0 || ev[k] != ev0;
                    }
                    const float phi = tx_atan2f(y[k], x[k]);
                    const uint32_t c = hit_cell_phi(g, z[k], phi);
                    if (use_rec) {
                        float4* r = rec + 2 * (size_t)(off + i);
                        r[0] = make_float4(x[k], y[k], z[k], phi);
                        r[1] = make_float4(__uint_as_float(la[k]), __uint_as_float(de[k]), __uint_as_float(id[k]),
                                           __uint_as_float(ev[k]));
                    }
                    if (c < g.cells) atomicAdd(&hist[c], 1u);
                    else bad = true;
                    if (local) L.lcell[i] = c < g.cells ? (uint16_t)c : kNoCell;
                }
            }
        }
        if (bad) atomicOr(&cnt->err, kErrCellRange);
        if (mixed) atomicOr(L.mixed, 1u);
        __syncthreads();
        lds_excl_scan(hist, g.cells, L.wsum);
        uint32_t* gout = grid + (size_t)item * g.cells;
        for (uint32_t i = threadIdx.x; i < g.cells; i += bs) gout[i] = off + hist[i];
        __syncthreads();
        if (fits) {
            // positions: hist becomes each cell's end
            if (local) {
                for (uint32_t i = threadIdx.x; i < n; i += bs) {
                    const uint16_t c = L.lcell[i];
                    if (c != kNoCell) L.lperm[atomicAdd(&hist[c], 1u)] = (uint16_t)i;
                }
                __syncthreads();
                sort_cells(L.lperm, hist, g.cells);
            } else {
                for (uint32_t i = threadIdx.x; i < n; i += bs) {
                    uint32_t c;
                    if (use_rec) {
                        const float4 r = rec[2 * (size_t)(off + i)];
                        c = hit_cell_phi(g, r.z, r.w);
                    } else {
                        c = hit_cell(g, in.x[off + i], in.y[off + i], in.z[off + i]);
                    }
                    if (c < g.cells) perm[off + atomicAdd(&hist[c], 1u)] = i;
                }
                __syncthreads();
                sort_cells(perm + off, hist, g.cells);
            }
            __syncthreads();
            const uint32_t placed = g.cells ? hist[g.cells - 1] : 0u;  // hits with a cell
            const bool from_lds = staged, cols_mixed = *L.mixed != 0u;
            for (uint32_t p0 = threadIdx.x; p0 < placed; p0 += kGridUnroll * bs) {
                uint32_t src[kGridUnroll];
                float x[kGridUnroll], y[kGridUnroll], z[kGridUnroll], ph[kGridUnroll];
                uint32_t la[kGridUnroll], de[kGridUnroll], id[kGridUnroll], ev[kGridUnroll];
#pragma unroll
                for (uint32_t k = 0; k < kGridUnroll; ++k) {
                    const uint32_t p = p0 + k * bs;
                    src[k] = p < placed ? (local ? (uint32_t)L.lperm[p] : perm[off + p]) : 0u;
                }
#pragma unroll
                for (uint32_t k = 0; k < kGridUnroll; ++k) {
                    if (p0 + k * bs < placed) {
                        const uint32_t i = src[k];
                        if (from_lds) {
                            x[k] = __uint_as_float(L.sx[i]); y[k] = __uint_as_float(L.sy[i]);
                            z[k] = __uint_as_float(L.sz[i]); de[k] = L.sdet[i]; id[k] = L.sid[i];
                            la[k] = cols_mixed ? in.layer[off + i] : la0;
                            ev[k] = cols_mixed ? in.evt[off + i] : ev0;
                        } else if (use_rec) {  // one 32-byte record per hit
                            const float4* r = rec + 2 * (size_t)(off + i);
                            const float4 a = r[0], b = r[1];
                            x[k] = a.x; y[k] = a.y; z[k] = a.z; ph[k] = a.w;
                            la[k] = __float_as_uint(b.x); de[k] = __float_as_uint(b.y);
                            id[k] = __float_as_uint(b.z); ev[k] = __float_as_uint(b.w);
                        } else {
                            const uint32_t j = off + i;
                            x[k] = in.x[j]; y[k] = in.y[j]; z[k] = in.z[j];
                            la[k] = in.layer[j]; de[k] = in.det[j]; id[k] = in.id[j]; ev[k] = in.evt[j];
                        }
                    }
                }
#pragma unroll
                for (uint32_t k = 0; k < kGridUnroll; ++k) {
                    const uint32_t p = p0 + k * bs;
                    if (p < placed) {
                        const uint32_t pos = off + p;
                        out.x[pos] = x[k]; out.y[pos] = y[k]; out.z[pos] = z[k];
                        out.layer[pos] = la[k];
                        out.det[pos] = de[k];
                        out.id[pos] = id[k];
                        out.evt[pos] = ev[k];
                        if (out.zphi) out.zphi[pos] = make_float2(z[k], use_rec ? ph[k] : tx_atan2f(y[k], x[k]));
                        if (out.xy) out.xy[pos] = make_float2(x[k], y[k]);
                    }
                }
            }
        }
        __syncthreads();
    }
}

// hits per item the LDS lists hold, and the items staged whole, for a block
// of `block` threads within an LDS budget of `budget` bytes
void grid_lds_plan(const GridGeom& g, uint32_t block, size_t budget, uint32_t& lds_hits, uint32_t& stage_cap) {
    lds_hits = stage_cap = 0;
    const size_t base = grid_lds_bytes(g.cells, block, 0, 0);
    if (g.cells >= kNoCell || base >= budget) return;
    if (budget > kGridLdsBase) {
        // staged: 20 B of columns + 4 B of lists per hit
        const size_t fit = (budget - base) / 24;
        stage_cap = lds_hits = (uint32_t)std::min<size_t>(fit & ~(size_t)63, 0xfff0u);
        return;
    }
    const size_t fit = (budget - base) / (2 * sizeof(uint16_t));
    lds_hits = (uint32_t)std::min<size_t>(fit & ~(size_t)63, kGridLdsHitsMax);
}

}  // namespace

hipError_t launch_grid_build(const GridGeom& g, uint32_t n_items_cap, const uint32_t* evt_offset,
                             const uint32_t* lay_nhits, const uint32_t* lay_offset, float4* rec,
                             const HitColumns& in, const HitColumns& out, uint32_t* grid, uint32_t* perm,
                             uint32_t n_hits_cap, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st,
                             bool reset_counts) {
    // defaults: 1024 threads, 160 KiB of LDS (items staged whole), largest items first
    const uint32_t bs = lc.grid_count_block ? lc.grid_count_block : 1024;
    uint32_t lds_hits = 0, stage_cap = 0;
    const size_t budget = lc.grid_stage_kib ? (size_t)lc.grid_stage_kib * 1024 : kGridLdsMax;
    grid_lds_plan(g, bs, std::min<size_t>(budget, kGridLdsMax), lds_hits, stage_cap);
    const size_t lds = grid_lds_bytes(g.cells, bs, stage_cap, lds_hits);
    // one block per item: the dispatcher hands the next item to the first block slot that frees up
    const uint32_t blocks = n_items_cap ? n_items_cap : 1;
    // perm holds n_hits_cap scatter entries, then the item order (n_items_cap)
    uint32_t* order = nullptr;
    uint32_t* reset = reinterpret_cast<uint32_t*>(cnt) + kCountsShapeWords;
    const uint32_t reset_words = reset_counts ? (uint32_t)(sizeof(DevCounts) / 4 - kCountsShapeWords) : 0u;
    if (lc.grid_order != 2) {
        order = perm + n_hits_cap;
        hipLaunchKernelGGL(k_grid_order, dim3(1), dim3(1024), 0, st, &cnt->n_events, g.n_layers, lay_nhits, order,
                           n_items_cap, reset, reset_words);
    } else if (reset_words) {
        const hipError_t e = launch_fill(reset, 0u, reset_words, st);
        if (e != hipSuccess) return e;
    }
    HitCols ic{in.x, in.y, in.z, in.layer, in.det, in.id, in.evt};
    HitColsOut oc{out.x, out.y, out.z, out.layer, out.det, out.id, out.evt, out.zphi, out.xy};
    hipLaunchKernelGGL(k_grid_build, dim3(blocks), dim3(bs), lds, st, g, &cnt->n_events, evt_offset, lay_nhits,
                       lay_offset, ic, oc, grid, perm, order, n_hits_cap, lds_hits, stage_cap, rec, cnt);
    return hipGetLastError();
}

// the build may take up to kGridLdsMax of LDS per block: set once, outside
// any graph capture
void grid_build_prepare() {
    static bool done = false;
    if (!done) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_grid_build),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kGridLdsMax);
        done = true;
    }
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
