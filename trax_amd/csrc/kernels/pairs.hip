// PairGeneratorBeamspot and TripletThetaPhiPredictor kernels
// (src/algorithms/PairGeneratorBeamspot.h:43-525, .cpp:5-183;
//  src/algorithms/TripletThetaPhiPredictor.h:49-447, .cpp:10-170).
//
// Pair order.  The reference gives each (event, layer pair) one work-group of
// nThreads work-items; item t handles outer hits t, t + nThreads, ... and
// writes its pairs contiguously, inner hits ascending (pairStore scans a dense
// N1 x N2 bitmap).  Here every outer hit is its own work item, and the count
// array is laid out "thread-major": segment s owns nThreads * K slots
// (K = ceil(N2 / nThreads)), slot t*K + k holding the count of outer hit
// k*nThreads + t.  One exclusive scan over the slots then gives every outer
// hit its output position in exactly the reference's order, and the store
// kernel enumerates the inner window in ascending index order (the order in
// which the reference reads its bitmap) -- without the O(N1*N2) bitmap.
//
// Predictor order.  One work item per pair, candidates in the reference's
// z-sector / wrap-around walk order, the store bounded by the count, both as
// in the reference (including the two different rOrigin expressions of the
// count and store kernels).  The batched path fuses count, scan and store into
// one pass (k_predict_fused): each pair's window is walked once with both
// expressions, the store-side matches are kept in LDS, tile offsets come from
// a decoupled look-back, and each block writes its candidates as one
// contiguous, coalesced run.
//
// Every window walk reads the hit's phi from the column the grid scatter
// wrote (the same atan2 of the same coordinates) instead of recomputing it.
#include <type_traits>

#include "filter_cuts.h"
#include "kernels.h"

namespace trax {
namespace {

struct Window {
    float zLow, zHigh, phiLow, phiHigh;
    uint32_t zLo, zHi, pLo, pHi;
    bool wrap;
};

// outer hit i of layer pair p: window in the inner layer (.h:108-163)
__device__ inline Window pair_window(const GridGeom& g, const TripletCols& tc, uint32_t p, uint32_t lay1,
                                     float X, float Y, float Z, float phi) {
    Window w;
    const float z0 = tc.z0[p];
    const float tip = tc.tip[p];
    const float sigmaZ = tc.sigma_z[p];
    const float sigmaPhi = tx_bits2f(tc.pair_spread_phi[p]);  // .cpp:59 passes pairSpreadPhi
    const float minR1 = tc.min_r[lay1], maxR1 = tc.max_r[lay1];
    const float mrc = tc.min_radius_curvature;
    const float signRadius = (Y > 0.0f) ? 1.0f : ((Y < 0.0f) ? -1.0f : (tx_isnan(Y) ? 0.0f : Y));
    const float r = signRadius * sqrtf(X * X + Y * Y);
    const float cotThetaHigh = (Z - z0) / r;
    const float cotThetaLow = (Z + z0) / r;
    float tmp = signRadius * maxR1 * cotThetaHigh + z0;
    float zHigh = signRadius * minR1 * cotThetaHigh + z0;
    zHigh = (float)(tmp < zHigh) * zHigh + (float)(tmp > zHigh) * tmp;
    tmp = signRadius * maxR1 * cotThetaLow - z0;
    float zLow = signRadius * minR1 * cotThetaLow - z0;
    zLow = (float)(tmp < zLow) * tmp + (float)(tmp > zLow) * zLow;
    zLow -= sigmaZ;
    zHigh += sigmaZ;
    const int zl = tx_f2i_sat(floorf((zLow - g.min_z) / g.sector_z));
    w.zLo = (uint32_t)(zl > 0 ? zl : 0);
    const uint32_t zh = tx_f2u_sat(floorf((zHigh - g.min_z) / g.sector_z)) + 1u;
    w.zHi = zh < g.n_z ? zh : g.n_z;
    const float aR = tx_acosf(r / (2 * mrc));
    tmp = fabsf(aR - tc.acos_max_r[lay1]);       // acos(maxR1 / (2 mrc))
    float dPhi = fabsf(aR - tc.acos_min_r[lay1]);  // acos(minR1 / (2 mrc))
    dPhi = (float)(tmp < dPhi) * dPhi + (float)(tmp > dPhi) * tmp;
    tmp = tx_atanf(tip * (r - minR1)) / (r * minR1);
    float phiHigh = tx_atanf(tip * (r - maxR1)) / (r * maxR1);
    dPhi += tmp > phiHigh ? tmp : phiHigh;
    float phiLow = phi - dPhi - sigmaPhi;
    phiHigh = phi + dPhi + sigmaPhi;
    const float PI = TX_PI_F;
    w.wrap = phiLow < -PI || phiHigh > PI || phiLow > PI || phiHigh < -PI;
    phiLow -= (float)((phiLow > PI) * 2) * PI;
    phiLow += (float)((phiLow < -PI) * 2) * PI;
    phiHigh -= (float)((phiHigh > PI) * 2) * PI;
    phiHigh += (float)((phiHigh < -PI) * 2) * PI;
    const uint32_t lim = g.n_phi + 1;  // keeps grid reads inside the layer (only NaN windows exceed it)
    const uint32_t pl = tx_f2u_sat(floorf((phiLow - g.min_phi) / g.sector_phi));
    const uint32_t ph = tx_f2u_sat(floorf((phiHigh - g.min_phi) / g.sector_phi) + 1.0f);
    w.pLo = pl < lim ? pl : lim;
    w.pHi = ph < lim ? ph : lim;
    w.zLow = zLow; w.zHigh = zHigh; w.phiLow = phiLow; w.phiHigh = phiHigh;
    return w;
}

__device__ inline bool in_window(const Window& w, float2 zphi) {
    const float PI = TX_PI_F;
    const float hz = zphi.x, actPhi = zphi.y;
    bool valid = w.zLow <= hz && hz <= w.zHigh;
    return valid && ((!w.wrap && (w.phiLow <= actPhi && actPhi <= w.phiHigh)) ||
                     (w.wrap && ((w.phiLow <= actPhi && actPhi <= PI) || (-PI <= actPhi && actPhi <= w.phiHigh))));
}

// the two hits of a pair: false for padding (sentinel) pairs and for a hit
// index out of range (reported)
__device__ inline bool pair_hits_ok(uint32_t h1, uint32_t h2, uint32_t nh, DevCounts* cnt) {
    if (h1 == kSentinel || h2 == kSentinel) return false;
    return index_ok(h1, nh, cnt) && index_ok(h2, nh, cnt);
}

// one pair segment (event, layer pair) -> number of count slots
__global__ void k_pair_plan(GridGeom g, TripletCols tc, const uint32_t* __restrict__ grid,
                            uint32_t* __restrict__ seg_base, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t nseg = cnt->n_events * tc.n_trip;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s <= nseg; s += gridDim.x * blockDim.x) {
        if (s == 0) cnt->n_segments = nseg;
        if (s == nseg) { seg_base[s] = 0; continue; }
        const uint32_t e = s / tc.n_trip, p = s % tc.n_trip;
        const uint32_t* g2 = grid + ((size_t)e * g.n_layers + tc.l2[p] - 1) * g.cells;
        const uint32_t n2 = g2[g.cells - 1] - g2[0];
        const uint32_t K = (n2 + tc.n_threads - 1) / tc.n_threads;
        seg_base[s] = K * tc.n_threads;
    }
}

struct SlotInfo {
    uint32_t seg, e, p, t, k;
};

// slot_seg (may be null) bounds the search: a slot's segment lies between the
// segments of the first slots of its block and of the next block
__device__ inline SlotInfo slot_info(const TripletCols& tc, const uint32_t* seg_base, uint32_t nseg,
                                     uint32_t slot, const uint32_t* __restrict__ slot_seg) {
    SlotInfo si;
    uint32_t lo = 0, hi = nseg;
    if (slot_seg) {
        const uint32_t b = slot >> kSlotSegShift;
        lo = slot_seg[b];
        const uint32_t h = slot_seg[b + 1] + 1u;
        hi = h < nseg ? h : nseg;
    }
    si.seg = find_segment_in(seg_base, lo, hi, slot);
    const uint32_t o = slot - seg_base[si.seg];
    const uint32_t span = seg_base[si.seg + 1] - seg_base[si.seg];
    const uint32_t K = span / tc.n_threads;
    si.t = o / K;
    si.k = o - si.t * K;
    si.e = si.seg / tc.n_trip;
    si.p = si.seg - si.e * tc.n_trip;
    return si;
}

// The entries [a, b) of the (z, phi) column in ascending order, read as
// aligned 16-byte pairs, four entries a step (half the loads of one entry a
// load; a wave's lanes each walk their own rows; the pair store's walk).  The column has
// kZphiSlack entries of room past its last hit, so a step's tail read stays
// inside the allocation.  f(j, entry) returns false to stop; returns whether
// the range was finished.
template <class F>
__device__ inline bool zphi_scan(const float2* __restrict__ hzp, uint32_t a, uint32_t b, F&& f) {
    const float4* p = reinterpret_cast<const float4*>(hzp);
    for (uint32_t jb = a & ~3u; jb < b; jb += 4) {
        const float4 v0 = p[jb >> 1], v1 = p[(jb >> 1) + 1];
        const float2 e[4] = {make_float2(v0.x, v0.y), make_float2(v0.z, v0.w), make_float2(v1.x, v1.y),
                             make_float2(v1.z, v1.w)};
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t j = jb + u;
            if (j >= a && j < b && !f(j, e[u])) return false;
        }
    }
    return true;
}

__global__ void k_pair_count(GridGeom g, TripletCols tc, const uint32_t* __restrict__ grid,
                             const float* __restrict__ hx, const float* __restrict__ hy,
                             const float* __restrict__ hz, const float2* __restrict__ hzp,
                             const uint32_t* __restrict__ seg_base, uint32_t* __restrict__ slot_counts,
                             DevCounts* __restrict__ cnt, const uint32_t* __restrict__ slot_seg, uint32_t xcds) {
    if (run_void(cnt)) return;
    const uint32_t nslot = bounded(cnt->n_pair_slots, cnt->cap_slots, cnt), nseg = cnt->n_segments;
    const uint32_t np1 = g.n_phi + 1;
    uint32_t overlap = 0;
    for (uint32_t slot = xcd_block(xcds) * blockDim.x + threadIdx.x; slot <= nslot; slot += gridDim.x * blockDim.x) {
        if (slot == nslot) { slot_counts[slot] = 0; continue; }
        const SlotInfo si = slot_info(tc, seg_base, nseg, slot, slot_seg);
        const uint32_t lay2 = tc.l2[si.p] - 1, lay1 = tc.l1[si.p] - 1;
        const uint32_t* g2 = grid + ((size_t)si.e * g.n_layers + lay2) * g.cells;
        const uint32_t layer2Offset = g2[0], end = g2[g.cells - 1];
        const uint32_t il = si.k * tc.n_threads + si.t;
        uint32_t nFound = 0;
        if (il < end - layer2Offset) {
            const uint32_t i = layer2Offset + il;
            const uint32_t* lg1 = grid + ((size_t)si.e * g.n_layers + lay1) * g.cells;
            const Window w = pair_window(g, tc, si.p, lay1, hx[i], hy[i], hz[i], hzp[i].y);
            if (w.wrap && w.pHi > w.pLo) overlap++;
            for (uint32_t zs = w.zLo; zs < w.zHi; ++zs) {  // (.h:172-230)
                const uint32_t zStart = lg1[zs * np1];
                const uint32_t zEnd = lg1[(zs + 1) * np1 - 1];
                const uint32_t zLen = zEnd - zStart;
                const uint32_t end2 = (uint32_t)w.wrap * zEnd + lg1[zs * np1 + w.pHi] - (uint32_t)w.wrap * zStart;
                for (uint32_t j = lg1[zs * np1 + w.pLo]; j < end2; j += 4) {  // four loads in flight
                    float2 q[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const uint32_t jj = j + u;
                        q[u] = jj < end2 ? hzp[jj - (uint32_t)(jj >= zEnd) * zLen] : make_float2(0.0f, 0.0f);
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) nFound += (j + u < end2) && in_window(w, q[u]);
                }
            }
        }
        slot_counts[slot] = nFound;
    }
    if (overlap) atomicAdd(&cnt->diag_wrap_overlap, overlap);
}

__global__ void k_pair_store(GridGeom g, TripletCols tc, const uint32_t* __restrict__ grid,
                             const float* __restrict__ hx, const float* __restrict__ hy,
                             const float* __restrict__ hz, const float2* __restrict__ hzp,
                             const uint32_t* __restrict__ seg_base, const uint32_t* __restrict__ slot_counts,
                             uint32_t* __restrict__ pairs, uint32_t n_pair_cap, DevCounts* __restrict__ cnt,
                             const uint32_t* __restrict__ slot_seg, uint32_t xcds) {
    if (run_void(cnt)) return;
    const uint32_t nslot = bounded(cnt->n_pair_slots, cnt->cap_slots, cnt), nseg = cnt->n_segments;
    const uint32_t np1 = g.n_phi + 1;
    uint2* __restrict__ pair2 = reinterpret_cast<uint2*>(pairs);  // (first hit, second hit), one 8-byte store
    uint32_t short_slots = 0;
    bool over = false;
    for (uint32_t slot = xcd_block(xcds) * blockDim.x + threadIdx.x; slot < nslot; slot += gridDim.x * blockDim.x) {
        uint32_t pos = slot_counts[slot];
        const uint32_t next_all = slot_counts[slot + 1];
        if (pos >= next_all) continue;
        // a slot that runs past the capacity writes the entries below it (the
        // run is void, but no entry below the capacity is left unwritten)
        if (next_all > n_pair_cap) over = true;
        const uint32_t next = next_all < n_pair_cap ? next_all : n_pair_cap;
        if (pos >= next) continue;
        const SlotInfo si = slot_info(tc, seg_base, nseg, slot, slot_seg);
        const uint32_t lay2 = tc.l2[si.p] - 1, lay1 = tc.l1[si.p] - 1;
        const uint32_t* g2 = grid + ((size_t)si.e * g.n_layers + lay2) * g.cells;
        const uint32_t i = g2[0] + si.k * tc.n_threads + si.t;
        const uint32_t* lg1 = grid + ((size_t)si.e * g.n_layers + lay1) * g.cells;
        const Window w = pair_window(g, tc, si.p, lay1, hx[i], hy[i], hz[i], hzp[i].y);
        // every z sector visits the index intervals [a1, b1) (j below the sector
        // end) and [a2, b2) (j past the end, wrapped back); the bitmap read of
        // the reference emits their union in ascending order
        for (uint32_t zs = w.zLo; zs < w.zHi && pos < next; ++zs) {
            const uint32_t zStart = lg1[zs * np1];
            const uint32_t zEnd = lg1[(zs + 1) * np1 - 1];
            const uint32_t zLen = zEnd - zStart;
            const uint32_t jStart = lg1[zs * np1 + w.pLo];
            const uint32_t end2 = (uint32_t)w.wrap * zEnd + lg1[zs * np1 + w.pHi] - (uint32_t)w.wrap * zStart;
            uint32_t a1 = jStart, b1 = end2 < zEnd ? end2 : zEnd;
            if (b1 < a1) b1 = a1;
            const uint32_t j2 = jStart > zEnd ? jStart : zEnd;
            uint32_t a2 = j2 - zLen, b2 = end2 > j2 ? end2 - zLen : a2;
            if (a2 < a1) {  // order the two intervals by start
                uint32_t ta = a1, tb = b1; a1 = a2; b1 = b2; a2 = ta; b2 = tb;
            }
            const uint32_t lo[2] = {a1, b1 > a2 ? b1 : a2};
            const uint32_t hi[2] = {b1, b2};
            for (int r = 0; r < 2 && pos < next; ++r)
                zphi_scan(hzp, lo[r], hi[r], [&](uint32_t index, float2 q) {
                    if (in_window(w, q)) {
                        pair2[pos] = make_uint2(index, i);
                        pos++;
                    }
                    return pos < next;
                });
        }
        for (uint32_t q = pos; q < next; ++q) pair2[q] = make_uint2(kSentinel, kSentinel);
        short_slots += next - pos;
    }
    if (short_slots) atomicAdd(&cnt->diag_pair_short, short_slots);
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

// slot_seg[b] = the segment of slot b << kSlotSegShift, for every block of
// slots up to the capacity (blocks past the last slot get the last segment)
__global__ void k_slot_segments(const uint32_t* __restrict__ seg_base, uint32_t n_entries,
                                uint32_t* __restrict__ slot_seg, const DevCounts* __restrict__ cnt) {
    const uint32_t nseg = cnt->n_segments;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n_entries; b += gridDim.x * blockDim.x)
        slot_seg[b] = find_segment(seg_base, nseg, b << kSlotSegShift);
}

__global__ void k_pair_offsets(const uint32_t* __restrict__ seg_base, const uint32_t* __restrict__ slot_counts,
                               uint32_t* __restrict__ pair_offsets, const DevCounts* __restrict__ cnt) {
    const uint32_t nseg = cnt->n_segments;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s <= nseg; s += gridDim.x * blockDim.x)
        pair_offsets[s] = (s == 0) ? 0u : slot_counts[seg_base[s]];
}

// ---------------------------------------------------------------------------
// predictor window (.h:77-188 count, .h:261-355 store)
struct PWin {
    Window w;
    const uint32_t* lg3;
    uint32_t ev, lt;
};

__device__ inline bool pred_window(const GridGeom& g, const TripletCols& tc, const uint32_t* grid,
                                   const float* hx, const float* hy, const float* hz, const float2* hzp,
                                   const uint32_t* hlayer, const uint32_t* hevt, uint32_t n_events,
                                   uint32_t firstHit, uint32_t secondHit, bool store_variant, PWin& pw) {
    const uint32_t ev = hevt[firstHit];
    const uint32_t lt = hlayer[firstHit] - 1;  // (.h:88): the triplet index is the inner layer - 1
    if (ev >= n_events || lt >= tc.n_trip) return false;
    const uint32_t layer = tc.l3[lt] - 1;
    if (layer >= g.n_layers || layer >= tc.n_geom) return false;
    pw.ev = ev;
    pw.lt = lt;
    pw.lg3 = grid + ((size_t)ev * g.n_layers + layer) * g.cells;
    Window& w = pw.w;
    const float sigmaZ = tc.sigma_z[lt], sigmaPhi = tc.sigma_phi[lt];
    const float minR = tc.min_r[layer], maxR = tc.max_r[layer];
    const float mrc = tc.min_radius_curvature;
    const float p1x = hx[firstHit], p1y = hy[firstHit], p1z = hz[firstHit];
    const float p2x = hx[secondHit], p2y = hy[secondHit], p2z = hz[secondHit];
    float tip2 = p1y * (p2x - p1x) - p1x * (p2y - p1y);
    tip2 *= tip2;
    tip2 /= ((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
    float rOrigin;
    if (!store_variant) {
        rOrigin = p1x * p1x + p1y * p1y;
        rOrigin = sqrtf(rOrigin - tip2);
    } else {
        rOrigin = sqrtf(p1x * p1x + p1y * p1y);
        rOrigin = sqrtf(rOrigin * rOrigin - tip2);
    }
    const float cotTheta = (p2z - p1z) / (sqrtf(p2x * p2x + p2y * p2y - tip2) - rOrigin);
    float tmp = p1z + (sqrtf(maxR * maxR - tip2) - rOrigin) * cotTheta;
    float zLow = p1z + (sqrtf(minR * minR - tip2) - rOrigin) * cotTheta;
    float zHigh = (float)(tmp > zLow) * tmp + (float)(tmp < zLow) * zLow;
    zLow = (float)(tmp > zLow) * zLow + (float)(tmp < zLow) * tmp;
    zLow -= sigmaZ;
    zHigh += sigmaZ;
    const int zl = tx_f2i_sat(floorf((zLow - g.min_z) / g.sector_z));
    w.zLo = (uint32_t)(zl > 0 ? zl : 0);
    const uint32_t zh = tx_f2u_sat(floorf((zHigh - g.min_z) / g.sector_z)) + 1u;
    w.zHi = zh < g.n_z ? zh : g.n_z;
    const float phi = hzp[secondHit].y;
    tmp = hzp[firstHit].y;
    const float PI = TX_PI_F;
    float dPhi = phi - tmp;
    dPhi += (dPhi > PI) ? -2 * PI : (dPhi < -PI) ? 2 * PI : 0;
    dPhi = fabsf(dPhi);
    const float dHits = sqrtf((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
    const float aD = tx_acosf(dHits / (2 * mrc));
    tmp = fabsf(aD - tc.acos_max_r[layer]);         // acos(maxR / (2 mrc))
    float phiLow = fabsf(aD - tc.acos_min_r[layer]);  // acos(minR / (2 mrc))
    const float m2 = (tmp < phiLow) ? phiLow : tmp;  // OpenCL max(x, y): y if x < y, else x
    dPhi = (dPhi < m2) ? m2 : dPhi;
    float phiHigh = phi + dPhi + sigmaPhi;
    phiLow = phi - dPhi - sigmaPhi;
    w.wrap = phiLow < -PI || phiHigh > PI || phiLow > PI || phiHigh < -PI;
    phiLow -= (float)((phiLow > PI) * 2) * PI;
    phiLow += (float)((phiLow < -PI) * 2) * PI;
    phiHigh -= (float)((phiHigh > PI) * 2) * PI;
    phiHigh += (float)((phiHigh < -PI) * 2) * PI;
    const uint32_t pl = tx_f2u_sat(floorf((phiLow - g.min_phi) / g.sector_phi));
    const uint32_t ph = tx_f2u_sat(floorf((phiHigh - g.min_phi) / g.sector_phi)) + 1u;
    w.pLo = pl < g.n_phi + 1 ? pl : g.n_phi + 1;  // in range for every finite window
    w.pHi = ph < g.n_phi ? ph : g.n_phi;
    w.zLow = zLow; w.zHigh = zHigh; w.phiLow = phiLow; w.phiHigh = phiHigh;
    return true;
}

// the reference's walk over the third-layer window (.h:190-217, .h:360-402):
// visit(index) for every hit of the z sectors and phi cells, in j order
template <class V>
__device__ inline void pred_walk(const Window& w, const uint32_t* lg3, uint32_t np1, V visit) {
    for (uint32_t zs = w.zLo; zs < w.zHi; ++zs) {
        const uint32_t zStart = lg3[zs * np1];
        const uint32_t zEnd = lg3[(zs + 1) * np1];
        const uint32_t zLen = zEnd - zStart;
        const uint32_t end2 = (uint32_t)w.wrap * zEnd + lg3[zs * np1 + w.pHi] - (uint32_t)w.wrap * zStart;
        for (uint32_t j = lg3[zs * np1 + w.pLo]; j < end2; ++j) {
            if (!visit(j - (uint32_t)(j >= zEnd) * zLen)) return;
        }
    }
}

__device__ inline bool same_cells(const Window& a, const Window& b) {
    return a.zLo == b.zLo && a.zHi == b.zHi && a.pLo == b.pLo && a.pHi == b.pHi && a.wrap == b.wrap;
}

__global__ void k_predict_count(GridGeom g, TripletCols tc, const uint32_t* __restrict__ grid,
                                const float* __restrict__ hx, const float* __restrict__ hy,
                                const float* __restrict__ hz, const float2* __restrict__ hzp,
                                const uint32_t* __restrict__ hlayer, const uint32_t* __restrict__ hevt,
                                const uint32_t* __restrict__ pairs, uint32_t* __restrict__ counts,
                                DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t n = bounded(cnt->n_pairs, cnt->cap_pairs, cnt), n_events = cnt->n_events, nh = cnt->n_hits;
    const uint32_t np1 = g.n_phi + 1;
    for (uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x; gid <= n; gid += gridDim.x * blockDim.x) {
        uint32_t nFound = 0;
        if (gid < n) {
            const uint32_t h1 = pairs[2 * (size_t)gid], h2 = pairs[2 * (size_t)gid + 1];
            PWin pw;
            if (pair_hits_ok(h1, h2, nh, cnt) &&
                pred_window(g, tc, grid, hx, hy, hz, hzp, hlayer, hevt, n_events, h1, h2, false, pw)) {
                pred_walk(pw.w, pw.lg3, np1, [&](uint32_t index) {
                    nFound += in_window(pw.w, hzp[index]);
                    return true;
                });
            }
        }
        counts[gid] = nFound;
    }
}

// triplet offsets at the end of each (event, layer triplet) group (.h:404-418)
// the offsets entry a pair writes as the last of its group, kSentinel when none
__device__ inline uint32_t pred_group_end_index(const TripletCols& tc, const uint32_t* hlayer, const uint32_t* hevt,
                                                const uint32_t* pairs, uint32_t n, uint32_t n_events, uint32_t gid,
                                                uint32_t h1, uint32_t nh) {
    if (h1 >= nh) return kSentinel;  // padding, or a bad index (reported by the pair's own check)
    const uint32_t ev = hevt[h1], lt = hlayer[h1] - 1;
    if (ev >= n_events || lt >= tc.n_trip) return kSentinel;
    // against the next pair that is not padding: padding pairs (short slots)
    // sit inside groups, and the reference's comparison with one (an
    // out-of-range hit read) lets several pairs of a group write its end, the
    // last one (in pair order) last.  Only that pair writes here, so the
    // result does not depend on the order in which pairs are processed.
    uint32_t j = gid + 1;
    while (j < n && pairs[2 * (size_t)j] == kSentinel) ++j;
    bool last = true;
    if (j < n) {
        const uint32_t h = pairs[2 * (size_t)j];
        if (h >= nh) return kSentinel;  // a bad index (reported by that pair's own check)
        last = (lt != hlayer[h] - 1) || (ev != hevt[h]);
    }
    return last ? ev * tc.n_trip + lt + 1 : kSentinel;
}

// returns the offsets entry written, kSentinel when none
__device__ inline uint32_t pred_group_end(const TripletCols& tc, const uint32_t* hlayer, const uint32_t* hevt,
                                          const uint32_t* pairs, uint32_t n, uint32_t n_events, uint32_t gid,
                                          uint32_t h1, uint32_t nh, uint32_t next, uint32_t* cand_offsets) {
    const uint32_t g = pred_group_end_index(tc, hlayer, hevt, pairs, n, n_events, gid, h1, nh);
    if (g != kSentinel) cand_offsets[g] = next;
    return g;
}

__global__ void k_predict_store(GridGeom g, TripletCols tc, const uint32_t* __restrict__ grid,
                                const float* __restrict__ hx, const float* __restrict__ hy,
                                const float* __restrict__ hz, const float2* __restrict__ hzp,
                                const uint32_t* __restrict__ hlayer, const uint32_t* __restrict__ hevt,
                                const uint32_t* __restrict__ pairs, const uint32_t* __restrict__ counts,
                                uint32_t* __restrict__ cands, uint32_t* __restrict__ cand_offsets,
                                uint32_t n_cand_cap, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t n = bounded(cnt->n_pairs, cnt->cap_pairs, cnt), n_events = cnt->n_events, nh = cnt->n_hits;
    const uint32_t np1 = g.n_phi + 1;
    uint32_t short_slots = 0, trunc = 0;
    bool over = false;
    for (uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x; gid < n; gid += gridDim.x * blockDim.x) {
        uint32_t pos = counts[gid];
        const uint32_t next_all = counts[gid + 1];
        // a pair whose candidates run past the capacity writes the entries
        // below it (the run is void, but none below the capacity is unwritten)
        if (next_all > n_cand_cap) over = true;
        const uint32_t next = next_all < n_cand_cap ? next_all : n_cand_cap;
        if (pos > next) pos = next;
        const uint32_t h1 = pairs[2 * (size_t)gid], h2 = pairs[2 * (size_t)gid + 1];
        PWin pw;
        const bool ok = pair_hits_ok(h1, h2, nh, cnt) &&
                        pred_window(g, tc, grid, hx, hy, hz, hzp, hlayer, hevt, n_events, h1, h2, true, pw);
        if (ok && pos < next) {
            bool more = false;
            pred_walk(pw.w, pw.lg3, np1, [&](uint32_t index) {
                const bool v = in_window(pw.w, hzp[index]);
                if (!(pos < next)) {
                    more = v;
                    return !v;
                }
                if (v) {
                    cands[2 * (size_t)pos] = gid;
                    cands[2 * (size_t)pos + 1] = index;
                    pos++;
                }
                return true;
            });
            trunc += more;
        }
        for (uint32_t q = pos; q < next; ++q) {
            cands[2 * (size_t)q] = kSentinel;
            cands[2 * (size_t)q + 1] = kSentinel;
        }
        short_slots += next - pos;
        pred_group_end(tc, hlayer, hevt, pairs, n, n_events, gid, h1, nh, next_all, cand_offsets);
    }
    if (short_slots) atomicAdd(&cnt->diag_pred_short, short_slots);
    if (trunc) atomicAdd(&cnt->diag_pred_trunc, trunc);
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

// Both predictor windows of one pair (.h:77-188 count, .h:261-355 store).
// They differ only in rOrigin (.h:130-131 vs .h:305-306), hence only in the z
// window; the phi window and every term without rOrigin are computed once
// (same operations in the same order, so the same bits as two separate
// evaluations).
struct PredWin2 {
    const uint32_t* lg3;
    float zLowA, zHighA, zLowB, zHighB, phiLow, phiHigh;
    uint32_t zLoA, zHiA, zLoB, zHiB, pLo, pHi;
    uint32_t ev, lt, el;  // el: the (event, layer) item of the third layer's grid
    bool wrap;
};

__device__ inline void pred_z(const GridGeom& g, float p1z, float p2z, float s2, float sMax, float sMin,
                              float sigmaZ, float rOrigin, float& zLow, float& zHigh, uint32_t& zLo, uint32_t& zHi) {
    const float cotTheta = (p2z - p1z) / (s2 - rOrigin);
    const float tmp = p1z + (sMax - rOrigin) * cotTheta;
    float zl = p1z + (sMin - rOrigin) * cotTheta;
    float zh = (float)(tmp > zl) * tmp + (float)(tmp < zl) * zl;
    zl = (float)(tmp > zl) * zl + (float)(tmp < zl) * tmp;
    zl -= sigmaZ;
    zh += sigmaZ;
    const int zlo = tx_f2i_sat(floorf((zl - g.min_z) / g.sector_z));
    zLo = (uint32_t)(zlo > 0 ? zlo : 0);
    const uint32_t zhi = tx_f2u_sat(floorf((zh - g.min_z) / g.sector_z)) + 1u;
    zHi = zhi < g.n_z ? zhi : g.n_z;
    zLow = zl;
    zHigh = zh;
}

__device__ inline bool pred_windows(const GridGeom& g, const TripletCols& tc, const uint32_t* grid,
                                    const float* hx, const float* hy, const float* hz, const float2* hzp,
                                    const uint32_t* hlayer, const uint32_t* hevt, uint32_t n_events,
                                    uint32_t firstHit, uint32_t secondHit, PredWin2& w) {
    const uint32_t ev = hevt[firstHit];
    const uint32_t lt = hlayer[firstHit] - 1;  // (.h:88): the triplet index is the inner layer - 1
    if (ev >= n_events || lt >= tc.n_trip) return false;
    const uint32_t layer = tc.l3[lt] - 1;
    if (layer >= g.n_layers || layer >= tc.n_geom) return false;
    w.ev = ev;
    w.lt = lt;
    w.el = ev * g.n_layers + layer;
    w.lg3 = grid + (size_t)w.el * g.cells;
    const float sigmaZ = tc.sigma_z[lt], sigmaPhi = tc.sigma_phi[lt];
    const float minR = tc.min_r[layer], maxR = tc.max_r[layer];
    const float mrc = tc.min_radius_curvature;
    const float p1x = hx[firstHit], p1y = hy[firstHit], p1z = hz[firstHit];
    const float p2x = hx[secondHit], p2y = hy[secondHit], p2z = hz[secondHit];
    float tip2 = p1y * (p2x - p1x) - p1x * (p2y - p1y);
    tip2 *= tip2;
    tip2 /= ((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
    const float s2 = sqrtf(p2x * p2x + p2y * p2y - tip2);
    const float sMax = sqrtf(maxR * maxR - tip2), sMin = sqrtf(minR * minR - tip2);
    const float r1sq = p1x * p1x + p1y * p1y;
    const float rOriginA = sqrtf(r1sq - tip2);  // count kernel (.h:130-131)
    const float r1 = sqrtf(r1sq);
    const float rOriginB = sqrtf(r1 * r1 - tip2);  // store kernel (.h:305-306)
    pred_z(g, p1z, p2z, s2, sMax, sMin, sigmaZ, rOriginA, w.zLowA, w.zHighA, w.zLoA, w.zHiA);
    pred_z(g, p1z, p2z, s2, sMax, sMin, sigmaZ, rOriginB, w.zLowB, w.zHighB, w.zLoB, w.zHiB);
    const float phi = hzp[secondHit].y;
    const float PI = TX_PI_F;
    float dPhi = phi - hzp[firstHit].y;
    dPhi += (dPhi > PI) ? -2 * PI : (dPhi < -PI) ? 2 * PI : 0;
    dPhi = fabsf(dPhi);
    const float dHits = sqrtf((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
    const float aD = tx_acosf(dHits / (2 * mrc));
    const float tmp = fabsf(aD - tc.acos_max_r[layer]);
    float phiLow = fabsf(aD - tc.acos_min_r[layer]);
    const float m2 = (tmp < phiLow) ? phiLow : tmp;  // OpenCL max(x, y): y if x < y, else x
    dPhi = (dPhi < m2) ? m2 : dPhi;
    float phiHigh = phi + dPhi + sigmaPhi;
    phiLow = phi - dPhi - sigmaPhi;
    w.wrap = phiLow < -PI || phiHigh > PI || phiLow > PI || phiHigh < -PI;
    phiLow -= (float)((phiLow > PI) * 2) * PI;
    phiLow += (float)((phiLow < -PI) * 2) * PI;
    phiHigh -= (float)((phiHigh > PI) * 2) * PI;
    phiHigh += (float)((phiHigh < -PI) * 2) * PI;
    const uint32_t pl = tx_f2u_sat(floorf((phiLow - g.min_phi) / g.sector_phi));
    const uint32_t ph = tx_f2u_sat(floorf((phiHigh - g.min_phi) / g.sector_phi)) + 1u;
    w.pLo = pl < g.n_phi + 1 ? pl : g.n_phi + 1;
    w.pHi = ph < g.n_phi ? ph : g.n_phi;
    w.phiLow = phiLow;
    w.phiHigh = phiHigh;
    return true;
}

__device__ inline bool phi_in(const PredWin2& w, float actPhi) {
    const float PI = TX_PI_F;
    return (!w.wrap && (w.phiLow <= actPhi && actPhi <= w.phiHigh)) ||
           (w.wrap && ((w.phiLow <= actPhi && actPhi <= PI) || (-PI <= actPhi && actPhi <= w.phiHigh)));
}

// One walk over the z sectors of both windows (their phi cells are the same):
// sector zs belongs to window A if zLoA <= zs < zHiA and to B likewise, and in
// every sector the hits are visited in the reference's j order, four loads in
// flight at a time.  visit(index, inA, inB) sees each hit once.
template <class V>
__device__ inline void pred_walk2(const PredWin2& w, uint32_t np1, const float2* __restrict__ hzp, V visit) {
    const bool ea = w.zLoA < w.zHiA, eb = w.zLoB < w.zHiB;
    if (!ea && !eb) return;
    const uint32_t z0 = !ea ? w.zLoB : (!eb ? w.zLoA : (w.zLoA < w.zLoB ? w.zLoA : w.zLoB));
    const uint32_t z1 = !ea ? w.zHiB : (!eb ? w.zHiA : (w.zHiA > w.zHiB ? w.zHiA : w.zHiB));
    const uint32_t* lg3 = w.lg3;
    for (uint32_t zs = z0; zs < z1; ++zs) {
        const bool sa = ea && zs >= w.zLoA && zs < w.zHiA;
        const bool sb = eb && zs >= w.zLoB && zs < w.zHiB;
        if (!sa && !sb) continue;
        const uint32_t zStart = lg3[zs * np1];
        const uint32_t zEnd = lg3[(zs + 1) * np1];
        const uint32_t zLen = zEnd - zStart;
        const uint32_t end2 = (uint32_t)w.wrap * zEnd + lg3[zs * np1 + w.pHi] - (uint32_t)w.wrap * zStart;
        for (uint32_t j = lg3[zs * np1 + w.pLo]; j < end2; j += 4) {
            float2 q[4];
            uint32_t idx[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t jj = j + u;
                idx[u] = jj - (uint32_t)(jj >= zEnd) * zLen;
                q[u] = jj < end2 ? hzp[idx[u]] : make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j + u < end2) visit(idx[u], q[u], sa, sb);
        }
    }
}

// The merged walk of both windows as one flat loop per lane: each iteration
// takes four hits of the current z sector (loads always issued, from a valid
// index, and masked), evaluates both windows with branch-free predicates,
// and moves to the next sector of either window when the current one ends.
// countA counts window A's matches; window B's matches go to `my` (16-bit
// offsets from `first`, at most cap of them) in the reference's visiting
// order, and nB counts them all.
__device__ inline void pred_walk_flat(const PredWin2& w, uint32_t np1, const float2* __restrict__ hzp, uint32_t cap,
                                      uint32_t first, uint16_t* my, uint32_t& countA, uint32_t& nB) {
    const bool ea = w.zLoA < w.zHiA, eb = w.zLoB < w.zHiB;
    uint32_t z0 = 0, z1 = 0;
    if (ea || eb) {
        z0 = !ea ? w.zLoB : (!eb ? w.zLoA : (w.zLoA < w.zLoB ? w.zLoA : w.zLoB));
        z1 = !ea ? w.zHiB : (!eb ? w.zHiA : (w.zHiA > w.zHiB ? w.zHiA : w.zHiB));
    }
    const uint32_t* lg3 = w.lg3;
    const float PI = TX_PI_F;
    const uint32_t wr = (uint32_t)w.wrap;
    uint32_t zs = z0, j = 0, end2 = 0, zEnd = 0, zLen = 0;
    bool sa = false, sb = false;
    // the next sector at or after zs that either window walks and that is not empty
    auto next_sector = [&]() {
        for (; zs < z1; ++zs) {
            sa = ea && zs >= w.zLoA && zs < w.zHiA;
            sb = eb && zs >= w.zLoB && zs < w.zHiB;
            if (!(sa || sb)) continue;
            const uint32_t zStart = lg3[zs * np1];
            zEnd = lg3[(zs + 1) * np1];
            zLen = zEnd - zStart;
            end2 = wr * zEnd + lg3[zs * np1 + w.pHi] - wr * zStart;
            j = lg3[zs * np1 + w.pLo];
            if (j < end2) return true;
        }
        return false;
    };
    bool active = next_sector();
    while (active) {
        float2 q[4];
        uint32_t idx[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t jj = j + u < end2 ? j + u : j;  // j < end2: always a hit of the sector
            idx[u] = jj - (uint32_t)(jj >= zEnd) * zLen;
            q[u] = hzp[idx[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float z = q[u].x, ph = q[u].y;
            const bool inr = j + u < end2;
            const bool p1 = (w.phiLow <= ph) & (ph <= w.phiHigh);
            const bool p2 = ((w.phiLow <= ph) & (ph <= PI)) | ((-PI <= ph) & (ph <= w.phiHigh));
            const bool pin = inr & ((w.wrap & p2) | (!w.wrap & p1));
            countA += (uint32_t)(pin & sa & (w.zLowA <= z) & (z <= w.zHighA));
            if (pin & sb & (w.zLowB <= z) & (z <= w.zHighB)) {
                if (nB < cap) my[nB] = (uint16_t)(idx[u] - first);
                nB++;
            }
        }
        j += 4;
        if (j >= end2) {
            ++zs;
            active = next_sector();
        }
    }
}

// Sub-row of a hit at height z in grid row zs: floor((z - row start) * F /
// sector_z) clamped to [0, F).  For a fixed row it is a non-decreasing function
// of z, which is all the fine walk relies on: every hit of row zs with
// zLow <= z <= zHigh lies in the sub-rows fine_bin(zLow) .. fine_bin(zHigh).
__device__ inline float fine_row_lo(const GridGeom& g, uint32_t zs) { return g.min_z + (float)zs * g.sector_z; }
__device__ inline uint32_t fine_bin(const FineGeom& fg, float row_lo, float z) {
    const int32_t s = tx_f2i_sat(floorf((z - row_lo) * fg.inv_sub));
    return s < 0 ? 0u : (s >= (int32_t)fg.F ? fg.F - 1u : (uint32_t)s);
}

// the fine rows [fLo, fHi) of a window whose grid rows are [zLo, zHi) and
// whose z test is zLow <= z <= zHigh: they hold every hit of those grid rows
// that passes the z test
__device__ inline void fine_range(const GridGeom& g, const FineGeom& fg, uint32_t zLo, uint32_t zHi, float zLow,
                                  float zHigh, uint32_t& fLo, uint32_t& fHi) {
    fLo = fHi = 0;
    if (zLo >= zHi) return;
    fLo = (zLo << fg.shift) + fine_bin(fg, fine_row_lo(g, zLo), zLow);
    fHi = ((zHi - 1) << fg.shift) + fine_bin(fg, fine_row_lo(g, zHi - 1), zHigh) + 1u;
    if (fHi < fLo) fHi = fLo;
}

// Fine structure of one (event, layer, grid row): the row's hits counted per
// (sub-row, phi cell) by the thread that owns the cell (cells hold their hits
// in input order, so a sequential pass keeps the order stable), an LDS scan
// in (sub-row, cell) order offset by the row's first hit, then the scatter.
// A block takes items in turn; items are (event, layer) * n_z + row.
__global__ __launch_bounds__(128) void k_fine_grid(GridGeom g, FineGeom fg, const uint32_t* __restrict__ grid,
                                                  const float2* __restrict__ hzp, uint32_t n_hit_cap,
                                                  const DevCounts* __restrict__ cnt, uint32_t* __restrict__ fgrid,
                                                  float4* __restrict__ fh) {
    extern __shared__ uint32_t s_cnt[];
    const uint32_t np1 = g.n_phi + 1, nb = fg.F * np1;
    uint32_t* wsum = s_cnt + nb;
    const uint32_t n_items = cnt->n_events * g.n_layers * g.n_z;
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const uint32_t el = item / g.n_z, zs = item - el * g.n_z;
        const uint32_t* lg = grid + (size_t)el * g.cells + (size_t)zs * np1;
        uint32_t* fo = fgrid + (size_t)el * fg.cells + (size_t)zs * nb;
        const float row_lo = fine_row_lo(g, zs);
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) s_cnt[i] = 0;
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < np1; c += blockDim.x) {
            const uint32_t p1 = lg[c + 1] < n_hit_cap ? lg[c + 1] : n_hit_cap;
            for (uint32_t p = lg[c]; p < p1; ++p) s_cnt[fine_bin(fg, row_lo, hzp[p].x) * np1 + c] += 1u;
        }
        __syncthreads();
        lds_excl_scan(s_cnt, nb, wsum);
        const uint32_t base = lg[0];
        for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) fo[i] = base + s_cnt[i];
        if (zs + 1 == g.n_z && threadIdx.x == 0) fo[nb] = lg[np1];  // the final entry: end of the last row
        __syncthreads();
        for (uint32_t c = threadIdx.x; c < np1; c += blockDim.x) {
            const uint32_t p1 = lg[c + 1] < n_hit_cap ? lg[c + 1] : n_hit_cap;
            for (uint32_t p = lg[c]; p < p1; ++p) {
                const float2 q = hzp[p];
                uint32_t& cur = s_cnt[fine_bin(fg, row_lo, q.x) * np1 + c];
                const uint32_t pos = base + cur;
                cur += 1u;
                if (pos < n_hit_cap) fh[pos] = make_float4(q.x, q.y, __uint_as_float(p), 0.0f);
            }
        }
        __syncthreads();
    }
}

// The merged walk of both windows over the fine rows (as pred_walk_flat over
// the grid rows, but each grid row is visited only in the sub-rows that can
// hold a hit passing the window's z test).  countA is the same number: the
// same hits pass.  Window B's matches are appended to `my` (at most cap, as
// 16-bit offsets of their grid indices from `first`) in walk order,
// (sub-row, phi segment, cell, grid index); the reference's order is (grid
// row, phi segment, grid index), the hits of the first phi segment (cells
// pLo..) before those of the wrapped one (cells ..pHi).  Entry k's key in that
// order is built from the bits the walk records: `rmask` bit k marks the first
// entry of a grid row (grid rows arrive in order), `smask` bit k an entry of
// the wrapped segment; entry k's grid-row ordinal is the number of rmask bits
// at or below k.  Matches beyond cap go to over(ordinal, wrapped, grid index).
// nB counts every match.
template <class M>
__device__ inline M fine_bit(uint32_t k) { return (M)1 << k; }
template <class M, class OV>
__device__ inline void pred_walk_fine(const PredWin2& w, const GridGeom& g, const FineGeom& fg,
                                      const uint32_t* __restrict__ fl, const float4* __restrict__ fh, uint32_t cap,
                                      uint32_t first, uint16_t* my, uint32_t& countA, uint32_t& nB, M& rmask,
                                      M& smask, OV&& over) {
    uint32_t aLo, aHi, bLo, bHi;
    fine_range(g, fg, w.zLoA, w.zHiA, w.zLowA, w.zHighA, aLo, aHi);
    fine_range(g, fg, w.zLoB, w.zHiB, w.zLowB, w.zHighB, bLo, bHi);
    const bool ea = aLo < aHi, eb = bLo < bHi;
    uint32_t f = 0, f1 = 0;
    if (ea || eb) {
        f = !ea ? bLo : (!eb ? aLo : (aLo < bLo ? aLo : bLo));
        f1 = !ea ? bHi : (!eb ? aHi : (aHi > bHi ? aHi : bHi));
    }
    const uint32_t np1 = g.n_phi + 1;
    const float PI = TX_PI_F;
    const uint32_t wr = (uint32_t)w.wrap;
    uint32_t row = 0xffffffffu, ord = 0;  // grid row of the last match, grid rows with matches so far
    // A sub-row's four prefix entries (its start, its end, the window's first
    // and last cell) are loaded one active sub-row ahead: the next sub-row's
    // loads are in flight while the current one's hits are tested.
    auto active_from = [&](uint32_t ff) {
        for (; ff < f1; ++ff)
            if ((ea && ff >= aLo && ff < aHi) || (eb && ff >= bLo && ff < bHi)) break;
        return ff;
    };
    struct RowPre {
        uint32_t f, start, end, lo, hi;
    };
    auto load_row = [&](uint32_t ff) {
        RowPre rp{ff, 0u, 0u, 0u, 0u};
        if (ff < f1) {
            const uint32_t* r = fl + (size_t)ff * np1;
            rp.start = r[0];
            rp.end = r[np1];
            rp.lo = r[w.pLo];
            rp.hi = r[w.pHi];
        }
        return rp;
    };
    RowPre cur = load_row(active_from(f));
    RowPre nxt = load_row(active_from(cur.f + 1));
    while (cur.f < f1) {
        const uint32_t fr = cur.f;
        const bool sa = ea && fr >= aLo && fr < aHi;
        const bool sb = eb && fr >= bLo && fr < bHi;
        const uint32_t fEnd = cur.end, fLen = fEnd - cur.start;
        const uint32_t end2 = wr * fEnd + cur.hi - wr * cur.start;
        uint32_t j = cur.lo;
        while (j < end2) {
            float4 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint32_t jj = j + u < end2 ? j + u : j;  // j < end2: always a hit of the row
                q[u] = fh[jj - (uint32_t)(jj >= fEnd) * fLen];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float z = q[u].x, ph = q[u].y;
                const bool inr = j + u < end2;
                const bool p1 = (w.phiLow <= ph) & (ph <= w.phiHigh);
                const bool p2 = ((w.phiLow <= ph) & (ph <= PI)) | ((-PI <= ph) & (ph <= w.phiHigh));
                const bool pin = inr & ((w.wrap & p2) | (!w.wrap & p1));
                countA += (uint32_t)(pin & sa & (w.zLowA <= z) & (z <= w.zHighA));
                if (pin & sb & (w.zLowB <= z) & (z <= w.zHighB)) {
                    const uint32_t rf = fr >> fg.shift;
                    const bool new_row = rf != row, seg2 = j + u >= fEnd;
                    row = rf;
                    ord += (uint32_t)new_row;
                    if (nB < cap) {
                        rmask |= new_row ? fine_bit<M>(nB) : (M)0;
                        smask |= seg2 ? fine_bit<M>(nB) : (M)0;
                        my[nB] = (uint16_t)(__float_as_uint(q[u].z) - first);
                    } else {
                        over(ord, seg2, __float_as_uint(q[u].z));
                    }
                    nB++;
                }
            }
            j += 4;
        }
        cur = nxt;
        if (cur.f < f1) nxt = load_row(active_from(cur.f + 1));
    }
}

// Bits lo .. hi-1 of a list mask (lo <= hi <= the mask's width).
template <class M>
__device__ inline M fine_bits(uint32_t lo, uint32_t hi) {
    constexpr uint32_t W = sizeof(M) * 8;
    const M upto = hi >= W ? ~(M)0 : (((M)1 << hi) - (M)1);
    const M below = lo >= W ? ~(M)0 : (((M)1 << lo) - (M)1);
    return upto & ~below;
}

// pred_walk_fine with the same results (countA, nB, the listed entries in the
// same order, rmask, smask and the overflow calls), in fewer instructions per
// visited hit:
//  - a row of a wrapped window is walked as its two index segments, [lo, row
//    end) and then [row start, hi), so a hit's index needs no wrap arithmetic;
//  - the phi test of a wrapped window is phiLow <= phi || phi <= phiHigh: every
//    hit's phi is an atan2 value in [-pi, pi] (as a float) or NaN, so the
//    reference's extra -pi <= phi and phi <= pi terms always hold;
//  - a match only appends its entry: rmask gets the bit of the first entry of
//    each grid row when that grid row is left (grid rows arrive in order),
//    smask the bits of the wrapped segment's entries when that segment ends,
//    and the grid-row ordinal that an overflowing entry carries is counted per
//    grid row.
template <class M, class OV>
__device__ inline void pred_walk_fine2(const PredWin2& w, const GridGeom& g, const FineGeom& fg,
                                       const uint32_t* __restrict__ fl, const float4* __restrict__ fh, uint32_t cap,
                                       uint32_t first, uint16_t* my, uint32_t& countA, uint32_t& nB, M& rmask,
                                       M& smask, OV&& over) {
    uint32_t aLo, aHi, bLo, bHi;
    fine_range(g, fg, w.zLoA, w.zHiA, w.zLowA, w.zHighA, aLo, aHi);
    fine_range(g, fg, w.zLoB, w.zHiB, w.zLowB, w.zHighB, bLo, bHi);
    const bool ea = aLo < aHi, eb = bLo < bHi;
    uint32_t f1 = 0, f = 0;
    if (ea || eb) {
        f = !ea ? bLo : (!eb ? aLo : (aLo < bLo ? aLo : bLo));
        f1 = !ea ? bHi : (!eb ? aHi : (aHi > bHi ? aHi : bHi));
    }
    const uint32_t np1 = g.n_phi + 1;
    const bool wrap = w.wrap;
    // the phi test as two intervals, (phiLow <= phi <= phiHigh) for an
    // unwrapped window and (phiLow <= phi) or (phi <= phiHigh) for a wrapped
    // one: [phiLow, hi1] or [lo2, phiHigh] (a NaN phi fails both)
    const float phiLow = w.phiLow, phiHigh = w.phiHigh;
    const float hi1 = wrap ? INFINITY : phiHigh, lo2 = wrap ? -INFINITY : INFINITY;
    const float zLowA = w.zLowA, zHighA = w.zHighA, zLowB = w.zLowB, zHighB = w.zHighB;
    // the grid row being walked, the entry count at its start, grid rows with matches before it
    uint32_t grow = 0xffffffffu, gstart = 0, ord_before = 0;
    auto close_grid_row = [&]() {
        if (nB > gstart) {
            if (gstart < cap) rmask |= fine_bit<M>(gstart);
            ++ord_before;
        }
    };
    auto active_from = [&](uint32_t ff) {
        for (; ff < f1; ++ff)
            if ((ea && ff >= aLo && ff < aHi) || (eb && ff >= bLo && ff < bHi)) break;
        return ff;
    };
    struct RowPre {
        uint32_t f, start, end, lo, hi;
    };
    auto load_row = [&](uint32_t ff) {
        RowPre rp{ff, 0u, 0u, 0u, 0u};
        if (ff < f1) {
            const uint32_t* r = fl + (size_t)ff * np1;
            rp.start = r[0];
            rp.end = r[np1];
            rp.lo = r[w.pLo];
            rp.hi = r[w.pHi];
        }
        return rp;
    };
    RowPre cur = load_row(active_from(f));
    RowPre nxt = load_row(active_from(cur.f + 1));
    while (cur.f < f1) {
        const uint32_t fr = cur.f;
        const bool sa = ea && fr >= aLo && fr < aHi;
        const bool sb = eb && fr >= bLo && fr < bHi;
        const uint32_t rf = fr >> fg.shift;
        if (rf != grow) {
            close_grid_row();
            grow = rf;
            gstart = nB;
        }
        const uint32_t ordc = ord_before + 1;
        // the row's first segment, then (wrapped window) its second
        uint32_t a = cur.lo, b = wrap ? cur.end : cur.hi;
        bool seg2 = false;
        while (true) {
            const uint32_t s0 = nB;
            for (uint32_t j = a; j < b; j += 4) {
                // entries j .. j + 3: those past b (at most three, within the
                // buffer's kFineSlack entries past its last hit) are masked below
                const float4* p = fh + j;
                float4 q[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) q[u] = p[u];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float z = q[u].x, ph = q[u].y;
                    const bool inr = j + u < b;
                    const bool pin = inr & (((phiLow <= ph) & (ph <= hi1)) | ((lo2 <= ph) & (ph <= phiHigh)));
                    countA += (uint32_t)(pin & sa & (zLowA <= z) & (z <= zHighA));
                    if (pin & sb & (zLowB <= z) & (z <= zHighB)) {
                        if (nB < cap) my[nB] = (uint16_t)(__float_as_uint(q[u].z) - first);
                        else over(ordc, seg2, __float_as_uint(q[u].z));
                        nB++;
                    }
                }
            }
            if (seg2) {
                const uint32_t s1 = nB < cap ? nB : cap;
                if (s0 < s1) smask |= fine_bits<M>(s0, s1);
                break;
            }
            if (!wrap) break;
            a = cur.start;
            b = cur.hi;
            seg2 = true;
        }
        cur = nxt;
        if (cur.f < f1) nxt = load_row(active_from(cur.f + 1));
    }
    close_grid_row();
}

// Count, scan and store of the predictor in one pass over tiles of
// kPredBlock pairs (batched path).  One merged walk counts the count-window
// matches and collects the store-window matches of every pair in LDS
// (kPredList per pair, as 16-bit offsets from the layer's first hit; a pair
// with more re-walks its window for the rest), the tile's output offset comes
// from a decoupled look-back, and the tile's candidates are written as one
// contiguous, coalesced run by all threads.
// Windows of a tile kept in LDS while the pairs are walked in another order.
template <int B>
struct TileWindows {
    float zLowA[B], zHighA[B], zLowB[B], zHighB[B], phiLow[B], phiHigh[B];
    uint32_t lg3[B];              // cell offset of the third layer's grid; kSentinel: no window
    uint32_t zA[B], zB[B], pr[B];  // zLo | zHi << 16 for A and B; pLo | pHi << 16 | wrap << 31
    uint32_t work[B];
};

template <int B>
__device__ inline void tile_store(TileWindows<B>& t, uint32_t i, const PredWin2& w, const uint32_t* grid) {
    t.zLowA[i] = w.zLowA; t.zHighA[i] = w.zHighA; t.zLowB[i] = w.zLowB; t.zHighB[i] = w.zHighB;
    t.phiLow[i] = w.phiLow; t.phiHigh[i] = w.phiHigh;
    t.lg3[i] = (uint32_t)(w.lg3 - grid);
    t.zA[i] = w.zLoA | w.zHiA << 16;
    t.zB[i] = w.zLoB | w.zHiB << 16;
    t.pr[i] = w.pLo | w.pHi << 16 | (uint32_t)w.wrap << 31;
}

template <int B>
__device__ inline void tile_load(const TileWindows<B>& t, uint32_t i, PredWin2& w, const uint32_t* grid,
                                 uint32_t g_cells) {
    w.zLowA = t.zLowA[i]; w.zHighA = t.zHighA[i]; w.zLowB = t.zLowB[i]; w.zHighB = t.zHighB[i];
    w.phiLow = t.phiLow[i]; w.phiHigh = t.phiHigh[i];
    w.lg3 = grid + t.lg3[i];
    w.el = t.lg3[i] / g_cells;
    w.zLoA = t.zA[i] & 0xffffu; w.zHiA = t.zA[i] >> 16;
    w.zLoB = t.zB[i] & 0xffffu; w.zHiB = t.zB[i] >> 16;
    w.pLo = t.pr[i] & 0xffffu; w.pHi = (t.pr[i] >> 16) & 0x7fffu; w.wrap = (t.pr[i] >> 31) != 0u;
}

// The store window's first `keep` matches in the reference's order (the grid
// walk of TripletThetaPhiPredictor's store kernel): f(k, grid index).
template <class F>
__device__ inline void grid_walk_store(const PredWin2& pw, uint32_t np1, const float2* __restrict__ hzp, uint32_t keep,
                                       F f) {
    uint32_t k = 0;
    for (uint32_t zs = pw.zLoB; zs < pw.zHiB && k < keep; ++zs) {
        const uint32_t zStart = pw.lg3[zs * np1];
        const uint32_t zEnd = pw.lg3[(zs + 1) * np1];
        const uint32_t zLen = zEnd - zStart;
        const uint32_t end2 = (uint32_t)pw.wrap * zEnd + pw.lg3[zs * np1 + pw.pHi] - (uint32_t)pw.wrap * zStart;
        for (uint32_t j = pw.lg3[zs * np1 + pw.pLo]; j < end2 && k < keep; ++j) {
            const uint32_t index = j - (uint32_t)(j >= zEnd) * zLen;
            const float2 q = hzp[index];
            if (phi_in(pw, q.y) && pw.zLowB <= q.x && q.x <= pw.zHighB) f(k++, index);
        }
    }
}

// hits a window walk visits (both windows' z sectors)
__device__ inline uint32_t walk_work(const PredWin2& w, uint32_t np1) {
    const bool ea = w.zLoA < w.zHiA, eb = w.zLoB < w.zHiB;
    if (!ea && !eb) return 0;
    const uint32_t z0 = !ea ? w.zLoB : (!eb ? w.zLoA : (w.zLoA < w.zLoB ? w.zLoA : w.zLoB));
    const uint32_t z1 = !ea ? w.zHiB : (!eb ? w.zHiA : (w.zHiA > w.zHiB ? w.zHiA : w.zHiB));
    uint32_t work = 0;
    for (uint32_t zs = z0; zs < z1; ++zs) {
        const uint32_t zStart = w.lg3[zs * np1], zEnd = w.lg3[(zs + 1) * np1];
        const uint32_t end2 = (uint32_t)w.wrap * zEnd + w.lg3[zs * np1 + w.pHi] - (uint32_t)w.wrap * zStart;
        const uint32_t j0 = w.lg3[zs * np1 + w.pLo];
        work += end2 > j0 ? end2 - j0 : 0u;
    }
    return work;
}

// kSort: the tile's pairs are walked in decreasing order of their walk
// length (16 logarithmic buckets), so that the lanes of a wave walk windows
// of similar length; each pair's results stay indexed by the pair.
// Six waves per SIMD: the walks are latency-bound, so the register budget is
// capped at 80 VGPRs (the variants that fit without spilling).
template <int kPredBlock, int kPredList, bool kSort>
__global__ __launch_bounds__(kPredBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_predict_fused(
    GridGeom g, TripletCols tc, const uint32_t* __restrict__ grid, const float* __restrict__ hx,
    const float* __restrict__ hy, const float* __restrict__ hz, const float2* __restrict__ hzp,
    const uint32_t* __restrict__ hlayer, const uint32_t* __restrict__ hevt, const uint32_t* __restrict__ pairs,
    uint32_t* __restrict__ cands, uint32_t* __restrict__ cand_offsets, uint32_t n_cand_cap,
    DevCounts* __restrict__ cnt, uint64_t* __restrict__ status, uint32_t* __restrict__ counter, FineGeom fg,
    const uint32_t* __restrict__ fgrid, const float4* __restrict__ fh) {
    constexpr int kSortB = kSort ? kPredBlock : 1;
    using Mask = typename std::conditional<(kPredList <= 32), uint32_t, uint64_t>::type;
    static_assert(kPredList <= 64, "list masks hold 64 entries");
    __shared__ uint16_t s_list[kPredBlock * kPredList];
    __shared__ uint32_t s_excl[kPredBlock + 1];  // output slots (count window) before each pair
    __shared__ uint32_t s_nb[kPredBlock];
    __shared__ uint32_t s_b3[kPredBlock];  // first hit of the pair's third layer; kSentinel: not listed
    __shared__ uint32_t s_wsum[kPredBlock / kWave + 1];
    __shared__ uint32_t s_tile, s_base;
    __shared__ TileWindows<kSortB> s_win;
    __shared__ uint32_t s_cnt[kSortB], s_perm[kSortB], s_hist[16], s_cur[16];
    const uint32_t tid = threadIdx.x;
    // a block that finds the run void claims no tile (every claimed tile is
    // processed in full, so no look-back waits on a tile that never publishes)
    if (block_run_void(cnt)) return;
    const uint32_t n = bounded(cnt->n_pairs, cnt->cap_pairs, cnt), n_events = cnt->n_events, nh = cnt->n_hits;
    const uint32_t n_tiles = (n + kPredBlock - 1) / kPredBlock;
    const uint32_t np1 = g.n_phi + 1;
    uint32_t short_slots = 0, trunc = 0;
    bool over = false;
    while (true) {
        if (tid == 0) s_tile = atomicAdd(counter, 1u);
        if (kSort && tid < 16) s_hist[tid] = 0;
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= (n_tiles ? n_tiles : 1u)) break;  // uniform: every thread read the same tile
        const uint32_t gid = tile * kPredBlock + tid;
        uint32_t countA = 0, nB = 0, b3 = kSentinel;
        uint32_t h1 = kSentinel;
        PredWin2 pw;
        bool have = false;
        if (gid < n) {
            h1 = pairs[2 * (size_t)gid];
            const uint32_t h2 = pairs[2 * (size_t)gid + 1];
            have = pair_hits_ok(h1, h2, nh, cnt) &&
                   pred_windows(g, tc, grid, hx, hy, hz, hzp, hlayer, hevt, n_events, h1, h2, pw);
        }
        // walk one pair's windows: count window A, list window B (by pair index)
        auto walk = [&](uint32_t pidx, const PredWin2& w, uint32_t& cA, uint32_t& nb, uint32_t& lb3) {
            const uint32_t first = w.lg3[0];
            // the list holds 16-bit offsets from the layer's first hit
            const bool listed = w.lg3[g.cells - 1] - first <= 0xffffu;
            lb3 = listed ? first : kSentinel;
            const uint32_t cap = listed ? (uint32_t)kPredList : 0u;
            uint16_t* my = s_list + pidx * kPredList;
            Mask rm = 0, sm = 0;
            pred_walk_fine<Mask>(w, g, fg, fgrid + (size_t)w.el * fg.cells, fh, cap, first, my, cA, nb, rm, sm,
                                 [](uint32_t, bool, uint32_t) {});
            // the list into the reference's order: each grid row's entries
            // (rows arrive in order) by (segment, grid index), insertion
            // sort carrying the segment bits; runs arrive sorted, so the
            // cost is the entries plus their inversions
            const uint32_t m = nb < cap ? nb : cap;
            uint32_t row0 = 0;
            for (uint32_t i = 1; i < m; ++i) {
                if ((rm >> i) & 1u) {
                    row0 = i;
                    continue;
                }
                const uint32_t x = my[i], xs = (uint32_t)((sm >> i) & 1u), kx = xs << 16 | x;
                uint32_t j = i;
                while (j > row0) {
                    const uint32_t y = my[j - 1], ys = (uint32_t)((sm >> (j - 1)) & 1u);
                    if ((ys << 16 | y) < kx) break;
                    my[j] = (uint16_t)y;
                    sm = (sm & ~fine_bit<Mask>(j)) | ((Mask)ys << j);
                    --j;
                }
                if (j != i) {
                    my[j] = (uint16_t)x;
                    sm = (sm & ~fine_bit<Mask>(j)) | ((Mask)xs << j);
                }
            }
            // a list without every match is not used: the grid walk writes them
            if (nb > cap) lb3 = kSentinel;
        };
        if (!kSort) {
            if (have) walk(tid, pw, countA, nB, b3);
        } else {
            // bucket the pairs by walk length, longest first
            uint32_t bucket = 0;
            if (have) {
                tile_store(s_win, tid, pw, grid);
                const uint32_t work = walk_work(pw, np1);
                bucket = work ? 31u - (uint32_t)__clz(work) + 1u : 0u;
                bucket = bucket > 15u ? 15u : bucket;
            } else {
                s_win.lg3[tid] = kSentinel;
            }
            atomicAdd(&s_hist[15 - bucket], 1u);
            __syncthreads();
            if (tid == 0) {
                uint32_t run = 0;
                for (int b = 0; b < 16; ++b) {
                    s_cur[b] = run;
                    run += s_hist[b];
                }
            }
            __syncthreads();
            s_perm[atomicAdd(&s_cur[15 - bucket], 1u)] = tid;
            __syncthreads();
            const uint32_t q = s_perm[tid];
            uint32_t cA = 0, nb = 0, lb3 = kSentinel;
            if (s_win.lg3[q] != kSentinel) {
                PredWin2 w;
                tile_load(s_win, q, w, grid, g.cells);
                walk(q, w, cA, nb, lb3);
            }
            s_cnt[q] = cA;
            s_nb[q] = nb;
            s_b3[q] = lb3;
            __syncthreads();
            countA = s_cnt[tid];
            nB = s_nb[tid];
            b3 = s_b3[tid];
        }
        uint32_t total;
        const uint32_t excl = block_excl_scan(countA, s_wsum, &total);
        if (tid < kWave) {
            const uint32_t b = tile_lookback(status, tile, total);
            if (tid == 0) s_base = b;
        }
        s_excl[tid] = excl;
        if (!kSort) {
            s_nb[tid] = nB;
            s_b3[tid] = b3;
        }
        if (tid == 0) s_excl[kPredBlock] = total;
        __syncthreads();
        const uint32_t base = s_base;
        if (tid == 0 && (tile + 1 >= n_tiles)) cnt->n_cands = base + total;
        const bool fits = (uint64_t)base + total <= n_cand_cap;
        over |= !fits;
        if (fits) {
            // coalesced write of the tile's run: slot s belongs to the pair
            // whose exclusive prefix is the last one <= s
            for (uint32_t s = tid; s < total; s += kPredBlock) {
                uint32_t lo = 0, hi = kPredBlock;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_excl[mid] <= s) lo = mid; else hi = mid;
                }
                const uint32_t k = s - s_excl[lo], nb = s_nb[lo], lb3 = s_b3[lo];
                const uint64_t out = 2 * ((uint64_t)base + s);
                if (k < nb) {
                    if (lb3 != kSentinel && k < (uint32_t)kPredList) {
                        cands[out] = tile * kPredBlock + lo;
                        cands[out + 1] = lb3 + s_list[lo * kPredList + k];
                    }
                } else {
                    cands[out] = kSentinel;
                    cands[out + 1] = kSentinel;
                }
            }
            const uint64_t o = (uint64_t)base + excl;
            // pairs whose matches are not all listed: the grid walk in the
            // reference's order
            const uint32_t keep = countA < nB ? countA : nB;
            if (b3 == kSentinel && keep > 0) {
                // the window again (rare path; keeps it out of registers during the walk)
                pred_windows(g, tc, grid, hx, hy, hz, hzp, hlayer, hevt, n_events, h1, pairs[2 * (size_t)gid + 1], pw);
                grid_walk_store(pw, np1, hzp, keep, [&](uint32_t k, uint32_t index) {
                    cands[2 * (o + k)] = gid;
                    cands[2 * (o + k) + 1] = index;
                });
            }
        }
        if (gid < n) {
            short_slots += countA > nB ? countA - nB : 0u;
            trunc += (countA > 0 && nB > countA);
            pred_group_end(tc, hlayer, hevt, pairs, n, n_events, gid, h1, nh, base + excl + countA, cand_offsets);
        }
        __syncthreads();  // the LDS lists are reused by the next tile
    }
    if (short_slots) atomicAdd(&cnt->diag_pred_short, short_slots);
    if (trunc) atomicAdd(&cnt->diag_pred_trunc, trunc);
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

// Predictor and filter in one pass (batched path with the filter fused).
// Per tile of kPredBlock pairs:
//  1. each pair walks its windows as in k_predict_fused: the count-window
//     matches are counted, the store-window matches (the candidates) are
//     listed in LDS (kPredList per pair, 16-bit offsets from the layer's first
//     hit) and the rest go to the block's pool in global memory (grid index,
//     grid-row ordinal, phi segment), so no pair walks twice;
//  2. a pair with more store-window matches than count-window slots keeps as
//     candidates only the first countA in the reference's order (a rank
//     over its matches; rare);
//  3. the tile's candidates are tested against the filter's cuts, by waves
//     of 64: the phi-cut prefilter first, the undecided ones queued in the
//     wave's LDS and tested 64 at a time with the full cuts;
//  4. every wave (a chunk of 64 consecutive pairs) takes room for its
//     tracklets in a scratch array from a range it holds (refilled with one
//     atomic per kChunkRange entries; chunks finish in any order, none waits
//     for another) and its pairs write their valid triplets there in the
//     reference's candidate order, (grid-row ordinal, phi segment, grid
//     index); the chunk records its candidate and tracklet counts and where
//     its tracklets went.
// The candidates are not stored.  Afterwards the chunk counts are scanned and
// each chunk's tracklets are copied to their place in pair order
// (launch_chunk_copy3), and the candidate group offsets, written chunk-local,
// get their chunk's offset (k_group_offsets).
template <int kPredBlock, int kPredList>
__global__ __launch_bounds__(kPredBlock) void k_predict_filter(
    GridGeom g, TripletCols tc, const uint32_t* __restrict__ grid, const float* __restrict__ hx,
    const float* __restrict__ hy, const float* __restrict__ hz, const float2* __restrict__ hzp,
    const uint32_t* __restrict__ hlayer, const uint32_t* __restrict__ hevt, const uint32_t* __restrict__ pairs,
    uint32_t* __restrict__ cand_offsets, uint32_t* __restrict__ th1, uint32_t* __restrict__ th2,
    uint32_t* __restrict__ th3, uint32_t scratch_cap, DevCounts* __restrict__ cnt, uint32_t* __restrict__ counter,
    FineGeom fg, const uint32_t* __restrict__ fgrid, const float4* __restrict__ fh, const float2* __restrict__ hxy,
    uint2* __restrict__ pools,
    uint32_t pool_cap, uint32_t* __restrict__ chunk_cands, uint32_t* __restrict__ chunk_trk,
    uint32_t* __restrict__ chunk_at, uint32_t* __restrict__ group_chunk) {
    using Mask = typename std::conditional<(kPredList <= 32), uint32_t, uint64_t>::type;
    static_assert(kPredList <= 64, "list masks hold 64 entries");
    constexpr uint32_t kWaves = kPredBlock / kWave, kQueue = 2 * kWave;
    constexpr uint32_t kValidBit = 0x80000000u, kDeadBit = 0x40000000u;  // pool entry flags (x word)
    __shared__ uint16_t s_list[kPredBlock * kPredList];
    __shared__ uint32_t s_eex[kPredBlock + 1];  // listed candidates before each pair
    __shared__ uint32_t s_first[kPredBlock];    // first hit of the pair's third layer
    __shared__ float4 s_pair[kPredBlock];       // x2, y2, x1, y1: the pair's hits for the prefilter and cuts
    __shared__ float2 s_z12[kPredBlock];        // z1, z2: with s_pair, the pair's hits for the angle cuts
    // blocks up to 256 pairs: the pair of each of the tile's first kOwnerCap
    // listed candidates (one byte each; later ones search the offsets)
    constexpr bool kOwner = kPredBlock <= 256;
    constexpr uint32_t kOwnerCap = kOwner ? kPredBlock * 16 : 1;
    __shared__ uint8_t s_owner[kOwnerCap];
    __shared__ uint32_t s_lt[kPredBlock];       // layer triplet | pool entries << 8
    __shared__ Mask s_valid[kPredBlock];        // listed candidates that pass the cuts
    __shared__ Mask s_dead[kPredBlock];         // listed matches that are not candidates
    __shared__ Mask s_rm[kPredBlock], s_sm[kPredBlock];
    __shared__ uint32_t s_pool_valid[kPredBlock];
    __shared__ uint32_t s_queue[kWaves * kQueue];
    __shared__ uint32_t s_queue2[kWaves * kQueue];
    __shared__ uint32_t s_wsum[kPredBlock / kWave + 1];
    __shared__ uint32_t s_tile, s_pool_n, s_dead_any;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
    const uint32_t n = bounded(cnt->n_pairs, cnt->cap_pairs, cnt), n_events = cnt->n_events, nh = cnt->n_hits;
    const uint32_t n_tiles = (n + kPredBlock - 1) / kPredBlock;
    if (blockIdx.x == 0 && tid == 0) cnt->n_pred_chunks = n_tiles * kWaves;
    uint2* pool = pools + (size_t)blockIdx.x * pool_cap;
    uint32_t short_slots = 0, trunc = 0;
    uint32_t range_at = 0, range_left = 0;  // the wave's scratch range (wave-uniform)
    bool over = false;
    if (tid == 0) {
        s_pool_n = 0;
        s_dead_any = 0;
    }
    // a run voided by an earlier kernel is not walked (this kernel's own pool
    // and scratch writes are bounded by their capacities)
    if (block_run_void(cnt)) return;
    // tiles from a work queue: tiles differ in cost, so blocks take the next one when free
    while (true) {
        if (tid == 0) s_tile = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t tile = s_tile;
        if (tile >= (n_tiles ? n_tiles : 1u)) break;  // uniform: every thread read the same tile
        uint32_t countA = 0, nB = 0, listed = 0, g_end = kSentinel;
        Mask rm = 0, sm = 0;
        {
            // the pair's hits and layer triplet go to LDS before the walk (fewer registers held across it)
            const uint32_t gid = tile * kPredBlock + tid;
            PredWin2 pw;
            bool have = false;
            uint32_t h1 = kSentinel, h2 = kSentinel, first = 0, lt = 0;
            if (gid < n) {
                h1 = pairs[2 * (size_t)gid];
                h2 = pairs[2 * (size_t)gid + 1];
                // whether the pair ends its candidate group (loads issued with the window's)
                g_end = pred_group_end_index(tc, hlayer, hevt, pairs, n, n_events, gid, h1, nh);
                have = pair_hits_ok(h1, h2, nh, cnt) &&
                       pred_windows(g, tc, grid, hx, hy, hz, hzp, hlayer, hevt, n_events, h1, h2, pw);
            }
            if (have) {
                lt = pw.lt;
                first = pw.lg3[0];
            }
            s_first[tid] = first;
            s_lt[tid] = lt;
            if (have) {
                const float2 p1 = hxy[h1], p2 = hxy[h2];
                s_pair[tid] = make_float4(p2.x, p2.y, p1.x, p1.y);
                s_z12[tid] = make_float2(hz[h1], hz[h2]);
            }
            if (have) {
                const uint32_t cap = pw.lg3[g.cells - 1] - first <= 0xffffu ? (uint32_t)kPredList : 0u;
                pred_walk_fine<Mask>(pw, g, fg, fgrid + (size_t)pw.el * fg.cells, fh, cap, first,
                                     s_list + tid * kPredList, countA, nB, rm, sm,
                                     [&](uint32_t ord, bool seg2, uint32_t h3) {
                                         const uint32_t slot = atomicAdd(&s_pool_n, 1u);
                                         if (slot < pool_cap)
                                             pool[slot] = make_uint2(tid | (uint32_t)seg2 << 16 | (ord & 0x1fffu) << 17,
                                                                     h3);
                                     });
                listed = nB < cap ? nB : cap;
            }
        }
        // the wave's candidates: chunk-local offsets and the chunk count
        const uint32_t chunk = tile * kWaves + wave;
        const uint32_t incl_c = wave_incl_scan(countA);
        const uint32_t excl_c = incl_c - countA;
        if (lane == kWave - 1 && tile < n_tiles) chunk_cands[chunk] = incl_c;
        uint32_t total_e;
        const uint32_t excl_e = block_excl_scan(listed, s_wsum, &total_e);
        if (kOwner)
            for (uint32_t i = excl_e; i < excl_e + listed && i < kOwnerCap; ++i) s_owner[i] = (uint8_t)tid;
        s_eex[tid] = excl_e;
        s_lt[tid] |= (nB - listed) << 8;
        s_valid[tid] = 0;
        s_dead[tid] = 0;
        s_rm[tid] = rm;
        s_sm[tid] = sm;
        s_pool_valid[tid] = 0;
        if (tid == 0) s_eex[kPredBlock] = total_e;
        if (nB > countA && nB > 0) s_dead_any = 1;
        __syncthreads();
        const uint32_t n_pool = s_pool_n < pool_cap ? s_pool_n : pool_cap;
        if (tid == 0 && s_pool_n > pool_cap) atomicOr(&cnt->err, kErrCapacity);
        if (tile * kPredBlock + tid < n) {
            short_slots += countA > nB ? countA - nB : 0u;
            trunc += (countA > 0 && nB > countA);
            // the group's end, chunk-local; k_group_offsets adds the chunk's offset
            if (g_end != kSentinel) {
                cand_offsets[g_end] = excl_c + countA;
                group_chunk[g_end] = chunk;
            }
        }
        // entry keys in the reference's order: (grid-row ordinal, phi segment, grid index)
        auto list_key = [&](uint32_t lo, uint32_t i) {
            const Mask r = s_rm[lo];
            const uint64_t ord = (uint64_t)__builtin_popcountll((uint64_t)(r & (((Mask)2 << i) - 1)));
            return ord << 33 | (uint64_t)((s_sm[lo] >> i) & 1u) << 32 | (s_first[lo] + s_list[lo * kPredList + i]);
        };
        auto pool_key = [](uint2 e) { return (uint64_t)((e.x >> 17) & 0x1fffu) << 33 | (uint64_t)((e.x >> 16) & 1u) << 32 | e.y; };
        // 2. only the first countA matches (in the reference's order) are candidates
        const bool prune = s_dead_any != 0;  // uniform: rarely any pair of a tile needs it
        if (prune && nB > countA && nB > 0) {
            const uint32_t n_own_pool = nB - listed;
            auto rank = [&](uint64_t k) {
                uint32_t r = 0;
                for (uint32_t i = 0; i < listed; ++i) r += list_key(tid, i) < k;
                if (n_own_pool)
                    for (uint32_t p = 0; p < n_pool; ++p) {
                        const uint2 e = pool[p];
                        r += (e.x & 0xffffu) == tid && pool_key(e) < k;
                    }
                return r;
            };
            Mask dead = 0;
            for (uint32_t i = 0; i < listed; ++i) dead |= rank(list_key(tid, i)) >= countA ? (Mask)1 << i : (Mask)0;
            s_dead[tid] = dead;
            if (n_own_pool)
                for (uint32_t p = 0; p < n_pool; ++p) {
                    const uint2 e = pool[p];
                    if ((e.x & 0xffffu) == tid && rank(pool_key(e)) >= countA) pool[p].x = e.x | kDeadBit;
                }
        }
        if (prune) __syncthreads();
        // 3. the candidates against the filter's cuts, by waves: listed entries
        //    [0, total_e), then the pool [total_e, total_e + n_pool)
        {
            uint32_t* q = s_queue + wave * kQueue;
            uint32_t qn = 0;
            const uint32_t n_all = total_e + n_pool;
            const uint32_t per = (n_all + kWaves - 1) / kWaves;
            const uint32_t e0 = wave * per, e1 = e0 + per < n_all ? e0 + per : n_all;
            // a candidate's pair (in the tile) and third hit: tag = lo << 16 | i
            // (listed) or kValidBit | p (pool)
            auto third_of = [&](uint32_t tag, uint32_t& c) {
                if (tag & kValidBit) {
                    const uint2 e = pool[tag & ~kValidBit];
                    c = e.y;
                    return e.x & 0xffffu;
                }
                const uint32_t lo = tag >> 16;
                c = s_first[lo] + s_list[lo * kPredList + (tag & 0xffffu)];
                return lo;
            };
            // The cuts are a conjunction of pure expressions, tested in two
            // compacted stages: the circle cuts first (they reject most of
            // what the phi prefilter keeps), then the phi and theta cuts (two
            // atan2) on the candidates left, 64 at a time from a second queue.
            uint32_t* q2 = s_queue2 + wave * kQueue;
            uint32_t qn2 = 0;
            auto test_angles = [&](uint32_t k0) {
                const uint32_t tag = q2[k0 + lane];
                uint32_t c;
                const uint32_t lo = third_of(tag, c);
                const float4 pr = s_pair[lo];  // x2, y2, x1, y1
                const float2 z12 = s_z12[lo];
                const float2 p3 = hxy[c];
                const uint32_t l = s_lt[lo] & 0xffu;
                // the pair's angles, for the few candidates that get this far
                const float phi12 = filter_pair_phi(pr.z, pr.w, pr.x, pr.y);
                const float theta12 = filter_pair_theta(pr.z, pr.w, z12.x, pr.x, pr.y, z12.y);
                if (filter_phi_cut(tc, l, pr.x, pr.y, p3.x, p3.y, phi12) &&
                    filter_theta_cut(tc, l, pr.x, pr.y, z12.y, p3.x, p3.y, hz[c], theta12)) {
                    if (tag & kValidBit) {
                        atomicAdd(&s_pool_valid[lo], 1u);
                        pool[tag & ~kValidBit].x |= kValidBit;
                    } else {
                        atomicOr(&s_valid[lo], (Mask)1 << (tag & 0xffffu));
                    }
                }
            };
            // called by the whole wave: the circle cuts on q[k0 + lane] for
            // the active lanes, those passing appended to the second queue
            auto test_queue = [&](uint32_t k0, bool active) {
                bool pass = false;
                uint32_t tag = 0;
                if (active) {
                    tag = q[k0 + lane];
                    uint32_t c;
                    const uint32_t lo = third_of(tag, c);
                    const float4 pr = s_pair[lo];  // the pair's hits, read once per tile from HBM
                    const float2 p3 = hxy[c];
                    pass = filter_circle_cuts(tc, s_lt[lo] & 0xffu, pr.z, pr.w, pr.x, pr.y, p3.x, p3.y);
                }
                const uint64_t m = __ballot(pass);
                if (pass) q2[qn2 + (uint32_t)__popcll(m & lanemask_lt())] = tag;
                qn2 += (uint32_t)__popcll(m);
                wave_sync();
                if (qn2 >= kWave) {
                    qn2 -= kWave;
                    test_angles(qn2);
                    wave_sync();
                }
            };
            for (uint32_t eb = e0; eb < e1; eb += kWave) {
                const uint32_t e = eb + lane;
                bool undecided = false;
                uint32_t tag = 0;
                if (e < e1) {
                    if (e < total_e) {
                        uint32_t lo = 0;
                        if (kOwner && e < kOwnerCap) {
                            lo = s_owner[e];
                        } else {
                            uint32_t hi = kPredBlock;
                            while (hi - lo > 1) {
                                const uint32_t mid = (lo + hi) >> 1;
                                if (s_eex[mid] <= e) lo = mid; else hi = mid;
                            }
                        }
                        const uint32_t i = e - s_eex[lo];
                        tag = lo << 16 | i;
                        undecided = !((s_dead[lo] >> i) & 1u);
                    } else {
                        tag = kValidBit | (e - total_e);
                        undecided = !(pool[e - total_e].x & kDeadBit);
                    }
                    if (undecided) {
                        uint32_t c;
                        const uint32_t lo = third_of(tag, c);
                        const float4 pr = s_pair[lo];
                        const float2 p3 = hxy[c];
                        undecided = !tx_phi_surely_outside(pr.x - pr.z, pr.y - pr.w, p3.x - pr.x, p3.y - pr.y,
                                                           tc.tan_dphi[s_lt[lo] & 0xffu]);
                    }
                }
                const uint64_t m = __ballot(undecided);
                if (undecided) q[qn + (uint32_t)__popcll(m & lanemask_lt())] = tag;
                qn += (uint32_t)__popcll(m);
                wave_sync();
                if (qn >= kWave) {
                    qn -= kWave;
                    test_queue(qn, true);
                    wave_sync();
                }
            }
            test_queue(0, lane < qn);
            if (lane < qn2) test_angles(0);
        }
        __syncthreads();
        // 4. the pair's valid triplets: the wave's room in the scratch, then
        //    every pair's triplets in the reference's order
        const Mask vm = s_valid[tid];
        const uint32_t nv = (uint32_t)__builtin_popcountll((uint64_t)vm) + s_pool_valid[tid];
        const uint32_t incl_t = wave_incl_scan(nv);
        const uint32_t total_t = __shfl(incl_t, kWave - 1, kWave);
        if (total_t > range_left) {  // uniform
            const uint32_t size = total_t > kChunkRange ? total_t : kChunkRange;
            uint32_t a = 0;
            if (lane == 0) a = atomicAdd(&cnt->trk_alloc, size);
            range_at = __shfl(a, 0, kWave);
            range_left = size;
        }
        const uint32_t at = range_at;
        range_at += total_t;
        range_left -= total_t;
        if (lane == kWave - 1 && tile < n_tiles) {
            chunk_trk[chunk] = total_t;
            chunk_at[chunk] = at;
        }
        if (nv) {
            const uint32_t o = at + incl_t - nv, pv = s_pool_valid[tid];
            auto write = [&](uint64_t k, uint32_t h3) {
                uint32_t r = 0;
                for (Mask m = vm; m; m &= m - 1) r += list_key(tid, (uint32_t)__builtin_ctzll((uint64_t)m)) < k;
                if (pv)
                    for (uint32_t p = 0; p < n_pool; ++p) {
                        const uint2 e = pool[p];
                        r += (e.x & 0xffffu) == tid && (e.x & kValidBit) && pool_key(e) < k;
                    }
                const uint64_t pos = (uint64_t)o + r;
                if (pos < scratch_cap) {
                    th1[pos] = pairs[2 * (size_t)(tile * kPredBlock + tid)];
                    th2[pos] = pairs[2 * (size_t)(tile * kPredBlock + tid) + 1];
                    th3[pos] = h3;
                } else {
                    over = true;
                }
            };
            for (Mask m = vm; m; m &= m - 1) {
                const uint32_t i = (uint32_t)__builtin_ctzll((uint64_t)m);
                write(list_key(tid, i), s_first[tid] + s_list[tid * kPredList + i]);
            }
            if (pv)
                for (uint32_t p = 0; p < n_pool; ++p) {
                    const uint2 e = pool[p];
                    if ((e.x & 0xffffu) == tid && (e.x & kValidBit)) write(pool_key(e), e.y);
                }
        }
        if (tid == 0) {  // read into registers above; cleared for the next tile
            s_pool_n = 0;
            s_dead_any = 0;
        }
        __syncthreads();  // the LDS lists and the pool are reused by the next tile
    }
    if (short_slots) atomicAdd(&cnt->diag_pred_short, short_slots);
    if (trunc) atomicAdd(&cnt->diag_pred_trunc, trunc);
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

// The same pass with every wave on its own: a wave claims a chunk of 64
// pairs, walks them, tests their candidates and writes their tracklets with
// wave-level synchronization only (its own LDS lists, candidate offsets and a
// quarter of the block's pool), then claims the next chunk.  No wave waits at
// a block barrier for the block's slowest walk.  Results are those of
// k_predict_filter: the chunk is the same 64 pairs, its candidates the same
// in the same order.
template <int kPredBlock, int kPredList>
__global__ __launch_bounds__(kPredBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_predict_filter_wave(
    GridGeom g, TripletCols tc, const uint32_t* __restrict__ grid, const float* __restrict__ hx,
    const float* __restrict__ hy, const float* __restrict__ hz, const float2* __restrict__ hzp,
    const uint32_t* __restrict__ hlayer, const uint32_t* __restrict__ hevt, const uint32_t* __restrict__ pairs,
    uint32_t* __restrict__ cand_offsets, uint32_t* __restrict__ th1, uint32_t* __restrict__ th2,
    uint32_t* __restrict__ th3, uint32_t scratch_cap, DevCounts* __restrict__ cnt, uint32_t* __restrict__ counter,
    FineGeom fg, const uint32_t* __restrict__ fgrid, const float4* __restrict__ fh, const float2* __restrict__ hxy,
    uint2* __restrict__ pools,
    uint32_t pool_cap, uint32_t* __restrict__ chunk_cands, uint32_t* __restrict__ chunk_trk,
    uint32_t* __restrict__ chunk_at, uint32_t* __restrict__ group_chunk, uint32_t per_claim) {
    using Mask = typename std::conditional<(kPredList <= 32), uint32_t, uint64_t>::type;
    static_assert(kPredList <= 64, "list masks hold 64 entries");
    constexpr uint32_t kWaves = kPredBlock / kWave, kQueue = 2 * kWave;
    constexpr uint32_t kValidBit = 0x80000000u, kDeadBit = 0x40000000u;  // pool entry flags (x word)
    __shared__ uint16_t s_list[kPredBlock * kPredList];
    __shared__ uint32_t s_eex[kPredBlock + kWaves];  // per wave: listed candidates before each pair, then the total
    __shared__ uint32_t s_first[kPredBlock];    // first hit of the pair's third layer
    __shared__ float4 s_pair[kPredBlock];       // x2, y2, x1, y1: the pair's hits for the prefilter and cuts
    __shared__ float2 s_z12[kPredBlock];        // z1, z2: with s_pair, the pair's hits for the angle cuts
    // per wave: the pair (lane) of each of the chunk's first kOwnerCap listed
    // candidates (one byte each; later ones search the offsets)
    constexpr uint32_t kOwnerCap = kWave * 7;
    __shared__ uint8_t s_owner[kWaves * kOwnerCap];
    __shared__ uint32_t s_lt[kPredBlock];       // layer triplet | pool entries << 8
    __shared__ Mask s_valid[kPredBlock];        // listed candidates that pass the cuts
    __shared__ Mask s_dead[kPredBlock];         // listed matches that are not candidates
    __shared__ Mask s_rm[kPredBlock], s_sm[kPredBlock];
    __shared__ uint32_t s_pool_valid[kPredBlock];
    __shared__ uint32_t s_queue[kWaves * kQueue];
    __shared__ float2 s_qxy[kWaves * kQueue];  // the third hit (x, y) of each queued candidate
    __shared__ uint32_t s_queue2[kWaves * kQueue];
    __shared__ uint32_t s_pool_n_w[kWaves], s_dead_any_w[kWaves];
    const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
    const uint32_t n = bounded(cnt->n_pairs, cnt->cap_pairs, cnt), n_events = cnt->n_events, nh = cnt->n_hits;
    const uint32_t n_chunks = (n + kWave - 1) / kWave;
    if (blockIdx.x == 0 && tid == 0) cnt->n_pred_chunks = n_chunks;
    // the wave's pool: its share of the block's
    const uint32_t pw_cap = pool_cap / kWaves;
    uint2* pool = pools + (size_t)blockIdx.x * pool_cap + (size_t)wave * pw_cap;
    uint32_t* eex = s_eex + wave * (kWave + 1);
    uint8_t* owner = s_owner + wave * kOwnerCap;
    uint32_t& s_pool_n = s_pool_n_w[wave];
    uint32_t& s_dead_any = s_dead_any_w[wave];
    uint32_t short_slots = 0, trunc = 0;
    uint32_t range_at = 0, range_left = 0;  // the wave's scratch range (wave-uniform)
    bool over = false;
    if (lane == 0) {
        s_pool_n = 0;
        s_dead_any = 0;
    }
    // a run voided by an earlier kernel is not walked (this kernel's own pool
    // and scratch writes are bounded by their capacities)
    if (block_run_void(cnt)) return;
    // chunks of 64 pairs from work queues: chunks differ in cost, so a wave
    // takes the next ones when free and never waits for the block's other
    // waves.  The chunks are split into kPredRegions contiguous regions, one
    // counter each (kPredCounterStride words apart); a block starts on the
    // region of its XCD (blocks go to the XCDs round-robin), so the waves of
    // one XCD work near one front of the pairs and its L2 holds their hits,
    // and a wave whose region is done moves on to the next.  A claim takes
    // per_claim consecutive chunks.
    const uint32_t n_work = n_chunks ? n_chunks : 1u;
    const uint32_t home = blockIdx.x % kPredRegions;
    uint32_t chunk = 0, left = 0, region = 0;
    while (true) {
        if (left == 0) {
            for (; region < kPredRegions; ++region) {
                const uint32_t r = (home + region) % kPredRegions;
                const uint32_t r0 = (uint32_t)((uint64_t)n_work * r / kPredRegions);
                const uint32_t r1 = (uint32_t)((uint64_t)n_work * (r + 1) / kPredRegions);
                uint32_t claimed = 0;
                if (lane == 0) claimed = atomicAdd(&counter[r * kPredCounterStride], per_claim);
                claimed = __shfl(claimed, 0, kWave);
                if (claimed < r1 - r0) {
                    chunk = r0 + claimed;
                    left = per_claim < r1 - chunk ? per_claim : r1 - chunk;
                    break;
                }
            }
            if (region >= kPredRegions) break;  // wave-uniform: every region is done
        }
        uint32_t countA = 0, nB = 0, listed = 0, g_end = kSentinel;
        Mask rm = 0, sm = 0;
        {
            // the pair's hits and layer triplet go to LDS before the walk (fewer registers held across it)
            const uint32_t gid = chunk * kWave + lane;
            PredWin2 pw;
            bool have = false;
            uint32_t h1 = kSentinel, h2 = kSentinel, first = 0, lt = 0;
            if (gid < n) {
                h1 = pairs[2 * (size_t)gid];
                h2 = pairs[2 * (size_t)gid + 1];
                // whether the pair ends its candidate group (loads issued with the window's)
                g_end = pred_group_end_index(tc, hlayer, hevt, pairs, n, n_events, gid, h1, nh);
                have = pair_hits_ok(h1, h2, nh, cnt) &&
                       pred_windows(g, tc, grid, hx, hy, hz, hzp, hlayer, hevt, n_events, h1, h2, pw);
            }
            if (have) {
                lt = pw.lt;
                first = pw.lg3[0];
            }
            s_first[tid] = first;
            s_lt[tid] = lt;
            if (have) {
                const float2 p1 = hxy[h1], p2 = hxy[h2];
                s_pair[tid] = make_float4(p2.x, p2.y, p1.x, p1.y);
                s_z12[tid] = make_float2(hz[h1], hz[h2]);
            }
            if (have) {
                const uint32_t cap = pw.lg3[g.cells - 1] - first <= 0xffffu ? (uint32_t)kPredList : 0u;
                pred_walk_fine2<Mask>(pw, g, fg, fgrid + (size_t)pw.el * fg.cells, fh, cap, first,
                                     s_list + tid * kPredList, countA, nB, rm, sm,
                                     [&](uint32_t ord, bool seg2, uint32_t h3) {
                                         const uint32_t slot = atomicAdd(&s_pool_n, 1u);
                                         if (slot < pw_cap)
                                             pool[slot] = make_uint2(tid | (uint32_t)seg2 << 16 | (ord & 0x1fffu) << 17,
                                                                     h3);
                                     });
                listed = nB < cap ? nB : cap;
            }
        }
        // the wave's candidates: chunk-local offsets and the chunk count
        const uint32_t incl_c = wave_incl_scan(countA);
        const uint32_t excl_c = incl_c - countA;
        if (lane == kWave - 1) chunk_cands[chunk] = incl_c;
        const uint32_t incl_e = wave_incl_scan(listed);
        const uint32_t excl_e = incl_e - listed, total_e = __shfl(incl_e, kWave - 1, kWave);
        for (uint32_t i = excl_e; i < excl_e + listed && i < kOwnerCap; ++i) owner[i] = (uint8_t)lane;
        eex[lane] = excl_e;
        s_lt[tid] |= (nB - listed) << 8;
        s_valid[tid] = 0;
        s_dead[tid] = 0;
        s_rm[tid] = rm;
        s_sm[tid] = sm;
        s_pool_valid[tid] = 0;
        if (lane == 0) eex[kWave] = total_e;
        if (nB > countA && nB > 0) s_dead_any = 1;
        wave_sync();
        const uint32_t n_pool = s_pool_n < pw_cap ? s_pool_n : pw_cap;
        if (lane == 0 && s_pool_n > pw_cap) atomicOr(&cnt->err, kErrCapacity);
        if (chunk * kWave + lane < n) {
            short_slots += countA > nB ? countA - nB : 0u;
            trunc += (countA > 0 && nB > countA);
            // the group's end, chunk-local; k_group_offsets adds the chunk's offset
            if (g_end != kSentinel) {
                cand_offsets[g_end] = excl_c + countA;
                group_chunk[g_end] = chunk;
            }
        }
        // entry keys in the reference's order: (grid-row ordinal, phi segment, grid index)
        auto list_key = [&](uint32_t lo, uint32_t i) {
            const Mask r = s_rm[lo];
            const uint64_t ord = (uint64_t)__builtin_popcountll((uint64_t)(r & (((Mask)2 << i) - 1)));
            return ord << 33 | (uint64_t)((s_sm[lo] >> i) & 1u) << 32 | (s_first[lo] + s_list[lo * kPredList + i]);
        };
        auto pool_key = [](uint2 e) { return (uint64_t)((e.x >> 17) & 0x1fffu) << 33 | (uint64_t)((e.x >> 16) & 1u) << 32 | e.y; };
        // 2. only the first countA matches (in the reference's order) are candidates
        const bool prune = s_dead_any != 0;  // uniform: rarely any pair of a chunk needs it
        if (prune && nB > countA && nB > 0) {
            const uint32_t n_own_pool = nB - listed;
            auto rank = [&](uint64_t k) {
                uint32_t r = 0;
                for (uint32_t i = 0; i < listed; ++i) r += list_key(tid, i) < k;
                if (n_own_pool)
                    for (uint32_t p = 0; p < n_pool; ++p) {
                        const uint2 e = pool[p];
                        r += (e.x & 0xffffu) == tid && pool_key(e) < k;
                    }
                return r;
            };
            Mask dead = 0;
            for (uint32_t i = 0; i < listed; ++i) dead |= rank(list_key(tid, i)) >= countA ? (Mask)1 << i : (Mask)0;
            s_dead[tid] = dead;
            if (n_own_pool)
                for (uint32_t p = 0; p < n_pool; ++p) {
                    const uint2 e = pool[p];
                    if ((e.x & 0xffffu) == tid && rank(pool_key(e)) >= countA) pool[p].x = e.x | kDeadBit;
                }
        }
        if (prune) wave_sync();
        // 3. the chunk's candidates against the filter's cuts: listed entries
        //    [0, total_e), then the pool [total_e, total_e + n_pool)
        {
            uint32_t* q = s_queue + wave * kQueue;
            float2* qxy = s_qxy + wave * kQueue;
            uint32_t qn = 0;
            const uint32_t n_all = total_e + n_pool;
            const uint32_t e0 = 0, e1 = n_all;
            // a candidate's pair (in the block) and third hit: tag = lo << 16 | i
            // (listed) or kValidBit | p (pool)
            auto third_of = [&](uint32_t tag, uint32_t& c) {
                if (tag & kValidBit) {
                    const uint2 e = pool[tag & ~kValidBit];
                    c = e.y;
                    return e.x & 0xffffu;
                }
                const uint32_t lo = tag >> 16;
                c = s_first[lo] + s_list[lo * kPredList + (tag & 0xffffu)];
                return lo;
            };
            // The cuts are a conjunction of pure expressions, tested in two
            // compacted stages: the circle cuts first (they reject most of
            // what the phi prefilter keeps), then the phi and theta cuts (two
            // atan2) on the candidates left, 64 at a time from a second queue.
            uint32_t* q2 = s_queue2 + wave * kQueue;
            uint32_t qn2 = 0;
            auto test_angles = [&](uint32_t k0) {
                const uint32_t tag = q2[k0 + lane];
                uint32_t c;
                const uint32_t lo = third_of(tag, c);
                const float4 pr = s_pair[lo];  // x2, y2, x1, y1
                const float2 z12 = s_z12[lo];
                const float2 p3 = hxy[c];
                const uint32_t l = s_lt[lo] & 0xffu;
                // the pair's angles, for the few candidates that get this far
                const float phi12 = filter_pair_phi(pr.z, pr.w, pr.x, pr.y);
                const float theta12 = filter_pair_theta(pr.z, pr.w, z12.x, pr.x, pr.y, z12.y);
                if (filter_phi_cut(tc, l, pr.x, pr.y, p3.x, p3.y, phi12) &&
                    filter_theta_cut(tc, l, pr.x, pr.y, z12.y, p3.x, p3.y, hz[c], theta12)) {
                    if (tag & kValidBit) {
                        atomicAdd(&s_pool_valid[lo], 1u);
                        pool[tag & ~kValidBit].x |= kValidBit;
                    } else {
                        atomicOr(&s_valid[lo], (Mask)1 << (tag & 0xffffu));
                    }
                }
            };
            // called by the whole wave: the circle cuts on q[k0 + lane] for
            // the active lanes, those passing appended to the second queue
            auto test_queue = [&](uint32_t k0, bool active) {
                bool pass = false;
                uint32_t tag = 0;
                if (active) {
                    tag = q[k0 + lane];
                    const uint32_t lo = (tag & kValidBit) ? pool[tag & ~kValidBit].x & 0xffffu : tag >> 16;
                    const float4 pr = s_pair[lo];  // the pair's hits, read once per chunk from HBM
                    const float2 p3 = qxy[k0 + lane];  // loaded by the prefilter
                    pass = filter_circle_cuts(tc, s_lt[lo] & 0xffu, pr.z, pr.w, pr.x, pr.y, p3.x, p3.y);
                }
                const uint64_t m = __ballot(pass);
                if (pass) q2[qn2 + (uint32_t)__popcll(m & lanemask_lt())] = tag;
                qn2 += (uint32_t)__popcll(m);
                wave_sync();
                if (qn2 >= kWave) {
                    qn2 -= kWave;
                    test_angles(qn2);
                    wave_sync();
                }
            };
            // the candidate at e: its tag, whether it is still undecided
            // (not pruned) and its pair (in the block) and third hit
            auto pick = [&](uint32_t e, uint32_t& tag, uint32_t& lo, uint32_t& c) {
                bool undecided = false;
                tag = 0;
                lo = 0;
                c = 0;
                if (e < e1) {
                    if (e < total_e) {
                        uint32_t ol = 0;
                        if (e < kOwnerCap) {
                            ol = owner[e];
                        } else {
                            uint32_t hi = kWave;
                            while (hi - ol > 1) {
                                const uint32_t mid = (ol + hi) >> 1;
                                if (eex[mid] <= e) ol = mid; else hi = mid;
                            }
                        }
                        const uint32_t i = e - eex[ol];
                        lo = wave * kWave + ol;
                        tag = lo << 16 | i;
                        undecided = !((s_dead[lo] >> i) & 1u);
                    } else {
                        tag = kValidBit | (e - total_e);
                        undecided = !(pool[e - total_e].x & kDeadBit);
                    }
                    if (undecided) lo = third_of(tag, c);
                }
                return undecided;
            };
            // whether the prefilter leaves the candidate undecided, given its third hit
            auto prefilter = [&](bool undecided, uint32_t lo, float2 p3) {
                if (!undecided) return false;
                const float4 pr = s_pair[lo];
                return !tx_phi_surely_outside(pr.x - pr.z, pr.y - pr.w, p3.x - pr.x, p3.y - pr.y,
                                              tc.tan_dphi[s_lt[lo] & 0xffu]);
            };
            auto push = [&](bool undecided, uint32_t tag, float2 p3) {
                const uint64_t m = __ballot(undecided);
                if (undecided) {
                    const uint32_t slot = qn + (uint32_t)__popcll(m & lanemask_lt());
                    q[slot] = tag;
                    qxy[slot] = p3;
                }
                qn += (uint32_t)__popcll(m);
                wave_sync();
                if (qn >= kWave) {
                    qn -= kWave;
                    test_queue(qn, true);
                    wave_sync();
                }
            };
            // two groups of 64 candidates at a time: both groups' third hits
            // are loaded before either is tested
            for (uint32_t eb = e0; eb < e1; eb += 2 * kWave) {
                uint32_t tagA, loA, cA, tagB, loB, cB;
                const bool uA = pick(eb + lane, tagA, loA, cA);
                const bool uB = pick(eb + kWave + lane, tagB, loB, cB);
                const float2 p3A = hxy[cA], p3B = hxy[cB];
                const bool vA = prefilter(uA, loA, p3A);
                const bool vB = prefilter(uB, loB, p3B);
                push(vA, tagA, p3A);
                if (eb + kWave < e1) push(vB, tagB, p3B);  // uniform
            }
            test_queue(0, lane < qn);
            if (lane < qn2) test_angles(0);
        }
        wave_sync();
        // 4. the pair's valid triplets: the wave's room in the scratch, then
        //    every pair's triplets in the reference's order
        const Mask vm = s_valid[tid];
        const uint32_t nv = (uint32_t)__builtin_popcountll((uint64_t)vm) + s_pool_valid[tid];
        const uint32_t incl_t = wave_incl_scan(nv);
        const uint32_t total_t = __shfl(incl_t, kWave - 1, kWave);
        if (total_t > range_left) {  // uniform
            const uint32_t size = total_t > kChunkRange ? total_t : kChunkRange;
            uint32_t a = 0;
            if (lane == 0) a = atomicAdd(&cnt->trk_alloc, size);
            range_at = __shfl(a, 0, kWave);
            range_left = size;
        }
        const uint32_t at = range_at;
        range_at += total_t;
        range_left -= total_t;
        if (lane == kWave - 1) {
            chunk_trk[chunk] = total_t;
            chunk_at[chunk] = at;
        }
        if (nv) {
            const uint32_t o = at + incl_t - nv, pv = s_pool_valid[tid];
            auto write = [&](uint64_t k, uint32_t h3) {
                uint32_t r = 0;
                for (Mask m = vm; m; m &= m - 1) r += list_key(tid, (uint32_t)__builtin_ctzll((uint64_t)m)) < k;
                if (pv)
                    for (uint32_t p = 0; p < n_pool; ++p) {
                        const uint2 e = pool[p];
                        r += (e.x & 0xffffu) == tid && (e.x & kValidBit) && pool_key(e) < k;
                    }
                const uint64_t pos = (uint64_t)o + r;
                if (pos < scratch_cap) {
                    th1[pos] = pairs[2 * (size_t)(chunk * kWave + lane)];
                    th2[pos] = pairs[2 * (size_t)(chunk * kWave + lane) + 1];
                    th3[pos] = h3;
                } else {
                    over = true;
                }
            };
            for (Mask m = vm; m; m &= m - 1) {
                const uint32_t i = (uint32_t)__builtin_ctzll((uint64_t)m);
                write(list_key(tid, i), s_first[tid] + s_list[tid * kPredList + i]);
            }
            if (pv)
                for (uint32_t p = 0; p < n_pool; ++p) {
                    const uint2 e = pool[p];
                    if ((e.x & 0xffffu) == tid && (e.x & kValidBit)) write(pool_key(e), e.y);
                }
        }
        if (lane == 0) {  // read into registers above; cleared for the next chunk
            s_pool_n = 0;
            s_dead_any = 0;
        }
        wave_sync();  // the wave's LDS lists and pool are reused by its next chunk
        ++chunk;
        --left;
    }
    if (short_slots) atomicAdd(&cnt->diag_pred_short, short_slots);
    if (trunc) atomicAdd(&cnt->diag_pred_trunc, trunc);
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

}  // namespace

hipError_t launch_pair_plan(const GridGeom& g, const TripletCols& tc, uint32_t n_seg_cap,
                            const uint32_t* grid, uint32_t* seg_base, DevCounts* cnt, hipStream_t st) {
    const uint32_t blocks = (n_seg_cap + 1 + 255) / 256;
    hipLaunchKernelGGL(k_pair_plan, dim3(blocks), dim3(256), 0, st, g, tc, grid, seg_base, cnt);
    return hipGetLastError();
}

hipError_t launch_slot_segments(const uint32_t* seg_base, uint64_t n_slot_cap, uint32_t* slot_seg,
                                const DevCounts* cnt, hipStream_t st) {
    const uint64_t n = slot_seg_entries(n_slot_cap);
    if (n >= 0xffffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_slot_segments, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, seg_base, (uint32_t)n,
                       slot_seg, cnt);
    return hipGetLastError();
}

hipError_t launch_pair_count(const GridGeom& g, const TripletCols& tc, uint32_t n_seg_cap,
                             uint32_t n_slot_cap, const uint32_t* grid, const float* hx, const float* hy,
                             const float* hz, const float2* hzp, const uint32_t* seg_base, uint32_t* slot_counts,
                             DevCounts* cnt, const LaunchCfg& lc, hipStream_t st, const uint32_t* slot_seg) {
    (void)n_seg_cap;
    const uint32_t b = lc.pair_block ? lc.pair_block : 256;
    hipLaunchKernelGGL(k_pair_count, dim3(stride_blocks((uint64_t)n_slot_cap + 1, b, lc)), dim3(b), 0, st, g,
                       tc, grid, hx, hy, hz, hzp, seg_base, slot_counts, cnt, slot_seg, lc.pair_xcd ? lc.pair_xcd : 8u);
    return hipGetLastError();
}

hipError_t launch_pair_store(const GridGeom& g, const TripletCols& tc, uint32_t n_seg_cap,
                             uint32_t n_slot_cap, uint32_t n_pair_cap, const uint32_t* grid,
                             const float* hx, const float* hy, const float* hz, const float2* hzp,
                             const uint32_t* seg_base, const uint32_t* slot_counts, uint32_t* pairs,
                             uint32_t* pair_offsets, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st,
                             const uint32_t* slot_seg) {
    const uint32_t b = lc.pair_block ? lc.pair_block : 256;
    if (reinterpret_cast<uintptr_t>(pairs) & 7u) return hipErrorInvalidValue;  // written as 8-byte pairs
    hipLaunchKernelGGL(k_pair_store, dim3(stride_blocks(n_slot_cap, b, lc)), dim3(b), 0, st, g, tc, grid, hx,
                       hy, hz, hzp, seg_base, slot_counts, pairs, n_pair_cap, cnt, slot_seg,
                       lc.pair_xcd ? lc.pair_xcd : 8u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_pair_offsets, dim3((n_seg_cap + 1 + 255) / 256), dim3(256), 0, st, seg_base,
                       slot_counts, pair_offsets, cnt);
    return hipGetLastError();
}

hipError_t launch_predict_count(const GridGeom& g, const TripletCols& tc, uint32_t n_pair_cap,
                                const uint32_t* grid, const float* hx, const float* hy, const float* hz,
                                const float2* hzp, const uint32_t* hlayer, const uint32_t* hevt,
                                const uint32_t* pairs, uint32_t* counts, DevCounts* cnt, const LaunchCfg& lc,
                                hipStream_t st) {
    const uint32_t b = lc.predict_block ? lc.predict_block : 256;
    hipLaunchKernelGGL(k_predict_count, dim3(stride_blocks((uint64_t)n_pair_cap + 1, b, lc)), dim3(b), 0, st,
                       g, tc, grid, hx, hy, hz, hzp, hlayer, hevt, pairs, counts, cnt);
    return hipGetLastError();
}

hipError_t launch_predict_store(const GridGeom& g, const TripletCols& tc, uint32_t n_pair_cap,
                                uint32_t n_cand_cap, const uint32_t* grid, const float* hx, const float* hy,
                                const float* hz, const float2* hzp, const uint32_t* hlayer, const uint32_t* hevt,
                                const uint32_t* pairs, const uint32_t* counts, uint32_t* cands,
                                uint32_t* cand_offsets, DevCounts* cnt, const LaunchCfg& lc, hipStream_t st) {
    const uint32_t b = lc.predict_block ? lc.predict_block : 256;
    hipLaunchKernelGGL(k_predict_store, dim3(stride_blocks(n_pair_cap, b, lc)), dim3(b), 0, st, g, tc, grid,
                       hx, hy, hz, hzp, hlayer, hevt, pairs, counts, cands, cand_offsets, n_cand_cap, cnt);
    return hipGetLastError();
}

uint32_t predict_fused_tiles(uint64_t n_pair_cap) { return (uint32_t)(n_pair_cap / 128) + 2; }  // every variant
uint64_t predict_fused_chunks(uint64_t n_pair_cap) { return n_pair_cap / kWave + 8; }  // 64-pair chunks, every variant
uint64_t fused_scratch_cap(uint64_t n_trk_cap) {
    // A wave that needs more than its range holds drops the rest (less than
    // that chunk's own tracklets) and takes a new one, so the dropped parts
    // add up to less than all tracklets; at the end every resident wave
    // (2048 blocks of at most 8 waves) may leave up to a range unused.
    return 2 * n_trk_cap + (uint64_t)kChunkRange * kFusedPoolBlocks * (512 / kWave);
}

namespace {

// blocks of a variant that stay resident on the device (persistent launch);
// first called outside any graph capture (predict_fused_prepare)
template <int B, int L, bool S>
uint32_t resident_blocks() {
    static uint32_t resident = 0;
    if (resident == 0) {
        int per_cu = 0, dev = 0, n_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_predict_fused<B, L, S>, B, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu < 1)
            n_cu = 256;
        resident = (uint32_t)(per_cu * n_cu);
    }
    return resident;
}

template <int B, int L, bool S>
hipError_t predict_fused_variant(const GridGeom& g, const TripletCols& tc, uint32_t n_pair_cap, uint32_t n_cand_cap,
                                 const uint32_t* grid, const float* hx, const float* hy, const float* hz,
                                 const float2* hzp, const uint32_t* hlayer, const uint32_t* hevt,
                                 const uint32_t* pairs, uint32_t* cands, uint32_t* cand_offsets, DevCounts* cnt,
                                 uint64_t* status, uint32_t* counter, const FineGeom& fg, const FineBufs& fb,
                                 hipStream_t st) {
    // persistent: as many blocks as stay resident, each loops over tiles
    const uint32_t resident = resident_blocks<B, L, S>();
    const uint64_t tiles = (uint64_t)n_pair_cap / B + 1;
    const uint32_t blocks = (uint32_t)(tiles < resident ? tiles : resident);
    hipLaunchKernelGGL((k_predict_fused<B, L, S>), dim3(blocks), dim3(B), 0, st, g, tc, grid, hx, hy, hz, hzp,
                       hlayer, hevt, pairs, cands, cand_offsets, n_cand_cap, cnt, status, counter, fg, fb.grid,
                       fb.hits);
    return hipGetLastError();
}

// candidate group ends written chunk-local get their chunk's offset (chunk_cands scanned)
__global__ void k_group_offsets(uint32_t* __restrict__ cand_offsets, const uint32_t* __restrict__ group_chunk,
                                const uint32_t* __restrict__ chunk_cands, uint32_t n_trip,
                                DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t n = cnt->n_events * n_trip + 1, n_chunks = cnt->n_pred_chunks;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = group_chunk[i];
        if (c != kSentinel && index_ok(c, n_chunks, cnt)) cand_offsets[i] += chunk_cands[c];
    }
}

template <int B, int L>
uint32_t resident_blocks_filter() {
    static uint32_t resident = 0;
    if (resident == 0) {
        int per_cu = 0, dev = 0, n_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_predict_filter<B, L>, B, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu < 1)
            n_cu = 256;
        resident = (uint32_t)(per_cu * n_cu);
    }
    return resident;
}

template <int B, int L>
uint32_t resident_blocks_filter_wave() {
    static uint32_t resident = 0;
    if (resident == 0) {
        int per_cu = 0, dev = 0, n_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_predict_filter_wave<B, L>, B, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu < 1)
            n_cu = 256;
        resident = (uint32_t)(per_cu * n_cu);
    }
    return resident;
}

template <int B, int L>
hipError_t predict_filter_variant(const GridGeom& g, const TripletCols& tc, uint32_t n_pair_cap, uint32_t n_trk_cap,
                                  const uint32_t* grid, const float* hx, const float* hy, const float* hz,
                                  const float2* hzp, const uint32_t* hlayer, const uint32_t* hevt,
                                  const uint32_t* pairs, uint32_t* cand_offsets, uint32_t* th1, uint32_t* th2,
                                  uint32_t* th3, DevCounts* cnt, uint32_t* counter, const FineGeom& fg,
                                  const FineBufs& fb, uint32_t per_wave, hipStream_t st) {
    uint32_t resident = per_wave ? resident_blocks_filter_wave<B, L>() : resident_blocks_filter<B, L>();
    resident = resident < fb.pool_blocks ? resident : fb.pool_blocks;  // one pool per block
    const uint64_t tiles = (uint64_t)n_pair_cap / B + 1;
    const uint32_t blocks = (uint32_t)(tiles < resident ? tiles : resident);
    if (per_wave)
        hipLaunchKernelGGL((k_predict_filter_wave<B, L>), dim3(blocks), dim3(B), 0, st, g, tc, grid, hx, hy, hz, hzp,
                           hlayer, hevt, pairs, cand_offsets, th1, th2, th3, n_trk_cap, cnt, counter, fg, fb.grid,
                           fb.hits, fb.xy, fb.pool, fb.pool_cap, fb.chunk_cands, fb.chunk_trk, fb.chunk_at,
                           fb.group_chunk, per_wave);
    else
        hipLaunchKernelGGL((k_predict_filter<B, L>), dim3(blocks), dim3(B), 0, st, g, tc, grid, hx, hy, hz, hzp,
                           hlayer, hevt, pairs, cand_offsets, th1, th2, th3, n_trk_cap, cnt, counter, fg, fb.grid,
                           fb.hits, fb.xy, fb.pool, fb.pool_cap, fb.chunk_cands, fb.chunk_trk, fb.chunk_at,
                           fb.group_chunk);
    return hipGetLastError();
}

// the variants built: (block, list entries, sorted walks)
#define TRAX_PRED_VARIANTS(X)                                                                                 \
    X(128, 32, false) X(128, 64, false) X(256, 24, false) X(256, 32, false) X(256, 48, false) X(256, 64, false) \
    X(512, 24, false) X(512, 32, false) X(256, 32, true)

}  // namespace

bool predict_variant_supported(uint32_t block, uint32_t list, uint32_t sort) {
    block = block ? block : 256;
    list = list ? list : 24;
#define TRAX_PRED_IS(B, L, S) if (block == B && list == L && (sort != 0) == S) return true;
    TRAX_PRED_VARIANTS(TRAX_PRED_IS)
#undef TRAX_PRED_IS
    return false;
}

void predict_fused_prepare(const LaunchCfg& lc) {
    const uint32_t b = lc.predict_block ? lc.predict_block : 256, l = lc.predict_list ? lc.predict_list : 24;
    const bool s = lc.predict_sort != 0;
#define TRAX_PRED_PREP(B, L, S)                              \
    if (b == B && l == L && s == S) {                        \
        (void)resident_blocks<B, L, S>();                    \
        if (!S) (void)resident_blocks_filter<B, L>();        \
        if (!S) (void)resident_blocks_filter_wave<B, L>();   \
    }
    TRAX_PRED_VARIANTS(TRAX_PRED_PREP)
#undef TRAX_PRED_PREP
}

FineGeom fine_geom(const GridGeom& g, uint32_t F) {
    FineGeom fg;
    fg.F = F;
    fg.shift = 0;
    while ((1u << fg.shift) < F) ++fg.shift;
    fg.inv_sub = (float)F / g.sector_z;
    fg.cells = g.n_z * F * (g.n_phi + 1) + 1;
    return fg;
}

uint32_t fine_max(const GridGeom& g) {
    uint32_t F = kFineMax;
    while (F > 1 && (uint64_t)F * (g.n_phi + 1) > 8192) F >>= 1;  // the build's LDS counts: at most 32 KiB
    return F;
}

namespace {

// the fine walk structure of every (event, layer), for the sub-row count of the launch shape
// the fine walk structure's geometry for the launch's sub-row count, and the
// structure itself unless the caller built it already (fb.prebuilt)
hipError_t fine_grid_for_predict(const GridGeom& g, uint32_t n_el_cap, uint32_t n_hit_cap, const uint32_t* grid,
                                 const float2* hzp, DevCounts* cnt, const FineBufs& fb, const LaunchCfg& lc,
                                 FineGeom& fg, hipStream_t st) {
    if (!fb.prebuilt) return launch_fine_grid(g, n_el_cap, n_hit_cap, grid, hzp, cnt, fb, lc, fg, st);
    const uint32_t F = lc.predict_fine ? lc.predict_fine : 8;
    if (F > fine_max(g) || (F & (F - 1)) != 0 || !fb.grid || !fb.hits) return hipErrorInvalidValue;
    fg = fine_geom(g, F);
    return hipSuccess;
}

}  // namespace

hipError_t launch_fine_grid(const GridGeom& g, uint32_t n_el_cap, uint32_t n_hit_cap, const uint32_t* grid,
                            const float2* hzp, DevCounts* cnt, const FineBufs& fb, const LaunchCfg& lc, FineGeom& fg,
                            hipStream_t st) {
    const uint32_t F = lc.predict_fine ? lc.predict_fine : 8;
    if (F > fine_max(g) || (F & (F - 1)) != 0 || !fb.grid || !fb.hits) return hipErrorInvalidValue;
    fg = fine_geom(g, F);
    // four 128-thread blocks per CU, each taking items in turn: the build runs
    // beside the pair phases, which keep most of the device (a block per item
    // of 256 threads, half of them idle over the 129 cells, filled the device
    // and slowed the pair count by more than the build itself takes)
    const uint64_t items = (uint64_t)n_el_cap * g.n_z;
    constexpr uint32_t kFineBlock = 128;
    LaunchCfg per_cu = lc;
    per_cu.blocks_per_cu = 4;
    const uint32_t blocks = stride_blocks(items * kFineBlock, kFineBlock, per_cu);
    const size_t lds = ((size_t)F * (g.n_phi + 1) + kFineBlock / kWave + 1) * sizeof(uint32_t);
    hipLaunchKernelGGL(k_fine_grid, dim3(blocks), dim3(kFineBlock), lds, st, g, fg, grid, hzp, n_hit_cap, cnt,
                       fb.grid, fb.hits);
    return hipGetLastError();
}

// chunks per claim of the wave-scheduled pass (0: the block's tiles)
uint32_t predict_wave_claim(const LaunchCfg& lc) {
    if (lc.predict_wave == kPredTiles) return 0;
    return lc.predict_wave ? lc.predict_wave : 2u;
}

hipError_t launch_predict_filter(const GridGeom& g, const TripletCols& tc, uint32_t n_el_cap, uint32_t n_hit_cap,
                                 uint32_t n_pair_cap, uint32_t n_trk_cap, const uint32_t* grid, const float* hx,
                                 const float* hy, const float* hz, const float2* hzp, const uint32_t* hlayer,
                                 const uint32_t* hevt, const uint32_t* pairs, uint32_t* cand_offsets, uint32_t* th1,
                                 uint32_t* th2, uint32_t* th3, DevCounts* cnt, const FineBufs& fb,
                                 const ScanScratch& ss, const LaunchCfg& lc, hipStream_t st) {
    const uint64_t chunks = predict_fused_chunks(n_pair_cap);
    if (!fb.pool || !fb.pool_cap || !fb.pool_blocks || !fb.chunk_cands || !fb.chunk_trk || !fb.chunk_at ||
        !fb.group_chunk || !fb.scratch || !fb.xy || fb.scratch_cap < fused_scratch_cap(n_trk_cap))
        return hipErrorInvalidValue;
    FineGeom fg;
    hipError_t e = fine_grid_for_predict(g, n_el_cap, n_hit_cap, grid, hzp, cnt, fb, lc, fg, st);
    if (e != hipSuccess) return e;
    // chunk counts (the entry after the last chunk stays 0 for the scans), group chunks, the tile counter
    for (uint32_t* a : {fb.chunk_cands, fb.chunk_trk})
        if ((e = launch_fill(a, 0u, chunks + 1, st)) != hipSuccess) return e;
    if ((e = launch_fill(fb.group_chunk, kSentinel, (uint64_t)n_el_cap / g.n_layers * tc.n_trip + 1, st)) !=
        hipSuccess)
        return e;
    // the tile counter, or the wave chunks' region counters
    uint32_t* counter = reinterpret_cast<uint32_t*>(ss.status);
    if ((uint64_t)ss.max_tiles + 1 < kPredCounterWords / 2) return hipErrorInvalidValue;
    if ((e = launch_fill(counter, 0u, kPredCounterWords, st)) != hipSuccess) return e;
    const uint32_t b = lc.predict_block ? lc.predict_block : 256, l = lc.predict_list ? lc.predict_list : 24;
    if (lc.predict_sort) return hipErrorInvalidConfiguration;
    e = hipErrorInvalidConfiguration;
#define TRAX_PF_RUN(B, L, S)                                                                                      \
    if (!S && b == B && l == L)                                                                                   \
        e = predict_filter_variant<B, L>(g, tc, n_pair_cap, fb.scratch_cap, grid, hx, hy, hz, hzp, hlayer, hevt,    \
                                         pairs, cand_offsets, fb.scratch, fb.scratch + fb.scratch_cap,               \
                                         fb.scratch + 2 * (size_t)fb.scratch_cap, cnt, counter, fg, fb,        \
                                         predict_wave_claim(lc), st);
    TRAX_PRED_VARIANTS(TRAX_PF_RUN)
#undef TRAX_PF_RUN
    return e;
}

hipError_t launch_predict_filter_place(const TripletCols& tc, uint32_t n_pair_cap, uint32_t n_trk_cap,
                                       uint32_t* cand_offsets, uint32_t* th1, uint32_t* th2, uint32_t* th3,
                                       DevCounts* cnt, const FineBufs& fb, const ScanScratch& ss, const LaunchCfg& lc,
                                       hipStream_t st) {
    const uint64_t chunks = predict_fused_chunks(n_pair_cap);
    if (!fb.chunk_cands || !fb.chunk_trk || !fb.chunk_at || !fb.group_chunk || !fb.scratch ||
        fb.scratch_cap < fused_scratch_cap(n_trk_cap))
        return hipErrorInvalidValue;
    hipError_t e;
    // the chunks' counts in pair order: candidate and tracklet offsets per chunk, and the totals
    if ((e = launch_scan(fb.chunk_cands, &cnt->n_pred_chunks, 1u, chunks + 1, &cnt->n_cands, ss, st)) != hipSuccess)
        return e;
    if ((e = launch_scan(fb.chunk_trk, &cnt->n_pred_chunks, 1u, chunks + 1, &cnt->n_tracklets, ss, st)) != hipSuccess)
        return e;
    if ((e = launch_chunk_copy3(&cnt->n_pred_chunks, chunks, fb.chunk_trk, fb.chunk_at, fb.scratch, fb.scratch_cap,
                                th1, th2, th3, n_trk_cap, cnt, lc, st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_group_offsets, dim3(16), dim3(256), 0, st, cand_offsets, fb.group_chunk, fb.chunk_cands,
                       tc.n_trip, cnt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_forward_fill(cand_offsets, cnt, tc.n_trip, st);
}

hipError_t launch_predict_fused(const GridGeom& g, const TripletCols& tc, uint32_t n_el_cap, uint32_t n_hit_cap,
                                uint32_t n_pair_cap, uint32_t n_cand_cap, const uint32_t* grid, const float* hx,
                                const float* hy, const float* hz, const float2* hzp, const uint32_t* hlayer,
                                const uint32_t* hevt, const uint32_t* pairs, uint32_t* cands, uint32_t* cand_offsets,
                                DevCounts* cnt, const FineBufs& fb, const ScanScratch& ss, const LaunchCfg& lc,
                                hipStream_t st) {
    const uint32_t tiles = predict_fused_tiles(n_pair_cap);
    if (tiles > ss.max_tiles) return hipErrorInvalidValue;
    FineGeom fg;
    hipError_t e = fine_grid_for_predict(g, n_el_cap, n_hit_cap, grid, hzp, cnt, fb, lc, fg, st);
    if (e != hipSuccess) return e;
    e = launch_fill(reinterpret_cast<uint32_t*>(ss.status), 0u, 2ull * tiles + 2, st);
    if (e != hipSuccess) return e;
    uint64_t* status = ss.status;
    uint32_t* counter = reinterpret_cast<uint32_t*>(ss.status + tiles);
    const uint32_t b = lc.predict_block ? lc.predict_block : 256, l = lc.predict_list ? lc.predict_list : 24;
    const bool s = lc.predict_sort != 0;
    e = hipErrorInvalidConfiguration;
#define TRAX_PRED_RUN(B, L, S)                                                                                   \
    if (b == B && l == L && s == S)                                                                              \
        e = predict_fused_variant<B, L, S>(g, tc, n_pair_cap, n_cand_cap, grid, hx, hy, hz, hzp, hlayer, hevt,    \
                                           pairs, cands, cand_offsets, cnt, status, counter, fg, fb, st);
    TRAX_PRED_VARIANTS(TRAX_PRED_RUN)
#undef TRAX_PRED_RUN
    if (e != hipSuccess) return e;
    return launch_forward_fill(cand_offsets, cnt, tc.n_trip, st);
}

}  // namespace trax
