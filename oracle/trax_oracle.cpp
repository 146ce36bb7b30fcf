// ORACLE — test infrastructure only.  CPU restatement of the OpenCL kernels
// of trax's triplet-seeding path, used to check the HIP implementation.  Only
// tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it; the
// product never does.
//
// Each stage follows the reference kernel it names (file:line relative to the
// reference's src/algorithms or src/datastructures).  Where the reference
// leaves an outcome to the OpenCL runtime, the oracle fixes one legal outcome
// and the HIP path produces the same one:
//   * the order of hits inside a grid cell (atomic_inc in gridStore,
//     GridBuilder.h:243-244) is the input order (stable scatter);
//   * CA basis lists (atomic_inc in followerBasisStore, CellularAutomaton.h:201)
//     are in connection-pair order, i.e. ascending basis index;
//   * the racy offset "monotonize" kernels (TripletThetaPhiPredictor.h:442-445,
//     TripletThetaPhiFilter.h:275-278) are applied as the sequential forward fill;
//   * slots the reference leaves uninitialised (a store kernel finding fewer
//     entries than its count kernel) hold 0xffffffff; such sentinel pairs /
//     candidates yield nothing downstream.
// The transcendental built-ins come from trax_amd/csrc/trax_math.h, the same
// implementation the device uses, so that comparisons isolate the pipeline
// logic; those built-ins are pinned separately against float64.
//
// Parity status of each function: see the header comment of each stage.

#include "../trax_amd/csrc/trax_math.h"

// Mutation hooks: a build with -DTRAX_ORACLE_PLANT=k plants mistake k in the
// linking half, so that the tests can show that each of its independent pins
// fails on it (tests/test_oracle_pins.py builds every variant).  The product
// library has no such hooks; the default build plants nothing.
//   1: the eta cut taken as <= instead of < (TripletConnectivityTight.h:137)
//   2: eta from the first to the second hit instead of the third (.h:60-75)
//   3: the circle plane offset c without the paraboloid lift (TrackletCircleFitter.h:80)
//   4: the best basis by the largest pT difference (CellularAutomaton.h:254-259)
#ifndef TRAX_ORACLE_PLANT
#define TRAX_ORACLE_PLANT 0
#endif

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

namespace {

using u32 = uint32_t;
constexpr u32 SENT = 0xffffffffu;

struct Arr {
    std::vector<u32> u;
    std::vector<float> f;
    bool is_float = false;
};

struct Result {
    std::map<std::string, Arr> arrays;
    std::map<std::string, double> timings_ms;
    std::vector<std::string> names;
    std::string err;
    void put(const std::string& k, std::vector<u32> v) {
        Arr a; a.u = std::move(v); a.is_float = false; arrays[k] = std::move(a);
    }
    void putf(const std::string& k, std::vector<float> v) {
        Arr a; a.f = std::move(v); a.is_float = true; arrays[k] = std::move(a);
    }
};

struct Input {
    // HitCollection columns (src/datastructures/HitCollection.h:23)
    const float* x; const float* y; const float* z;
    const u32* layer; const u32* det_id; const u32* hit_id; const u32* event;
    u32 n_hits;
    // EventSupplement / LayerSupplement (EventSupplement.h:14, LayerSupplement.h:13)
    u32 n_events, n_layers;
    const u32* evt_nhits; const u32* evt_offset;
    const u32* lay_nhits; const u32* lay_offset;
    // GridConfig (Grid.h:5-39)
    u32 n_z, n_phi;
    // GeometrySupplement (GeometrySupplement.h:21): radius per layer index
    const float* min_r; const float* max_r; u32 n_geom;
    // TripletConfigurations (TripletConfiguration.h:33)
    u32 n_trip;
    const u32* l1; const u32* l2; const u32* l3;
    const float* dtheta; const float* sigma_z; const float* dphi; const float* sigma_phi;
    const float* tip; const float* z0; const u32* pair_spread_phi;
    float min_radius_curvature;
    u32 n_threads;
    float d_eta_cut;
    int sparse_pairs;   // 0: literal dense pair bitmap, 1: per-outer-hit sparse bit lists
    int stop_after;     // 0 all stages, else the last stage: 1 grid, 2 pairs, 3 predict, 4 filter, 5 CA
    int n_workers;      // host threads for the NDRange emulation
};

// ---------------------------------------------------------------------------
// NDRange helper: run fn(g) for g in [0, n) over a thread pool; work-groups are
// independent in every kernel this is used for.
template <class F>
void par_for(u32 n, int workers, F fn) {
    if (workers <= 1 || n < 2) {
        for (u32 g = 0; g < n; ++g) fn(g);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<u32> next{0};
    for (int w = 0; w < workers; ++w)
        th.emplace_back([&] {
            for (u32 g = next.fetch_add(1); g < n; g = next.fetch_add(1)) fn(g);
        });
    for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------
// PrefixSum (PrefixSum.h:58-248, PrefixSum.cpp:8-140): exclusive in-place
// Blelloch scan, 2*wg elements per work-group, bank-padded local memory,
// recursive over partial sums.  Restated with its work-group structure.
// Pinned by: the reference KAT (trax_test.cpp:473-479), closed form, brute force.
u32 bank(u32 i) { return i + (i >> 4); }

void wg_blelloch(u32* input, u32 size, u32 gid, u32 ls, u32* partial) {
    std::vector<u32> data(2 * ls + (2 * ls) / 16 + 2, 0);
    const u32 g_off = gid * (ls << 1);
    for (u32 lid = 0; lid < ls; ++lid) {
        data[bank(lid)] = (g_off + lid) < size ? input[g_off + lid] : 0;
        data[bank(lid + ls)] = (g_off + lid + ls) < size ? input[g_off + lid + ls] : 0;
    }
    u32 offset = 1;
    for (u32 d = ls; d > 0; d >>= 1) {  // up-sweep (:80-96)
        for (u32 lid = 0; lid < d; ++lid) {
            u32 ai = offset * ((lid << 1) + 1) - 1, bi = offset * ((lid << 1) + 2) - 1;
            data[bank(bi)] += data[bank(ai)];
        }
        offset <<= 1;
    }
    const u32 last = bank((ls << 1) - 1);
    if (partial) partial[gid] = data[last];  // (:189)
    data[last] = 0;
    for (u32 d = 1; d <= ls; d <<= 1) {  // down-sweep (:110-127)
        offset >>= 1;
        for (u32 lid = 0; lid < d; ++lid) {
            u32 ai = offset * ((lid << 1) + 1) - 1, bi = offset * ((lid << 1) + 2) - 1;
            u32 t = data[bank(ai)];
            data[bank(ai)] = data[bank(bi)];
            data[bank(bi)] += t;
        }
    }
    for (u32 lid = 0; lid < ls; ++lid) {
        if (g_off + lid < size) input[g_off + lid] = data[bank(lid)];
        if (g_off + lid + ls < size) input[g_off + lid + ls] = data[bank(lid + ls)];
    }
}

u32 next_pow2(u32 n) {
    n--; n |= n >> 1; n |= n >> 2; n |= n >> 4; n |= n >> 8; n |= n >> 16; return n + 1;
}

u32 ngroups_f(u32 size, u32 wg) {  // (uint) max(1.0f, ceil((float)size / (wg << 1)))
    float v = ceilf((float)size / (float)(wg << 1));
    return (u32)std::max(1.0f, v);
}

void recursive_scan(u32* input, u32 size, u32 wg) {
    const u32 n_groups = ngroups_f(size, wg);
    if (n_groups == 1) {  // base case (:84-98)
        u32 local = (u32)std::max(1.0f, ceilf((float)size / 2.0f));
        local = next_pow2(local);
        wg_blelloch(input, size, 0, local, nullptr);
        return;
    }
    std::vector<u32> partial(n_groups, 0);
    for (u32 g = 0; g < n_groups; ++g) wg_blelloch(input, size, g, wg, partial.data());
    recursive_scan(partial.data(), n_groups, wg);
    for (u32 g = 0; g < n_groups; ++g) {  // prefixSumUniformAdd (:223-248)
        const u32 g_off = g * (wg << 1);
        for (u32 lid = 0; lid < wg; ++lid) {
            if (g_off + lid < size) input[g_off + lid] += partial[g];
            if (g_off + lid + wg < size) input[g_off + lid + wg] += partial[g];
        }
    }
}

void prefix_sum(std::vector<u32>& v, u32 wg) {
    if (!v.empty()) recursive_scan(v.data(), (u32)v.size(), wg);
}

// sequential reading of the offset "monotonize" kernels
void forward_fill(std::vector<u32>& off) {
    for (size_t g = 1; g < off.size(); ++g)
        if (off[g] == 0) off[g] = off[g - 1];
}

// ---------------------------------------------------------------------------
struct Ctx {
    const Input& in;
    Result& res;
    u32 C, T, L, E;
    float min_z, sz, min_phi, sphi;
    // sorted hit columns
    std::vector<float> hx, hy, hz;
    std::vector<u32> hlayer, hdet, hid, hevt;
    std::vector<u32> grid;
    std::vector<u32> pairs;  // interleaved (x, y)
    std::vector<u32> pair_off;
    std::vector<u32> cands;  // interleaved (pairIdx, hit3)
    std::vector<u32> cand_off;
    std::vector<u32> th1, th2, th3, t_off;
    u32 diag_pair_short = 0, diag_pred_short = 0, diag_pred_trunc = 0, diag_wrap_overlap = 0;

    explicit Ctx(const Input& i, Result& r) : in(i), res(r) {
        C = (in.n_z + 1) * (in.n_phi + 1);
        T = in.n_trip; L = in.n_layers; E = in.n_events;
        min_z = -300.0f;
        sz = (300.0f - -300.0f) / (float)in.n_z;                 // Grid.h:11-13
        min_phi = -(float)M_PI;
        sphi = ((float)M_PI - -(float)M_PI) / (float)in.n_phi;  // Grid.h:15-17
    }
};

u32 hit_cell(const Ctx& c, float x, float y, float z) {
    u32 zs = tx_f2u_sat(floorf((z - c.min_z) / c.sz));
    float phi = tx_atan2f(y, x);
    u32 ps = tx_f2u_sat(floorf((phi - c.min_phi) / c.sphi));
    if (zs > c.in.n_z || ps > c.in.n_phi) return SENT;  // outside the grid: reference UB
    return zs * (c.in.n_phi + 1) + ps;
}

// GridBuilder::run (GridBuilder.cpp:6-161): gridCount (GridBuilder.h:48-115),
// PrefixSum over the whole grid, gridStore (:172-257) with stable in-cell order.
// Pinned by: verifyGrid invariants (GridBuilder.cpp:188-225), brute-force histogram,
// permutation check.
bool grid_build(Ctx& c) {
    const Input& in = c.in;
    c.grid.assign((size_t)c.E * c.L * c.C, 0);
    for (u32 e = 0; e < c.E; ++e)
        for (u32 l = 0; l < c.L; ++l) {
            const u32 hits = in.lay_nhits[e * c.L + l];
            const u32 off = in.evt_offset[e] + in.lay_offset[e * c.L + l];
            u32* g = &c.grid[((size_t)e * c.L + l) * c.C];
            for (u32 i = 0; i < hits; ++i) {
                u32 cell = hit_cell(c, in.x[off + i], in.y[off + i], in.z[off + i]);
                if (cell >= c.C) { c.res.err = "hit outside the grid"; return false; }
                g[cell]++;
            }
        }
    prefix_sum(c.grid, in.n_threads);
    const u32 n = in.n_hits;
    c.hx.assign(n, 0); c.hy.assign(n, 0); c.hz.assign(n, 0);
    c.hlayer.assign(n, 0); c.hdet.assign(n, 0); c.hid.assign(n, 0); c.hevt.assign(n, 0);
    std::vector<u32> written(c.C);
    for (u32 e = 0; e < c.E; ++e)
        for (u32 l = 0; l < c.L; ++l) {
            std::fill(written.begin(), written.end(), 0);
            const u32* g = &c.grid[((size_t)e * c.L + l) * c.C];
            const u32 off = in.evt_offset[e] + in.lay_offset[e * c.L + l];
            const u32 end = off + in.lay_nhits[e * c.L + l];
            for (u32 i = off; i < end; ++i) {
                u32 cell = hit_cell(c, in.x[i], in.y[i], in.z[i]);
                u32 w = g[cell] + written[cell]++;
                if (w >= n) { c.res.err = "grid scatter out of range"; return false; }
                c.hx[w] = in.x[i]; c.hy[w] = in.y[i]; c.hz[w] = in.z[i];
                c.hlayer[w] = in.layer[i]; c.hdet[w] = in.det_id[i];
                c.hid[w] = in.hit_id[i]; c.hevt[w] = in.event[i];
            }
        }
    return true;
}

// ---------------------------------------------------------------------------
// PairGeneratorBeamspot (PairGeneratorBeamspot.cpp:5-183, .h:43-246 pairCount,
// .h:448-525 pairStore).  Restated with the work-item structure that fixes the
// output order: work-group (layerPair, event), work-item t strides over the
// outer-layer hits, the inner hits come out in ascending order.
// Pinned by: brute force over all (inner, outer) combinations with the same
// window predicate on tiny events; sparse and dense bitmap modes agree.
struct PairWindow {
    float zLow, zHigh, phiLow, phiHigh;
    bool wrap;
    u32 zLowSector, zHighSector, phiLowSector, phiHighSector;
};

PairWindow pair_window(const Ctx& c, u32 p, u32 lay1, u32 i) {
    const Input& in = c.in;
    PairWindow w;
    const float z0 = in.z0[p];
    const float tip = in.tip[p];
    const float sigmaZ = in.sigma_z[p];
    const float sigmaPhi = tx_bits2f(in.pair_spread_phi[p]);  // .cpp:59 passes pairSpreadPhi
    const float minR1 = in.min_r[lay1], maxR1 = in.max_r[lay1];
    const float mrc = in.min_radius_curvature;
    const float X = c.hx[i], Y = c.hy[i], Z = c.hz[i];
    // OpenCL sign(): 1, -1, or the argument itself for +-0, 0 for NaN
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
    const int zl = tx_f2i_sat(floorf((zLow - c.min_z) / c.sz));
    w.zLowSector = (u32)(zl > 0 ? zl : 0);
    const u32 zh = tx_f2u_sat(floorf((zHigh - c.min_z) / c.sz)) + 1u;
    w.zHighSector = zh < in.n_z ? zh : in.n_z;
    const float phi = tx_atan2f(Y, X);
    tmp = fabsf(tx_acosf(r / (2 * mrc)) - tx_acosf(maxR1 / (2 * mrc)));
    float dPhi = fabsf(tx_acosf(r / (2 * mrc)) - tx_acosf(minR1 / (2 * mrc)));
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
    w.phiLowSector = tx_f2u_sat(floorf((phiLow - c.min_phi) / c.sphi));
    w.phiHighSector = tx_f2u_sat(floorf((phiHigh - c.min_phi) / c.sphi) + 1.0f);
    w.zLow = zLow; w.zHigh = zHigh; w.phiLow = phiLow; w.phiHigh = phiHigh;
    return w;
}

// walk the inner-layer window of outer hit i; visit(index) for every valid hit,
// in the reference's loop order (z sector, then j with wrap-around)
template <class V>
void pair_walk(const Ctx& c, const u32* lGrid1, const PairWindow& w, V visit) {
    const u32 np1 = c.in.n_phi + 1;
    const float PI = TX_PI_F;
    for (u32 zs = w.zLowSector; zs < w.zHighSector; ++zs) {
        const u32 zStart = lGrid1[zs * np1];
        const u32 zEnd = lGrid1[(zs + 1) * np1 - 1];
        const u32 zLen = zEnd - zStart;
        u32 j = lGrid1[zs * np1 + w.phiLowSector];
        const u32 end2 = (u32)w.wrap * zEnd + lGrid1[zs * np1 + w.phiHighSector] - (u32)w.wrap * zStart;
        for (; j < end2; ++j) {
            const u32 index = j - (u32)(j >= zEnd) * zLen;
            bool valid = w.zLow <= c.hz[index] && c.hz[index] <= w.zHigh;
            const float actPhi = tx_atan2f(c.hy[index], c.hx[index]);
            valid = valid && ((!w.wrap && (w.phiLow <= actPhi && actPhi <= w.phiHigh)) ||
                              (w.wrap && ((w.phiLow <= actPhi && actPhi <= PI) ||
                                          (-PI <= actPhi && actPhi <= w.phiHigh))));
            if (valid) visit(index);
        }
    }
}

bool pair_gen(Ctx& c) {
    const Input& in = c.in;
    const u32 nt = in.n_threads, T = c.T, E = c.E, C = c.C;
    // oracle offsets (.cpp:14-28): 32*ceil(N1*N2/32) bits per (event, layerPair)
    std::vector<uint64_t> o_off(E * T);
    uint64_t total_bits = 0;
    for (u32 e = 0; e < E; ++e)
        for (u32 p = 0; p < T; ++p) {
            const u32* g1 = &c.grid[((size_t)e * c.L + in.l1[p] - 1) * C];
            const u32* g2 = &c.grid[((size_t)e * c.L + in.l2[p] - 1) * C];
            uint64_t n1 = g1[C - 1] - g1[0], n2 = g2[C - 1] - g2[0];
            uint64_t nm = (uint64_t)(n1 * n2);
            nm = 32 * (uint64_t)std::ceil(nm / 32.0);
            o_off[e * T + p] = total_bits;
            total_bits += nm;
        }
    const bool sparse = in.sparse_pairs || total_bits > (uint64_t)8 << 33;
    std::vector<u32> bitmap;
    if (!sparse) bitmap.assign((size_t)((total_bits + 31) / 32), 0);
    // sparse mode: per (segment, outer hit) the sorted unique inner indices
    std::vector<std::vector<std::vector<u32>>> lists(E * T);
    std::vector<u32> prefix((size_t)E * T * nt + 1, 0);
    std::vector<u32> wrap_overlap(E * T, 0);

    par_for(E * T, in.n_workers, [&](u32 seg) {
        const u32 e = seg / T, p = seg % T;
        const u32 lay2 = in.l2[p] - 1, lay1 = in.l1[p] - 1;
        const u32* g2 = &c.grid[((size_t)e * c.L + lay2) * C];
        const u32 layer2Offset = g2[0];
        const u32 end = g2[C - 1];
        const u32 nHits2 = end - layer2Offset;
        const u32* lGrid1 = &c.grid[((size_t)e * c.L + lay1) * C];
        const u32 offset = lGrid1[0];
        if (sparse) lists[seg].assign(nHits2, {});
        for (u32 t = 0; t < nt; ++t) {
            u32 nFound = 0;
            for (u32 i = layer2Offset + t; i < end; i += nt) {
                PairWindow w = pair_window(c, p, lay1, i);
                if (w.wrap && w.phiHighSector > w.phiLowSector) wrap_overlap[seg]++;
                pair_walk(c, lGrid1, w, [&](u32 index) {
                    nFound++;
                    const uint64_t bit = (uint64_t)(index - offset) * nHits2 + (i - layer2Offset);
                    if (sparse) {
                        lists[seg][i - layer2Offset].push_back(index - offset);
                    } else {
                        const uint64_t b = o_off[seg] + bit;
                        bitmap[b / 32] |= (1u << (b % 32));
                    }
                });
            }
            prefix[(size_t)seg * nt + t] = nFound;
        }
        if (sparse)
            for (auto& v : lists[seg]) {
                std::sort(v.begin(), v.end());
                v.erase(std::unique(v.begin(), v.end()), v.end());
            }
    });
    prefix_sum(prefix, nt);
    const u32 nPairs = prefix.back();
    c.pairs.assign((size_t)nPairs * 2, SENT);
    c.pair_off.assign(E * T + 1, 0);
    std::vector<u32> written_per_seg(E * T, 0);
    par_for(E * T, in.n_workers, [&](u32 seg) {
        const u32 e = seg / T, p = seg % T;
        const u32 lay2 = in.l2[p] - 1, lay1 = in.l1[p] - 1;
        const u32* g2 = &c.grid[((size_t)e * c.L + lay2) * C];
        const u32 layer2Offset = g2[0], end = g2[C - 1], nHits2 = end - layer2Offset;
        const u32* g1 = &c.grid[((size_t)e * c.L + lay1) * C];
        const u32 nHits1 = g1[C - 1] - g1[0], offset = g1[0];
        u32 wr = 0;
        for (u32 t = 0; t < nt; ++t) {
            u32 pos = prefix[(size_t)seg * nt + t];
            const u32 next = prefix[(size_t)seg * nt + t + 1];
            for (u32 i = layer2Offset + t; i < end; i += nt) {
                if (sparse) {
                    for (u32 j : lists[seg][i - layer2Offset]) {
                        if (!(pos < next)) break;
                        c.pairs[2 * (size_t)pos] = j + offset;
                        c.pairs[2 * (size_t)pos + 1] = i;
                        pos++; wr++;
                    }
                } else {
                    for (u32 j = 0; j < nHits1 && pos < next; ++j) {  // (.h:489-516)
                        const uint64_t index = (uint64_t)j * nHits2 + i - layer2Offset;
                        const uint64_t b = o_off[seg] + index;
                        if (bitmap[b / 32] & (1u << (b % 32))) {
                            c.pairs[2 * (size_t)pos] = j + offset;
                            c.pairs[2 * (size_t)pos + 1] = i;
                            pos++; wr++;
                        }
                    }
                }
            }
            if (t == nt - 1) c.pair_off[seg + 1] = next;  // (.h:519-521)
        }
        written_per_seg[seg] = wr;
    });
    u32 written = 0;
    for (u32 v : written_per_seg) written += v;
    c.diag_pair_short = nPairs - written;
    for (u32 v : wrap_overlap) c.diag_wrap_overlap += v;
    return true;
}

// ---------------------------------------------------------------------------
// TripletThetaPhiPredictor (.cpp:10-170; predictCount .h:49-230, predictStore
// .h:232-430, predictOffsetMonotonizeStore .h:432-447).  One work-item per pair.
// The count kernel uses rOrigin = sqrt(x1^2 + y1^2 - tip2) (.h:130-131), the store
// kernel rOrigin = sqrt(sqrt(x1^2 + y1^2)^2 - tip2) (.h:305-306); the store is
// bounded by the count.  Pinned by: brute force over all third hits with the same
// window predicate on tiny events.
struct PredWindow {
    float zLow, zHigh, phiLow, phiHigh;
    bool wrap;
    u32 zLowSector, zHighSector, phiLowSector, phiHighSector;
    u32 lay3;
};

bool pred_window(const Ctx& c, u32 firstHit, u32 secondHit, bool storeVariant, PredWindow& w, u32& event) {
    const Input& in = c.in;
    event = c.hevt[firstHit];
    const u32 lt = c.hlayer[firstHit] - 1;  // (.h:88)
    if (event >= c.E || lt >= c.T) return false;
    const u32 layer = in.l3[lt] - 1;
    if (layer >= c.L || layer >= in.n_geom) return false;
    w.lay3 = layer;
    const float sigmaZ = in.sigma_z[lt], sigmaPhi = in.sigma_phi[lt];
    const float minR = in.min_r[layer], maxR = in.max_r[layer];
    const float mrc = in.min_radius_curvature;
    const float p1x = c.hx[firstHit], p1y = c.hy[firstHit], p1z = c.hz[firstHit];
    const float p2x = c.hx[secondHit], p2y = c.hy[secondHit], p2z = c.hz[secondHit];
    float tip2 = p1y * (p2x - p1x) - p1x * (p2y - p1y);
    tip2 *= tip2;
    tip2 /= ((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
    float rOrigin;
    if (!storeVariant) {
        rOrigin = p1x * p1x + p1y * p1y;
        rOrigin = sqrtf(rOrigin - tip2);
    } else {
        rOrigin = sqrtf(p1x * p1x + p1y * p1y);
        rOrigin = sqrtf(rOrigin * rOrigin - tip2);
    }
    const float cotTheta = (p2z - p1z) / (sqrtf(p2x * p2x + p2y * p2y - tip2) - rOrigin);
    float tmp = p1z + (sqrtf(maxR * maxR - tip2) - rOrigin) * cotTheta;
    float zLow = p1z + (sqrtf(minR * minR - tip2) - rOrigin) * cotTheta;
    const float zHigh0 = (float)(tmp > zLow) * tmp + (float)(tmp < zLow) * zLow;
    zLow = (float)(tmp > zLow) * zLow + (float)(tmp < zLow) * tmp;
    float zHigh = zHigh0;
    zLow -= sigmaZ;
    zHigh += sigmaZ;
    const int zl = tx_f2i_sat(floorf((zLow - c.min_z) / c.sz));
    w.zLowSector = (u32)(zl > 0 ? zl : 0);
    const u32 zh = tx_f2u_sat(floorf((zHigh - c.min_z) / c.sz)) + 1u;
    w.zHighSector = zh < in.n_z ? zh : in.n_z;
    const float phi = tx_atan2f(p2y, p2x);
    tmp = tx_atan2f(p1y, p1x);
    const float PI = TX_PI_F;
    float dPhi = phi - tmp;
    dPhi += (dPhi > PI) ? -2 * PI : (dPhi < -PI) ? 2 * PI : 0;
    dPhi = fabsf(dPhi);
    const float dHits = sqrtf((p2x - p1x) * (p2x - p1x) + (p2y - p1y) * (p2y - p1y));
    tmp = fabsf(tx_acosf(dHits / (2 * mrc)) - tx_acosf(maxR / (2 * mrc)));
    float phiLow = fabsf(tx_acosf(dHits / (2 * mrc)) - tx_acosf(minR / (2 * mrc)));
    // OpenCL max(x, y): y if x < y, else x
    auto fmx = [](float a, float b) { return (a < b) ? b : a; };
    dPhi = fmx(dPhi, fmx(tmp, phiLow));
    float phiHigh = phi + dPhi + sigmaPhi;
    phiLow = phi - dPhi - sigmaPhi;
    w.wrap = phiLow < -PI || phiHigh > PI || phiLow > PI || phiHigh < -PI;
    phiLow -= (float)((phiLow > PI) * 2) * PI;
    phiLow += (float)((phiLow < -PI) * 2) * PI;
    phiHigh -= (float)((phiHigh > PI) * 2) * PI;
    phiHigh += (float)((phiHigh < -PI) * 2) * PI;
    w.phiLowSector = tx_f2u_sat(floorf((phiLow - c.min_phi) / c.sphi));
    const u32 ph = tx_f2u_sat(floorf((phiHigh - c.min_phi) / c.sphi)) + 1u;
    w.phiHighSector = ph < in.n_phi ? ph : in.n_phi;
    w.zLow = zLow; w.zHigh = zHigh; w.phiLow = phiLow; w.phiHigh = phiHigh;
    return true;
}

template <class V>
void pred_walk(const Ctx& c, const u32* lGrid3, const PredWindow& w, V visit) {
    const u32 np1 = c.in.n_phi + 1;
    const float PI = TX_PI_F;
    for (u32 zs = w.zLowSector; zs < w.zHighSector; ++zs) {
        const u32 zStart = lGrid3[zs * np1];
        const u32 zEnd = lGrid3[(zs + 1) * np1];  // (.h:193)
        const u32 zLen = zEnd - zStart;
        const u32 end2 = (u32)w.wrap * zEnd + lGrid3[zs * np1 + w.phiHighSector] - (u32)w.wrap * zStart;
        for (u32 j = lGrid3[zs * np1 + w.phiLowSector]; j < end2; ++j) {
            const u32 index = j - (u32)(j >= zEnd) * zLen;
            bool valid = w.zLow <= c.hz[index] && c.hz[index] <= w.zHigh;
            const float actPhi = tx_atan2f(c.hy[index], c.hx[index]);
            valid = valid && ((!w.wrap && (w.phiLow <= actPhi && actPhi <= w.phiHigh)) ||
                              (w.wrap && ((w.phiLow <= actPhi && actPhi <= PI) ||
                                          (-PI <= actPhi && actPhi <= w.phiHigh))));
            if (!visit(index, valid)) return;
        }
    }
}

bool predict(Ctx& c) {
    const Input& in = c.in;
    const u32 nPairs = (u32)(c.pairs.size() / 2);
    std::vector<u32> prefix(nPairs + 1, 0);
    par_for((nPairs + 255) / 256, in.n_workers, [&](u32 blk) {
        for (u32 gid = blk * 256; gid < std::min(nPairs, blk * 256 + 256); ++gid) {
            const u32 h1 = c.pairs[2 * (size_t)gid], h2 = c.pairs[2 * (size_t)gid + 1];
            if (h1 == SENT || h2 == SENT) continue;
            PredWindow w; u32 ev;
            if (!pred_window(c, h1, h2, false, w, ev)) continue;
            const u32* lGrid3 = &c.grid[((size_t)ev * c.L + w.lay3) * c.C];
            u32 n = 0;
            pred_walk(c, lGrid3, w, [&](u32, bool v) { n += v; return true; });
            prefix[gid] = n;
        }
    });
    prefix_sum(prefix, in.n_threads);
    const u32 nCand = prefix.back();
    c.cands.assign((size_t)nCand * 2, SENT);
    c.cand_off.assign(c.E * c.T + 1, 0);
    std::vector<u32> short_by(nPairs, 0);
    par_for((nPairs + 255) / 256, in.n_workers, [&](u32 blk) {
        for (u32 gid = blk * 256; gid < std::min(nPairs, blk * 256 + 256); ++gid) {
            const u32 h1 = c.pairs[2 * (size_t)gid], h2 = c.pairs[2 * (size_t)gid + 1];
            u32 pos = prefix[gid];
            const u32 next = prefix[gid + 1];
            if (h1 == SENT || h2 == SENT) continue;
            // a pair whose count is zero writes nothing (.h:376-377); the
            // truncation diagnostic covers the pairs with a nonzero count
            if (pos == next) continue;
            PredWindow w; u32 ev;
            if (!pred_window(c, h1, h2, true, w, ev)) { short_by[gid] = next - pos; continue; }
            const u32* lGrid3 = &c.grid[((size_t)ev * c.L + w.lay3) * c.C];
            bool trunc = false;
            pred_walk(c, lGrid3, w, [&](u32 index, bool v) {
                if (!(pos < next)) { if (v) trunc = true; return !trunc; }
                if (v) {
                    c.cands[2 * (size_t)pos] = gid;
                    c.cands[2 * (size_t)pos + 1] = index;
                }
                pos += v;
                return true;
            });
            short_by[gid] = (next - pos) | (trunc ? 0x80000000u : 0u);
        }
    });
    for (u32 gid = 0; gid < nPairs; ++gid) {
        c.diag_pred_short += short_by[gid] & 0x7fffffffu;
        c.diag_pred_trunc += (short_by[gid] >> 31);
    }
    // tripletOffsets (.h:404-418)
    for (u32 gid = 0; gid < nPairs; ++gid) {
        const u32 h1 = c.pairs[2 * (size_t)gid];
        if (h1 == SENT) continue;
        const u32 ev = c.hevt[h1], lt = c.hlayer[h1] - 1;
        if (ev >= c.E || lt >= c.T) continue;
        bool last = true;
        if (gid < nPairs - 1) {
            const u32 nh = c.pairs[2 * (size_t)(gid + 1)];
            if (nh != SENT) last = (lt != c.hlayer[nh] - 1) || (ev != c.hevt[nh]);
        }
        if (last) c.cand_off[ev * c.T + lt + 1] = prefix[gid + 1];
    }
    forward_fill(c.cand_off);
    return true;
}

// ---------------------------------------------------------------------------
// TripletThetaPhiFilter (.cpp:10-208; filterCount .h:50-162, filterPopCount
// :164-179, filterStore :181-219, filterOffsetStore :221-261, monotonize :267-280).
// Pinned by: FP64 recomputation of each cut on every candidate (cut decisions may
// differ only within a few ulp of a cut); brute force on tiny events.
bool filter_valid(const Ctx& c, u32 h1, u32 h2, u32 h3, u32 lt) {
    const Input& in = c.in;
    const float dThetaCut = in.dtheta[lt], dPhiCut = in.dphi[lt], tipCut = in.tip[lt];
    const float minRadius = in.min_radius_curvature;
    const float x1 = c.hx[h1], y1 = c.hy[h1], z1 = c.hz[h1];
    const float x2 = c.hx[h2], y2 = c.hy[h2], z2 = c.hz[h2];
    const float x3 = c.hx[h3], y3 = c.hy[h3], z3 = c.hz[h3];
    bool valid = true;
    float angle1 = tx_atan2f(sqrtf((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1)), (z2 - z1));
    float angle2 = tx_atan2f(sqrtf((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2)), (z3 - z2));
    float delta = fabsf(angle2 / angle1);
    valid = valid && (1 - dThetaCut <= delta && delta <= 1 + dThetaCut);
    angle1 = tx_atan2f((y2 - y1), (x2 - x1));
    angle2 = tx_atan2f((y3 - y2), (x3 - x2));
    delta = angle2 - angle1;
    const float PI = TX_PI_F;
    delta += (delta > PI) ? -2 * PI : (delta < -PI) ? 2 * PI : 0;
    valid = valid && (fabsf(delta) <= dPhiCut);
    const float p1x = x1, p1y = y1, p1z = x1 * x1 + y1 * y1;
    const float p2x = x2, p2y = y2, p2z = x2 * x2 + y2 * y2;
    const float p3x = x3, p3y = y3, p3z = x3 * x3 + y3 * y3;
    const float ax = p2x - p1x, ay = p2y - p1y, az = p2z - p1z;
    const float bx = p3x - p1x, by = p3y - p1y, bz = p3z - p1z;
    float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
    tx_normalize3(nx, ny, nz);
    const float cox = (-nx) / (2 * nz), coy = (-ny) / (2 * nz);
    const float cc = -(nx * p1x + ny * p1y + nz * p1z);
    const float cR = sqrtf((1 - nz * nz - 4 * cc * nz) / (4 * nz * nz));
    float vx = -cox, vy = -coy;
    tx_normalize2(vx, vy);
    const float pcax = cox + cR * vx, pcay = coy + cR * vy;
    const float tip_squared = pcax * pcax + pcay * pcay;
    valid = valid && (tip_squared <= (tipCut * tipCut));
    valid = valid && (cR >= minRadius);
    return valid;
}

void tracklet_offsets(Ctx& c);

bool filter(Ctx& c) {
    const Input& in = c.in;
    const u32 nCand = (u32)(c.cands.size() / 2);
    const u32 nOracle = (u32)std::ceil(nCand / 32.0);
    std::vector<u32> oracle(nOracle, 0);
    par_for(nOracle, in.n_workers, [&](u32 word) {
        u32 bits = 0;
        for (u32 b = 0; b < 32; ++b) {
            const u32 gid = word * 32 + b;
            if (gid >= nCand) break;
            const u32 pi = c.cands[2 * (size_t)gid], h3 = c.cands[2 * (size_t)gid + 1];
            if (pi == SENT || h3 == SENT) continue;
            const u32 h1 = c.pairs[2 * (size_t)pi], h2 = c.pairs[2 * (size_t)pi + 1];
            if (h1 == SENT || h2 == SENT) continue;
            const u32 lt = c.hlayer[h1] - 1;
            if (lt >= c.T) continue;
            if (filter_valid(c, h1, h2, h3, lt)) bits |= (1u << b);
        }
        oracle[word] = bits;
    });
    std::vector<u32> prefix(nOracle + 1, 0);
    for (u32 g = 0; g < nOracle; ++g) prefix[g] = (u32)__builtin_popcount(oracle[g]);
    prefix_sum(prefix, in.n_threads);
    const u32 R = prefix.back();
    c.th1.assign(R, 0); c.th2.assign(R, 0); c.th3.assign(R, 0);
    for (u32 g = 0; g < nOracle; ++g) {
        u32 pos = prefix[g];
        for (u32 b = 0; b < 32; ++b)
            if (oracle[g] & (1u << b)) {
                const u32 gid = g * 32 + b;
                const u32 pi = c.cands[2 * (size_t)gid];
                c.th1[pos] = c.pairs[2 * (size_t)pi];
                c.th2[pos] = c.pairs[2 * (size_t)pi + 1];
                c.th3[pos] = c.cands[2 * (size_t)gid + 1];
                pos++;
            }
    }
    tracklet_offsets(c);
    return true;
}

// filterOffsetStore (.h:233-256) and filterOffsetMonotonizeStore (.h:267-280)
void tracklet_offsets(Ctx& c) {
    const u32 R = (u32)c.th1.size();
    c.t_off.assign(c.E * c.T + 1, 0);
    const size_t n_off = c.t_off.size();
    for (u32 gid = 0; gid < R; ++gid) {
        const u32 hitId = c.th1[gid];
        const u32 layerTriplet = c.hlayer[hitId];
        const u32 event = c.hevt[hitId];
        auto put = [&](size_t idx, u32 v) { if (idx < n_off) c.t_off[idx] = v; };
        if (gid > 0) {
            const u32 prev = c.th1[gid - 1];
            if (layerTriplet != c.hlayer[prev] || event != c.hevt[prev])
                put((size_t)event * c.T + layerTriplet - 1, gid);
            if (gid + 1 == R) put((size_t)event * c.T + layerTriplet, gid);
        } else {
            put((size_t)event * c.T + layerTriplet - 1, gid);
        }
    }
    forward_fill(c.t_off);
}

// ---------------------------------------------------------------------------
// TripletConnectivityTight (.cpp:5-474; tripletEtaCalculatorStore .h:47-79,
// count :83-167, store :169-243, streamCompaction :251-273, inversion :327-346),
// TrackletCircleFitter (.cpp:4-48, .h:32-101) and CellularAutomaton (.cpp:7-464,
// .h:52-327).  Pinned by: hand-built chains (the reference's
// TripletConnectivityTest.h:66-98 pattern), brute-force linking, CA on path graphs.
float circle_radius(const Ctx& c, u32 h0, u32 h1, u32 h2) {
    const float x0 = c.hx[h0], y0 = c.hy[h0];
    const float x1 = c.hx[h1], y1 = c.hy[h1];
    const float x2 = c.hx[h2], y2 = c.hy[h2];
    const float p0x = x0, p0y = y0, p0z = x0 * x0 + y0 * y0;
    const float p1x = x1, p1y = y1, p1z = x1 * x1 + y1 * y1;
    const float p2x = x2, p2y = y2, p2z = x2 * x2 + y2 * y2;
    const float ax = p1x - p0x, ay = p1y - p0y, az = p1z - p0z;
    const float bx = p2x - p0x, by = p2y - p0y, bz = p2z - p0z;
    float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
    tx_normalize3(nx, ny, nz);
    const float cc = (TRAX_ORACLE_PLANT == 3) ? -(nx * p0x + ny * p0y) : -(nx * p0x + ny * p0y + nz * p0z);
    const float nz2 = nz * nz;
    return sqrtf((1 - nz2 - 4 * cc * nz) / (4 * nz2));
}

bool connectivity_ca(Ctx& c) {
    const Input& in = c.in;
    const u32 R = (u32)c.th1.size();
    const u32 nOff = (u32)c.t_off.size();
    std::vector<float> eta(R);
    for (u32 t = 0; t < R; ++t) {
        const u32 a = c.th1[t], b = (TRAX_ORACLE_PLANT == 2) ? c.th2[t] : c.th3[t];
        const float px = c.hx[b] - c.hx[a], py = c.hy[b] - c.hy[a], pz = c.hz[b] - c.hz[a];
        eta[t] = tx_asinhf(pz / sqrtf(px * px + py * py));
    }
    c.res.putf("eta", eta);
    std::vector<u32> counts(R + 1, 0), conn(R, 0);
    const u32 limit = (nOff >= 2) ? c.t_off[nOff - 2] : 0;
    auto range_of = [&](u32 t, u32& begin, u32& end, u32& ev) -> bool {
        if (t >= limit) return false;  // (.h:113-115)
        const u32 inner = c.th2[t];
        ev = c.hevt[inner];
        const u32 ln = c.hlayer[inner] - 1;
        const size_t ib = (size_t)ev * c.T + ln;
        if (ib + 1 >= nOff) return false;
        begin = c.t_off[ib]; end = c.t_off[ib + 1];
        if (begin >= R) return false;
        if (ev != c.hevt[c.th1[begin]]) return false;  // (.h:130-132)
        return true;
    };
    auto link = [&](u32 t, u32 i) {
        const float de = fabsf(eta[t] - eta[i]);
        const bool close = (TRAX_ORACLE_PLANT == 1) ? de <= in.d_eta_cut : de < in.d_eta_cut;
        return (c.th2[t] == c.th1[i]) && (c.th3[t] == c.th2[i]) && close;
    };
    for (u32 t = 0; t < R; ++t) {
        u32 begin, end, ev;
        if (!range_of(t, begin, end, ev)) continue;
        u32 n = 0;
        for (u32 i = begin; i < end; ++i)
            if (link(t, i)) { n++; conn[i] = 1; }
        counts[t] = n;
        if (n) conn[t] = 1;
    }
    prefix_sum(counts, in.n_threads);
    const u32 K = counts.back();
    c.res.put("conn.count_prefix", counts);
    std::vector<u32> basis(K), foll(K), pev(K);
    for (u32 t = 0; t < R && K; ++t) {
        u32 begin, end, ev;
        if (!range_of(t, begin, end, ev)) continue;
        u32 so = counts[t];
        const u32 sn = counts[t + 1];
        for (u32 i = begin; i < end && so < sn; ++i) {
            basis[so] = t; foll[so] = i; pev[so] = ev;
            so += link(t, i);
        }
    }
    std::vector<u32> connectable, nonconnectable;
    if (K) {
        for (u32 t = 0; t + 1 < R; ++t) (conn[t] ? connectable : nonconnectable).push_back(t);
    }
    c.res.put("conn.basis", basis);
    c.res.put("conn.followers", foll);
    c.res.put("conn.event", pev);
    c.res.put("conn.connectable", connectable);
    c.res.put("conn.nonconnectable", nonconnectable);
    if (!K) return true;  // TripletBuilder.cpp:414-422: group skipped

    std::vector<float> pt(R, 0.0f);
    for (u32 t : connectable) pt[t] = circle_radius(c, c.th1[t], c.th2[t], c.th3[t]);
    c.res.putf("pt", pt);

    // CA forward phase (.cpp:37-98, iteration .h:52-100)
    std::vector<u32> st(R, 1), nx(R, 1), living(R + 1, 0);
    u32 iterations = 0, alive;
    do {
        iterations++;
        std::fill(living.begin(), living.begin() + R, 0);
        for (u32 p = 0; p < K; ++p) {
            const u32 f = foll[p];
            const u32 fs = st[f];
            const u32 same = st[basis[p]] == fs;
            nx[f] = std::max(nx[f], fs + same);
            if (same) living[f] = 1;
        }
        alive = 0;
        for (u32 t = 0; t < R; ++t) alive += living[t];
        std::swap(st, nx);
    } while (alive > 0);
    c.res.put("ca.states", st);
    c.res.put("ca.iterations", {iterations});
    // backward phase (.cpp:107-352)
    std::vector<u32> fbc(R + 1, 0);
    for (u32 p = 0; p < K; ++p)
        if (st[basis[p]] + 1 == st[foll[p]]) fbc[foll[p]]++;
    prefix_sum(fbc, in.n_threads);
    const u32 nb = fbc.back();
    std::vector<u32> bidx(nb), cursor(fbc.begin(), fbc.end());
    std::vector<float> bdiff(nb);
    for (u32 p = 0; p < K; ++p) {  // connection-pair order
        const u32 b = basis[p], f = foll[p];
        if (st[b] + 1 != st[f]) continue;
        const u32 so = cursor[f]++;
        bidx[so] = b;
        bdiff[so] = fabsf(pt[b] - pt[f]);
    }
    // after the atomic increments the array holds the inclusive sums
    std::vector<u32> best(R, 0xffffffffu), isbest(R, 0);
    for (u32 t = 0; t < R; ++t) {
        const u32 begin = fbc[t], end = fbc[t + 1];
        if (begin == end) continue;
        float minPt = bdiff[begin];
        u32 minIdx = bidx[begin];
        for (u32 i = begin + 1; i < end; ++i) {  // (.h:254-259)
            const float d = bdiff[i];
            const u32 test = (TRAX_ORACLE_PLANT == 4) ? d > minPt : d < minPt;
            minPt = minPt * (float)(!test) + d * (float)test;
            minIdx = minIdx * (u32)(!test) + bidx[i] * test;
        }
        best[t] = minIdx;
        isbest[minIdx] = 1;
    }
    std::vector<u32> handler(R + 1, 0);
    for (u32 t = 0; t < R; ++t) handler[t] = (isbest[t] == 0) * st[t];
    prefix_sum(handler, in.n_threads);
    std::vector<u32> tracks(handler.back());
    for (u32 t = 0; t < R; ++t) {
        const u32 so = handler[t];
        const u32 len = handler[t + 1] - so;
        u32 b = t;
        for (u32 i = 0; i < len; ++i) {
            tracks[so + i] = b;
            if (b >= R) break;
            b = best[b];
        }
    }
    c.res.put("ca.best_basis", best);
    c.res.put("tracks", tracks);
    c.res.put("track_offsets", handler);
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

struct oracle_result;

oracle_result* oracle_run(const Input* in) {
    Result* r = new Result();
    Ctx c(*in, *r);
    auto t0 = std::chrono::steady_clock::now();
    auto lap = [&](const char* k) {
        auto t1 = std::chrono::steady_clock::now();
        r->timings_ms[k] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        t0 = t1;
    };
    bool ok = grid_build(c);
    lap("grid");
    auto want = [&](int stage) { return ok && (in->stop_after == 0 || in->stop_after >= stage); };
    if (ok) {
        r->put("grid", c.grid);
        r->putf("hits.x", c.hx); r->putf("hits.y", c.hy); r->putf("hits.z", c.hz);
        r->put("hits.layer", c.hlayer); r->put("hits.det_id", c.hdet);
        r->put("hits.hit_id", c.hid); r->put("hits.event", c.hevt);
    }
    if (want(2)) {
        ok = pair_gen(c);
        lap("pairs");
        r->put("pairs", c.pairs); r->put("pair_offsets", c.pair_off);
    }
    if (want(3)) {
        ok = predict(c);
        lap("predict");
        r->put("cands", c.cands); r->put("cand_offsets", c.cand_off);
    }
    if (want(4)) {
        ok = filter(c);
        lap("filter");
        r->put("tracklets.h1", c.th1); r->put("tracklets.h2", c.th2); r->put("tracklets.h3", c.th3);
        r->put("tracklet_offsets", c.t_off);
    }
    if (want(5)) {
        ok = connectivity_ca(c);
        lap("connectivity_ca");
    }
    r->put("diag", {c.diag_pair_short, c.diag_pred_short, c.diag_pred_trunc, c.diag_wrap_overlap});
    if (!ok && r->err.empty()) r->err = "failed";
    return reinterpret_cast<oracle_result*>(r);
}

const char* oracle_error(oracle_result* h) {
    Result* r = reinterpret_cast<Result*>(h);
    return r->err.empty() ? nullptr : r->err.c_str();
}

// returns element count, or -1 if absent; *is_float set; *data points to storage
long long oracle_get(oracle_result* h, const char* name, const void** data, int* is_float) {
    Result* r = reinterpret_cast<Result*>(h);
    auto it = r->arrays.find(name);
    if (it == r->arrays.end()) return -1;
    *is_float = it->second.is_float ? 1 : 0;
    if (it->second.is_float) { *data = it->second.f.data(); return (long long)it->second.f.size(); }
    *data = it->second.u.data();
    return (long long)it->second.u.size();
}

double oracle_timing(oracle_result* h, const char* stage) {
    Result* r = reinterpret_cast<Result*>(h);
    auto it = r->timings_ms.find(stage);
    return it == r->timings_ms.end() ? -1.0 : it->second;
}

void oracle_free(oracle_result* h) { delete reinterpret_cast<Result*>(h); }

// connectivity + circle fit + CA on given tracklets; the Input's hit columns
// are taken as already grid-ordered (hand-built fixtures)
oracle_result* oracle_run_link(const Input* in, const uint32_t* h1, const uint32_t* h2, const uint32_t* h3,
                               uint32_t n) {
    Result* r = new Result();
    Ctx c(*in, *r);
    c.hx.assign(in->x, in->x + in->n_hits);
    c.hy.assign(in->y, in->y + in->n_hits);
    c.hz.assign(in->z, in->z + in->n_hits);
    c.hlayer.assign(in->layer, in->layer + in->n_hits);
    c.hevt.assign(in->event, in->event + in->n_hits);
    c.th1.assign(h1, h1 + n);
    c.th2.assign(h2, h2 + n);
    c.th3.assign(h3, h3 + n);
    tracklet_offsets(c);
    r->put("tracklet_offsets", c.t_off);
    connectivity_ca(c);
    return reinterpret_cast<oracle_result*>(r);
}

// the reference's PrefixSum on a host array (KAT and brute-force tests)
void oracle_prefix_sum(uint32_t* data, uint32_t n, uint32_t wg) {
    if (n) recursive_scan(data, n, wg);
}

}  // extern "C"
