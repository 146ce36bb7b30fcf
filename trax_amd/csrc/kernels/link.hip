// TripletThetaPhiFilter, TripletConnectivityTight, TrackletCircleFitter and
// CellularAutomaton kernels (src/algorithms/TripletThetaPhiFilter.h:50-280,
// TripletConnectivityTight.h:47-346, TrackletCircleFitter.h:32-101,
// CellularAutomaton.h:52-327 and their .cpp drivers).
//
// Filter: one lane per candidate; the cut results of a wave become one 64-bit
//   word of the validity bitmap (ballot) and its popcount, the popcounts are
//   scanned and each wave writes its tracklets in candidate order.
// Connectivity: the reference compares every base triplet with every triplet
//   of the next layer triplet (O(R_l * R_l+1)).  A follower must start with the
//   base's middle hit, so here the tracklets are bucketed by their first hit
//   (a CSR built with one atomic count, one scan and one fill) and each base
//   tests only the tracklets of its middle hit, restricted to the same
//   [begin, end) offsets window the reference scans; matches are emitted in
//   ascending follower index, the reference's order.
// CA: the forward iterations run as a fixed number of launches with a device
//   "alive" flag (an iteration after the one that found no living cell returns
//   at once), so no host read-back is needed per iteration.
#include "filter_cuts.h"
#include "kernels.h"

namespace trax {
namespace {

__device__ inline uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline uint32_t gstride() { return gridDim.x * blockDim.x; }

// A candidate's hits (false for sentinel slots, which yield nothing, and for
// a pair or hit index out of range, which is reported)
__device__ inline bool cand_hits(const TripletCols& tc, const uint32_t* hlayer, const uint32_t* pairs,
                                 const uint32_t* cands, uint32_t gid, uint32_t n_pairs, uint32_t nh, DevCounts* cnt,
                                 uint32_t& h1, uint32_t& h2, uint32_t& h3, uint32_t& lt) {
    const uint32_t pi = cands[2 * (size_t)gid];
    h3 = cands[2 * (size_t)gid + 1];
    if (pi == kSentinel || h3 == kSentinel) return false;
    if (!index_ok(pi, n_pairs, cnt) || !index_ok(h3, nh, cnt)) return false;
    h1 = pairs[2 * (size_t)pi];
    h2 = pairs[2 * (size_t)pi + 1];
    if (h1 == kSentinel || h2 == kSentinel) return false;
    if (!index_ok(h1, nh, cnt) || !index_ok(h2, nh, cnt)) return false;
    lt = hlayer[h1] - 1;
    return lt < tc.n_trip;
}

// the three hits of tracklet t, checked against the hit count
__device__ inline bool trk_hits_ok(const uint32_t* th1, const uint32_t* th2, const uint32_t* th3, uint32_t t,
                                   uint32_t nh, DevCounts* cnt, uint32_t& a, uint32_t& m, uint32_t& b) {
    a = th1[t];
    m = th2[t];
    b = th3[t];
    return index_ok(a, nh, cnt) && index_ok(m, nh, cnt) && index_ok(b, nh, cnt);
}

// filterCount (TripletThetaPhiFilter.h:50-162) with the validity bitmap built
// per 64-candidate word.  Each wave takes kFilterWords words at a time: the
// phi-cut prefilter (tx_phi_surely_outside) settles most candidates, the
// rest are queued in the wave's LDS and evaluated with the reference's cuts
// 64 at a time, so that the full evaluation runs on whole waves of
// candidates that need it.  Dynamic LDS: (kFilterWords * 64 + 2 * kFilterWords)
// words per wave.
constexpr int kFilterWords = 8;
constexpr uint32_t kFilterWaveLds = kFilterWords * 64 + 2 * kFilterWords;

__global__ void k_filter_count(TripletCols tc, const float* __restrict__ hx, const float* __restrict__ hy,
                               const float* __restrict__ hz, const uint32_t* __restrict__ hlayer,
                               const uint32_t* __restrict__ pairs, const uint32_t* __restrict__ cands,
                               uint64_t* __restrict__ bitmap, uint32_t* __restrict__ word_counts,
                               DevCounts* __restrict__ cnt) {
    extern __shared__ uint32_t lds[];
    if (run_void(cnt)) return;
    const uint32_t n = bounded(cnt->n_cands, cnt->cap_cands, cnt);
    const uint32_t n_pairs = bounded(cnt->n_pairs, cnt->cap_pairs, cnt), nh = cnt->n_hits;
    const uint32_t nwords = (n + 63) / 64;
    if (gtid() == 0) {
        cnt->n_filter_words = nwords;
        word_counts[nwords] = 0;
    }
    const uint32_t lane = lane_id();
    uint32_t* queue = lds + (threadIdx.x / kWave) * kFilterWaveLds;
    uint64_t* masks = reinterpret_cast<uint64_t*>(queue + kFilterWords * 64);
    const uint32_t n_batches = (nwords + kFilterWords - 1) / kFilterWords;
    for (uint32_t bt = gtid() / kWave; bt < n_batches; bt += gstride() / kWave) {
        const uint32_t w0 = bt * kFilterWords;
        uint32_t qn = 0;
        for (int w = 0; w < kFilterWords; ++w) {
            const uint32_t gid = (w0 + w) * 64 + lane;
            bool undecided = false;
            uint32_t h1, h2, h3, lt;
            if (gid < n && cand_hits(tc, hlayer, pairs, cands, gid, n_pairs, nh, cnt, h1, h2, h3, lt))
                undecided = !tx_phi_surely_outside(hx[h2] - hx[h1], hy[h2] - hy[h1], hx[h3] - hx[h2],
                                                   hy[h3] - hy[h2], tc.tan_dphi[lt]);
            const uint64_t m = __ballot(undecided);
            if (undecided) queue[qn + (uint32_t)__popcll(m & lanemask_lt())] = gid;
            qn += (uint32_t)__popcll(m);
        }
        if (lane < kFilterWords) masks[lane] = 0ull;
        wave_sync();
        for (uint32_t k = 0; k < qn; k += 64) {
            if (k + lane < qn) {
                const uint32_t gid = queue[k + lane];
                uint32_t h1, h2, h3, lt;
                if (cand_hits(tc, hlayer, pairs, cands, gid, n_pairs, nh, cnt, h1, h2, h3, lt) &&
                    filter_valid(tc, hx, hy, hz, h1, h2, h3, lt))
                    atomicOr(reinterpret_cast<unsigned long long*>(&masks[gid / 64 - w0]), 1ull << (gid & 63));
            }
        }
        wave_sync();
        if (lane < kFilterWords && w0 + lane < nwords) {
            const uint64_t m = masks[lane];
            bitmap[w0 + lane] = m;
            word_counts[w0 + lane] = (uint32_t)__popcll(m);
        }
        wave_sync();  // the queue and masks are reused by the next batch
    }
}

__global__ void k_filter_store(const uint32_t* __restrict__ pairs, const uint32_t* __restrict__ cands,
                               const uint64_t* __restrict__ bitmap, const uint32_t* __restrict__ word_counts,
                               uint32_t* __restrict__ th1, uint32_t* __restrict__ th2, uint32_t* __restrict__ th3,
                               uint32_t n_trk_cap, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t nwords = cnt->n_filter_words;
    const uint32_t n_pairs = bounded(cnt->n_pairs, cnt->cap_pairs, cnt);
    const uint32_t lane = lane_id();
    bool over = false;
    for (uint32_t word = gtid() / kWave; word < nwords; word += gstride() / kWave) {
        const uint64_t m = bitmap[word];
        if ((m >> lane) & 1ull) {
            const uint32_t pos = word_counts[word] + (uint32_t)__popcll(m & lanemask_lt());
            const uint32_t gid = word * 64 + lane;
            const uint32_t pi = cands[2 * (size_t)gid];
            if (!index_ok(pi, n_pairs, cnt)) continue;
            if (pos < n_trk_cap) {
                th1[pos] = pairs[2 * (size_t)pi];
                th2[pos] = pairs[2 * (size_t)pi + 1];
                th3[pos] = cands[2 * (size_t)gid + 1];
            } else {
                over = true;
            }
        }
    }
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

// filterOffsetStore (TripletThetaPhiFilter.h:221-261), literal: the entry after
// the last tracklet's group receives its index R - 1
__global__ void k_filter_offsets(TripletCols tc, const uint32_t* __restrict__ hlayer,
                                 const uint32_t* __restrict__ hevt, const uint32_t* __restrict__ th1,
                                 uint32_t* __restrict__ t_off, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), nh = cnt->n_hits;
    const uint64_t n_off = (uint64_t)cnt->n_events * tc.n_trip + 1;
    for (uint32_t gid = gtid(); gid < R; gid += gstride()) {
        const uint32_t hitId = th1[gid];
        if (!index_ok(hitId, nh, cnt)) continue;
        const uint32_t layerTriplet = hlayer[hitId];
        const uint32_t event = hevt[hitId];
        const uint64_t base = (uint64_t)event * tc.n_trip + layerTriplet;
        if (gid > 0) {
            const uint32_t prev = th1[gid - 1];
            if (!index_ok(prev, nh, cnt)) continue;
            if (layerTriplet != hlayer[prev] || event != hevt[prev])
                if (base - 1 < n_off) t_off[base - 1] = gid;
            if (gid + 1 == R && base < n_off) t_off[base] = gid;
        } else if (base - 1 < n_off) {
            t_off[base - 1] = gid;
        }
    }
}

// ---------------------------------------------------------------------------
__global__ void k_eta(const float* __restrict__ hx, const float* __restrict__ hy, const float* __restrict__ hz,
                      const uint32_t* __restrict__ th1, const uint32_t* __restrict__ th3,
                      float* __restrict__ eta, uint32_t* __restrict__ hit_count,
                      DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), nh = cnt->n_hits;
    for (uint32_t t = gtid(); t < R; t += gstride()) {  // tripletEtaCalculatorStore (.h:60-75)
        const uint32_t a = th1[t], b = th3[t];
        if (!(index_ok(a, nh, cnt) && index_ok(b, nh, cnt))) continue;
        const float px = hx[b] - hx[a], py = hy[b] - hy[a], pz = hz[b] - hz[a];
        const float tt = pz / sqrtf(px * px + py * py);
        eta[t] = tx_asinhf(tt);
        atomicAdd(&hit_count[a], 1u);
    }
}

// η per tracklet as in k_eta, and the run of every middle hit: run_start[h],
// run_end[h] (tracklets of one middle hit are consecutive when each layer is
// the middle layer of at most one layer triplet)
__global__ void k_eta_runs(const float* __restrict__ hx, const float* __restrict__ hy, const float* __restrict__ hz,
                           const uint32_t* __restrict__ th1, const uint32_t* __restrict__ th2,
                           const uint32_t* __restrict__ th3, float* __restrict__ eta,
                           uint32_t* __restrict__ run_start, uint32_t* __restrict__ run_end,
                           DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), nh = cnt->n_hits;
    for (uint32_t t = gtid(); t < R; t += gstride()) {  // tripletEtaCalculatorStore (.h:60-75)
        uint32_t a, m, b;
        if (!trk_hits_ok(th1, th2, th3, t, nh, cnt, a, m, b)) continue;
        const float px = hx[b] - hx[a], py = hy[b] - hy[a], pz = hz[b] - hz[a];
        const float tt = pz / sqrtf(px * px + py * py);
        eta[t] = tx_asinhf(tt);
        if (t == 0 || th2[t - 1] != m) run_start[m] = t;
        if (t + 1 == R || th2[t + 1] != m) run_end[m] = t + 1;
    }
}

__global__ void k_csr_fill(const uint32_t* __restrict__ th1, const uint32_t* __restrict__ hit_start,
                           uint32_t* __restrict__ hit_cursor, uint32_t* __restrict__ csr,
                           DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), nh = cnt->n_hits;
    for (uint32_t t = gtid(); t < R; t += gstride()) {
        const uint32_t h = th1[t];
        if (!index_ok(h, nh, cnt)) continue;
        const uint32_t q = hit_start[h] + atomicAdd(&hit_cursor[h], 1u);
        if (index_ok(q, R, cnt)) csr[q] = t;
    }
}

struct LinkRange {
    uint32_t begin, end, ev;
};

// guards of tripletConnectivityTightCount/Store (.h:109-132).  The offsets are
// the group's E*T + 1 tracklet offsets, or with each event its own group the
// event's T + 1 offsets as a run of that event alone has them.
__device__ inline bool link_range(const TripletCols& tc, const uint32_t* hlayer, const uint32_t* hevt,
                                  const uint32_t* th1, const uint32_t* th2, const uint32_t* t_off,
                                  uint32_t n_off, uint32_t R, uint32_t t, const LinkBufs& b, LinkRange& r,
                                  DevCounts* cnt) {
    const uint32_t inner = th2[t], nh = cnt->n_hits;
    if (!index_ok(inner, nh, cnt)) return false;
    r.ev = hevt[inner];
    if (!index_ok(r.ev, cnt->n_events, cnt)) return false;
    const uint32_t* off = t_off;
    uint64_t ib = (uint64_t)r.ev * tc.n_trip + (hlayer[inner] - 1);
    if (b.per_event) {
        off = b.ev_off + (uint64_t)r.ev * (tc.n_trip + 1);
        n_off = tc.n_trip + 1;
        ib = hlayer[inner] - 1;
    }
    const uint32_t limit = n_off >= 2 ? off[n_off - 2] : 0u;
    if (t >= limit) return false;
    if (ib + 1 >= n_off) return false;
    r.begin = off[ib];
    r.end = off[ib + 1];
    if (r.begin >= R) return false;
    if (!index_ok(r.end, R + 1, cnt)) return false;  // offsets are tracklet positions, at most R
    const uint32_t h = th1[r.begin];
    if (!index_ok(h, nh, cnt)) return false;
    return r.ev == hevt[h];
}

// first and last tracklet of each event (tracklets are ordered by event)
__global__ void k_event_bounds(const uint32_t* __restrict__ hevt, const uint32_t* __restrict__ th1, LinkBufs b,
                               DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), nh = cnt->n_hits, E = cnt->n_events;
    // the event of tracklet t (kSentinel for a bad index, which is reported)
    auto event_of = [&](uint32_t t) {
        const uint32_t h = th1[t];
        if (!index_ok(h, nh, cnt)) return kSentinel;
        const uint32_t e = hevt[h];
        return index_ok(e, E, cnt) ? e : kSentinel;
    };
    for (uint32_t t = gtid(); t < R; t += gstride()) {
        const uint32_t e = event_of(t);
        if (e == kSentinel) continue;
        if (t == 0 || event_of(t - 1) != e) b.ev_first[e] = t;
        if (t + 1 == R || event_of(t + 1) != e) b.ev_last[e] = t;
    }
}

// The tracklet offsets a run of event e alone produces (filterOffsetStore and
// its monotonize, TripletThetaPhiFilter.h:221-280), shifted by the event's
// first tracklet f: groups before the first non-empty one hold f, the
// non-empty range holds the group starts (empty groups inside it the
// previous start), and from one past the last non-empty group on the value
// is the event's last tracklet l (the run's R - 1).
__global__ void k_event_offsets(TripletCols tc, const uint32_t* __restrict__ hlayer,
                                const uint32_t* __restrict__ th1, const uint32_t* __restrict__ t_off, LinkBufs b,
                                DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t T = tc.n_trip, E = cnt->n_events;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), nh = cnt->n_hits;
    const uint64_t n = (uint64_t)E * (T + 1);
    for (uint64_t i = gtid(); i < n; i += gstride()) {
        const uint32_t e = (uint32_t)(i / (T + 1)), k = (uint32_t)(i % (T + 1));
        const uint32_t f = b.ev_first[e], l = b.ev_last[e];
        if (f == kSentinel) {
            b.ev_off[i] = kSentinel;
            continue;
        }
        if (!(index_ok(f, R, cnt) && index_ok(l, R, cnt)) || !(index_ok(th1[f], nh, cnt) && index_ok(th1[l], nh, cnt)))
            continue;
        const uint32_t lf = hlayer[th1[f]] - 1, ll = hlayer[th1[l]] - 1;
        b.ev_off[i] = k < lf ? f : (k <= ll ? t_off[(uint64_t)e * T + k] : l);
    }
}

__global__ void k_link_count(TripletCols tc, const uint32_t* __restrict__ hlayer, const uint32_t* __restrict__ hevt,
                             const uint32_t* __restrict__ th1, const uint32_t* __restrict__ th2,
                             const uint32_t* __restrict__ th3, const uint32_t* __restrict__ t_off, LinkBufs b,
                             DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    const uint32_t n_off = cnt->n_events * tc.n_trip + 1;
    for (uint32_t t = gtid(); t <= R; t += gstride()) {
        uint32_t n = 0;
        LinkRange r;
        if (t < R && link_range(tc, hlayer, hevt, th1, th2, t_off, n_off, R, t, b, r, cnt)) {
            const uint32_t h = th2[t], h3 = th3[t];
            const float et = b.eta[t];
            for (uint32_t q = b.hit_count[h]; q < b.hit_count[h + 1]; ++q) {
                const uint32_t i = b.csr[q];
                if (!index_ok(i, R, cnt)) continue;
                if (i >= r.begin && i < r.end && th2[i] == h3 && fabsf(et - b.eta[i]) < tc.d_eta_cut) {
                    n++;
                    b.conn[i] = 1u;
                }
            }
            if (n) {
                b.conn[t] = 1u;
                if (b.per_event) b.ev_has[r.ev] = 1u;
            }
        }
        b.counts[t] = n;
    }
}

__global__ void k_link_store(TripletCols tc, const uint32_t* __restrict__ hlayer, const uint32_t* __restrict__ hevt,
                             const uint32_t* __restrict__ th1, const uint32_t* __restrict__ th2,
                             const uint32_t* __restrict__ th3, const uint32_t* __restrict__ t_off, LinkBufs b,
                             uint32_t n_conn_cap, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    const uint32_t n_off = cnt->n_events * tc.n_trip + 1;
    bool over = false;
    for (uint32_t t = gtid(); t < R; t += gstride()) {
        uint32_t so = b.counts[t];
        const uint32_t sn_all = b.counts[t + 1];
        if (so >= sn_all) continue;
        // a base whose connections run past the capacity writes the entries
        // below it (the run is void, but none below the capacity is unwritten)
        if (sn_all > n_conn_cap) over = true;
        const uint32_t sn = sn_all < n_conn_cap ? sn_all : n_conn_cap;
        if (so >= sn) continue;
        LinkRange r;
        if (!link_range(tc, hlayer, hevt, th1, th2, t_off, n_off, R, t, b, r, cnt)) continue;
        const uint32_t h = th2[t], h3 = th3[t];
        const float et = b.eta[t];
        const uint32_t q0 = b.hit_count[h], q1 = b.hit_count[h + 1];
        auto follows = [&](uint32_t i) {
            return index_ok(i, R, cnt) && i >= r.begin && i < r.end && th2[i] == h3 &&
                   fabsf(et - b.eta[i]) < tc.d_eta_cut;
        };
        // the bucket is in fill (atomic) order: collect the followers, sort
        // them and emit them in ascending index, the reference's order
        constexpr uint32_t kFew = 16;
        if (sn - so <= kFew) {
            uint32_t f[kFew];
            uint32_t k = 0;
            for (uint32_t q = q0; q < q1 && k < sn - so; ++q) {
                const uint32_t i = b.csr[q];
                if (follows(i)) {
                    uint32_t p = k++;
                    while (p > 0 && f[p - 1] > i) {
                        f[p] = f[p - 1];
                        --p;
                    }
                    f[p] = i;
                }
            }
            for (uint32_t x = 0; x < k; ++x, ++so) {
                b.basis[so] = t;
                b.foll[so] = f[x];
                b.pev[so] = r.ev;
            }
            for (; so < sn; ++so) b.basis[so] = b.foll[so] = b.pev[so] = kSentinel;
        } else {
            uint32_t last = 0;
            bool first = true;
            for (; so < sn; ++so) {  // many followers: select them one by one in ascending index
                uint32_t best = kSentinel;
                for (uint32_t q = q0; q < q1; ++q) {
                    const uint32_t i = b.csr[q];
                    if ((first || i > last) && i < best && follows(i)) best = i;
                }
                if (best == kSentinel) break;
                b.basis[so] = t;
                b.foll[so] = best;
                b.pev[so] = r.ev;
                last = best;
                first = false;
            }
            for (; so < sn; ++so) b.basis[so] = b.foll[so] = b.pev[so] = kSentinel;
        }
    }
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

// tripletConnectivityTightCount and Store in one pass (batched path): a wave
// takes a chunk of 64 consecutive bases, each lane counts its base's
// followers (setting the connectivity flags as the count kernel does), the
// wave takes room for the chunk's connections in the scratch from a range it
// holds, and each lane writes its (base, follower, event) entries there in
// ascending follower index, the reference's order: the bucket (in fill order)
// is read once and the followers are kept sorted in the lane's kLinkFew LDS
// slots; a base with more is emitted by selection, smallest first.  The
// chunks finish in any order; launch_link_fused scans their counts and
// launch_link_fused_place copies them into base order.  Dynamic LDS:
// kLinkFew words per thread.
constexpr uint32_t kLinkFew = 16;
__global__ void k_link_fused(TripletCols tc, const uint32_t* __restrict__ hlayer, const uint32_t* __restrict__ hevt,
                             const uint32_t* __restrict__ th1, const uint32_t* __restrict__ th2,
                             const uint32_t* __restrict__ th3, const uint32_t* __restrict__ t_off, LinkBufs b,
                             DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), nh = cnt->n_hits;
    const uint32_t n_off = cnt->n_events * tc.n_trip + 1;
    const uint32_t n_chunks = (R + kWave - 1) / kWave;
    if (gtid() == 0) cnt->n_link_chunks = n_chunks;
    extern __shared__ uint32_t s_few[];
    uint32_t* few = s_few + threadIdx.x * kLinkFew;
    const uint32_t lane = lane_id(), first_wave = gtid() / kWave, n_waves = gstride() / kWave;
    uint32_t range_at = 0, range_left = 0;  // the wave's scratch range (wave-uniform)
    bool over = false;
    for (uint32_t chunk = first_wave; chunk < n_chunks; chunk += n_waves) {
        const uint32_t t = chunk * kWave + lane;
        uint32_t n = 0, q0 = 0, q1 = 0, h3 = 0;
        float et = 0.0f;
        LinkRange r{0, 0, 0};
        if (t < R && b.runs && link_range(tc, hlayer, hevt, th1, th2, t_off, n_off, R, t, b, r, cnt)) {
            // the run of the third hit: tracklets i with th2[i] == h3, in ascending index
            const uint32_t h = th2[t];
            h3 = th3[t];
            et = b.eta[t];
            const bool h3_ok = index_ok(h3, nh, cnt);
            const uint32_t i0 = h3_ok ? b.hit_count[h3] : 0u, i1 = h3_ok ? b.hit_cursor[h3] : 0u;
            const uint32_t lo = i0 > r.begin ? i0 : r.begin, hi = i1 < r.end ? i1 : r.end;
            for (uint32_t i = lo; i < hi; ++i) {
                if (th1[i] == h && fabsf(et - b.eta[i]) < tc.d_eta_cut) {
                    if (n < kLinkFew) few[n] = i;
                    n++;
                    b.conn[i] = 1u;
                }
            }
            if (n) {
                b.conn[t] = 1u;
                if (b.per_event) b.ev_has[r.ev] = 1u;
            }
            if (n > kLinkFew) {  // emitted by the selection below: the same tests over the CSR-less run
                q0 = lo;
                q1 = hi;
            }
        } else if (t < R && !b.runs && link_range(tc, hlayer, hevt, th1, th2, t_off, n_off, R, t, b, r, cnt)) {
            const uint32_t h = th2[t];
            h3 = th3[t];
            et = b.eta[t];
            q0 = b.hit_count[h];
            q1 = b.hit_count[h + 1];
            for (uint32_t q = q0; q < q1; ++q) {
                const uint32_t i = b.csr[q];
                if (!index_ok(i, R, cnt)) continue;
                if (i >= r.begin && i < r.end && th2[i] == h3 && fabsf(et - b.eta[i]) < tc.d_eta_cut) {
                    if (n < kLinkFew) {  // insertion into the sorted slots
                        uint32_t p = n;
                        while (p > 0 && few[p - 1] > i) {
                            few[p] = few[p - 1];
                            --p;
                        }
                        few[p] = i;
                    }
                    n++;
                    b.conn[i] = 1u;
                }
            }
            if (n) {
                b.conn[t] = 1u;
                if (b.per_event) b.ev_has[r.ev] = 1u;
            }
        }
        const uint32_t incl = wave_incl_scan(n);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        if (total > range_left) {  // uniform
            const uint32_t size = total > kChunkRange ? total : kChunkRange;
            uint32_t a = 0;
            if (lane == 0) a = atomicAdd(&cnt->link_alloc, size);
            range_at = __shfl(a, 0, kWave);
            range_left = size;
        }
        const uint32_t at = range_at;
        range_at += total;
        range_left -= total;
        if (lane == kWave - 1) {
            b.chunk_n[chunk] = total;
            b.chunk_at[chunk] = at;
        }
        uint32_t o = at + incl - n;
        auto emit = [&](uint32_t f) {
            if (o < b.scratch_cap) {
                b.scratch[o] = t;
                b.scratch[(size_t)b.scratch_cap + o] = f;
                b.scratch[2 * (size_t)b.scratch_cap + o] = r.ev;
            } else {
                over = true;
            }
            ++o;
        };
        if (n <= kLinkFew) {
            for (uint32_t k = 0; k < n; ++k) emit(few[k]);
        } else {
            uint32_t last = 0;
            if (b.runs) {  // the run is in ascending index: every follower in turn
                const uint32_t h = th2[t];
                for (uint32_t i = q0; i < q1; ++i)
                    if (th1[i] == h && fabsf(et - b.eta[i]) < tc.d_eta_cut) emit(i);
            } else {
                for (uint32_t k = 0; k < n; ++k) {
                    uint32_t best = kSentinel;
                    for (uint32_t q = q0; q < q1; ++q) {
                        const uint32_t i = b.csr[q];
                        if ((k == 0 || i > last) && i < best && i >= r.begin && i < r.end && i < R && th2[i] == h3 &&
                            fabsf(et - b.eta[i]) < tc.d_eta_cut)
                            best = i;
                    }
                    emit(best);
                    last = best;
                }
            }
        }
    }
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

// predicates over the first R - 1 tracklets (TripletConnectivityTight.cpp:184-256);
// nothing when no connection exists (the reference returns null lists, .cpp:93-95).
// With each event its own group: over every event's tracklets but its last,
// for the events that have a connection.
__global__ void k_link_flags(const uint32_t* __restrict__ hevt, const uint32_t* __restrict__ th1, LinkBufs b,
                             DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), K = bounded(cnt->n_conn, cnt->cap_conn, cnt);
    const uint32_t nh = cnt->n_hits, E = cnt->n_events;
    for (uint32_t t = gtid(); t <= R; t += gstride()) {
        bool in = K > 0 && t + 1 < R;
        if (b.per_event && t < R) {
            const uint32_t h = th1[t];
            const uint32_t e = index_ok(h, nh, cnt) ? hevt[h] : kSentinel;
            in = e != kSentinel && index_ok(e, E, cnt) && b.ev_has[e] != 0u && t != b.ev_last[e];
        }
        const uint32_t c = (t < R) ? b.conn[t] : 0u;
        b.flag_a[t] = in ? (c != 0u) : 0u;
        b.flag_b[t] = in ? (c == 0u) : 0u;
    }
}

__global__ void k_link_compact(LinkBufs b, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    for (uint32_t t = gtid(); t + 1 < R; t += gstride()) {
        const uint32_t a = b.flag_a[t], c = b.flag_b[t];  // scanned flags: list positions below R
        if (a != b.flag_a[t + 1] && index_ok(a, R, cnt)) b.connectable[a] = t;
        if (c != b.flag_b[t + 1] && index_ok(c, R, cnt)) b.nonconnectable[c] = t;
    }
}

// trackletCircleFitterStore (TrackletCircleFitter.h:47-97)
__global__ void k_circle(const float* __restrict__ hx, const float* __restrict__ hy,
                         const uint32_t* __restrict__ th1, const uint32_t* __restrict__ th2,
                         const uint32_t* __restrict__ th3, LinkBufs b, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t n = bounded(cnt->n_connectable, cnt->cap_tracklets, cnt);
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), nh = cnt->n_hits;
    for (uint32_t j = gtid(); j < n; j += gstride()) {
        const uint32_t t = b.connectable[j];
        if (!index_ok(t, R, cnt)) continue;
        uint32_t h0, h1, h2;
        if (!trk_hits_ok(th1, th2, th3, t, nh, cnt, h0, h1, h2)) continue;
        const float x0 = hx[h0], y0 = hy[h0], x1 = hx[h1], y1 = hy[h1], x2 = hx[h2], y2 = hy[h2];
        const float p0x = x0, p0y = y0, p0z = x0 * x0 + y0 * y0;
        const float p1x = x1, p1y = y1, p1z = x1 * x1 + y1 * y1;
        const float p2x = x2, p2y = y2, p2z = x2 * x2 + y2 * y2;
        const float ax = p1x - p0x, ay = p1y - p0y, az = p1z - p0z;
        const float bx = p2x - p0x, by = p2y - p0y, bz = p2z - p0z;
        float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
        tx_normalize3(nx, ny, nz);
        const float cc = -(nx * p0x + ny * p0y + nz * p0z);
        const float nz2 = nz * nz;
        b.pt[t] = sqrtf((1 - nz2 - 4 * cc * nz) / (4 * nz2));
    }
}

// ---------------------------------------------------------------------------
__global__ void k_ca_init(CABufs b, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    for (uint32_t t = gtid(); t < R; t += gstride()) { b.st_a[t] = 1u; b.st_b[t] = 1u; }
    if (gtid() == 0) {
        b.flags[0] = 0u; b.flags[1] = 0u; b.flags[2] = 1u;
        cnt->ca_iterations = 0;
    }
}

// CellularAutomaton iteration (.h:52-100) number k of the forward phase
__global__ void k_ca_iter(const uint32_t* __restrict__ basis, const uint32_t* __restrict__ foll, CABufs b,
                          uint32_t k, DevCounts* __restrict__ cnt) {
    const uint32_t prev = b.flags[(k + 2) % 3];
    if (gtid() == 0) b.flags[(k + 1) % 3] = 0u;
    if (run_void(cnt)) return;  // no living cell is reported: the iterations stop
    const uint32_t K = bounded(cnt->n_conn, cnt->cap_conn, cnt), R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    if (!prev || K == 0) return;
    if (gtid() == 0) atomicAdd(&cnt->ca_iterations, 1u);
    const uint32_t* cur = (k & 1u) ? b.st_b : b.st_a;
    uint32_t* nxt = (k & 1u) ? b.st_a : b.st_b;
    bool alive = false;
    for (uint32_t p = gtid(); p < K; p += gstride()) {
        const uint32_t f = foll[p], bs = basis[p];
        if (!(index_ok(f, R, cnt) && index_ok(bs, R, cnt))) continue;
        const uint32_t fs = cur[f];
        const uint32_t same = cur[bs] == fs;
        atomicMax(&nxt[f], fs + same);
        alive |= same != 0u;
    }
    if (__ballot(alive) != 0ull && lane_id() == 0) b.flags[k % 3] = 1u;
}

// The forward phase by layers.  From all states 1, the synchronous update
// (.h:52-100) raises a follower by one while some basis has its state, so at
// iteration k a tracklet holds min(k, L) + 1, L the longest chain of
// connections ending at it, and the iterations stop at max L + 1.  When every
// connection goes to a higher layer triplet, st[f] = max(st[f], st[b] + 1)
// over the connections taken in layer-triplet order of their bases gives the
// same final states: each pass reads bases that no later pass writes.
// each connection's basis layer triplet, and per block of a fixed grid (each
// a contiguous range of connections) its connections per layer triplet, in a
// table tab[lt * gridDim.x + block] that a scan turns into bucket positions
__device__ inline void ca_block_range(uint32_t K, uint32_t& lo, uint32_t& hi) {
    const uint32_t per = (K + gridDim.x - 1) / gridDim.x;
    lo = blockIdx.x * per < K ? blockIdx.x * per : K;
    hi = lo + per < K ? lo + per : K;
}

__global__ void k_ca_conn_layer(const uint32_t* __restrict__ basis, const uint32_t* __restrict__ th1,
                                const uint32_t* __restrict__ hlayer, uint32_t n_trip, uint8_t* __restrict__ clt,
                                uint32_t* __restrict__ tab, DevCounts* __restrict__ cnt) {
    __shared__ uint32_t hist[256];
    if (block_run_void(cnt)) return;
    for (uint32_t i = threadIdx.x; i < n_trip; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const uint32_t K = bounded(cnt->n_conn, cnt->cap_conn, cnt), R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    const uint32_t nh = cnt->n_hits;
    uint32_t lo, hi;
    ca_block_range(K, lo, hi);
    bool bad = false;
    for (uint32_t p = lo + threadIdx.x; p < hi; p += blockDim.x) {
        const uint32_t bs = basis[p];
        const uint32_t h = bs < R ? th1[bs] : kSentinel;
        const uint32_t lt = h < nh ? hlayer[h] - 1u : kSentinel;
        clt[p] = (uint8_t)lt;  // 0xff for an entry that is skipped (lt >= n_trip, and n_trip < 256)
        if (lt < n_trip) atomicAdd(&hist[lt], 1u);
        else bad = true;
    }
    if (bad) atomicOr(&cnt->err, kErrIndexRange);  // a basis outside the configured layer triplets
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n_trip; i += blockDim.x) tab[i * gridDim.x + blockIdx.x] = hist[i];
}

// the block's connections, as (basis, follower), to their bucket positions
// (order within a bucket free)
__global__ void k_ca_conn_scatter(const uint32_t* __restrict__ basis, const uint32_t* __restrict__ foll,
                                  const uint8_t* __restrict__ clt, uint32_t n_trip, const uint32_t* __restrict__ tab,
                                  uint2* __restrict__ lt_conn, DevCounts* __restrict__ cnt) {
    __shared__ uint32_t cur[256];
    if (block_run_void(cnt)) return;
    for (uint32_t i = threadIdx.x; i < n_trip; i += blockDim.x) cur[i] = tab[i * gridDim.x + blockIdx.x];
    __syncthreads();
    const uint32_t K = bounded(cnt->n_conn, cnt->cap_conn, cnt);
    uint32_t lo, hi;
    ca_block_range(K, lo, hi);
    for (uint32_t p = lo + threadIdx.x; p < hi; p += blockDim.x) {
        const uint32_t lt = clt[p];
        if (lt < n_trip) lt_conn[atomicAdd(&cur[lt], 1u)] = make_uint2(basis[p], foll[p]);
    }
}

__global__ void k_ca_layer(const uint2* __restrict__ lt_conn, const uint32_t* __restrict__ tab, uint32_t tab_blocks,
                           uint32_t n_trip, uint32_t lt, CABufs b, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), K = bounded(cnt->n_conn, cnt->cap_conn, cnt);
    // the bucket: from its first table entry to the next bucket's (the total after the last)
    const uint32_t lo = tab[lt * tab_blocks], hi0 = tab[lt + 1 < n_trip ? (lt + 1) * tab_blocks : n_trip * tab_blocks];
    const uint32_t hi = hi0 < K ? hi0 : K;
    for (uint32_t v = lo + gtid(); v < hi; v += gstride()) {
        const uint2 c = lt_conn[v];
        if (index_ok(c.x, R, cnt) && index_ok(c.y, R, cnt)) atomicMax(&b.st_a[c.y], b.st_a[c.x] + 1u);
    }
}

// the largest state: per-block maxima (part, gridDim.x entries), combined by k_ca_layers_done
__global__ void k_ca_state_max(CABufs b, uint32_t* __restrict__ part, DevCounts* __restrict__ cnt) {
    __shared__ uint32_t s_w[1024 / kWave];
    if (block_run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    uint32_t lo, hi;
    ca_block_range(R, lo, hi);
    uint32_t mx = 0;
    for (uint32_t t = lo + threadIdx.x; t < hi; t += blockDim.x) mx = b.st_a[t] > mx ? b.st_a[t] : mx;
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(mx, d, kWave);
        mx = o > mx ? o : mx;
    }
    if (lane_id() == 0) s_w[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / kWave; ++w) mx = s_w[w] > mx ? s_w[w] : mx;
        part[blockIdx.x] = mx;
    }
}

// one block: the iteration count max L + 1 (the largest state) when there are connections
__global__ void k_ca_layers_done(CABufs b, const uint32_t* __restrict__ part, uint32_t n_part,
                                 DevCounts* __restrict__ cnt) {
    __shared__ uint32_t s_w[1024 / kWave];
    if (block_run_void(cnt)) {  // nothing for the host to iterate
        if (threadIdx.x == 0) b.flags[0] = b.flags[1] = b.flags[2] = 0u;
        return;
    }
    uint32_t mx = 0;
    for (uint32_t i = threadIdx.x; i < n_part; i += blockDim.x) mx = part[i] > mx ? part[i] : mx;
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t o = __shfl_xor(mx, d, kWave);
        mx = o > mx ? o : mx;
    }
    if (lane_id() == 0) s_w[threadIdx.x / kWave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / kWave; ++w) mx = s_w[w] > mx ? s_w[w] : mx;
        const uint32_t K = bounded(cnt->n_conn, cnt->cap_conn, cnt);
        cnt->ca_iterations = K ? mx : 0u;
        b.flags[0] = 0u; b.flags[1] = 0u; b.flags[2] = 0u;  // nothing left for the host to iterate
    }
}

__global__ void k_ca_basis_count(const uint32_t* __restrict__ basis, const uint32_t* __restrict__ foll, CABufs b,
                                 DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t K = bounded(cnt->n_conn, cnt->cap_conn, cnt), R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    for (uint32_t p = gtid(); p < K; p += gstride()) {  // followerBasisCount (.h:152-168)
        const uint32_t f = foll[p], bs = basis[p];
        if (!(index_ok(f, R, cnt) && index_ok(bs, R, cnt))) continue;
        if (b.st_a[bs] + 1u == b.st_a[f]) atomicAdd(&b.fbc[f], 1u);
    }
}

__global__ void k_ca_basis_store(const uint32_t* __restrict__ basis, const uint32_t* __restrict__ foll,
                                 const float* __restrict__ pt, CABufs b, uint32_t n_basis_cap,
                                 DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t K = bounded(cnt->n_conn, cnt->cap_conn, cnt), R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    bool over = false;
    for (uint32_t p = gtid(); p < K; p += gstride()) {  // followerBasisStore (.h:188-205)
        const uint32_t bs = basis[p], f = foll[p];
        if (!(index_ok(f, R, cnt) && index_ok(bs, R, cnt))) continue;
        if (b.st_a[bs] + 1u != b.st_a[f]) continue;
        const uint32_t so = b.fbc[f] + atomicAdd(&b.cursor[f], 1u);
        if (so >= n_basis_cap) { over = true; continue; }
        b.bidx[so] = bs;
        b.bdiff[so] = fabsf(pt[bs] - pt[f]);
    }
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

// followerBestBasisStore (.h:224-263) over a basis list in connection-pair
// (= ascending basis) order: the first entry, replaced by each later entry of
// a strictly smaller pT difference.  That is, with the list in any order: the
// entry of the smallest basis index if its difference is NaN (nothing
// compares below a NaN), else the smallest difference among the non-NaN ones,
// the smallest basis index among equal ones (a NaN never replaces).  One pass
// over the stored list, which stays as stored.
__global__ void k_ca_best(CABufs b, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    for (uint32_t t = gtid(); t < R; t += gstride()) {
        const uint32_t begin = b.fbc[t], end = b.fbc[t + 1];
        if (begin == end) continue;
        uint32_t first_idx = kSentinel, best_idx = kSentinel;
        float first_d = 0.0f, best_d = 0.0f;
        bool have = false;
        for (uint32_t i = begin; i < end; ++i) {
            const uint32_t k = b.bidx[i];
            const float d = b.bdiff[i];
            if (k < first_idx || i == begin) {
                first_idx = k;
                first_d = d;
            }
            if (!tx_isnan(d) && (!have || d < best_d || (d == best_d && k < best_idx))) {
                best_d = d;
                best_idx = k;
                have = true;
            }
        }
        const uint32_t minIdx = (tx_isnan(first_d) || !have) ? first_idx : best_idx;
        b.best[t] = minIdx;
        if (minIdx < R) b.isbest[minIdx] = 1u;
    }
}

__global__ void k_ca_handler(CABufs b, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt), K = bounded(cnt->n_conn, cnt->cap_conn, cnt);
    for (uint32_t t = gtid(); t <= R; t += gstride())  // handlersStateStore (.h:280-288)
        b.handler[t] = (t < R && K > 0) ? (b.isbest[t] == 0u) * b.st_a[t] : 0u;
}

__global__ void k_ca_tracks(CABufs b, uint32_t n_track_cap, DevCounts* __restrict__ cnt) {
    if (run_void(cnt)) return;
    const uint32_t R = bounded(cnt->n_tracklets, cnt->cap_tracklets, cnt);
    bool over = false;
    for (uint32_t t = gtid(); t < R; t += gstride()) {  // trackCollectionStore (.h:305-323)
        const uint32_t so = b.handler[t];
        const uint32_t len = b.handler[t + 1] - so;
        if (len == 0) continue;
        if (so + len > n_track_cap) { over = true; continue; }
        uint32_t bi = t;
        for (uint32_t i = 0; i < len; ++i) {
            b.tracks[so + i] = bi;
            if (bi >= R) break;
            bi = b.best[bi];
        }
    }
    if (over) atomicOr(&cnt->err, kErrCapacity);
}

}  // namespace

#define TRAX_LAUNCH(kern, work, blk, ...)                                                          \
    do {                                                                                           \
        hipLaunchKernelGGL(kern, dim3(stride_blocks((work), (blk), lc)), dim3(blk), 0, st, __VA_ARGS__); \
        hipError_t e_ = hipGetLastError();                                                         \
        if (e_ != hipSuccess) return e_;                                                           \
    } while (0)

hipError_t launch_filter_count(const TripletCols& tc, uint32_t n_cand_cap, const float* hx, const float* hy,
                               const float* hz, const uint32_t* hlayer, const uint32_t* pairs,
                               const uint32_t* cands, uint64_t* bitmap, uint32_t* word_counts, DevCounts* cnt,
                               const LaunchCfg& lc, hipStream_t st) {
    const uint32_t b = lc.filter_block ? lc.filter_block : 256;
    const size_t lds = (size_t)(b / kWave) * kFilterWaveLds * sizeof(uint32_t);
    const uint64_t words = ((uint64_t)n_cand_cap + 63) / 64;
    hipLaunchKernelGGL(k_filter_count, dim3(stride_blocks((words / kFilterWords + 1) * kWave, b, lc)), dim3(b), lds,
                       st, tc, hx, hy, hz, hlayer, pairs, cands, bitmap, word_counts, cnt);
    return hipGetLastError();
}

hipError_t launch_filter_store(const TripletCols& tc, uint32_t n_cand_cap, uint32_t n_trk_cap,
                               const uint32_t* hlayer, const uint32_t* hevt, const uint32_t* pairs,
                               const uint32_t* cands, const uint64_t* bitmap, const uint32_t* word_counts,
                               uint32_t* th1, uint32_t* th2, uint32_t* th3, uint32_t* t_off, DevCounts* cnt,
                               const LaunchCfg& lc, hipStream_t st) {
    const uint32_t b = lc.filter_block ? lc.filter_block : 256;
    TRAX_LAUNCH(k_filter_store, (uint64_t)n_cand_cap + 64, b, pairs, cands, bitmap, word_counts, th1, th2, th3,
                n_trk_cap, cnt);
    return hipSuccess;
}

hipError_t launch_filter_offset_store(const TripletCols& tc, uint32_t n_trk_cap, const uint32_t* hlayer,
                                      const uint32_t* hevt, const uint32_t* th1, uint32_t* t_off, DevCounts* cnt,
                                      hipStream_t st) {
    LaunchCfg lc;
    TRAX_LAUNCH(k_filter_offsets, n_trk_cap, 256u, tc, hlayer, hevt, th1, t_off, cnt);
    return hipSuccess;
}

hipError_t launch_filter_offsets(const TripletCols& tc, uint32_t n_trk_cap, const uint32_t* hlayer,
                                 const uint32_t* hevt, const uint32_t* th1, uint32_t* t_off, DevCounts* cnt,
                                 const LaunchCfg& lc, hipStream_t st) {
    (void)lc;
    const hipError_t e = launch_filter_offset_store(tc, n_trk_cap, hlayer, hevt, th1, t_off, cnt, st);
    if (e != hipSuccess) return e;
    return launch_forward_fill(t_off, cnt, tc.n_trip, st);
}

hipError_t launch_link_prepare(const TripletCols& tc, uint32_t n_hits_cap, uint32_t n_trk_cap,
                               const float* hx, const float* hy, const float* hz, const uint32_t* th1,
                               const uint32_t* th2, const uint32_t* th3, const LinkBufs& b, DevCounts* cnt,
                               const ScanScratch& ss, const LaunchCfg& lc, hipStream_t st) {
    (void)tc;
    const uint32_t blk = lc.link_block ? lc.link_block : 256;
    if (b.runs) {  // runs per middle hit: no CSR (empty runs for hits without tracklets)
        hipError_t e = launch_fill_dn(b.hit_count, 0u, &cnt->n_hits, 0u, n_hits_cap, st);
        if (e != hipSuccess) return e;
        e = launch_fill_dn(b.hit_cursor, 0u, &cnt->n_hits, 0u, n_hits_cap, st);
        if (e != hipSuccess) return e;
        TRAX_LAUNCH(k_eta_runs, n_trk_cap, blk, hx, hy, hz, th1, th2, th3, b.eta, b.hit_count, b.hit_cursor, cnt);
        return launch_fill_dn(b.conn, 0u, &cnt->n_tracklets, 1u, (uint64_t)n_trk_cap + 1, st);
    }
    // fills cover the sizes on the device, not the capacities
    hipError_t e = launch_fill_dn(b.hit_count, 0u, &cnt->n_hits, 1u, (uint64_t)n_hits_cap + 1, st);
    if (e != hipSuccess) return e;
    e = launch_fill_dn(b.hit_cursor, 0u, &cnt->n_hits, 0u, n_hits_cap, st);
    if (e != hipSuccess) return e;
    TRAX_LAUNCH(k_eta, n_trk_cap, blk, hx, hy, hz, th1, th3, b.eta, b.hit_count, cnt);
    e = launch_scan(b.hit_count, &cnt->n_hits, 1u, (uint64_t)n_hits_cap + 1, nullptr, ss, st);
    if (e != hipSuccess) return e;
    TRAX_LAUNCH(k_csr_fill, n_trk_cap, blk, th1, b.hit_count, b.hit_cursor, b.csr, cnt);
    return launch_fill_dn(b.conn, 0u, &cnt->n_tracklets, 1u, (uint64_t)n_trk_cap + 1, st);
}

hipError_t launch_link_count(const TripletCols& tc, uint32_t n_trk_cap, const uint32_t* hlayer,
                             const uint32_t* hevt, const uint32_t* th1, const uint32_t* th2, const uint32_t* th3,
                             const uint32_t* t_off, const LinkBufs& b, DevCounts* cnt, const LaunchCfg& lc,
                             hipStream_t st) {
    const uint32_t blk = lc.link_block ? lc.link_block : 256;
    TRAX_LAUNCH(k_link_count, (uint64_t)n_trk_cap + 1, blk, tc, hlayer, hevt, th1, th2, th3, t_off, b, cnt);
    return hipSuccess;
}

uint64_t link_chunks(uint64_t n_trk_cap) { return n_trk_cap / kWave + 2; }
constexpr uint32_t kLinkMaxWaves = 65536;  // waves of the single-pass launch, at most
uint64_t link_scratch_cap(uint64_t n_conn_cap) {
    // a refill drops less than the refilling chunk's own connections, so the
    // dropped parts add up to less than all connections; at the end every
    // wave may leave up to a range unused
    return 2 * n_conn_cap + (uint64_t)kChunkRange * kLinkMaxWaves;
}

hipError_t launch_link_fused(const TripletCols& tc, uint32_t n_trk_cap, const uint32_t* hlayer, const uint32_t* hevt,
                             const uint32_t* th1, const uint32_t* th2, const uint32_t* th3, const uint32_t* t_off,
                             const LinkBufs& b, DevCounts* cnt, const ScanScratch& ss, const LaunchCfg& lc,
                             hipStream_t st) {
    if (!b.chunk_n || !b.chunk_at || !b.scratch || b.scratch_cap == 0) return hipErrorInvalidValue;
    const uint64_t chunks = link_chunks(n_trk_cap);
    hipError_t e = launch_fill(b.chunk_n, 0u, chunks + 1, st);
    if (e != hipSuccess) return e;
    const uint32_t blk = lc.link_block ? lc.link_block : 256;
    uint32_t blocks = stride_blocks(chunks * kWave, blk, lc);
    if ((uint64_t)blocks * (blk / kWave) > kLinkMaxWaves) blocks = kLinkMaxWaves / (blk / kWave);
    hipLaunchKernelGGL(k_link_fused, dim3(blocks), dim3(blk), (size_t)blk * kLinkFew * sizeof(uint32_t), st, tc, hlayer,
                       hevt, th1, th2, th3, t_off, b, cnt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_scan(b.chunk_n, &cnt->n_link_chunks, 1u, chunks + 1, &cnt->n_conn, ss, st);
}

hipError_t launch_link_fused_place(uint32_t n_trk_cap, uint32_t n_conn_cap, const LinkBufs& b, DevCounts* cnt,
                                   const LaunchCfg& lc, hipStream_t st) {
    return launch_chunk_copy3(&cnt->n_link_chunks, link_chunks(n_trk_cap), b.chunk_n, b.chunk_at, b.scratch,
                              b.scratch_cap, b.basis, b.foll, b.pev, n_conn_cap, cnt, lc, st);
}

hipError_t launch_link_store(const TripletCols& tc, uint32_t n_trk_cap, uint32_t n_conn_cap,
                             const uint32_t* hlayer, const uint32_t* hevt, const uint32_t* th1,
                             const uint32_t* th2, const uint32_t* th3, const uint32_t* t_off, const LinkBufs& b,
                             DevCounts* cnt, const LaunchCfg& lc, hipStream_t st) {
    const uint32_t blk = lc.link_block ? lc.link_block : 256;
    TRAX_LAUNCH(k_link_store, n_trk_cap, blk, tc, hlayer, hevt, th1, th2, th3, t_off, b, n_conn_cap, cnt);
    return hipSuccess;
}

hipError_t launch_link_events(const TripletCols& tc, uint32_t n_events_cap, uint32_t n_trk_cap,
                              const uint32_t* hlayer, const uint32_t* hevt, const uint32_t* th1,
                              const uint32_t* t_off, const LinkBufs& b, DevCounts* cnt, hipStream_t st) {
    if (!b.per_event) return hipSuccess;
    LaunchCfg lc;
    hipError_t e = launch_fill(b.ev_first, kSentinel, n_events_cap, st);
    if (e != hipSuccess) return e;
    e = launch_fill(b.ev_last, kSentinel, n_events_cap, st);
    if (e != hipSuccess) return e;
    e = launch_fill(b.ev_has, 0u, n_events_cap, st);
    if (e != hipSuccess) return e;
    TRAX_LAUNCH(k_event_bounds, n_trk_cap, 256u, hevt, th1, b, cnt);
    TRAX_LAUNCH(k_event_offsets, (uint64_t)n_events_cap * (tc.n_trip + 1), 256u, tc, hlayer, th1, t_off, b, cnt);
    return hipSuccess;
}

hipError_t launch_link_flags(uint32_t n_trk_cap, const uint32_t* hevt, const uint32_t* th1, const LinkBufs& b,
                             DevCounts* cnt, const LaunchCfg& lc, hipStream_t st) {
    TRAX_LAUNCH(k_link_flags, (uint64_t)n_trk_cap + 1, 256u, hevt, th1, b, cnt);
    return hipSuccess;
}

hipError_t launch_link_compact(uint32_t n_trk_cap, const LinkBufs& b, DevCounts* cnt, const LaunchCfg& lc,
                               hipStream_t st) {
    TRAX_LAUNCH(k_link_compact, n_trk_cap, 256u, b, cnt);
    return hipSuccess;
}

hipError_t launch_circle_fit(uint32_t n_trk_cap, const float* hx, const float* hy, const uint32_t* th1,
                             const uint32_t* th2, const uint32_t* th3, const LinkBufs& b, DevCounts* cnt,
                             const LaunchCfg& lc, hipStream_t st) {
    hipError_t e = launch_fill_dn(reinterpret_cast<uint32_t*>(b.pt), 0u, &cnt->n_tracklets, 0u, n_trk_cap, st);
    if (e != hipSuccess) return e;
    TRAX_LAUNCH(k_circle, n_trk_cap, 256u, hx, hy, th1, th2, th3, b, cnt);
    return hipSuccess;
}

hipError_t launch_ca_init(uint32_t n_trk_cap, const CABufs& b, DevCounts* cnt, hipStream_t st) {
    LaunchCfg lc;
    TRAX_LAUNCH(k_ca_init, n_trk_cap, 256u, b, cnt);
    return hipSuccess;
}

hipError_t launch_ca_iterations(uint32_t n_conn_cap, const uint32_t* basis, const uint32_t* foll,
                                const CABufs& b, DevCounts* cnt, uint32_t first_iter, uint32_t n_iter,
                                const LaunchCfg& lc, hipStream_t st) {
    for (uint32_t k = first_iter; k < first_iter + n_iter; ++k) TRAX_LAUNCH(k_ca_iter, n_conn_cap, 256u, basis, foll, b, k, cnt);
    return hipSuccess;
}

hipError_t launch_ca_layers(uint32_t n_conn_cap, uint32_t n_trip, const uint32_t* basis, const uint32_t* foll,
                            const uint32_t* th1, const uint32_t* hlayer, const CABufs& b, DevCounts* cnt,
                            const ScanScratch& ss, const LaunchCfg& lc, hipStream_t st) {
    if (!b.clt || !b.lt_conn || !b.lt_tab || n_trip == 0 || n_trip > 256) return hipErrorInvalidValue;
    const uint32_t nb = kCaTableBlocks;
    hipLaunchKernelGGL(k_ca_conn_layer, dim3(nb), dim3(256), 0, st, basis, th1, hlayer, n_trip, b.clt, b.lt_tab, cnt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // bucket positions: the table scanned in (layer triplet, block) order; its total after the last entry
    e = launch_scan(b.lt_tab, nullptr, n_trip * nb, (uint64_t)n_trip * nb, b.lt_tab + (size_t)n_trip * nb, ss, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_ca_conn_scatter, dim3(nb), dim3(256), 0, st, basis, foll, b.clt, n_trip, b.lt_tab, b.lt_conn,
                       cnt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (uint32_t lt = 0; lt < n_trip; ++lt)
        TRAX_LAUNCH(k_ca_layer, n_conn_cap, 256u, b.lt_conn, b.lt_tab, nb, n_trip, lt, b, cnt);
    // the table's first nb entries are free now: the per-block maxima
    hipLaunchKernelGGL(k_ca_state_max, dim3(nb), dim3(256), 0, st, b, b.lt_tab, cnt);
    hipLaunchKernelGGL(k_ca_layers_done, dim3(1), dim3(1024), 0, st, b, b.lt_tab, nb, cnt);
    return hipGetLastError();
}

hipError_t launch_ca_basis_count(uint32_t n_trk_cap, uint32_t n_conn_cap, const uint32_t* basis,
                                 const uint32_t* foll, const CABufs& b, DevCounts* cnt, const LaunchCfg& lc,
                                 hipStream_t st) {
    hipError_t e = launch_fill_dn(b.fbc, 0u, &cnt->n_tracklets, 1u, (uint64_t)n_trk_cap + 1, st);
    if (e != hipSuccess) return e;
    e = launch_fill_dn(b.cursor, 0u, &cnt->n_tracklets, 0u, n_trk_cap, st);
    if (e != hipSuccess) return e;
    TRAX_LAUNCH(k_ca_basis_count, n_conn_cap, 256u, basis, foll, b, cnt);
    return hipSuccess;
}

hipError_t launch_ca_basis_store(uint32_t n_trk_cap, uint32_t n_conn_cap, uint32_t n_basis_cap,
                                 const uint32_t* basis, const uint32_t* foll, const float* pt, const CABufs& b,
                                 DevCounts* cnt, const LaunchCfg& lc, hipStream_t st) {
    TRAX_LAUNCH(k_ca_basis_store, n_conn_cap, 256u, basis, foll, pt, b, n_basis_cap, cnt);
    hipError_t e = launch_fill_dn(b.best, kSentinel, &cnt->n_tracklets, 0u, n_trk_cap, st);
    if (e != hipSuccess) return e;
    e = launch_fill_dn(b.isbest, 0u, &cnt->n_tracklets, 0u, n_trk_cap, st);
    if (e != hipSuccess) return e;
    TRAX_LAUNCH(k_ca_best, n_trk_cap, 256u, b, cnt);
    TRAX_LAUNCH(k_ca_handler, (uint64_t)n_trk_cap + 1, 256u, b, cnt);
    return hipSuccess;
}

hipError_t launch_ca_tracks(uint32_t n_trk_cap, uint32_t n_track_cap, const CABufs& b, DevCounts* cnt,
                            const ScanScratch& ss, const LaunchCfg& lc, hipStream_t st, int part) {
    (void)ss;
    (void)part;
    TRAX_LAUNCH(k_ca_tracks, n_trk_cap, 256u, b, n_track_cap, cnt);
    return hipSuccess;
}

}  // namespace trax
