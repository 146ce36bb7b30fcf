// Physics evaluation of a result against the simulation truth, on the host:
//
// - trax_physics_eval: per (event, layer triplet) counts and rates of
//   PhysicsRecord::fillData (src/triplet_builder/PhysicsRecord.cpp:86-255);
// - trax_physics_*: PhysicsRecord / PhysicsRecords whole -- the eta and pT
//   histograms with their per-bin rates, merge and csvDump
//   (PhysicsRecord.h:48-214, PhysicsRecord.cpp:10-427), filled from the
//   findable tracks HitCollection::addEvent returned for every event;
// - trax_track_eval: the track-level evaluation of the CA's chains
//   (TripletBuilder.cpp:840-1066): extractTracks, cloneRemoval,
//   splitRealFakeTracks, findUniqueRealTracks and printTracks' counts.
//
// A tracklet is valid when its three hits carry the same simulated track id
// (TrackletCollection.h:122-133).  A valid tracklet of a findable track is
// found the first time its track appears in the record, a clone afterwards;
// a valid tracklet of another track counts as misc; an invalid one is a fake.
// A record's tracklets are [off[e*T + p], off[e*T + p + 1]) of the
// reference's tracklet offsets, as the reference reads them.
//
// Arithmetic follows the reference's types: hit coordinates are float and
// the sums the reference forms from them in float stay float; circle and eta
// in double; a tracklet's eta and pT are rounded to float before binning,
// a missed track's are not; the running means and variances are long double,
// the per-bin and per-record rates are formed in double.  Text is written by
// the C++ stream's default formatting (six significant digits).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "capi_internal.h"

using namespace trax;
using namespace trax_capi;

namespace {

// Utils::clamp (utils.h:65-67): at least one, of the int the reference passes
uint32_t clamp1(int64_t n) { return n > 0 ? (uint32_t)n : 1u; }

// tBinnedData (PhysicsRecord.h:48-78)
struct BinnedData {
    uint32_t valid = 0, clones = 0, fake = 0, missed = 0, misc = 0;
    long double eff_mean = 0, eff_var = 0, fr_mean = 0, fr_var = 0, cr_mean = 0, cr_var = 0;
    uint32_t n = 1;

    long double sd(long double m2) const { return n > 1 ? std::sqrt(m2 / (long double)clamp1((int64_t)n - 1)) : 0; }

    // tBinnedData::fill (PhysicsRecord.cpp:34-42): the bin's own rates, in double
    void fill() {
        const uint32_t all = valid + fake + clones + misc;
        eff_mean = (double)valid / (double)clamp1(valid + missed);
        fr_mean = (double)fake / (double)clamp1(all);
        cr_mean = (double)clones / (double)clamp1(all);
        n = all;
    }

    // tBinnedData::operator+= (PhysicsRecord.cpp:10-32): counts summed (misc
    // is not), the rates merged as weighted running means and sums of squared
    // deviations, the efficiency weighted by the truth (valid + missed)
    void add(const BinnedData& c) {
        valid += c.valid;
        clones += c.clones;
        fake += c.fake;
        missed += c.missed;
        n += c.n;
        const long double w_eff = (long double)(c.valid + c.missed);
        long double d = c.eff_mean - eff_mean;
        eff_mean += w_eff * d / (long double)clamp1(valid + missed);
        eff_var += w_eff * d * (c.eff_mean - eff_mean) + c.eff_var;
        d = c.fr_mean - fr_mean;
        fr_mean += (long double)c.n * d / (long double)clamp1(n);
        fr_var += (long double)c.n * d * (c.fr_mean - fr_mean) + c.fr_var;
        d = c.cr_mean - cr_mean;
        cr_mean += (long double)c.n * d / (long double)clamp1(n);
        cr_var += (long double)c.n * d * (c.cr_mean - cr_mean) + c.cr_var;
    }
};

// tHistogram (PhysicsRecord.h:80-100): bins keyed by their lower edge
struct Histogram {
    double width = 0.1;
    std::map<double, BinnedData> bins;
    double bin_of(double v) const { return width * std::floor(v / width); }
    void fill() {
        for (auto& b : bins) b.second.fill();
    }
    // mergeHistograms (PhysicsRecord.cpp:257-268); a bin the target lacks
    // starts as a default bin (n = 1) before the merge, as std::map's []
    void merge(const Histogram& o) {
        for (const auto& b : o.bins) {
            double key = b.first;
            if (width != o.width) key = width * std::floor((b.first * o.width + o.width / 2) / width);
            bins[key].add(b.second);
        }
    }
};

// PhysicsRecord::getEta (PhysicsRecord.h:149-160): the direction inner->outer,
// its components differences of float coordinates
double eta_of(float xi, float yi, float zi, float xo, float yo, float zo) {
    const float fx = xo - xi, fy = yo - yi, fz = zo - zi;
    const double px = fx, py = fy, pz = fz;
    return std::asinh(pz / std::sqrt(px * px + py * py));
}

// PhysicsRecord::getCircleParams (PhysicsRecord.cpp:44-84), radius only: the
// three points lifted onto the paraboloid (x, y, x^2 + y^2), the lift's
// square sum formed in float as the reference's float accessors give it
double circle_radius(const float* x, const float* y, const uint32_t (&h)[3]) {
    double P[3][3];
    for (int k = 0; k < 3; ++k) {
        const float fx = x[h[k]], fy = y[h[k]];
        const float q = fx * fx + fy * fy;
        P[k][0] = fx;
        P[k][1] = fy;
        P[k][2] = q;
    }
    const double ax = P[1][0] - P[0][0], ay = P[1][1] - P[0][1], az = P[1][2] - P[0][2];
    const double bx = P[2][0] - P[0][0], by = P[2][1] - P[0][1], bz = P[2][2] - P[0][2];
    double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
    const double len = std::sqrt(nx * nx + ny * ny + nz * nz);
    nx /= len;
    ny /= len;
    nz /= len;
    const double c = -(nx * P[0][0] + ny * P[0][1] + nz * P[0][2]);
    return std::sqrt((1 - nz * nz - 4 * c * nz) / (4 * nz * nz));
}

// PhysicsRecord::getPt (PhysicsRecord.h:168-172) in GeV/c, with its constants
// (PhysicsRecord.h:135-143): radius in cm to m, times the charge and field
double pt_of_radius(double r) {
    const double BZ = 3.8112, Q = 1.602177E-19, GEV_C = 5.344286E-19;
    return Q * BZ * (r * 1E-2) / GEV_C;
}

std::string fmt_row(std::initializer_list<long double> v) {
    std::ostringstream s;
    bool first = true;
    for (long double x : v) {
        if (!first) s << ' ';
        s << x;
        first = false;
    }
    return s.str();
}

// PhysicsRecord (PhysicsRecord.h:106-192)
struct Record {
    uint32_t event = 0, events_in_group = 0, layer_triplet = 0, total_layers = 0;
    long double eff_mean = 0, eff_var = 0, fr_mean = 0, fr_var = 0, cr_mean = 0, cr_var = 0;
    uint32_t n = 1;
    Histogram pt, eta;

    long double sd(long double m2) const { return n > 1 ? std::sqrt(m2 / (long double)clamp1((int64_t)n - 1)) : 0; }

    // PhysicsRecord::merge (PhysicsRecord.cpp:270-286)
    void merge(const Record& c) {
        eta.merge(c.eta);
        pt.merge(c.pt);
        n += c.n;
        long double d = c.eff_mean - eff_mean;
        eff_mean += (long double)c.n * d / (long double)clamp1(n);
        eff_var += (long double)c.n * d * (c.eff_mean - eff_mean) + c.eff_var;
        d = c.fr_mean - fr_mean;
        fr_mean += (long double)c.n * d / (long double)clamp1(n);
        fr_var += (long double)c.n * d * (c.fr_mean - fr_mean) + c.fr_var;
        d = c.cr_mean - cr_mean;
        cr_mean += (long double)c.n * d / (long double)clamp1(n);
        cr_var += (long double)c.n * d * (c.cr_mean - cr_mean) + c.cr_var;
    }

    // PhysicsRecord::csvDump's row (PhysicsRecord.cpp:288-296)
    std::string row() const {
        std::ostringstream s;
        s << layer_triplet << ' ' << n << ' '
          << fmt_row({eff_mean, sd(eff_var), fr_mean, sd(fr_var), cr_mean, sd(cr_var)});
        return s.str();
    }
};

// tHistogram::csvDump (PhysicsRecord.cpp:412-425); layout 1 is the five
// columns of the published histogram files (plotting/data/physics/ttbar/*.csv)
std::string histogram_csv(const Histogram& h, int layout) {
    std::ostringstream s;
    if (layout == 1) {
        s << "bin valid fake clones missed\n";
        for (const auto& b : h.bins)
            s << b.first << ' ' << b.second.valid << ' ' << b.second.fake << ' ' << b.second.clones << ' '
              << b.second.missed << '\n';
        return s.str();
    }
    s << "bin valid fake clones missed misc efficiencyMean efficiencyVar fakeRateMean fakeRateVar cloneRateMean "
         "cloneRateVar\n";
    for (const auto& b : h.bins) {
        const BinnedData& d = b.second;
        s << b.first << ' ' << d.valid << ' ' << d.fake << ' ' << d.clones << ' ' << d.missed << ' ' << d.misc << ' '
          << fmt_row({d.eff_mean, d.sd(d.eff_var), d.fr_mean, d.sd(d.fr_var), d.cr_mean, d.sd(d.cr_var)}) << '\n';
    }
    return s.str();
}

using Truth = HitCollection::tTrackList;

struct ResultView {
    uint32_t n_hits;
    const float *x, *y, *z;
    const uint32_t* hit_id;
    uint32_t n_tracklets;
    const uint32_t *h1, *h2, *h3, *off;
};

// PhysicsRecord::fillData (PhysicsRecord.cpp:86-255) of event e, layer triplet p
void fill_record(Record& r, const Truth& truth, const ResultView& v, uint32_t T) {
    const uint64_t g = (uint64_t)r.event * T + r.layer_triplet;
    const uint32_t lo = v.off[g], hi = v.off[g + 1];
    if (hi < lo || hi > v.n_tracklets) throw InputError("tracklet offsets out of range");
    std::set<uint32_t> found;
    uint32_t fakes = 0, clones = 0;
    for (uint32_t t = lo; t < hi; ++t) {
        const uint32_t h[3] = {v.h1[t], v.h2[t], v.h3[t]};
        if (h[0] >= v.n_hits || h[1] >= v.n_hits || h[2] >= v.n_hits)
            throw InputError("tracklet hit out of range");
        const float t_eta =
            (float)eta_of(v.x[h[0]], v.y[h[0]], v.z[h[0]], v.x[h[2]], v.y[h[2]], v.z[h[2]]);
        const float t_pt = (float)pt_of_radius(circle_radius(v.x, v.y, h));
        BinnedData& be = r.eta.bins[r.eta.bin_of(t_eta)];
        BinnedData& bp = r.pt.bins[r.pt.bin_of(t_pt)];
        const uint32_t id = v.hit_id[h[0]];
        if (id == v.hit_id[h[1]] && id == v.hit_id[h[2]]) {
            if (truth.count(id)) {
                if (found.insert(id).second) {
                    ++be.valid;
                    ++bp.valid;
                } else {
                    ++clones;
                    ++be.clones;
                    ++bp.clones;
                }
            } else {
                ++be.misc;
                ++bp.misc;
            }
        } else {
            ++fakes;
            ++be.fake;
            ++bp.fake;
        }
    }
    // findable tracks not found: binned by their innermost, middle and
    // outermost truth hits, in double
    for (const auto& tr : truth) {
        if (found.count(tr.first) || tr.second.empty()) continue;
        const std::vector<PHit>& hs = tr.second;
        const PHit &a = hs.front(), &m = hs[hs.size() / 2], &b = hs.back();
        const float xs[3] = {a.x, m.x, b.x}, ys[3] = {a.y, m.y, b.y};
        const uint32_t idx[3] = {0, 1, 2};
        ++r.eta.bins[r.eta.bin_of(eta_of(a.x, a.y, a.z, b.x, b.y, b.z))].missed;
        ++r.pt.bins[r.pt.bin_of(pt_of_radius(circle_radius(xs, ys, idx)))].missed;
    }
    const uint32_t n_found = hi - lo;
    r.eff_mean = (double)found.size() / (double)clamp1((int64_t)truth.size());
    r.fr_mean = (double)fakes / (double)clamp1(n_found);
    r.cr_mean = (double)clones / (double)clamp1(n_found);
    r.eta.fill();
    r.pt.fill();
}

}  // namespace

// PhysicsRecords (PhysicsRecord.h:205-214)
struct trax_physics {
    std::vector<Record> records;

    // PhysicsRecords::addRecord (PhysicsRecord.cpp:349-360): merged into the
    // record of the same layer triplet, else appended
    void add(const Record& r) {
        for (auto& x : records)
            if (x.layer_triplet == r.layer_triplet) {
                x.merge(r);
                return;
            }
        records.push_back(r);
    }

    // the total row of PhysicsRecords::csvDump (PhysicsRecord.cpp:369-392):
    // the first record merged with the others, numbered records + 1
    Record total() const {
        Record t = records.at(0);
        t.layer_triplet = (uint32_t)records.size() + 1;
        for (size_t i = 1; i < records.size(); ++i) t.merge(records[i]);
        return t;
    }
};

extern "C" int trax_physics_create(trax_physics** out) {
    return guarded([&] {
        if (!out) return fail(TRAX_ERR_ARG, "null argument");
        *out = new trax_physics;
        return TRAX_OK;
    });
}

extern "C" void trax_physics_destroy(trax_physics* p) { delete p; }

extern "C" int trax_physics_add(trax_physics* p, const trax_config* cfg, const trax_hits* hits, uint32_t n_hits,
                                const float* x, const float* y, const float* z, const uint32_t* hit_id,
                                uint32_t n_tracklets, const uint32_t* h1, const uint32_t* h2, const uint32_t* h3,
                                const uint32_t* tracklet_offsets) {
    return guarded([&] {
        if (!p || !cfg || !hits || (n_hits && (!x || !y || !z || !hit_id)) || (n_tracklets && (!h1 || !h2 || !h3)) ||
            !tracklet_offsets)
            return fail(TRAX_ERR_ARG, "null argument");
        const uint32_t T = cfg->trip.size(), E = hits->n_events;
        if (hits->truth.size() != E) return fail(TRAX_ERR_INPUT, "collection holds no truth for its events");
        const ResultView v{n_hits, x, y, z, hit_id, n_tracklets, h1, h2, h3, tracklet_offsets};
        // buildTriplets' evaluation loop (TripletBuilder.cpp:515-524): every
        // event, every layer triplet, one record each, added in that order
        trax_physics staged = *p;
        for (uint32_t e = 0; e < E; ++e)
            for (uint32_t lt = 0; lt < T; ++lt) {
                Record r;
                r.event = e;
                r.events_in_group = E;
                r.layer_triplet = lt;
                r.total_layers = T;
                fill_record(r, hits->truth[e], v, T);
                staged.add(r);
            }
        p->records.swap(staged.records);
        return TRAX_OK;
    });
}

extern "C" int trax_physics_merge(trax_physics* into, const trax_physics* from) {
    return guarded([&] {
        if (!into || !from) return fail(TRAX_ERR_ARG, "null argument");
        // PhysicsRecords::merge (PhysicsRecord.cpp:362-367)
        const std::vector<Record> src = from->records;
        for (const auto& r : src) into->add(r);
        return TRAX_OK;
    });
}

extern "C" int trax_physics_count(const trax_physics* p, uint32_t* n) {
    return guarded([&] {
        if (!p || !n) return fail(TRAX_ERR_ARG, "null argument");
        *n = (uint32_t)p->records.size();
        return TRAX_OK;
    });
}

extern "C" int trax_physics_csv(const trax_physics* p, int table, uint32_t layer_triplet, int layout, char* buf,
                                uint64_t size, uint64_t* needed) {
    return guarded([&] {
        if (!p || (!buf && size)) return fail(TRAX_ERR_ARG, "null argument");
        if (p->records.empty()) return fail(TRAX_ERR_INPUT, "no physics records");
        if (table < 0 || table > 2 || layout < 0 || layout > 1) return fail(TRAX_ERR_ARG, "unknown table or layout");
        std::string out;
        if (table == 0) {
            // PhysicsRecords::csvDump (PhysicsRecord.cpp:369-392)
            std::ostringstream s;
            s << "layerTriplet n eff effVar fr frVar cr crVar\n";
            for (const auto& r : p->records) s << r.row() << '\n';
            s << p->total().row() << '\n';
            out = s.str();
        } else {
            const Record* r = nullptr;
            for (const auto& x : p->records)
                if (x.layer_triplet == layer_triplet) {
                    r = &x;
                    break;
                }
            Record t;
            if (!r && layer_triplet == p->records.size() + 1) {
                t = p->total();
                r = &t;
            }
            if (!r) return fail(TRAX_ERR_ARG, "no record of that layer triplet");
            out = histogram_csv(table == 1 ? r->eta : r->pt, layout);
        }
        if (needed) *needed = out.size() + 1;
        if (size) {
            const size_t k = std::min<uint64_t>(out.size(), size - 1);
            std::memcpy(buf, out.data(), k);
            buf[k] = '\0';
            if (out.size() + 1 > size) return fail(TRAX_ERR_CAPACITY, "buffer too small");
        }
        return TRAX_OK;
    });
}

extern "C" int trax_hits_truth(const trax_hits* h, uint32_t event, uint32_t* n_tracks, uint32_t* n_hits,
                               uint32_t* track_id, uint32_t* hit_offsets, float* x, float* y, float* z,
                               uint32_t* layer) {
    return guarded([&] {
        if (!h || !n_tracks || !n_hits) return fail(TRAX_ERR_ARG, "null argument");
        if (event != UINT32_MAX && event >= h->truth.size()) return fail(TRAX_ERR_ARG, "event out of range");
        const uint32_t e0 = event == UINT32_MAX ? 0 : event;
        const uint32_t e1 = event == UINT32_MAX ? (uint32_t)h->truth.size() : event + 1;
        uint64_t nt = 0, nh = 0;
        for (uint32_t e = e0; e < e1; ++e)
            for (const auto& tr : h->truth[e]) {
                ++nt;
                nh += tr.second.size();
            }
        if (nt > UINT32_MAX || nh > UINT32_MAX) return fail(TRAX_ERR_CAPACITY, "truth too large for 32-bit counts");
        const bool fill = track_id || hit_offsets || x || y || z || layer;
        if (fill && (*n_tracks < nt || *n_hits < nh)) return fail(TRAX_ERR_CAPACITY, "output arrays too small");
        *n_tracks = (uint32_t)nt;
        *n_hits = (uint32_t)nh;
        if (!fill) return TRAX_OK;
        uint32_t t = 0, k = 0;
        for (uint32_t e = e0; e < e1; ++e)
            for (const auto& tr : h->truth[e]) {
                if (track_id) track_id[t] = tr.first;
                if (hit_offsets) hit_offsets[t] = k;
                ++t;
                for (const PHit& p : tr.second) {
                    if (x) x[k] = p.x;
                    if (y) y[k] = p.y;
                    if (z) z[k] = p.z;
                    if (layer) layer[k] = p.layer;
                    ++k;
                }
            }
        if (hit_offsets) hit_offsets[t] = k;
        return TRAX_OK;
    });
}

// ---- track-level evaluation ---------------------------------------------------

extern "C" int trax_track_eval(uint32_t n_hits, const uint32_t* hit_id, uint32_t n_tracklets, const uint32_t* h1,
                               const uint32_t* h2, const uint32_t* h3, uint32_t n_chains,
                               const uint32_t* track_offsets, const uint32_t* tracks, uint64_t simulated,
                               trax_track_summary* out, uint8_t* kept) {
    return guarded([&] {
        if (!out || (n_hits && !hit_id) || (n_tracklets && (!h1 || !h2 || !h3)) || !track_offsets ||
            (n_chains && !tracks && track_offsets[n_chains] > 0))
            return fail(TRAX_ERR_ARG, "null argument");
        // extractTracks (TripletBuilder.cpp:955-988): chains of at least three
        // tracklets, stored outermost first, as hit lists innermost first
        std::vector<uint32_t> trk_off{0}, trk_hits, trk_chain;
        for (uint32_t i = 0; i < n_chains; ++i) {
            const uint32_t a = track_offsets[i], b = track_offsets[i + 1];
            if (b < a) return fail(TRAX_ERR_INPUT, "track offsets decrease");
            if (b - a < 3) continue;
            for (uint32_t j = 0; j < b - a; ++j) {
                const uint32_t t = tracks[b - 1 - j];
                if (t >= n_tracklets) return fail(TRAX_ERR_INPUT, "track tracklet out of range");
                if (j == 0) {
                    trk_hits.push_back(h1[t]);
                    trk_hits.push_back(h2[t]);
                }
                trk_hits.push_back(h3[t]);
            }
            trk_off.push_back((uint32_t)trk_hits.size());
            trk_chain.push_back(i);
        }
        for (uint32_t hh : trk_hits)
            if (hh >= n_hits) return fail(TRAX_ERR_INPUT, "track hit out of range");
        const uint32_t R = (uint32_t)trk_chain.size();
        auto len = [&](uint32_t r) { return trk_off[r + 1] - trk_off[r]; };
        // TripletTrackComparator (TripletBuilder.cpp:845-867): the hits of the
        // longer track found in the shorter, as a share of the shorter's
        // length, at least 0.19; the new track is the base, which counts as
        // the longer on equal lengths
        std::vector<uint32_t> sorted_hits(trk_hits);
        for (uint32_t r = 0; r < R; ++r) std::sort(sorted_hits.begin() + trk_off[r], sorted_hits.begin() + trk_off[r + 1]);
        auto similar = [&](uint32_t base, uint32_t other) {
            const uint32_t longer = len(base) < len(other) ? other : base;
            const uint32_t shorter = len(base) >= len(other) ? other : base;
            const uint32_t* s0 = sorted_hits.data() + trk_off[shorter];
            const uint32_t* s1 = sorted_hits.data() + trk_off[shorter + 1];
            uint32_t shared = 0;
            for (uint32_t k = trk_off[longer]; k < trk_off[longer + 1]; ++k)
                if (std::binary_search(s0, s1, trk_hits[k])) ++shared;
            const float f = shared / float(len(shorter));
            return f >= 0.19;
        };
        // cloneRemoval (TripletBuilder.cpp:880-913) over the list in order:
        // the first kept track similar to the new one decides.  Only a track
        // that shares a hit can be similar (0 shared is a share of 0), so the
        // candidates are the kept tracks on the new one's hits, tried in list
        // order (the order they were kept in; a removed one is skipped)
        std::vector<uint32_t> kept_track;  // list position -> track
        std::vector<uint8_t> alive;
        std::unordered_map<uint32_t, std::vector<uint32_t>> on_hit;
        std::vector<uint32_t> cand;
        auto keep = [&](uint32_t r) {
            const uint32_t pos = (uint32_t)kept_track.size();
            kept_track.push_back(r);
            alive.push_back(1);
            for (uint32_t k = trk_off[r]; k < trk_off[r + 1]; ++k) {
                auto& v = on_hit[trk_hits[k]];
                if (v.empty() || v.back() != pos) v.push_back(pos);
            }
        };
        for (uint32_t r = 0; r < R; ++r) {
            cand.clear();
            for (uint32_t k = trk_off[r]; k < trk_off[r + 1]; ++k) {
                auto it = on_hit.find(trk_hits[k]);
                if (it != on_hit.end()) cand.insert(cand.end(), it->second.begin(), it->second.end());
            }
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            int64_t found = -1;
            for (uint32_t pos : cand)
                if (alive[pos] && similar(r, kept_track[pos])) {
                    found = pos;
                    break;
                }
            if (found < 0 || len(kept_track[found]) == len(r)) {
                keep(r);
            } else if (len(kept_track[found]) < len(r)) {
                alive[found] = 0;
                keep(r);
            }
        }
        // splitRealFakeTracks / findUniqueRealTracks (TripletBuilder.cpp:915-1006)
        trax_track_summary s{};
        s.total = R;
        std::set<uint32_t> unique;
        if (kept) std::memset(kept, 0, n_chains);
        for (size_t pos = 0; pos < kept_track.size(); ++pos) {
            if (!alive[pos]) continue;
            const uint32_t r = kept_track[pos];
            ++s.filtered;
            const uint32_t id = hit_id[trk_hits[trk_off[r]]];
            bool real = true;
            for (uint32_t k = trk_off[r]; k < trk_off[r + 1]; ++k) real = real && hit_id[trk_hits[k]] == id;
            if (real) {
                ++s.valids;
                unique.insert(id);
            } else {
                ++s.fakes;
            }
            if (kept) kept[trk_chain[r]] = (uint8_t)(1u | (real ? 2u : 0u));
        }
        s.valid_unique = unique.size();
        s.simulated = simulated;
        // printTracks' rates (TripletBuilder.cpp:1049-1056), in float
        s.fake_rate = (float)s.fakes / float(s.filtered);
        s.clone_rate = (float)(s.valids - s.valid_unique) / float(s.filtered);
        s.efficiency = (float)s.valid_unique / float(s.simulated);
        *out = s;
        return TRAX_OK;
    });
}

extern "C" int trax_physics_eval(const trax_config* cfg, uint32_t n_events, uint32_t n_hits, const uint32_t* hit_id,
                                 const uint32_t* hit_layer, const uint32_t* hit_event, uint32_t n_tracklets,
                                 const uint32_t* h1, const uint32_t* h2, const uint32_t* h3,
                                 const uint32_t* tracklet_offsets, int only_tracks, trax_physics_row* rows) {
    return guarded([&] {
        if (!cfg || !rows || (n_hits && (!hit_id || !hit_layer || !hit_event)) ||
            (n_tracklets && (!h1 || !h2 || !h3)) || !tracklet_offsets)
            return fail(TRAX_ERR_ARG, "null argument");
        const uint32_t T = cfg->trip.size();
        std::set<uint32_t> used;
        for (uint32_t i = 0; i < T; ++i) {
            used.insert(cfg->trip.layer1[i]);
            used.insert(cfg->trip.layer2[i]);
            used.insert(cfg->trip.layer3[i]);
        }
        const uint32_t n_used = (uint32_t)used.size();
        // findable tracks of every event: ids whose hits cover every used layer
        // (the hits of the collection are those HitCollection::addEvent keeps:
        // a track missing a used layer never covers them all)
        std::vector<std::unordered_map<uint32_t, std::set<uint32_t>>> layers_of(n_events);
        for (uint32_t i = 0; i < n_hits; ++i) {
            if (hit_event[i] >= n_events) return fail(TRAX_ERR_INPUT, "hit event out of range");
            if (used.count(hit_layer[i])) layers_of[hit_event[i]][hit_id[i]].insert(hit_layer[i]);
        }
        std::vector<std::set<uint32_t>> findable(n_events);
        for (uint32_t e = 0; e < n_events; ++e)
            for (const auto& kv : layers_of[e])
                if (kv.second.size() == n_used && !(only_tracks && kv.first == 0)) findable[e].insert(kv.first);
        const uint64_t n_off = (uint64_t)n_events * T + 1;
        for (uint32_t e = 0; e < n_events; ++e)
            for (uint32_t p = 0; p < T; ++p) {
                trax_physics_row& r = rows[(size_t)e * T + p];
                std::memset(&r, 0, sizeof(r));
                r.event = e;
                r.layer_triplet = p;
                const uint64_t g = (uint64_t)e * T + p;
                const uint32_t lo = tracklet_offsets[g], hi = g + 1 < n_off ? tracklet_offsets[g + 1] : lo;
                std::set<uint32_t> found;
                for (uint32_t t = lo; t < hi && t < n_tracklets; ++t) {
                    if (h1[t] >= n_hits || h2[t] >= n_hits || h3[t] >= n_hits)
                        return fail(TRAX_ERR_INPUT, "tracklet hit out of range");
                    const uint32_t a = hit_id[h1[t]], b = hit_id[h2[t]], c = hit_id[h3[t]];
                    if (a == b && a == c) {
                        if (findable[e].count(a)) {
                            if (found.insert(a).second) ++r.found;
                            else ++r.clones;
                        } else {
                            ++r.misc;
                        }
                    } else {
                        ++r.fakes;
                    }
                }
                r.n_tracklets = hi > lo ? hi - lo : 0;
                r.n_findable = (uint32_t)findable[e].size();
                r.efficiency = (double)r.found / (double)clamp1(r.n_findable);
                r.fake_rate = (double)r.fakes / (double)clamp1(r.n_tracklets);
                r.clone_rate = (double)r.clones / (double)clamp1(r.n_tracklets);
            }
        return TRAX_OK;
    });
}
