// ORACLE -- test infrastructure only: the checker of the library's physics
// evaluation, never part of the product.  Loaded by tests/ through
// oracle/__init__.py.
//
// CPU restatement, in the reference's own structure and types, of
//   - PhysicsRecord / PhysicsRecords (src/triplet_builder/PhysicsRecord.h:
//     48-214, PhysicsRecord.cpp:10-427): per record (event, layer triplet)
//     the tracklets classified against the findable tracks, eta and pT
//     histograms of 0.1, the record and bin rates, merge and csvDump;
//   - the track-level evaluation (TripletBuilder.cpp:840-1066):
//     extractTracks, cloneRemoval as a std::list scanned by find_if,
//     splitRealFakeTracks, findUniqueRealTracks and printTracks' counts.
// The findable tracks come from oracle/__init__.py's restatement of
// HitCollection::addEvent's selection.
//
// Mutation hooks (-DTRAX_ORACLE_PLANT=k, built by oracle.build(k)):
//   20: a tracklet's eta from its first to its second hit (PhysicsRecord.cpp:106)
//   21: the spread of a record over n instead of n - 1 (PhysicsRecord.h:125-128)
//   22: cloneRemoval keeps the shorter of two similar tracks (TripletBuilder.cpp:892)
//   23: a bin's fake rate without misc in its denominator (PhysicsRecord.cpp:37)
#ifndef TRAX_ORACLE_PLANT
#define TRAX_ORACLE_PLANT 0
#endif

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace {

typedef unsigned int uint;

struct fVector2 {
    double x, y;
};
struct fVector3 {
    double x, y, z;
};
struct HitXYZ {  // the reference's Hit / PHitWrapper: float coordinates
    float x, y, z;
};

typedef std::map<uint, std::vector<HitXYZ>> tTrackList;

uint clamp(int n) { return n > 0 ? n : 1; }

std::string csv(std::initializer_list<long double> a) {
    std::stringstream s;
    for (auto it = a.begin(); it != a.end(); ++it) {
        s << *it;
        if (it + 1 != a.end()) s << ' ';
    }
    return s.str();
}
std::string csv(std::initializer_list<uint> a) {
    std::stringstream s;
    for (auto it = a.begin(); it != a.end(); ++it) {
        s << *it;
        if (it + 1 != a.end()) s << ' ';
    }
    return s.str();
}

struct tBinnedData {
    uint valid = 0, clones = 0, fake = 0, missed = 0, misc = 0;
    long double efficiencyMean = 0, efficiencyVar = 0, fakeRateMean = 0, fakeRateVar = 0, cloneRateMean = 0,
                cloneRateVar = 0;
    uint n = 1;

    long double toVar(long double x) const { return n > 1 ? std::sqrt(x / clamp(n - 1)) : 0; }

    void fill() {
        efficiencyMean = ((double)valid) / clamp(valid + missed);
        fakeRateMean = (TRAX_ORACLE_PLANT == 23) ? ((double)fake / clamp(valid + fake + clones))
                                                  : ((double)fake / clamp(valid + fake + clones + misc));
        cloneRateMean = ((double)clones / clamp(valid + fake + clones + misc));
        n = valid + fake + clones + misc;
    }

    void operator+=(const tBinnedData c) {
        valid += c.valid;
        clones += c.clones;
        fake += c.fake;
        missed += c.missed;
        n += c.n;
        long double delta = c.efficiencyMean - efficiencyMean;
        efficiencyMean += (c.valid + c.missed) * delta / clamp((valid + missed));
        efficiencyVar += (c.valid + c.missed) * delta * (c.efficiencyMean - efficiencyMean) + c.efficiencyVar;
        delta = c.fakeRateMean - fakeRateMean;
        fakeRateMean += c.n * delta / clamp(n);
        fakeRateVar += c.n * delta * (c.fakeRateMean - fakeRateMean) + c.fakeRateVar;
        delta = c.cloneRateMean - cloneRateMean;
        cloneRateMean += c.n * delta / clamp(n);
        cloneRateVar += c.n * delta * (c.cloneRateMean - cloneRateMean) + c.cloneRateVar;
    }

    std::string csvDump(int layout) const {
        std::stringstream s;
        if (layout == 1) {
            s << csv({valid, fake, clones, missed});
            return s.str();
        }
        s << csv({valid, fake, clones, missed, misc}) << ' '
          << csv({efficiencyMean, toVar(efficiencyVar), fakeRateMean, toVar(fakeRateVar), cloneRateMean,
                  toVar(cloneRateVar)});
        return s.str();
    }
};

struct tHistogram : public std::map<double, tBinnedData> {
    explicit tHistogram(double w) : binWidth(w) {}
    double binWidth;
    void fill() {
        for (auto& t : *this) t.second.fill();
    }
    std::string csvDump(int layout) const {
        std::stringstream s;
        s << (layout == 1 ? "bin valid fake clones missed"
                          : "bin valid fake clones missed misc efficiencyMean efficiencyVar fakeRateMean "
                            "fakeRateVar cloneRateMean cloneRateVar")
          << std::endl;
        for (auto it = begin(); it != end(); ++it) s << it->first << ' ' << it->second.csvDump(layout) << std::endl;
        return s.str();
    }
};

void mergeHistograms(tHistogram& t1, const tHistogram& t2) {
    for (auto t : t2) {
        if (t1.binWidth == t2.binWidth) {
            t1[t.first] += t.second;
        } else {
            double bin = t.first * t2.binWidth + t2.binWidth / 2;
            bin = t1.binWidth * floor(bin / t1.binWidth);
            t1[bin] += t.second;
        }
    }
}

struct Result {
    uint n_hits;
    const float *x, *y, *z;
    const uint* hit_id;
    uint n_tracklets;
    const uint *h1, *h2, *h3, *off;
    HitXYZ hit(uint i) const { return HitXYZ{x[i], y[i], z[i]}; }
};

class PhysicsRecord {
public:
    PhysicsRecord(uint e, uint total, uint lt, uint layers)
        : event(e), eventsInGroup(total), layerTriplet(lt), totalLayers(layers), pt(0.1), eta(0.1) {}

    uint event, eventsInGroup, layerTriplet, totalLayers;
    long double efficiencyMean = 0, efficiencyVar = 0, fakeRateMean = 0, fakeRateVar = 0, cloneRateMean = 0,
                cloneRateVar = 0;
    uint n = 1;
    tHistogram pt, eta;

    const double BZ = 3.8112;
    const double Q = 1.602177E-19;
    const double C = 2.998E8;
    const double GEV_C = 5.344286E-19;

    long double toVar(long double x) const {
        return n > 1 ? std::sqrt(x / clamp((TRAX_ORACLE_PLANT == 21) ? n : n - 1)) : 0;
    }

    double getEta(const HitXYZ& innerHit, const HitXYZ& outerHit) const {
        fVector3 p{outerHit.x - innerHit.x, outerHit.y - innerHit.y, outerHit.z - innerHit.z};
        double t(p.z / std::sqrt(p.x * p.x + p.y * p.y));
        return asinh(t);
    }
    double getEtaBin(double e) const { return eta.binWidth * floor(e / eta.binWidth); }
    double getPtBin(double p) const { return pt.binWidth * floor(p / pt.binWidth); }

    double getPt(const HitXYZ& p1, const HitXYZ& p2, const HitXYZ& p3) const {
        fVector3 pP1{p1.x, p1.y, p1.x * p1.x + p1.y * p1.y};
        fVector3 pP2{p2.x, p2.y, p2.x * p2.x + p2.y * p2.y};
        fVector3 pP3{p3.x, p3.y, p3.x * p3.x + p3.y * p3.y};
        fVector3 a{pP2.x - pP1.x, pP2.y - pP1.y, pP2.z - pP1.z};
        fVector3 b{pP3.x - pP1.x, pP3.y - pP1.y, pP3.z - pP1.z};
        fVector3 nn{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
        double value = sqrt(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
        nn.x /= value;
        nn.y /= value;
        nn.z /= value;
        double c = -(nn.x * pP1.x + nn.y * pP1.y + nn.z * pP1.z);
        double radius = sqrt((1 - nn.z * nn.z - 4 * c * nn.z) / (4 * nn.z * nn.z));
        return Q * BZ * (radius * 1E-2) / GEV_C;
    }

    void fillData(const Result& r, const tTrackList& mcTruth, uint nLayerTriplets) {
        std::set<uint> foundTracks;
        uint fakeTracks = 0, foundClones = 0;
        const uint lo = r.off[event * nLayerTriplets + layerTriplet];
        const uint hi = r.off[event * nLayerTriplets + layerTriplet + 1];
        uint nFoundTracklets = hi - lo;
        for (uint i = lo; i < hi; ++i) {
            const uint a = r.h1[i], b = r.h2[i], c = r.h3[i];
            float tEta = getEta(r.hit(a), (TRAX_ORACLE_PLANT == 20) ? r.hit(b) : r.hit(c));
            float tPt = getPt(r.hit(a), r.hit(b), r.hit(c));
            const bool valid = r.hit_id[a] == r.hit_id[b] && r.hit_id[a] == r.hit_id[c];
            if (valid) {
                const uint id = r.hit_id[a];
                if (mcTruth.find(id) != mcTruth.end()) {
                    bool inserted = foundTracks.insert(id).second;
                    if (inserted) {
                        eta[getEtaBin(tEta)].valid++;
                        pt[getPtBin(tPt)].valid++;
                    } else {
                        ++foundClones;
                        eta[getEtaBin(tEta)].clones++;
                        pt[getPtBin(tPt)].clones++;
                    }
                } else {
                    eta[getEtaBin(tEta)].misc++;
                    pt[getPtBin(tPt)].misc++;
                }
            } else {
                ++fakeTracks;
                eta[getEtaBin(tEta)].fake++;
                pt[getPtBin(tPt)].fake++;
            }
        }
        for (auto vTrack : mcTruth) {
            if (foundTracks.find(vTrack.first) == foundTracks.end()) {
                HitXYZ innerHit(vTrack.second[0]);
                HitXYZ middleHit(vTrack.second[floor(vTrack.second.size() / 2)]);
                HitXYZ outerHit(vTrack.second[vTrack.second.size() - 1]);
                eta[getEtaBin(getEta(innerHit, outerHit))].missed++;
                pt[getPtBin(getPt(innerHit, middleHit, outerHit))].missed++;
            }
        }
        efficiencyMean = ((double)foundTracks.size()) / clamp(mcTruth.size());
        fakeRateMean = ((double)fakeTracks) / clamp(nFoundTracklets);
        cloneRateMean = ((double)foundClones) / clamp(nFoundTracklets);
        eta.fill();
        pt.fill();
    }

    void merge(const PhysicsRecord& c) {
        mergeHistograms(eta, c.eta);
        mergeHistograms(pt, c.pt);
        n += c.n;
        long double delta = c.efficiencyMean - efficiencyMean;
        efficiencyMean += c.n * delta / clamp(n);
        efficiencyVar += c.n * delta * (c.efficiencyMean - efficiencyMean) + c.efficiencyVar;
        delta = c.fakeRateMean - fakeRateMean;
        fakeRateMean += c.n * delta / clamp(n);
        fakeRateVar += c.n * delta * (c.fakeRateMean - fakeRateMean) + c.fakeRateVar;
        delta = c.cloneRateMean - cloneRateMean;
        cloneRateMean += c.n * delta / clamp(n);
        cloneRateVar += c.n * delta * (c.cloneRateMean - cloneRateMean) + c.cloneRateVar;
    }

    std::string csvDump() const {
        std::stringstream s;
        s << csv({layerTriplet, n}) << ' ' << csv({efficiencyMean, toVar(efficiencyVar)}) << ' '
          << csv({fakeRateMean, toVar(fakeRateVar)}) << ' ' << csv({cloneRateMean, toVar(cloneRateVar)});
        return s.str();
    }
};

struct PhysicsRecords {
    std::vector<PhysicsRecord> records;
    void addRecord(const PhysicsRecord& r) {
        for (uint i = 0; i < records.size(); ++i)
            if (records[i].layerTriplet == r.layerTriplet) {
                records[i].merge(r);
                return;
            }
        records.push_back(r);
    }
    PhysicsRecord total() const {
        PhysicsRecord t(records[0]);
        t.layerTriplet = records.size() + 1;
        for (uint i = 1; i < records.size(); ++i) t.merge(records[i]);
        return t;
    }
    std::string csvDump() const {
        std::stringstream s;
        s << "layerTriplet n eff effVar fr frVar cr crVar" << std::endl;
        for (uint i = 0; i < records.size(); ++i) s << records[i].csvDump() << std::endl;
        s << total().csvDump() << std::endl;
        return s.str();
    }
};

struct PhysicsHandle {
    PhysicsRecords recs;
    std::vector<PhysicsRecord> fills;  // every record as filled, before merging
    std::string text;
};

// ---- track level -------------------------------------------------------------

typedef std::pair<uint, std::shared_ptr<std::vector<uint>>> Track;
typedef std::list<Track> TrackCollection;

bool similar(const Track& baseTrack, const Track& other) {
    const Track& longer = baseTrack.second->size() < other.second->size() ? other : baseTrack;
    const Track& shorter = baseTrack.second->size() >= other.second->size() ? other : baseTrack;
    uint sharedHits = 0;
    for (auto hit : *(longer.second))
        if (std::find(shorter.second->begin(), shorter.second->end(), hit) != shorter.second->end()) sharedHits++;
    const float fshared = sharedHits / float(shorter.second->size());
    return fshared >= 0.19;
}

bool trackIsValid(const Track& t, const uint* hit_id) {
    const uint id = hit_id[t.second->front()];
    for (auto h : *(t.second))
        if (hit_id[h] != id) return false;
    return true;
}

TrackCollection cloneRemoval(const TrackCollection& in) {
    TrackCollection out;
    for (auto& track : in) {
        auto found = std::find_if(out.begin(), out.end(), [&](const Track& o) { return similar(track, o); });
        if (found == out.end()) {
            out.push_back(track);
        } else if (found->second->size() == track.second->size()) {
            out.push_back(track);
        } else if ((TRAX_ORACLE_PLANT == 22) ? found->second->size() > track.second->size()
                                              : found->second->size() < track.second->size()) {
            out.erase(found);
            out.push_back(track);
        }
    }
    return out;
}

}  // namespace

extern "C" {

void* oracle_physics(uint n_events, uint n_lt, const uint* truth_event_off, const uint* truth_tid,
                     const uint* truth_hit_off, const float* tx, const float* ty, const float* tz, uint n_hits,
                     const float* x, const float* y, const float* z, const uint* hit_id, uint n_tracklets,
                     const uint* h1, const uint* h2, const uint* h3, const uint* off) {
    PhysicsHandle* ph = new PhysicsHandle;
    const Result r{n_hits, x, y, z, hit_id, n_tracklets, h1, h2, h3, off};
    for (uint e = 0; e < n_events; ++e) {
        tTrackList truth;
        for (uint t = truth_event_off[e]; t < truth_event_off[e + 1]; ++t) {
            std::vector<HitXYZ>& v = truth[truth_tid[t]];
            for (uint k = truth_hit_off[t]; k < truth_hit_off[t + 1]; ++k) v.push_back(HitXYZ{tx[k], ty[k], tz[k]});
        }
        for (uint p = 0; p < n_lt; ++p) {
            PhysicsRecord rec(e, n_events, p, n_lt);
            rec.fillData(r, truth, n_lt);
            ph->fills.push_back(rec);
            ph->recs.addRecord(rec);
        }
    }
    return ph;
}

// table 0: PhysicsRecords::csvDump; 1 / 2: the eta / pT histogram of the
// record of layer triplet lt (records + 1: the total), layout 0 or 1
const char* oracle_physics_text(void* h, int table, uint lt, int layout) {
    PhysicsHandle* ph = static_cast<PhysicsHandle*>(h);
    if (ph->recs.records.empty()) return nullptr;
    if (table == 0) {
        ph->text = ph->recs.csvDump();
        return ph->text.c_str();
    }
    for (auto& rec : ph->recs.records)
        if (rec.layerTriplet == lt) {
            ph->text = (table == 1 ? rec.eta : rec.pt).csvDump(layout);
            return ph->text.c_str();
        }
    if (lt == ph->recs.records.size() + 1) {
        PhysicsRecord t = ph->recs.total();
        ph->text = (table == 1 ? t.eta : t.pt).csvDump(layout);
        return ph->text.c_str();
    }
    return nullptr;
}

// the record filled i-th (event-major): event, layer triplet, then its
// efficiency, fake and clone rates
int oracle_physics_fill(void* h, uint i, uint* ids, double* rates) {
    PhysicsHandle* ph = static_cast<PhysicsHandle*>(h);
    if (i >= ph->fills.size()) return -1;
    const PhysicsRecord& r = ph->fills[i];
    ids[0] = r.event;
    ids[1] = r.layerTriplet;
    rates[0] = (double)r.efficiencyMean;
    rates[1] = (double)r.fakeRateMean;
    rates[2] = (double)r.cloneRateMean;
    return 0;
}

// the eta (1) or pT (2) histogram of the record filled i-th, before merging
const char* oracle_physics_fill_text(void* h, uint i, int table, int layout) {
    PhysicsHandle* ph = static_cast<PhysicsHandle*>(h);
    if (i >= ph->fills.size()) return nullptr;
    ph->text = (table == 1 ? ph->fills[i].eta : ph->fills[i].pt).csvDump(layout);
    return ph->text.c_str();
}

void oracle_physics_free(void* h) { delete static_cast<PhysicsHandle*>(h); }

// printTracks' counts: total, filtered, valids, valid unique, fakes,
// simulated; the fake, clone and efficiency rates; kept[i] per chain as the
// library reports it (bit 0 kept, bit 1 real)
void oracle_tracks(uint n_hits, const uint* hit_id, const uint* h1, const uint* h2, const uint* h3, uint n_chains,
                   const uint* vPSum, const uint* vTrackCollection, unsigned long long simulated,
                   unsigned long long* counts, float* rates, unsigned char* kept) {
    (void)n_hits;
    TrackCollection reconstructed;
    uint trackID = 0;
    std::vector<uint> chain_of;
    for (uint i = 0; i < n_chains; i++) {
        uint trackOffset = vPSum[i];
        uint trackLength = vPSum[i + 1] - trackOffset;
        if (trackLength < 3) continue;
        std::shared_ptr<std::vector<uint>> track(new std::vector<uint>());
        for (uint j = 0; j < trackLength; j++) {
            uint triplet = vTrackCollection[trackOffset + (trackLength - 1 - j)];
            if (track->size() == 0) {
                track->push_back(h1[triplet]);
                track->push_back(h2[triplet]);
            }
            track->push_back(h3[triplet]);
        }
        reconstructed.push_back(std::make_pair(trackID++, track));
        chain_of.push_back(i);
    }
    TrackCollection filtered = cloneRemoval(reconstructed);
    TrackCollection real, fake;
    for (auto& t : filtered) (trackIsValid(t, hit_id) ? real : fake).push_back(t);
    std::map<uint, Track> unique;
    for (auto& t : real) unique[hit_id[t.second->front()]] = t;
    counts[0] = reconstructed.size();
    counts[1] = filtered.size();
    counts[2] = real.size();
    counts[3] = unique.size();
    counts[4] = fake.size();
    counts[5] = simulated;
    rates[0] = fake.size() / float(filtered.size());
    rates[1] = (real.size() - unique.size()) / float(filtered.size());
    rates[2] = unique.size() / float(simulated);
    if (kept) {
        std::memset(kept, 0, n_chains);
        for (auto& t : filtered) kept[chain_of[t.first]] = (unsigned char)(1u | (trackIsValid(t, hit_id) ? 2u : 0u));
    }
}

}  // extern "C"
