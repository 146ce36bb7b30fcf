// Physics evaluation of tracklets against the simulation truth, per event
// and layer triplet: PhysicsRecord::fillData
// (src/triplet_builder/PhysicsRecord.cpp:86-255) on plain result arrays.
//
// A tracklet is valid when its three hits carry the same simulated track id
// (TrackletCollection.h:122-133).  A valid tracklet of a findable track is
// found the first time its track appears in the group, a clone afterwards;
// a valid tracklet of another track counts as misc; an invalid one is a fake.
// Findable tracks (the mcTruth of the reference: HitCollection::addEvent's
// findableTracks, HitCollection.cpp:56-96) are the track ids whose hits cover
// every layer the layer triplets use (id 0, noise, excluded with onlyTracks).
// The group's tracklets are [off[e*T + p], off[e*T + p + 1]) of the
// reference's tracklet offsets, as the reference reads them.
#include <cstring>
#include <set>
#include <unordered_map>

#include "capi_internal.h"

using namespace trax;
using namespace trax_capi;

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
                auto clamp1 = [](uint32_t n) { return (double)(n > 0 ? n : 1u); };  // Utils::clamp (utils.h:65-67)
                r.efficiency = (double)r.found / clamp1(r.n_findable);
                r.fake_rate = (double)r.fakes / clamp1(r.n_tracklets);
                r.clone_rate = (double)r.clones / clamp1(r.n_tracklets);
            }
        return TRAX_OK;
    });
}
