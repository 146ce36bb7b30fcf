// Host classes of the triplet-seeding path (see trax_host.h).
#include "trax_host.h"

#include <algorithm>
#include <cstring>
#include <set>

namespace trax {

// ---------------------------------------------------------------------------
Context::Context(int device) : device_(device) {
    hip_check(hipSetDevice(device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&own_stream_, hipStreamNonBlocking), "hipStreamCreate");
    stream_ = own_stream_;
    hip_check(hipMalloc(&d_cnt_, sizeof(DevCounts)), "hipMalloc(counts)");
    DevCounts init{};
    // the plain path sizes every buffer exactly from the device counts
    init.cap_slots = init.cap_pairs = init.cap_cands = init.cap_tracklets = init.cap_conn = init.cap_basis =
        init.cap_tracks = 0xffffffffu;
    hip_check(hipMemcpy(d_cnt_, &init, sizeof(DevCounts), hipMemcpyHostToDevice), "h2d counts");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_pinned_), 256, hipHostMallocDefault), "hipHostMalloc");
    int v = 0;
    hip_check(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device), "attr");
    lds_bytes_ = (uint32_t)v;
    hip_check(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device), "attr");
    cu_count_ = (uint32_t)v;
}

Context::~Context() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (scan_.status) (void)hipFree(scan_.status);
    if (d_cnt_) (void)hipFree(d_cnt_);
    if (h_pinned_) (void)hipHostFree(h_pinned_);
    if (own_stream_) (void)hipStreamDestroy(own_stream_);
}

const ScanScratch& Context::scan_scratch(uint64_t n_cap) {
    const uint32_t tiles = scan_tiles(n_cap);
    if (tiles > scan_.max_tiles) {
        if (scan_.status) {  // a scan on any stream may still use it
            hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            hip_check(hipFree(scan_.status), "hipFree(scan)");
        }
        const uint32_t t = std::max<uint32_t>(tiles, 1024);
        hip_check(hipMalloc(&scan_.status, (size_t)(t + 1) * sizeof(uint64_t)), "hipMalloc(scan)");
        scan_.max_tiles = t;
    }
    return scan_;
}

uint32_t Context::download_count(const uint32_t* field) {
    hip_check(hipMemcpyAsync(h_pinned_, field, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_), "d2h count");
    finish_default_queue();
    return h_pinned_[0];
}

void Context::set_batch_shape(uint32_t n_events, uint32_t n_hits) {
    h_pinned_[8] = n_events;
    h_pinned_[9] = n_hits;
    hip_check(hipMemcpyAsync(d_cnt_, h_pinned_ + 8, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, stream_),
              "h2d shape");
    finish_default_queue();
}

void Context::reset_counts() {
    hip_check(launch_fill(reinterpret_cast<uint32_t*>(d_cnt_) + kCountsShapeWords, 0u,
                          sizeof(DevCounts) / 4 - kCountsShapeWords, stream_),
              "reset counts");
}

void Context::set_count(uint32_t* field, uint32_t v) {
    h_pinned_[16] = v;
    hip_check(hipMemcpyAsync(field, h_pinned_ + 16, sizeof(uint32_t), hipMemcpyHostToDevice, stream_), "h2d count");
    finish_default_queue();
}

void Context::check_errors() {
    const uint32_t e = download_count(&d_cnt_->err);
    if (e & kErrCellRange) throw InputError("a hit lies outside the z-phi grid");
    if (e & kErrCapacity) throw CapacityError("an output exceeded its capacity");
    if (e & kErrIndexRange) throw InputError("an index is out of range");
}

// ---------------------------------------------------------------------------
void TripletConfigurations::addWithValue(uint l1, uint l2, uint l3, float dTheta, float sZ, float dPhi,
                                         float sPhi, float tip, uint spreadZ, float z0_, uint spreadPhi,
                                         uint nCand) {
    layer1.push_back(l1); layer2.push_back(l2); layer3.push_back(l3);
    dThetaCut.push_back(dTheta); sigmaZ.push_back(sZ); dPhiCut.push_back(dPhi); sigmaPhi.push_back(sPhi);
    tipCut.push_back(tip); pairSpreadZ.push_back(spreadZ); z0.push_back(z0_);
    pairSpreadPhi.push_back(spreadPhi); nCandidates.push_back(nCand);
}

float TripletConfigurations::minRadiusCurvature() const {
    const double Q = 1.602177E-19;
    const double GEV_C = 5.344286E-19;
    const float Bz = (float)IDEAL_BZ;  // the float parameter of minRadiusCurvature(float, float)
    return (float)(minPt_ * GEV_C / (Bz * Q) * 1E2);
}

uint TripletConfigurations::maxLayer() const {
    uint m = 0;
    for (uint l : layer3) m = std::max(m, l);
    return m;
}

// ---- protocol-buffer wire format (PEventContainer, Event.proto:1-37) ------------
namespace {

struct WireReader {
    const uint8_t* p;
    const uint8_t* end;
    bool done() const { return p >= end; }
    uint64_t varint() {
        uint64_t v = 0;
        for (int shift = 0; shift < 64; shift += 7) {
            if (p >= end) throw InputError("truncated varint");
            const uint8_t b = *p++;
            v |= (uint64_t)(b & 0x7f) << shift;
            if (b < 0x80) return v;
        }
        throw InputError("varint longer than 10 bytes");
    }
    float fixed32f() {
        if (end - p < 4) throw InputError("truncated fixed32");
        uint32_t u = (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
        p += 4;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    }
    WireReader sub() {  // a length-delimited field's payload
        const uint64_t n = varint();
        if ((uint64_t)(end - p) < n) throw InputError("truncated length-delimited field");
        WireReader r{p, p + n};
        p += n;
        return r;
    }
    void skip(uint32_t wire) {
        switch (wire) {
            case 0: (void)varint(); break;
            case 1: if (end - p < 8) throw InputError("truncated fixed64"); p += 8; break;
            case 2: (void)sub(); break;
            case 5: if (end - p < 4) throw InputError("truncated fixed32"); p += 4; break;
            default: throw InputError("unsupported wire type");
        }
    }
};

// PHit: 1 position (PGlobalPoint: 1 x, 2 y, 3 z as float), 2 detectorType,
// 3 layer, 4 detectorId, 5 hitId, 6 simTrackID, 7 simTrackPt
PHit decodeHit(WireReader r) {
    PHit h{0.0f, 0.0f, 0.0f, 0u, 0u, 0u, 0.0f};
    while (!r.done()) {
        const uint64_t key = r.varint();
        const uint32_t field = (uint32_t)(key >> 3), wire = (uint32_t)(key & 7);
        if (field == 1 && wire == 2) {
            WireReader q = r.sub();
            while (!q.done()) {
                const uint64_t k = q.varint();
                const uint32_t f = (uint32_t)(k >> 3), w = (uint32_t)(k & 7);
                if (w == 5 && f >= 1 && f <= 3) {
                    const float v = q.fixed32f();
                    (f == 1 ? h.x : f == 2 ? h.y : h.z) = v;
                } else {
                    q.skip(w);
                }
            }
        } else if (field == 3 && wire == 0) {
            h.layer = (uint32_t)r.varint();
        } else if (field == 4 && wire == 0) {
            h.detectorId = (uint32_t)r.varint();
        } else if (field == 6 && wire == 0) {
            h.simTrackId = (uint32_t)r.varint();
        } else if (field == 7 && wire == 5) {
            h.simTrackPt = r.fixed32f();
        } else {
            r.skip(wire);
        }
    }
    return h;
}

}  // namespace

namespace {

// PEvent: 1 runNumber, 2 lumiSection, 3 eventNumber, 4 hits (PHit)
PEventData decodeEvent(WireReader e) {
    PEventData ev;
    while (!e.done()) {
        const uint64_t k = e.varint();
        const uint32_t f = (uint32_t)(k >> 3), w = (uint32_t)(k & 7);
        if (f == 4 && w == 2) ev.hits.push_back(decodeHit(e.sub()));
        else if (f == 1 && w == 0) ev.runNumber = e.varint();
        else if (f == 2 && w == 0) ev.lumiSection = e.varint();
        else if (f == 3 && w == 0) ev.eventNumber = e.varint();
        else e.skip(w);
    }
    return ev;
}

}  // namespace

std::vector<PEventData> decodeEventContainer(const uint8_t* data, size_t size) {
    std::vector<PEventData> out;
    WireReader r{data, data + size};
    while (!r.done()) {
        const uint64_t key = r.varint();
        if ((key >> 3) != 1 || (key & 7) != 2) {
            r.skip((uint32_t)(key & 7));
            continue;
        }
        out.push_back(decodeEvent(r.sub()));
    }
    return out;
}

std::vector<PEventData> decodeEventStore(const uint8_t* data, size_t size, uint32_t skip, uint32_t max_events,
                                         uint32_t* header_events) {
    std::vector<PEventData> out;
    if (header_events) *header_events = 0;
    if (size < 4) return out;
    const uint32_t n = (uint32_t)data[0] | (uint32_t)data[1] << 8 | (uint32_t)data[2] << 16 | (uint32_t)data[3] << 24;
    if (header_events) *header_events = n;
    WireReader r{data + 4, data + size};
    for (uint32_t read = 0; read < n && out.size() < max_events; ++read) {
        // readNextElement: the entry type (1 = an event), its size, its bytes;
        // a missing or other type, a zero size or a short entry ends the store
        uint64_t type, len;
        try {
            type = r.varint();
            if (type != 1) break;
            len = r.varint();
        } catch (const InputError&) {
            break;
        }
        if (len == 0 || (uint64_t)(r.end - r.p) < len) break;
        WireReader e{r.p, r.p + len};
        r.p += len;
        if (read < skip) continue;
        out.push_back(decodeEvent(e));
    }
    return out;
}

HitCollection::tTrackList HitCollection::addEvent(const std::vector<PHit>& ev, EventSupplement& eventSupplement,
                                                  uint evtInGroup, LayerSupplement& layerSupplement,
                                                  const TripletConfigurations& layerTriplets, int numTracks,
                                                  bool onlyTracks) {
    const float minPt = layerTriplets.minPt_;
    std::set<uint> usedLayers;
    uint maxLayer = 0;
    for (uint i = 0; i < layerTriplets.size(); ++i) {
        usedLayers.insert(layerTriplets.layer1[i]);
        usedLayers.insert(layerTriplets.layer2[i]);
        usedLayers.insert(layerTriplets.layer3[i]);
        maxLayer = std::max(maxLayer, layerTriplets.layer3[i]);
    }
    tTrackList allTracks;  // associate hits to tracks (HitCollection.cpp:41-54)
    for (const PHit& h : ev) {
        if (onlyTracks && h.simTrackId == 0) continue;
        if (h.simTrackPt < minPt) continue;
        allTracks[h.simTrackId].push_back(h);
    }
    // sort hits of each track by layer; std::sort leaves equal layers in any
    // order, the stable sort keeps the event order (one of those outcomes)
    for (auto& t : allTracks)
        std::stable_sort(t.second.begin(), t.second.end(),
                         [](const PHit& a, const PHit& b) { return a.layer < b.layer; });
    tTrackList layers, findableTracks;
    int foundTracks = 0;
    for (auto it = allTracks.begin(); it != allTracks.end() && foundTracks != numTracks; ++it) {
        if (onlyTracks && it->first == 0) continue;
        std::vector<bool> covLayers(maxLayer);
        for (const PHit& h : it->second)
            if (h.layer >= 1 && h.layer <= maxLayer) covLayers[h.layer - 1] = true;
        bool missingHits = false;
        for (uint l : usedLayers)
            if (!covLayers[l - 1]) missingHits = true;
        if (it->second[0].simTrackPt < minPt) continue;
        for (const PHit& h : it->second) {
            if (!missingHits) findableTracks[it->first].push_back(h);
            else if (!onlyTracks) layers[h.layer].push_back(h);
        }
        if (!missingHits) ++foundTracks;
    }
    for (auto& t : findableTracks)
        for (const PHit& h : t.second)
            if (h.layer <= maxLayer) layers[h.layer].push_back(h);
    uint offset = 0;
    for (uint i = 1; i <= maxLayer; ++i) {
        layerSupplement.nHits[evtInGroup * maxLayer + i - 1] = (uint)layers[i].size();
        layerSupplement.offset[evtInGroup * maxLayer + i - 1] = offset;
        offset += (uint)layers[i].size();
    }
    eventSupplement.nHits[evtInGroup] = offset;
    eventSupplement.offset[evtInGroup] =
        evtInGroup > 0 ? eventSupplement.offset[evtInGroup - 1] + eventSupplement.nHits[evtInGroup - 1] : 0;
    for (uint i = 1; i <= maxLayer; ++i)
        for (const PHit& h : layers[i]) {
            x.push_back(h.x); y.push_back(h.y); z.push_back(h.z);
            layer.push_back(h.layer); detId.push_back(h.detectorId);
            hitId.push_back(h.simTrackId); event.push_back(evtInGroup);
        }
    return findableTracks;
}

GridGeom GridConfig::geom() const {
    GridGeom g;
    g.min_z = MIN_Z;
    g.sector_z = sectorSizeZ();
    g.n_z = nSectorsZ;
    g.min_phi = MIN_PHI;
    g.sector_phi = sectorSizePhi();
    g.n_phi = nSectorsPhi;
    g.n_layers = nLayers;
    g.cells = cells();
    return g;
}

ConfigMirror::ConfigMirror(Context& ctx, const TripletConfigurations& t, const GeometrySupplement& g,
                           uint nThreads, float dEtaCut) {
    const uint T = t.size(), G = g.size();
    const size_t dbl_at = (10 * (size_t)T + 4 * (size_t)G + 1) & ~(size_t)1;  // 8-byte aligned doubles
    const size_t words = dbl_at + 2 * (size_t)T;
    std::vector<uint32_t> host(words);
    auto put_u = [&](size_t at, const std::vector<uint>& v) { std::memcpy(&host[at], v.data(), v.size() * 4); };
    auto put_f = [&](size_t at, const std::vector<float>& v) { std::memcpy(&host[at], v.data(), v.size() * 4); };
    put_u(0 * T, t.layer1); put_u(1 * T, t.layer2); put_u(2 * T, t.layer3);
    put_f(3 * T, t.dThetaCut); put_f(4 * T, t.sigmaZ); put_f(5 * T, t.dPhiCut); put_f(6 * T, t.sigmaPhi);
    put_f(7 * T, t.tipCut); put_f(8 * T, t.z0); put_u(9 * T, t.pairSpreadPhi);
    put_f(10 * T, g.minRadius); put_f(10 * T + G, g.maxRadius);
    const float mrc = t.minRadiusCurvature();
    std::vector<float> acos_min(G), acos_max(G);
    for (uint l = 0; l < G; ++l) {
        acos_min[l] = tx_acosf(g.minRadius[l] / (2 * mrc));
        acos_max[l] = tx_acosf(g.maxRadius[l] / (2 * mrc));
    }
    put_f(10 * T + 2 * G, acos_min); put_f(10 * T + 3 * G, acos_max);
    std::vector<double> tan_dphi(T);
    for (uint i = 0; i < T; ++i) tan_dphi[i] = tx_phi_prefilter_tan(t.dPhiCut[i]);
    if (T) std::memcpy(&host[dbl_at], tan_dphi.data(), T * sizeof(double));
    blob_.reset(new DeviceVector<uint8_t>(ctx, words * 4 + 4));
    hip_check(hipMemcpyAsync(blob_->get_mem(), host.data(), words * 4, hipMemcpyHostToDevice, ctx.stream()), "h2d cfg");
    ctx.finish_default_queue();
    const uint32_t* d = reinterpret_cast<const uint32_t*>(blob_->get_mem());
    const float* f = reinterpret_cast<const float*>(d);
    cols_.l1 = d; cols_.l2 = d + T; cols_.l3 = d + 2 * T;
    cols_.dtheta = f + 3 * T; cols_.sigma_z = f + 4 * T; cols_.dphi = f + 5 * T; cols_.sigma_phi = f + 6 * T;
    cols_.tip = f + 7 * T; cols_.z0 = f + 8 * T; cols_.pair_spread_phi = d + 9 * T;
    cols_.min_r = f + 10 * T; cols_.max_r = f + 10 * T + G;
    cols_.acos_min_r = f + 10 * T + 2 * G; cols_.acos_max_r = f + 10 * T + 3 * G;
    cols_.tan_dphi = reinterpret_cast<const double*>(d + dbl_at);
    cols_.n_trip = T;
    cols_.n_geom = G;
    cols_.min_radius_curvature = t.minRadiusCurvature();
    cols_.n_threads = nThreads;
    cols_.d_eta_cut = dEtaCut;
}

// ten words per hit: seven columns, one spare word and the (z, phi) pair at
// a 16-byte aligned offset, then kZphiSlack pairs of room past its end
HitMirror::HitMirror(Context& ctx, uint n)
    : n_(n), buf_(new DeviceVector<uint32_t>(ctx, 10ull * (n ? n : 1) + 2ull * kZphiSlack)) {}

HitColumns HitMirror::cols() const {
    HitColumns c;
    uint32_t* b = buf_->get_mem();
    const size_t n = n_ ? n_ : 1;
    c.x = reinterpret_cast<float*>(b); c.y = reinterpret_cast<float*>(b + n); c.z = reinterpret_cast<float*>(b + 2 * n);
    c.layer = b + 3 * n; c.det = b + 4 * n; c.id = b + 5 * n; c.evt = b + 6 * n;
    c.zphi = reinterpret_cast<float2*>(b + 8 * n);
    return c;
}

void HitMirror::upload(Context& ctx, const HitCollection& h) {
    const HitColumns c = cols();
    const size_t bytes = (size_t)n_ * 4;
    if (!n_) return;
    auto cp = [&](void* dst, const void* src) {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx.stream()), "h2d hits");
    };
    cp(c.x, h.x.data()); cp(c.y, h.y.data()); cp(c.z, h.z.data());
    cp(c.layer, h.layer.data()); cp(c.det, h.detId.data()); cp(c.id, h.hitId.data()); cp(c.evt, h.event.data());
}

void HitMirror::download(Context& ctx, HitCollection& h) const {
    const HitColumns c = cols();
    const size_t bytes = (size_t)n_ * 4;
    h.x.resize(n_); h.y.resize(n_); h.z.resize(n_);
    h.layer.resize(n_); h.detId.resize(n_); h.hitId.resize(n_); h.event.resize(n_);
    if (!n_) return;
    auto cp = [&](void* dst, const void* src) {
        hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx.stream()), "d2h hits");
    };
    cp(h.x.data(), c.x); cp(h.y.data(), c.y); cp(h.z.data(), c.z);
    cp(h.layer.data(), c.layer); cp(h.detId.data(), c.det); cp(h.hitId.data(), c.id); cp(h.event.data(), c.evt);
    ctx.finish_default_queue();
}

// ---------------------------------------------------------------------------
void PrefixSum::run(uint32_t* input, uint size, uint nThreads, uint32_t* d_total) {
    (void)nThreads;  // the single-pass scan does not depend on the work-group size
    hip_check(launch_scan(input, nullptr, size, size, d_total, ctx.scan_scratch(size), ctx.stream()), "scan");
}

void HitTransfer::initBuffers(Context& ctx, const HitCollection& hits) { m.reset(new HitMirror(ctx, hits.size())); }

void HitTransfer::toDevice(Context& ctx, const HitCollection& hits) {
    if (!m || m->n_ != hits.size()) initBuffers(ctx, hits);
    m->upload(ctx, hits);
}

void HitTransfer::fromDevice(Context& ctx, HitCollection& hits) const { mirror().download(ctx, hits); }

HitMirror& HitTransfer::mirror() const {
    if (!m) throw InputError("the hit collection has no device buffers (transfer.initBuffers / toDevice first)");
    return *m;
}

uint64_t GridBuilder::run(HitCollection& hits, uint nThreads, const EventSupplement& es, const LayerSupplement& ls,
                          Grid& grid, bool printProlix) {
    (void)nThreads;
    (void)printProlix;
    build(hits, hits.transfer.mirror(), es, ls, grid);
    grid.host = grid.boundary.download(ctx);   // GridBuilder.cpp:64
    hits.transfer.fromDevice(ctx, hits);       // GridBuilder.cpp:151
    return 0;
}

void GridBuilder::build(HitCollection& hits, HitMirror& dhits, const EventSupplement& es, const LayerSupplement& ls,
                        Grid& grid) {

    const GridConfig& gc = grid.config;
    const uint E = gc.nEvents, L = gc.nLayers, N = hits.size();
    if (es.nEvents != E || ls.nLayers != L) throw InputError("supplement shape does not match the grid");
    // the grid values are global hit indices: the supplements must be running sums
    uint running = 0;
    for (uint e = 0; e < E; ++e) {
        if (es.offset[e] != running) throw InputError("event offsets are not running sums");
        uint lr = 0;
        for (uint l = 0; l < L; ++l) {
            if (ls.offset[e * L + l] != lr) throw InputError("layer offsets are not running sums");
            lr += ls.nHits[e * L + l];
        }
        if (lr != es.nHits[e]) throw InputError("layer hit counts do not add up to the event's");
        running += es.nHits[e];
    }
    if (running != N) throw InputError("event hit counts do not add up to the collection");
    std::vector<uint32_t> sup(E + 2 * (size_t)E * L);
    std::copy(es.offset.begin(), es.offset.end(), sup.begin());
    std::copy(ls.nHits.begin(), ls.nHits.end(), sup.begin() + E);
    std::copy(ls.offset.begin(), ls.offset.end(), sup.begin() + E + (size_t)E * L);
    UVec dsup(ctx, sup.size());
    dsup.upload(ctx, sup);
    ctx.set_batch_shape(E, N);
    const GridGeom g = gc.geom();
    const uint32_t* d_eo = dsup.get_mem();
    const uint32_t* d_ln = d_eo + E;
    const uint32_t* d_lo = d_ln + (size_t)E * L;
    hip_check(launch_grid_count(g, E * L, d_eo, d_ln, d_lo, dhits.cols().x, dhits.cols().y, dhits.cols().z,
                                grid.boundary.get_mem(), ctx.counts(), ctx.launch(), ctx.stream()),
              "grid count");
    ctx.finish_default_queue();  // GridBuilder.cpp:52
    HitMirror out(ctx, N);       // hits.transfer.initBuffers (GridBuilder.cpp:70)
    hip_check(launch_grid_store(g, E * L, d_eo, d_ln, d_lo, grid.boundary.get_mem(), dhits.cols(), out.cols(), N,
                                ctx.counts(), ctx.launch(), ctx.stream()),
              "grid store");
    ctx.finish_default_queue();  // GridBuilder.cpp:146
    ctx.check_errors();
    dhits.swap(out);  // the old buffers are released with `out`
}


Pairing* PairGeneratorBeamspot::run(HitCollection& hits, const GeometrySupplement& geomSupplement, uint nThreads,
                                    const TripletConfigurations& layerTriplets, const Grid& grid, bool printProlix) {
    (void)printProlix;
    const ConfigMirror cfg(ctx, layerTriplets, geomSupplement, nThreads, 0.0f);
    const uint E = grid.config.nEvents, T = layerTriplets.size();
    const GridGeom g = grid.config.geom();
    const TripletCols tc = cfg.cols();
    DevCounts* cnt = ctx.counts();
    const uint nseg = E * T;
    UVec seg_base(ctx, nseg + 1ull);
    hip_check(launch_pair_plan(g, tc, nseg, grid.boundary.get_mem(), seg_base.get_mem(), cnt, ctx.stream()), "pair plan");
    PrefixSum(ctx).run(seg_base.get_mem(), nseg + 1, nThreads, &cnt->n_pair_slots);
    const uint nslot = ctx.download_count(&cnt->n_pair_slots);
    UVec slot_counts(ctx, nslot + 1ull);
    const HitColumns hc = hits.transfer.mirror().cols();
    hip_check(launch_pair_count(g, tc, nseg, nslot, grid.boundary.get_mem(), hc.x, hc.y, hc.z, hc.zphi, seg_base.get_mem(),
                                slot_counts.get_mem(), cnt, ctx.launch(), ctx.stream()),
              "pair count");
    PrefixSum(ctx).run(slot_counts.get_mem(), nslot + 1, nThreads, &cnt->n_pairs);
    const uint nPairs = ctx.download_count(&cnt->n_pairs);  // PairGeneratorBeamspot.cpp:124
    Pairing* p = new Pairing(ctx, nPairs, E, T);
    hip_check(launch_pair_store(g, tc, nseg, nslot, nPairs, grid.boundary.get_mem(), hc.x, hc.y, hc.z, hc.zphi,
                                seg_base.get_mem(), slot_counts.get_mem(), p->pairing.get_mem(),
                                p->pairingOffsets.get_mem(), cnt, ctx.launch(), ctx.stream()),
              "pair store");
    ctx.finish_default_queue();  // the temporaries are released on return
    return p;
}

Pairing* TripletThetaPhiPredictor::run(HitCollection& hits, const DetectorGeometry&,
                                       const GeometrySupplement& geomSupplement, const Dictionary&, int nThreads,
                                       const TripletConfigurations& layerTriplets, const Grid& grid,
                                       const Pairing& pairs, bool printProlix) {
    (void)printProlix;
    const ConfigMirror cfg(ctx, layerTriplets, geomSupplement, (uint)nThreads, 0.0f);
    const uint E = grid.config.nEvents, T = layerTriplets.size();
    const GridGeom g = grid.config.geom();
    const TripletCols tc = cfg.cols();
    DevCounts* cnt = ctx.counts();
    const uint nPairs = pairs.n;
    ctx.set_count(&cnt->n_pairs, pairs.n);
    UVec counts(ctx, nPairs + 1ull);
    const HitColumns hc = hits.transfer.mirror().cols();
    hip_check(launch_predict_count(g, tc, nPairs, grid.boundary.get_mem(), hc.x, hc.y, hc.z, hc.zphi, hc.layer, hc.evt,
                                   pairs.pairing.get_mem(), counts.get_mem(), cnt, ctx.launch(), ctx.stream()),
              "predict count");
    PrefixSum(ctx).run(counts.get_mem(), nPairs + 1, (uint)nThreads, &cnt->n_cands);
    const uint nCand = ctx.download_count(&cnt->n_cands);  // TripletThetaPhiPredictor.cpp:68
    Pairing* c = new Pairing(ctx, nCand, E, T);
    hip_check(launch_fill(c->pairingOffsets.get_mem(), 0u, c->pairingOffsets.get_count(), ctx.stream()), "fill");
    hip_check(launch_predict_store(g, tc, nPairs, nCand, grid.boundary.get_mem(), hc.x, hc.y, hc.z, hc.zphi, hc.layer,
                                   hc.evt, pairs.pairing.get_mem(), counts.get_mem(), c->pairing.get_mem(),
                                   c->pairingOffsets.get_mem(), cnt, ctx.launch(), ctx.stream()),
              "predict store");
    // predictOffsetMonotonizeStore (TripletThetaPhiPredictor.h:432-447)
    hip_check(launch_forward_fill(c->pairingOffsets.get_mem(), cnt, T, ctx.stream()), "offsets fill");
    ctx.finish_default_queue();
    return c;
}

TrackletCollection* TripletThetaPhiFilter::run(HitCollection& hits, const Grid& grid, const Pairing& pairs,
                                               const Pairing& cands, int nThreads,
                                               const TripletConfigurations& layerTriplets, bool printProlix) {
    (void)printProlix;
    const ConfigMirror cfg(ctx, layerTriplets, GeometrySupplement(), (uint)nThreads, 0.0f);
    const uint E = grid.config.nEvents, T = layerTriplets.size();
    const TripletCols tc = cfg.cols();
    DevCounts* cnt = ctx.counts();
    const uint nCand = cands.n;
    const uint nWords = (nCand + 63) / 64;
    ctx.set_count(&cnt->n_cands, cands.n);
    DeviceVector<uint64_t> bitmap(ctx, nWords ? nWords : 1);
    UVec word_counts(ctx, nWords + 1ull);
    const HitColumns hc = hits.transfer.mirror().cols();
    hip_check(launch_filter_count(tc, nCand, hc.x, hc.y, hc.z, hc.layer, pairs.pairing.get_mem(),
                                  cands.pairing.get_mem(), bitmap.get_mem(), word_counts.get_mem(), cnt,
                                  ctx.launch(), ctx.stream()),
              "filter count");
    PrefixSum(ctx).run(word_counts.get_mem(), nWords + 1, (uint)nThreads, &cnt->n_tracklets);
    const uint R = ctx.download_count(&cnt->n_tracklets);  // TripletThetaPhiFilter.cpp:85
    TrackletCollection* t = new TrackletCollection(ctx, R, E, T);
    hip_check(launch_fill(t->trackletOffsets.get_mem(), 0u, t->trackletOffsets.get_count(), ctx.stream()), "fill");
    hip_check(launch_filter_store(tc, nCand, R, hc.layer, hc.evt, pairs.pairing.get_mem(), cands.pairing.get_mem(),
                                  bitmap.get_mem(), word_counts.get_mem(), t->hit1.get_mem(), t->hit2.get_mem(),
                                  t->hit3.get_mem(), t->trackletOffsets.get_mem(), cnt, ctx.launch(), ctx.stream()),
              "filter store");
    // filterOffsetStore + filterOffsetMonotonizeStore (TripletThetaPhiFilter.h:221-280)
    hip_check(launch_filter_offsets(tc, R, hc.layer, hc.evt, t->hit1.get_mem(), t->trackletOffsets.get_mem(), cnt,
                                    ctx.launch(), ctx.stream()),
              "filter offsets");
    ctx.finish_default_queue();
    t->h1 = t->hit1.download(ctx);  // TripletThetaPhiFilter.cpp:184
    t->h2 = t->hit2.download(ctx);
    t->h3 = t->hit3.download(ctx);
    return t;
}

std::tuple<UVec*, UVec*, UVec*, UVec*> TripletConnectivityTight::run(const HitCollection& hits, TrackletCollection& tr,
                                                                    const TripletConfigurations& layerTriplets,
                                                                    float dEtaCut, uint nThreads, bool printPROLIX,
                                                                    ConnectivityDiag* diag) const {
    (void)printPROLIX;
    TripletCols tc{};  // the connectivity kernels read only the triplet count and dEtaCut
    tc.n_trip = layerTriplets.size();
    tc.d_eta_cut = dEtaCut;
    tc.n_threads = nThreads;
    DevCounts* cnt = ctx.counts();
    const uint R = tr.size();
    const HitMirror& dh = hits.transfer.mirror();
    const uint N = dh.n_;
    ctx.set_count(&cnt->n_tracklets, tr.n);
    std::tuple<UVec*, UVec*, UVec*, UVec*> none(nullptr, nullptr, nullptr, nullptr);
    if (R == 0) return none;
    std::unique_ptr<FVec> eta(new FVec(ctx, R));
    UVec hit_count(ctx, N + 1ull), hit_cursor(ctx, N ? N : 1), csr(ctx, R), conn(ctx, R + 1ull),
        counts(ctx, R + 1ull), flag_a(ctx, R + 1ull), flag_b(ctx, R + 1ull);
    LinkBufs b;
    b.eta = eta->get_mem();
    b.hit_count = hit_count.get_mem(); b.hit_cursor = hit_cursor.get_mem(); b.csr = csr.get_mem();
    b.conn = conn.get_mem(); b.counts = counts.get_mem(); b.flag_a = flag_a.get_mem(); b.flag_b = flag_b.get_mem();
    const HitColumns hc = dh.cols();
    // tripletEtaCalculatorStore (.h:47-79) and the per-hit buckets of the tracklets
    hip_check(launch_link_prepare(tc, N, R, hc.x, hc.y, hc.z, tr.hit1.get_mem(), tr.hit2.get_mem(), tr.hit3.get_mem(),
                                  b, cnt, ctx.scan_scratch(N + 1ull), ctx.launch(), ctx.stream()),
              "link prepare");
    if (diag) diag->eta = std::move(eta);
    hip_check(launch_link_count(tc, R, hc.layer, hc.evt, tr.hit1.get_mem(), tr.hit2.get_mem(), tr.hit3.get_mem(),
                                tr.trackletOffsets.get_mem(), b, cnt, ctx.launch(), ctx.stream()),
              "link count");
    PrefixSum(ctx).run(counts.get_mem(), R + 1, nThreads, &cnt->n_conn);
    const uint nConn = ctx.download_count(&cnt->n_conn);  // TripletConnectivityTight.cpp:89
    if (nConn == 0) return none;                           // .cpp:93-95: null tuple
    std::unique_ptr<UVec> basis(new UVec(ctx, nConn)), foll(new UVec(ctx, nConn)), pev(new UVec(ctx, nConn));
    b.basis = basis->get_mem(); b.foll = foll->get_mem(); b.pev = pev->get_mem();
    hip_check(launch_link_store(tc, R, nConn, hc.layer, hc.evt, tr.hit1.get_mem(), tr.hit2.get_mem(),
                                tr.hit3.get_mem(), tr.trackletOffsets.get_mem(), b, cnt, ctx.launch(), ctx.stream()),
              "link store");
    // predicateInversionStore + streamCompactionGetValidIndexesStore (.h:251-346, .cpp:184-256)
    hip_check(launch_link_flags(R, hc.evt, tr.hit1.get_mem(), b, cnt, ctx.launch(), ctx.stream()), "link flags");
    PrefixSum(ctx).run(flag_a.get_mem(), R + 1, nThreads, &cnt->n_connectable);
    const uint nConnectable = ctx.download_count(&cnt->n_connectable);  // .cpp:200
    PrefixSum(ctx).run(flag_b.get_mem(), R + 1, nThreads, &cnt->n_nonconnectable);
    const uint nNonConnectable = ctx.download_count(&cnt->n_nonconnectable);  // .cpp:211
    std::unique_ptr<UVec> connectable(new UVec(ctx, nConnectable)), nonconnectable(new UVec(ctx, nNonConnectable));
    b.connectable = connectable->get_mem();
    b.nonconnectable = nonconnectable->get_mem();
    hip_check(launch_link_compact(R, b, cnt, ctx.launch(), ctx.stream()), "link compact");
    ctx.finish_default_queue();
    ctx.check_errors();
    if (diag) diag->event = std::move(pev);
    return std::make_tuple(basis.release(), foll.release(), connectable.release(), nonconnectable.release());
}

FVec* TrackletCircleFitter::run(const HitCollection& hits, const TrackletCollection& tr, const UVec& valid,
                                uint nThreads, bool printPROLIX) const {
    (void)nThreads;
    (void)printPROLIX;
    DevCounts* cnt = ctx.counts();
    std::unique_ptr<FVec> pt(new FVec(ctx, tr.size()));
    ctx.set_count(&cnt->n_tracklets, tr.n);
    ctx.set_count(&cnt->n_connectable, (uint)valid.get_count());
    LinkBufs b;
    b.connectable = valid.get_mem();
    b.pt = pt->get_mem();
    const HitColumns hc = hits.transfer.mirror().cols();
    hip_check(launch_circle_fit(tr.size(), hc.x, hc.y, tr.hit1.get_mem(), tr.hit2.get_mem(), tr.hit3.get_mem(), b,
                                cnt, ctx.launch(), ctx.stream()),
              "circle fit");
    ctx.finish_default_queue();
    return pt.release();
}

std::tuple<const UVec*, const UVec*> CellularAutomaton::run(const UVec& basis, const UVec& followers, const FVec& pt,
                                                           const UVec& connectableTriplets, uint nThreads,
                                                           bool printPROLIX, CADiag* diag) const {
    (void)connectableTriplets;  // the reference uses it only for debugging output (CellularAutomaton.cpp:11)
    (void)printPROLIX;
    const uint nConn = (uint)basis.get_count(), R = (uint)pt.get_count();  // CellularAutomaton.cpp:29-30
    DevCounts* cnt = ctx.counts();
    ctx.set_count(&cnt->n_tracklets, R);
    ctx.set_count(&cnt->n_conn, nConn);
    std::unique_ptr<UVec> states(new UVec(ctx, R)), bestBasis(new UVec(ctx, R)), trackOffsets(new UVec(ctx, R + 1ull));
    UVec st_b(ctx, R), flags(ctx, 3), fbc(ctx, R + 1ull), cursor(ctx, R), isbest(ctx, R);
    CABufs b;
    b.st_a = states->get_mem(); b.st_b = st_b.get_mem(); b.flags = flags.get_mem();
    b.fbc = fbc.get_mem(); b.cursor = cursor.get_mem(); b.best = bestBasis->get_mem();
    b.isbest = isbest.get_mem(); b.handler = trackOffsets->get_mem();
    hip_check(launch_ca_init(R, b, cnt, ctx.stream()), "ca init");
    // forward phase: one iteration per launch, living cells read back each time
    // (CellularAutomaton.cpp:40-76)
    uint k = 0;
    while (true) {
        hip_check(launch_ca_iterations(nConn, basis.get_mem(), followers.get_mem(), b, cnt, k, 1, ctx.launch(),
                                       ctx.stream()),
                  "ca iteration");
        const uint alive = ctx.download_count(flags.get_mem() + (k % 3));
        ++k;
        if (!alive) break;
    }
    const uint iterations = ctx.download_count(&cnt->ca_iterations);
    hip_check(launch_ca_basis_count(R, nConn, basis.get_mem(), followers.get_mem(), b, cnt, ctx.launch(), ctx.stream()),
              "ca basis count");
    PrefixSum(ctx).run(fbc.get_mem(), R + 1, nThreads, &cnt->n_basis);
    const uint nb = ctx.download_count(&cnt->n_basis);  // CellularAutomaton.cpp:138
    UVec bidx(ctx, nb ? nb : 1);
    FVec bdiff(ctx, nb ? nb : 1);
    b.bidx = bidx.get_mem();
    b.bdiff = bdiff.get_mem();
    hip_check(launch_ca_basis_store(R, nConn, nb, basis.get_mem(), followers.get_mem(), pt.get_mem(), b, cnt,
                                    ctx.launch(), ctx.stream()),
              "ca basis store");
    PrefixSum(ctx).run(b.handler, R + 1, nThreads, &cnt->n_track_entries);
    const uint nTrackEntries = ctx.download_count(&cnt->n_track_entries);  // CellularAutomaton.cpp:284
    std::unique_ptr<UVec> tracks(new UVec(ctx, nTrackEntries));
    b.tracks = tracks->get_mem();
    hip_check(launch_ca_tracks(R, nTrackEntries, b, cnt, ctx.scan_scratch(R + 1ull), ctx.launch(), ctx.stream(), 0),
              "ca tracks");
    ctx.finish_default_queue();
    ctx.check_errors();
    if (diag) {
        diag->states = std::move(states);
        diag->bestBasis = std::move(bestBasis);
        diag->iterations = iterations;
    }
    return std::make_tuple(tracks.release(), trackOffsets.release());
}

}  // namespace trax
