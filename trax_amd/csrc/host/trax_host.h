// C++ host classes of the triplet-seeding path, named and shaped after the
// reference's containers (src/datastructures) and algorithm classes
// (src/algorithms), running on the gfx950 launchers instead of clever/OpenCL.
//
// Ownership follows the reference: algorithm run() calls return heap objects
// owned by the caller (Pairing*, TrackletCollection*, device vectors); the
// GridBuilder rewrites the hit columns of the HitCollection in place.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../kernels/kernels.h"

namespace trax {

using uint = uint32_t;

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string& where)
        : std::runtime_error(where + ": " + hipGetErrorString(c)), code(c) {}
};
struct InputError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct CapacityError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* where) {
    if (e != hipSuccess) throw HipError(e, where);
}

// clever::context: one device, one in-order queue, scan scratch, the
// device-resident size record and the launch tuning
class Context {
public:
    explicit Context(int device);
    ~Context();
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }
    void set_stream(hipStream_t s) { stream_ = s ? s : own_stream_; }
    DevCounts* counts() const { return d_cnt_; }
    const ScanScratch& scan_scratch(uint64_t n_cap);  // grows on demand
    LaunchCfg& launch() { return lc_; }
    uint32_t local_mem_size() const { return lds_bytes_; }  // getLocalMemSize
    uint32_t cu_count() const { return cu_count_; }
    void finish_default_queue() { hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }
    // transfer::downloadScalar of one DevCounts field (blocking)
    uint32_t download_count(const uint32_t* field);
    // host -> device copy of the batch shape into the size record
    void set_batch_shape(uint32_t n_events, uint32_t n_hits);
    void reset_counts();  // zero every size / flag of the record except the shape
    void set_count(uint32_t* field, uint32_t v);  // blocking host -> device of one field
    void check_errors();  // raise on device-side error bits

private:
    int device_;
    hipStream_t own_stream_ = nullptr, stream_ = nullptr;
    DevCounts* d_cnt_ = nullptr;
    uint32_t* h_pinned_ = nullptr;
    ScanScratch scan_;
    uint32_t lds_bytes_ = 0, cu_count_ = 0;
    LaunchCfg lc_;
};

// clever::vector<T, 1>
template <class T>
class DeviceVector {
public:
    DeviceVector(Context& ctx, uint64_t count) : count_(count) {
        hip_check(hipMalloc(&ptr_, bytes() ? bytes() : sizeof(T)), "hipMalloc");
        (void)ctx;
    }
    ~DeviceVector() {
        if (ptr_) (void)hipFree(ptr_);
    }
    DeviceVector(const DeviceVector&) = delete;
    DeviceVector& operator=(const DeviceVector&) = delete;
    T* get_mem() const { return ptr_; }
    uint64_t get_count() const { return count_; }
    uint64_t bytes() const { return count_ * sizeof(T); }
    std::vector<T> download(Context& ctx) const {
        std::vector<T> v(count_);
        if (count_) {
            hip_check(hipMemcpyAsync(v.data(), ptr_, bytes(), hipMemcpyDeviceToHost, ctx.stream()), "d2h");
            ctx.finish_default_queue();
        }
        return v;
    }
    void upload(Context& ctx, const std::vector<T>& v) {
        if (v.size() && count_)
            hip_check(hipMemcpyAsync(ptr_, v.data(), std::min<uint64_t>(v.size(), count_) * sizeof(T),
                                     hipMemcpyHostToDevice, ctx.stream()),
                      "h2d");
    }

private:
    T* ptr_ = nullptr;
    uint64_t count_;
};

// -- containers -------------------------------------------------------------

struct PHit {  // PB_Event::PHit (src/datastructures/serialize/Event.proto)
    float x, y, z;
    uint32_t layer, detectorId, simTrackId;
    float simTrackPt;
};

// One decoded PEvent of a PEventContainer (Event.proto:1-37): its header
// numbers and hits.
struct PEventData {
    uint64_t runNumber = 0, lumiSection = 0, eventNumber = 0;
    std::vector<PHit> hits;
};
// Decodes a serialized PB_Event::PEventContainer (the input the reference's
// EventLoader parses, src/triplet_builder/EventLoader.h:15-67) from the
// protocol-buffer wire format, without libprotobuf.  Throws InputError on
// malformed input.
std::vector<PEventData> decodeEventContainer(const uint8_t* data, size_t size);
// Decodes an event store as GenericOutputStore writes it and GenericInputStore
// reads it (src/datastructures/serialize/PEventStore.h:29-248): the event count
// as a little-endian u32, then per event its entry type (a varint, 1 =
// event), its size (a varint) and the serialized PB_Event::PEvent.  Reading
// stops at the count, at an entry of another type, at a zero size or at an
// entry the data cuts short (readNextElement's NULL); the first skip events
// are read and dropped (EventStoreLoader, src/triplet_builder/EventLoader.h:
// 171-206), at most max_events are returned.  *header_events receives the
// count.  Throws InputError on a malformed event.
std::vector<PEventData> decodeEventStore(const uint8_t* data, size_t size, uint32_t skip, uint32_t max_events,
                                         uint32_t* header_events);

// TripletConfigurations (src/datastructures/TripletConfiguration.h:39-72)
class TripletConfigurations {
public:
    explicit TripletConfigurations(float minPt) : minPt_(minPt) {}
    uint size() const { return (uint)layer1.size(); }
    void addWithValue(uint l1, uint l2, uint l3, float dTheta, float sZ, float dPhi, float sPhi, float tip,
                      uint spreadZ, float z0_, uint spreadPhi, uint nCand);
    float minRadiusCurvature() const;  // TripletConfiguration.cpp:36-43
    uint maxLayer() const;
    std::vector<uint> layer1, layer2, layer3;
    std::vector<float> dThetaCut, sigmaZ, dPhiCut, sigmaPhi, tipCut;
    std::vector<uint> pairSpreadZ;
    std::vector<float> z0;
    std::vector<uint> pairSpreadPhi, nCandidates;
    float minPt_;
    static constexpr double IDEAL_BZ = 3.8112;
};

// GeometrySupplement (src/datastructures/GeometrySupplement.h:21-81)
class GeometrySupplement {
public:
    void addWithValue(uint layer, float minR, float maxR) {
        detectorLayer.push_back(layer); minRadius.push_back(minR); maxRadius.push_back(maxR);
    }
    uint size() const { return (uint)minRadius.size(); }
    std::vector<uint> detectorLayer;
    std::vector<float> minRadius, maxRadius;
};

// inert arguments of TripletThetaPhiPredictor::run (never read by its kernels,
// src/algorithms/TripletThetaPhiPredictor.h:54)
struct DetectorGeometry {};
struct Dictionary {};

// EventSupplement / LayerSupplement (EventSupplement.h:14, LayerSupplement.h:13)
struct EventSupplement {
    explicit EventSupplement(uint n = 1) : nEvents(n), nHits(n, 0), offset(n, 0) {}
    uint nEvents;
    std::vector<uint> nHits, offset;
};
struct LayerSupplement {
    LayerSupplement(uint layers, uint events = 1)
        : nLayers(layers), nHits((size_t)layers * events, 0), offset((size_t)layers * events, 0) {}
    uint nLayers;
    std::vector<uint> nHits, offset;
};

struct HitMirror;
class HitCollection;
// clever::OpenCLTransfer<...> of a HitCollection (HitCollection.h:58): the
// device buffers of its columns.  initBuffers allocates them for the
// collection's size, toDevice / fromDevice copy; the GridBuilder replaces
// them with the grid-ordered columns (GridBuilder.cpp:70,157).
struct HitTransfer {
    void initBuffers(Context& ctx, const HitCollection& hits);
    void toDevice(Context& ctx, const HitCollection& hits);
    void fromDevice(Context& ctx, HitCollection& hits) const;
    HitMirror& mirror() const;  // InputError before initBuffers
    std::shared_ptr<HitMirror> m;
};

// HitCollection (src/datastructures/HitCollection.h:28-62)
class HitCollection {
public:
    typedef std::map<uint, std::vector<PHit>> tTrackList;
    uint size() const { return (uint)x.size(); }
    // HitCollection.cpp:21-180
    tTrackList addEvent(const std::vector<PHit>& event, EventSupplement& eventSupplement, uint evtInGroup,
                        LayerSupplement& layerSupplement, const TripletConfigurations& layerTriplets,
                        int numTracks = -1, bool onlyTracks = false);
    std::vector<float> x, y, z;
    std::vector<uint> layer, detId, hitId, event;
    HitTransfer transfer;
};

// GridConfig / Grid (src/datastructures/Grid.h:5-131)
struct GridConfig {
    uint nEvents = 1, nLayers = 1, nSectorsZ = 1, nSectorsPhi = 1;
    static constexpr float MIN_Z = -300;
    static constexpr float MAX_Z = 300;
    static constexpr float MIN_PHI = -(float)M_PI;
    static constexpr float MAX_PHI = (float)M_PI;
    float sectorSizeZ() const { return (MAX_Z - MIN_Z) / nSectorsZ; }
    float sectorSizePhi() const { return (MAX_PHI - MIN_PHI) / nSectorsPhi; }
    uint cells() const { return (nSectorsZ + 1) * (nSectorsPhi + 1); }
    GridGeom geom() const;
};

// device mirrors of the configuration columns
struct ConfigMirror {
    ConfigMirror(Context& ctx, const TripletConfigurations& t, const GeometrySupplement& g, uint nThreads,
                 float dEtaCut);
    TripletCols cols() const { return cols_; }

private:
    std::unique_ptr<DeviceVector<uint8_t>> blob_;
    TripletCols cols_;
};

// hit columns on the device (clever::OpenCLTransfer of the HitCollection)
struct HitMirror {
    HitMirror(Context& ctx, uint n);
    void upload(Context& ctx, const HitCollection& h);
    void download(Context& ctx, HitCollection& h) const;
    HitColumns cols() const;
    void swap(HitMirror& o) { buf_.swap(o.buf_); n_ = o.n_; }
    uint n_;
    std::unique_ptr<DeviceVector<uint32_t>> buf_;  // 7 columns of n entries
};

struct Grid {
    Grid(Context& ctx, const GridConfig& cfg)
        : config(cfg), boundary(ctx, (uint64_t)cfg.nEvents * cfg.nLayers * cfg.cells()) {}
    GridConfig config;
    DeviceVector<uint32_t> boundary;
    std::vector<uint> host;  // downloaded after the build (GridBuilder.cpp:64)
};

// Pairing (src/datastructures/Pairings.h:9-59): uint2 pairs + E*T+1 offsets
struct Pairing {
    Pairing(Context& ctx, uint nTotalPairings, uint nEvents, uint nLayerTriplets)
        : pairing(ctx, 2ull * nTotalPairings), pairingOffsets(ctx, (uint64_t)nEvents * nLayerTriplets + 1),
          n(nTotalPairings) {}
    DeviceVector<uint32_t> pairing;
    DeviceVector<uint32_t> pairingOffsets;
    uint n;
};

// TrackletCollection (src/datastructures/TrackletCollection.h:36-78)
struct TrackletCollection {
    TrackletCollection(Context& ctx, uint items, uint nEvents, uint nLayerTriplets)
        : hit1(ctx, items), hit2(ctx, items), hit3(ctx, items),
          trackletOffsets(ctx, (uint64_t)nEvents * nLayerTriplets + 1), n(items) {}
    uint size() const { return n; }
    DeviceVector<uint32_t> hit1, hit2, hit3;
    DeviceVector<uint32_t> trackletOffsets;
    std::vector<uint> h1, h2, h3;  // host copies (TripletThetaPhiFilter.cpp:184)
    uint n;
};

using UVec = DeviceVector<uint32_t>;
using FVec = DeviceVector<float>;

// -- algorithms -----------------------------------------------------------

class PrefixSum {  // src/algorithms/PrefixSum.h:16-56
public:
    explicit PrefixSum(Context& c) : ctx(c) {}
    // exclusive in-place scan of `size` entries; the sum goes to *d_total
    void run(uint32_t* input, uint size, uint nThreads, uint32_t* d_total = nullptr);
    Context& ctx;
};

class GridBuilder {  // src/algorithms/GridBuilder.h:25-46
public:
    explicit GridBuilder(Context& c) : ctx(c) {}
    // GridBuilder.h:42: count, scan and stable scatter of hits.transfer's
    // columns into the grid; the transfer then holds the grid-ordered
    // columns, and the grid and the hits are downloaded (GridBuilder.cpp:64,151)
    uint64_t run(HitCollection& hits, uint nThreads, const EventSupplement& eventSupplement,
                 const LayerSupplement& layerSupplement, Grid& grid, bool printProlix = false);
    Context& ctx;

private:
    void build(HitCollection& hits, HitMirror& dhits, const EventSupplement& es, const LayerSupplement& ls,
               Grid& grid);
};

class PairGeneratorBeamspot {  // src/algorithms/PairGeneratorBeamspot.h:24-41
public:
    explicit PairGeneratorBeamspot(Context& c) : ctx(c) {}
    // PairGeneratorBeamspot.h:40; nThreads fixes the pair order (the
    // reference's work-item order)
    Pairing* run(HitCollection& hits, const GeometrySupplement& geomSupplement, uint nThreads,
                 const TripletConfigurations& layerTriplets, const Grid& grid, bool printProlix = false);
    Context& ctx;
};

class TripletThetaPhiPredictor {  // src/algorithms/TripletThetaPhiPredictor.h:44-47
public:
    explicit TripletThetaPhiPredictor(Context& c) : ctx(c) {}
    Pairing* run(HitCollection& hits, const DetectorGeometry& geom, const GeometrySupplement& geomSupplement,
                 const Dictionary& dict, int nThreads, const TripletConfigurations& layerTriplets, const Grid& grid,
                 const Pairing& pairs, bool printProlix = false);
    Context& ctx;
};

class TripletThetaPhiFilter {  // src/algorithms/TripletThetaPhiFilter.h:46-48
public:
    explicit TripletThetaPhiFilter(Context& c) : ctx(c) {}
    // the tracklets' hit columns are downloaded into h1..h3 (TripletThetaPhiFilter.cpp:184)
    TrackletCollection* run(HitCollection& hits, const Grid& grid, const Pairing& pairs,
                            const Pairing& tripletCandidates, int nThreads, const TripletConfigurations& layerTriplets,
                            bool printProlix = false);
    Context& ctx;
};

// what the reference computes inside TripletConnectivityTight::run and
// CellularAutomaton::run but does not return (the eta of every tracklet, the
// event of every connection; the CA states, best bases and iteration count),
// kept when a caller asks for it
struct ConnectivityDiag {
    std::unique_ptr<FVec> eta;
    std::unique_ptr<UVec> event;
};
struct CADiag {
    std::unique_ptr<UVec> states, bestBasis;
    uint iterations = 0;
};

class TripletConnectivityTight {  // src/algorithms/TripletConnectivityTight.h:25-44
public:
    explicit TripletConnectivityTight(Context& c) : ctx(c) {}
    // TripletConnectivityTight.h:42: (basis, followers, connectable indices,
    // non-connectable indices), owned by the caller, each sized exactly
    // (get_count()); all null when no connection exists (.cpp:93-95)
    std::tuple<UVec*, UVec*, UVec*, UVec*> run(const HitCollection& hits, TrackletCollection& tracklets,
                                               const TripletConfigurations& layerTriplets, float dEtaCut,
                                               uint nThreads, bool printPROLIX = false,
                                               ConnectivityDiag* diag = nullptr) const;
    Context& ctx;
};

class TrackletCircleFitter {  // src/algorithms/TrackletCircleFitter.h:17-29
public:
    explicit TrackletCircleFitter(Context& c) : ctx(c) {}
    // TrackletCircleFitter.h:25: the radius of every tracklet listed in
    // validTrackletsIndices (all of its get_count() entries); one entry per
    // tracklet, the others 0
    FVec* run(const HitCollection& hits, const TrackletCollection& tracklets, const UVec& validTrackletsIndices,
              uint nThreads, bool printPROLIX = false) const;
    Context& ctx;
};

class CellularAutomaton {  // src/algorithms/CellularAutomaton.h:22-49
public:
    explicit CellularAutomaton(Context& c) : ctx(c) {}
    // CellularAutomaton.h:43-49: (tracks, track offsets), owned by the
    // caller; the connections are tripletsBasis / tripletsFollowers (all of
    // their entries), the tracklets those of tripletsPt
    std::tuple<const UVec*, const UVec*> run(const UVec& tripletsBasis, const UVec& tripletsFollowers,
                                             const FVec& tripletsPt, const UVec& connectableTriplets, uint nThreads,
                                             bool printPROLIX = false, CADiag* diag = nullptr) const;
    Context& ctx;
};

}  // namespace trax
