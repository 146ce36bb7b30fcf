// Internal definitions shared by the C ABI translation units (capi.cpp,
// serve.cpp): the opaque handle types of include/trax_hip.h and the batched
// pipeline's workspace.
#pragma once

#include "../../include/trax_hip.h"

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "host/trax_host.h"
#include "kernels/kernels.h"

namespace trax_capi {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

// runs f, turning exceptions into status codes (no exception crosses the ABI)
template <class F>
int guarded(F f) {
    try {
        return f();
    } catch (const trax::HipError& e) {
        return fail(TRAX_ERR_HIP, e.what());
    } catch (const trax::InputError& e) {
        return fail(TRAX_ERR_INPUT, e.what());
    } catch (const trax::CapacityError& e) {
        return fail(TRAX_ERR_CAPACITY, e.what());
    } catch (const std::exception& e) {
        return fail(TRAX_ERR_ARG, e.what());
    }
}

struct Arr {
    std::vector<uint32_t> u;
    std::vector<float> f;
    bool is_float = false;
};

// bump allocator over one device allocation; every buffer 256-byte aligned
struct Arena {
    uint8_t* base = nullptr;
    size_t used = 0, cap = 0;
    template <class T>
    T* take(uint64_t n) {
        size_t bytes = (size_t)std::max<uint64_t>(n, 1) * sizeof(T);
        bytes = (bytes + 255) & ~(size_t)255;
        T* p = reinterpret_cast<T*>(base + used);
        used += bytes;
        if (used > cap) throw trax::CapacityError("arena overflow");
        return p;
    }
};

// phases of the batched launch sequence: one graph each, HIP events between
// them time every phase on the batch's stream
constexpr int kNumStages = 16;
constexpr int kGridStage = 0, kPredictStage = 6;  // the grid build, and the stage that reads the fine walk structure
extern const char* const kPhases[kNumStages];
// the reference's algorithm classes as ranges of phases
struct StageRange {
    const char* name;
    int first, last;
};
extern const StageRange kStageRanges[6];

void validate_config(const struct ::trax_config* cfg);

// launch shapes of a trax_tuning (NULL: the defaults)
inline trax::LaunchCfg launch_cfg(const trax_tuning* t) {
    trax::LaunchCfg lc;
    if (!t) return lc;
    lc.grid_count_block = t->grid_count_block;
    lc.grid_store_block = t->grid_store_block;
    lc.pair_block = t->pair_block;
    lc.predict_block = t->predict_block;
    lc.filter_block = t->filter_block;
    lc.link_block = t->link_block;
    lc.blocks_per_cu = t->blocks_per_cu;
    lc.predict_list = t->predict_list;
    lc.predict_sort = t->predict_sort;
    lc.predict_fine = t->predict_fine;
    lc.grid_order = t->grid_order;
    lc.grid_stage_kib = t->grid_stage_kib;
    lc.predict_wave = t->predict_wave;
    lc.pair_xcd = t->pair_xcd;
    return lc;
}

}  // namespace trax_capi

struct trax_result {
    std::map<std::string, trax_capi::Arr> arrays;
    std::map<std::string, double> timings;
    void put(const std::string& k, std::vector<uint32_t> v) {
        trax_capi::Arr a; a.u = std::move(v); arrays[k] = std::move(a);
    }
    void putf(const std::string& k, std::vector<float> v) {
        trax_capi::Arr a; a.f = std::move(v); a.is_float = true; arrays[k] = std::move(a);
    }
};

struct trax_ctx {
    std::unique_ptr<trax::Context> c;
};
struct trax_stream {
    hipStream_t s = nullptr;
};
struct trax_timer {
    hipEvent_t start = nullptr, stop = nullptr;
};

struct trax_config {
    trax::GridConfig grid;
    trax::TripletConfigurations trip{1.0f};
    trax::GeometrySupplement geom;
    uint32_t n_threads = 256;
    float d_eta_cut = 0.0256f;
};

struct trax_hits {
    uint32_t max_layer = 0;
    uint32_t n_events = 0;
    trax::HitCollection hits;
    std::vector<uint32_t> evt_nhits, evt_offset, lay_nhits, lay_offset;
    // the findable tracks of every event with their hits in layer order, as
    // HitCollection::addEvent returns them (the mcTruth of PhysicsRecord::fillData)
    std::vector<trax::HitCollection::tTrackList> truth;
};

inline hipStream_t pick_stream(trax_ctx* ctx, trax_stream* s) { return s ? s->s : ctx->c->stream(); }

// The batched pipeline's workspace: every buffer of the path for up to
// caps.max_events events in one arena, sizes kept on the device (DevCounts),
// the launch sequence captured into one graph (or one per phase).
struct trax_batch {
    trax_ctx* owner = nullptr;
    hipStream_t st = nullptr;
    trax_config cfg;
    trax_batch_caps caps{};
    trax::LaunchCfg lc;
    uint32_t E = 0, L = 0, T = 0, C = 0, slot_cap = 0, scan_cap = 0, ca_iter = 0;
    bool ca_layered = false;  // CA forward phase by layer triplets (launch_ca_layers)
    std::unique_ptr<trax::ConfigMirror> cm;
    trax::DevCounts* cnt = nullptr;
    trax::ScanScratch scan;
    trax_capi::Arena arena;
    uint8_t* arena_mem = nullptr;
    uint32_t* staging = nullptr;  // pinned: 7 hit columns, supplements, shape
    size_t staging_words = 0;
    trax::DevCounts* h_counts = nullptr;  // pinned copy after a run
    uint32_t* h_flags = nullptr;          // pinned: the CA's last three iteration flags
    // device buffers
    trax::HitColumns hin, hout;
    uint32_t *sup = nullptr, *grid = nullptr, *seg_base = nullptr, *slot_seg = nullptr, *slot_counts = nullptr,
             *pairs = nullptr,
             *pair_off = nullptr, *pred_counts = nullptr, *cands = nullptr, *cand_off = nullptr,
             *word_counts = nullptr, *th1 = nullptr, *th2 = nullptr, *th3 = nullptr, *t_off = nullptr;
    uint64_t* bitmap = nullptr;
    trax::LinkBufs lb;
    trax::CABufs cb;
    trax::FineBufs fine;  // the fused predictor's walk structure
    uint32_t* grid_perm = nullptr;  // the grid build's scatter scratch (N), then its item order (E * L)
    float4* grid_rec = nullptr;     // the grid build's 32-byte hit records (2 N float4)
    hipGraph_t graphs[trax_capi::kNumStages] = {};
    hipGraphExec_t execs[trax_capi::kNumStages] = {};
    bool captured = false;
    hipEvent_t ev[trax_capi::kNumStages + 1] = {};
    // the predictor's fine walk structure is built on a second stream, beside
    // the pair stages: it needs only the grid build (ev_grid), and the
    // predict stage waits for it (ev_fine).  Without one (single_stream())
    // it is built on st after the grid build.
    hipStream_t st2 = nullptr;
    hipStream_t side() const { return st2 ? st2 : st; }
    // drop the second stream (before the first launch): every kernel of the
    // batch then runs on st
    void single_stream();
    hipEvent_t ev_grid = nullptr, ev_fine = nullptr;
    hipGraph_t fine_graph = nullptr;
    hipGraphExec_t fine_exec = nullptr;
    // the whole sequence (every phase, the fork to st2 and the join) as one
    // graph, launched with one call between the first and last phase events
    // (caps.use_graph == 2); the per-phase graphs above are the fallback when
    // the runtime does not capture that form
    hipGraph_t whole_graph = nullptr;
    hipGraphExec_t whole_exec = nullptr;
    bool whole = false;
    uint32_t n_events = 0, n_hits = 0;
    bool submitted = false, have_counts = false;

    void enqueue_stage(int s);
    void enqueue_fine(hipStream_t s);
    void enqueue_ca_tail();
    // carve every buffer out of the arena (run on a measuring arena first)
    void layout(trax_capi::Arena& a);
    // enqueue the whole sequence on the resident inputs (graph replay or eager)
    void launch();
    // the sequence as launch() enqueues it eagerly: phases, fork and join, and
    // (phase_events) an event before every phase and after the last
    void enqueue_all(bool phase_events);
    // capture enqueue_all() into whole_exec; false (and nothing kept) when the capture fails
    bool capture_whole();
    // after the stream has drained: continue the CA forward phase on the host
    // while the fixed number of iterations left living cells, then rerun the
    // backward phase; returns the status of the run from the device counts
    int finish();
};
