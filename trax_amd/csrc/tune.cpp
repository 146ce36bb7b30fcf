// Autotuner of the batched pipeline's launch shapes.
//
// Every launch shape is a performance choice only: the kernels read sizes
// from the device, tile boundaries come from scans, and the output order is
// fixed by the reference's work-item order, not by the launch.  The tuner
// therefore picks each shape by timing alone, by coordinate descent in a
// fixed order (each parameter in turn, the others at their best so far).
// Each candidate is timed on the phases that parameter drives, with the HIP
// events recorded between the phases on the batch's stream, as the best of
// `repeats` eager runs.  As a guard it checks that every candidate reproduces
// the first run's tracklets and counts, and it fails if one does not.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.h"

using namespace trax;
using namespace trax_capi;

namespace {

struct Param {
    const char* name;
    std::vector<uint32_t> values;        // predictor: (block << 16) | list
    std::vector<const char*> phases;     // empty: the whole sequence
};

void set_param(trax_tuning& t, const char* name, uint32_t v) {
    const std::string n(name);
    if (n == "grid") {
        t.grid_count_block = v >> 16;
        t.grid_stage_kib = (v >> 2) & 0x3fffu;
        t.grid_order = v & 3u;
    }
    else if (n == "grid_count_block") t.grid_count_block = v;
    else if (n == "predict_wave") t.predict_wave = v;
    else if (n == "pair_xcd") t.pair_xcd = v;
    else if (n == "grid_store_block") t.grid_store_block = v;
    else if (n == "pair_block") t.pair_block = v;
    else if (n == "predict") {
        t.predict_block = (v >> 16) & 0x7fffu;
        t.predict_list = v & 0xffffu;
        t.predict_sort = v >> 31;
    }
    else if (n == "filter_block") t.filter_block = v;
    else if (n == "link_block") t.link_block = v;
    else if (n == "blocks_per_cu") t.blocks_per_cu = v;
    else if (n == "predict_fine") t.predict_fine = v;

}

LaunchCfg to_launch(const trax_tuning& t) { return launch_cfg(&t); }

uint64_t mix(uint64_t h, uint64_t v) {
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h * 0xff51afd7ed558ccdull;
}

// digest of a completed run: the counts and every tracklet
uint64_t run_digest(trax_batch* b) {
    const DevCounts& c = *b->h_counts;
    uint64_t h = 0;
    for (uint32_t v : {c.n_pairs, c.n_cands, c.n_tracklets, c.n_conn, c.n_track_entries, c.err}) h = mix(h, v);
    const uint32_t R = std::min(c.n_tracklets, b->caps.max_tracklets);
    std::vector<uint32_t> t(R);
    for (const uint32_t* col : {b->th1, b->th2, b->th3}) {
        if (R) hip_check(hipMemcpy(t.data(), col, R * 4ull, hipMemcpyDeviceToHost), "d2h");
        for (uint32_t i = 0; i < R; i += 1) h = mix(h, t[i]);
    }
    return h;
}

float phases_ms(trax_batch* b, const std::vector<const char*>& phases) {
    float sum = 0.0f;
    if (phases.empty()) {
        hip_check(hipEventElapsedTime(&sum, b->ev[0], b->ev[kNumStages]), "event time");
        return sum;
    }
    for (const char* p : phases)
        for (int k = 0; k < kNumStages; ++k)
            if (std::strcmp(p, kPhases[k]) == 0) {
                float ms = 0.0f;
                hip_check(hipEventElapsedTime(&ms, b->ev[k], b->ev[k + 1]), "event time");
                sum += ms;
            }
    return sum;
}

}  // namespace

extern "C" int trax_autotune(trax_ctx* ctx, const trax_config* cfg, const trax_hits* sample,
                             const trax_batch_caps* caps_in, uint32_t repeats, trax_tuning* best, char* report,
                             uint64_t report_bytes) {
    return guarded([&] {
        if (!ctx || !cfg || !sample || !caps_in || !best) return fail(TRAX_ERR_ARG, "null argument");
        if (repeats == 0) repeats = 1;
        trax_batch_caps caps = *caps_in;
        caps.use_graph = 0;  // the launch shape changes per trial; eager launches read it each time
        trax_batch* b = nullptr;
        trax_tuning cur{};
        int rc = trax_batch_create(ctx, cfg, &caps, &cur, nullptr, &b);
        if (rc != TRAX_OK) return rc;
        std::unique_ptr<trax_batch, void (*)(trax_batch*)> hold(b, trax_batch_destroy);
        rc = trax_batch_submit(b, sample);
        if (rc != TRAX_OK) return rc;
        rc = b->finish();
        if (rc != TRAX_OK) return rc;
        const uint64_t ref = run_digest(b);
        const std::vector<Param> params = {
            {"predict", {(128u << 16) | 32u, (128u << 16) | 64u, (256u << 16) | 24u, (256u << 16) | 32u,
                         (256u << 16) | 48u, (256u << 16) | 64u, (512u << 16) | 24u, (512u << 16) | 32u,
                         1u << 31 | (256u << 16) | 32u}, {"predict", "filter_count"}},
            {"predict_fine", {4, 8, 16}, {"predict", "filter_count"}},
            {"predict_wave", {1, 2, 4, 128}, {"predict", "filter_count"}},
            {"filter_block", {64, 128, 256, 512, 1024}, {"filter_count", "filter_store"}},
            {"pair_block", {64, 128, 256, 512}, {"pair_count", "pair_store"}},
            {"pair_xcd", {1, 8}, {"pair_count", "pair_store"}},
            {"link_block", {64, 128, 256, 512}, {"link_prepare", "link_count", "link_store"}},
            // the grid build: (block, LDS KiB, item order) together
            {"grid", {1024u << 16 | 160u << 2 | 1u, 512u << 16 | 160u << 2 | 1u, 512u << 16 | 80u << 2 | 1u,
                      1024u << 16 | 64u << 2 | 1u, 512u << 16 | 64u << 2 | 1u, 256u << 16 | 64u << 2 | 2u},
             {"grid_count"}},
            {"blocks_per_cu", {2, 4, 8, 16}, {}},
        };
        std::string log;
        bool changed = false;
        char first_change[96] = "";
        for (const Param& p : params) {
            uint32_t best_v = 0;
            float best_ms = 0.0f;
            // the fused predictor + filter has no sorted-walk variant and no separate filter kernels
            if (caps.fuse_filter && std::strcmp(p.name, "filter_block") == 0) continue;
            for (uint32_t v : p.values) {
                if (std::strcmp(p.name, "predict_fine") == 0 && v > fine_max(b->cfg.grid.geom())) continue;
                if (caps.fuse_filter && std::strcmp(p.name, "predict") == 0 && (v >> 31)) continue;
                trax_tuning t = cur;
                set_param(t, p.name, v);
                b->lc = to_launch(t);
                predict_fused_prepare(b->lc);
                grid_build_prepare();
                float ms = 0.0f;
                bool same = true;
                for (uint32_t r = 0; r < repeats; ++r) {
                    b->launch();
                    rc = b->finish();
                    if (rc != TRAX_OK) return rc;
                    const float m = phases_ms(b, p.phases);
                    ms = r == 0 ? m : std::min(ms, m);
                    if (r == 0) same = run_digest(b) == ref;
                }
                if (!same && !changed) {
                    std::snprintf(first_change, sizeof(first_change), "%s = %u", p.name,
                                  p.name == std::string("predict")
                                      ? (v >> 31) * 1000000 + ((v >> 16) & 0x7fffu) * 1000 + (v & 0xffffu)
                                      : v);
                }
                changed |= !same;
                char line[256];
                std::snprintf(line, sizeof(line),
                              "{\"param\": \"%s\", \"value\": %u, \"ms\": %.4f, \"same_results\": %s}\n", p.name,
                              p.name == std::string("predict")
                                  ? (v >> 31) * 1000000 + ((v >> 16) & 0x7fffu) * 1000 + (v & 0xffffu)
                                  : v,
                              ms,
                              same ? "true" : "false");
                log += line;
                if (best_v == 0 || ms < best_ms) {
                    best_v = v;
                    best_ms = ms;
                }
            }
            set_param(cur, p.name, best_v);
        }
        b->lc = to_launch(cur);
        *best = cur;
        if (report && report_bytes) {
            const size_t n = std::min<size_t>(log.size(), report_bytes - 1);
            std::memcpy(report, log.data(), n);
            report[n] = 0;
        }
        if (changed) {
            const std::string msg = std::string("a launch shape changed the results (first: ") + first_change + ")";
            return fail(TRAX_ERR_STATE, msg.c_str());
        }
        return TRAX_OK;
    });
}
