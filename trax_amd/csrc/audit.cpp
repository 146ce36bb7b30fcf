// Float64 audit of the filter's float32 decisions, on the host: every stored
// candidate's four cuts of TripletThetaPhiFilter (src/algorithms/
// TripletThetaPhiFilter.h:83-150) evaluated in float64 from the float32 hit
// coordinates, on the host's threads.  The caller compares them with the
// device's decisions; the disagreements are then examined one by one.
//
// Quantities (float64): the theta ratio |theta(h2->h3) / theta(h1->h2)|,
// |dphi| of the two transverse directions wrapped into [-pi, pi], the
// distance of closest approach of the circle through the three transverse
// points to the origin (| |centre| - radius |), and the radius; the circle
// from the paraboloid lift as the filter forms it.  Thresholds are the
// float32 values the kernels compare against: 1 -+ dThetaCut (in float32),
// dPhiCut, tipCut, the minimum radius of curvature.  The layer triplet of a
// candidate is the layer of its first hit minus one, as the filter reads it.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "capi_internal.h"

using namespace trax;
using namespace trax_capi;

extern "C" int trax_filter_audit_f64(const trax_config* cfg, uint64_t n_cands, const uint32_t* cands,
                                     uint64_t n_pairs, const uint32_t* pairs, uint32_t n_hits, const float* x,
                                     const float* y, const float* z, const uint32_t* layer,
                                     float min_radius_curvature, int n_threads, uint8_t* pass) {
    return guarded([&] {
        if (!cfg || !pass || (n_cands && !cands) || (n_pairs && !pairs) || (n_hits && (!x || !y || !z || !layer)))
            return fail(TRAX_ERR_ARG, "null argument");
        const TripletConfigurations& tc = cfg->trip;
        const uint32_t T = tc.size();
        std::vector<double> lo(T), hi(T), phic(T), tipc(T);
        for (uint32_t t = 0; t < T; ++t) {
            lo[t] = (double)(1.0f - tc.dThetaCut[t]);
            hi[t] = (double)(1.0f + tc.dThetaCut[t]);
            phic[t] = (double)tc.dPhiCut[t];
            tipc[t] = (double)tc.tipCut[t];
        }
        const double rmin = (double)min_radius_curvature;
        const int nt = std::max(1, std::min(n_threads, 64));
        std::vector<int> bad(nt, 0);
        auto work = [&](int w) {
            const uint64_t a = n_cands * (uint64_t)w / nt, b = n_cands * (uint64_t)(w + 1) / nt;
            for (uint64_t i = a; i < b; ++i) {
                const uint32_t pi = cands[2 * i], h3 = cands[2 * i + 1];
                if (pi == kSentinel || h3 == kSentinel) {
                    pass[i] = 0x80;  // an empty candidate slot
                    continue;
                }
                if (pi >= n_pairs || h3 >= n_hits) {
                    bad[w] = 1;
                    pass[i] = 0x40;
                    continue;
                }
                const uint32_t h1 = pairs[2 * (uint64_t)pi], h2 = pairs[2 * (uint64_t)pi + 1];
                if (h1 >= n_hits || h2 >= n_hits || layer[h1] == 0 || layer[h1] > T) {
                    bad[w] = 1;
                    pass[i] = 0x40;
                    continue;
                }
                const uint32_t t = layer[h1] - 1;
                const double x1 = x[h1], y1 = y[h1], z1 = z[h1];
                const double x2 = x[h2], y2 = y[h2], z2 = z[h2];
                const double x3 = x[h3], y3 = y[h3], z3 = z[h3];
                const double th1 = std::atan2(std::hypot(x2 - x1, y2 - y1), z2 - z1);
                const double th2 = std::atan2(std::hypot(x3 - x2, y3 - y2), z3 - z2);
                const double ratio = std::fabs(th2 / th1);
                double d = std::atan2(y3 - y2, x3 - x2) - std::atan2(y2 - y1, x2 - x1);
                if (d > M_PI) d -= 2 * M_PI;
                else if (d < -M_PI) d += 2 * M_PI;
                const double q1 = x1 * x1 + y1 * y1, q2 = x2 * x2 + y2 * y2, q3 = x3 * x3 + y3 * y3;
                const double ax = x2 - x1, ay = y2 - y1, az = q2 - q1;
                const double bx = x3 - x1, by = y3 - y1, bz = q3 - q1;
                double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
                const double len = std::sqrt(nx * nx + ny * ny + nz * nz);
                nx /= len;
                ny /= len;
                nz /= len;
                const double cx = -nx / (2 * nz), cy = -ny / (2 * nz);
                const double c = -(nx * x1 + ny * y1 + nz * q1);
                const double radius = std::sqrt((1 - nz * nz - 4 * c * nz) / (4 * (nz * nz)));
                const double tip = std::fabs(std::hypot(cx, cy) - radius);
                uint8_t m = 0;
                if (lo[t] <= ratio && ratio <= hi[t]) m |= 1;
                if (std::fabs(d) <= phic[t]) m |= 2;
                if (tip <= tipc[t]) m |= 4;
                if (radius >= rmin) m |= 8;
                pass[i] = m;
            }
        };
        std::vector<std::thread> th;
        for (int w = 1; w < nt; ++w) th.emplace_back(work, w);
        work(0);
        for (auto& t : th) t.join();
        for (int v : bad)
            if (v) return fail(TRAX_ERR_INPUT, "candidate index out of range");
        return TRAX_OK;
    });
}
