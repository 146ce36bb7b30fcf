// The cuts of TripletThetaPhiFilter (src/algorithms/TripletThetaPhiFilter.h:81-152)
// on one candidate, shared by the filter kernels and the fused predictor.
#pragma once

#include "common.h"
#include "kernels.h"

namespace trax {

// The pair terms of the cuts (TripletThetaPhiFilter.h:99-100, :84-88): the phi and
// theta angles of the first hit pair, the same for every candidate of a pair.
__device__ inline float filter_pair_phi(float x1, float y1, float x2, float y2) {
    return tx_atan2f((y2 - y1), (x2 - x1));
}
__device__ inline float filter_pair_theta(float x1, float y1, float z1, float x2, float y2, float z2) {
    return tx_atan2f(sqrtf((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1)), (z2 - z1));
}

// The phi cut, TripletThetaPhiFilter.h:98-107, given the pair's phi angle.
__device__ inline bool filter_phi_cut(const TripletCols& tc, uint32_t lt, float x2, float y2, float x3, float y3,
                                      float phi1) {
    const float angle2 = tx_atan2f((y3 - y2), (x3 - x2));
    float delta = angle2 - phi1;
    const float PI = TX_PI_F;
    delta += (delta > PI) ? -2 * PI : (delta < -PI) ? 2 * PI : 0;
    return fabsf(delta) <= tc.dphi[lt];
}

// The theta cut, TripletThetaPhiFilter.h:83-96, given the pair's theta angle.
__device__ inline bool filter_theta_cut(const TripletCols& tc, uint32_t lt, float x2, float y2, float z2, float x3,
                                        float y3, float z3, float theta1) {
    const float dThetaCut = tc.dtheta[lt];
    const float angle2 = tx_atan2f(sqrtf((x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2)), (z3 - z2));
    const float delta = fabsf(angle2 / theta1);
    return 1 - dThetaCut <= delta && delta <= 1 + dThetaCut;
}

// The circle cuts, TripletThetaPhiFilter.h:109-150: the transverse impact
// parameter and the radius of the circle through the three hits (from the
// plane through their lifts onto the paraboloid z = x^2 + y^2).
__device__ inline bool filter_circle_cuts(const TripletCols& tc, uint32_t lt, float x1, float y1, float x2, float y2,
                                          float x3, float y3) {
    const float tipCut = tc.tip[lt];
    const float minRadius = tc.min_radius_curvature;
    bool valid = true;
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

// The theta and circle cuts, TripletThetaPhiFilter.h:83-96 and 109-150, given the
// pair's theta angle.  With the phi cut these are a conjunction of pure
// expressions: a candidate is rejected at the first cut it fails, and the
// cuts may be tested in any order (the result is the reference's).
__device__ inline bool filter_theta_circle_cuts(const TripletCols& tc, uint32_t lt, float x1, float y1, float x2,
                                                float y2, float z2, float x3, float y3, float z3, float theta1) {
    return filter_theta_cut(tc, lt, x2, y2, z2, x3, y3, z3, theta1) &&
           filter_circle_cuts(tc, lt, x1, y1, x2, y2, x3, y3);
}

// TripletThetaPhiFilter.h:81-152 on one candidate: the phi prefilter (most
// candidates fail the phi cut by far, which it proves without atan2), then
// the cuts
__device__ inline bool filter_valid(const TripletCols& tc, const float* hx, const float* hy, const float* hz,
                                    uint32_t h1, uint32_t h2, uint32_t h3, uint32_t lt) {
    const float x1 = hx[h1], y1 = hy[h1], z1 = hz[h1];
    const float x2 = hx[h2], y2 = hy[h2], z2 = hz[h2];
    const float x3 = hx[h3], y3 = hy[h3], z3 = hz[h3];
    if (tx_phi_surely_outside(x2 - x1, y2 - y1, x3 - x2, y3 - y2, tc.tan_dphi[lt])) return false;
    if (!filter_phi_cut(tc, lt, x2, y2, x3, y3, filter_pair_phi(x1, y1, x2, y2))) return false;
    return filter_theta_circle_cuts(tc, lt, x1, y1, x2, y2, z2, x3, y3, z3,
                                    filter_pair_theta(x1, y1, z1, x2, y2, z2));
}

}  // namespace trax
