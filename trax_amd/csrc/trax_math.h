// Float32 built-ins shared by the gfx950 kernels and the host code.
//
// The triplet-seeding kernels of trax call the OpenCL built-ins atan2, acos,
// atan, asinh, sqrt, floor and fabs (e.g. src/algorithms/GridBuilder.h:81,
// PairGeneratorBeamspot.h:138-145, TripletThetaPhiFilter.h:84-152,
// TripletConnectivityTight.h:73-75).  OpenCL leaves the accuracy of the
// transcendental ones implementation-defined (atan2 <= 6 ulp, acos/atan <= 4-5
// ulp).  This header is one such implementation, written to be evaluated
// bit-identically by the device compiler (hipcc, gfx950) and by a host C++
// compiler: every operation is a single IEEE-754 binary32 add, multiply,
// fused multiply-add, divide or square root (all correctly rounded on both
// sides: fmaf is the C99 function, one rounding of x*y+z, and v_fma_f32 on the
// device), with implicit FMA contraction disabled (-ffp-contract=off on both
// compilers, and the pragma below for clang), so that only the fmaf calls
// written here fuse.  sqrt and floor are the correctly rounded / exact C
// functions.
//
// Each function reduces its argument to a small interval and evaluates
// s + s*t*P(t), t = s*s, P by Horner's rule with one fmaf per step.  The coefficients of P are weighted-minimax fits
// derived by fit_math_coeffs.py in this directory (Lawson iteration on a dense
// Chebyshev grid, relative-error weight).  Measured accuracy over dense sweeps
// against float64: atan <= 1.2 ulp, atan2 <= 2 ulp, acos <= 1.5 ulp,
// asinh <= 4 ulp (asserted by the unit tests).
//
// The float -> integer conversions of the reference kernels ((uint)floor(x),
// (int)floor(x)) are undefined in C for out-of-range values; the AMD (and
// NVIDIA) GPU conversion instructions saturate and send NaN to 0.  tx_f2u_sat
// and tx_f2i_sat spell that behaviour out so that host and device agree.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define TX_HD __host__ __device__ inline
#else
#define TX_HD static inline
#endif

// float nearest to pi (OpenCL M_PI_F) and the remainder pi - kPi
#define TX_PI_F 3.14159274101257324e+00f
#define TX_PI_LO (-8.74227766e-08f)
#define TX_PIO2_F 1.57079637050628662e+00f
#define TX_PIO2_LO (-4.37113883e-08f)
#define TX_LN2_HI 6.93145751953125e-01f
#define TX_LN2_LO 1.42860677e-06f
#define TX_LN2_F 6.93147182e-01f

TX_HD float tx_bits2f(uint32_t u) { return __builtin_bit_cast(float, u); }
TX_HD uint32_t tx_f2bits(float f) { return __builtin_bit_cast(uint32_t, f); }
TX_HD float tx_copysign(float mag, float sgn) {
    return tx_bits2f((tx_f2bits(mag) & 0x7fffffffu) | (tx_f2bits(sgn) & 0x80000000u));
}
TX_HD bool tx_signbit(float f) { return (tx_f2bits(f) >> 31) != 0u; }
TX_HD bool tx_isnan(float f) { return (tx_f2bits(f) & 0x7fffffffu) > 0x7f800000u; }

// saturating float -> uint32 / int32 conversion, NaN -> 0 (GPU semantics)
TX_HD uint32_t tx_f2u_sat(float f) {
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)f;
}
TX_HD int32_t tx_f2i_sat(float f) {
    if (tx_isnan(f)) return 0;
    if (f >= 2147483648.0f) return 2147483647;
    if (f <= -2147483648.0f) return (-2147483647 - 1);
    return (int32_t)f;
}

// atan(a) for a in [0, 1]: a + a*t*P(t), P of degree 8
TX_HD float tx_atan_core(float a) {
    const float t = a * a;
    float p = -1.793621108e-03f;
    p = fmaf(p, t, 1.091460325e-02f);
    p = fmaf(p, t, -3.117784299e-02f);
    p = fmaf(p, t, 5.795758963e-02f);
    p = fmaf(p, t, -8.403450251e-02f);
    p = fmaf(p, t, 1.095218584e-01f);
    p = fmaf(p, t, -1.426424235e-01f);
    p = fmaf(p, t, 1.999854892e-01f);
    p = fmaf(p, t, -3.333329856e-01f);
    return fmaf(a * t, p, a);
}

// asin(s) for s in [0, 0.5], with t = s*s supplied: s + s*t*P(t), degree 4
TX_HD float tx_asin_core(float s, float t) {
    float p = 4.216632247e-02f;
    p = fmaf(p, t, 2.417950705e-02f);
    p = fmaf(p, t, 4.547037557e-02f);
    p = fmaf(p, t, 7.495297492e-02f);
    p = fmaf(p, t, 1.666675210e-01f);
    return fmaf(s * t, p, s);
}

// atanh(s) for |s| <= 3 - 2*sqrt(2): s + s*t*P(t), degree 2
TX_HD float tx_atanh_core(float s) {
    const float t = s * s;
    float p = 1.493553668e-01f;
    p = fmaf(p, t, 1.998878419e-01f);
    p = fmaf(p, t, 3.333338797e-01f);
    return fmaf(s * t, p, s);
}

// OpenCL normalize (its accuracy is implementation-defined): the squared
// length, one correctly rounded square root and one correctly rounded
// reciprocal of it, then one multiply per component.  A zero vector is left
// as it is.
TX_HD void tx_normalize3(float& x, float& y, float& z) {
    const float d = x * x + y * y + z * z;
    if (d != 0.0f) {
        const float r = 1.0f / sqrtf(d);
        x = x * r;
        y = y * r;
        z = z * r;
    }
}
TX_HD void tx_normalize2(float& x, float& y) {
    const float d = x * x + y * y;
    if (d != 0.0f) {
        const float r = 1.0f / sqrtf(d);
        x = x * r;
        y = y * r;
    }
}

TX_HD float tx_atan2f(float y, float x) {
    if (tx_isnan(x) || tx_isnan(y)) return x + y;
    const float ax = fabsf(x), ay = fabsf(y);
    const bool swap = ay > ax;
    const float mx = swap ? ay : ax;
    const float mn = swap ? ax : ay;
    float r;
    if (mx == 0.0f) {
        r = 0.0f;  // atan2(+-0, +-0)
    } else if (mn == INFINITY) {
        r = 7.85398185253143311e-01f;  // both infinite: pi/4
    } else {
        r = tx_atan_core(mn / mx);
    }
    if (swap) r = (TX_PIO2_F - r) + TX_PIO2_LO;
    if (tx_signbit(x)) r = (TX_PI_F - r) + TX_PI_LO;
    return tx_copysign(r, y);
}

TX_HD float tx_atanf(float x) {
    if (tx_isnan(x)) return x;
    const float ax = fabsf(x);
    float r;
    if (ax <= 1.0f) {
        r = tx_atan_core(ax);
    } else {
        r = (TX_PIO2_F - tx_atan_core(1.0f / ax)) + TX_PIO2_LO;
    }
    return tx_copysign(r, x);
}

TX_HD float tx_acosf(float x) {
    const float ax = fabsf(x);
    if (!(ax <= 1.0f)) return (x - x) / (x - x);  // NaN for |x| > 1 and for NaN
    if (ax <= 0.5f) {
        const float as = tx_asin_core(x, x * x);
        return (TX_PIO2_F - as) + TX_PIO2_LO;
    }
    const float z = (1.0f - ax) * 0.5f;
    const float s = sqrtf(z);
    const float as = tx_asin_core(s, z);
    if (x > 0.0f) return 2.0f * as;
    return (TX_PI_F - 2.0f * as) + TX_PI_LO;
}

// natural log of a positive, finite, normal float
TX_HD float tx_log_pos(float w) {
    const uint32_t u = tx_f2bits(w);
    int k = (int)((u >> 23) & 0xffu) - 127;
    float m = tx_bits2f((u & 0x007fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 1.41421354f) {
        m = m * 0.5f;
        k += 1;
    }
    const float f = m - 1.0f;  // exact
    const float s = f / (2.0f + f);
    const float lm = 2.0f * tx_atanh_core(s);
    const float fk = (float)k;
    return fk * TX_LN2_HI + (fk * TX_LN2_LO + lm);
}

// log(1 + u) for u >= 0 (Goldberg 1991, theorem 4: log(w) * u / (w - 1))
TX_HD float tx_log1p_nonneg(float u) {
    const float w = 1.0f + u;
    if (w == 1.0f) return u;
    return tx_log_pos(w) * (u / (w - 1.0f));
}

TX_HD float tx_asinhf(float x) {
    if (tx_isnan(x)) return x;
    const float a = fabsf(x);
    if (a < 2.44140625e-04f) return x;  // |x| < 2^-12: asinh(x) = x to < 0.5 ulp
    float r;
    if (a > 4096.0f) {
        r = (a == INFINITY) ? a : tx_log_pos(a) + TX_LN2_F;
    } else {
        const float a2 = a * a;
        r = tx_log1p_nonneg(a + a2 / (1.0f + sqrtf(1.0f + a2)));
    }
    return tx_copysign(r, x);
}

// ---- filter phi-cut prefilter ------------------------------------------------
// The filter's phi cut (TripletThetaPhiFilter.h:95-103) accepts a candidate
// when |wrap(atan2(v) - atan2(u))| <= dPhiCut for the float hit differences
// u = p2 - p1, v = p3 - p2.  That float expression differs from the exact
// angle from u to v by less than 2e-6 rad (two atan2 of <= 2 ulp of pi, the
// subtraction and the wrap by 2 pi).  In double the cross and dot products of
// u and v are exact up to one rounding each, so
//     d <= 0  or  |c| > tan(dPhiCut + 1e-5) * d
// proves the exact angle exceeds the cut by more than 1e-5 rad, and with it
// that the reference's expression rejects the candidate.  Candidates it does
// not decide (near the cut, tiny or non-finite vectors, wide cuts) go
// through the reference's expression; the decision is the same either way.
// tan_cut: tan(dPhiCut + 1e-5) in double, or a non-positive value to disable
// (cuts with dPhiCut + 1e-5 >= 1.5 rad).
TX_HD bool tx_phi_surely_outside(float ux, float uy, float vx, float vy, double tan_cut) {
    if (!(tan_cut > 0.0)) return false;
    const float big = 3.0e38f;
    if (!(fabsf(ux) < big && fabsf(uy) < big && fabsf(vx) < big && fabsf(vy) < big)) return false;
    // zero, denormal or tiny vectors: left to the reference expression
    const float tiny = 1.0e-30f;
    if (fmaxf(fabsf(ux), fabsf(uy)) < tiny || fmaxf(fabsf(vx), fabsf(vy)) < tiny) return false;
    const double c = (double)ux * (double)vy - (double)uy * (double)vx;
    const double d = (double)ux * (double)vx + (double)uy * (double)vy;
    return d <= 0.0 || fabs(c) > tan_cut * d;
}

// host-side value of tan_cut for a phi cut (see tx_phi_surely_outside)
static inline double tx_phi_prefilter_tan(float d_phi_cut) {
    const double a = (double)d_phi_cut + 1e-5;
    return (a > 0.0 && a < 1.5) ? tan(a) : -1.0;
}
