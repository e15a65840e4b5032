// Scalar/vector arithmetic of the resolve2d step, usable on host and device.
//
// Bit-exactness contract: every expression keeps the operation order of the
// reference (src/core/nmath.zig, collision.zig, Bodies/*, Constraints/*), the
// library is compiled with -ffp-contract=off, and f32 division / sqrt are the
// correctly rounded ones.  sin/cos follow the musl/msun method that Zig's
// compiler_rt ports (double-precision reduction + __sindf/__cosdf kernels),
// so host and GPU produce the same bits.
//
// The sin/cos kernels, their coefficients and the reduction thresholds below
// follow FreeBSD msun (k_sinf.c, k_cosf.c, e_rem_pio2f.c, s_sinf.c, s_cosf.c;
// as carried by musl), which carry this notice:
//
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this
//   software is freely granted, provided that this notice
//   is preserved.
//
//   (float versions: conversion to float by Ian Lance Taylor, Cygnus
//   Support; optimized by Bruce D. Evans.)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define R2D_HD __host__ __device__ __forceinline__

namespace r2d {

struct v2 {
    float x, y;
};

R2D_HD v2 mk(float x, float y) { v2 r; r.x = x; r.y = y; return r; }
// nmath.zig:58-93
R2D_HD v2 add2(v2 a, v2 b) { return mk(a.x + b.x, a.y + b.y); }
R2D_HD v2 sub2(v2 a, v2 b) { return mk(a.x - b.x, a.y - b.y); }
R2D_HD v2 scale2(v2 a, float s) { return mk(a.x * s, a.y * s); }
R2D_HD float dot2(v2 a, v2 b) { return a.x * b.x + a.y * b.y; }
R2D_HD float cross2(v2 a, v2 b) { return a.x * b.y - a.y * b.x; }
R2D_HD float length2sq(v2 a) { return dot2(a, a); }
R2D_HD float sqrt_rn(float x) { return __builtin_sqrtf(x); }
R2D_HD float length2(v2 a) { return sqrt_rn(length2sq(a)); }
// nmath.zig:95-103
R2D_HD v2 normalize2(v2 a)
{
    float len = length2(a);
    if (len < 1e-3f) return mk(0.0f, 0.0f);
    return scale2(a, 1.0f / len);
}
R2D_HD v2 negate2(v2 a) { return mk(-a.x, -a.y); }
// Vector2.negate (nmath.zig:48-51) multiplies by -1
R2D_HD v2 negself(v2 a) { return mk(a.x * -1.0f, a.y * -1.0f); }
R2D_HD v2 addmult2(v2 a, v2 b, float s) { return add2(a, scale2(b, s)); }
R2D_HD bool approx(float a, float b, float eps) { return a > b - eps && a < b + eps; }
R2D_HD bool approx2(v2 a, v2 b, float eps) { return approx(a.x, b.x, eps) && approx(a.y, b.y, eps); }
R2D_HD v2 rot90cw(v2 a) { return mk(a.y, -a.x); }
R2D_HD v2 rot90ccw(v2 a) { return mk(-a.y, a.x); }
// rotate2 with a precomputed (cos, sin) pair of the same angle (nmath.zig:129-136)
R2D_HD v2 rotcs(v2 a, float c, float s) { return mk(a.x * c - a.y * s, a.x * s + a.y * c); }

// @max / @min: LLVM maxnum/minnum; NaN-suppressing; equal operands return the first.
R2D_HD bool isnanf_(float a) { return a != a; }
// One select each: a NaN gives b, a NaN b fails the comparison and gives a,
// and equal operands give a.
R2D_HD float maxn(float a, float b) { return (b > a || isnanf_(a)) ? b : a; }
R2D_HD float minn(float a, float b) { return (b < a || isnanf_(a)) ? b : a; }
// std.math.clamp
R2D_HD float clampn(float v, float lo, float hi) { return maxn(lo, minn(v, hi)); }
R2D_HD float absf(float a) { return __builtin_fabsf(a); }

// simulation_constants.zig:3-22
constexpr float MIN_MANIFOLD_IMPULSE = 1e-4f;
constexpr float BAUMGARTE = 0.02f;
constexpr float BAUMGARTE_SLOP = 0.005f;
constexpr float SAT_OVERLAP_THRESHOLD = 1e-4f;
constexpr float COLLISION_MARGIN = 0.01f;
constexpr float AABB_EPS_OVERLAP = 0.01f;
constexpr float GRAD_DIV_LIMIT = 1e-4f;
constexpr float ALLOWED_CONSTRAINT_VALUE = 1e-6f;

// ------------------------------------------------------------------ trig
R2D_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }

// msun __sindf: |sin(x)/x - s(x)| < 2**-37.5 on [-pi/4, pi/4]
R2D_HD float k_sindf(double x)
{
    const double S1 = -0x15555554cbac77.0p-55, S2 = 0x111110896efbb2.0p-59,
                 S3 = -0x1a00f9e2cae774.0p-65, S4 = 0x16cd878c3b46a7.0p-71;
    double z = x * x;
    double w = z * z;
    double r = S3 + z * S4;
    double s = z * x;
    return (float)((x + s * (S1 + z * S2)) + s * w * r);
}

// msun __cosdf: |cos(x) - c(x)| < 2**-34.1 on [-pi/4, pi/4]
R2D_HD float k_cosdf(double x)
{
    const double C0 = -0x1ffffffd0c5e81.0p-54, C1 = 0x155553e1053a42.0p-57,
                 C2 = -0x16c087e80f1e27.0p-62, C3 = 0x199342e0ee5069.0p-68;
    double z = x * x;
    double w = z * z;
    double r = C2 + z * C3;
    return (float)(((1.0 + z * C0) + w * C1) + (w * z) * r);
}

// high 64 bits of a 64x64-bit product, from 32-bit halves
R2D_HD uint64_t mulhi64(uint64_t a, uint64_t b)
{
    uint64_t a0 = a & 0xffffffffu, a1 = a >> 32, b0 = b & 0xffffffffu, b1 = b >> 32;
    uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    uint64_t mid = (p00 >> 32) + (p01 & 0xffffffffu) + (p10 & 0xffffffffu);
    return p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// Huge |x| (>= 2^28*pi/2): exact reduction with a 128-bit window of 2/pi.
// Out of line: it is rare and large.
static __host__ __device__ __noinline__ int rem_pio2_huge(float x, double* y)
{
    // bits of 2/pi after the binary point, 32 per word
    const uint32_t tp[12] = {0xa2f9836eu, 0x4e441529u, 0xfc2757d1u, 0xf534ddc0u, 0xdb629599u, 0x3c439041u,
                             0xfe5163abu, 0xdebbc561u, 0xb7246e3au, 0x424dd2e0u, 0x06492eeau, 0x09d1921cu};
    uint32_t u = f2u(x);
    int sign = (int)(u >> 31);
    uint32_t ix = u & 0x7fffffffu;
    uint64_t m = (uint64_t)((ix & 0x7fffffu) | 0x800000u);
    int E = (int)(ix >> 23) - 150;
    int s = E - 1;
    uint64_t whi = 0, wlo = 0;
    for (int i = 0; i < 128; ++i) {
        int bit = s + i;
        int word = (bit - 1) >> 5, off = (bit - 1) & 31;
        uint32_t b = (tp[word] >> (31 - off)) & 1u;
        if (i < 64) whi = (whi << 1) | b; else wlo = (wlo << 1) | b;
    }
    // P = m * (whi:wlo), fraction bits below bit 126
    uint64_t lo_hi = mulhi64(m, wlo);
    uint64_t hi_lo = m * whi;
    uint64_t p1 = hi_lo + lo_hi;   // bits 64..127 (carry into 128+ irrelevant mod 4)
    int n = (int)((p1 >> 62) & 3u);
    uint64_t frac = p1 << 2;
    double f = (double)(frac >> 11) * 0x1p-53;
    if (f >= 0.5) { f -= 1.0; n = (n + 1) & 3; }
    double r = f * 1.57079632679489661923;
    if (sign) { *y = -r; return -n; }
    *y = r;
    return n;
}

// msun __rem_pio2f
R2D_HD int rem_pio2f(float x, double* y)
{
    const double toint = 6755399441055744.0;   // 1.5 / DBL_EPSILON
    const double pio4 = 0x1.921fb6p-1;
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079631090164184570e+00;
    const double pio2_1t = 1.58932547735281966916e-08;
    uint32_t ix = f2u(x) & 0x7fffffffu;
    if (ix < 0x4dc90fdbu) {
        double fn = ((double)x * invpio2 + toint) - toint;
        int n = (int)fn;
        *y = ((double)x - fn * pio2_1) - fn * pio2_1t;
        if (*y < -pio4) {
            n--; fn--;
            *y = ((double)x - fn * pio2_1) - fn * pio2_1t;
        } else if (*y > pio4) {
            n++; fn++;
            *y = ((double)x - fn * pio2_1) - fn * pio2_1t;
        }
        return n;
    }
    if (ix >= 0x7f800000u) { *y = (double)(x - x); return 0; }
    return rem_pio2_huge(x, y);
}

R2D_HD float sinf_r(float x)
{
    const double p1 = 1.57079632679489655800e+00, p2 = 3.14159265358979311600e+00,
                 p3 = 4.71238898038468967400e+00, p4 = 6.28318530717958623200e+00;
    uint32_t ix = f2u(x);
    int sign = (int)(ix >> 31);
    ix &= 0x7fffffffu;
    if (ix <= 0x3f490fdau) {
        if (ix < 0x39800000u) return x;
        return k_sindf(x);
    }
    if (ix <= 0x407b53d1u) {
        if (ix <= 0x4016cbe3u) {
            if (sign) return -k_cosdf((double)x + p1);
            return k_cosdf((double)x - p1);
        }
        return k_sindf(sign ? -((double)x + p2) : -((double)x - p2));
    }
    if (ix <= 0x40e231d5u) {
        if (ix <= 0x40afeddfu) {
            if (sign) return k_cosdf((double)x + p3);
            return -k_cosdf((double)x - p3);
        }
        return k_sindf(sign ? (double)x + p4 : (double)x - p4);
    }
    if (ix >= 0x7f800000u) return x - x;
    double y;
    int n = rem_pio2f(x, &y);
    switch (n & 3) {
    case 0: return k_sindf(y);
    case 1: return k_cosdf(y);
    case 2: return k_sindf(-y);
    default: return -k_cosdf(y);
    }
}

R2D_HD float cosf_r(float x)
{
    const double p1 = 1.57079632679489655800e+00, p2 = 3.14159265358979311600e+00,
                 p3 = 4.71238898038468967400e+00, p4 = 6.28318530717958623200e+00;
    uint32_t ix = f2u(x);
    int sign = (int)(ix >> 31);
    ix &= 0x7fffffffu;
    if (ix <= 0x3f490fdau) {
        if (ix < 0x39800000u) return 1.0f;
        return k_cosdf(x);
    }
    if (ix <= 0x407b53d1u) {
        if (ix > 0x4016cbe3u) return -k_cosdf(sign ? (double)x + p2 : (double)x - p2);
        if (sign) return k_sindf((double)x + p1);
        return k_sindf(p1 - (double)x);
    }
    if (ix <= 0x40e231d5u) {
        if (ix > 0x40afeddfu) return k_cosdf(sign ? (double)x + p4 : (double)x - p4);
        if (sign) return k_sindf(-(double)x - p3);
        return k_sindf((double)x - p3);
    }
    if (ix >= 0x7f800000u) return x - x;
    double y;
    int n = rem_pio2f(x, &y);
    switch (n & 3) {
    case 0: return k_cosdf(y);
    case 1: return k_sindf(-y);
    case 2: return -k_cosdf(y);
    default: return k_sindf(y);
    }
}

R2D_HD v2 rotate2(v2 a, float angle)
{
    float c = cosf_r(angle);
    float s = sinf_r(angle);
    return rotcs(a, c, s);
}

// SpatialHash.hash (SpatialHash.zig:78-81): wrapping i64 products, xor, u64 modulo
R2D_HD uint64_t cell_hash(uint64_t table_size, int64_t xi, int64_t yi)
{
    uint64_t h = ((uint64_t)xi * 92837111ull) ^ ((uint64_t)yi * 689287499ull);
    return h % table_size;
}

}  // namespace r2d
