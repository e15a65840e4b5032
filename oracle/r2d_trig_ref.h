/*
 * ORACLE (test infrastructure only) -- single-precision sine/cosine.
 *
 * The reference rotates vectors with `std.math.cos/sin` on f32
 * (src/core/nmath.zig:129-136).  On targets without libc (the shipped wasm32
 * build, src/wasm_root.zig) Zig resolves those to its compiler_rt `sinf` /
 * `cosf`, which are ports of the FreeBSD msun / musl routines: a double-
 * precision argument reduction (__rem_pio2f) followed by the double kernels
 * __sindf / __cosdf and a rounding to float.  This header restates that
 * published method so that the oracle and the GPU path share one well-defined
 * set of trig bits.
 *
 * Parity note: for |x| >= 2^28 * pi/2 (~4.2e8 rad) musl uses a 24-bit
 * Payne-Hanek reduction (__rem_pio2_large, prec=0).  This restatement uses an
 * exact 128-bit reduction there instead, so results in that range are checked
 * only against FP64 sin/cos (<= 1 ulp) -- parity unpinned for that range.
 *
 * The sin/cos kernels, their coefficients and the reduction thresholds below
 * follow FreeBSD msun (k_sinf.c, k_cosf.c, e_rem_pio2f.c, s_sinf.c, s_cosf.c;
 * as carried by musl), which carry this notice:
 *
 *   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
 *
 *   Developed at SunPro, a Sun Microsystems, Inc. business.
 *   Permission to use, copy, modify, and distribute this
 *   software is freely granted, provided that this notice
 *   is preserved.
 *
 *   (float versions: conversion to float by Ian Lance Taylor, Cygnus
 *   Support; optimized by Bruce D. Evans.)
 */
#ifndef R2D_TRIG_REF_H
#define R2D_TRIG_REF_H

#include <stdint.h>
#include <string.h>

static inline uint32_t r2dref_f2u(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }

/* __sindf: |sin(x)/x - s(x)| < 2**-37.5 on [-pi/4, pi/4] */
static inline float r2dref_sindf(double x)
{
    const double S1 = -0x15555554cbac77.0p-55, S2 = 0x111110896efbb2.0p-59,
                 S3 = -0x1a00f9e2cae774.0p-65, S4 = 0x16cd878c3b46a7.0p-71;
    double z = x * x;
    double w = z * z;
    double r = S3 + z * S4;
    double s = z * x;
    return (float)((x + s * (S1 + z * S2)) + s * w * r);
}

/* __cosdf: |cos(x) - c(x)| < 2**-34.1 on [-pi/4, pi/4] */
static inline float r2dref_cosdf(double x)
{
    const double C0 = -0x1ffffffd0c5e81.0p-54, C1 = 0x155553e1053a42.0p-57,
                 C2 = -0x16c087e80f1e27.0p-62, C3 = 0x199342e0ee5069.0p-68;
    double z = x * x;
    double w = z * z;
    double r = C2 + z * C3;
    return (float)(((1.0 + z * C0) + w * C1) + (w * z) * r);
}

/* 2/pi as 32-bit words, most significant first (bits after the binary point). */
static const uint32_t r2dref_two_over_pi[] = {
    0xa2f9836eu, 0x4e441529u, 0xfc2757d1u, 0xf534ddc0u, 0xdb629599u, 0x3c439041u,
    0xfe5163abu, 0xdebbc561u, 0xb7246e3au, 0x424dd2e0u, 0x06492eeau, 0x09d1921cu,
};

/* Exact reduction for huge |x|: returns quadrant, writes y in [-pi/4, pi/4]. */
static inline int r2dref_rem_pio2_huge(float x, double *y)
{
    uint32_t u = r2dref_f2u(x);
    int sign = (int)(u >> 31);
    uint32_t ix = u & 0x7fffffffu;
    uint64_t m = (uint64_t)((ix & 0x7fffffu) | 0x800000u);
    int E = (int)(ix >> 23) - 150;           /* |x| = m * 2^E, E >= 6 here */
    int s = E - 1;                           /* first 2/pi bit index (1-based) that matters mod 4 */
    /* gather 128-bit window of 2/pi starting at bit s */
    uint64_t w_hi = 0, w_lo = 0;
    for (int i = 0; i < 128; ++i) {
        int bit = s + i;                     /* 1-based */
        int word = (bit - 1) >> 5, off = (bit - 1) & 31;
        uint32_t b = (r2dref_two_over_pi[word] >> (31 - off)) & 1u;
        if (i < 64) w_hi = (w_hi << 1) | b; else w_lo = (w_lo << 1) | b;
    }
    /* product P = m * (w_hi:w_lo); value mod 4 = P * 2^(2-128) */
    unsigned __int128 lo = (unsigned __int128)m * w_lo;
    unsigned __int128 hi = (unsigned __int128)m * w_hi + (lo >> 64);
    uint64_t p1 = (uint64_t)hi;              /* bits 64..127 of P */
    uint64_t p2 = (uint64_t)(hi >> 64);      /* bits 128..  of P */
    /* integer part (mod 4) = bits 126..127 of P; fraction = bits below 126 */
    int n = (int)((p1 >> 62) & 3u);
    (void)p2;
    uint64_t frac = p1 << 2;                 /* top 64 fraction bits */
    double f = (double)(frac >> 11) * 0x1p-53;
    if (f >= 0.5) { f -= 1.0; n = (n + 1) & 3; }
    double r = f * 1.57079632679489661923;
    if (sign) { *y = -r; return -n; }
    *y = r;
    return n;
}

/* __rem_pio2f: x = n*pi/2 + y, |y| <~ pi/4 */
static inline int r2dref_rem_pio2f(float x, double *y)
{
    const double toint = 1.5 / 2.220446049250313080847e-16;
    const double pio4 = 0x1.921fb6p-1;
    const double invpio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079631090164184570e+00;
    const double pio2_1t = 1.58932547735281966916e-08;
    uint32_t ix = r2dref_f2u(x) & 0x7fffffffu;
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
    return r2dref_rem_pio2_huge(x, y);
}

static inline float r2dref_sinf(float x)
{
    const double s1pio2 = 1.57079632679489655800e+00, s2pio2 = 3.14159265358979311600e+00,
                 s3pio2 = 4.71238898038468967400e+00, s4pio2 = 6.28318530717958623200e+00;
    uint32_t ix = r2dref_f2u(x);
    int sign = (int)(ix >> 31);
    ix &= 0x7fffffffu;
    if (ix <= 0x3f490fdau) {
        if (ix < 0x39800000u) return x;
        return r2dref_sindf(x);
    }
    if (ix <= 0x407b53d1u) {
        if (ix <= 0x4016cbe3u) {
            if (sign) return -r2dref_cosdf((double)x + s1pio2);
            return r2dref_cosdf((double)x - s1pio2);
        }
        return r2dref_sindf(sign ? -((double)x + s2pio2) : -((double)x - s2pio2));
    }
    if (ix <= 0x40e231d5u) {
        if (ix <= 0x40afeddfu) {
            if (sign) return r2dref_cosdf((double)x + s3pio2);
            return -r2dref_cosdf((double)x - s3pio2);
        }
        return r2dref_sindf(sign ? (double)x + s4pio2 : (double)x - s4pio2);
    }
    if (ix >= 0x7f800000u) return x - x;
    double y;
    int n = r2dref_rem_pio2f(x, &y);
    switch (n & 3) {
    case 0: return r2dref_sindf(y);
    case 1: return r2dref_cosdf(y);
    case 2: return r2dref_sindf(-y);
    default: return -r2dref_cosdf(y);
    }
}

static inline float r2dref_cosf(float x)
{
    const double c1pio2 = 1.57079632679489655800e+00, c2pio2 = 3.14159265358979311600e+00,
                 c3pio2 = 4.71238898038468967400e+00, c4pio2 = 6.28318530717958623200e+00;
    uint32_t ix = r2dref_f2u(x);
    int sign = (int)(ix >> 31);
    ix &= 0x7fffffffu;
    if (ix <= 0x3f490fdau) {
        if (ix < 0x39800000u) return 1.0f;
        return r2dref_cosdf(x);
    }
    if (ix <= 0x407b53d1u) {
        if (ix > 0x4016cbe3u) return -r2dref_cosdf(sign ? (double)x + c2pio2 : (double)x - c2pio2);
        if (sign) return r2dref_sindf((double)x + c1pio2);
        return r2dref_sindf(c1pio2 - (double)x);
    }
    if (ix <= 0x40e231d5u) {
        if (ix > 0x40afeddfu) return r2dref_cosdf(sign ? (double)x + c4pio2 : (double)x - c4pio2);
        if (sign) return r2dref_sindf(-(double)x - c3pio2);
        return r2dref_sindf((double)x - c3pio2);
    }
    if (ix >= 0x7f800000u) return x - x;
    double y;
    int n = r2dref_rem_pio2f(x, &y);
    switch (n & 3) {
    case 0: return r2dref_cosdf(y);
    case 1: return r2dref_sindf(-y);
    case 2: return -r2dref_cosdf(y);
    default: return r2dref_sindf(y);
    }
}

#endif
