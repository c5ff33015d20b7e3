// detmath.h -- deterministic elementary functions and XLA-semantics helpers
// for the DPDEM kernels (gfx950).
//
// sin/cos/pow are evaluated in binary64 with a fixed sequence of + - * /,
// rint, frexp and ldexp (the file is compiled with -ffp-contract=off, so no
// fused multiply-add is formed) and rounded once to float32:
//   sin/cos: Cody-Waite reduction by pi/2 (30 + 30 + 53 bit split), Taylor
//            series of sin through r^19 and of cos through r^20 on |r|<=pi/4.
//   log:     x = m 2^e, m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//            log m = 2 atanh(s) through s^27, plus e ln2 (32 + 53 bit split).
//   exp:     y = k ln2 + r, Taylor series of exp(r) through r^17, ldexp.
//   pow:     x >= 0 only; p == 1 -> x; x == 0 / p == 0 special cases;
//            otherwise exp(p log x).
// The coefficients are 1/n! and 2/(2k+1) rounded to binary64.  These results
// are correctly rounded float32 values except in rare double-rounding ties.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace icedem {

// pi/2 = PIO2_1 + PIO2_2 + PIO2_3 (first two parts carry 30 significant bits)
#define DM_PIO2_1 0x1.921fb54000000p+0
#define DM_PIO2_2 0x1.10b4611800000p-30
#define DM_PIO2_3 0x1.313198a2e0370p-61
#define DM_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define DM_LN2_HI 0x1.62e42fee00000p-1
#define DM_LN2_LO 0x1.a39ef35793c76p-33
#define DM_INV_LN2 0x1.71547652b82fep+0
#define DM_SQRT_HALF 0x1.6a09e667f3bcdp-1

// 1/n!
#define DM_F2 0x1.0000000000000p-1
#define DM_F3 0x1.5555555555555p-3
#define DM_F4 0x1.5555555555555p-5
#define DM_F5 0x1.1111111111111p-7
#define DM_F6 0x1.6c16c16c16c17p-10
#define DM_F7 0x1.a01a01a01a01ap-13
#define DM_F8 0x1.a01a01a01a01ap-16
#define DM_F9 0x1.71de3a556c734p-19
#define DM_F10 0x1.27e4fb7789f5cp-22
#define DM_F11 0x1.ae64567f544e4p-26
#define DM_F12 0x1.1eed8eff8d898p-29
#define DM_F13 0x1.6124613a86d09p-33
#define DM_F14 0x1.93974a8c07c9dp-37
#define DM_F15 0x1.ae7f3e733b81fp-41
#define DM_F16 0x1.ae7f3e733b81fp-45
#define DM_F17 0x1.952c77030ad4ap-49
#define DM_F18 0x1.6827863b97d97p-53
#define DM_F19 0x1.2f49b46814157p-57
#define DM_F20 0x1.e542ba4020225p-62

__device__ __forceinline__ double dm_sin_poly(double r) {
    double z = r * r;
    double p = -DM_F19;
    p = DM_F17 + z * p;
    p = -DM_F15 + z * p;
    p = DM_F13 + z * p;
    p = -DM_F11 + z * p;
    p = DM_F9 + z * p;
    p = -DM_F7 + z * p;
    p = DM_F5 + z * p;
    p = -DM_F3 + z * p;
    return r + (r * z) * p;
}

__device__ __forceinline__ double dm_cos_poly(double r) {
    double z = r * r;
    double p = DM_F20;
    p = -DM_F18 + z * p;
    p = DM_F16 + z * p;
    p = -DM_F14 + z * p;
    p = DM_F12 + z * p;
    p = -DM_F10 + z * p;
    p = DM_F8 + z * p;
    p = -DM_F6 + z * p;
    p = DM_F4 + z * p;
    return (1.0 - 0.5 * z) + (z * z) * p;
}

__device__ __forceinline__ void dm_sincos64(double x, double* sv, double* cv) {
    double k = rint(x * DM_TWO_OVER_PI);
    double r = ((x - k * DM_PIO2_1) - k * DM_PIO2_2) - k * DM_PIO2_3;
    int q = (int)(((long long)k) & 3LL);   // floor-mod 4 (two's complement)
    double s = dm_sin_poly(r);
    double c = dm_cos_poly(r);
    double so, co;
    if (q == 0) { so = s; co = c; }
    else if (q == 1) { so = c; co = -s; }
    else if (q == 2) { so = -s; co = -c; }
    else { so = -c; co = s; }
    *sv = so; *cv = co;
}

__device__ __forceinline__ float dm_sinf(float x) {
    double s, c; dm_sincos64((double)x, &s, &c); return (float)s;
}
__device__ __forceinline__ float dm_cosf(float x) {
    double s, c; dm_sincos64((double)x, &s, &c); return (float)c;
}
__device__ __forceinline__ void dm_sincosf(float x, float* s, float* c) {
    double sd, cd; dm_sincos64((double)x, &sd, &cd); *s = (float)sd; *c = (float)cd;
}

// 2/(2k+1)
#define DM_O3 0x1.5555555555555p-1
#define DM_O5 0x1.999999999999ap-2
#define DM_O7 0x1.2492492492492p-2
#define DM_O9 0x1.c71c71c71c71cp-3
#define DM_O11 0x1.745d1745d1746p-3
#define DM_O13 0x1.3b13b13b13b14p-3
#define DM_O15 0x1.1111111111111p-3
#define DM_O17 0x1.e1e1e1e1e1e1ep-4
#define DM_O19 0x1.af286bca1af28p-4
#define DM_O21 0x1.8618618618618p-4
#define DM_O23 0x1.642c8590b2164p-4
#define DM_O25 0x1.47ae147ae147bp-4
#define DM_O27 0x1.2f684bda12f68p-4

__device__ __forceinline__ double dm_log64(double x) {   // finite x > 0
    int e;
    double m = frexp(x, &e);
    if (m < DM_SQRT_HALF) { m = m * 2.0; e = e - 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = DM_O27;
    p = DM_O25 + z * p;
    p = DM_O23 + z * p;
    p = DM_O21 + z * p;
    p = DM_O19 + z * p;
    p = DM_O17 + z * p;
    p = DM_O15 + z * p;
    p = DM_O13 + z * p;
    p = DM_O11 + z * p;
    p = DM_O9 + z * p;
    p = DM_O7 + z * p;
    p = DM_O5 + z * p;
    p = DM_O3 + z * p;
    double logm = (2.0 * s) + (s * z) * p;
    double ed = (double)e;
    return ed * DM_LN2_HI + (logm + ed * DM_LN2_LO);
}

__device__ __forceinline__ double dm_exp64(double y) {
    if (y > 709.8) return __builtin_huge_val();
    if (y < -745.2) return 0.0;
    double yc = y;
    double k = rint(yc * DM_INV_LN2);
    double r = (yc - k * DM_LN2_HI) - k * DM_LN2_LO;
    double p = DM_F17;
    p = DM_F16 + r * p;
    p = DM_F15 + r * p;
    p = DM_F14 + r * p;
    p = DM_F13 + r * p;
    p = DM_F12 + r * p;
    p = DM_F11 + r * p;
    p = DM_F10 + r * p;
    p = DM_F9 + r * p;
    p = DM_F8 + r * p;
    p = DM_F7 + r * p;
    p = DM_F6 + r * p;
    p = DM_F5 + r * p;
    p = DM_F4 + r * p;
    p = DM_F3 + r * p;
    p = DM_F2 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    return ldexp(p, (int)k);
}

// x**p for x >= 0 (float32 in, float32 out), p a float32 exponent.
__device__ __forceinline__ float dm_powf(float x, float p) {
    if (p == 1.0f) return x;
    if (x != x) return x;
    if (p == 0.0f) return 1.0f;
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 0.0f) return p > 0.0f ? 0.0f : __builtin_huge_valf();
    if (__builtin_isinf(x)) return p > 0.0f ? __builtin_huge_valf() : 0.0f;
    double v = dm_exp64((double)p * dm_log64((double)x));
    return (float)v;
}

// ── XLA-semantics helpers (NaN-propagating max/min, clip, sign) ─────────────
__device__ __forceinline__ float vmax(float a, float b) { return (a != a || a >= b) ? a : b; }
__device__ __forceinline__ float vmin(float a, float b) { return (a != a || a <= b) ? a : b; }
__device__ __forceinline__ float vclip(float x, float lo, float hi) { return vmin(vmax(x, lo), hi); }
__device__ __forceinline__ float vsign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : x); }
__device__ __forceinline__ bool visfinite(float x) { return __builtin_isfinite(x); }

}  // namespace icedem
