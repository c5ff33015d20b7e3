"""Deterministic transcendental functions for the float32 oracle.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports this module;
it is used by the oracle (the checker) and by tests.

The reference evaluates ``jnp.cos``, ``jnp.sin`` and ``jnp.power`` through
XLA's CPU lowering, whose last-ulp behaviour cannot be pinned without running
JAX.  The float32 oracle therefore fixes its own definition of these functions
and the HIP kernels implement the *same* operation sequence
(``mbd-icenav_amd/csrc/detmath.h``), so the GPU and the oracle agree bit for bit:

* every function is evaluated in IEEE binary64 with plain ``+ - * /`` (no fused
  multiply-add; numpy never contracts and the HIP build uses
  ``-ffp-contract=off``), ``rint`` (round-half-even), ``frexp`` and ``ldexp``;
* the float32 result is the binary64 result rounded once to float32.

Methods (all coefficients derived here from the mathematics, see the hex
literals' comments):

* sin / cos: Cody-Waite reduction by pi/2 with a three-part split of pi/2
  (30 + 30 + 53 bits), then the Taylor series of sin (through r^19) and
  cos (through r^20) on |r| <= pi/4, quadrant selection by k mod 4.
* log: x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
  log m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^27/27), plus e ln2 with a
  32-bit / 53-bit split of ln2.
* exp: y = k ln2 + r with k = rint(y / ln2), Taylor series of exp(r) through
  r^17, then ldexp.
* pow(x, p) for x >= 0: exact special cases p == 1 (returns x), x == 0, p == 0;
  otherwise exp(p log x) in binary64.

Accuracy (checked by ``tests/test_detmath.py`` against numpy's binary64 libm):
better than 4 ulp of binary64 over the tested ranges, hence the float32 results
are correctly rounded except in (untested-for) double-rounding ties.
"""
from __future__ import annotations

import numpy as np

# pi/2 split: PIO2_1 and PIO2_2 carry 30 significant bits so k*PIO2_1 and
# k*PIO2_2 are exact for |k| < 2^23.
PIO2_1 = float.fromhex("0x1.921fb54000000p+0")
PIO2_2 = float.fromhex("0x1.10b4611800000p-30")
PIO2_3 = float.fromhex("0x1.313198a2e0370p-61")
TWO_OVER_PI = float.fromhex("0x1.45f306dc9c883p-1")

# ln2 split: LN2_HI has 32 significant bits so e*LN2_HI is exact for |e| < 2^21.
LN2_HI = float.fromhex("0x1.62e42fee00000p-1")
LN2_LO = float.fromhex("0x1.a39ef35793c76p-33")
INV_LN2 = float.fromhex("0x1.71547652b82fep+0")

# 1/n! rounded to binary64, n = 2..21
INV_FACT = {
    2: float.fromhex("0x1.0000000000000p-1"),
    3: float.fromhex("0x1.5555555555555p-3"),
    4: float.fromhex("0x1.5555555555555p-5"),
    5: float.fromhex("0x1.1111111111111p-7"),
    6: float.fromhex("0x1.6c16c16c16c17p-10"),
    7: float.fromhex("0x1.a01a01a01a01ap-13"),
    8: float.fromhex("0x1.a01a01a01a01ap-16"),
    9: float.fromhex("0x1.71de3a556c734p-19"),
    10: float.fromhex("0x1.27e4fb7789f5cp-22"),
    11: float.fromhex("0x1.ae64567f544e4p-26"),
    12: float.fromhex("0x1.1eed8eff8d898p-29"),
    13: float.fromhex("0x1.6124613a86d09p-33"),
    14: float.fromhex("0x1.93974a8c07c9dp-37"),
    15: float.fromhex("0x1.ae7f3e733b81fp-41"),
    16: float.fromhex("0x1.ae7f3e733b81fp-45"),
    17: float.fromhex("0x1.952c77030ad4ap-49"),
    18: float.fromhex("0x1.6827863b97d97p-53"),
    19: float.fromhex("0x1.2f49b46814157p-57"),
    20: float.fromhex("0x1.e542ba4020225p-62"),
}
# 2/(2k+1) rounded to binary64, k = 1..13  (atanh series of log)
TWO_OVER_ODD = {
    3: float.fromhex("0x1.5555555555555p-1"),
    5: float.fromhex("0x1.999999999999ap-2"),
    7: float.fromhex("0x1.2492492492492p-2"),
    9: float.fromhex("0x1.c71c71c71c71cp-3"),
    11: float.fromhex("0x1.745d1745d1746p-3"),
    13: float.fromhex("0x1.3b13b13b13b14p-3"),
    15: float.fromhex("0x1.1111111111111p-3"),
    17: float.fromhex("0x1.e1e1e1e1e1e1ep-4"),
    19: float.fromhex("0x1.af286bca1af28p-4"),
    21: float.fromhex("0x1.8618618618618p-4"),
    23: float.fromhex("0x1.642c8590b2164p-4"),
    25: float.fromhex("0x1.47ae147ae147bp-4"),
    27: float.fromhex("0x1.2f684bda12f68p-4"),
}
SQRT_HALF = float.fromhex("0x1.6a09e667f3bcdp-1")

_F64 = np.float64


def _sin_poly(r):
    """sin(r) for |r| <= pi/4: r + r^3 (-1/3! + z (1/5! - z (...)))."""
    z = r * r
    p = _F64(-INV_FACT[19])
    p = _F64(INV_FACT[17]) + z * p
    p = _F64(-INV_FACT[15]) + z * p
    p = _F64(INV_FACT[13]) + z * p
    p = _F64(-INV_FACT[11]) + z * p
    p = _F64(INV_FACT[9]) + z * p
    p = _F64(-INV_FACT[7]) + z * p
    p = _F64(INV_FACT[5]) + z * p
    p = _F64(-INV_FACT[3]) + z * p
    return r + (r * z) * p


def _cos_poly(r):
    """cos(r) for |r| <= pi/4: 1 - z/2 + z^2 (1/4! - z (...))."""
    z = r * r
    p = _F64(INV_FACT[20])
    p = _F64(-INV_FACT[18]) + z * p
    p = _F64(INV_FACT[16]) + z * p
    p = _F64(-INV_FACT[14]) + z * p
    p = _F64(INV_FACT[12]) + z * p
    p = _F64(-INV_FACT[10]) + z * p
    p = _F64(INV_FACT[8]) + z * p
    p = _F64(-INV_FACT[6]) + z * p
    p = _F64(INV_FACT[4]) + z * p
    return (_F64(1.0) - _F64(0.5) * z) + (z * z) * p


def _reduce(x):
    x = np.asarray(x, dtype=np.float64)
    k = np.rint(x * _F64(TWO_OVER_PI))
    r = ((x - k * _F64(PIO2_1)) - k * _F64(PIO2_2)) - k * _F64(PIO2_3)
    q = np.mod(k, 4.0).astype(np.int64)   # k is integral; mod 4 in {0,1,2,3}
    return r, q


def sincos64(x):
    """(sin x, cos x) in binary64 by the fixed operation sequence above."""
    r, q = _reduce(x)
    s = _sin_poly(r)
    c = _cos_poly(r)
    sin_v = np.select([q == 0, q == 1, q == 2], [s, c, -s], -c)
    cos_v = np.select([q == 0, q == 1, q == 2], [c, -s, -c], s)
    return sin_v, cos_v


def sin32(x):
    """float32 sin: binary64 evaluation of the float32 argument, rounded once."""
    s, _ = sincos64(np.asarray(x, dtype=np.float32).astype(np.float64))
    return np.asarray(s).astype(np.float32)


def cos32(x):
    _, c = sincos64(np.asarray(x, dtype=np.float32).astype(np.float64))
    return np.asarray(c).astype(np.float32)


def log64(x):
    """Natural log in binary64 for finite x > 0 (callers guard the rest)."""
    x = np.asarray(x, dtype=np.float64)
    m, e = np.frexp(x)                      # x = m 2^e, m in [0.5, 1)
    small = m < _F64(SQRT_HALF)
    m = np.where(small, m * _F64(2.0), m)   # exact
    e = np.where(small, e - 1, e).astype(np.float64)
    f = m - _F64(1.0)                       # exact (Sterbenz)
    s = f / (_F64(2.0) + f)
    z = s * s
    p = _F64(TWO_OVER_ODD[27])
    for k in (25, 23, 21, 19, 17, 15, 13, 11, 9, 7, 5, 3):
        p = _F64(TWO_OVER_ODD[k]) + z * p
    logm = (_F64(2.0) * s) + (s * z) * p
    return e * _F64(LN2_HI) + (logm + e * _F64(LN2_LO))


def exp64(y):
    """exp in binary64 for finite y; saturates to inf / 0 outside the range."""
    y = np.asarray(y, dtype=np.float64)
    yc = np.clip(y, -746.0, 710.0)
    k = np.rint(yc * _F64(INV_LN2))
    r = (yc - k * _F64(LN2_HI)) - k * _F64(LN2_LO)
    p = _F64(INV_FACT[17])
    for n in (16, 15, 14, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2):
        p = _F64(INV_FACT[n]) + r * p
    p = _F64(1.0) + r * p
    p = _F64(1.0) + r * p
    with np.errstate(over="ignore", under="ignore"):
        out = np.ldexp(p, k.astype(np.int64))
    out = np.where(y > 709.8, np.inf, out)
    out = np.where(y < -745.2, 0.0, out)
    return out


def pow32(x, p):
    """float32 x**p for x >= 0 (NaN propagates), scalar float32 exponent p.

    p == 1 returns x unchanged; x == 0 gives 0 (p > 0), 1 (p == 0), inf (p < 0);
    p == 0 gives 1; otherwise exp(p log x) in binary64 rounded to float32.
    """
    x = np.asarray(x, dtype=np.float32)
    p32 = np.float32(p)
    if p32 == np.float32(1.0):
        return x.copy()
    if p32 == np.float32(0.0):
        return np.where(np.isnan(x), x, np.float32(1.0)).astype(np.float32)
    xd = x.astype(np.float64)
    pd = np.float64(p32)
    pos = xd > 0.0
    safe = np.where(pos & np.isfinite(xd), xd, 1.0)
    with np.errstate(over="ignore", invalid="ignore"):
        v = exp64(pd * log64(safe))
    zero_val = 0.0 if pd > 0 else np.inf
    inf_val = np.inf if pd > 0 else 0.0
    v = np.where(xd == 0.0, zero_val, v)
    v = np.where(np.isinf(xd) & (xd > 0), inf_val, v)
    v = np.where(np.isnan(xd) | (xd < 0), np.nan, v)
    with np.errstate(over="ignore"):
        return v.astype(np.float32)
