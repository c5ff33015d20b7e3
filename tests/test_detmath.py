"""The oracle's deterministic transcendentals against numpy's binary64 libm.

These functions define the float32 oracle's sin/cos/pow (and the kernels
implement the same operation sequence), so they are pinned here against the
library implementations over the ranges the step uses.
"""
import numpy as np

from oracle import detmath


def ulp64(x):
    return np.spacing(np.abs(x).astype(np.float64))


def test_sincos64_accuracy():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-4, 4, 200000), rng.uniform(-300, 300, 100000),
                        np.array([0.0, -0.0, np.pi / 4, -np.pi / 4, np.pi / 2, np.pi, 1e-12, 5e-300])])
    s, c = detmath.sincos64(x)
    ref_s, ref_c = np.sin(x), np.cos(x)
    es = np.abs(s - ref_s) / np.maximum(ulp64(ref_s), 1e-300)
    ec = np.abs(c - ref_c) / np.maximum(ulp64(ref_c), 1e-300)
    # absolute error near the zeros of sin/cos is bounded by the reduction
    assert np.max(np.where(np.abs(ref_s) > 1e-3, es, 0)) < 4
    assert np.max(np.where(np.abs(ref_c) > 1e-3, ec, 0)) < 4
    assert np.max(np.abs(s - ref_s)) < 1e-15
    assert np.max(np.abs(c - ref_c)) < 1e-15


def test_sincos32_mostly_correctly_rounded():
    rng = np.random.default_rng(1)
    x = rng.uniform(-10, 10, 200000).astype(np.float32)
    s = detmath.sin32(x); c = detmath.cos32(x)
    rs = np.sin(x.astype(np.float64)).astype(np.float32)
    rc = np.cos(x.astype(np.float64)).astype(np.float32)
    assert np.mean(s == rs) > 0.9999 and np.mean(c == rc) > 0.9999
    assert np.max(np.abs(s.astype(np.float64) - np.sin(x.astype(np.float64)))) <= 2 ** -24
    # symmetry used by the hull test: cos(-h) == cos(h), sin(-h) == -sin(h)
    assert np.array_equal(detmath.cos32(-x), c) and np.array_equal(detmath.sin32(-x), -s)


def test_log_exp64_accuracy():
    rng = np.random.default_rng(2)
    x = np.exp(rng.uniform(-80, 80, 200000))
    lg = detmath.log64(x)
    assert np.max(np.abs(lg - np.log(x)) / ulp64(np.log(x))) < 4
    y = rng.uniform(-700, 700, 200000)
    ex = detmath.exp64(y)
    assert np.max(np.abs(ex - np.exp(y)) / ulp64(np.exp(y))) < 8


def test_pow32():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(0, 1, 100000), np.exp(rng.uniform(-20, 5, 100000))]).astype(np.float32)
    for p in (1.5, 1.75, 2.0, 0.5, 3.3):
        got = detmath.pow32(x, p)
        ref = np.power(x.astype(np.float64), np.float64(np.float32(p))).astype(np.float32)
        assert np.mean(got == ref) > 0.9999, p
        assert np.max(np.abs(got.astype(np.float64) / np.maximum(ref.astype(np.float64), 1e-38) - 1)) < 2e-7
    # exact special cases
    z = np.array([0.0, -0.0, 1.0, 7.5, np.inf, np.nan], np.float32)
    np.testing.assert_array_equal(detmath.pow32(z, 1.0), z)
    out = detmath.pow32(z, 1.5)
    assert out[0] == 0 and out[1] == 0 and out[2] == 1 and np.isinf(out[4]) and np.isnan(out[5])
    assert detmath.pow32(np.float32(2.0), 0.0) == 1.0
