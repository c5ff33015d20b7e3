"""CPU checks that pin the surrogate oracle: central differences of the
float64 network against its hand-written backward, im2col/col2im adjointness,
and the bf16 rounding rule."""
import numpy as np
import pytest

from oracle import surrogate as osur


def _tiny():
    arch = osur.Arch(G=8, cin=2, chans=(3, 4), extra=3, hidden=(5,), out=2)
    rng = np.random.default_rng(0)
    params = osur.init_params(arch, 1)
    for k in params:
        if k.startswith(("cb", "fb")):
            params[k] = rng.normal(0, 0.1, params[k].shape).astype(np.float32)
    raster = rng.uniform(0, 1, (3, 8, 8, 2))
    extra = rng.normal(0, 1, (3, 3))
    dy = rng.normal(0, 1, (3, 2))
    return arch, params, raster, extra, dy


def test_backward_matches_central_differences():
    arch, params, raster, extra, dy = _tiny()
    net = osur.Net(arch, params, round_bf16=False)
    net.forward(raster, extra)
    g = net.backward(dy)
    rng = np.random.default_rng(2)
    checked = 0
    for key, shp in arch.param_shapes().items():
        for _ in range(4):
            idx = tuple(int(rng.integers(0, s)) for s in shp)
            num = osur.numeric_grad(arch, params, raster, extra, dy, key, idx)
            assert abs(g[key][idx] - num) <= 1e-6 * max(1.0, abs(num)), (key, idx, g[key][idx], num)
            checked += 1
    assert checked == 4 * len(arch.param_shapes())


def test_im2col_col2im_adjoint():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 7, 6, 3))
    col = osur.im2col(x)
    assert col.shape == (2 * 4 * 3, 27)
    d = rng.normal(size=col.shape)
    # <im2col(x), d> == <x, col2im(d)>
    assert np.isclose((col * d).sum(), (x * osur.col2im(d, x.shape)).sum(), rtol=1e-12)
    # a single patch: centre tap of output (0, 0) reads input (0, 0); tap (0, 0) is padding
    assert col[0, 4 * 3 + 1] == x[0, 0, 0, 1] and col[0, 0] == 0.0


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -3.14159, 65504.0], np.float32)
    r = osur.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even (1.0)
    assert r[2] == 1.0 + 4 * 2 ** -8        # tie -> even (mantissa ...10)
    assert r[3] == 1.0                      # below half ulp
    assert abs(r[4] - (-3.140625)) < 1e-7
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)
