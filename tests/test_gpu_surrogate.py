"""Surrogate kernels (BF16 MFMA GEMM, im2col/col2im, reductions, Adam, raster)
and the whole network's forward and backward against the NumPy oracle."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import surrogate as osur

pytestmark = pytest.mark.gpu


def _lib():
    from icedem import _lib
    return _lib, _lib.load()


def _bf(a):
    return torch.as_tensor(osur.bf16(a)).to(device="cuda", dtype=torch.bfloat16)


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _s():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K,split", [(16, 16, 32, 1), (70, 45, 100, 1), (128, 64, 2134, 1), (37, 29, 5000, 4),
                                         (18, 32, 40000, 70)])
def test_gemm_matches_float64_of_bf16_inputs(ta, tb, M, N, K, split):
    L, lib = _lib()
    rng = np.random.default_rng(M * 7 + N * 3 + K + ta * 2 + tb)
    A = osur.bf16(rng.normal(size=(K, M) if ta else (M, K)))
    B = osur.bf16(rng.normal(size=(N, K) if tb else (K, N)))
    bias = rng.normal(size=N).astype(np.float32)
    opA = A.T.astype(np.float64) if ta else A.astype(np.float64)
    opB = B.T.astype(np.float64) if tb else B.astype(np.float64)
    ref = opA @ opB + bias
    bound = (np.abs(opA) @ np.abs(opB)) * (K * 2.0 ** -23) + 1e-6
    dA, dB = _bf(A), _bf(B)
    Cg = torch.full((M, N), 7.0, dtype=torch.float32, device="cuda")
    Cb = torch.zeros((M, N), dtype=torch.bfloat16, device="cuda")
    work = torch.zeros((split + (split + 31) // 32) * M * N, dtype=torch.float32, device="cuda")
    db = torch.as_tensor(bias, device="cuda")
    rc = lib.icesur_gemm(ta, tb, M, N, K, _p(dA), M if ta else K, _p(dB), K if tb else N, _p(Cg), N, _p(Cb), N,
                         _p(db), 4, split, _p(work), _s())
    assert rc == 0, lib.icedem_last_error(None)
    got = Cg.cpu().numpy().astype(np.float64)
    err = np.abs(got - ref)
    print(f"gemm ta={ta} tb={tb} {M}x{N}x{K} split={split}: max err/bound {np.max(err / bound):.3g}", flush=True)
    assert np.all(err <= bound), np.max(err / bound)
    np.testing.assert_array_equal(Cb.float().cpu().numpy(), osur.bf16(Cg.cpu().numpy()))
    # relu + accumulate
    C2 = Cg.clone()
    rc = lib.icesur_gemm(ta, tb, M, N, K, _p(dA), M if ta else K, _p(dB), K if tb else N, _p(C2), N, None, 0,
                         None, 1 | 2, split, _p(work), _s())
    assert rc == 0
    ref2 = np.maximum(got + (opA @ opB), 0.0)
    assert np.all(np.abs(C2.cpu().numpy() - ref2) <= 2 * bound + 1e-5 * np.abs(ref2))


def test_im2col_col2im_relu_colsum():
    L, lib = _lib()
    rng = np.random.default_rng(5)
    B, H, W, Cc = 3, 9, 8, 5
    x = osur.bf16(rng.normal(size=(B, H, W, Cc)))
    dx_ = _bf(x)
    Ho, Wo = (H + 1) // 2, (W + 1) // 2
    col = torch.zeros((B * Ho * Wo, 9 * Cc), dtype=torch.bfloat16, device="cuda")
    assert lib.icesur_im2col(_p(dx_), B, H, W, Cc, _p(col), _s()) == 0
    np.testing.assert_array_equal(col.float().cpu().numpy(), osur.im2col(x))
    d = rng.normal(size=col.shape).astype(np.float32)
    dd = torch.as_tensor(d, device="cuda")
    out = torch.zeros((B, H, W, Cc), dtype=torch.float32, device="cuda")
    assert lib.icesur_col2im(_p(dd), B, H, W, Cc, _p(out), _s()) == 0
    ref = osur.col2im(d.astype(np.float64), (B, H, W, Cc))
    assert np.max(np.abs(out.cpu().numpy() - ref)) <= 1e-5
    # relu backward
    g = torch.as_tensor(d.copy(), device="cuda")
    y = _bf(rng.normal(size=d.shape))
    gb = torch.zeros(d.shape, dtype=torch.bfloat16, device="cuda")
    assert lib.icesur_relu_back(_p(g), d.shape[1], _p(y), d.shape[1], d.shape[0], d.shape[1], _p(gb), d.shape[1],
                                _s()) == 0
    mref = np.where(y.float().cpu().numpy() > 0, d, 0.0).astype(np.float32)
    np.testing.assert_array_equal(g.cpu().numpy(), mref)
    np.testing.assert_array_equal(gb.float().cpu().numpy(), osur.bf16(mref))
    # column sums in the documented float32 order; and within float32 rounding of the float64 sum
    for R, Cn in [(700, 13), (5000, 70)]:
        xs = rng.normal(size=(R, Cn)).astype(np.float32)
        out = torch.zeros(Cn, dtype=torch.float32, device="cuda")
        nb = (R + 255) // 256
        work = torch.zeros(nb * Cn, dtype=torch.float32, device="cuda")
        assert lib.icesur_colsum(_p(torch.as_tensor(xs, device="cuda")), Cn, R, Cn, _p(out), 0, _p(work), _s()) == 0
        parts = []
        for b0 in range(0, R, 256):
            ph = [np.zeros(Cn, np.float32) for _ in range(4)]
            for p in range(4):
                for r in range(b0 + p, min(R, b0 + 256), 4):
                    ph[p] = (ph[p] + xs[r]).astype(np.float32)
            parts.append(((ph[0] + ph[1]).astype(np.float32) + (ph[2] + ph[3]).astype(np.float32)).astype(np.float32))
        v = []
        for g in range(16):
            acc = np.zeros(Cn, np.float32)
            for b in range(g, nb, 16):
                acc = (acc + parts[b]).astype(np.float32)
            v.append(acc)
        w = 1
        while w < 16:
            for i in range(0, 16 - w, 2 * w):
                v[i] = (v[i] + v[i + w]).astype(np.float32)
            w *= 2
        np.testing.assert_array_equal(out.cpu().numpy(), v[0])
        np.testing.assert_allclose(out.cpu().numpy(), xs.astype(np.float64).sum(0), rtol=1e-4, atol=1e-3)


@pytest.mark.parametrize("B,H,C,N", [(3, 16, 2, 32), (5, 7, 32, 64), (2, 9, 64, 128), (3, 4, 128, 128),
                                     (130, 8, 32, 32), (1, 5, 3, 40)])
def test_implicit_conv_equals_im2col_gemm(B, H, C, N):
    """The implicit-GEMM convolution adds the same products in the same
    order as im2col + GEMM: bit-identical, with and without ReLU; and both
    match the float64 convolution of the bf16 inputs."""
    L, lib = _lib()
    rng = np.random.default_rng(B * 1000 + H * 10 + C)
    x = rng.normal(size=(B, H, H, C)).astype(np.float32)
    w = rng.normal(size=(9 * C, N)).astype(np.float32) / np.sqrt(9 * C)
    bias = rng.normal(size=N).astype(np.float32)
    dx, dw = _bf(x), _bf(w)
    db = torch.as_tensor(bias, device="cuda")
    Ho = (H + 1) // 2
    M = B * Ho * Ho
    wt = torch.empty(9 * C * N, dtype=torch.bfloat16, device="cuda")
    col = torch.empty(M * 9 * C, dtype=torch.bfloat16, device="cuda")
    assert lib.icesur_im2col(_p(dx), B, H, H, C, _p(col), _s()) == 0
    ref_col = osur.im2col(osur.bf16(x).astype(np.float64))
    ref = ref_col @ osur.bf16(w).astype(np.float64) + bias
    for flags in (4, 5):     # BIAS, BIAS | RELU
        out = torch.empty(M * N, dtype=torch.bfloat16, device="cuda")
        assert lib.icesur_conv(_p(dx), B, H, H, C, _p(dw), N, _p(db), flags, _p(wt), _p(out), _s()) == 0
        out2 = torch.empty(M * N, dtype=torch.bfloat16, device="cuda")
        assert lib.icesur_gemm(0, 0, M, N, 9 * C, _p(col), 9 * C, _p(dw), N, None, 0, _p(out2), N, _p(db), flags, 1,
                               None, _s()) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out.view(torch.int16).cpu().numpy(), out2.view(torch.int16).cpu().numpy())
        r = np.maximum(ref, 0.0) if flags & 1 else ref
        got = out.float().cpu().numpy().reshape(M, N)
        assert np.all(np.abs(got - r) <= 2 ** -8 * np.abs(r) + 1e-5 * (1 + np.abs(r).max())), (flags, np.abs(got - r).max())



def test_mse_and_adam():
    L, lib = _lib()
    rng = np.random.default_rng(9)
    Bn, D = 300, 3
    y = rng.normal(size=(Bn, D)).astype(np.float32)
    t = rng.normal(size=(Bn, D)).astype(np.float32)
    w = np.array([1.0, 0.5, 2.0], np.float32)
    dy = torch.zeros((Bn, D), dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    assert lib.icesur_mse(_p(torch.as_tensor(y, device="cuda")), _p(torch.as_tensor(t, device="cuda")),
                          _p(torch.as_tensor(w, device="cuda")), Bn, D, _p(dy), None, _p(loss), _s()) == 0
    ref_loss = float((w * (y - t) ** 2).sum() / (Bn * D))
    assert abs(loss.item() - ref_loss) <= 1e-5 * ref_loss
    np.testing.assert_allclose(dy.cpu().numpy(), 2 * w * (y - t) / (Bn * D), rtol=1e-6)
    n = 1000
    wt = rng.normal(size=n).astype(np.float32)
    g = rng.normal(size=n).astype(np.float32)
    dw = torch.as_tensor(wt.copy(), device="cuda")
    dm = torch.zeros(n, device="cuda")
    dv = torch.zeros(n, device="cuda")
    wb = torch.zeros(n, dtype=torch.bfloat16, device="cuda")
    for step in (1, 2):
        assert lib.icesur_adam(_p(dw), _p(torch.as_tensor(g, device="cuda")), _p(dm), _p(dv), n, 1e-2, 0.9, 0.999,
                               1e-8, step, _p(wb), _s()) == 0
    m = np.zeros(n); v = np.zeros(n); ww = wt.astype(np.float64)
    for step in (1, 2):
        m = 0.9 * m + 0.1 * g; v = 0.999 * v + 0.001 * g * g
        ww = ww - 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    np.testing.assert_allclose(dw.cpu().numpy(), ww, rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(wb.float().cpu().numpy(), osur.bf16(dw.cpu().numpy()))
    # decoupled weight decay: one step from zero moments with g = 0 shrinks w by (1 - lr wd)
    dw2 = torch.as_tensor(wt.copy(), device="cuda")
    dm.zero_(); dv.zero_()
    assert lib.icesur_adamw(_p(dw2), _p(torch.zeros(n, device="cuda")), _p(dm), _p(dv), n, 1e-2, 0.9, 0.999, 1e-8,
                            0.5, 1, _p(wb), _s()) == 0
    np.testing.assert_allclose(dw2.cpu().numpy(), wt.astype(np.float64) * (1 - 1e-2 * 0.5), rtol=1e-6)


def test_raster_of_known_discs():
    from icedem import surrogate as sur
    L, lib = _lib()
    # two discs of area pi (r = 1) and 4 pi (r = 2), ship at the origin heading +x
    pos = torch.tensor([[5.0, 10.0], [0.0, 3.0]], dtype=torch.float32, device="cuda").reshape(-1)
    area = torch.tensor([np.pi, 4 * np.pi], dtype=torch.float32, device="cuda")
    thick = torch.tensor([1.0, 2.0], dtype=torch.float32, device="cuda")
    poses = torch.tensor([[0.0, 0.0, 0.0]], dtype=torch.float32, device="cuda")
    r = sur.raster(lib, pos, area, thick, poses, G=16, cell=1.0, x_lo=0.0).float().cpu().numpy()[0]
    conc, th = r[..., 0], r[..., 1]
    # disc 1 centred at (5, 0): covers cell centres within 1 m -> cells (4,7),(4,8),(5,7),(5,8)
    assert conc[4, 8] == 1.0 and conc[5, 7] == 1.0 and th[4, 8] == 1.0
    # disc 2 centred at (10, 3): thickness 2 where only it covers
    assert conc[10, 11] == 1.0 and th[10, 11] == 2.0
    assert conc[0, 0] == 0.0 and th[0, 0] == 0.0
    # covered cell count = cell centres inside the discs
    u = np.arange(16) + 0.5
    cu, cv = np.meshgrid(u, u - 8.0, indexing="ij")
    inside = ((cu - 5) ** 2 + cv ** 2 <= 1.0) | ((cu - 10) ** 2 + (cv - 3) ** 2 <= 4.0)
    np.testing.assert_array_equal(conc > 0, inside)


def _poly_counts(poly, G, cell, x_lo, ss=4):
    """Per cell, the number of the ss x ss sample points strictly inside the
    convex polygon (lo) and inside or on its boundary (hi), float64."""
    P = np.asarray(poly, np.float64)
    a2 = sum(P[k - 1, 0] * P[k, 1] - P[k, 0] * P[k - 1, 1] for k in range(len(P)))
    o = 1.0 if a2 >= 0 else -1.0
    h = cell / ss
    off = (np.arange(ss) + 0.5) * h
    qu = (x_lo + np.arange(G)[:, None] * cell + off[None, :]).reshape(-1)
    qv = (-0.5 * G * cell + np.arange(G)[:, None] * cell + off[None, :]).reshape(-1)
    U, V = np.meshgrid(qu, qv, indexing="ij")
    cr = np.stack([o * ((P[k, 0] - P[k - 1, 0]) * (V - P[k - 1, 1]) - (P[k, 1] - P[k - 1, 1]) * (U - P[k - 1, 0]))
                   for k in range(len(P))])
    lo = np.all(cr > 1e-9, axis=0).reshape(G, ss, G, ss).sum(axis=(1, 3))
    hi = np.all(cr > -1e-9, axis=0).reshape(G, ss, G, ss).sum(axis=(1, 3))
    return lo, hi


def test_raster_of_known_polygons():
    from icedem import surrogate as sur
    L, lib = _lib()
    # a 2 x 2 m square at (5, 0) (clockwise) and a right triangle of area 2
    # (counter-clockwise, padded to 4 vertices), ship at the origin heading +x
    sq = [[4.0, -1.0], [4.0, 1.0], [6.0, 1.0], [6.0, -1.0]]
    tri = [[10.0, 2.0], [12.0, 2.0], [10.0, 4.0], [0.0, 0.0]]
    verts = torch.tensor([sq, tri], dtype=torch.float32, device="cuda")
    nv = torch.tensor([4, 3], dtype=torch.int32, device="cuda")
    thick = torch.tensor([1.0, 2.0], dtype=torch.float32, device="cuda")
    poses = torch.tensor([[0.0, 0.0, 0.0]], dtype=torch.float32, device="cuda")
    r = sur.raster_poly(lib, verts, nv, thick, poses, G=16, cell=1.0, x_lo=0.0).float().cpu().numpy()[0]
    conc, th = r[..., 0], r[..., 1]
    # the square covers cells i = 4, 5 and j = 7, 8 exactly (v from -8)
    assert np.all(conc[4:6, 7:9] == 1.0) and np.all(th[4:6, 7:9] == 1.0)
    # every cell's coverage = its share of the 16 sample points inside (brute force)
    lo1, hi1 = _poly_counts(sq, 16, 1.0, 0.0)
    lo2, hi2 = _poly_counts(tri[:3], 16, 1.0, 0.0)
    got = np.rint(conc * 16).astype(int)
    assert np.all(got >= np.minimum(lo1 + lo2, 16)) and np.all(got <= np.minimum(hi1 + hi2, 16))
    assert np.all(th[(hi2 > 0) & (hi1 == 0) & (got > 0)] == 2.0)
    assert conc[0, 0] == 0.0 and th[0, 0] == 0.0
    # rotating the world by +90 degrees and the ship with it reproduces the raster
    rot = torch.tensor([[0.0, -1.0], [1.0, 0.0]], device="cuda")
    pose2 = torch.tensor([[0.0, 0.0, np.pi / 2]], dtype=torch.float32, device="cuda")
    r2 = sur.raster_poly(lib, verts @ rot.T, nv, thick, pose2, G=16, cell=1.0, x_lo=0.0).float().cpu().numpy()[0]
    assert np.abs(r2[..., 0] - conc).max() <= 4 / 16 and abs(r2[..., 0].sum() - conc.sum()) <= 8 / 16


def test_network_forward_backward_matches_oracle():
    from icedem import surrogate as sur
    arch = sur.Arch(G=16, cin=2, chans=(8, 16, 16), extra=10, hidden=(32, 24), out=3)
    net = sur.Surrogate(arch, seed=4)
    params = net.params()
    rng = np.random.default_rng(11)
    Bn = 37
    raster = osur.bf16(rng.uniform(0, 1, (Bn, 16, 16, 2)))
    extra = rng.normal(0, 1, (Bn, 10)).astype(np.float32)
    oarch = osur.Arch(G=16, cin=2, chans=(8, 16, 16), extra=10, hidden=(32, 24), out=3)
    onet = osur.Net(oarch, params, round_bf16=True)
    yo = onet.forward(raster, extra)
    y = net.forward(_bf(raster), torch.as_tensor(extra, device="cuda")).cpu().numpy()
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print("forward rel err", rel, flush=True)
    assert rel < 2e-3
    dy = rng.normal(0, 1, (Bn, 3)).astype(np.float32)
    ddy = torch.as_tensor(dy, device="cuda")
    net.backward(ddy, ddy.to(torch.bfloat16))
    g = net.grads()
    go = onet.backward(dy.astype(np.float64))
    for k in g:
        e = np.linalg.norm(g[k] - go[k]) / max(np.linalg.norm(go[k]), 1e-12)
        print(f"grad {k}: rel err {e:.2e}", flush=True)
        assert e < 2e-2, k


def test_network_implicit_conv_equals_im2col_path():
    from icedem import surrogate as sur
    net = sur.Surrogate(sur.Arch(), seed=3)
    rng = np.random.default_rng(5)
    raster = _bf(rng.uniform(0, 1, (33, 64, 64, 2)))
    extra = torch.as_tensor(rng.normal(0, 1, (33, 9)).astype(np.float32), device="cuda")
    y1 = net.forward(raster, extra).clone()
    net.implicit_conv = False
    y2 = net.forward(raster, extra).clone()
    np.testing.assert_array_equal(y1.cpu().numpy(), y2.cpu().numpy())


def test_training_reduces_loss():
    from icedem import surrogate as sur
    arch = sur.Arch(G=16, cin=2, chans=(8, 16, 16), extra=10, hidden=(32, 24), out=3)
    net = sur.Surrogate(arch, seed=1)
    rng = np.random.default_rng(12)
    Bn = 256
    raster = _bf(rng.uniform(0, 1, (Bn, 16, 16, 2)))
    extra_np = rng.normal(0, 1, (Bn, 10)).astype(np.float32)
    extra = torch.as_tensor(extra_np, device="cuda")
    # a smooth target of the inputs
    r = raster.float().cpu().numpy()
    tgt = np.stack([r[..., 0].mean((1, 2)) * 3, extra_np[:, 0] * 0.5, (r[..., 1] * r[..., 0]).mean((1, 2))], 1)
    t = torch.as_tensor(tgt.astype(np.float32), device="cuda")
    first = net.train_step(raster, extra, t, lr=3e-3).item()
    for _ in range(200):
        loss = net.train_step(raster, extra, t, lr=3e-3)
    last = loss.item()
    print("loss", first, "->", last, flush=True)
    assert last < 0.2 * first


def test_shared_first_layer_equals_full_product():
    """predict_shared's first layer: the shared feature row's product as the
    accumulators' start, then the extra inputs only -- bit-identical to the
    full product over the broadcast features, and the row start used alone
    (ICESUR_ROWINIT with K = 0) returns it unchanged."""
    from icedem.surrogate import Arch, Surrogate
    arch = Arch()
    net = Surrogate(arch, seed=3)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    ras = torch.rand((1, arch.G, arch.G, arch.cin), device="cuda", generator=g).to(torch.bfloat16)
    ext = torch.randn((3000, arch.extra), device="cuda", generator=g)
    fast = net.predict_shared(ras, ext, chunk=1024)
    full = net.predict_shared(ras, ext, chunk=1024, shared_first_layer=False)
    torch.cuda.synchronize()
    assert torch.equal(fast, full)
    assert torch.isfinite(fast).all() and fast.abs().max() > 0
    # row start alone
    L, lib = _lib()
    init = torch.randn(45, device="cuda", generator=g)
    a = torch.zeros((7, 8), device="cuda", dtype=torch.bfloat16)
    b = torch.zeros((8, 45), device="cuda", dtype=torch.bfloat16)
    c = torch.empty((7, 45), device="cuda", dtype=torch.float32)
    assert lib.icesur_gemm(0, 0, 7, 45, 0, _p(a), 8, _p(b), 45, _p(c), 45, None, 0, None, 8, 1, _p(init), _s()) == 0
    torch.cuda.synchronize()
    assert torch.equal(c, init.expand(7, 45))


@pytest.mark.parametrize("B", [1, 33, 300])
def test_fused_first_convolutions_equal_separate(B):
    """icesur_conv12 (inference forward) == the two icesur_conv launches, bit
    for bit, and the network's predict equals its training-mode forward."""
    from icedem import surrogate as sur
    net = sur.Surrogate(sur.Arch(), seed=7)
    rng = np.random.default_rng(B)
    raster = _bf(rng.uniform(0, 1, (B, 64, 64, 2)))
    extra = torch.as_tensor(rng.normal(0, 1, (B, 9)).astype(np.float32), device="cuda")
    fused = net.predict(raster, extra, chunk=max(B, 1)).clone()
    full = net.forward(raster, extra).clone()          # keep=True: separate launches
    np.testing.assert_array_equal(fused.cpu().numpy(), full.cpu().numpy())
    net._cnn(raster, keep=False)
    a12 = net._buf("act1", B * 16 * 16 * 64, torch.bfloat16).clone()
    net._cnn(raster, keep=True)
    np.testing.assert_array_equal(a12.view(torch.int16).cpu().numpy(), net._acts[1].view(torch.int16).cpu().numpy())
