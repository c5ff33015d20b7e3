"""Rollout-cost surrogate: CNN over an ice raster + MLP (no reference counterpart).

Predicts a DEM rollout's (impulse, KE loss, peak |F|) from
  - a G x G x 2 raster of the ice ahead of the ship (area fraction, mean
    thickness; bf16), built on the GPU by ``icesur_raster``,
  - the ship's initial state (6 values) and the rollout's control knots
    (n_knots x 2) seen through the ice-free ship model: 3 values of the
    controls' ship-only rollout (KE loss, KE-costmap integral, path length),
    fp32; the raw knots can be appended as well.
Network: 3x3 stride-2 convolutions (32/64/128/128 channels, ReLU) -> 4x4x128
features, concatenated with the 9 extra inputs -> MLP 512 -> 512 -> 3.

Every floating-point operation of the forward pass, the backward pass and
the Adam update runs in this package's HIP kernels (icesur.h): GEMMs in
BF16 on the matrix cores with fp32 accumulation, fp32 master weights.
Torch tensors are only device buffers here.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib

RELU, ACCUM, BIAS, ROWINIT = 1, 2, 4, 8
TK = 32   # K step of the GEMM kernels


def _p(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _rup(x: int, m: int) -> int:
    return (x + m - 1) // m * m


@dataclass
class Arch:
    G: int = 64
    cin: int = 2
    chans: Tuple[int, ...] = (32, 64, 128, 128)
    extra: int = 9
    hidden: Tuple[int, ...] = (512, 512)
    out: int = 3

    def conv_shapes(self):
        s, c, shapes = self.G, self.cin, []
        for co in self.chans:
            shapes.append((s, c, co))
            s, c = (s + 1) // 2, co
        return shapes, s, c

    def feat(self) -> int:
        _, s, c = self.conv_shapes()
        return s * s * c

    def param_shapes(self) -> Dict[str, tuple]:
        p = {}
        shapes, _, _ = self.conv_shapes()
        for i, (_, c, co) in enumerate(shapes):
            p[f"cw{i}"] = (9 * c, co)
            p[f"cb{i}"] = (co,)
        dims = [self.feat() + self.extra] + list(self.hidden) + [self.out]
        for i in range(len(dims) - 1):
            p[f"fw{i}"] = (dims[i], dims[i + 1])
            p[f"fb{i}"] = (dims[i + 1],)
        return p

    def flops_per_sample(self) -> float:
        """Multiply-adds x 2 of one forward pass."""
        shapes, _, _ = self.conv_shapes()
        mac = sum(((s + 1) // 2) ** 2 * 9 * c * co for s, c, co in shapes)
        dims = [self.feat() + self.extra] + list(self.hidden) + [self.out]
        mac += sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
        return 2.0 * mac


class Surrogate:
    """Parameters, Adam state and work buffers of the surrogate on one device."""

    def __init__(self, arch: Optional[Arch] = None, device=None, seed: int = 0,
                 params: Optional[Dict[str, np.ndarray]] = None):
        if not torch.cuda.is_available():
            raise _lib.IceDemError("the surrogate needs a ROCm GPU")
        self.lib = _lib.load()
        self.arch = arch or Arch()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        shapes = self.arch.param_shapes()
        self.names = list(shapes)
        self.offsets = {}
        off = 0
        for k, shp in shapes.items():
            n = int(np.prod(shp))
            self.offsets[k] = (off, shp)
            off = _rup(off + n, 8)          # every parameter 16-byte aligned in bf16
        self.nparam = off
        dev = self.device
        self.w = torch.zeros(off, dtype=torch.float32, device=dev)
        self.g = torch.zeros(off, dtype=torch.float32, device=dev)
        self.m = torch.zeros(off, dtype=torch.float32, device=dev)
        self.v = torch.zeros(off, dtype=torch.float32, device=dev)
        self.wb = torch.zeros(off, dtype=torch.bfloat16, device=dev)
        self.t = 0
        self.load(params if params is not None else self.init_params(seed))
        self._bufs: Dict[Tuple[str, int], torch.Tensor] = {}
        self.out_w = torch.ones(self.arch.out, dtype=torch.float32, device=dev)
        self.implicit_conv = True     # False: im2col + GEMM (the same sums, for tests)
        self.fuse_convs = True        # inference: conv 1 + 2 in one launch (the same sums)
        self.loss = torch.zeros(1, dtype=torch.float32, device=dev)

    # ── parameters ──────────────────────────────────────────────────────────
    def init_params(self, seed: int) -> Dict[str, np.ndarray]:
        """He-uniform weights, zero biases."""
        rng = np.random.default_rng(seed)
        out = {}
        for k, (_, shp) in self.offsets.items():
            if len(shp) == 2:
                lim = math.sqrt(6.0 / shp[0])
                out[k] = rng.uniform(-lim, lim, shp).astype(np.float32)
            else:
                out[k] = np.zeros(shp, np.float32)
        return out

    def load(self, params: Dict[str, np.ndarray]):
        host = np.zeros(self.nparam, np.float32)
        for k, (o, shp) in self.offsets.items():
            host[o:o + int(np.prod(shp))] = np.asarray(params[k], np.float32).reshape(-1)
        self.w.copy_(torch.as_tensor(host))
        self._chk(self.lib.icesur_to_bf16(_p(self.w), self.nparam, _p(self.wb), self._s()), "to_bf16")

    def params(self) -> Dict[str, np.ndarray]:
        host = self.w.cpu().numpy()
        return {k: host[o:o + int(np.prod(shp))].reshape(shp).copy() for k, (o, shp) in self.offsets.items()}

    def grads(self) -> Dict[str, np.ndarray]:
        host = self.g.cpu().numpy()
        return {k: host[o:o + int(np.prod(shp))].reshape(shp).copy() for k, (o, shp) in self.offsets.items()}

    def _view(self, t: torch.Tensor, k: str) -> torch.Tensor:
        o, shp = self.offsets[k]
        return t[o:o + int(np.prod(shp))]

    # ── plumbing ────────────────────────────────────────────────────────────
    def _s(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _chk(self, rc, what):
        _lib.check(rc, None, what)

    def _buf(self, name: str, n: int, dtype) -> torch.Tensor:
        key = (name, 0)
        b = self._bufs.get(key)
        if b is None or b.numel() < n or b.dtype != dtype:
            b = torch.empty(max(n, 1), dtype=dtype, device=self.device)
            self._bufs[key] = b
        return b[:n]

    def _gemm(self, ta, tb, M, N, K, A, lda, B, ldb, Cf=None, ldc=0, Cb=None, ldcb=0, bias=None, flags=0,
              split=1):
        work = None
        if split > 1:
            extra = ((split + 31) // 32) * M * N if split > 32 else 0
            work = self._buf("splitk", split * M * N + extra, torch.float32)
        self._chk(self.lib.icesur_gemm(int(ta), int(tb), M, N, K, _p(A), lda, _p(B), ldb, _p(Cf), ldc, _p(Cb), ldcb,
                                       _p(bias), flags, split, _p(work), self._s()), "icesur_gemm")

    @staticmethod
    def _split_for(M: int, N: int, K: int) -> int:
        """Split-K factor for a long reduction with few output tiles: about
        1024 workgroups, at least 512 of K per split."""
        tiles = ((M + 63) // 64) * ((N + 63) // 64)
        return int(max(1, min((1024 + tiles - 1) // tiles, K // 512, 1024)))

    # ── forward ─────────────────────────────────────────────────────────────
    def _cnn(self, raster: torch.Tensor, keep: bool = True) -> torch.Tensor:
        """(B, G, G, cin) bf16 -> (B, F) bf16 features (the last activation).
        keep: hold every layer's input and activation for backward; without
        it the first two convolutions of the default shapes run fused
        (icesur_conv12, bit-identical)."""
        a = self.arch
        B = int(raster.shape[0])
        assert raster.dtype == torch.bfloat16 and tuple(raster.shape[1:]) == (a.G, a.G, a.cin)
        x = raster.contiguous()
        shapes, _, _ = a.conv_shapes()
        self._acts: List[torch.Tensor] = []
        self._ins: List[torch.Tensor] = []
        i0 = 0
        if (not keep and self.fuse_convs and self.implicit_conv and len(shapes) >= 2
                and tuple(shapes[0]) == (64, 2, 32) and tuple(shapes[1]) == (32, 32, 64)):
            act = self._buf("act1", B * 16 * 16 * 64, torch.bfloat16)
            wt0 = self._buf("wt0", 18 * 32, torch.bfloat16)
            wt1 = self._buf("wt1", 288 * 64, torch.bfloat16)
            self._chk(self.lib.icesur_conv12(_p(x), B, _p(self._view(self.wb, "cw0")), _p(self._view(self.w, "cb0")),
                                             _p(self._view(self.wb, "cw1")), _p(self._view(self.w, "cb1")),
                                             BIAS | RELU, _p(wt0), _p(wt1), _p(act), self._s()), "conv12")
            x = act
            i0 = 2
        for i, (s, c, co) in enumerate(shapes):
            if i < i0:
                continue
            so = (s + 1) // 2
            rows = B * so * so
            act = self._buf(f"act{i}", rows * co, torch.bfloat16)
            if self.implicit_conv:
                wt = self._buf(f"wt{i}", 9 * c * co, torch.bfloat16)
                self._chk(self.lib.icesur_conv(_p(x), B, s, s, c, _p(self._view(self.wb, f"cw{i}")), co,
                                               _p(self._view(self.w, f"cb{i}")), BIAS | RELU, _p(wt), _p(act),
                                               self._s()), "conv")
            else:
                col = self._buf("col", rows * 9 * c, torch.bfloat16)
                self._chk(self.lib.icesur_im2col(_p(x), B, s, s, c, _p(col), self._s()), "im2col")
                self._gemm(0, 0, rows, co, 9 * c, col, 9 * c, self._view(self.wb, f"cw{i}"), co, Cb=act, ldcb=co,
                           bias=self._view(self.w, f"cb{i}"), flags=BIAS | RELU)
            self._ins.append(x)
            self._acts.append(act)
            x = act
        return x

    def _mlp(self, feat: torch.Tensor, ldf: int, extra: torch.Tensor, B: int) -> torch.Tensor:
        """Concatenate features (row stride ldf; 0 = one row for all) with the
        extra inputs, then the MLP -> (B, out) fp32."""
        a = self.arch
        F = a.feat()
        assert extra.dtype == torch.float32 and tuple(extra.shape) == (B, a.extra)
        extra = extra.contiguous()
        ldz = _rup(F + a.extra, 8)
        z = self._buf("z0", B * ldz, torch.bfloat16)
        self._chk(self.lib.icesur_concat(_p(feat), ldf, F, _p(extra), a.extra, a.extra, B, _p(z), ldz, self._s()),
                  "concat")
        dims = [F + a.extra] + list(a.hidden) + [a.out]
        self._zs = [(z, ldz)]
        nl = len(dims) - 1
        y = None
        for i in range(nl):
            zin, ldin = self._zs[-1]
            N = dims[i + 1]
            if i < nl - 1:
                h = self._buf(f"z{i + 1}", B * N, torch.bfloat16)
                self._gemm(0, 0, B, N, dims[i], zin, ldin, self._view(self.wb, f"fw{i}"), N, Cb=h, ldcb=N,
                           bias=self._view(self.w, f"fb{i}"), flags=BIAS | RELU)
                self._zs.append((h, N))
            else:
                y = self._buf("y", B * N, torch.float32)
                self._gemm(0, 0, B, N, dims[i], zin, ldin, self._view(self.wb, f"fw{i}"), N, Cf=y, ldc=N,
                           bias=self._view(self.w, f"fb{i}"), flags=BIAS)
        self._B = B
        return y.view(B, a.out)

    def forward(self, raster: torch.Tensor, extra: torch.Tensor, keep: bool = True) -> torch.Tensor:
        """raster (B, G, G, cin) bf16, extra (B, arch.extra) fp32 -> (B, out) fp32.
        keep=False: inference only (no backward after it)."""
        B = int(raster.shape[0])
        feat = self._cnn(raster, keep)
        return self._mlp(feat, self.arch.feat(), extra, B)

    def predict_shared(self, raster1: torch.Tensor, extra: torch.Tensor, chunk: int = 16384,
                       shared_first_layer: bool = True) -> torch.Tensor:
        """Many candidates that share one raster (1, G, G, cin): the CNN runs
        once, the MLP per candidate; -> (S, out) fp32.  With
        shared_first_layer (and F a multiple of the GEMM's K step) the first
        layer's feature part is one 1-row product that starts every
        candidate's accumulators (icesur_gemm ICESUR_ROWINIT); the results are
        bit-identical to the full product."""
        S = int(extra.shape[0])
        a = self.arch
        F = a.feat()
        feat = self._cnn(raster1).clone()
        out = torch.empty((S, a.out), dtype=torch.float32, device=self.device)
        fast = shared_first_layer and F % TK == 0 and len(a.hidden) > 0
        p0 = None
        if fast:
            N0 = a.hidden[0]
            p0 = torch.empty(N0, dtype=torch.float32, device=self.device)
            self._gemm(0, 0, 1, N0, F, feat, F, self._view(self.wb, "fw0"), N0, Cf=p0, ldc=N0)
        for s0 in range(0, S, chunk):
            e = min(S, s0 + chunk)
            out[s0:e] = self._mlp_rowinit(p0, extra[s0:e], e - s0) if fast else self._mlp(feat, 0, extra[s0:e], e - s0)
        return out

    def _mlp_rowinit(self, p0: torch.Tensor, extra: torch.Tensor, B: int) -> torch.Tensor:
        """The MLP of B candidates whose first-layer feature product is p0
        (fp32, one row): the first layer runs over the extra inputs only,
        its accumulators starting from p0."""
        a = self.arch
        F = a.feat()
        assert extra.dtype == torch.float32 and tuple(extra.shape) == (B, a.extra)
        extra = extra.contiguous()
        lde = _rup(a.extra, 8)
        ze = self._buf("ze", B * lde, torch.bfloat16)
        self._chk(self.lib.icesur_concat(_p(extra), 0, 0, _p(extra), a.extra, a.extra, B, _p(ze), lde, self._s()),
                  "concat")
        dims = [F + a.extra] + list(a.hidden) + [a.out]
        N0 = dims[1]
        w0 = self._view(self.wb, "fw0")[F * N0:]
        h = self._buf("z1", B * N0, torch.bfloat16)
        self._chk(self.lib.icesur_gemm(0, 0, B, N0, a.extra, _p(ze), lde, _p(w0), N0, None, 0, _p(h), N0,
                                       _p(self._view(self.w, "fb0")), BIAS | RELU | ROWINIT, 1, _p(p0), self._s()),
                  "icesur_gemm")
        zin, ldin = h, N0
        nl = len(dims) - 1
        y = None
        for i in range(1, nl):
            N = dims[i + 1]
            if i < nl - 1:
                h = self._buf(f"z{i + 1}", B * N, torch.bfloat16)
                self._gemm(0, 0, B, N, dims[i], zin, ldin, self._view(self.wb, f"fw{i}"), N, Cb=h, ldcb=N,
                           bias=self._view(self.w, f"fb{i}"), flags=BIAS | RELU)
                zin, ldin = h, N
            else:
                y = self._buf("y", B * N, torch.float32)
                self._gemm(0, 0, B, N, dims[i], zin, ldin, self._view(self.wb, f"fw{i}"), N, Cf=y, ldc=N,
                           bias=self._view(self.w, f"fb{i}"), flags=BIAS)
        return y.view(B, a.out)

    # ── backward ────────────────────────────────────────────────────────────
    def backward(self, dy: torch.Tensor, dyb: torch.Tensor):
        """Gradients (into self.g) of sum(dy * y) for the last forward;
        dyb = bf16(dy)."""
        a = self.arch
        B = self._B
        F = a.feat()
        dims = [F + a.extra] + list(a.hidden) + [a.out]
        nl = len(dims) - 1
        d, db = dy.reshape(B, -1), dyb.reshape(B, -1)
        dz = None
        for i in reversed(range(nl)):
            zin, ldin = self._zs[i]
            K, N = dims[i], dims[i + 1]
            self._gemm(1, 0, K, N, B, zin, ldin, db, N, Cf=self._view(self.g, f"fw{i}"), ldc=N,
                       split=self._split_for(K, N, B))
            self._chk(self.lib.icesur_colsum(_p(d), N, B, N, _p(self._view(self.g, f"fb{i}")), 0,
                                             _p(self._buf("csum", ((B + 255) // 256) * N, torch.float32)),
                                             self._s()), "colsum")
            ldk = ldin
            dz = self._buf(f"dz{i}", B * ldk, torch.float32)
            self._gemm(0, 1, B, K, N, db, N, self._view(self.wb, f"fw{i}"), N, Cf=dz, ldc=ldk)
            if i > 0:
                dzb = self._buf(f"dzb{i}", B * ldk, torch.bfloat16)
                self._chk(self.lib.icesur_relu_back(_p(dz), ldk, _p(zin), ldin, B, K, _p(dzb), ldk, self._s()),
                          "relu_back")
                d, db = dz, dzb
        # into the CNN: the first F columns of dz, masked by the last activation
        shapes, _, _ = a.conv_shapes()
        dfeat = self._buf("dfeat", B * F, torch.float32)
        self._chk(self.lib.icesur_split(_p(dz), self._zs[0][1], F, B, _p(dfeat), self._s()), "split")
        co_last = shapes[-1][2]
        rows_last = B * F // co_last
        dfb = self._buf("dfeatb", B * F, torch.bfloat16)
        self._chk(self.lib.icesur_relu_back(_p(dfeat), co_last, _p(self._acts[-1]), co_last, rows_last, co_last,
                                            _p(dfb), co_last, self._s()), "relu_back")
        d, db = dfeat, dfb
        for i in reversed(range(len(shapes))):
            s, c, co = shapes[i]
            so = (s + 1) // 2
            rows = B * so * so
            col = self._buf("col", rows * 9 * c, torch.bfloat16)
            self._chk(self.lib.icesur_im2col(_p(self._ins[i]), B, s, s, c, _p(col), self._s()), "im2col")
            self._gemm(1, 0, 9 * c, co, rows, col, 9 * c, db, co, Cf=self._view(self.g, f"cw{i}"), ldc=co,
                       split=self._split_for(9 * c, co, rows))
            self._chk(self.lib.icesur_colsum(_p(d), co, rows, co, _p(self._view(self.g, f"cb{i}")), 0,
                                             _p(self._buf("csum", ((rows + 255) // 256) * co, torch.float32)),
                                             self._s()), "colsum")
            if i > 0:
                dcol = self._buf("dcol", rows * 9 * c, torch.float32)
                self._gemm(0, 1, rows, 9 * c, co, db, co, self._view(self.wb, f"cw{i}"), co, Cf=dcol, ldc=9 * c)
                dx = self._buf(f"dx{i}", B * s * s * c, torch.float32)
                self._chk(self.lib.icesur_col2im(_p(dcol), B, s, s, c, _p(dx), self._s()), "col2im")
                dxb = self._buf(f"dxb{i}", B * s * s * c, torch.bfloat16)
                self._chk(self.lib.icesur_relu_back(_p(dx), c, _p(self._acts[i - 1]), c, B * s * s, c, _p(dxb), c,
                                                    self._s()), "relu_back")
                d, db = dx, dxb

    # ── training ────────────────────────────────────────────────────────────
    def train_step(self, raster, extra, target, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                   eps: float = 1e-8, weight_decay: float = 0.0) -> torch.Tensor:
        """One Adam step (decoupled weight decay when ``weight_decay`` > 0) on
        the weighted MSE; returns the loss (device tensor)."""
        y = self.forward(raster, extra)
        B = y.shape[0]
        dy = self._buf("dy", B * self.arch.out, torch.float32)
        dyb = self._buf("dyb", B * self.arch.out, torch.bfloat16)
        self._chk(self.lib.icesur_mse(_p(y), _p(target.contiguous()), _p(self.out_w), B, self.arch.out, _p(dy),
                                      _p(dyb), _p(self.loss), self._s()), "mse")
        self.backward(dy, dyb)
        self.t += 1
        self._chk(self.lib.icesur_adamw(_p(self.w), _p(self.g), _p(self.m), _p(self.v), self.nparam, lr, beta1, beta2,
                                        eps, weight_decay, self.t, _p(self.wb), self._s()), "adamw")
        return self.loss

    def predict(self, raster, extra, chunk: int = 4096) -> torch.Tensor:
        """Forward pass over any number of samples, `chunk` at a time."""
        B = raster.shape[0]
        out = torch.empty((B, self.arch.out), dtype=torch.float32, device=self.device)
        for s in range(0, B, chunk):
            e = min(B, s + chunk)
            out[s:e] = self.forward(raster[s:e], extra[s:e], keep=False)
        return out


def raster(lib, pos: torch.Tensor, area: torch.Tensor, thick: torch.Tensor, poses: torch.Tensor, G: int = 64,
           cell: float = 1.0, x_lo: float = 20.0, acc: Optional[torch.Tensor] = None) -> torch.Tensor:
    """(B, G, G, 2) bf16 rasters of one scene (pos (2, n) SoA fp32) seen from
    B ship poses (B, 3) = x, y, heading; the window starts x_lo ahead of the
    ship's centre."""
    B = int(poses.shape[0])
    n = int(area.shape[0])
    dev = poses.device
    if acc is None or acc.numel() < B * G * G * 2:
        acc = torch.zeros(B * G * G * 2, dtype=torch.int32, device=dev)
    out = torch.empty((B, G, G, 2), dtype=torch.bfloat16, device=dev)
    s = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.icesur_raster(_p(pos), _p(area), _p(thick), n, _p(poses.contiguous()), B, G, float(cell),
                                 float(x_lo), _p(acc), _p(out), s), None, "icesur_raster")
    return out


def raster_poly(lib, verts: torch.Tensor, nv: torch.Tensor, thick: torch.Tensor, poses: torch.Tensor, G: int = 64,
                cell: float = 0.25, x_lo: float = 26.0, acc: Optional[torch.Tensor] = None) -> torch.Tensor:
    """(B, G, G, 2) bf16 rasters of one scene's convex polygons (verts (n,
    max_v, 2) world vertices fp32, nv (n,) int32 vertex counts), 4 x 4 samples
    per cell, seen from B ship poses (B, 3); the window starts x_lo ahead of
    the ship's centre and spans +-G cell / 2 to either side."""
    B = int(poses.shape[0])
    n, max_v = int(verts.shape[0]), int(verts.shape[1])
    dev = poses.device
    if acc is None or acc.numel() < B * G * G * 2:
        acc = torch.zeros(B * G * G * 2, dtype=torch.int32, device=dev)
    out = torch.empty((B, G, G, 2), dtype=torch.bfloat16, device=dev)
    s = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _lib.check(lib.icesur_raster_poly(_p(verts.contiguous()), _p(nv.contiguous()), _p(thick), n, max_v,
                                      _p(poses.contiguous()), B, G, float(cell), float(x_lo), _p(acc), _p(out), s),
               None, "icesur_raster_poly")
    return out
