"""NumPy restatement of the rollout-cost surrogate (test infrastructure only).

There is no reference counterpart: the reference has no learned model.  This
module states the network the GPU kernels compute -- 3x3 stride-2 padding-1
convolutions with ReLU over an NHWC raster, the flattened features
concatenated with extra inputs, then ReLU MLP layers and a linear head --
with bf16 rounding (round to nearest even) at the same points as the device
path (inputs, weights, every activation that feeds a matrix product, and the
gradients that feed one), and float64 accumulation.  Backward is written out
by hand; `numeric_grad` checks it by central differences.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even); returns float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).copy()
    out[nan] = np.nan
    return out.reshape(a.shape)


def im2col(x: np.ndarray) -> np.ndarray:
    """(B, H, W, C) -> (B*Ho*Wo, 9*C) patches of a 3x3 stride-2 padding-1 conv."""
    B, H, W, C = x.shape
    Ho, Wo = (H + 1) // 2, (W + 1) // 2
    xp = np.zeros((B, H + 2, W + 2, C), dtype=x.dtype)
    xp[:, 1:H + 1, 1:W + 1] = x
    cols = np.empty((B, Ho, Wo, 9, C), dtype=x.dtype)
    for kh in range(3):
        for kw in range(3):
            cols[:, :, :, kh * 3 + kw] = xp[:, kh:kh + 2 * Ho:2, kw:kw + 2 * Wo:2]
    return cols.reshape(B * Ho * Wo, 9 * C)


def col2im(dcol: np.ndarray, shape) -> np.ndarray:
    """Adjoint of im2col: (B*Ho*Wo, 9*C) -> (B, H, W, C)."""
    B, H, W, C = shape
    Ho, Wo = (H + 1) // 2, (W + 1) // 2
    d = dcol.reshape(B, Ho, Wo, 9, C)
    dxp = np.zeros((B, H + 2, W + 2, C), dtype=np.float64)
    for kh in range(3):
        for kw in range(3):
            dxp[:, kh:kh + 2 * Ho:2, kw:kw + 2 * Wo:2] += d[:, :, :, kh * 3 + kw]
    return dxp[:, 1:H + 1, 1:W + 1]


@dataclass
class Arch:
    G: int = 64                    # raster side
    cin: int = 2
    chans: Tuple[int, ...] = (32, 64, 128, 128)
    extra: int = 86                # ship state 6 + controls 80
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


def init_params(arch: Arch, seed: int = 0) -> Dict[str, np.ndarray]:
    """He-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    out = {}
    for k, shp in arch.param_shapes().items():
        if len(shp) == 2:
            lim = np.sqrt(6.0 / shp[0])
            out[k] = rng.uniform(-lim, lim, shp).astype(np.float32)
        else:
            out[k] = np.zeros(shp, np.float32)
    return out


class Net:
    """Forward and backward of the surrogate; `round_bf16` applies the
    device's bf16 rounding points (False: exact float64 network)."""

    def __init__(self, arch: Arch, params: Dict[str, np.ndarray], round_bf16: bool = True):
        self.arch = arch
        self.p = {k: np.asarray(v, np.float64) for k, v in params.items()}
        self.rb = round_bf16

    def _r(self, x):
        return bf16(x).astype(np.float64) if self.rb else np.asarray(x, np.float64)

    def forward(self, raster: np.ndarray, extra: np.ndarray) -> np.ndarray:
        a = self.arch
        B = raster.shape[0]
        x = self._r(raster)
        self.cache = {"x": [], "cols": [], "acts": []}
        for i in range(len(a.chans)):
            col = im2col(x)
            y = np.maximum(col @ self._r(self.p[f"cw{i}"]) + self.p[f"cb{i}"], 0.0)
            self.cache["x"].append(x.shape)
            self.cache["cols"].append(col)
            x = self._r(y).reshape(B, (x.shape[1] + 1) // 2, (x.shape[2] + 1) // 2, a.chans[i])
            self.cache["acts"].append(x)
        z = np.concatenate([x.reshape(B, -1), self._r(extra)], axis=1)
        self.cache["z"] = [z]
        nl = len(a.hidden) + 1
        for i in range(nl):
            y = z @ self._r(self.p[f"fw{i}"]) + self.p[f"fb{i}"]
            if i < nl - 1:
                y = np.maximum(y, 0.0)
                z = self._r(y)
                self.cache["z"].append(z)
        self.y = y
        return y

    def backward(self, dy: np.ndarray) -> Dict[str, np.ndarray]:
        """Gradients of sum(dy * y) for the last forward."""
        a = self.arch
        g = {}
        nl = len(a.hidden) + 1
        d = np.asarray(dy, np.float64)
        for i in reversed(range(nl)):
            zin = self.cache["z"][i]
            db = self._r(d)
            g[f"fw{i}"] = zin.T @ db
            g[f"fb{i}"] = d.sum(axis=0)
            dz = db @ self._r(self.p[f"fw{i}"]).T
            if i > 0:
                d = dz * (self.cache["z"][i] > 0)
        F = a.feat()
        B = dy.shape[0]
        d = dz[:, :F]
        for i in reversed(range(len(a.chans))):
            act = self.cache["acts"][i]
            d = d.reshape(act.shape) * (act > 0)
            dflat = d.reshape(-1, a.chans[i])
            db = self._r(dflat)
            g[f"cw{i}"] = self.cache["cols"][i].T @ db
            g[f"cb{i}"] = dflat.sum(axis=0)
            if i > 0:
                dcol = db @ self._r(self.p[f"cw{i}"]).T
                d = col2im(dcol, self.cache["x"][i])
        return g


def numeric_grad(arch: Arch, params, raster, extra, dy, key, idx, h=1e-5) -> float:
    """Central difference of sum(dy * y) w.r.t. params[key][idx] (float64 net)."""
    p1 = {k: np.array(v, np.float64) for k, v in params.items()}
    p2 = {k: np.array(v, np.float64) for k, v in params.items()}
    p1[key][idx] += h
    p2[key][idx] -= h
    f1 = (Net(arch, p1, False).forward(raster, extra) * dy).sum()
    f2 = (Net(arch, p2, False).forward(raster, extra) * dy).sum()
    return (f1 - f2) / (2 * h)
