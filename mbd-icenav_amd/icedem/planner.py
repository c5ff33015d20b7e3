"""Surrogate-plus-DEM rollout scoring for sampling planners (no reference counterpart).

The reference's MPPI/MBD planners score candidates with a 3-DOF ship model
on a costmap (``jax_mppi_planner.py:34-115``); here candidates are scored by
the DEM itself (``rollouts``) and, at planner scale, by a learned surrogate
(``surrogate``) trained on DEM rollouts:

  * ``make_dataset``: DEM rollouts over rollout patches with random start
    states and random control sequences; the surrogate's inputs are the
    ship-frame ice raster, a 6-value ship state and the control knots.
  * ``train``: Adam on the weighted MSE of standardised targets, on the GPU.
  * ``Planner.iterate``: one sampling iteration -- S candidates scored by the
    surrogate (one CNN pass for the shared raster, the MLP per candidate), the
    k best re-scored by the batched DEM as judge; when the surrogate's error
    on the judged set exceeds a threshold, the next k by surrogate rank are
    judged as well (fallback), up to a budget.

Rollout cost is ``rollouts.rollout_cost``: impulse / 1e5 N s + KE loss / 1e6 J.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .rollouts import J_REF, KE_REF, RolloutBatch, ShipSpec, nominal_controls, rollout_cost
from .scenario import Scene
from .surrogate import Arch, Surrogate, raster, raster_poly

PEAK_REF = 1.0e6     # N
# raster: 64 x 64 cells of 0.5 m from 26 m ahead of the ship's centre (4 m
# behind the bow of the 60 m hull), +-16 m to either side
RASTER_G, RASTER_CELL, RASTER_X0, RASTER_KIND = 64, 0.5, 26.0, "disc"


@dataclass
class StartSpread:
    """Random start states around the patch's nominal start."""
    u: tuple = (1.0, 3.0)          # surge speed, m/s
    dpsi: float = 0.15             # heading offset, rad
    dx: float = 8.0                # lateral offset, m
    dy: float = 2.0                # along-track offset, m


def ship_features(ship: torch.Tensor, start_xy: torch.Tensor) -> torch.Tensor:
    """(R, 15) ship states -> (R, 6): u/3, v, 10 r, 5 (psi - pi/2), dx/10, dy/2
    relative to the patch's nominal start."""
    f = torch.empty((ship.shape[0], 6), dtype=torch.float32, device=ship.device)
    f[:, 0] = ship[:, 3] / 3.0
    f[:, 1] = ship[:, 4]
    f[:, 2] = ship[:, 5] * 10.0
    f[:, 3] = (ship[:, 2] - math.pi / 2) * 5.0
    f[:, 4] = (ship[:, 0] - start_xy[:, 0]) / 10.0
    f[:, 5] = (ship[:, 1] - start_xy[:, 1]) / 2.0
    return f


class PatchRasters:
    """Device copies of each patch's initial element polygons (or centroids
    and areas for ``kind="disc"``) and thicknesses, for ship-frame rasters."""

    def __init__(self, rb: RolloutBatch, kind: str = None, cell: float = None, x0: float = None):
        dev = rb.device
        self.lib = _lib.load()
        self.kind = kind or RASTER_KIND
        self.cell = RASTER_CELL if cell is None else cell
        self.x0 = RASTER_X0 if x0 is None else x0
        self.items = []
        for r, eng in enumerate(rb.engines):
            sc = rb.scenes[rb.inst[r]]
            thick = torch.as_tensor(eng._host["thick"], device=dev)
            if self.kind == "poly":
                mv = max(len(p) for p in sc.polygons)
                verts = np.zeros((len(sc.polygons), mv, 2), np.float32)
                for i, p in enumerate(sc.polygons):
                    verts[i, :len(p)] = p
                nv = np.array([len(p) for p in sc.polygons], np.int32)
                self.items.append((torch.as_tensor(verts, device=dev), torch.as_tensor(nv, device=dev), thick))
            else:
                c = np.asarray(sc.centroids, np.float32)
                pos = torch.as_tensor(np.concatenate([c[:, 0], c[:, 1]]), device=dev)
                area = torch.as_tensor(eng._host["areas"], device=dev)
                self.items.append((pos, area, thick))
        self.acc = torch.zeros(RASTER_G * RASTER_G * 2 * 64, dtype=torch.int32, device=dev)

    def of(self, r: int, poses: torch.Tensor) -> torch.Tensor:
        need = poses.shape[0] * RASTER_G * RASTER_G * 2
        if self.acc.numel() < need:
            self.acc = torch.zeros(need, dtype=torch.int32, device=poses.device)
        a, b, thick = self.items[r]
        if self.kind == "poly":
            return raster_poly(self.lib, a, b, thick, poses, RASTER_G, self.cell, self.x0, self.acc)
        return raster(self.lib, a, b, thick, poses, RASTER_G, self.cell, self.x0, self.acc)


def random_starts(rb: RolloutBatch, nominal: torch.Tensor, rng: np.random.Generator, spread: StartSpread):
    """(R, 15) start states around each instance's nominal start (R, 15)."""
    R = rb.R
    s = nominal.clone()
    s[:, 0] += torch.as_tensor(rng.uniform(-spread.dx, spread.dx, R), dtype=torch.float32, device=s.device)
    s[:, 1] += torch.as_tensor(rng.uniform(-spread.dy, spread.dy, R), dtype=torch.float32, device=s.device)
    s[:, 2] += torch.as_tensor(rng.uniform(-spread.dpsi, spread.dpsi, R), dtype=torch.float32, device=s.device)
    s[:, 3] = torch.as_tensor(rng.uniform(*spread.u, R), dtype=torch.float32, device=s.device)
    return s


def random_controls(R: int, n_knots: int, rng: np.random.Generator) -> np.ndarray:
    """Nominal (u ~ U[0.2, 1], r ~ U[-0.5, 0.5]) per rollout plus N(0, 0.2^2) per knot, clipped."""
    nom = np.stack([rng.uniform(0.2, 1.0, R), rng.uniform(-0.5, 0.5, R)], 1)[:, None, :]
    c = nom + rng.normal(0.0, 0.2, (R, n_knots, 2))
    return np.clip(c, -1.0, 1.0).astype(np.float32)


def targets_of(out: np.ndarray) -> np.ndarray:
    """(R, 4) rollout outputs -> (R, 3) targets [log1p(impulse / J_REF),
    KE loss / KE_REF, log1p(peak / PEAK_REF)]."""
    out = np.asarray(out, np.float64)
    return np.stack([np.log1p(np.maximum(out[:, 0], 0.0) / J_REF), out[:, 1] / KE_REF,
                     np.log1p(np.maximum(out[:, 2], 0.0) / PEAK_REF)], 1).astype(np.float32)


def cost_of_targets(t) -> np.ndarray:
    """Rollout cost (impulse / J_REF + KE loss / KE_REF) from targets."""
    t = np.asarray(t, np.float64)
    return np.expm1(t[:, 0]) + t[:, 1]


COST_FLOOR = 1.0   # cost units; relative errors use max(|cost|, COST_FLOOR)


def make_dataset(rb: RolloutBatch, rounds: int, frames: int = 400, hold: int = 10, n_knots: int = 40,
                 seed: int = 0, spread: Optional[StartSpread] = None, progress=None) -> Dict[str, torch.Tensor]:
    """`rounds` x R DEM rollouts (one per patch instance per round)."""
    spread = spread or StartSpread()
    rng = np.random.default_rng(seed)
    dev = rb.device
    pr = PatchRasters(rb)
    nominal = rb.ship_init.clone()
    start_xy = nominal[:, :2].clone()
    R = rb.R
    N = rounds * R
    ras = torch.empty((N, RASTER_G, RASTER_G, 2), dtype=torch.bfloat16, device=dev)
    extra = torch.empty((N, 6 + 2 * n_knots), dtype=torch.float32, device=dev)
    tgt = np.empty((N, 3), np.float32)
    for k in range(rounds):
        starts = random_starts(rb, nominal, rng, spread)
        ctl = random_controls(R, n_knots, rng)
        rb.ship_init.copy_(starts)
        rb.reset()
        out = rb.run(torch.as_tensor(ctl), hold, frames).cpu().numpy()
        sl = slice(k * R, (k + 1) * R)
        for r in range(R):
            ras[k * R + r] = pr.of(r, starts[r:r + 1, :3])[0]
        extra[sl, :6] = ship_features(starts, start_xy)
        extra[sl, 6:] = torch.as_tensor(ctl.reshape(R, -1), device=dev)
        tgt[sl] = targets_of(out)
        if progress and (k + 1) % max(1, rounds // 10) == 0:
            progress(k + 1, rounds)
    rb.ship_init.copy_(nominal)
    return {"raster": ras, "extra": extra, "target": torch.as_tensor(tgt, device=dev)}


@dataclass
class Norm:
    mean: np.ndarray
    std: np.ndarray

    def fwd(self, t: torch.Tensor) -> torch.Tensor:
        m = torch.as_tensor(self.mean, dtype=torch.float32, device=t.device)
        s = torch.as_tensor(self.std, dtype=torch.float32, device=t.device)
        return (t - m) / s

    def inv(self, y: torch.Tensor) -> torch.Tensor:
        m = torch.as_tensor(self.mean, dtype=torch.float32, device=y.device)
        s = torch.as_tensor(self.std, dtype=torch.float32, device=y.device)
        return y * s + m


def mirror_signs(n_knots: int = 40, device=None) -> torch.Tensor:
    """Signs of the extra inputs under the mirror image about the ship's
    centre line: sway, yaw rate, heading offset, lateral offset and the yaw
    commands change sign; surge, along-track offset and thrust do not."""
    s = torch.ones(6 + 2 * n_knots, dtype=torch.float32, device=device)
    s[1:5] = -1.0
    s[7::2] = -1.0
    return s


def train(net: Surrogate, data: Dict[str, torch.Tensor], epochs: int = 10, batch: int = 512, lr: float = 1e-3,
          holdout: float = 0.05, seed: int = 0, weight_decay: float = 0.0, mirror: bool = False, progress=None):
    """Adam on standardised targets; returns (norm, history) with the
    held-out relative cost error after every epoch.  ``mirror`` replaces a
    random half of every batch by its mirror image (the hull, the contact
    model and the drag are symmetric about the ship's centre line, so the
    mirrored input has the same outcome distribution)."""
    N = data["raster"].shape[0]
    rng = np.random.default_rng(seed)
    perm = rng.permutation(N)
    nval = max(1, int(N * holdout))
    val, tr = perm[:nval], perm[nval:]
    t_np = data["target"].cpu().numpy()
    norm = Norm(t_np[tr].mean(0), t_np[tr].std(0) + 1e-6)
    tgt = norm.fwd(data["target"])
    dev = data["raster"].device
    signs = mirror_signs((data["extra"].shape[1] - 6) // 2, dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    hist = []
    steps = max(1, len(tr) // batch)
    for ep in range(epochs):
        order = torch.as_tensor(rng.permutation(tr), device=dev)
        lr_ep = lr * 0.5 * (1.0 + math.cos(math.pi * ep / max(1, epochs)))
        for s in range(steps):
            idx = order[s * batch:(s + 1) * batch]
            ras, ext = data["raster"][idx], data["extra"][idx]
            if mirror:
                flip = torch.rand(idx.shape[0], device=dev, generator=gen) < 0.5
                ras = torch.where(flip[:, None, None, None], ras.flip(2), ras)
                ext = torch.where(flip[:, None], ext * signs, ext)
            loss = net.train_step(ras, ext, tgt[idx], lr=lr_ep, weight_decay=weight_decay)
        err = relative_cost_error(net, norm, data, val)
        hist.append({"epoch": ep + 1, "loss": float(loss.item()), "val_rel_cost_err": err})
        if progress:
            progress(hist[-1])
    return norm, hist


def relative_cost_error(net: Surrogate, norm: Norm, data, idx) -> Dict[str, float]:
    """Relative error of the predicted rollout cost on the given samples."""
    ii = torch.as_tensor(np.asarray(idx), device=data["raster"].device)
    y = norm.inv(net.predict(data["raster"][ii], data["extra"][ii])).cpu().numpy()
    c_pred = cost_of_targets(y)
    c_true = cost_of_targets(data["target"][ii].cpu().numpy())
    rel = np.abs(c_pred - c_true) / np.maximum(np.abs(c_true), COST_FLOOR)
    return {"median": float(np.median(rel)), "mean": float(np.mean(rel)), "p90": float(np.percentile(rel, 90))}


class Planner:
    """One patch, one start state: S candidates scored by the surrogate, the
    best k judged by the batched DEM (k instances of the patch)."""

    def __init__(self, net: Surrogate, norm: Norm, scene: Scene, start: torch.Tensor, k: int = 64,
                 frames: int = 400, hold: int = 10, n_knots: int = 40, err_threshold: float = 0.05,
                 max_judged: int = 256, cell_size_factor: float = 1.0):
        self.net, self.norm = net, norm
        self.k, self.frames, self.hold, self.nk = k, frames, hold, n_knots
        self.thr, self.max_judged = err_threshold, max_judged
        self.cell_size_factor = cell_size_factor
        self._ex = None
        self.judge = RolloutBatch([scene], [0] * k, cell_size_factor=cell_size_factor)
        self.judge.set_mode("frames")
        self.nominal_start = self.judge.ship_init[0].clone()
        self.start = start.to(self.judge.device).float()
        self.judge.ship_init.copy_(self.start.expand(k, 15))
        pr = PatchRasters(self.judge)
        self.ras = pr.of(0, self.start[None, :3]).contiguous()
        self.feat = ship_features(self.start[None], self.nominal_start[None, :2])

    def surrogate_costs(self, ctl: torch.Tensor) -> torch.Tensor:
        """(S, n_knots, 2) -> (S,) predicted costs (one CNN pass for the shared raster)."""
        S = ctl.shape[0]
        extra = torch.empty((S, 6 + 2 * self.nk), dtype=torch.float32, device=ctl.device)
        extra[:, :6] = self.feat
        extra[:, 6:] = ctl.reshape(S, -1)
        y = self.norm.inv(self.net.predict_shared(self.ras, extra))
        return torch.expm1(y[:, 0]) + y[:, 1]

    def dem_costs(self, ctl: torch.Tensor) -> np.ndarray:
        """DEM costs of up to k candidates (padded with the first)."""
        m = ctl.shape[0]
        full = torch.empty((self.k, self.nk, 2), dtype=torch.float32, device=ctl.device)
        full[:m] = ctl
        full[m:] = ctl[:1]
        self.judge.reset()
        out = self.judge.run(full, self.hold, self.frames).cpu().numpy()
        return rollout_cost(out[:m])

    def iterate(self, nominal: np.ndarray, S: int = 16384, sigma: float = 0.2, seed: int = 0) -> Dict:
        """One sampling iteration around `nominal` (n_knots, 2)."""
        rng = np.random.default_rng(seed)
        dev = self.judge.device
        cand = np.clip(nominal[None] + rng.normal(0.0, sigma, (S, self.nk, 2)), -1.0, 1.0).astype(np.float32)
        cand[0] = nominal
        ctl = torch.as_tensor(cand, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pred = self.surrogate_costs(ctl)
        order = torch.argsort(pred).cpu().numpy()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        judged, dem, prd = [], [], []
        err = float("inf")
        pos = 0
        while pos < len(order) and len(judged) < self.max_judged:
            sel = order[pos:pos + self.k]
            pos += self.k
            c = self.dem_costs(ctl[torch.as_tensor(sel, device=dev)])
            p = pred[torch.as_tensor(sel, device=dev)].cpu().numpy()
            judged.extend(sel.tolist()); dem.extend(c.tolist()); prd.extend(p.tolist())
            d, q = np.asarray(dem), np.asarray(prd)
            err = float(np.median(np.abs(q - d) / np.maximum(np.abs(d), COST_FLOOR)))
            if err <= self.thr:
                break
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        dem = np.asarray(dem)
        b = int(np.argmin(dem))
        return {"best_index": int(judged[b]), "best_cost": float(dem[b]), "best_controls": cand[judged[b]],
                "judged": len(judged), "median_rel_err_judged": err,
                "surrogate_s": t1 - t0, "judge_s": t2 - t1, "scored": S + len(judged),
                "rollouts_per_s": (S + len(judged)) / (t2 - t0)}

    def random_baseline(self, nominal: np.ndarray, S: int = 16384, sigma: float = 0.2, seed: int = 0) -> Dict:
        """The same DEM budget without the surrogate: k candidates of the same
        sample set (the nominal and k - 1 random ones) judged by the DEM."""
        rng = np.random.default_rng(seed)
        cand = np.clip(nominal[None] + rng.normal(0.0, sigma, (S, self.nk, 2)), -1.0, 1.0).astype(np.float32)
        cand[0] = nominal
        pick = np.concatenate([[0], np.random.default_rng(seed + 7).choice(np.arange(1, S), self.k - 1, False)])
        c = self.dem_costs(torch.as_tensor(cand[pick], device=self.judge.device))
        b = int(np.argmin(c))
        return {"best_cost": float(c[b]), "best_index": int(pick[b])}

    def exhaustive(self, cand: np.ndarray, batch: int = 256) -> np.ndarray:
        """DEM costs of every candidate (S, n_knots, 2), ``batch`` instances
        of the patch per launch: the plain batched-DEM planner's scores."""
        if self._ex is None or self._ex.R != batch:
            self._ex = RolloutBatch([self.judge.scenes[0]], [0] * batch, cell_size_factor=self.cell_size_factor)
            self._ex.set_mode("frames")
            self._ex.ship_init.copy_(self.start.expand(batch, 15))
        S = cand.shape[0]
        out = np.empty(S, np.float64)
        ctl = torch.as_tensor(cand, device=self.judge.device)
        for s0 in range(0, S, batch):
            m = min(batch, S - s0)
            full = torch.empty((batch, self.nk, 2), dtype=torch.float32, device=ctl.device)
            full[:m] = ctl[s0:s0 + m]
            full[m:] = ctl[s0:s0 + 1]
            self._ex.reset()
            out[s0:s0 + m] = rollout_cost(self._ex.run(full, self.hold, self.frames).cpu().numpy()[:m])
        return out

    def close(self):
        self.judge.close()
        if self._ex is not None:
            self._ex.close()
            self._ex = None
