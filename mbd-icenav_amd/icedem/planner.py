"""Surrogate-plus-DEM rollout scoring for sampling planners (no reference counterpart).

The reference's MPPI/MBD planners score candidates with a 3-DOF ship model
on a costmap (``jax_mppi_planner.py:34-115``); here candidates are scored by
the DEM itself (``rollouts``) and, at planner scale, by a learned surrogate
(``surrogate``) trained on DEM rollouts:

  * ``make_dataset``: DEM rollouts over rollout patches with random start
    states and random control sequences; the surrogate's inputs are the
    ship-frame ice raster, a 6-value ship state and the control knots seen
    through the ice-free ship model (``ship_rollouts``).
  * ``train``: Adam on the weighted MSE of standardised targets, on the GPU.
  * ``KECostmap`` / ``ship_rollouts`` / ``costmap_baseline``: the reference
    planners' scorer, the method without the learned part.
  * ``Planner.iterate``: one sampling iteration -- S candidates scored by the
    surrogate (one CNN pass for the shared raster, the ice-free rollouts and
    the MLP per candidate), the k best re-scored by the batched DEM as judge;
    when the surrogate's error on the judged set exceeds a threshold, the next
    k by surrogate rank are judged as well (fallback), up to a budget; an
    iteration that needed the fallback makes the next one judge the whole
    budget in one launch.

Rollout cost is ``rollouts.rollout_cost``: impulse / 1e5 N s + KE loss / 1e6 J.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .engine import _ptr
from .properties import ShipProperties
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


@dataclass
class CostmapSpec:
    """build_costmap_jax's defaults (jax_costmap.py:177-195), with a grid of
    0.5 m instead of 2 m (the rollouts cover a few metres)."""
    resolution: float = 0.5
    margin: float = 20.0
    base_cost: float = 0.1
    collision_cost: float = 5.0
    obs_cost_buffer_factor: float = 0.1
    ke_cost_scale: float = 1e-6
    planning_clearance_margin: float = 0.0
    oob_penalty: float = 0.0


class KECostmap:
    """The reference planners' KE collision costmap of one patch's initial
    ice (``icesur_ke_costmap``), rows along y; element masses are the static
    masses, the KE speed is the nominal start speed."""

    def __init__(self, eng, scene: Scene, ship: ShipSpec, spec: Optional[CostmapSpec] = None, speed: float = None):
        self.spec = spec = spec or CostmapSpec()
        self.lib = _lib.load()
        dev = eng.device
        c = np.asarray(scene.centroids, np.float64)
        radii = np.asarray(eng._host["radii"], np.float64)
        lo, hi = c.min(0) - radii.max(), c.max(0) + radii.max()
        res = spec.resolution
        self.x0, self.y0, self.res = float(lo[0] - spec.margin), float(lo[1] - spec.margin), float(res)
        self.cols = max(1, int(math.ceil((hi[0] - lo[0] + 2 * spec.margin) / res)))
        self.rows = max(1, int(math.ceil((hi[1] - lo[1] + 2 * spec.margin) / res)))
        pos = torch.as_tensor(np.concatenate([c[:, 0], c[:, 1]]).astype(np.float32), device=dev)
        rad = torch.as_tensor(eng._host["radii"], device=dev)
        mass = torch.as_tensor(eng._host["mass"], device=dev)
        half = int(math.ceil((0.5 * ship.beam + spec.planning_clearance_margin) / res))
        work = torch.empty(self.rows * self.cols, dtype=torch.int32, device=dev)
        self.grid = torch.empty(self.rows * self.cols, dtype=torch.float32, device=dev)
        sp = ShipProperties(ship.length, ship.beam, ship.draft)
        s = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.check(self.lib.icesur_ke_costmap(_ptr(pos), _ptr(rad), _ptr(mass), len(c), self.rows, self.cols, self.x0,
                                              self.y0, res, spec.collision_cost,
                                              float(ship.speed0 if speed is None else speed), float(sp.mass),
                                              spec.ke_cost_scale, spec.obs_cost_buffer_factor, spec.base_cost, half,
                                              _ptr(work), _ptr(self.grid), s), None, "icesur_ke_costmap")
        self._keep = (pos, rad, mass, work)


def ship_rollouts(eng, starts: torch.Tensor, controls: torch.Tensor, hold: int, frames: int,
                  costmap: Optional[KECostmap] = None, n_snap: int = 0):
    """Ship-only rollouts (no ice) of (S, n_knots, 2) controls from one start
    (15,) or one start each (S, 15), scored on the costmap; -> (S, 4)
    [costmap integral, KE loss J, path length, final heading], and with
    n_snap > 0 also (S, n_snap, 4) snapshots every frames / n_snap frames
    [along, across, heading change, speed]."""
    lib = _lib.load()
    ctl = torch.as_tensor(controls).to(device=eng.device, dtype=torch.float32).contiguous()
    st = starts.to(device=eng.device, dtype=torch.float32).contiguous()
    S, nk = int(ctl.shape[0]), int(ctl.shape[1])
    stride = 0 if st.dim() == 1 else 15
    out = torch.empty((S, 4), dtype=torch.float32, device=eng.device)
    snap = torch.empty((S, max(n_snap, 1), 4), dtype=torch.float32, device=eng.device) if n_snap else None
    every = max(1, int(frames) // max(n_snap, 1))
    cm = costmap
    s = C.c_void_p(torch.cuda.current_stream(eng.device).cuda_stream)
    _lib.check(lib.icedem_ship_rollouts(eng.ctx, _ptr(st), stride, _ptr(ctl), S, nk, int(hold), int(frames),
                                        _ptr(cm.grid) if cm else None, cm.rows if cm else 0, cm.cols if cm else 0,
                                        cm.x0 if cm else 0.0, cm.y0 if cm else 0.0, cm.res if cm else 1.0,
                                        cm.spec.oob_penalty if cm else 0.0, _ptr(out), every, int(n_snap),
                                        _ptr(snap) if snap is not None else None, s), eng.ctx, "icedem_ship_rollouts")
    return (out, snap) if n_snap else out


# Network inputs computed from the controls by the ice-free ship model: 3
# summary values, plus 4 per path snapshot.  With the path snapshots the
# network ranked the candidates of one start worse (more inputs to fit noise
# to), so none are used.
N_SNAP = 0
N_PHYS = 3 + 4 * N_SNAP


def phys_features(out: torch.Tensor, snap: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Ship-rollout results (S, 4) and snapshots (S, N_SNAP, 4) -> (S,
    N_PHYS) network inputs: ice-free KE loss / KE_REF, costmap integral / 10,
    path length / 5, then per snapshot along / 5, across, 10 x heading
    change, speed / 3."""
    f = torch.empty((out.shape[0], N_PHYS), dtype=torch.float32, device=out.device)
    f[:, 0] = out[:, 1] / KE_REF
    f[:, 1] = out[:, 0] / 10.0
    f[:, 2] = out[:, 2] / 5.0
    if N_SNAP:
        sc = torch.tensor([0.2, 1.0, 10.0, 1.0 / 3.0], dtype=torch.float32, device=out.device)
        f[:, 3:] = (snap * sc).reshape(out.shape[0], -1)
    return f


def _phys(eng, starts, ctl, hold, frames, cm):
    if N_SNAP:
        return ship_rollouts(eng, starts, ctl, hold, frames, cm, N_SNAP)
    return ship_rollouts(eng, starts, ctl, hold, frames, cm), None


def costmap_features(out: torch.Tensor) -> torch.Tensor:
    """(S, 4) ship-rollout results -> (S, 3) [costmap integral, ice-free KE
    loss / KE_REF, 1] for the calibrated costmap cost."""
    f = torch.ones((out.shape[0], 3), dtype=torch.float64, device=out.device)
    f[:, 0] = out[:, 0].double()
    f[:, 1] = out[:, 1].double() / KE_REF
    return f


def fit_costmap_cost(feat: np.ndarray, cost: np.ndarray) -> np.ndarray:
    """Least-squares weights of the costmap cost a integral + b KE_free + c on
    DEM costs (the reference weights its costmap term by a calibrated
    normaliser, jax_mbd_planner.py:696-699)."""
    w, *_ = np.linalg.lstsq(np.asarray(feat, np.float64), np.asarray(cost, np.float64), rcond=None)
    return w


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


def jitter_offsets(jitters: int, radius: float, phase: float = 0.0) -> np.ndarray:
    """(J, 2) start offsets (x, y) in metres: J points on a circle of
    `radius` from angle `phase`, or one zero offset for J <= 1.  A rollout's
    jitter-mean cost is the mean of its costs from these starts."""
    if jitters <= 1:
        return np.zeros((1, 2))
    a = phase + 2.0 * math.pi * np.arange(jitters) / jitters
    return radius * np.stack([np.cos(a), np.sin(a)], 1)


def run_jittered(rb: RolloutBatch, starts: torch.Tensor, ctl, hold: int, frames: int,
                 offsets: np.ndarray) -> np.ndarray:
    """Mean rollout outputs (R, 4) of `ctl` over the start offsets."""
    acc = None
    for off in offsets:
        s = starts.clone()
        if off[0] != 0.0:
            s[:, 0] += float(off[0])
        if off[1] != 0.0:
            s[:, 1] += float(off[1])
        rb.ship_init.copy_(s)
        rb.reset()
        o = rb.run(ctl, hold, frames).cpu().numpy().astype(np.float64)
        acc = o if acc is None else acc + o
    return acc / len(offsets)


def make_dataset(rb: RolloutBatch, rounds: int, frames: int = 400, hold: int = 10, n_knots: int = 40,
                 seed: int = 0, spread: Optional[StartSpread] = None, raw_controls: bool = False,
                 progress=None, jitters: int = 1, jitter_m: float = 0.05) -> Dict[str, torch.Tensor]:
    """`rounds` x R DEM rollouts (one per patch instance per round).  The
    extra inputs are the ship state (6) and the controls seen through the
    ice-free ship model (N_PHYS); ``raw_controls`` appends the knots
    themselves (2 n_knots).  With ``jitters`` > 1 each target is the mean
    over that many starts on a circle of ``jitter_m`` metres around the
    sampled one (jitter_offsets): the jitter-mean cost."""
    spread = spread or StartSpread()
    rng = np.random.default_rng(seed)
    dev = rb.device
    pr = PatchRasters(rb)
    nominal = rb.ship_init.clone()
    start_xy = nominal[:, :2].clone()
    R = rb.R
    N = rounds * R
    ras = torch.empty((N, RASTER_G, RASTER_G, 2), dtype=torch.bfloat16, device=dev)
    extra = torch.empty((N, 6 + N_PHYS), dtype=torch.float32, device=dev)
    tgt = np.empty((N, 3), np.float32)
    starts_all = torch.empty((N, 15), dtype=torch.float32, device=dev)
    ctl_all = torch.empty((N, n_knots, 2), dtype=torch.float32, device=dev)
    for k in range(rounds):
        starts = random_starts(rb, nominal, rng, spread)
        ctl = random_controls(R, n_knots, rng)
        out = run_jittered(rb, starts, torch.as_tensor(ctl), hold, frames, jitter_offsets(jitters, jitter_m))
        sl = slice(k * R, (k + 1) * R)
        for r in range(R):
            ras[k * R + r] = pr.of(r, starts[r:r + 1, :3])[0]
        extra[sl, :6] = ship_features(starts, start_xy)
        tgt[sl] = targets_of(out)
        starts_all[sl] = starts
        ctl_all[sl] = torch.as_tensor(ctl, device=dev)
        if progress and (k + 1) % max(1, rounds // 10) == 0:
            progress(k + 1, rounds)
    rb.ship_init.copy_(nominal)
    # the costmap planners' view of the same rollouts: ship-only paths on each patch's KE costmap
    cmf = torch.empty((N, 3), dtype=torch.float64, device=dev)
    for r in range(R):
        eng = rb.engines[r]
        cm = KECostmap(eng, rb.scenes[rb.inst[r]], rb.ship)
        ii = torch.arange(r, N, R, device=dev)
        so, snap = _phys(eng, starts_all[ii], ctl_all[ii], hold, frames, cm)
        cmf[ii] = costmap_features(so)
        extra[ii, 6:] = phys_features(so, snap)
    if raw_controls:
        extra = torch.cat([extra, ctl_all.reshape(N, -1)], dim=1)
    return {"raster": ras, "extra": extra, "target": torch.as_tensor(tgt, device=dev), "costmap_feat": cmf}


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
    centre line: sway, yaw rate, heading offset, lateral offset, the
    ice-free path's across offsets and heading changes, and the yaw commands
    change sign; the rest do not."""
    s = torch.ones(6 + N_PHYS + 2 * n_knots, dtype=torch.float32, device=device)
    s[1:5] = -1.0
    for k in range(N_SNAP):
        s[6 + 3 + 4 * k + 1] = -1.0     # across
        s[6 + 3 + 4 * k + 2] = -1.0     # heading change
    s[6 + N_PHYS + 1::2] = -1.0
    return s


def split_indices(N: int, holdout: float = 0.05, seed: int = 0):
    """(validation, training) sample indices of a dataset of N."""
    perm = np.random.default_rng(seed).permutation(N)
    nval = max(1, int(N * holdout))
    return perm[:nval], perm[nval:]


def costmap_baseline(data: Dict[str, torch.Tensor], holdout: float = 0.05, seed: int = 0):
    """The method without the learned part: the costmap cost calibrated on
    the training split; returns (weights, relative cost error on the
    validation split)."""
    val, tr = split_indices(data["raster"].shape[0], holdout, seed)
    f = data["costmap_feat"].cpu().numpy()
    c = cost_of_targets(data["target"].cpu().numpy())
    w = fit_costmap_cost(f[tr], c[tr])
    pred = f[val] @ w
    rel = np.abs(pred - c[val]) / np.maximum(np.abs(c[val]), COST_FLOOR)
    return w, {"median": float(np.median(rel)), "mean": float(np.mean(rel)), "p90": float(np.percentile(rel, 90))}


def train(net: Surrogate, data: Dict[str, torch.Tensor], epochs: int = 10, batch: int = 512, lr: float = 1e-3,
          holdout: float = 0.05, seed: int = 0, weight_decay: float = 0.0, mirror: bool = False, progress=None):
    """Adam on standardised targets; returns (norm, history) with the
    held-out relative cost error after every epoch.  ``mirror`` replaces a
    random half of every batch by its mirror image (the hull, the contact
    model and the drag are symmetric about the ship's centre line, so the
    mirrored input has the same outcome distribution)."""
    N = data["raster"].shape[0]
    val, tr = split_indices(N, holdout, seed)
    rng = np.random.default_rng(seed + 1)
    t_np = data["target"].cpu().numpy()
    norm = Norm(t_np[tr].mean(0), t_np[tr].std(0) + 1e-6)
    tgt = norm.fwd(data["target"])
    dev = data["raster"].device
    signs = mirror_signs((data["extra"].shape[1] - 6 - N_PHYS) // 2, dev)
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


def mppi_update(costs: torch.Tensor, samples: torch.Tensor, temperature: float) -> torch.Tensor:
    """run_mppi's update (jax_mppi_planner.py:95-101) on the GPU
    (`icesur_mppi_update`): costs (S,), samples (S, ...) -> the clipped
    exp(-(c - min c) / temperature) weighted mean, shaped like one sample."""
    lib = _lib.load()
    c = costs.to(torch.float32).contiguous()
    u = samples.to(torch.float32).contiguous()
    S = int(c.shape[0])
    D = int(u[0].numel())
    work = torch.empty(2 + ((S + 255) // 256) * D, dtype=torch.float32, device=c.device)
    out = torch.empty(D, dtype=torch.float32, device=c.device)
    s = C.c_void_p(torch.cuda.current_stream(c.device).cuda_stream)
    _lib.check(lib.icesur_mppi_update(_ptr(c), _ptr(u), S, D, float(temperature), _ptr(work), _ptr(out), s), None,
               "icesur_mppi_update")
    return out.view(u.shape[1:])


def mbd_schedule(n_diffuse: int, beta0: float, betaT: float):
    """run_mbd's float32 noise schedule (jax_mbd_planner.py:1358-1361):
    betas = linspace(beta0, betaT), alphas = 1 - betas, alphas_bar their
    running product, sigmas = sqrt(1 - alphas_bar)."""
    betas = np.linspace(beta0, betaT, n_diffuse).astype(np.float32)
    alphas = (np.float32(1.0) - betas).astype(np.float32)
    alphas_bar = np.cumprod(alphas, dtype=np.float32)
    sigmas = np.sqrt(np.float32(1.0) - alphas_bar).astype(np.float32)
    return betas, alphas, alphas_bar, sigmas


def mbd_update(costs: torch.Tensor, samples: torch.Tensor, temperature: float, ybar: torch.Tensor,
               alpha_bar: float, alpha: float, alpha_bar_prev: float) -> torch.Tensor:
    """One reverse-diffusion step of run_mbd (jax_mbd_planner.py:1363-1396) on
    the GPU (`icesur_mbd_update`): costs (S,), samples (S, ...) the step's
    clipped samples, ybar the current mean -> the next mean, shaped like one
    sample."""
    lib = _lib.load()
    c = costs.to(torch.float32).contiguous()
    u = samples.to(torch.float32).contiguous()
    S = int(c.shape[0])
    D = int(u[0].numel())
    yb = ybar.to(device=c.device, dtype=torch.float32).contiguous().reshape(-1)
    if yb.numel() != D:
        raise ValueError("ybar must have the shape of one sample")
    work = torch.empty(2 + ((S + 255) // 256) * D, dtype=torch.float32, device=c.device)
    out = torch.empty(D, dtype=torch.float32, device=c.device)
    s = C.c_void_p(torch.cuda.current_stream(c.device).cuda_stream)
    _lib.check(lib.icesur_mbd_update(_ptr(c), _ptr(u), S, D, float(temperature), _ptr(yb), float(alpha_bar),
                                     float(alpha), float(alpha_bar_prev), _ptr(work), _ptr(out), s), None,
               "icesur_mbd_update")
    return out.view(u.shape[1:])


class Planner:
    """One patch, one start state: S candidates scored by the surrogate, the
    best k judged by the batched DEM (k instances of the patch)."""

    def __init__(self, net: Surrogate, norm: Norm, scene: Scene, start: torch.Tensor, k: int = 64,
                 frames: int = 400, hold: int = 10, n_knots: int = 40, err_threshold: float = 0.05,
                 max_judged: int = 256, cell_size_factor: float = 1.0, costmap_weights=None):
        self.net, self.norm = net, norm
        self.k, self.frames, self.hold, self.nk = k, frames, hold, n_knots
        self.thr, self.max_judged = err_threshold, max_judged
        self.cell_size_factor = cell_size_factor
        self._ex = None
        self.width = k
        self.judge = RolloutBatch([scene], [0] * k, cell_size_factor=cell_size_factor)
        self.judge.set_mode("frames")
        self.nominal_start = self.judge.ship_init[0].clone()
        self.start = start.to(self.judge.device).float()
        self.judge.ship_init.copy_(self.start.expand(k, 15))
        pr = PatchRasters(self.judge)
        self.ras = pr.of(0, self.start[None, :3]).contiguous()
        self.feat = ship_features(self.start[None], self.nominal_start[None, :2])
        self.cm_w = None if costmap_weights is None else np.asarray(costmap_weights, np.float64)
        self.costmap = KECostmap(self.judge.engines[0], scene, self.judge.ship)

    def surrogate_costs(self, ctl: torch.Tensor) -> torch.Tensor:
        """(S, n_knots, 2) -> (S,) predicted costs (one CNN pass for the
        shared raster, the ice-free ship rollouts, the MLP per candidate)."""
        S = ctl.shape[0]
        raw = self.net.arch.extra > 6 + N_PHYS
        extra = torch.empty((S, 6 + N_PHYS + (2 * self.nk if raw else 0)), dtype=torch.float32, device=ctl.device)
        extra[:, :6] = self.feat
        so, snap = _phys(self.judge.engines[0], self.start, ctl, self.hold, self.frames, self.costmap)
        extra[:, 6:6 + N_PHYS] = phys_features(so, snap)
        if raw:
            extra[:, 6 + N_PHYS:] = ctl.reshape(S, -1)
        y = self.norm.inv(self.net.predict_shared(self.ras, extra))
        return torch.expm1(y[:, 0]) + y[:, 1]

    def costmap_costs(self, ctl: torch.Tensor) -> torch.Tensor:
        """(S, n_knots, 2) -> (S,) calibrated costmap costs of the ship-only
        rollouts (the reference planners' scoring, without the network)."""
        out = ship_rollouts(self.judge.engines[0], self.start, ctl, self.hold, self.frames, self.costmap)
        w = torch.as_tensor(self.cm_w, device=out.device)
        return costmap_features(out) @ w

    def _batch(self, width: int) -> RolloutBatch:
        """A judge batch of `width` instances of the patch (k, or the
        exhaustive batch's 256)."""
        if width == self.k:
            return self.judge
        if self._ex is None or self._ex.R != width:
            if self._ex is not None:
                self._ex.close()
            self._ex = RolloutBatch([self.judge.scenes[0]], [0] * width, cell_size_factor=self.cell_size_factor)
            self._ex.set_mode("frames")
            self._ex.ship_init.copy_(self.start.expand(width, 15))
        return self._ex

    def dem_costs(self, ctl: torch.Tensor, width: Optional[int] = None) -> np.ndarray:
        """DEM costs of up to `width` (default k) candidates in one launch
        (padded with the first)."""
        m = ctl.shape[0]
        width = width or self.k
        rb = self._batch(width)
        full = torch.empty((width, self.nk, 2), dtype=torch.float32, device=ctl.device)
        full[:m] = ctl
        full[m:] = ctl[:1]
        rb.reset()
        out = rb.run(full, self.hold, self.frames).cpu().numpy()
        return rollout_cost(out[:m])

    def iterate(self, nominal: np.ndarray, S: int = 16384, sigma: float = 0.2, seed: int = 0,
                keep: bool = False, ranker: str = "surrogate") -> Dict:
        """One sampling iteration around `nominal` (n_knots, 2).  With `keep`
        the result also holds the candidates (S, n_knots, 2) and their costs
        (S,): the ranker's, replaced by the DEM's where judged.  ``ranker``
        "costmap" ranks the candidates by the calibrated costmap cost instead
        of the surrogate, "ensemble" by the mean of the two estimates (both
        are fitted to the same rollout cost), with the same DEM judge and
        fallback."""
        if ranker not in ("surrogate", "costmap", "ensemble"):
            raise ValueError(f"ranker must be 'surrogate', 'costmap' or 'ensemble', not {ranker!r}")
        rng = np.random.default_rng(seed)
        dev = self.judge.device
        cand = np.clip(nominal[None] + rng.normal(0.0, sigma, (S, self.nk, 2)), -1.0, 1.0).astype(np.float32)
        cand[0] = nominal
        ctl = torch.as_tensor(cand, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if ranker == "surrogate":
            pred = self.surrogate_costs(ctl)
        elif ranker == "costmap":
            pred = self.costmap_costs(ctl).float()
        else:
            pred = 0.5 * (self.surrogate_costs(ctl).float() + self.costmap_costs(ctl).float())
        order = torch.argsort(pred).cpu().numpy()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        judged, dem, prd = [], [], []
        err = float("inf")
        pos = 0
        while pos < len(order) and len(judged) < self.max_judged:
            # the first judge launch is as wide as the last iteration needed
            w = self.width if pos == 0 else self.k
            sel = order[pos:pos + w]
            pos += w
            c = self.dem_costs(ctl[torch.as_tensor(sel, device=dev)], w)
            p = pred[torch.as_tensor(sel, device=dev)].cpu().numpy()
            judged.extend(sel.tolist()); dem.extend(c.tolist()); prd.extend(p.tolist())
            d, q = np.asarray(dem), np.asarray(prd)
            err = float(np.median(np.abs(q - d) / np.maximum(np.abs(d), COST_FLOOR)))
            if err <= self.thr:
                break
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        # judge this many at once next time (the fallback rounds this one
        # needed); the wide batch is built and its graph captured now
        width = 256 if len(judged) > self.k and self.max_judged >= 256 else self.k
        if width != self.width:
            self.width = width
            self.dem_costs(ctl[:1], width)
        dem = np.asarray(dem)
        b = int(np.argmin(dem))
        out = {"best_index": int(judged[b]), "best_cost": float(dem[b]), "best_controls": cand[judged[b]],
               "judged": len(judged), "judged_indices": list(judged), "judged_costs": dem.copy(),
               "median_rel_err_judged": err,
               "surrogate_s": t1 - t0, "judge_s": t2 - t1, "scored": S + len(judged),
               "rollouts_per_s": (S + len(judged)) / (t2 - t0)}
        if keep:
            costs = pred.float().clone()
            costs[torch.as_tensor(judged, device=dev)] = torch.as_tensor(dem, dtype=torch.float32, device=dev)
            out["costs"], out["candidates"] = costs, ctl
        return out

    def mppi(self, nominal: np.ndarray, iterations: int = 3, temperature: float = 0.1, S: int = 16384,
             sigma: float = 0.2, seed: int = 0) -> Dict:
        """The reference's MPPI loop (run_mppi, jax_mppi_planner.py:34-115)
        with this planner's scores: each iteration samples S candidates
        around the nominal, scores them (surrogate, DEM on the judged ones)
        and moves the nominal to their exp(-(c - min c) / temperature)
        weighted mean (`mppi_update`).  Returns the final nominal, its DEM
        cost, and per-iteration (mean, min, best DEM) costs."""
        v = np.asarray(nominal, np.float32)
        hist = []
        for it in range(iterations):
            r = self.iterate(v, S=S, sigma=sigma, seed=seed + it, keep=True)
            c = r["costs"]
            hist.append({"mean": float(c.mean()), "min": float(c.min()), "best_dem": r["best_cost"],
                         "judged": r["judged"]})
            v = mppi_update(c, r["candidates"], temperature).cpu().numpy()
        final = float(self.dem_costs(torch.as_tensor(v[None], device=self.judge.device))[0])
        return {"nominal": v, "dem_cost": final, "history": hist}

    def mbd(self, nominal: Optional[np.ndarray] = None, n_diffuse: int = 8, temperature: float = 0.1,
            beta0: float = 1e-5, betaT: float = 0.05, S: int = 16384, seed: int = 0,
            init_surge_bias: float = 0.7) -> Dict:
        """The reference's MBD loop (run_mbd, jax_mbd_planner.py:1314-1411)
        over this planner's knots (n_knots, 2) with this planner's scores:
        for i = n_diffuse - 1 .. 1 the step samples S candidates
        clip(Ybar + sigma_i eps) (candidate 0 is Ybar itself), scores them
        (surrogate, DEM on the judged ones) and moves Ybar by the diffusion
        score step (`mbd_update`).  Starts from `nominal` (clipped) or the
        constant surge bias.  Returns Ybar, the last step's mean and min
        cost (the reference's three results), the final Ybar's DEM cost and
        per-step history."""
        if nominal is None:
            ybar = np.zeros((self.nk, 2), np.float32)
            ybar[:, 0] = init_surge_bias
        else:
            ybar = np.clip(np.asarray(nominal, np.float32), -1.0, 1.0)
        if n_diffuse <= 1:
            return {"ybar": ybar, "cost_mean": 0.0, "cost_min": 0.0, "history": []}
        _, alphas, alphas_bar, sigmas = mbd_schedule(n_diffuse, beta0, betaT)
        hist = []
        cmean = cmin = 0.0
        for step, i in enumerate(range(n_diffuse - 1, 0, -1)):
            r = self.iterate(ybar, S=S, sigma=float(sigmas[i]), seed=seed + step, keep=True)
            c = r["costs"]
            cmean, cmin = float(c.mean()), float(c.min())
            hist.append({"i": i, "sigma": float(sigmas[i]), "mean": cmean, "min": cmin, "best_dem": r["best_cost"],
                         "judged": r["judged"]})
            yb = torch.as_tensor(ybar, device=self.judge.device)
            ybar = mbd_update(c, r["candidates"], temperature, yb, float(alphas_bar[i]), float(alphas[i]),
                              float(alphas_bar[max(i - 1, 0)])).cpu().numpy()
        final = float(self.dem_costs(torch.as_tensor(ybar[None], device=self.judge.device))[0])
        return {"ybar": ybar, "cost_mean": cmean, "cost_min": cmin, "dem_cost": final, "history": hist}

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

    def exhaustive(self, cand: np.ndarray, batch: int = 256, offsets: Optional[np.ndarray] = None) -> np.ndarray:
        """DEM costs of every candidate (S, n_knots, 2), ``batch`` instances
        of the patch per launch: the plain batched-DEM planner's scores.
        With ``offsets`` (jitter_offsets) the jitter-mean costs instead."""
        ex = self._batch(batch)
        S = cand.shape[0]
        out = np.empty(S, np.float64)
        ctl = torch.as_tensor(cand, device=self.judge.device)
        starts = self.start.expand(batch, 15)
        offs = np.zeros((1, 2)) if offsets is None else np.asarray(offsets, np.float64)
        for s0 in range(0, S, batch):
            m = min(batch, S - s0)
            full = torch.empty((batch, self.nk, 2), dtype=torch.float32, device=ctl.device)
            full[:m] = ctl[s0:s0 + m]
            full[m:] = ctl[s0:s0 + 1]
            out[s0:s0 + m] = rollout_cost(run_jittered(ex, starts, full, self.hold, self.frames, offs)[:m])
        ex.ship_init.copy_(starts)
        return out

    def close(self):
        self.judge.close()
        if self._ex is not None:
            self._ex.close()
            self._ex = None
