"""Batched DEM rollouts for sampling planners (no reference counterpart).

The reference's MPPI/MBD planners (``jax_mppi_planner.py:34-115``,
``jax_mbd_planner.py:1314-1411``) score candidate controls with a 3-DOF ship
model against a costmap; they never roll out the DEM.  This module scores
candidates with the DEM itself: every candidate is an instance (a scene
context + its own state) and all instances advance ``frames`` frames in one
graph launch (``icedem_batch_rollouts``): either every frame as two
kernels spanning all instances (mode "frames") or one workgroup per instance
running all frames (mode "scenes").  ``autotune`` picks the mode, the grid
cell size, the force pass's lanes per element and the partner-list skin by
timing the batch.

Ship control for rollouts is direct (ship_mode 1): a control sequence of
``n_knots`` knots (u, r_cmd) in [-1, 1], each held ``hold`` frames, gives the
thrust tau = [u X_max, 0, Kd[2,2] (r_cmd r_max - r)].

Rollout cost (this build's definition):
``cost = impulse / J_REF + KE_loss / KE_REF`` with impulse = sum_f |F_ice| dt
(jax_simulation.py:695-696 per frame) and KE_loss = KE(start) - KE(end),
KE = 0.5 m (u^2 + v^2) + 0.5 Izz r^2.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .config import SimulationConfig, benchmark_config
from .engine import DemEngine, _ptr
from .ice_state import (IceState, build_ship_dynamics_data, build_static_data, ship_hull_polygon, soa_pair)
from .properties import ShipProperties
from .scenario import Scene

J_REF = 1.0e5      # N s
KE_REF = 1.0e6     # J


def rollout_cost(out: np.ndarray) -> np.ndarray:
    """(R, 4) [impulse, KE loss, peak |F|, KE0] -> (R,) cost."""
    out = np.asarray(out, dtype=np.float64)
    return out[:, 0] / J_REF + out[:, 1] / KE_REF


@dataclass
class ShipSpec:
    length: float = 60.0
    beam: float = 12.0
    draft: float = 5.0
    hull_points: int = 12
    thrust_max: float = 2e6
    yaw_rate_max: float = 0.15
    speed0: float = 2.0


def ship_start(scene: Scene, ship: ShipSpec, gap: float = 1.0):
    """Bow (heading +y) ``gap`` m short of the lowest ice vertex, centred on it."""
    lowest = min(range(len(scene.polygons)), key=lambda i: scene.polygons[i][:, 1].min())
    p = scene.polygons[lowest]
    return float(p[:, 0].mean()), float(p[:, 1].min()) - gap - ship.length / 2.0


class RolloutBatch:
    """R rollout instances over a list of scenes, advanced in one launch."""

    def __init__(self, scenes: Sequence[Scene], instance_scene: Sequence[int], cfg: Optional[SimulationConfig] = None,
                 ship: Optional[ShipSpec] = None, device=None, cell_size_factor: float = 1.0,
                 small_quantile: float = 0.98, force_lanes: int = 0, list_skin: Optional[float] = None):
        """force_lanes: lanes per element of the batched force pass (8 or 16;
        0 chooses by the batch's total element count); list_skin: reach of
        the contact partner lists beyond contact (None: the engine default,
        0: the grids every frame)."""
        if not torch.cuda.is_available():
            raise _lib.IceDemError("icedem needs a ROCm GPU")
        self.cfg = cfg or benchmark_config()
        self.ship = ship or ShipSpec()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.scenes = list(scenes)
        self.inst = [int(i) for i in instance_scene]
        self.R = len(self.inst)
        sp = ShipProperties(self.ship.length, self.ship.beam, self.ship.draft)
        hull = ship_hull_polygon(self.ship.length, self.ship.beam, self.ship.hull_points // 2 - 1)
        self.ship_dyn = build_ship_dynamics_data(sp, self.cfg, self.ship.length, self.ship.beam, ship_hull_local=hull)
        from .simulation import _thruster_config, _time_stepping
        B, T_n = _thruster_config(sp, self.ship.length)
        statics = {}
        first = {}     # scene -> its first engine: later instances share its static tables
        self.engines: List[DemEngine] = []
        for s_idx in self.inst:
            if s_idx in first:
                self.engines.append(DemEngine.like(first[s_idx]))
                continue
            if s_idx not in statics:
                sc = self.scenes[s_idx]
                statics[s_idx] = build_static_data(sc.polygons, sc.centroids, sc.bonds, sc.thicknesses,
                                                   sc.parent_ids, self.cfg)
            st = statics[s_idx]
            sc = self.scenes[s_idx]
            dt_ice, nsteps = _time_stepping(st, self.cfg, self.cfg.dt)
            if nsteps != 1:
                raise ValueError("batched rollouts need one ice sub-step per frame")
            rmin = 2.0 * float(np.quantile(st.radii, small_quantile)) * 1.01
            eng = DemEngine(st, self.cfg, self.ship_dyn, self.ship.length, self.ship.beam, 1, self.cfg.dt, dt_ice,
                            False, B, T_n, centroids=sc.centroids, device=self.device,
                            cell_size=rmin * cell_size_factor, ship_mode=1,
                            ctrl_thrust_max=self.ship.thrust_max, ctrl_yaw_rate_max=self.ship.yaw_rate_max,
                            grid_margin=25.0, small_quantile=small_quantile, force_lanes=force_lanes,
                            list_skin=list_skin)
            first[s_idx] = eng
            self.engines.append(eng)
        self._alloc_states()
        ctxs = (C.c_void_p * self.R)(*[e.ctx.value for e in self.engines])
        h = C.c_void_p()
        lib = _lib.load()
        _lib.check(lib.icedem_batch_create(ctxs, self.R, C.byref(h)), None, "icedem_batch_create")
        self.handle = h
        self.lib = lib
        self.mode = "frames"
        self.out = torch.zeros((self.R, 4), dtype=torch.float32, device=self.device)

    def _alloc_states(self):
        """All instances' state in flat buffers (one reset = two copies)."""
        dev = self.device
        fl, u8 = 0, 0
        lay = []
        for eng in self.engines:
            n, nb, nf = eng.n, eng.n_bonds, eng.n_floes
            lay.append((fl, u8, n, nb, nf))
            fl += 7 * n + 4 * nb
            u8 += n + nb + nf
        self.f_init = torch.zeros(fl, dtype=torch.float32, device=dev)
        self.u_init = torch.zeros(max(u8, 1), dtype=torch.uint8, device=dev)
        self.ship_init = torch.zeros((self.R, 15), dtype=torch.float32, device=dev)
        for r, (eng, (f0, b0, n, nb, nf)) in enumerate(zip(self.engines, lay)):
            sc = self.scenes[self.inst[r]]
            c = torch.as_tensor(np.asarray(sc.centroids, dtype=np.float32), device=dev)
            self.f_init[f0:f0 + n] = c[:, 0]
            self.f_init[f0 + n:f0 + 2 * n] = c[:, 1]
            self.u_init[b0 + n:b0 + n + nb] = 1          # bonds intact
            x0, y0 = ship_start(sc, self.ship)
            self.ship_init[r, 0], self.ship_init[r, 1], self.ship_init[r, 2] = x0, y0, math.pi / 2
            self.ship_init[r, 3] = self.ship.speed0
        self.f = self.f_init.clone()
        self.u = self.u_init.clone()
        self.ships = self.ship_init.clone()
        self.layout = lay
        self.states = (_lib.IceDemState * self.R)()
        self.views: List[IceState] = []
        for r, (eng, (f0, b0, n, nb, nf)) in enumerate(zip(self.engines, lay)):
            f, u = self.f, self.u
            pos = f[f0:f0 + 2 * n].view(2, n).t()
            vel = f[f0 + 2 * n:f0 + 4 * n].view(2, n).t()
            rot = f[f0 + 4 * n:f0 + 5 * n]
            om = f[f0 + 5 * n:f0 + 6 * n]
            sub = f[f0 + 6 * n:f0 + 7 * n]
            bd = f[f0 + 7 * n:f0 + 7 * n + 4 * nb].view(4, nb)
            act = u[b0:b0 + n].view(torch.bool)
            bi = u[b0 + n:b0 + n + nb].view(torch.bool)
            el = u[b0 + n + nb:b0 + n + nb + nf].view(torch.bool)
            st = IceState(pos, vel, rot, om, sub, bi, bd[0], bd[1], bd[2], bd[3], act, el)
            self.views.append(st)
            self.states[r] = eng.state_struct(st, self.ships[r])

    def reset(self):
        self.f.copy_(self.f_init)
        self.u.copy_(self.u_init)
        self.ships.copy_(self.ship_init)

    def set_mode(self, mode: str):
        m = {"frames": _lib.BATCH_FRAMES, "scenes": _lib.BATCH_SCENES}[mode]
        _lib.check(self.lib.icedem_batch_mode(self.handle, m), None, "icedem_batch_mode")
        self.mode = mode

    def run(self, controls: torch.Tensor, hold: int, frames: int, use_graph: bool = True, sync: bool = True):
        """controls (R, n_knots, 2) -> (R, 4) [impulse, KE loss, peak |F|, KE0]."""
        ctl = torch.as_tensor(controls).to(device=self.device, dtype=torch.float32).contiguous()
        if ctl.shape[0] != self.R or ctl.dim() != 3 or ctl.shape[2] != 2:
            raise ValueError("controls must be (R, n_knots, 2)")
        self._ctl = ctl
        nk = int(ctl.shape[1])
        inputs = (_lib.IceDemInputs * self.R)()
        diags = (_lib.IceDemDiag * self.R)()
        for r in range(self.R):
            inputs[r].d_controls = _ptr(ctl[r])
            inputs[r].n_knots = nk
            inputs[r].hold = int(hold)
            inputs[r].d_setpoints = _ptr(ctl[r])   # not read in ship_mode 1
            diags[r].d_rollout = _ptr(self.out[r])
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(self.lib.icedem_batch_rollouts(self.handle, self.states, inputs, int(frames), diags,
                                                  1 if use_graph else 0, stream), None, "icedem_batch_rollouts")
        if sync:
            torch.cuda.synchronize(self.device)
        return self.out

    def run_sequential(self, controls: torch.Tensor, hold: int, frames: int, use_graph: bool = True):
        """The plain per-rollout path: each instance in turn through
        icedem_frames (K = frames), one launch sequence per rollout."""
        ctl = torch.as_tensor(controls).to(device=self.device, dtype=torch.float32).contiguous()
        sp = torch.zeros((frames, 6), dtype=torch.float32, device=self.device)
        for r, eng in enumerate(self.engines):
            eng.run_frames(self.views[r], self.ships[r], sp, hist=None,
                           diag={"rollout": self.out[r]}, use_graph=use_graph, sync=False,
                           controls=ctl[r], hold=hold)
        torch.cuda.synchronize(self.device)
        return self.out

    def close(self):
        if getattr(self, "handle", None):
            self.lib.icedem_batch_destroy(self.handle)
            self.handle = None
        for e in getattr(self, "engines", []):
            e.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def patch_params(r: int, n_lo: int = 50, n_hi: int = 800, c_lo: float = 0.2, c_hi: float = 0.6):
    """Patch of rollout r: seed 1000 + r, ~U{n_lo..n_hi} elements, concentration U[c_lo, c_hi]."""
    rng = np.random.default_rng(1000 + r)
    return 1000 + r, int(rng.integers(n_lo, n_hi + 1)), float(rng.uniform(c_lo, c_hi))


def _patch(args):
    from .scenario import scene_patch
    return scene_patch(*args)


def rollout_patches(R: int, workers: int = 1) -> List[Scene]:
    """The R rollout patches (deterministic per seed), built in a process pool."""
    params = [patch_params(r) for r in range(R)]
    if workers <= 1:
        return [_patch(a) for a in params]
    import multiprocessing as mp
    with mp.get_context("spawn").Pool(workers) as pool:
        return pool.map(_patch, params, chunksize=max(1, R // (4 * workers)))


def autotune(scenes: Sequence[Scene], cfg=None, ship: Optional[ShipSpec] = None, sample: Optional[int] = None,
             frames: int = 100, cell_factors=(1.0, 1.25), lanes=(0, 16), skins=(None, 3.0),
             modes=("frames", "scenes"), seed: int = 0):
    """Pick the batch's launch configuration by timing `frames` frames of the
    batch (the first `sample` instances; None: all of them, so the timed
    launches have the real batch's workgroup counts) for every combination
    of grid cell size factor, lanes per element of the force pass (0: the
    library's choice for the batch's size) and partner-list skin (None: the
    engine default) in frames mode, and of cell size factor in scenes mode
    (one workgroup per instance: no lists, fixed lanes).  Returns (best,
    times): best = dict(mode, cell_size_factor, force_lanes, list_skin), times
    = {label: seconds}.  Every configuration gives the same results bit for
    bit; only the time differs."""
    import time
    n = len(scenes) if sample is None else min(sample, len(scenes))
    sub = list(range(n))
    ctl = torch.as_tensor(nominal_controls(n, 4, seed=seed))
    times, confs = {}, {}

    def timed(rb):
        rb.reset(); rb.run(ctl, max(1, frames // 4), frames)          # capture + warm
        best = float("inf")
        for _ in range(2):
            rb.reset()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rb.run(ctl, max(1, frames // 4), frames)
            best = min(best, time.perf_counter() - t0)
        return best

    for cf in cell_factors:
        for ln in (lanes if "frames" in modes else (0,)):
            for sk in (skins if "frames" in modes else (None,)):
                rb = RolloutBatch([scenes[i] for i in sub], list(range(n)), cfg, ship, cell_size_factor=cf,
                                  force_lanes=ln, list_skin=sk)
                try:
                    for m in modes:
                        if m == "scenes" and (ln, sk) != (lanes[0], skins[0]):
                            continue   # scenes mode uses neither setting
                        rb.set_mode(m)
                        label = (f"{m}/cell{cf}" if m == "scenes" else
                                 f"{m}/cell{cf}/lanes{ln or 'auto'}/skin{sk if sk is not None else 'default'}")
                        times[label] = timed(rb)
                        confs[label] = dict(mode=m, cell_size_factor=cf, force_lanes=ln, list_skin=sk)
                finally:
                    rb.close()
    best = min(times, key=times.get)
    return confs[best], times


def nominal_controls(R: int, n_knots: int = 40, u: float = 0.7, r: float = 0.0, sigma: float = 0.2,
                     seed: int = 0) -> np.ndarray:
    """MPPI-style candidates: nominal plus N(0, sigma^2) noise, clipped to [-1, 1]."""
    rng = np.random.default_rng(seed)
    nom = np.tile(np.array([u, r], np.float32), (n_knots, 1))
    c = nom[None] + rng.normal(0.0, sigma, (R, n_knots, 2)).astype(np.float32)
    return np.clip(c, -1.0, 1.0).astype(np.float32)
