"""Shared test fixtures: scenes, the oracle, the device engine, comparisons.

Tests (and ``__graft_entry__.smoke``) only.  A ``Case`` holds one scene with
both the device engine and the float32 oracle built from the same inputs.
"""
from __future__ import annotations

import math
import sys
from dataclasses import dataclass, field, fields
from pathlib import Path
from typing import Dict, List, Optional

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
for p in (str(ROOT / "mbd-icenav_amd"), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)

from icedem import scenario  # noqa: E402
from icedem.config import SimulationConfig, benchmark_config  # noqa: E402
from icedem.properties import ShipProperties  # noqa: E402
from oracle import model as omodel  # noqa: E402
from oracle import step as ostep  # noqa: E402

F32 = np.float32


def oracle_config(cfg: SimulationConfig) -> omodel.OracleConfig:
    names = {f.name for f in fields(omodel.OracleConfig)}
    return omodel.OracleConfig(**{k: getattr(cfg, k) for k in names if hasattr(cfg, k)})


@dataclass
class Case:
    name: str
    scene: scenario.Scene
    cfg: SimulationConfig
    ship_length: float
    ship_beam: float
    ship_mass: float
    hull: np.ndarray
    traj: np.ndarray
    dt: float
    use_auto: bool
    warm_frames: int = 0
    v_water: Optional[np.ndarray] = None
    _engine: object = None
    _oracle: object = None
    _oracle64: object = None
    pair_mode: str = "all"
    force_lanes: int = 0          # lanes per element of the GPU force pass (0: the library's choice)
    list_skin: Optional[float] = None   # GPU contact partner-list skin (None: the library's default, 0: off)
    cell_factor: float = 1.0      # GPU fine grid cell / 2 (r_s + delta)
    bucket_caps: Optional[tuple] = None   # GPU (fine, coarse) bucket capacities in place of the sized ones
    near_cand_margin: float = -1.0   # GPU near-candidate margin (< 0: the library's)

    @staticmethod
    def make(name: str, n_frames: int = 200, seed: int = 0, cfg: Optional[SimulationConfig] = None,
             use_auto: bool = True, hull_points: int = 12, current: Optional[tuple] = None,
             force_lanes: int = 0, cell_factor: float = 1.0) -> "Case":
        if name == "hexagon":
            sc = scenario.hexagon_scene()
            cfg = cfg or SimulationConfig(n_edge_samples=3)
            L, B = 15.0, 5.0
            mass = ShipProperties(L, B, 2.5).mass
            traj = np.array([[-25.0 + i * 0.5, 0.0, 0.0] for i in range(n_frames + 2)])
            hull = omodel.hull_polygon(L, B)
            dt = 0.05
            warm = 22
        elif name == "c0":
            sc = scenario.scene_c0(seed)
            cfg = cfg or benchmark_config()
            L, B = 60.0, 12.0
            mass = ShipProperties(L, B, 5.0).mass
            dt = cfg.dt
            x0, y0 = scenario.ship_start_below(sc, L)
            traj = scenario.straight_trajectory(x0, y0, math.pi / 2, 2.0, dt, n_frames)
            hull = omodel.hull_polygon(L, B, n_pts=hull_points // 2 - 1)
            warm = 780
        elif name == "c1small":
            sc = scenario.scene_field(seed, 120.0, 160.0, 0.30, 40)
            cfg = cfg or benchmark_config()
            L, B = 60.0, 12.0
            mass = ShipProperties(L, B, 5.0).mass
            dt = cfg.dt
            x0, y0 = scenario.ship_start_below(sc, L, gap=-3.0)
            traj = scenario.straight_trajectory(x0, y0, math.pi / 2, 2.0, dt, n_frames)
            hull = omodel.hull_polygon(L, B, n_pts=hull_points // 2 - 1)
            warm = 0
        elif name in ("c1", "c2"):
            # the benchmark fields at full size (bench.py build_case)
            sc = scenario.scene_c1(seed) if name == "c1" else scenario.scene_c2(seed)
            cfg = cfg or benchmark_config()
            L, B = 60.0, 12.0
            mass = ShipProperties(L, B, 5.0).mass
            dt = cfg.dt
            x0, y0 = scenario.ship_start_below(sc, L, gap=1.0)
            traj = scenario.straight_trajectory(x0, y0, math.pi / 2, 2.0 if name == "c1" else 3.0, dt, n_frames)
            hull = omodel.hull_polygon(L, B, n_pts=hull_points // 2 - 1)
            warm = 0
        elif name == "ring":
            # one big square ringed by 60 small squares: >48 contacts on one element
            polys = [np.array([[-10.0, -10.0], [10.0, -10.0], [10.0, 10.0], [-10.0, 10.0]])]
            for k in range(60):
                a = 2 * np.pi * k / 60
                c = 11.6 * np.array([np.cos(a), np.sin(a)])
                polys.append(c + np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]]))
            cents = np.array([p.mean(axis=0) for p in polys])
            n = len(polys)
            sc = scenario.Scene(polys, cents, [(0, 1), (1, 2), (2, 3)], np.ones(n), np.arange(n, dtype=np.int32),
                                n, 0.0, (-30.0, 30.0, -30.0, 30.0))
            cfg = cfg or benchmark_config()
            L, B = 60.0, 12.0
            mass = ShipProperties(L, B, 5.0).mass
            dt = cfg.dt
            traj = scenario.straight_trajectory(0.0, -60.0, math.pi / 2, 2.0, dt, n_frames)
            hull = omodel.hull_polygon(L, B, n_pts=hull_points // 2 - 1)
            warm = 0
        else:
            raise ValueError(name)
        case = Case(name, sc, cfg, L, B, mass, hull, traj, dt, use_auto, warm)
        case.force_lanes = force_lanes
        case.cell_factor = cell_factor
        if case.n > 20000:
            case.pair_mode = "cells"      # the oracle's cell-list pairs (same pairs and order as all pairs)
        if current is not None:
            speed, direction = current
            v = np.array([speed * math.cos(direction), speed * math.sin(direction)], np.float32)
            case.v_water = np.tile(v, (len(sc.polygons), 1))
        return case

    # ── builders ────────────────────────────────────────────────────────────
    @property
    def n(self):
        return len(self.scene.polygons)

    def oracle(self, dtype=np.float32):
        attr = "_oracle" if dtype == np.float32 else "_oracle64"
        if getattr(self, attr) is None:
            ocfg = oracle_config(self.cfg)
            sc = self.scene
            st = omodel.build_static_data(sc.polygons, sc.centroids, sc.bonds, sc.thicknesses, sc.parent_ids, ocfg)
            sd = omodel.build_ship_dynamics_data(self.ship_mass, self.ship_length, self.ship_beam, self.hull)
            fos = omodel.Fossen(self.ship_mass, self.ship_length)
            dt_ice, nsteps = omodel.time_stepping(st, ocfg, self.dt)
            o = ostep.Oracle(st, ocfg, sd, self.ship_length, self.ship_beam, nsteps, self.dt, dt_ice,
                             self.use_auto, fos.B, fos.T_n, dtype=dtype, pair_mode=self.pair_mode)
            setattr(self, attr, o)
        return getattr(self, attr)

    def engine(self):
        if self._engine is None:
            from icedem.engine import DemEngine
            from icedem.ice_state import build_static_data, build_ship_dynamics_data
            from icedem.simulation import _time_stepping, _thruster_config
            sc = self.scene
            static = build_static_data(sc.polygons, sc.centroids, sc.bonds, sc.thicknesses, sc.parent_ids, self.cfg)
            sp = ShipProperties(self.ship_length, self.ship_beam, 1.0, displacement=self.ship_mass)
            sd = build_ship_dynamics_data(sp, self.cfg, self.ship_length, self.ship_beam, ship_hull_local=self.hull)
            dt_ice, nsteps = _time_stepping(static, self.cfg, self.dt)
            B, T_n = _thruster_config(sp, self.ship_length)
            grid = None
            if self.bucket_caps is not None:
                from icedem.engine import choose_grid
                grid = list(choose_grid(np.asarray(sc.centroids, dtype=np.float64), static.radii,
                                        float(getattr(self.cfg, "contact_dilation", 0.0)),
                                        cell_factor=self.cell_factor))
                grid[6], grid[7] = self.bucket_caps
            self._engine = DemEngine(static, self.cfg, sd, self.ship_length, self.ship_beam, nsteps, self.dt, dt_ice,
                                     self.use_auto, B, T_n, centroids=sc.centroids, force_lanes=self.force_lanes,
                                     list_skin=self.list_skin, near_cand_margin=self.near_cand_margin,
                                     cell_factor=self.cell_factor, grid=grid)
        return self._engine

    # ── states ──────────────────────────────────────────────────────────────
    def oracle_initial(self, dtype=np.float32):
        sc = self.scene
        st = ostep.initial_ice_state(sc.centroids, len(sc.bonds), sc.parent_ids, dtype)
        ship = ostep.initial_ship_state(self.traj[0, :2], self.traj[0, 2], dtype)
        return st, ship

    def gpu_initial(self):
        from icedem.ice_state import build_initial_ice_state
        eng = self.engine()
        sc = self.scene
        st = build_initial_ice_state(self.n, sc.centroids, len(sc.bonds), sc.parent_ids, eng.device)
        import torch
        ship = torch.zeros(15, dtype=torch.float32, device=eng.device)
        ship[0] = float(self.traj[0, 0]); ship[1] = float(self.traj[0, 1]); ship[2] = float(self.traj[0, 2])
        return st, ship

    def setpoint_row(self, k):
        i = min(k, len(self.traj) - 1)
        j = min(k + 1, len(self.traj) - 1)
        return np.concatenate([self.traj[i], self.traj[j]]).astype(np.float32)

    def gpu_frame(self, st, ship, k, diag=None):
        import torch
        eng = self.engine()
        sp = torch.as_tensor(self.setpoint_row(k)[None, :], device=eng.device)
        vw = None if self.v_water is None else torch.as_tensor(self.v_water, device=eng.device)
        eng.run_frames(st, ship, sp, vw, None, diag=diag, use_graph=False)
        return st, ship

    def to_oracle(self, st, ship):
        """Device state -> oracle (AoS numpy float32) state."""
        c = lambda t: t.detach().cpu().numpy()
        ost = ostep.IceState(c(st.positions).astype(F32).copy(), c(st.velocities).astype(F32).copy(),
                             c(st.rotations).copy(), c(st.angular_vels).copy(), c(st.submergence).copy(),
                             c(st.bond_intact).astype(bool).copy(), c(st.bond_damage).copy(),
                             c(st.bond_delta_m_0).copy(), c(st.bond_delta_m_max).copy(), c(st.bond_delta_s).copy(),
                             c(st.elem_active).astype(bool).copy(), c(st.dpdem_floe_eligible).astype(bool).copy())
        s = c(ship)
        oship = ostep.ShipState(s[0:3].copy(), s[3:6].copy(), s[6:12].copy(), s[12:15].copy())
        return ost, oship


@dataclass
class Report:
    exact: Dict[str, bool] = field(default_factory=dict)
    rel: Dict[str, float] = field(default_factory=dict)
    bits: Dict[str, float] = field(default_factory=dict)
    tol: float = 1e-4

    def add_exact(self, name, a, b):
        a = np.asarray(a); b = np.asarray(b)
        self.exact[name] = bool(a.shape == b.shape and np.array_equal(a, b))

    def add_float(self, name, a, b, floor_frac=1e-3, abs_floor=0.0):
        a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
        if a.size == 0:
            self.rel[name] = 0.0; self.bits[name] = 1.0
            return
        scale = max(float(np.max(np.abs(b))) * floor_frac, abs_floor, 1e-30)
        den = np.maximum(np.abs(b), scale)
        both_nan = np.isnan(a) & np.isnan(b)
        err = np.where(both_nan, 0.0, np.abs(a - b) / den)
        self.rel[name] = float(np.max(err))
        self.bits[name] = float(np.mean((a == b) | both_nan))

    def ok(self):
        return all(self.exact.values()) and all(v <= self.tol for v in self.rel.values())

    def summary(self):
        bad = [k for k, v in self.exact.items() if not v] + [k for k, v in self.rel.items() if v > self.tol]
        worst = max(self.rel.values()) if self.rel else 0.0
        exact_frac = min(self.bits.values()) if self.bits else 1.0
        return (f"flags_exact={all(self.exact.values())} max_rel={worst:.3e} "
                f"min_bit_identical_frac={exact_frac:.4f} bad={bad}")

    def assert_ok(self):
        assert self.ok(), self.summary() + f"\nrel={self.rel}\nexact={self.exact}"


def compare_frame(case: Case, st, ship, k: int, report: Optional[Report] = None) -> Report:
    """Run frame k from the same input state on the GPU and the float32 oracle."""
    import torch
    eng = case.engine()
    o = case.oracle()
    ost, oship = case.to_oracle(st, ship)
    vw = case.v_water if case.v_water is not None else np.zeros((case.n, 2), F32)
    sp = case.setpoint_row(k)
    ost2, oship2, od = o.frame_step(ost, oship, sp[:3], sp[3:], vw, ost.dpdem_floe_eligible)
    gst = eng.clone_state(st)
    gship = ship.clone()
    diag = eng.alloc_diag()
    case.gpu_frame(gst, gship, k, diag=diag)
    g, gs = case.to_oracle(gst, gship)
    r = report or Report()
    r.add_exact("bond_intact", g.bond_intact, ost2.bond_intact)
    r.add_exact("elem_active", g.elem_active, ost2.elem_active)
    r.add_exact("floe_eligible", g.dpdem_floe_eligible, ost2.dpdem_floe_eligible)
    gpen = diag["pen_max"].cpu().numpy()
    r.add_exact("has_collision", gpen > 0, od["pen_max"] > 0)
    r.add_float("pen_max", gpen, od["pen_max"])
    r.add_float("force_mag", diag["force_mag"].cpu().numpy(), od["force_mag"])
    r.add_float("positions", g.positions, ost2.positions)
    r.add_float("displacement", g.positions - ost.positions, ost2.positions - ost.positions)
    r.add_float("velocities", g.velocities, ost2.velocities)
    r.add_float("rotations", g.rotations, ost2.rotations, abs_floor=1e-6)
    r.add_float("angular_vels", g.angular_vels, ost2.angular_vels, abs_floor=1e-6)
    r.add_float("submergence", g.submergence, ost2.submergence, abs_floor=1e-6)
    r.add_float("bond_damage", g.bond_damage, ost2.bond_damage, abs_floor=1e-6)
    r.add_float("bond_delta_s", g.bond_delta_s, ost2.bond_delta_s, abs_floor=1e-9)
    r.add_float("bond_delta_m_max", g.bond_delta_m_max, ost2.bond_delta_m_max, abs_floor=1e-9)
    r.add_float("bond_delta_m_0", g.bond_delta_m_0, ost2.bond_delta_m_0, abs_floor=1e-9)
    r.add_float("ship_eta", gs.eta, oship2.eta)
    r.add_float("ship_nu", gs.nu, oship2.nu, abs_floor=1e-6)
    r.add_float("ship_n", gs.n_actual, oship2.n_actual, abs_floor=1e-6)
    r.add_float("F_ice_on_ship", diag["F_ice_on_ship"].cpu().numpy(), od["F_ice_on_ship"], abs_floor=1.0)
    return r


def oracle_run(case: "Case", frames: int, dtype, checkpoints=()):
    """Oracle trajectory of `case` (setpoints in the dtype: float64 rows come
    straight from the float64 trajectory); returns the final (state, ship)
    and {frame: (state, ship)} at the checkpoints (frames completed)."""
    o = case.oracle(dtype)
    st, ship = case.oracle_initial(dtype)
    vw = np.zeros((case.n, 2), dtype)
    snaps = {}
    for k in range(frames):
        i, j = min(k, len(case.traj) - 1), min(k + 1, len(case.traj) - 1)
        sp = np.asarray(case.traj[i], dtype)
        spn = np.asarray(case.traj[j], dtype)
        st, ship, _ = o.frame_step(st, ship, sp, spn, vw, st.dpdem_floe_eligible)
        if k + 1 in checkpoints:
            snaps[k + 1] = (st, ship)
    return st, ship, snaps


def drift_stats(a_state, a_ship, b_state, b_ship) -> Dict[str, float]:
    """Differences of a float32 run (a) from a float64 run (b)."""
    dp = np.linalg.norm(np.asarray(a_state.positions, np.float64) - np.asarray(b_state.positions, np.float64), axis=1)
    dv = np.linalg.norm(np.asarray(a_state.velocities, np.float64) - np.asarray(b_state.velocities, np.float64), axis=1)
    de = np.abs(np.asarray(a_ship.eta, np.float64) - np.asarray(b_ship.eta, np.float64))
    return {"pos_max_m": float(dp.max()), "pos_median_m": float(np.median(dp)),
            "vel_max_mps": float(dv.max()), "ship_pos_m": float(np.hypot(de[0], de[1])), "ship_heading_rad": float(de[2]),
            "bond_flags_differing": int(np.sum(np.asarray(a_state.bond_intact) != np.asarray(b_state.bond_intact))),
            "active_flags_differing": int(np.sum(np.asarray(a_state.elem_active) != np.asarray(b_state.elem_active)))}
