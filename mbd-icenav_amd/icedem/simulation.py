"""Reference-shaped simulation API on the MI355X step.

Drop-in for ``medea_ice_ship/jax_simulation.py``: the same function names,
positional signatures and return structures, running the frame step in
``libicedem.so`` (float32, on the GPU):

* ``_make_frame_step`` -> ``frame_step(state, ship_state, setpoint,
  setpoint_next, v_water, dpdem_eligible, v_water_ship=zeros(2))``
  (jax_simulation.py:302-701)
* ``build_jax_sim_context`` / ``JaxSimContext.scan_segment`` /
  ``_run_jax_segment`` (:1211-1384)
* drivers ``jax_simulate_ice_ship_interaction`` (:706-973),
  ``simulate_full_jax_scan`` (:978-1203), ``simulate_jax_chunked_replan``
  (:1438-1793) with the 11-argument positional signature and 6-tuple results.

Neutral aliases without the ``jax`` prefix are exported as well.  Device
arrays are torch tensors; the step is functional (inputs are not modified).
"""
from __future__ import annotations

import time
from typing import NamedTuple, Optional

import numpy as np
import torch

from .config import SimulationConfig, get_default_simulation_config
from .engine import DemEngine
from .ice_state import (IceState, IceStaticData, ShipDynamicsData, ShipState, FossenSupply3DOF,
                        build_initial_ice_state, build_initial_ship_state, build_ship_dynamics_data,
                        build_static_data, ship_state_from_buffer)
from .ocean_current import NoCurrentModel
from .properties import IceProperties


def _time_stepping(static: IceStaticData, cfg, dt: float):
    """dt_ice and n_ice_steps (jax_simulation.py:771-785)."""
    nu_r = getattr(cfg, "poisson_ratio", 0.3)
    G_shear = IceProperties.YOUNGS_MODULUS / (2.0 * (1.0 + nu_r))
    rho_ice = getattr(cfg, "ice_density", IceProperties.ICE_DENSITY)
    r_min = float(np.array(static.radii).min())
    b_safe = getattr(cfg, "rayleigh_safety_factor", 0.1)
    dt_rayleigh = b_safe * np.pi * r_min / (0.163 * nu_r + 0.877) * np.sqrt(rho_ice / G_shear)
    dt_ice = min(dt, dt_rayleigh) if getattr(cfg, "use_rayleigh_dt_cap", False) else dt
    n_ice_steps = max(1, int(round(dt / dt_ice))) if dt_ice > 0 else 1
    return dt_ice, n_ice_steps


def _thruster_config(ship_props, ship_length):
    s = FossenSupply3DOF(mass=ship_props.mass, length=ship_length)
    return np.asarray(s.B), float(s.T_n)


def _as_dev(x, device, dtype=torch.float32):
    t = torch.as_tensor(np.asarray(x) if not torch.is_tensor(x) else x)
    return t.to(device=device, dtype=dtype)


class FrameStep:
    """The compiled single-frame step returned by :func:`_make_frame_step`.

    The broad-phase grid covers the centroids given at construction, or, when
    none were given, the positions of the first state the step receives; a
    state with an element outside the grid's extent (positions outside still
    give exact results, but every such element lands in a border cell) makes
    the step rebuild its grid around that state (``regrids`` counts it)."""

    def __init__(self, make_engine, centroids=None):
        self._make = make_engine
        self._engine = make_engine(centroids) if centroids is not None else None
        self.regrids = 0

    @property
    def engine(self) -> DemEngine:
        if self._engine is None:
            self._engine = self._make(None)
        return self._engine

    def _grid_for(self, state: IceState) -> DemEngine:
        pos = state.positions
        if self._engine is None:
            self._engine = self._make(torch.as_tensor(pos).detach().cpu().numpy().astype(np.float64))
            return self._engine
        g = self._engine.grid
        lo = torch.as_tensor(pos).detach().amin(dim=0).cpu().numpy()
        hi = torch.as_tensor(pos).detach().amax(dim=0).cpu().numpy()
        if lo[0] < g[0] or lo[1] < g[1] or hi[0] > g[0] + g[4] * g[2] or hi[1] > g[1] + g[5] * g[2]:
            self._engine.close()
            self._engine = self._make(torch.as_tensor(pos).detach().cpu().numpy().astype(np.float64))
            self.regrids += 1
        return self._engine

    def __call__(self, state: IceState, ship_state: ShipState, setpoint, setpoint_next, v_water,
                 dpdem_eligible, v_water_ship=None):
        eng = self._grid_for(state)
        st = eng.clone_state(state)
        st.dpdem_floe_eligible.copy_(_as_dev(dpdem_eligible, eng.device, torch.bool))
        ship_buf = eng.ship_buffer(ship_state)
        sp = torch.cat([_as_dev(setpoint, eng.device).reshape(3), _as_dev(setpoint_next, eng.device).reshape(3)])
        vw = None if v_water is None else _as_dev(v_water, eng.device)
        vws = None if v_water_ship is None else _as_dev(v_water_ship, eng.device).reshape(1, 2)
        diag = eng.alloc_diag()
        eng.run_frames(st, ship_buf, sp.reshape(1, 6), vw, vws, hist=None, diag=diag, use_graph=False)
        out_diag = {"pen_max": diag["pen_max"], "force_mag": diag["force_mag"],
                    "F_ice_on_ship": diag["F_ice_on_ship"], "J_cumul_incr": diag["J_cumul_incr"][0]}
        return st, ship_state_from_buffer(ship_buf), out_diag


def _make_frame_step(static: IceStaticData, cfg, ship_dyn: ShipDynamicsData, ship_length, ship_beam, n_ice_steps,
                     dt, dt_ice, use_auto_icenav, B_mat, T_n, centroids=None, device=None, cell_size=None):
    """jax_simulation.py:302-701.  ``centroids`` (not a reference argument)
    sets the broad-phase grid's extent; without it the grid is sized from the
    first state the returned step receives (see :class:`FrameStep`)."""

    def make(cen):
        if cen is None:
            cen = np.zeros((len(static.radii), 2))
        return DemEngine(static, cfg, ship_dyn, ship_length, ship_beam, n_ice_steps, dt, dt_ice, use_auto_icenav,
                         B_mat, T_n, centroids=cen, device=device, cell_size=cell_size)
    return FrameStep(make, centroids)


class JaxSimContext(NamedTuple):
    """jax_simulation.py:1211-1226."""
    frame_step: object
    scan_segment: object
    static: IceStaticData
    ship_dyn: object
    n: int
    n_bonds: int
    dt: float
    chunk_size: int


def build_jax_sim_context(polygons, centroids, bonds, ship_pos, ship_length, ship_beam, n_frames, dt, drift_vel,
                          thicknesses, ship_props, config=None, parent_ids=None, use_auto_icenav_ship: bool = False,
                          chunk_size: int = 50, verbose: bool = True, device=None, use_graph: bool = True,
                          ship_dyn: Optional[ShipDynamicsData] = None) -> JaxSimContext:
    """jax_simulation.py:1229-1358: static data, parameters, device context."""
    cfg = config or get_default_simulation_config()
    if verbose:
        print("  [icedem-ctx] Building static data...", flush=True)
    static = build_static_data(polygons, centroids, bonds, thicknesses, parent_ids, cfg)
    if ship_dyn is None:
        ship_dyn = build_ship_dynamics_data(ship_props, cfg, ship_length, ship_beam)
    dt_ice, n_ice_steps = _time_stepping(static, cfg, dt)
    B_mat, T_n = _thruster_config(ship_props, ship_length)
    t0 = time.perf_counter()
    fs = _make_frame_step(static, cfg, ship_dyn, ship_length, ship_beam, n_ice_steps, dt, dt_ice,
                          use_auto_icenav_ship, B_mat, T_n, centroids=centroids, device=device)
    eng = fs.engine

    def scan_segment(state, ship_state, setpoints_K6, v_water_Kn2, v_water_ship_K2):
        K = int(np.shape(setpoints_K6)[0])
        st = eng.clone_state(state)
        ship_buf = eng.ship_buffer(ship_state)
        hist = eng.alloc_history(K)
        eng.run_frames(st, ship_buf, _as_dev(setpoints_K6, eng.device),
                       None if v_water_Kn2 is None else _as_dev(v_water_Kn2, eng.device),
                       None if v_water_ship_K2 is None else _as_dev(v_water_ship_K2, eng.device),
                       hist=hist, use_graph=use_graph)
        histories = (hist["positions"].permute(0, 2, 1), hist["rotations"], hist["bond_intact"], hist["eta"],
                     hist["submergence"], hist["J"], hist["F"])
        return st, ship_state_from_buffer(ship_buf), histories

    if verbose:
        print(f"  [icedem-ctx] Ready in {time.perf_counter() - t0:.2f} s "
              f"({eng.n} elements, {eng.n_bonds} bonds, grid {eng.grid[4]}x{eng.grid[5]})", flush=True)
    return JaxSimContext(fs, scan_segment, static, ship_dyn, eng.n, eng.n_bonds, dt, chunk_size)


def _run_jax_segment(ctx: JaxSimContext, state: IceState, ship_state: ShipState, setpoints_K6, v_water_Kn2,
                     v_water_ship_K2=None):
    """jax_simulation.py:1361-1384."""
    if v_water_ship_K2 is None:
        v_water_ship_K2 = np.zeros((np.shape(setpoints_K6)[0], 2), dtype=np.float32)
    return ctx.scan_segment(state, ship_state, setpoints_K6, v_water_Kn2, v_water_ship_K2)


def _trajectory_list_to_arr(traj_list) -> np.ndarray:
    """jax_simulation.py:1387-1402."""
    return np.array([[float(p[0]), float(p[1]), float(h)] for p, h in traj_list], dtype=np.float64)


def _extract_world_polygons_np(state: IceState, static: IceStaticData):
    """jax_simulation.py:1405-1435 (host)."""
    pos = state.positions.detach().cpu().numpy().astype(np.float64)
    rot = state.rotations.detach().cpu().numpy().astype(np.float64)
    lp = static.local_polys_padded
    c, s = np.cos(rot), np.sin(rot)
    R = np.stack([np.stack([c, -s], axis=1), np.stack([s, c], axis=1)], axis=1)
    world = np.einsum("nvi,nji->nvj", lp, R) + pos[:, None, :]
    return [world[i, : int(static.poly_n_verts[i])] for i in range(len(static.poly_n_verts))]


def _setpoints_for(traj, frame_start, K, eta=None):
    if traj is None:
        row = np.concatenate([eta, eta])[None, :]
        return np.repeat(row, K, axis=0)
    frames = np.arange(frame_start, frame_start + K)
    idx = np.minimum(frames, len(traj) - 1)
    idx_next = np.minimum(frames + 1, len(traj) - 1)
    return np.concatenate([traj[idx], traj[idx_next]], axis=1)


def _currents(ocean_current, K, state, ship_xy, device):
    """Water velocity at the elements (K, n, 2) (None: still water) and at the
    ship (K, 2) for a chunk of K frames, advancing the model K steps
    (jax_simulation.py:1610-1638): a model with a chunk sampler is sampled
    once and held over the chunk (static fields); any other is sampled at the
    chunk's start positions and stepped frame by frame, for the ice and the
    ship alike."""
    ship_pt = np.asarray(ship_xy, dtype=np.float64)[None, :]
    if ocean_current is None or isinstance(ocean_current, NoCurrentModel) or hasattr(ocean_current, "sample_chunk"):
        vw = None
        if ocean_current is not None and not isinstance(ocean_current, NoCurrentModel):
            vw = ocean_current.sample_chunk(K, state.positions.shape[0], device=device)
        vs = np.zeros(2, np.float32) if ocean_current is None else ocean_current.sample_batch(ship_pt)[0]
        if ocean_current is not None:
            for _ in range(K):
                ocean_current.step()
        return vw, np.tile(np.asarray(vs, dtype=np.float32), (K, 1))
    pos = state.positions.detach().cpu().numpy()
    rows, ship_rows = [], []
    for _ in range(K):
        rows.append(ocean_current.sample_batch(pos))
        ship_rows.append(ocean_current.sample_batch(ship_pt)[0])
        ocean_current.step()
    return np.stack(rows).astype(np.float32), np.stack(ship_rows).astype(np.float32)


def simulate_jax_chunked_replan(polygons, centroids, bonds, ship_pos, ship_length, ship_beam, n_frames, dt,
                                drift_vel, thicknesses, ship_props, ship_trajectory=None,
                                return_ship_diagnostics: bool = False, config=None, replan_every_n_frames=None,
                                replan_callback=None, frame_callback=None, parent_ids=None, ocean_current=None,
                                goal_point=None, goal_y_line=None, goal_tolerance: float = 15.0,
                                use_auto_icenav_ship: bool = False, skip_polygon_extraction: bool = False,
                                verbose: bool = True, device=None):
    """jax_simulation.py:1438-1793: fixed-length chunks with replanning."""
    cfg = config or get_default_simulation_config()
    if ocean_current is None:
        ocean_current = NoCurrentModel()
    n = len(polygons)
    chunk = replan_every_n_frames if (replan_every_n_frames and replan_every_n_frames > 0) else n_frames
    ctx = build_jax_sim_context(polygons, centroids, bonds, ship_pos, ship_length, ship_beam, n_frames, dt, drift_vel,
                                thicknesses, ship_props, config=cfg, parent_ids=parent_ids,
                                use_auto_icenav_ship=use_auto_icenav_ship, chunk_size=chunk, verbose=verbose,
                                device=device)
    eng = ctx.frame_step.engine
    if ship_trajectory is not None:
        if isinstance(ship_trajectory, (list, tuple)):
            ship_trajectory = _trajectory_list_to_arr(ship_trajectory)
        else:
            ship_trajectory = np.asarray(ship_trajectory, dtype=np.float64)
    ship_start_pos = np.array(ship_pos)
    ship_start_head = 0.0
    if ship_trajectory is not None and len(ship_trajectory) > 0:
        ship_start_pos = ship_trajectory[0, :2].copy()
        ship_start_head = float(ship_trajectory[0, 2])
    state = build_initial_ice_state(n, centroids, len(bonds), np.array(ctx.static.initial_floe_of), eng.device)
    ship_state = build_initial_ship_state(ship_start_pos, ship_start_head, eng.device)
    positions_history = [state.positions.cpu().numpy()]
    rotations_history = [state.rotations.cpu().numpy()]
    bonds_history = [state.bond_intact.cpu().numpy()]
    submergence_history = [state.submergence.cpu().numpy()]
    ship_history = []
    ship_diagnostics = [] if return_ship_diagnostics else None
    J_cumul = 0.0
    frame_start = 0
    use_ship_dynamics = ship_trajectory is not None
    while frame_start < n_frames:
        frame_end = min(frame_start + chunk, n_frames)
        K = frame_end - frame_start
        if use_ship_dynamics and replan_callback is not None:
            world_polygons = None if skip_polygon_extraction else _extract_world_polygons_np(state, ctx.static)
            pos_np = state.positions.cpu().numpy().astype(np.float64)
            eta_np = ship_state.eta.cpu().numpy().astype(np.float64)
            new_traj = replan_callback(frame_start, world_polygons, pos_np.copy(), np.array(ctx.static.thicknesses),
                                       eta_np[:2].copy(), float(eta_np[2]), n_frames - frame_start)
            if new_traj is not None and len(new_traj) >= 1:
                new_traj = (_trajectory_list_to_arr(new_traj) if isinstance(new_traj, (list, tuple))
                            else np.asarray(new_traj, dtype=np.float64))
                new_traj = new_traj[: len(ship_trajectory) - frame_start]
                end_slice = frame_start + len(new_traj)
                ship_trajectory[frame_start:end_slice] = new_traj
                if end_slice < len(ship_trajectory):
                    ship_trajectory[end_slice:] = new_traj[-1]
        sp = _setpoints_for(ship_trajectory, frame_start, K, ship_state.eta.cpu().numpy())
        vw, vws = _currents(ocean_current, K, state, ship_state.eta.cpu().numpy()[:2], eng.device)
        pad = chunk - K
        if pad > 0:
            sp = np.concatenate([sp, np.repeat(sp[-1:], pad, axis=0)], axis=0)
            vws = np.concatenate([vws, np.repeat(vws[-1:], pad, axis=0)], axis=0)
            if vw is not None:
                vw_t = torch.as_tensor(vw).to(eng.device)
                vw = torch.cat([vw_t, vw_t[-1:].expand(pad, -1, -1)], dim=0)
        final_state, final_ship, hist = _run_jax_segment(ctx, state, ship_state, sp, vw, vws)
        pos_K, rot_K, bond_K, eta_K, sub_K, j_K, fship_K = [h[:K].cpu().numpy() for h in hist]
        for f in range(K):
            positions_history.append(pos_K[f]); rotations_history.append(rot_K[f])
            bonds_history.append(bond_K[f]); submergence_history.append(sub_K[f])
            ship_history.append((eta_K[f, :2].copy(), float(eta_K[f, 2])))
            J_cumul += float(j_K[f])
            if ship_diagnostics is not None:
                ship_diagnostics.append({"t": (frame_start + f) * dt, "F_ice_ship": fship_K[f],
                                         "ship_pos": eta_K[f, :2].copy(), "impulse_cumul": J_cumul})
        if frame_callback is not None:
            frame_callback(frame_end - 1, pos_K[-1], rot_K[-1], bond_K[-1], eta_K[-1, :2], float(eta_K[-1, 2]))
        ship_xy = eta_K[-1, :2]
        if goal_y_line is not None and ship_xy[1] >= goal_y_line - goal_tolerance:
            crossed = eta_K[:, 1] >= goal_y_line - goal_tolerance
            actual = frame_start + int(np.argmax(crossed)) + 1
            positions_history = positions_history[: actual + 1]
            rotations_history = rotations_history[: actual + 1]
            bonds_history = bonds_history[: actual + 1]
            submergence_history = submergence_history[: actual + 1]
            ship_history = ship_history[:actual]
            if ship_diagnostics:
                ship_diagnostics = ship_diagnostics[:actual]
            state, ship_state = final_state, final_ship
            break
        if goal_point is not None and np.linalg.norm(ship_xy - np.array(goal_point)) <= goal_tolerance:
            state, ship_state = final_state, final_ship
            break
        state, ship_state = final_state, final_ship
        frame_start = frame_end
    damage_summary = {"impulse_total": J_cumul}
    if return_ship_diagnostics:
        if ship_diagnostics:
            ship_diagnostics[-1]["damage_summary"] = damage_summary
        return (positions_history, rotations_history, bonds_history, ship_history, submergence_history,
                ship_diagnostics)
    return positions_history, rotations_history, bonds_history, ship_history, submergence_history, damage_summary


def simulate_full_jax_scan(polygons, centroids, bonds, ship_pos, ship_length, ship_beam, n_frames, dt, drift_vel,
                           thicknesses, ship_props, ship_trajectory=None, config=None, parent_ids=None,
                           v_water_all=None, v_water_ship_all=None, use_auto_icenav_ship: bool = False,
                           device=None):
    """jax_simulation.py:978-1203: every frame in one device segment."""
    cfg = config or get_default_simulation_config()
    n = len(polygons)
    ctx = build_jax_sim_context(polygons, centroids, bonds, ship_pos, ship_length, ship_beam, n_frames, dt, drift_vel,
                                thicknesses, ship_props, config=cfg, parent_ids=parent_ids,
                                use_auto_icenav_ship=use_auto_icenav_ship and ship_trajectory is not None,
                                chunk_size=n_frames, verbose=False, device=device)
    eng = ctx.frame_step.engine
    ship_start_pos = np.array(ship_pos)
    ship_start_head = 0.0
    traj = None
    if ship_trajectory is not None and len(ship_trajectory) > 0:
        traj = (_trajectory_list_to_arr(ship_trajectory) if isinstance(ship_trajectory, (list, tuple))
                else np.asarray(ship_trajectory, dtype=np.float64))
        ship_start_pos = traj[0, :2].copy()
        ship_start_head = float(traj[0, 2])
    state = build_initial_ice_state(n, centroids, len(bonds), np.array(ctx.static.initial_floe_of), eng.device)
    ship_state = build_initial_ship_state(ship_start_pos, ship_start_head, eng.device)
    sp = _setpoints_for(traj, 0, n_frames, ship_state.eta.cpu().numpy())
    init_positions = state.positions.cpu().numpy()
    _, final_ship, hist = ctx.scan_segment(state, ship_state, sp, v_water_all, v_water_ship_all)
    pos_all, rot_all, bond_all, eta_all, sub_all, j_all, _ = [h.cpu().numpy() for h in hist]
    positions_history = [init_positions] + [pos_all[f] for f in range(n_frames)]
    rotations_history = [np.zeros(n, np.float32)] + [rot_all[f] for f in range(n_frames)]
    bonds_history = [np.ones(len(bonds), bool)] + [bond_all[f] for f in range(n_frames)]
    submergence_history = [np.zeros(n, np.float32)] + [sub_all[f] for f in range(n_frames)]
    ship_history = [(eta_all[f, :2].copy(), float(eta_all[f, 2])) for f in range(n_frames)]
    return (positions_history, rotations_history, bonds_history, ship_history, submergence_history,
            {"impulse_total": float(np.sum(j_all))})


def jax_simulate_ice_ship_interaction(polygons, centroids, bonds, ship_pos, ship_length, ship_beam, n_frames, dt,
                                      drift_vel, thicknesses, ship_props, ship_trajectory=None,
                                      return_ship_diagnostics=False, config=None, replan_every_n_frames=None,
                                      replan_callback=None, frame_callback=None, kinematic_ship=False,
                                      parent_ids=None, ocean_current=None, ship_drift_only=False, goal_point=None,
                                      goal_y_line=None, goal_tolerance=15.0, use_auto_icenav_ship: bool = False,
                                      device=None):
    """jax_simulation.py:706-973: Python loop over frames, one device frame each."""
    cfg = config or get_default_simulation_config()
    if ocean_current is None:
        ocean_current = NoCurrentModel()
    n = len(polygons)
    static = build_static_data(polygons, centroids, bonds, thicknesses, parent_ids, cfg)
    ship_dyn = build_ship_dynamics_data(ship_props, cfg, ship_length, ship_beam)
    ship_start_pos = np.array(ship_pos)
    ship_start_head = 0.0
    if ship_trajectory is not None and len(ship_trajectory) > 0:
        ship_start_pos = np.array(ship_trajectory[0][0])
        ship_start_head = float(ship_trajectory[0][1])
    dt_ice, n_ice_steps = _time_stepping(static, cfg, dt)
    B_mat, T_n = _thruster_config(ship_props, ship_length)
    use_auto = use_auto_icenav_ship and ship_trajectory is not None and not kinematic_ship
    frame_step = _make_frame_step(static, cfg, ship_dyn, ship_length, ship_beam, n_ice_steps, dt, dt_ice, use_auto,
                                  B_mat, T_n, centroids=centroids, device=device)
    eng = frame_step.engine
    state = build_initial_ice_state(n, centroids, len(bonds), np.array(static.initial_floe_of), eng.device)
    ship_state = build_initial_ship_state(ship_start_pos, ship_start_head, eng.device)
    positions_history = [state.positions.cpu().numpy()]
    rotations_history = [state.rotations.cpu().numpy()]
    bonds_history = [state.bond_intact.cpu().numpy()]
    ship_history = []
    submergence_history = [state.submergence.cpu().numpy()]
    ship_diagnostics = [] if return_ship_diagnostics else None
    J_cumul = 0.0
    use_ship_dynamics = ship_trajectory is not None and not kinematic_ship
    for frame in range(n_frames):
        if (use_ship_dynamics and replan_callback is not None and replan_every_n_frames
                and replan_every_n_frames > 0 and frame % replan_every_n_frames == 0):
            world_polygons = _extract_world_polygons_np(state, static)
            eta_np = ship_state.eta.cpu().numpy().astype(np.float64)
            new_traj = replan_callback(frame, world_polygons, state.positions.cpu().numpy().copy(),
                                       np.array(static.thicknesses), eta_np[:2].copy(), float(eta_np[2]),
                                       n_frames - frame)
            if new_traj is not None and len(new_traj) >= 1:
                end_slice = frame + len(new_traj)
                ship_trajectory[frame:end_slice] = new_traj
                if end_slice < len(ship_trajectory):
                    ship_trajectory[end_slice:] = [new_traj[-1]] * (len(ship_trajectory) - end_slice)
        if ship_trajectory is not None:
            idx = min(frame, len(ship_trajectory) - 1)
            idx_next = min(frame + 1, len(ship_trajectory) - 1)
            sp_pos, sp_heading = ship_trajectory[idx]
            sp_pos_next, sp_heading_next = ship_trajectory[idx_next]
            setpoint = [float(sp_pos[0]), float(sp_pos[1]), float(sp_heading)]
            setpoint_next = [float(sp_pos_next[0]), float(sp_pos_next[1]), float(sp_heading_next)]
        else:
            setpoint = setpoint_next = ship_state.eta
        pos_np = state.positions.cpu().numpy()
        v_water = ocean_current.sample_batch(pos_np)
        ship_xy = ship_state.eta.cpu().numpy()[:2]
        v_water_ship = ocean_current.sample_batch(ship_xy[None, :])[0]
        state, ship_state, diag = frame_step(state, ship_state, setpoint, setpoint_next, v_water,
                                             state.dpdem_floe_eligible, v_water_ship)
        J_cumul += float(diag["J_cumul_incr"])
        eta_np = ship_state.eta.cpu().numpy()
        ship_history.append((eta_np[:2].copy(), float(eta_np[2])))
        positions_history.append(state.positions.cpu().numpy())
        rotations_history.append(state.rotations.cpu().numpy())
        bonds_history.append(state.bond_intact.cpu().numpy())
        submergence_history.append(state.submergence.cpu().numpy())
        if ship_diagnostics is not None:
            ship_diagnostics.append({"t": frame * dt, "F_ice_ship": diag["F_ice_on_ship"].cpu().numpy(),
                                     "ship_pos": eta_np[:2].copy(), "impulse_cumul": J_cumul})
        if frame_callback is not None:
            frame_callback(frame, state.positions.cpu().numpy(), state.rotations.cpu().numpy(),
                           state.bond_intact.cpu().numpy(), eta_np[:2], float(eta_np[2]))
        ship_xy = eta_np[:2]
        if goal_y_line is not None:
            if ship_xy[1] >= goal_y_line - goal_tolerance:
                break
        elif goal_point is not None and np.linalg.norm(ship_xy - np.array(goal_point)) <= goal_tolerance:
            break
        ocean_current.step()
    damage_summary = {"impulse_total": J_cumul}
    if return_ship_diagnostics:
        if ship_diagnostics:
            ship_diagnostics[-1]["damage_summary"] = damage_summary
        return (positions_history, rotations_history, bonds_history, ship_history, submergence_history,
                ship_diagnostics)
    return positions_history, rotations_history, bonds_history, ship_history, submergence_history, damage_summary


# neutral names
make_frame_step = _make_frame_step
build_sim_context = build_jax_sim_context
run_segment = _run_jax_segment
simulate_full_scan = simulate_full_jax_scan
simulate_chunked_replan = simulate_jax_chunked_replan
simulate_ice_ship_interaction = jax_simulate_ice_ship_interaction
