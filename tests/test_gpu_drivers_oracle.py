"""The drivers' histories against the float32 oracle, bit for bit.

simulate_full_jax_scan (jax_simulation.py:978-1203) on the reference's
three-hexagon scene, and simulate_jax_chunked_replan (:1438-1793) on the
c0 scene in chunks of 100 frames (the last one padded) with a uniform
current, compared frame by frame with the oracle's frame_step loop fed the
same float32 setpoints and water velocities: positions, rotations, bond
flags, submergence, ship pose and the impulse.
"""
import math

import numpy as np
import pytest

from helpers import oracle_config
from icedem import scenario, simulation as sim
from icedem.config import SimulationConfig, benchmark_config
from icedem.ocean_current import UniformCurrentModel
from icedem.properties import ShipProperties
from oracle import model as om, step as ostep

pytestmark = pytest.mark.gpu


def _oracle(polygons, centroids, bonds, thicknesses, parent_ids, cfg, ship_props, L, B, hull, dt, use_auto):
    ocfg = oracle_config(cfg)
    st64 = om.build_static_data(polygons, centroids, bonds, thicknesses, parent_ids, ocfg)
    sd = om.build_ship_dynamics_data(ship_props.mass, L, B, np.asarray(hull, np.float64))
    fos = om.Fossen(ship_props.mass, L)
    dt_ice, nsteps = om.time_stepping(st64, ocfg, dt)
    return ostep.Oracle(st64, ocfg, sd, L, B, nsteps, dt, dt_ice, use_auto, fos.B, fos.T_n, dtype=np.float32)


def _oracle_run(o, centroids, n_bonds, parent_ids, traj, n_frames, vw=None, vws=None):
    st = ostep.initial_ice_state(centroids, n_bonds, parent_ids, np.float32)
    ship = ostep.initial_ship_state(traj[0, :2], traj[0, 2], np.float32)
    sp = sim._setpoints_for(traj, 0, n_frames).astype(np.float32)
    n = len(centroids)
    vw = np.zeros((n, 2), np.float32) if vw is None else vw
    vws = np.zeros(2, np.float32) if vws is None else vws
    out = {"pos": [], "rot": [], "bond": [], "sub": [], "eta": [], "J": []}
    for k in range(n_frames):
        st, ship, d = o.frame_step(st, ship, sp[k, :3], sp[k, 3:], vw, st.dpdem_floe_eligible, vws)
        out["pos"].append(st.positions.copy()); out["rot"].append(st.rotations.copy())
        out["bond"].append(st.bond_intact.copy()); out["sub"].append(st.submergence.copy())
        out["eta"].append(ship.eta.copy()); out["J"].append(d["J_cumul_incr"])
    return out


def _compare(res, ref, n_frames):
    pos_h, rot_h, bond_h, ship_h, sub_h, summary = res
    assert len(pos_h) == n_frames + 1 and len(ship_h) == n_frames
    for k in range(n_frames):
        np.testing.assert_array_equal(np.asarray(pos_h[k + 1]), ref["pos"][k], err_msg=f"positions, frame {k}")
        np.testing.assert_array_equal(np.asarray(rot_h[k + 1]), ref["rot"][k], err_msg=f"rotations, frame {k}")
        np.testing.assert_array_equal(np.asarray(bond_h[k + 1]), ref["bond"][k], err_msg=f"bonds, frame {k}")
        np.testing.assert_array_equal(np.asarray(sub_h[k + 1]), ref["sub"][k], err_msg=f"submergence, frame {k}")
        eta = np.array([ship_h[k][0][0], ship_h[k][0][1], ship_h[k][1]], np.float32)
        np.testing.assert_array_equal(eta, ref["eta"][k], err_msg=f"ship, frame {k}")
    J = np.array(ref["J"], np.float32)
    assert summary["impulse_total"] == pytest.approx(float(np.sum(J.astype(np.float64))), rel=1e-6, abs=1e-6)


def test_full_scan_equals_oracle_hexagons():
    sc = scenario.hexagon_scene()
    n_frames = 60
    L, B = 15.0, 5.0
    props = ShipProperties(length=L, beam=B, draft=2.5)
    traj = np.array([[-25.0 + 0.5 * i, 0.0, 0.0] for i in range(n_frames + 2)])
    cfg = SimulationConfig()
    res = sim.simulate_full_jax_scan(sc.polygons, sc.centroids, sc.bonds, np.array([-25.0, 0.0]), L, B, n_frames,
                                     0.05, np.zeros(2), sc.thicknesses, props, ship_trajectory=traj, config=cfg,
                                     parent_ids=sc.parent_ids)
    hull = sim.build_ship_dynamics_data(props, cfg, L, B).ship_hull_local
    o = _oracle(sc.polygons, sc.centroids, sc.bonds, sc.thicknesses, sc.parent_ids, cfg, props, L, B, hull, 0.05,
                False)
    ref = _oracle_run(o, sc.centroids, len(sc.bonds), sc.parent_ids, traj, n_frames)
    assert np.max(np.abs(np.array(ref["J"]))) > 0, "the ship must touch the ice"
    _compare(res, ref, n_frames)


def test_chunked_replan_equals_oracle_c0_current():
    sc = scenario.scene_c0(0)
    n_frames = 930
    L, B = 60.0, 12.0
    props = ShipProperties(L, B, 5.0)
    cfg = benchmark_config()
    x0, y0 = scenario.ship_start_below(sc, L)
    traj = scenario.straight_trajectory(x0, y0, math.pi / 2, 2.0, cfg.dt, n_frames)
    current = UniformCurrentModel(0.3, math.pi / 6)
    res = sim.simulate_jax_chunked_replan(sc.polygons, sc.centroids, sc.bonds, np.array([x0, y0]), L, B, n_frames,
                                          cfg.dt, np.zeros(2), sc.thicknesses, props, ship_trajectory=traj.copy(),
                                          config=cfg, replan_every_n_frames=100, parent_ids=sc.parent_ids,
                                          ocean_current=current, verbose=False)
    hull = sim.build_ship_dynamics_data(props, cfg, L, B).ship_hull_local
    o = _oracle(sc.polygons, sc.centroids, sc.bonds, sc.thicknesses, sc.parent_ids, cfg, props, L, B, hull, cfg.dt,
                False)
    v = current.velocity.astype(np.float32)
    ref = _oracle_run(o, sc.centroids, len(sc.bonds), sc.parent_ids, traj, n_frames,
                      vw=np.tile(v, (len(sc.polygons), 1)), vws=v)
    assert np.max(np.abs(np.array(ref["J"]))) > 0, "the ship must touch the ice"
    assert np.sum(~ref["bond"][-1]) > 0, "some bonds must break"
    _compare(res, ref, n_frames)


def test_drop_in_frame_step_sizes_its_grid_from_the_state():
    """_make_frame_step with the reference's arguments only (no centroids):
    the grid is built around the first state it steps, so a real field is
    not clamped into border cells; the frame equals the one of a step built
    with the centroids and takes about as long; a state that leaves the grid
    makes the step rebuild it."""
    import time
    import torch
    from icedem.ice_state import build_initial_ice_state, build_static_data
    sc = scenario.scene_c1(0)
    L, B = 60.0, 12.0
    props = ShipProperties(L, B, 5.0)
    cfg = benchmark_config()
    static = build_static_data(sc.polygons, sc.centroids, sc.bonds, sc.thicknesses, sc.parent_ids, cfg)
    sd = sim.build_ship_dynamics_data(props, cfg, L, B)
    dt_ice, nsteps = sim._time_stepping(static, cfg, cfg.dt)
    Bm, T_n = sim._thruster_config(props, L)
    fs = sim._make_frame_step(static, cfg, sd, L, B, nsteps, cfg.dt, dt_ice, True, Bm, T_n)
    ref = sim._make_frame_step(static, cfg, sd, L, B, nsteps, cfg.dt, dt_ice, True, Bm, T_n, centroids=sc.centroids)
    dev = ref.engine.device
    st = build_initial_ice_state(len(sc.polygons), sc.centroids, len(sc.bonds), sc.parent_ids, dev)
    x0, y0 = scenario.ship_start_below(sc, L, gap=1.0)
    ship = sim.build_initial_ship_state(np.array([x0, y0]), math.pi / 2, dev)
    sp, spn = np.array([x0, y0, math.pi / 2]), np.array([x0, y0 + 0.01, math.pi / 2])
    elig = torch.zeros(sc.n_floes, dtype=torch.bool, device=dev)

    def run(step, s, sh, frames):
        t0 = time.perf_counter()
        for _ in range(frames):
            s, sh, _ = step(s, sh, sp, spn, None, elig)
        torch.cuda.synchronize()
        return s, sh, (time.perf_counter() - t0) / frames
    a, ash, _ = run(fs, st, ship, 3)
    b, bsh, _ = run(ref, st, ship, 3)
    g = fs.engine.grid
    c = np.asarray(sc.centroids)
    assert g[0] <= c[:, 0].min() and g[1] <= c[:, 1].min()
    assert g[0] + g[4] * g[2] >= c[:, 0].max() and g[1] + g[5] * g[2] >= c[:, 1].max()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
    _, _, t_fs = run(fs, a, ash, 20)
    _, _, t_ref = run(ref, b, bsh, 20)
    print(f"drop-in frame_step {1e3 * t_fs:.3f} ms per frame, with centroids {1e3 * t_ref:.3f} ms", flush=True)
    assert t_fs < 1.5 * t_ref
    assert fs.regrids == 0
    far = a._replace(positions=a.positions + torch.tensor([500.0, 0.0], device=dev))
    run(fs, far, ash, 1)
    assert fs.regrids == 1
    g = fs.engine.grid
    assert g[0] + g[4] * g[2] >= c[:, 0].max() + 500.0
