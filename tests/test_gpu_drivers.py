"""Self-consistency of the three drivers and the segment API on the GPU.

The reference pins its drivers against each other (test_jax_full_scan.py:
233-324, test_jax_chunked_replan.py:68-418, test_jax_end_to_end.py:244-268):
per-frame == full scan == one segment == chunked (with and without padding
of the last chunk, with and without a replan callback) to 1e-8 in float64,
bond flags equal.  Here every driver runs the same kernels in the same order
on the same float32 inputs, so the gate is bit equality.  Scene: the
reference's three hexagons in a line, ship 15 x 5 m moving east 0.5 m per
frame (test_jax_full_scan.py:187-230).
"""
import numpy as np
import pytest

from icedem import simulation as sim
from icedem.ice_state import build_initial_ice_state, build_initial_ship_state
from icedem.properties import ShipProperties

pytestmark = pytest.mark.gpu


def _scenario(n_frames=30, dt=0.05):
    angles = np.linspace(0, 2 * np.pi, 7)[:-1]
    hex_pts = np.column_stack([3.0 * np.cos(angles), 3.0 * np.sin(angles)])
    centroids = np.array([[-10.0, 0.0], [0.0, 0.0], [10.0, 0.0]])
    return dict(polygons=[hex_pts + c for c in centroids], centroids=centroids, bonds=[(0, 1), (1, 2)],
                ship_pos=np.array([-25.0, 0.0]), ship_length=15.0, ship_beam=5.0, n_frames=n_frames, dt=dt,
                drift_vel=np.array([0.0, 0.0]), thicknesses=np.array([1.0, 1.0, 1.0]),
                ship_props=ShipProperties(length=15.0, beam=5.0, draft=2.5),
                ship_trajectory=[(np.array([-25.0 + i * 0.5, 0.0]), 0.0) for i in range(n_frames + 2)],
                parent_ids=np.arange(3, dtype=np.int32))


def _args(sc):
    return (sc["polygons"], sc["centroids"], sc["bonds"], sc["ship_pos"], sc["ship_length"], sc["ship_beam"],
            sc["n_frames"], sc["dt"], sc["drift_vel"], sc["thicknesses"], sc["ship_props"])


def _eta(entry):
    return np.array([entry[0][0], entry[0][1], entry[1]])


def _same_final(a, b):
    np.testing.assert_array_equal(np.asarray(a[0][-1]), np.asarray(b[0][-1]))     # positions
    np.testing.assert_array_equal(np.asarray(a[1][-1]), np.asarray(b[1][-1]))     # rotations
    np.testing.assert_array_equal(np.asarray(a[2][-1]), np.asarray(b[2][-1]))     # bond flags
    np.testing.assert_array_equal(_eta(a[3][-1]), _eta(b[3][-1]))                 # ship pose
    assert len(a[0]) == len(b[0]) and len(a[3]) == len(b[3])


def test_full_scan_matches_per_frame():
    sc = _scenario(n_frames=20)
    pf = sim.jax_simulate_ice_ship_interaction(*_args(sc), ship_trajectory=list(sc["ship_trajectory"]),
                                               parent_ids=sc["parent_ids"])
    fs = sim.simulate_full_jax_scan(*_args(sc), ship_trajectory=sc["ship_trajectory"], parent_ids=sc["parent_ids"])
    _same_final(pf, fs)
    assert len(pf[0]) == sc["n_frames"] + 1 and len(fs[0]) == sc["n_frames"] + 1
    # the ship has moved east along the setpoints
    assert _eta(fs[3][-1])[0] > -25.0
    assert fs[5]["impulse_total"] == pytest.approx(pf[5]["impulse_total"], rel=1e-6)


def test_full_scan_no_trajectory():
    sc = _scenario(n_frames=5)
    res = sim.simulate_full_jax_scan(*_args(sc), ship_trajectory=None, parent_ids=sc["parent_ids"])
    assert len(res[0]) == sc["n_frames"] + 1
    # without a trajectory the ship receives its own pose as setpoint and stays put
    assert np.linalg.norm(res[3][-1][0] - sc["ship_pos"]) < 5.0


def test_run_jax_segment_matches_full_scan():
    sc = _scenario(n_frames=20)
    ref = sim.simulate_full_jax_scan(*_args(sc), ship_trajectory=sc["ship_trajectory"],
                                     parent_ids=sc["parent_ids"])
    ctx = sim.build_jax_sim_context(*_args(sc), parent_ids=sc["parent_ids"], chunk_size=sc["n_frames"],
                                    verbose=False)
    dev = ctx.frame_step.engine.device
    state = build_initial_ice_state(ctx.n, sc["centroids"], ctx.n_bonds, np.array(ctx.static.initial_floe_of), dev)
    ship = build_initial_ship_state(np.array(sc["ship_trajectory"][0][0]), 0.0, dev)
    K = sc["n_frames"]
    traj = sc["ship_trajectory"]
    rows = []
    for f in range(K):
        (p, h), (pn, hn) = traj[min(f, len(traj) - 1)], traj[min(f + 1, len(traj) - 1)]
        rows.append([p[0], p[1], h, pn[0], pn[1], hn])
    final_state, final_ship, hist = sim._run_jax_segment(ctx, state, ship, np.array(rows),
                                                         np.zeros((K, ctx.n, 2), np.float32))
    assert len(hist) == 7 and hist[0].shape == (K, ctx.n, 2) and hist[2].shape == (K, ctx.n_bonds)
    np.testing.assert_array_equal(hist[0][-1].cpu().numpy(), np.asarray(ref[0][-1]))
    np.testing.assert_array_equal(hist[1][-1].cpu().numpy(), np.asarray(ref[1][-1]))
    np.testing.assert_array_equal(hist[2][-1].cpu().numpy(), np.asarray(ref[2][-1]))
    np.testing.assert_array_equal(hist[3][-1].cpu().numpy(), _eta(ref[3][-1]).astype(np.float32))
    # the returned state is the last history row
    np.testing.assert_array_equal(final_state.positions.cpu().numpy(), hist[0][-1].cpu().numpy())
    np.testing.assert_array_equal(final_ship.eta.cpu().numpy(), hist[3][-1].cpu().numpy())


def test_chunked_no_replan_matches_full_scan():
    sc = _scenario(n_frames=20)
    ref = sim.simulate_full_jax_scan(*_args(sc), ship_trajectory=sc["ship_trajectory"],
                                     parent_ids=sc["parent_ids"])
    chk = sim.simulate_jax_chunked_replan(*_args(sc), ship_trajectory=list(sc["ship_trajectory"]),
                                          replan_every_n_frames=sc["n_frames"], replan_callback=None,
                                          parent_ids=sc["parent_ids"], verbose=False)
    _same_final(ref, chk)


def test_chunked_multi_chunk_matches_full_scan():
    """13 frames in chunks of 4: the last chunk is padded (13 = 3 * 4 + 1)."""
    sc = _scenario(n_frames=13)
    ref = sim.simulate_full_jax_scan(*_args(sc), ship_trajectory=sc["ship_trajectory"],
                                     parent_ids=sc["parent_ids"])
    chk = sim.simulate_jax_chunked_replan(*_args(sc), ship_trajectory=list(sc["ship_trajectory"]),
                                          replan_every_n_frames=4, replan_callback=None,
                                          parent_ids=sc["parent_ids"], verbose=False)
    _same_final(ref, chk)
    for f in range(sc["n_frames"] + 1):
        np.testing.assert_array_equal(np.asarray(ref[0][f]), np.asarray(chk[0][f]))


def test_chunked_with_replan_matches_per_frame():
    sc = _scenario(n_frames=20)
    calls = [0]

    def callback(frame, world_polygons, centroids, thicknesses, ship_pos, ship_heading, n_remaining):
        calls[0] += 1
        assert len(world_polygons) == 3 and world_polygons[0].shape == (6, 2)
        return [(np.array([-25.0 + (frame + i) * 0.5, 0.0]), 0.0) for i in range(n_remaining + 1)]

    pf = sim.jax_simulate_ice_ship_interaction(*_args(sc), ship_trajectory=list(sc["ship_trajectory"]),
                                               replan_every_n_frames=5, replan_callback=callback,
                                               parent_ids=sc["parent_ids"])
    n_pf, calls[0] = calls[0], 0
    chk = sim.simulate_jax_chunked_replan(*_args(sc), ship_trajectory=list(sc["ship_trajectory"]),
                                          replan_every_n_frames=5, replan_callback=callback,
                                          parent_ids=sc["parent_ids"], verbose=False)
    assert n_pf == calls[0] == 4        # frames 0, 5, 10, 15
    _same_final(pf, chk)


def test_history_lengths_and_goal_line():
    sc = _scenario(n_frames=15)
    res = sim.simulate_jax_chunked_replan(*_args(sc), ship_trajectory=list(sc["ship_trajectory"]),
                                          replan_every_n_frames=5, parent_ids=sc["parent_ids"], verbose=False)
    for k in (0, 1, 2, 4):
        assert len(res[k]) == sc["n_frames"] + 1
    assert len(res[3]) == sc["n_frames"]
    # stopping at a goal line: both drivers cut the histories after the same
    # frame.  Ship heading north; the line lies between the ship's y at the
    # frame before its first new maximum after frame 5 and its y there.
    sc = _scenario(n_frames=20)
    traj_n = [(np.array([0.0, -25.0 + i * 0.5]), np.pi / 2) for i in range(sc["n_frames"] + 2)]
    free = sim.jax_simulate_ice_ship_interaction(*_args(sc), ship_trajectory=list(traj_n),
                                                 parent_ids=sc["parent_ids"])
    ys = np.array([e[0][1] for e in free[3]], np.float64)
    f = next(k for k in range(5, len(ys)) if ys[k] > ys[:k].max())
    goal = 0.5 * (ys[:f].max() + ys[f])
    pf = sim.jax_simulate_ice_ship_interaction(*_args(sc), ship_trajectory=list(traj_n), parent_ids=sc["parent_ids"],
                                               goal_y_line=goal, goal_tolerance=0.0)
    ck = sim.simulate_jax_chunked_replan(*_args(sc), ship_trajectory=list(traj_n), parent_ids=sc["parent_ids"],
                                         replan_every_n_frames=6, goal_y_line=goal, goal_tolerance=0.0,
                                         verbose=False)
    assert len(pf[3]) == len(ck[3]) == f + 1
    assert len(pf[0]) == len(ck[0]) == f + 2
    np.testing.assert_array_equal(np.asarray(pf[0][-1]), np.asarray(ck[0][-1]))
    np.testing.assert_array_equal(_eta(pf[3][-1]), _eta(ck[3][-1]))


def test_engine_like_shares_tables_and_outlives_base():
    """icedem_create_like: a second context of the same scene runs the same
    frames bit for bit, and keeps working after the context whose tables it
    shares is destroyed."""
    import numpy as np
    import torch
    from helpers import Case
    from icedem.engine import DemEngine
    case = Case.make("c1small", n_frames=70)
    base = case.engine()
    twin = DemEngine.like(base)
    sa, sha = case.gpu_initial()
    sb, shb = case.gpu_initial()
    sp1 = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(30)]), device=base.device)
    sp2 = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(30, 60)]), device=base.device)
    base.run_frames(sa, sha, sp1, use_graph=False)
    twin.run_frames(sb, shb, sp1, use_graph=True)
    assert torch.equal(sa.positions, sb.positions) and torch.equal(sha, shb)
    assert torch.equal(sa.bond_intact, sb.bond_intact)
    base.close()
    case._engine = None
    fresh = case.engine()
    fresh.run_frames(sa, sha, sp2, use_graph=False)
    twin.run_frames(sb, shb, sp2, use_graph=True)
    assert torch.equal(sa.positions, sb.positions) and torch.equal(sha, shb)
    assert torch.equal(sa.velocities, sb.velocities) and torch.equal(sa.bond_intact, sb.bond_intact)
    twin.close()
