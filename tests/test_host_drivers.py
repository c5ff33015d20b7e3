"""CPU checks of the drivers' host logic: (K, 6) setpoint packaging against
the per-frame loop it replaces (test_jax_trajectory_opt.py:214-290), world
polygon extraction against the per-element loop
(test_jax_chunked_replan.py:163-222), trajectory list conversion."""
import numpy as np
import torch

from icedem import simulation as sim
from icedem.config import get_default_simulation_config
from icedem.ice_state import build_initial_ice_state, build_static_data


def _loop_rows(traj_list, f0, f1):
    rows = []
    for f in range(f0, f1):
        (p, h), (pn, hn) = traj_list[min(f, len(traj_list) - 1)], traj_list[min(f + 1, len(traj_list) - 1)]
        rows.append([p[0], p[1], h, pn[0], pn[1], hn])
    return np.array(rows, np.float64)


def _traj(n):
    return [(np.array([float(i), 0.5 * i]), 0.1 * i) for i in range(n)]


def test_setpoint_packaging_matches_loop():
    for n, f0, f1 in [(50, 10, 20), (50, 0, 15), (10, 8, 20), (30, 5, 6)]:
        tl = _traj(n)
        got = sim._setpoints_for(sim._trajectory_list_to_arr(tl), f0, f1 - f0)
        assert got.shape == (f1 - f0, 6)
        np.testing.assert_array_equal(got, _loop_rows(tl, f0, f1))


def test_setpoints_without_trajectory_hold_pose():
    eta = np.array([3.0, -2.0, 0.25])
    got = sim._setpoints_for(None, 7, 4, eta)
    assert got.shape == (4, 6)
    np.testing.assert_array_equal(got, np.tile(np.concatenate([eta, eta]), (4, 1)))


def test_extract_world_polygons_matches_loop():
    angles = np.linspace(0, 2 * np.pi, 7)[:-1]
    hex_pts = np.column_stack([3.0 * np.cos(angles), 3.0 * np.sin(angles)])
    centroids = np.array([[-10.0, 0.0], [0.0, 0.0], [10.0, 0.0]])
    polys = [hex_pts + c for c in centroids]
    static = build_static_data(polys, centroids, [(0, 1), (1, 2)], np.ones(3), np.arange(3, dtype=np.int32),
                               get_default_simulation_config())
    st = build_initial_ice_state(3, centroids, 2, np.array(static.initial_floe_of), torch.device("cpu"))
    st.positions[0] += torch.tensor([1.5, -0.7])
    st.rotations[1] = 0.4
    got = sim._extract_world_polygons_np(st, static)
    pos, rot = st.positions.numpy().astype(np.float64), st.rotations.numpy().astype(np.float64)
    for i in range(3):
        nv = int(static.poly_n_verts[i])
        c, s = np.cos(rot[i]), np.sin(rot[i])
        ref = static.local_polys_padded[i, :nv] @ np.array([[c, -s], [s, c]]).T + pos[i]
        np.testing.assert_allclose(got[i], ref, atol=1e-12)
    # an unrotated element is its input polygon again
    np.testing.assert_allclose(got[2], polys[2], atol=1e-5)
