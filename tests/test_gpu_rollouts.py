"""Batched rollouts (both batch modes) == per-rollout frame calls == oracle.

The rollout ship control (direct thrust) and the cost terms are this build's
extension; the oracle restates them (oracle.step.rollout) with the same
operation order, and the ice physics is the reference path.
"""
import math

import numpy as np
import pytest
import torch

from helpers import oracle_config
from oracle import model as om, step as ostep

pytestmark = pytest.mark.gpu


def _batch(seeds_sizes, inst, force_lanes=0):
    from icedem import scenario
    from icedem.rollouts import RolloutBatch
    scenes = [scenario.scene_patch(s, n, c) for s, n, c in seeds_sizes]
    return RolloutBatch(scenes, inst, force_lanes=force_lanes), scenes


def _oracle_for(rb, scene):
    cfg = rb.cfg
    ocfg = oracle_config(cfg)
    st64 = om.build_static_data(scene.polygons, scene.centroids, scene.bonds, scene.thicknesses, scene.parent_ids,
                                ocfg)
    sd = om.build_ship_dynamics_data(rb.ship_dyn.mass, rb.ship.length, rb.ship.beam, rb.ship_dyn.ship_hull_local)
    fos = om.Fossen(rb.ship_dyn.mass, rb.ship.length)
    return om, ostep.Oracle(st64, ocfg, sd, rb.ship.length, rb.ship.beam, 1, cfg.dt, cfg.dt, False, fos.B, fos.T_n,
                            dtype=np.float32, ship_mode=1, ctrl_thrust_max=rb.ship.thrust_max,
                            ctrl_yaw_rate_max=rb.ship.yaw_rate_max)


@pytest.mark.parametrize("mode,lanes", [("frames", 0), ("scenes", 0), ("frames", 8)])
def test_batched_equals_sequential_and_oracle(mode, lanes):
    from icedem.rollouts import nominal_controls
    rb, scenes = _batch([(1000, 60, 0.3), (1001, 150, 0.4), (1002, 90, 0.5)], [0, 1, 2, 0], force_lanes=lanes)
    rb.set_mode(mode)
    frames, hold, nk = 120, 5, 24
    ctl = nominal_controls(rb.R, nk, seed=3)
    rb.reset()
    out_b = rb.run(torch.as_tensor(ctl), hold, frames).cpu().numpy().copy()
    pos_b = [v.positions.cpu().numpy().copy() for v in rb.views]
    bonds_b = [v.bond_intact.cpu().numpy().copy() for v in rb.views]
    ship_b = rb.ships.cpu().numpy().copy()
    print("batched costs", out_b, flush=True)
    assert np.all(out_b[:, 3] > 0) and np.any(out_b[:, 0] > 0), "ship must reach the ice"
    rb.reset()
    out_s = rb.run_sequential(torch.as_tensor(ctl), hold, frames).cpu().numpy().copy()
    np.testing.assert_array_equal(out_s, out_b)
    for r, v in enumerate(rb.views):
        np.testing.assert_array_equal(v.positions.cpu().numpy(), pos_b[r])
        np.testing.assert_array_equal(v.bond_intact.cpu().numpy(), bonds_b[r])
    np.testing.assert_array_equal(rb.ships.cpu().numpy(), ship_b)
    # oracle, instances 0 (scene 0) and 1 (scene 1)
    for r in (0, 1):
        sc = scenes[rb.inst[r]]
        _, o = _oracle_for(rb, sc)
        st = ostep.initial_ice_state(sc.centroids, len(sc.bonds), sc.parent_ids, np.float32)
        s0 = rb.ship_init[r].cpu().numpy()
        ship = ostep.ShipState(s0[0:3].copy(), s0[3:6].copy(), s0[6:12].copy(), s0[12:15].copy())
        st, ship, out_o = ostep.rollout(o, st, ship, ctl[r], hold, frames)
        print("oracle", r, out_o, "gpu", out_b[r], flush=True)
        np.testing.assert_array_equal(out_b[r], out_o)
        np.testing.assert_array_equal(pos_b[r], st.positions)
        np.testing.assert_array_equal(bonds_b[r], st.bond_intact)
    rb.close()
