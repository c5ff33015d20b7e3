"""2000 frames of the three-floe scene on the GPU (graph-replayed segments of
500): bit-identical to the float32 oracle's run; the drift of both against
the float64 oracle is reported (the gate on drift is sanity only)."""
import numpy as np
import pytest
import torch

from helpers import Case, Report, drift_stats, oracle_run

pytestmark = pytest.mark.gpu


def test_gpu_2000_frames_bit_identical_and_drift():
    case = Case.make("c0", n_frames=2000)
    eng = case.engine()
    st, ship = case.gpu_initial()
    for k0 in range(0, 2000, 500):
        sp = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(k0, k0 + 500)]), device=eng.device)
        eng.run_frames(st, ship, sp, use_graph=True)
    g, gs = case.to_oracle(st, ship)
    o32, s32, _ = oracle_run(case, 2000, np.float32)
    rep = Report()
    rep.add_float("positions", g.positions, o32.positions)
    rep.add_float("velocities", g.velocities, o32.velocities)
    rep.add_float("eta", gs.eta, s32.eta)
    rep.add_exact("bond_intact", g.bond_intact, o32.bond_intact)
    rep.add_exact("elem_active", g.elem_active, o32.elem_active)
    print("GPU vs float32 oracle after 2000 frames:", rep.summary(), flush=True)
    rep.assert_ok()
    assert np.array_equal(g.positions, o32.positions)
    o64, s64, _ = oracle_run(case, 2000, np.float64)
    d = drift_stats(g, gs, o64, s64)
    print("GPU float32 vs float64 oracle after 2000 frames:", d, flush=True)
    assert d["pos_median_m"] < 1e-3 and d["ship_pos_m"] < 1.0
