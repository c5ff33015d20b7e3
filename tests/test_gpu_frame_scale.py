"""Frame parity at the benchmark sizes: the full c1 field (7 640 elements,
all-pairs oracle) and the full c2 field (99 943 elements, 279 351 bonds,
5 000 floes, bimodal thickness, a 0.3 m/s current at 30 degrees; the
oracle's cell-list pairs, which enumerate the same pairs in the same order as
all pairs).  The GPU runs the approach in graph-replayed segments, then one
frame is compared from identical input state with the same gate as the
small-scene frame tests."""
import math

import numpy as np
import pytest
import torch

from helpers import Case, compare_frame

pytestmark = pytest.mark.gpu


def _advance(case, st, ship, frames, seg=100):
    eng = case.engine()
    vw = None if case.v_water is None else torch.as_tensor(case.v_water, device=eng.device)
    for k0 in range(0, frames, seg):
        sp = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(k0, k0 + seg)]), device=eng.device)
        eng.run_frames(st, ship, sp, vw, use_graph=True)


@pytest.mark.parametrize("name,frames,current", [("c1", 300, None), ("c2", 200, (0.3, math.pi / 6))])
def test_frame_parity_full_size(name, frames, current):
    case = Case.make(name, n_frames=frames + 2, current=current)
    st, ship = case.gpu_initial()
    _advance(case, st, ship, frames)
    rep = compare_frame(case, st, ship, frames)
    print(f"[{name}] {case.n} elements, frame {frames}: {rep.summary()}", flush=True)
    rep.assert_ok()
