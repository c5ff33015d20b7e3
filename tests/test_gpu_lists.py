"""GPU parity of the contact partner lists against the float32 oracle.

With one ice sub-step per frame the force pass keeps, per element, the
partners within r_i + r_j + 2 delta + skin, rebuilt from the grids when an
element may have moved skin / 2; the frames in between evaluate only the
listed partners.  A call's first frame always builds.  These tests run many
frames in ONE call (graph-replayed), so list frames, build frames and the
switches between them all occur, and compare every frame's positions
(histories) and the final state with the oracle, bit for bit.
"""
import numpy as np
import pytest
import torch

from helpers import Case

pytestmark = pytest.mark.gpu


def _gpu_one_call(case, frames):
    eng = case.engine()
    st, ship = case.gpu_initial()
    rows = np.stack([case.setpoint_row(k) for k in range(frames)])
    sp = torch.as_tensor(rows, device=eng.device)
    vw = None if case.v_water is None else torch.as_tensor(case.v_water, device=eng.device)
    hist = eng.alloc_history(frames)
    before = eng.list_stats()
    eng.run_frames(st, ship, sp, vw, hist=hist, use_graph=True)
    after = eng.list_stats()
    # partners listed for the active elements: a subset, all of them for an all-true mask
    act = eng.list_stats(st.elem_active)
    assert act["active_elements"] == int(st.elem_active.sum())
    assert act["listed_partners"] == after["listed_partners"]
    assert 0 <= act["listed_partners_active"] <= after["listed_partners"]
    every = eng.list_stats(torch.ones_like(st.elem_active))
    assert every["listed_partners_active"] == after["listed_partners"] and every["active_elements"] == case.n
    return st, ship, hist, after["rebuilt_frames"] - before["rebuilt_frames"], after


def _oracle_frames(case, frames):
    o = case.oracle()
    ost, oship = case.oracle_initial()
    vw = case.v_water if case.v_water is not None else np.zeros((case.n, 2), np.float32)
    pos_hist = []
    for k in range(frames):
        r = case.setpoint_row(k)
        ost, oship, _ = o.frame_step(ost, oship, r[:3], r[3:], vw, ost.dpdem_floe_eligible)
        pos_hist.append(ost.positions.copy())
    return ost, oship, pos_hist


def _check(case, frames, expect_lists=True, expect_overflow=False):
    st, ship, hist, rebuilt, stats = _gpu_one_call(case, frames)
    ost, oship, opos = _oracle_frames(case, frames)
    g, gs = case.to_oracle(st, ship)
    print(f"[{case.name}] {frames} frames, {rebuilt} list builds, stats {stats}", flush=True)
    for name in ("positions", "velocities", "rotations", "angular_vels", "submergence", "bond_intact",
                 "bond_damage", "bond_delta_s", "bond_delta_m_0", "bond_delta_m_max", "elem_active",
                 "dpdem_floe_eligible"):
        np.testing.assert_array_equal(getattr(g, name), getattr(ost, name), err_msg=name)
    np.testing.assert_array_equal(gs.eta, oship.eta)
    np.testing.assert_array_equal(gs.nu, oship.nu)
    hp = hist["positions"].cpu().numpy()
    for k in range(frames):
        np.testing.assert_array_equal(hp[k].T, opos[k], err_msg=f"frame {k}")
    assert stats["lists"] == expect_lists
    assert stats["overflowed"] == expect_overflow
    if expect_lists:
        assert 1 <= rebuilt < frames
    return rebuilt


@pytest.mark.parametrize("lanes", [0, 8])
def test_lists_field(lanes):
    case = Case.make("c1small", n_frames=300, force_lanes=lanes)
    _check(case, 300)


def test_lists_field_small_cells_bucket_overflow():
    """Reach boxes over 5 x 5 cells and capacity-1 buckets (most candidates
    through the overflow list) on the build frames."""
    case = Case.make("c1small", n_frames=160, cell_factor=0.55)
    case.bucket_caps = (1, 1)
    _check(case, 160)


def test_lists_c0_ship_contact():
    """The ship ploughs through a floe: fast elements force frequent builds."""
    case = Case.make("c0", n_frames=1100)
    rebuilt = _check(case, 1100)
    assert rebuilt > 1


def test_lists_c0_current_dilation_small_skin():
    from icedem.config import benchmark_config
    cfg = benchmark_config(contact_dilation=0.2)
    case = Case.make("c0", n_frames=950, cfg=cfg, current=(0.3, np.pi / 6))
    case.list_skin = 0.05
    _check(case, 950)


def test_lists_overflow_falls_back_to_the_grids():
    """60 small squares around one big one: the big element has more
    candidates than a list holds and scans the grids every frame."""
    case = Case.make("ring", n_frames=200)
    _check(case, 200, expect_overflow=True)


def test_lists_off_and_substeps_use_the_grids():
    case = Case.make("c1small", n_frames=60)
    case.list_skin = 0.0
    _check(case, 60, expect_lists=False)
    from icedem.config import benchmark_config
    cfg = benchmark_config(use_rayleigh_dt_cap=True, rayleigh_safety_factor=0.5)
    case2 = Case.make("c0", n_frames=820, cfg=cfg)
    assert case2.engine().params.n_ice_steps == 3
    _check(case2, 820, expect_lists=False)


@pytest.mark.parametrize("margin", [0.0, 25.0])
def test_near_candidates_missed_or_many(margin):
    """The post pass tests the next frame's near elements among candidates the
    previous cull listed.  Margin 0 leaves the elements within the near
    test's 1 m band unlisted, so they take the test where they are culled;
    25 m lists a wide band, most of which fails the exact test.  Either way
    every frame equals the oracle bit for bit."""
    case = Case.make("c0", n_frames=1100)
    case.near_cand_margin = margin
    _check(case, 1100)
