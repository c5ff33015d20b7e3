"""GPU parity of the contact partner lists against the float32 oracle.

With one ice sub-step per frame the force pass keeps, per element, the
partners within r_i + r_j + 2 delta + skin, rebuilt from the grids when an
element may have moved skin / 2; the frames in between evaluate only the
listed partners.  A call's first frame builds when an element lies farther
than that from its build position; a build keeps the lists of frozen
elements no live element came near since the last one.  These tests run many
frames in ONE call (graph-replayed), so list frames, build frames and the
switches between them all occur, and compare every frame's positions
(histories) and the final state with the oracle, bit for bit.
"""
import numpy as np
import pytest
import torch

from helpers import Case
from icedem.engine import DEFAULT_LIST_SKIN

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


def _chained(case, sizes, edit=None):
    """Frames in calls of the given sizes (graph-replayed); edit(call index,
    st) may change the state between calls.  Every frame and the final state
    are compared with the oracle run over the same frames from the same
    edited states."""
    eng = case.engine()
    st, ship = case.gpu_initial()
    o = case.oracle()
    ost, oship = case.oracle_initial()
    vw = case.v_water if case.v_water is not None else np.zeros((case.n, 2), np.float32)
    before = eng.list_stats()
    k0 = 0
    for c, K in enumerate(sizes):
        if edit is not None and edit(c, st):
            ost, oship = case.to_oracle(st, ship)   # the oracle continues from the edited state
        rows = np.stack([case.setpoint_row(k) for k in range(k0, k0 + K)])
        hist = eng.alloc_history(K)
        eng.run_frames(st, ship, torch.as_tensor(rows, device=eng.device), None if case.v_water is None
                       else torch.as_tensor(vw, device=eng.device), hist=hist, use_graph=True)
        hp = hist["positions"].cpu().numpy()
        for k in range(K):
            r = case.setpoint_row(k0 + k)
            ost, oship, _ = o.frame_step(ost, oship, r[:3], r[3:], vw, ost.dpdem_floe_eligible)
            np.testing.assert_array_equal(hp[k].T, ost.positions, err_msg=f"call {c} frame {k0 + k}")
        k0 += K
        g, gs = case.to_oracle(st, ship)
        for name in ("positions", "velocities", "rotations", "angular_vels", "submergence", "bond_intact",
                     "bond_damage", "elem_active", "dpdem_floe_eligible"):
            np.testing.assert_array_equal(getattr(g, name), getattr(ost, name), err_msg=f"call {c} {name}")
        np.testing.assert_array_equal(gs.eta, oship.eta)
    after = eng.list_stats()
    return after["rebuilt_frames"] - before["rebuilt_frames"], after


def test_lists_chained_calls_build_only_when_needed():
    """A call's first frame builds the lists only when an element lies
    farther than the threshold from its build position: 40 calls of 1-17
    frames on a field the hull enters equal the oracle frame by frame, with
    fewer builds than calls."""
    case = Case.make("c1small", n_frames=400)
    sizes = [1, 7, 3, 1, 1, 17, 2, 5] * 5
    rebuilt, stats = _chained(case, sizes)
    print(f"{len(sizes)} calls, {sum(sizes)} frames, {rebuilt} list builds, stats {stats}", flush=True)
    assert stats["lists"]
    assert 1 <= rebuilt < len(sizes)


def test_lists_state_edited_between_calls():
    """The caller moves elements between calls, by more than the threshold
    (the next call builds) and by less (the lists stay valid), and flips
    elements active and frozen: every frame still equals the oracle."""
    case = Case.make("c1small", n_frames=200)
    n = case.n
    rng = np.random.default_rng(5)
    far = rng.choice(n, 6, replace=False)
    near = rng.choice(n, 12, replace=False)

    def edit(c, st):
        if c == 3:      # a far move: two elements swap places
            with torch.no_grad():
                a, b = int(far[0]), int(far[1])
                pa = st.positions[a].clone()
                st.positions[a] = st.positions[b]
                st.positions[b] = pa
            return True
        if c == 5:      # small moves, inside the threshold
            with torch.no_grad():
                idx = torch.as_tensor(near, device=st.positions.device)
                st.positions[idx] += 0.004
            return True
        if c == 7:      # activity flipped both ways, and a 3 m jump
            with torch.no_grad():
                st.elem_active[torch.as_tensor(far[:3], device=st.elem_active.device)] = True
                st.elem_active[torch.as_tensor(far[3:], device=st.elem_active.device)] = False
                st.positions[int(far[2])] += 3.0
            return True
        return False

    rebuilt, stats = _chained(case, [4, 9, 2, 5, 6, 3, 8, 6, 20, 1], edit)
    print(f"edited calls: {rebuilt} list builds, stats {stats}", flush=True)
    assert rebuilt >= 3     # the first call, and the calls after the far moves


def test_lists_overflow_chained_calls():
    """An overflowed list forces the next call's first frame to build (the
    grids path keeps its buckets filled); chained calls equal the oracle."""
    case = Case.make("ring", n_frames=200)
    rebuilt, stats = _chained(case, [3, 1, 5, 2, 9])
    print(f"ring chained: {rebuilt} list builds, stats {stats}", flush=True)
    assert stats["overflowed"]
    assert rebuilt >= 5


def test_lists_caller_moves_beside_live_elements():
    """Partial builds against the caller's edits.  Call after call the caller
    moves frozen elements next to the live ones by 0.9 of the threshold in
    random directions (builds in between rebuild some lists and keep others),
    and activates frozen elements with and without a move: every frame
    equals the oracle."""
    from scipy.spatial import cKDTree
    case = Case.make("c1small", n_frames=200)
    rng = np.random.default_rng(11)
    step = 0.9 * (DEFAULT_LIST_SKIN / 2 - 0.005)
    moved = []

    def edit(c, st):
        if c == 0:
            return False
        pos = st.positions.detach().cpu().numpy()
        act = st.elem_active.cpu().numpy().astype(bool)
        live, frozen = np.flatnonzero(act), np.flatnonzero(~act)
        if len(live) == 0 or len(frozen) == 0:
            return False
        d, _ = cKDTree(pos[live]).query(pos[frozen])
        cand = frozen[d < 20.0]
        if len(cand) < 8:
            return False
        pick = rng.choice(cand, min(30, len(cand)), replace=False)
        ang = rng.uniform(0.0, 2.0 * np.pi, len(pick))
        delta = (step * np.stack([np.cos(ang), np.sin(ang)], axis=1)).astype(np.float32)
        dev = st.positions.device
        with torch.no_grad():
            st.positions[torch.as_tensor(pick, device=dev)] += torch.as_tensor(delta, device=dev)
            if c % 3 == 2:   # activity: three moved elements, three unmoved ones
                rest = np.setdiff1d(cand, pick)[:3]
                st.elem_active[torch.as_tensor(np.concatenate([pick[:3], rest]), device=dev)] = True
        moved.append(len(pick))
        return True

    rebuilt, stats = _chained(case, [8, 10, 6, 12, 8, 10, 6, 12, 8, 10], edit)
    print(f"caller moves: {moved} elements moved per call, {rebuilt} list builds, stats {stats}", flush=True)
    assert len(moved) >= 5
