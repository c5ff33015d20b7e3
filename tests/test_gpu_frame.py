"""GPU parity of the full frame step against the float32 oracle.

Gate (per frame, from identical input state): bond_intact, elem_active,
dpdem_floe_eligible and has_collision flags bit-exact; every float state
within 1e-4 relative (floor 1e-3 of the field's max magnitude).  The kernels
mirror the oracle's operation order, so the floats are expected to be
bit-identical as well; the report records the bit-identical fraction.
"""
import numpy as np
import pytest
import torch

from helpers import Case, Report, compare_frame

pytestmark = pytest.mark.gpu


def _run(case, frames, check_every=1, start_check=0):
    st, ship = case.gpu_initial()
    worst = Report()
    for k in range(frames):
        if k >= start_check and (k - start_check) % check_every == 0:
            rep = compare_frame(case, st, ship, k)
            print(f"[{case.name}] frame {k}: {rep.summary()}", flush=True)
            rep.assert_ok()
        st, ship = case.gpu_frame(st, ship, k)
    return st, ship


def test_frame_parity_hexagon():
    case = Case.make("hexagon", n_frames=60)
    _run(case, 60)


@pytest.mark.parametrize("lanes", [0, 8])
def test_frame_parity_c0_contact_window(lanes):
    case = Case.make("c0", n_frames=1100, force_lanes=lanes)
    _run(case, 1100, check_every=5, start_check=700)


def test_frame_parity_c0_current_dilation():
    from icedem.config import benchmark_config
    cfg = benchmark_config(contact_dilation=0.2)
    case = Case.make("c0", n_frames=900, cfg=cfg, current=(0.3, np.pi / 6))
    _run(case, 900, check_every=10, start_check=700)


@pytest.mark.parametrize("cap", [None, 2e4])
def test_frame_parity_c0_thrust_cap(cap):
    """The ship-force totals without a thrust cap, and with a cap small
    enough that the scale is below 1 in contact (the scaled partials, the
    scaled force magnitudes of the fracture pass and floe eligibility)."""
    from icedem.config import benchmark_config
    cfg = benchmark_config(max_ship_ice_force=cap)
    case = Case.make("c0", n_frames=1010, cfg=cfg)
    st, ship = _run(case, 1000, check_every=10, start_check=700)
    if cap is not None:
        # the cap is active: the ship's reaction is the capped total
        eng = case.engine()
        peak = 0.0
        for k in range(1000, 1010):
            diag = eng.alloc_diag()
            st, ship = case.gpu_frame(st, ship, k, diag=diag)
            peak = max(peak, float(torch.linalg.vector_norm(diag["F_ice_on_ship"])))
        print("peak ship reaction", peak, "cap", cap, flush=True)
        assert 0.5 * cap < peak <= cap * (1 + 1e-4)


@pytest.mark.parametrize("lanes", [0, 8])
def test_frame_parity_field_bucket_overflow(lanes):
    """Buckets of capacity 1: most insertions spill into the overflow list,
    which every element scans; the contact set and sums stay the oracle's."""
    case = Case.make("c1small", n_frames=120, force_lanes=lanes)
    case.bucket_caps = (1, 1)
    _run(case, 120, check_every=4)


@pytest.mark.parametrize("factor,lanes", [(0.5, 0), (0.36, 0), (0.36, 8)])
def test_frame_parity_c0_ice_substeps(factor, lanes):
    """Ice sub-steps (the Rayleigh time-step cap: 3 and 4 sub-steps per
    frame): cell rebuilds between sub-steps, the kept ship forces, the
    next-state buffers' parity."""
    from icedem.config import benchmark_config
    cfg = benchmark_config(use_rayleigh_dt_cap=True, rayleigh_safety_factor=factor)
    case = Case.make("c0", n_frames=900, cfg=cfg, force_lanes=lanes)
    assert case.engine().params.n_ice_steps == (3 if factor == 0.5 else 4)
    _run(case, 900, check_every=10, start_check=700)


@pytest.mark.parametrize("lanes,cell", [(0, 1.0), (8, 1.0), (0, 0.55), (8, 0.55)])
def test_frame_parity_field(lanes, cell):
    """cell 0.55: reach boxes of up to 5x5 fine cells, scan lists longer
    than one round of the force pass's lanes"""
    case = Case.make("c1small", n_frames=300, force_lanes=lanes, cell_factor=cell)
    _run(case, 300, check_every=3)


def test_scan_bit_identical_to_oracle_trajectory():
    """200 frames in one graph-replayed segment vs an independent oracle run."""
    case = Case.make("c0", n_frames=1000)
    eng = case.engine()
    o = case.oracle()
    ost, oship = case.oracle_initial()
    st, ship = case.gpu_initial()
    K0 = 700
    sp = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(K0)]), device=eng.device)
    eng.run_frames(st, ship, sp, use_graph=True)
    vw = np.zeros((case.n, 2), np.float32)
    for k in range(K0):
        r = case.setpoint_row(k)
        ost, oship, _ = o.frame_step(ost, oship, r[:3], r[3:], vw, ost.dpdem_floe_eligible)
    g, gs = case.to_oracle(st, ship)
    rep = Report()
    rep.add_float("positions", g.positions, ost.positions)
    rep.add_float("eta", gs.eta, oship.eta)
    rep.add_exact("bond_intact", g.bond_intact, ost.bond_intact)
    print("after", K0, "frames:", rep.summary(), flush=True)
    rep.assert_ok()
    K = 200
    sp = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(K0, K0 + K)]), device=eng.device)
    hist = eng.alloc_history(K)
    eng.run_frames(st, ship, sp, hist=hist, use_graph=True)
    for k in range(K0, K0 + K):
        r = case.setpoint_row(k)
        ost, oship, od = o.frame_step(ost, oship, r[:3], r[3:], vw, ost.dpdem_floe_eligible)
    g, gs = case.to_oracle(st, ship)
    rep = Report()
    rep.add_float("positions", g.positions, ost.positions)
    rep.add_float("velocities", g.velocities, ost.velocities)
    rep.add_float("eta", gs.eta, oship.eta)
    rep.add_exact("bond_intact", g.bond_intact, ost.bond_intact)
    rep.add_exact("hist_bond_last", hist["bond_intact"][-1].cpu().numpy(), ost.bond_intact)
    rep.add_float("hist_pos_last", hist["positions"][-1].cpu().numpy().T, ost.positions)
    print("after", K0 + K, "frames:", rep.summary(), flush=True)
    rep.assert_ok()


def test_ice_substeps_graph_equals_launches():
    """With 3 ice sub-steps per frame the K-frame graph (cell rebuilds
    between sub-steps captured with the frame kernels) gives the same state
    and histories, bit for bit, as per-kernel launches and the oracle."""
    from icedem.config import benchmark_config
    cfg = benchmark_config(use_rayleigh_dt_cap=True, rayleigh_safety_factor=0.5)
    case = Case.make("c0", n_frames=900, cfg=cfg)
    eng = case.engine()
    assert eng.params.n_ice_steps == 3
    st, ship = case.gpu_initial()
    K0 = 760
    sp = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(K0)]), device=eng.device)
    eng.run_frames(st, ship, sp, use_graph=False)
    K = 60
    sp = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(K0, K0 + K)]), device=eng.device)
    out = []
    for use_graph in (True, False, True):
        s2, sh2 = eng.clone_state(st), ship.clone()
        hist = eng.alloc_history(K)
        eng.run_frames(s2, sh2, sp, hist=hist, use_graph=use_graph)
        out.append((case.to_oracle(s2, sh2), {k: v.cpu().numpy() for k, v in hist.items()}))
    (g, gs), gh = out[0]
    for (a, as_), ah in out[1:]:
        np.testing.assert_array_equal(a.positions, g.positions)
        np.testing.assert_array_equal(a.velocities, g.velocities)
        np.testing.assert_array_equal(a.bond_intact, g.bond_intact)
        np.testing.assert_array_equal(as_.eta, gs.eta)
        for k in gh:
            np.testing.assert_array_equal(ah[k], gh[k], err_msg=k)
    o = case.oracle()
    ost, oship = case.to_oracle(st, ship)
    vw = np.zeros((case.n, 2), np.float32)
    for k in range(K0, K0 + K):
        r = case.setpoint_row(k)
        ost, oship, _ = o.frame_step(ost, oship, r[:3], r[3:], vw, ost.dpdem_floe_eligible)
    rep = Report()
    rep.add_float("positions", g.positions, ost.positions)
    rep.add_float("eta", gs.eta, oship.eta)
    rep.add_exact("bond_intact", g.bond_intact, ost.bond_intact)
    print("sub-step graph vs oracle after", K0 + K, "frames:", rep.summary(), flush=True)
    rep.assert_ok()


def test_profiled_frames_equal_graph_replays():
    """icedem_frames_profiled (a start/stop event pair per dispatch) leaves
    the same state, bit for bit, as the captured graph of the same frames,
    and reports positive kernel times."""
    import ctypes as C
    from icedem import _lib
    from icedem.engine import _ptr
    case = Case.make("c0", n_frames=900)
    eng = case.engine()
    st, ship = case.gpu_initial()
    K0, K = 700, 60
    sp = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(K0)]), device=eng.device)
    eng.run_frames(st, ship, sp, use_graph=True)
    sp = torch.as_tensor(np.stack([case.setpoint_row(k) for k in range(K0, K0 + K)]), device=eng.device)
    a, ash = eng.clone_state(st), ship.clone()
    eng.run_frames(a, ash, sp, use_graph=True)
    b, bsh = eng.clone_state(st), ship.clone()
    strm = torch.cuda.current_stream().cuda_stream
    s = eng.state_struct(b, bsh)
    inp = _lib.IceDemInputs()
    inp.d_setpoints = _ptr(sp)
    ms = (C.c_float * 4)()
    torch.cuda.synchronize()
    _lib.check(eng.lib.icedem_frames_profiled(eng.ctx, C.byref(s), C.byref(inp), K, None, ms, strm),
               eng.ctx, "profiled")
    torch.cuda.synchronize()
    (ga, gas), (gb, gbs) = case.to_oracle(a, ash), case.to_oracle(b, bsh)
    np.testing.assert_array_equal(gb.positions, ga.positions)
    np.testing.assert_array_equal(gb.velocities, ga.velocities)
    np.testing.assert_array_equal(gb.rotations, ga.rotations)
    np.testing.assert_array_equal(gb.bond_intact, ga.bond_intact)
    np.testing.assert_array_equal(gb.elem_active, ga.elem_active)
    np.testing.assert_array_equal(gbs.eta, gas.eta)
    np.testing.assert_array_equal(gbs.nu, gas.nu)
    print("profiled kernel ms:", list(ms), flush=True)
    assert ms[1] > 0.0 and ms[2] > 0.0
