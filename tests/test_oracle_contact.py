"""Ice-ice contact, bonds, drag and integration of the oracle (a9-a12):
closed forms and invariants (no reference test pins absolute values)."""
import numpy as np
import pytest

from helpers import Case
from oracle import step as ostep
from oracle.cells import candidate_pairs


def _two_circle_oracle(ft, hertz=1e8, p=1.0, mu=0.3, dil=0.0):
    from icedem.config import benchmark_config
    cfg = benchmark_config(hertz_exponent=p, hertz_factor=hertz, ice_ice_friction=mu, contact_dilation=dil)
    return Case.make("hexagon", cfg=cfg).oracle(ft)


def test_head_on_contact_closed_form():
    for p in (1.0, 1.5):
        o = _two_circle_oracle(np.float64, p=p)
        r = o.radii[0]
        pos = np.array([[0.0, 0.0], [2 * r - 0.1, 0.0], [100.0, 0.0]])
        vel = np.zeros((3, 2))
        f, m = o.ice_ice_forces(pos, vel)
        fn = 1e8 * 0.1 ** p
        np.testing.assert_allclose(f[0], [-fn, 0.0], rtol=1e-9)
        np.testing.assert_allclose(f[1], [fn, 0.0], rtol=1e-9)
        assert np.all(f[2] == 0)
        np.testing.assert_allclose(m, 0.0, atol=1e-6)


def test_friction_direction_quirk():
    """t_hat = -v_t/|v_t| with v_rel = v_j - v_i; i receives -f_n + f_t
    (jax_simulation.py:141-154)."""
    o = _two_circle_oracle(np.float64, mu=0.3)
    r = o.radii[0]
    pos = np.array([[0.0, 0.0], [2 * r - 0.1, 0.0], [100.0, 0.0]])
    vel = np.array([[0.0, 0.0], [0.0, 1.0], [0.0, 0.0]])
    f, m = o.ice_ice_forces(pos, vel)
    fn = 1e8 * 0.1
    ft_ = min(0.3 * fn, 1e5)
    np.testing.assert_allclose(f[0], [-fn, -ft_], rtol=1e-9)
    np.testing.assert_allclose(f[1], [fn, ft_], rtol=1e-9)


def test_dilation_extends_contact():
    o0 = _two_circle_oracle(np.float64, dil=0.0)
    o2 = _two_circle_oracle(np.float64, dil=0.2)
    r = o0.radii[0]
    pos = np.array([[0.0, 0.0], [2 * r + 0.1, 0.0], [100.0, 0.0]])
    f0, _ = o0.ice_ice_forces(pos, np.zeros((3, 2)))
    f2, _ = o2.ice_ice_forces(pos, np.zeros((3, 2)))
    assert np.all(f0 == 0)
    np.testing.assert_allclose(f2[1, 0], 1e8 * 0.3, rtol=1e-9)


@pytest.mark.parametrize("ft", [np.float32, np.float64])
def test_forces_sum_to_zero_and_cells_match_all_pairs(ft):
    case = Case.make("c1small")
    o_all = case.oracle(ft)
    rng = np.random.default_rng(0)
    n = case.n
    pos = (case.scene.centroids + rng.normal(0, 0.3, (n, 2))).astype(ft)
    vel = rng.normal(0, 0.5, (n, 2)).astype(ft)
    fi, mi = o_all.ice_ice_forces(pos, vel)
    assert abs(float(np.sum(fi[:, 0], dtype=np.float64))) < 1e-3 * np.abs(fi).max() * n ** 0.5
    assert abs(float(np.sum(fi[:, 1], dtype=np.float64))) < 1e-3 * np.abs(fi).max() * n ** 0.5
    act = rng.random(n) < 0.5
    intact = rng.random(len(case.scene.bonds)) < 0.8
    dmg = rng.uniform(0, 0.5, len(case.scene.bonds)).astype(ft)
    fb, mb = o_all.bond_forces(pos, act, intact, dmg)
    assert abs(float(np.sum(fb[:, 0], dtype=np.float64))) <= 1e-4 * np.abs(fb).max() * n ** 0.5
    # bond torque 0.5 d x F is rounding noise because F is parallel to d
    assert np.max(np.abs(mb)) <= 1e-5 * np.max(np.abs(fb)) * 10
    # the cell-list candidate enumeration gives the identical pair set and forces
    o_all.pair_mode = "cells"
    fc, mc = o_all.ice_ice_forces(pos, vel)
    np.testing.assert_array_equal(fc, fi)
    np.testing.assert_array_equal(mc, mi)
    o_all.pair_mode = "all"


def test_cell_candidates_superset_brute_force():
    rng = np.random.default_rng(1)
    for trial in range(10):
        n = 300
        pos = rng.uniform(0, 60, (n, 2))
        radii = rng.uniform(0.5, 3.0, n)
        pi, pj = candidate_pairs(pos, radii, 0.0)
        cand = set(zip(pi.tolist(), pj.tolist()))
        d = np.linalg.norm(pos[:, None] - pos[None], axis=2)
        touch = (radii[:, None] + radii[None] - d > 0) & (d > 1e-6)
        ii, jj = np.nonzero(np.triu(touch, 1))
        assert set(zip(ii.tolist(), jj.tolist())) <= cand


def test_drag_closed_forms_and_terminal_drift():
    o = Case.make("hexagon").oracle(np.float64)
    vel = np.array([[2.0, 0.0], [0.0, 0.0], [1e-8, 0.0]])
    vw = np.zeros((3, 2))
    d = o.drag(vel, vw)
    A = o.areas[0] * o.thick[0]
    np.testing.assert_allclose(d[0], [0.5 * 1025 * 0.4 * A * 4.0, 0.0], rtol=1e-12)
    np.testing.assert_allclose(d[2], [0.5 * 1025 * 0.4 * A * 0.1 * 1e-8, 0.0], rtol=1e-9)
    assert np.all(d[1] == 0)
    # an active free element in a uniform current drifts to the current speed
    st, _ = Case.make("hexagon").oracle_initial(np.float64)
    st = st._replace(elem_active=np.ones(3, bool), positions=np.array([[0.0, 0.0], [50.0, 0.0], [100.0, 0.0]]))
    vw = np.tile([0.3, 0.1], (3, 1))
    steps = 4000
    for _ in range(steps):
        f = -o.drag(st.velocities, vw)
        st = o.integrate(st, f, np.zeros(3))
    # quadratic drag: du/dt = -k u^2 with k = 0.5 rho_w c_d / (rho_ice h) -> u = u0 / (1 + k u0 t)
    k = 0.5 * 1025 * 0.4 / (o.masses[0] / (o.areas[0] * o.thick[0]) * o.thick[0])
    u0 = np.hypot(0.3, 0.1)
    u = np.linalg.norm(vw - st.velocities, axis=1)
    np.testing.assert_allclose(u, u0 / (1 + k * u0 * steps * o.dt_ice), rtol=0.02)
    # the relative velocity stays parallel to the current
    np.testing.assert_allclose(st.velocities[:, 0] / st.velocities[:, 1], 3.0, rtol=1e-9)


def test_integration_frozen_and_sanitise():
    o = Case.make("hexagon").oracle(np.float64)
    st, _ = Case.make("hexagon").oracle_initial(np.float64)
    st = st._replace(elem_active=np.array([True, False, True]))
    f = np.array([[1e6, 0.0], [1e6, 1e6], [np.nan, 0.0]])
    out = o.integrate(st, f, np.array([1e6, 1e6, 0.0]))
    assert np.all(out.positions[1] == st.positions[1]) and np.all(out.velocities[1] == 0)
    assert out.velocities[2, 0] == 0.0 and out.positions[2, 0] == st.positions[2, 0]   # NaN -> v=0, x kept
    assert out.velocities[0, 0] == min(1e6 / o.masses[0] * o.dt_ice, 20.0)
    assert out.angular_vels[0] == min(1e6 / o.inertia[0] * o.dt_ice, 0.2)
