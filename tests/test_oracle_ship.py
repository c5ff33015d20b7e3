"""Ship path of the oracle (a14) against the reference's own golden values.

Constants and closed forms restated from the reference's
``test_ship_dynamics.py:22-63`` (Fossen bis-scaled matrices, pole-placement
gains, thruster matrix, n_max), ``:261-330`` (linear Euler step, no Coriolis)
and ``:335-478`` (DP at the setpoint gives zero, integral error, heading wrap,
allocation reconstructs tau).
"""
import math

import numpy as np

from oracle import model, step

_MASS, _L, _G = 6000.0e3, 76.2, 9.81
_SQRT_G_L = math.sqrt(_G / _L)
_Tbis_inv = np.diag([1.0, 1.0, _L])
_Mbis = np.array([[1.1274, 0.0, 0.0], [0.0, 1.8902, -0.0744], [0.0, -0.0744, 0.1278]])
_Dbis = np.array([[0.0358, 0.0, 0.0], [0.0, 0.1183, -0.0124], [0.0, -0.0041, 0.0308]])
_M3_ref = _MASS * (_Tbis_inv @ _Mbis @ _Tbis_inv)
_D3_ref = _MASS * _SQRT_G_L * (_Tbis_inv @ _Dbis @ _Tbis_inv)
_wn = np.diag([0.1, 0.1, 0.2])
_zeta = np.diag([1.0 / math.sqrt(2.0)] * 3)
_Kp_ref = _wn @ _wn @ np.diag(np.diag(_M3_ref))
_Kd_ref = 2.0 * _zeta @ _wn @ np.diag(np.diag(_M3_ref)) - np.diag(np.diag(_D3_ref))
_B_ref = np.array([[0, 0, 0, 0, 1, 1], [1, 1, 1, 1, 0, 0], [30, 22, -22, -30, -8, 8]], float) @ np.diag(
    [3.2, 3.2, 3.2, 3.2, 31.2, 31.2])
_B_pinv_ref = _B_ref.T @ np.linalg.inv(_B_ref @ _B_ref.T)
_n_max_ref = np.array([250., 250., 250., 250., 160., 160.])


def test_fossen_golden_matrices():
    s = model.Fossen()
    np.testing.assert_allclose(s.M3, _M3_ref, rtol=1e-10)
    np.testing.assert_allclose(s.D3, _D3_ref, rtol=1e-10)
    np.testing.assert_allclose(s.M3 @ s.M3inv, np.eye(3), atol=1e-8)
    np.testing.assert_allclose(s.B, _B_ref, rtol=1e-10)
    np.testing.assert_array_equal(s.n_max, _n_max_ref)
    assert s.T_n == 1.0
    assert abs(3.2 * 250.0 ** 2 - 200000.0) < 1e-9
    sd = model.build_ship_dynamics_data(_MASS, _L, 10.0)
    np.testing.assert_allclose(sd.Kp, _Kp_ref, rtol=1e-12)
    np.testing.assert_allclose(sd.Kd, _Kd_ref, rtol=1e-12)
    np.testing.assert_allclose(sd.B_pinv, _B_pinv_ref, rtol=1e-12)
    assert np.max(np.abs(_B_ref @ _B_pinv_ref - np.eye(3))) < 1e-10
    assert sd.ship_hull_local.shape == (62, 2)


def _oracle(ft=np.float64, use_auto=False, dt=0.05):
    sd = model.build_ship_dynamics_data(_MASS, _L, 10.0)
    sd = sd._replace(M3inv=np.linalg.inv(_M3_ref), D3=_D3_ref, Kp=_Kp_ref, Kd=_Kd_ref)
    st = model.StaticF64(np.zeros((1, 3, 2)), np.ones((1, 3), bool), np.array([3]), np.ones(1), np.ones(1),
                         np.ones(1), np.ones(1), np.ones(1), np.zeros(0, np.int32), np.zeros(0, np.int32),
                         np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(1, np.int32), 1.0, 1.0)
    cfg = model.OracleConfig(ship_velocity_clip=(1e9, 1e9, 1e9))
    return step.Oracle(st, cfg, sd, _L, 10.0, 1, dt, dt, use_auto, _B_ref, 1.0, dtype=ft)


def test_euler_step_is_linear_no_coriolis():
    """test_ship_dynamics.py:281-309: nu' = nu + dt M^-1 (tau - D nu)."""
    o = _oracle()
    nu0 = np.array([2.0, 0.5, 0.05])
    tau0 = np.array([1e5, 0.0, 0.0])
    dt = 0.05
    # drive ship_step with a zero controller: setpoint at the pose, zero gains
    o.Kp[:] = 0; o.Kd[:] = 0; o.D3 = _D3_ref.copy()
    eta = np.zeros(3)
    out = o.ship_step(eta, nu0, np.zeros(6), np.zeros(3), eta, eta, tau0, np.zeros(2))
    nu_expected = nu0 + dt * (np.linalg.inv(_M3_ref) @ (tau0 - _D3_ref @ nu0))
    np.testing.assert_allclose(out[1], nu_expected, rtol=1e-8)
    # kinematics: heading 0, surge -> +x
    assert out[0][0] > 0 and abs(out[0][1] - dt * dt * 0) < 1.0


def test_dp_at_setpoint_zero_and_integral():
    o = _oracle()
    eta = np.array([10., 20., 0.5])
    tau, _ = o.dp_control(eta, np.zeros(3), np.zeros(3), eta.copy())
    np.testing.assert_allclose(tau, 0.0, atol=1e-6)
    eta = np.array([5., 3., 0.1])
    _, e_int = o.dp_control(eta, np.zeros(3), np.zeros(3), np.zeros(3))
    e_exp = np.array([5., 3., (0.1 + math.pi) % (2 * math.pi) - math.pi])
    np.testing.assert_allclose(e_int, 0.05 * e_exp, rtol=1e-8)
    # heading wrap across +-pi gives a small yaw torque
    tau, _ = o.dp_control(np.array([0., 0., -math.pi + 0.1]), np.zeros(3), np.zeros(3),
                          np.array([0., 0., math.pi - 0.1]))
    assert abs(tau[2]) < 1e10


def test_dp_matches_closed_form():
    """tau = -R^T Kp e - Kd nu (test_ship_dynamics.py:379-398 reference law)."""
    o = _oracle()
    eta = np.array([10., 5., math.pi / 4]); nu = np.array([1.0, 0.3, 0.02])
    tau, _ = o.dp_control(eta, nu, np.zeros(3), np.zeros(3))
    c, s = math.cos(eta[2]), math.sin(eta[2])
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    e = np.array([10., 5., math.pi / 4])
    np.testing.assert_allclose(tau, -(R.T @ _Kp_ref @ e) - _Kd_ref @ nu, rtol=1e-10)


def test_allocation_reconstructs_tau():
    """test_ship_dynamics.py:423-470: B (|n| n) with n from B+ tau returns tau."""
    for tau in (np.array([1e5, 0., 0.]), np.array([0., 5e4, 0.])):
        u = _B_pinv_ref @ tau
        n_cmd = np.clip(np.sign(u) * np.sqrt(np.abs(u) + 1e-12), -_n_max_ref, _n_max_ref)
        out = _B_ref @ (np.abs(n_cmd) * n_cmd)
        np.testing.assert_allclose(out[np.argmax(np.abs(tau))], tau[np.argmax(np.abs(tau))], rtol=1e-4)


def test_float32_ship_step_tracks_float64():
    o32 = _oracle(np.float32, use_auto=True)
    o64 = _oracle(np.float64, use_auto=True)
    rng = np.random.default_rng(0)
    for _ in range(50):
        s = np.concatenate([rng.normal(0, 50, 2), rng.uniform(-3, 3, 1), rng.normal(0, 1, 3) * [1, .3, .02],
                            rng.normal(0, 100, 6), rng.normal(0, 1, 3)])
        sp = np.concatenate([s[:2] + rng.normal(0, 5, 2), rng.uniform(-3, 3, 1)])
        spn = sp + np.array([0.1, 0.0, 0.0])
        tau = rng.normal(0, 1e6, 3)
        a = o32.ship_step(s[0:3].astype(np.float32), s[3:6].astype(np.float32), s[6:12].astype(np.float32),
                          s[12:15].astype(np.float32), sp.astype(np.float32), spn.astype(np.float32),
                          tau.astype(np.float32), np.zeros(2, np.float32))
        b = o64.ship_step(s[0:3], s[3:6], s[6:12], s[12:15], sp, spn, tau, np.zeros(2))
        np.testing.assert_allclose(a[0], b[0], rtol=1e-5, atol=1e-4)
