"""DPDEM fracture algebra (a13).

Known answers from the reference's ``dpdem/validation/tests.py:100-146``
(same formulas as ``jax_bond_fracture.py:29-128``): criterion below failure for
sigma = 0.3 MPa, tau = 0.5 MPa; criterion ~1 on the failure envelope;
compression gives 0; B-K energy between the modes, pure mode I -> G_c^I.
Hand-derived cases: delta_s = 0 -> G_ratio = 0 -> G_c = G_c^I; damage rate cap
0.15 per call; damage monotone; broken once D reaches 1.
"""
import numpy as np

from helpers import Case


def _o(ft=np.float64):
    return Case.make("hexagon").oracle(ft)


def test_reference_known_answers():
    for ft in (np.float64, np.float32):
        o = _o(ft)
        sigma_bar, C, mu = 0.6e6, 1.8e6, 0.7
        tb = o.mohr_coulomb(C, mu, np.array([0.3e6], ft))
        assert o.strength_criterion(np.array([0.3e6], ft), np.array([0.5e6], ft), np.array([sigma_bar], ft), tb)[0] < 1
        sf = sigma_bar * 0.8
        tbf = o.mohr_coulomb(C, mu, np.array([sf], ft))
        tf = tbf * np.sqrt(1 - (sf / sigma_bar) ** 2)
        cf = o.strength_criterion(np.array([sf], ft), tf.astype(ft), np.array([sigma_bar], ft), tbf)
        assert abs(cf[0] - 1.0) < 0.05
        cc = o.strength_criterion(np.array([-0.5e6], ft), np.array([0.0], ft), np.array([sigma_bar], ft), tb)
        assert cc[0] == 0.0
        g = o.bk_energy(np.array([6.0], ft), np.array([6.0], ft), 12.0, 12.0, 1.75)
        assert 12.0 <= g[0] <= 12.0
        g1 = o.bk_energy(np.array([1.0], ft), np.array([0.0], ft), 12.0, 12.0, 1.75)
        assert abs(g1[0] - 12.0) < 1.2
        # mixed energies with distinct modes stay between them
        gm = o.bk_energy(np.array([0.3], ft), np.array([0.7], ft), 10.0, 30.0, 1.75)
        assert 10.0 < gm[0] < 30.0
        assert abs(gm[0] - (10.0 + 20.0 * (0.7 / (1.0 + 1e-10)) ** 1.75)) < 1e-3


def test_pure_mode_one_and_rate_cap():
    o = _o()
    dn = np.array([1e-3]); ds = np.array([0.0])
    gr = o.fracture_energy_ratio(dn, ds)
    assert gr[0] == 0.0
    # strongly overloaded bond: criterion >= 1, damage initiates, cap 0.15 per step
    d, dmx, dm0, it = o.update_bond_damage_state(np.array([5e-3]), np.array([0.0]), np.array([5e7]),
                                                 np.array([0.0]), np.array([1e6]), 1.8e6, 0.7, 12.0, 12.0, 1.75,
                                                 np.array([0.0]), np.array([0.0]), np.array([0.0]),
                                                 np.array([True]))
    assert dm0[0] == dmx[0] > 0
    damages = [d[0]]
    intact = it
    for _ in range(10):
        d, dmx, dm0, intact = o.update_bond_damage_state(np.array([5e-2]), np.array([0.0]), np.array([5e7]),
                                                         np.array([0.0]), np.array([1e6]), 1.8e6, 0.7, 12.0,
                                                         12.0, 1.75, d, dmx, dm0, intact)
        damages.append(d[0])
    inc = np.diff(damages)
    assert np.all(inc >= 0) and np.all(inc <= 0.15 + 1e-12)
    assert damages[-1] >= 1.0 - 1e-12 and not intact[0]


def test_no_damage_below_stress_floor():
    o = _o()
    d, dmx, dm0, it = o.update_bond_damage_state(np.array([1e-3]), np.array([0.0]), np.array([5e3]),
                                                 np.array([0.0]), np.array([1e6]), 1.8e6, 0.7, 12.0, 12.0, 1.75,
                                                 np.array([0.2]), np.array([0.0]), np.array([0.0]),
                                                 np.array([True]))
    assert d[0] == 0.2 and it[0]


def test_damage_variable_closed_form():
    o = _o()
    D = o.damage_variable(np.array([2e-3]), np.array([1e-3]), np.array([4e-3]), np.array([2e-3]))
    expected = 4e-3 * (2e-3 - 1e-3) / (2e-3 * (4e-3 - 1e-3 + 1e-10))
    assert abs(D[0] - expected) < 1e-12
