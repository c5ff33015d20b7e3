"""GPU parity of the leaf entry points against the oracle."""
import numpy as np
import pytest
import torch

from helpers import Case, Report
from oracle import bonds as obonds

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,caps", [("c1small", None), ("ring", None), ("c1small", (1, 1))])
def test_contact_pairs_exact(name, caps):
    case = Case.make(name)
    case.bucket_caps = caps
    eng = case.engine()
    o = case.oracle()
    rng = np.random.default_rng(1)
    pos = case.scene.centroids.astype(np.float32) + rng.normal(0, 0.3, (case.n, 2)).astype(np.float32)
    gi, gj = eng.contact_pairs(torch.as_tensor(pos))
    oi, oj = o.contact_pairs(pos)
    print("pairs", len(oi), flush=True)
    assert len(oi) > 0
    np.testing.assert_array_equal(gi.cpu().numpy(), oi)
    np.testing.assert_array_equal(gj.cpu().numpy(), oj)


@pytest.mark.parametrize("name,lanes", [("c1small", 0), ("ring", 0), ("c1small", 8), ("ring", 8)])
def test_substep_forces(name, lanes):
    case = Case.make(name, force_lanes=lanes)
    eng = case.engine()
    o = case.oracle()
    st, ship = case.gpu_initial()
    rng = np.random.default_rng(2)
    n = case.n
    pos = case.scene.centroids.astype(np.float32) + rng.normal(0, 0.2, (n, 2)).astype(np.float32)
    vel = rng.normal(0, 0.5, (n, 2)).astype(np.float32)
    st.positions.copy_(torch.as_tensor(pos)); st.velocities.copy_(torch.as_tensor(vel))
    st.angular_vels.copy_(torch.as_tensor(rng.normal(0, 0.05, n).astype(np.float32)))
    st.submergence.copy_(torch.as_tensor(rng.uniform(0, 0.5, n).astype(np.float32)))
    act = rng.random(n) < 0.5
    st.elem_active.copy_(torch.as_tensor(act))
    nb = len(case.scene.bonds)
    st.bond_intact.copy_(torch.as_tensor(rng.random(nb) < 0.8))
    st.bond_damage.copy_(torch.as_tensor(rng.uniform(0, 0.5, nb).astype(np.float32)))
    shipf = rng.normal(0, 1e5, (n, 2)).astype(np.float32)
    vw = np.tile(np.array([0.2, -0.1], np.float32), (n, 1))
    f, m = eng.substep_forces(st, ship, torch.as_tensor(shipf), torch.as_tensor(vw))
    ost, _ = case.to_oracle(st, ship)
    fb, mb = o.bond_forces(ost.positions, ost.elem_active, ost.bond_intact, ost.bond_damage)
    fi, mi = o.ice_ice_forces(ost.positions, ost.velocities)
    fd = o.drag(ost.velocities, vw)
    md = o.rot_drag(ost.angular_vels, ost.submergence)
    F = fb + fi - fd + shipf
    M = mb + mi + md
    rep = Report()
    rep.add_float("forces", f.cpu().numpy(), F)
    rep.add_float("moments", m.cpu().numpy(), M)
    print(rep.summary(), rep.bits, flush=True)
    rep.assert_ok()


def test_collisions_combined():
    from icedem.geometry import compute_all_collisions_combined
    case = Case.make("c0")
    o = case.oracle()
    rng = np.random.default_rng(3)
    lp = o.local_polys
    for trial in range(6):
        sx, sy = case.scene.centroids[rng.integers(case.n)] + rng.normal(0, 10.0, 2)
        sh = rng.uniform(-np.pi, np.pi)
        pos = case.scene.centroids.astype(np.float32)
        rot = rng.uniform(-3, 3, case.n).astype(np.float32)
        g = compute_all_collisions_combined(lp, o.poly_mask, pos, rot, sx, sy, sh, 30.0, 6.0, o.hull_local,
                                            case.cfg.n_edge_samples)
        r = o.collisions_combined(pos, rot, np.float32(sx), np.float32(sy), np.float32(sh))
        g = [x.cpu().numpy() for x in g]
        print("trial", trial, "contacts", int((r[0] > 0).sum()), flush=True)
        for a, b in zip(g, r):
            np.testing.assert_array_equal(a, b)


def test_bond_damage_leaf():
    from icedem.bond_fracture import update_bond_damage_state
    case = Case.make("c0")
    o = case.oracle()
    rng = np.random.default_rng(4)
    nb = 4096
    f = np.float32
    dn = rng.normal(0, 1e-3, nb).astype(f); ds = np.abs(rng.normal(0, 1e-3, nb)).astype(f)
    sn = rng.normal(0, 2e6, nb).astype(f); stt = rng.normal(0, 2e6, nb).astype(f)
    sb = rng.uniform(5e5, 2e6, nb).astype(f)
    dmg = rng.uniform(0, 0.9, nb).astype(f); dmx = rng.uniform(0, 1e-3, nb).astype(f)
    dm0 = np.where(rng.random(nb) < 0.5, 0, rng.uniform(0, 5e-4, nb)).astype(f)
    it = rng.random(nb) < 0.9
    g = update_bond_damage_state(dn, ds, sn, stt, sb, 1.8e6, 0.7, 12.0, 12.0, 1.75, dmg, dmx, dm0, it)
    r = o.update_bond_damage_state(dn, ds, sn, stt, sb, 1.8e6, 0.7, 12.0, 12.0, 1.75, dmg, dmx, dm0, it)
    for a, b in zip(g, r):
        np.testing.assert_array_equal(a.cpu().numpy(), b)
    # and with a non-trivial Benzeggagh-Kenane branch (G_c_I != G_c_II)
    g = update_bond_damage_state(dn, ds, sn, stt, sb, 1.8e6, 0.7, 12.0, 30.0, 1.75, dmg, dmx, dm0, it)
    r = o.update_bond_damage_state(dn, ds, sn, stt, sb, 1.8e6, 0.7, 12.0, 30.0, 1.75, dmg, dmx, dm0, it)
    for a, b in zip(g, r):
        np.testing.assert_array_equal(a.cpu().numpy(), b)


def test_connected_components_and_com():
    from icedem.bonds import find_connected_components_jax, component_com_velocities_jax
    rng = np.random.default_rng(5)
    cases = [(4, [(0, 1), (1, 2), (2, 3)], None), (4, [(0, 1), (0, 2), (0, 3)], None),
             (4, [(0, 1), (2, 3)], None), (1, [], None), (3, [(0, 1), (1, 2)], [False, False]),
             (20, [(i, i + 1) for i in range(19)], None)]
    n = 3000
    e = rng.integers(0, n, (2500, 2))
    e = e[e[:, 0] != e[:, 1]]
    cases.append((n, [tuple(map(int, x)) for x in e], list(rng.random(len(e)) < 0.8)))
    for n, edges, intact in cases:
        bi = np.array([a for a, b in edges], np.int32); bj = np.array([b for a, b in edges], np.int32)
        it = np.ones(len(edges), bool) if intact is None else np.array(intact, bool)
        g = find_connected_components_jax(n, bi, bj, it).cpu().numpy()
        r = obonds.find_connected_components(n, bi, bj, it)
        np.testing.assert_array_equal(g, r)
    labels = np.array([0, 0, 2, 2], np.int32)
    masses = np.array([1.0, 1.0, 2.0, 2.0], np.float32)
    vels = np.array([[1.0, 0.0], [3.0, 0.0], [0.0, 1.0], [0.0, 3.0]], np.float32)
    v = component_com_velocities_jax(labels, masses, vels).cpu().numpy()
    np.testing.assert_allclose(v, [[2, 0], [2, 0], [0, 2], [0, 2]], atol=1e-6)
    lab = obonds.find_connected_components(n, bi, bj, it)
    m = rng.uniform(1, 100, n).astype(np.float32)
    vv = rng.normal(0, 1, (n, 2)).astype(np.float32)
    g = component_com_velocities_jax(lab, m, vv).cpu().numpy()
    r = obonds.component_com_velocities(lab, m, vv, ft=np.float32)
    np.testing.assert_array_equal(g, r)


def test_ship_step_leaf():
    case = Case.make("c0")
    eng = case.engine()
    o = case.oracle()
    rng = np.random.default_rng(6)
    for trial in range(20):
        s = np.concatenate([rng.normal(0, 50, 2), rng.uniform(-3, 3, 1), rng.normal(0, 1, 3) * [1, 0.3, 0.02],
                            rng.normal(0, 100, 6), rng.normal(0, 1, 3)]).astype(np.float32)
        sp = np.concatenate([s[:2] + rng.normal(0, 5, 2), rng.uniform(-3, 3, 1), s[:2] + rng.normal(0, 5, 2),
                             rng.uniform(-3, 3, 1)]).astype(np.float32)
        tau = rng.normal(0, 1e6, 3).astype(np.float32)
        vws = rng.normal(0, 0.3, 2).astype(np.float32)
        g = eng.ship_step(torch.as_tensor(s, device=eng.device), sp, tau, vws).cpu().numpy()
        eta, nu, nn, ei = o.ship_step(s[0:3], s[3:6], s[6:12], s[12:15], sp[:3], sp[3:], tau, vws)
        np.testing.assert_array_equal(g, np.concatenate([eta, nu, nn, ei]))
