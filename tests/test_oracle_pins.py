"""Independent pins of the oracle's bond forces, rotational drag, contact
pre-pass, floe activation, DPDEM eligibility and fracture gates.

Each check builds a few elements by hand and compares the oracle with
numbers worked out from the reference's formulas and constants (cited per
check), plus physical facts that do not depend on how the formula is
written: a stretched bond pulls its ends together, drag opposes spin, a
rotated configuration gives rotated forces.  Every check is then run against
copies of the oracle with one planted mistake each (a flipped sign, a dropped
term, a wrong index or gate, the source text of one oracle method edited) and
must fail for every one of them: a check that passes a mutant pins nothing.

Constants from the reference: E = 5e9 Pa, rho_w = 1025 kg/m^3, flexural
strength base 0.3 MPa (properties.py:8-11); bond_stiffness_factor 0.001,
drag_coefficient_moment 0.4, pen_crit_ref 0.5 m at 1 m, w_contact factor 2 and
cap 4 m, side_push_factor 0.2, submergence rate 0.4, pen cap 0.5 m, rise time
0.5 s, effective_strength_collision 0.1, max_fracture_thickness 15 m
(config.py:58-170).
"""
import copy
import inspect
import math
import textwrap

import numpy as np
import pytest

from helpers import Case
from icedem.config import SimulationConfig
from oracle import step as ostep

F64 = np.float64


def _oracle(n, cfg=None, dtype=F64, **arrays):
    """A float64 oracle whose per-element and per-bond tables are replaced by
    the given arrays (the checked methods read only these)."""
    case = Case.make("hexagon", cfg=cfg or SimulationConfig(n_edge_samples=3))
    o = case.oracle(dtype)
    ft = o.ft
    defaults = dict(radii=np.full(n, 2.0), areas=np.full(n, 12.0), thick=np.ones(n), masses=np.full(n, 1e4),
                    inertia=np.full(n, 1e4), floe=np.zeros(n, np.int64), bond_i=np.zeros(0, np.int64),
                    bond_j=np.zeros(0, np.int64), rest=np.zeros(0), bstr=np.zeros(0), sigma_bar=np.zeros(0))
    defaults.update(arrays)
    for k, v in defaults.items():
        if k in ("floe", "bond_i", "bond_j"):
            setattr(o, k, np.asarray(v, np.int64))
        else:
            setattr(o, k, np.asarray(v, ft))
    return o


def mutant(o, method, *edits):
    """A copy of oracle `o` whose `method` is the oracle's source text with
    each (old, new) substitution applied (every `old` must occur)."""
    src = textwrap.dedent(inspect.getsource(getattr(ostep.Oracle, method)))
    for old, new in edits:
        assert old in src, (method, old)
        src = src.replace(old, new)
    ns = {}
    exec(compile(src, f"<mutant of Oracle.{method}>", "exec"), dict(vars(ostep)), ns)
    m = copy.copy(o)
    m.__class__ = type(f"Mutant_{method}", (type(o),), {method: ns[method]})
    return m


def _close(a, b, rtol=1e-9, atol=0.0):
    np.testing.assert_allclose(np.asarray(a, F64), np.asarray(b, F64), rtol=rtol, atol=atol)


# ── bond forces, jax_simulation.py:63-108 ────────────────────────────────────

def check_bond_forces(mut=None):
    # elements 0 (x=0), 1 (x=20), 2 (x=10); bond 0 = (0, 2) stretched by
    # 1 cm with damage 0.2, bond 1 = (1, 2) compressed by 2 cm, undamaged
    o = _oracle(3, thick=[0.5, 1.0, 2.0], bond_i=[0, 1], bond_j=[2, 2], rest=[9.99, 10.02],
                bstr=[1e6, 1e6], sigma_bar=[1e6, 1e6])
    if mut:
        o = mut(o)
    pos = np.array([[0.0, 0.0], [20.0, 0.0], [10.0, 0.0]])
    act = np.ones(3, bool)
    intact = np.ones(2, bool)
    dmg = np.array([0.2, 0.0])
    f, m = o.bond_forces(pos, act, intact, dmg)
    # k = E * mean(h) * 0.001: bond 0 mean h 1.25 -> 6.25e6 N/m, F = k * 0.01 * (1 - 0.2) = 5e4 N;
    # bond 1 mean h 1.5 -> 7.5e6 N/m, F = k * 0.02 = 1.5e5 N
    F0, F1 = 5e9 * 1.25 * 1e-3 * 0.01 * 0.8, 5e9 * 1.5 * 1e-3 * 0.02
    _close(F0, 5e4); _close(F1, 1.5e5)
    # the stretched bond pulls 0 towards 2 (+x), the compressed one pushes 1 away from 2 (+x)
    _close(f, [[F0, 0.0], [F1, 0.0], [-F0 - F1, 0.0]], atol=1e-9)
    assert np.max(np.abs(m)) < 1e-9 * F1
    # rotated by 30 degrees: the same magnitudes along the rotated axis
    c, s = math.cos(math.pi / 6), math.sin(math.pi / 6)
    R = np.array([[c, -s], [s, c]])
    fr, _ = o.bond_forces(pos @ R.T + np.array([3.0, -7.0]), act, intact, dmg)
    _close(fr, f @ R.T, rtol=1e-6, atol=1e-6 * F1)
    # a broken bond carries nothing
    fb, _ = o.bond_forces(pos, act, np.array([False, True]), dmg)
    _close(fb, [[0.0, 0.0], [F1, 0.0], [-F1, 0.0]], atol=1e-9)
    # mixed active/frozen ends carry nothing; both frozen carry the force
    fm_, _ = o.bond_forces(pos, np.array([True, False, True]), intact, dmg)
    _close(fm_, [[F0, 0.0], [0.0, 0.0], [-F0, 0.0]], atol=1e-9)
    ff, _ = o.bond_forces(pos, np.zeros(3, bool), intact, dmg)
    _close(ff, f, atol=1e-9)
    # fully damaged bond: no force
    fd, _ = o.bond_forces(pos, act, intact, np.array([1.0, 0.0]))
    _close(fd[0], [0.0, 0.0], atol=1e-9)


BOND_MUTANTS = [
    ("stretch sign", ("stretch = dist - self.rest", "stretch = self.rest - dist")),
    ("(1 + D)", ("(ft(1.0) - damage)", "(ft(1.0) + damage)")),
    ("h_i not mean h", ("(self.thick[bi] + self.thick[bj]) * ft(0.5)", "self.thick[bi]")),
    ("no stiffness factor", (" * ft(self.cfg.bond_stiffness_factor)", "")),
    ("broken bonds kept", ("Fx = np.where(intact, Fx, ft(0.0))", "Fx = Fx"),
     ("Fy = np.where(intact, Fy, ft(0.0))", "Fy = Fy")),
    ("frozen pairs dropped", ("ok = (both_active | both_frozen)", "ok = (both_active)")),
    ("mixed pairs kept", ("ok = (both_active | both_frozen)", "ok = (both_active | ~both_frozen)")),
    ("scatter index", ("np.add.at(fx, bj, -Fx)", "np.add.at(fx, bi, -Fx)")),
]


def test_bond_forces_pin():
    check_bond_forces()


@pytest.mark.parametrize("name,edits", [(m[0], m[1:]) for m in BOND_MUTANTS], ids=[m[0] for m in BOND_MUTANTS])
def test_bond_forces_pin_catches(name, edits):
    with pytest.raises(AssertionError):
        check_bond_forces(lambda o: mutant(o, "bond_forces", *edits))


# ── rotational drag, jax_simulation.py:195-212 ───────────────────────────────

def check_rot_drag(mut=None):
    o = _oracle(3, radii=[2.0, 2.0, 1.0], areas=[10.0, 10.0, 10.0], thick=[1.0, 1.0, 1.0])
    if mut:
        o = mut(o)
    w = np.array([0.1, -0.1, 0.1])
    # M = -0.5 * 0.4 * 1025 * r^3 * w|w| * A * min(sub/h, 1): r = 2, w = 0.1, A = 10, sub/h = 0.5 -> -82 N m
    _close(o.rot_drag(w, np.full(3, 0.5)), [-82.0, 82.0, -82.0 / 8.0])
    # a fully submerged element feels the whole area; deeper changes nothing; a dry one feels nothing
    _close(o.rot_drag(w, np.array([1.0, 3.0, 0.0])), [-164.0, 164.0, 0.0], atol=1e-12)
    # quadratic in the spin
    _close(o.rot_drag(2 * w, np.full(3, 0.5)), 4 * o.rot_drag(w, np.full(3, 0.5)))


ROT_MUTANTS = [
    ("w^2 not w|w|", ("omega * np.abs(omega)", "omega * omega")),
    ("no submergence cap", ("vmin(sub / vmax(self.thick, ft(1e-10)), ft(1.0))", "sub / vmax(self.thick, ft(1e-10))")),
    ("r^2", ("self.radii * (self.radii * self.radii)", "self.radii * self.radii")),
    ("sign", ("ft(-0.5)", "ft(0.5)")),
    ("no submergence", ("* A_sub", "* self.areas")),
]


def test_rot_drag_pin():
    check_rot_drag()


@pytest.mark.parametrize("name,edits", [(m[0], m[1:]) for m in ROT_MUTANTS], ids=[m[0] for m in ROT_MUTANTS])
def test_rot_drag_pin_catches(name, edits):
    with pytest.raises(AssertionError):
        check_rot_drag(lambda o: mutant(o, "rot_drag", *edits))


# ── contact pre-pass, jax_simulation.py:431-498 ──────────────────────────────

def _prepass_case(cap=None, mut=None):
    cfg = SimulationConfig(n_edge_samples=3, max_ship_ice_force=cap)
    # element 0: r 1.5 (w = 3 m), h 1; 1: r 3 (w = 4 m cap), h 1; 2: r 3, h 2; 3: no contact
    o = _oracle(4, cfg=cfg, radii=[1.5, 3.0, 3.0, 2.0], areas=[7.0, 28.0, 28.0, 12.0], thick=[1.0, 1.0, 2.0, 1.0])
    o.dt_ice = o.dt / 4.0          # the submergence update uses the frame dt, not the sub-step's
    if mut:
        o = mut(o)
    sx, sy = 1.0, 2.0
    pos = np.array([[sx - 5.0, sy + 2.0], [sx + 5.0, sy - 3.0], [sx + 0.5, sy], [sx + 9.0, sy + 9.0]])
    pen = np.array([0.25, 1.0, 0.8, 0.0])
    nx = np.array([1.0, 0.0, 0.6, 0.0])
    ny = np.array([0.0, 1.0, 0.8, 0.0])
    sub = np.array([0.1, 0.999, 1.9, 0.3])
    return o, pos, sub, pen, nx, ny, sx, sy


def _prepass_expected(o, pos, sub, pen, nx, ny, sx, sy):
    """The pre-pass worked out element by element from jax_simulation.py:436-498."""
    dt = o.dt
    sig1 = 0.3e6                       # sigma_f = 0.3 MPa * sqrt(h)
    fm = np.array([sig1 * 3.0 * 1.0 * min(0.25 / 0.5, 1.0),             # 4.5e5
                   sig1 * 4.0 * 1.0 * min(1.0 / 0.5, 1.0),              # 1.2e6
                   sig1 * math.sqrt(2.0) * 4.0 * 4.0 * min(0.8 / 1.0, 1.0),
                   0.0])
    side = np.array([+1.0, -1.0, 0.0, 0.0])        # sign of world y - ship y
    soi = np.stack([fm * nx, fm * ny + 0.2 * fm * side], axis=1)
    rx, ry = pos[:, 0] - sx, pos[:, 1] - sy
    M = float(np.sum(rx * -soi[:, 1] - ry * -soi[:, 0]))
    new_sub = np.array([min(0.1 + 0.4 * 0.25 * dt, 1.0), min(0.999 + 0.4 * 0.5 * dt, 1.0),
                        min(1.9 + 0.4 * 0.5 * dt, 2.0), 0.3 - (1.0 / 0.5) * 0.3 * dt])
    buoy = 1025.0 * 9.81 * o.areas * np.minimum(new_sub, o.thick) * (pen > 0)
    return fm, soi, M, new_sub, buoy


def check_prepass(mut=None):
    o, pos, sub, pen, nx, ny, sx, sy = _prepass_case(None, mut)
    hc, fm, soi_out, tau, new_sub = o.contact_prepass(pos, sub, pen, nx, ny, sx, sy)
    efm, esoi, eM, esub, ebuoy = _prepass_expected(o, pos, sub, pen, nx, ny, sx, sy)
    assert list(hc) == [True, True, True, False]
    _close(fm, efm)
    _close(new_sub, esub)
    assert new_sub[1] == 1.0                       # clipped at the thickness
    # the ship feels minus the sum (no buoyancy) and the moment of the pushes
    _close(tau[:2], -esoi.sum(axis=0))
    _close(tau[2], eM)
    # the ice feels the pushes plus buoyancy rho_w g A min(sub, h) along the normal
    _close(soi_out, esoi + np.stack([ebuoy * nx, ebuoy * ny], axis=1), atol=1e-9)


def check_prepass_thrust_cap(mut=None):
    cap = 1.0e6
    o, pos, sub, pen, nx, ny, sx, sy = _prepass_case(cap, mut)
    hc, fm, soi_out, tau, new_sub = o.contact_prepass(pos, sub, pen, nx, ny, sx, sy)
    efm, esoi, eM, esub, ebuoy = _prepass_expected(o, pos, sub, pen, nx, ny, sx, sy)
    total = esoi.sum(axis=0)
    scale = cap / float(np.hypot(*total))
    assert scale < 0.5
    # the ship feels the capped total; the force magnitudes (fracture, eligibility) are scaled too
    _close(np.hypot(tau[0], tau[1]), cap)
    _close(tau[:2], -scale * total)
    _close(fm, scale * efm)
    # quirks: the moment is formed from the unscaled pushes, and the ice keeps
    # its unscaled pushes plus buoyancy (jax_simulation.py:467-496)
    _close(tau[2], eM)
    _close(soi_out, esoi + np.stack([ebuoy * nx, ebuoy * ny], axis=1), atol=1e-9)


PREPASS_MUTANTS = [
    ("sigma_f without sqrt(h)", ("ft(cfg.flexural_strength_base) * np.sqrt(self.thick)", "ft(cfg.flexural_strength_base)")),
    ("pen_crit not scaled by h", ("(self.thick / ft(cfg.pen_crit_thickness_ref))", "ft(1.0)")),
    ("no width cap", ("vmin(ft(cfg.w_contact_radius_factor) * self.radii, ft(cfg.w_contact_cap))",
                      "ft(cfg.w_contact_radius_factor) * self.radii")),
    ("no saturation", ("vmin(pen / vmax(pen_crit, ft(1e-10)), ft(1.0))", "(pen / vmax(pen_crit, ft(1e-10)))")),
    ("side push from x", ("rel_y = pos[:, 1] - sy", "rel_y = pos[:, 0] - sx")),
    ("side push sign", ("soi_y = soi_y + side_push", "soi_y = soi_y - side_push")),
    ("moment sign", ("r_x * (-soi_y) - r_y * (-soi_x)", "r_x * soi_y - r_y * soi_x")),
    ("moment from scaled pushes", ("r_x * (-soi_y) - r_y * (-soi_x)", "r_x * (-soi_sy) - r_y * (-soi_sx)")),
    ("ice gets scaled pushes", ("soi_x = soi_x + np.where", "soi_x = soi_sx + np.where"),
     ("soi_y = soi_y + np.where", "soi_y = soi_sy + np.where")),
    ("force magnitude unscaled", ("fm = fm * scale", "fm = fm")),
    ("ship feels buoyancy", ("F_ship_x = -tree_sum(soi_sx, ft)",
                             "F_ship_x = -tree_sum(soi_sx + ft(WATER_DENSITY * 9.81) * self.areas * sub * nx, ft)")),
    ("submergence with dt_ice", ("pen_c * ft(self.dt)", "pen_c * ft(self.dt_ice)")),
    ("no pen cap", ("vmin(pen, ft(cfg.submergence_pen_cap))", "pen")),
    ("no rise", ("sub - ft(1.0 / cfg.rise_time_const) * sub * ft(self.dt)", "sub")),
    ("buoyancy from old submergence", ("z_sub = vmin(new_sub, self.thick)", "z_sub = vmin(sub, self.thick)")),
]


def test_prepass_pins():
    check_prepass()
    check_prepass_thrust_cap()


@pytest.mark.parametrize("name,edits", [(m[0], m[1:]) for m in PREPASS_MUTANTS], ids=[m[0] for m in PREPASS_MUTANTS])
def test_prepass_pins_catch(name, edits):
    mut = lambda o: mutant(o, "contact_prepass", *edits)
    failed = 0
    for chk in (check_prepass, check_prepass_thrust_cap):
        try:
            chk(mut)
        except AssertionError:
            failed += 1
    assert failed > 0, f"mutant '{name}' passed both pre-pass checks"


# ── activation and eligibility, jax_simulation.py:500-524 ────────────────────

def check_activation_eligibility(mut_act=None, mut_elig=None):
    # floes: {0, 1} -> 0, {2} -> 1, {3} -> 2; bond strengths: the weakest bond
    # (5e5 N) lies in floe 2, so the threshold is 0.1 * 5e5 = 5e4 N everywhere
    o = _oracle(4, floe=[0, 0, 1, 2], bond_i=[0, 2, 3], bond_j=[1, 3, 3], rest=[1.0, 1.0, 1.0],
                bstr=[2e6, 8e6, 5e5], sigma_bar=[1e6] * 3)
    if mut_act:
        o = mut_act(o)
    if mut_elig:
        o = mut_elig(o)
    active = np.array([False, False, False, True])
    hc = np.array([False, True, False, False])
    new_active, collided = o.activation(active, hc, 3)
    assert list(new_active) == [True, True, False, True]      # floe 0 wholly active, 3 stays active
    assert list(collided) == [True, False, False]
    elig = np.array([False, False, True])
    # 7.2e5 N on floe 0 exceeds 5e4 (but not 0.1 * the strongest bond, 8e5)
    assert list(o.eligibility(elig, collided, np.array([0.0, 7.2e5, 0.0, 0.0]))) == [True, False, True]
    assert list(o.eligibility(elig, collided, np.array([0.0, 1.0e5, 0.0, 0.0]))) == [True, False, True]
    # below the threshold: active but not eligible
    assert list(o.eligibility(elig, collided, np.array([0.0, 4.9e4, 0.0, 0.0]))) == [False, False, True]
    # a large force on a floe that had no contact does not make it eligible
    assert list(o.eligibility(np.zeros(3, bool), collided, np.array([0.0, 0.0, 9e5, 0.0]))) == [False, False, False]


ACT_MUTANTS = [
    ("element-wise activation", ("active | collided[self.floe]", "active | hc")),
    ("activation replaces", ("active | collided[self.floe]", "collided[self.floe]")),
]
ELIG_MUTANTS = [
    ("max bond strength", ("self.bstr.min()", "self.bstr.max()")),
    ("no collided gate", ("elig | (collided & (mx > thr))", "elig | (mx > thr)")),
    ("no 0.1 factor", (" * ft(self.cfg.effective_strength_collision)", "")),
    ("eligibility not kept", ("return elig | (collided & (mx > thr))", "return collided & (mx > thr)")),
]


def test_activation_eligibility_pin():
    check_activation_eligibility()


@pytest.mark.parametrize("name,edits", [(m[0], m[1:]) for m in ACT_MUTANTS], ids=[m[0] for m in ACT_MUTANTS])
def test_activation_pin_catches(name, edits):
    with pytest.raises(AssertionError):
        check_activation_eligibility(mut_act=lambda o: mutant(o, "activation", *edits))


@pytest.mark.parametrize("name,edits", [(m[0], m[1:]) for m in ELIG_MUTANTS], ids=[m[0] for m in ELIG_MUTANTS])
def test_eligibility_pin_catches(name, edits):
    with pytest.raises(AssertionError):
        check_activation_eligibility(mut_elig=lambda o: mutant(o, "eligibility", *edits))


# ── fracture gates, jax_simulation.py:563-640 ────────────────────────────────

K_N = 5e9 * 0.7 / (1.3 * 0.4)            # E (1 - nu) / ((1 + nu)(1 - 2 nu)), nu = 0.3 (jax_ice_state.py)
K_S = K_N * 0.4 / (2 * 0.7)              # k_n (1 - 2 nu) / (2 (1 - nu))


def _fracture_case(mut=None, thick3=1.0):
    """A chain 0-1-2-3 with bond (0,1) broken (so bond (1,2) is at the crack
    front), bond (2,3) next to the ship (element 3 has a hull contact), and a
    separate pair 4-5 far from both; bonds 1, 2 and 3 stretched by 1 cm;
    elements 2 and 5 slide sideways at 1 m/s."""
    o = _oracle(6, thick=[1.0, 1.0, 1.0, thick3, 1.0, 1.0], floe=[0, 0, 0, 1, 2, 2],
                bond_i=[0, 1, 2, 4], bond_j=[1, 2, 3, 5], rest=[2.0, 2.0, 2.0, 2.0],
                bstr=[1e6] * 4, sigma_bar=[1e6] * 4)
    o.k_n, o.k_s = K_N, K_S
    o.dt_ice = o.dt / 3.0          # delta_s grows with the sub-step's dt
    if mut:
        o = mut(o)
    pos = np.array([[0.0, 0.0], [2.0, 0.0], [4.01, 0.0], [6.02, 0.0], [0.0, 50.0], [2.01, 50.0]])
    vel = np.zeros((6, 2))
    vel[2, 1] = 1.0
    vel[5, 1] = 1.0
    nb = 4
    st = ostep.IceState(pos, vel, np.zeros(6), np.zeros(6), np.zeros(6), np.array([False, True, True, True]),
                        np.zeros(nb), np.zeros(nb), np.zeros(nb), np.zeros(nb), np.ones(6, bool), np.ones(3, bool))
    return o, st


def check_fracture_gates(mut=None):
    o, st = _fracture_case(mut)
    fm = np.array([0.0, 0.0, 0.0, 5e4, 0.0, 0.0])       # below 0.1 * F_b = 1e5: no collision break
    out = o._fracture(st, fm, np.ones(6, bool), np.ones(3, bool))
    dt_i = o.dt_ice
    # delta_s grows by |v_t| dt_ice only on the bond next to the ship
    _close(out.bond_delta_s, [0.0, 0.0, 1.0 * dt_i, 0.0], atol=1e-15)
    # crack-front bond (1,2): delta_n enters the stresses as 0, so no damage
    # initiates, but delta_m_max records the raw 1 cm stretch
    assert out.bond_delta_m_0[1] == 0.0 and out.bond_damage[1] == 0.0
    _close(out.bond_delta_m_max[1], 0.01, rtol=1e-6)
    # bond (2,3) next to the ship: sigma_n = k_n 0.01/2 >> 1 MPa initiates damage at
    # delta_m = sqrt(0.01^2 + delta_s^2); the critical deformation 2 G_c / sigma_m0
    # (G_c = 12 J/m^2, sigma_m0 ~ 6e7 Pa) is far below delta_m_0, so D clips to 1
    # and the step's damage is the +0.15 per-step cap (jax_bond_fracture.py:24, 186-205)
    dm = math.sqrt(0.01 ** 2 + dt_i ** 2)
    _close(out.bond_delta_m_0[2], dm, rtol=1e-6)
    _close(out.bond_delta_m_max[2], dm, rtol=1e-6)
    sm0 = math.hypot(K_N * 0.01 / 2.0, K_S * dt_i / 2.0)
    assert 2.0 * 12.0 / sm0 < dm
    _close(out.bond_damage[2], 0.15)
    # the far pair is neither near the ship nor at the crack: untouched
    assert out.bond_delta_m_max[3] == 0.0 and out.bond_delta_s[3] == 0.0
    assert list(out.bond_intact) == [False, True, True, True]
    # collision break: max(force_mag) > 0.1 F_b breaks an intact bond regardless of eligibility
    for f3, broken in ((1.0001e5, True), (0.9999e5, False)):
        fm2 = fm.copy()
        fm2[3] = f3
        o2 = o._fracture(st, fm2, np.ones(6, bool), np.array([True, False, True]))
        assert bool(~o2.bond_intact[2]) == broken, f3
        # floe 1 not eligible: no damage-state update on bond (2,3)
        assert o2.bond_delta_s[2] == 0.0 and o2.bond_delta_m_0[2] == 0.0
    # an unbreakable bond (one end thicker than 15 m) never breaks
    o3, st3 = _fracture_case(mut, thick3=16.0)
    fm3 = fm.copy()
    fm3[3] = 1e7
    out3 = o3._fracture(st3, fm3, np.ones(6, bool), np.ones(3, bool))
    assert out3.bond_intact[2] and out3.bond_delta_s[2] == 0.0
    # an element not yet active keeps its bonds out of the damage update
    act = np.ones(6, bool)
    act[3] = False
    out4 = o._fracture(st, fm, act, np.ones(3, bool))
    assert out4.bond_delta_s[2] == 0.0 and out4.bond_delta_m_0[2] == 0.0


FRACTURE_MUTANTS = [
    ("delta_s at the crack front", ("gate = near & both", "gate = (near | at_crack) & both")),
    ("delta_n not zeroed", ("dn_d = np.where(near, delta_n, ft(0.0))", "dn_d = delta_n")),
    ("delta_m from zeroed delta_n", ("delta_n, new_ds, sigma_n, sigma_t, self.sigma_bar",
                                     "dn_d, new_ds, sigma_n, sigma_t, self.sigma_bar")),
    ("break at F_b", ("eff = self.bstr * ft(cfg.effective_strength_collision)", "eff = self.bstr")),
    ("unbreakable bonds break", ("& (cf > eff) & breakable", "& (cf > eff)")),
    ("no floe gate", ("dpdem_ok = both & floe_ok &", "dpdem_ok = both &")),
    ("crack from intact bonds", ("broken = (~st.bond_intact)", "broken = (st.bond_intact)")),
    ("no activity gate", ("gate = near & both & floe_ok & breakable", "gate = near & floe_ok & breakable"),
     ("dpdem_ok = both & floe_ok", "dpdem_ok = floe_ok")),
    ("delta_s with dt frame", ("self.norm2(vtx, vty) * ft(self.dt_ice)", "self.norm2(vtx, vty) * ft(self.dt)")),
]


def test_fracture_gates_pin():
    check_fracture_gates()


@pytest.mark.parametrize("name,edits", [(m[0], m[1:]) for m in FRACTURE_MUTANTS], ids=[m[0] for m in FRACTURE_MUTANTS])
def test_fracture_gates_pin_catches(name, edits):
    with pytest.raises(AssertionError):
        check_fracture_gates(lambda o: mutant(o, "_fracture", *edits))


@pytest.mark.parametrize("dtype", [np.float32, F64])
def test_pins_hold_in_both_precisions(dtype):
    """The float32 oracle (the GPU's comparison target) satisfies the sign and
    gate facts above too (magnitudes to float32 rounding)."""
    o = _oracle(3, dtype=dtype, thick=[0.5, 1.0, 2.0], bond_i=[0, 1], bond_j=[2, 2], rest=[9.99, 10.02],
                bstr=[1e6, 1e6], sigma_bar=[1e6, 1e6])
    ft = o.ft
    f, _ = o.bond_forces(np.array([[0.0, 0.0], [20.0, 0.0], [10.0, 0.0]], ft), np.ones(3, bool),
                         np.ones(2, bool), np.array([0.2, 0.0], ft))
    np.testing.assert_allclose(f[:, 0], [5e4, 1.5e5, -2e5], rtol=2e-3)
    r = o.rot_drag(np.array([0.1, -0.1, 0.1], ft), np.full(3, 0.5, ft))
    assert r[0] < 0 < r[1]
