"""Oracle: configuration, constants and one-time builders (binary64 NumPy).

TEST INFRASTRUCTURE ONLY (the checker; see ``oracle/__init__.py``).

Restates ``medea_ice_ship/config.py:46-227`` (the fields the frame step reads),
``properties.py:6-43``, ``auto_icenav_ship.py:100-160`` (FossenSupply3DOF) and
``jax_ice_state.py:121-355`` (builders).  Pinned by the golden constants of the
reference's ``test_ship_dynamics.py:22-63`` (``tests/test_oracle_ship.py``) and
by closed forms in ``tests/test_oracle_static.py``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import NamedTuple, Optional, Tuple

import numpy as np

# properties.py:8-11 (class constants; density and strength are exposed as
# parameters here because the benchmark configurations change them)
ICE_DENSITY = 920.0
WATER_DENSITY = 1025.0
FLEXURAL_STRENGTH_BASE = 0.3e6
YOUNGS_MODULUS = 5e9


@dataclass(frozen=True)
class OracleConfig:
    """Fields of ``SimulationConfig`` (config.py:46-227) read by the step."""
    use_rayleigh_dt_cap: bool = False
    rayleigh_safety_factor: float = 0.1
    pen_crit_ref: float = 0.5
    pen_crit_thickness_ref: float = 1.0
    max_fracture_thickness: float = 15.0
    n_edge_samples: int = 4
    w_contact_radius_factor: float = 2.0
    w_contact_cap: float = 4.0
    side_push_factor: float = 0.2
    submergence_rate: float = 0.4
    submergence_pen_cap: float = 0.5
    rise_time_const: float = 0.5
    max_bond_width: float = 4.0
    bond_stiffness_factor: float = 0.001
    poisson_ratio: float = 0.3
    bond_cohesion: float = 1.8e6
    bond_internal_friction: float = 0.7
    bond_G_c_I: float = 12.0
    bond_G_c_II: float = 12.0
    bond_eta_bk: float = 1.75
    ice_ice_friction: float = 0.1
    hertz_factor: float = 2e5
    hertz_exponent: float = 1.5
    friction_cap: float = 1e5
    c_d_drag: float = 0.4
    drag_ref_thickness: float = 1.0
    linear_drag_factor: float = 0.1
    drag_coefficient_moment: float = 0.4
    velocity_clip: Tuple[float, float] = (-20.0, 20.0)
    effective_strength_collision: float = 0.1
    angular_damping: float = 0.85
    angular_cap: Tuple[float, float] = (-0.2, 0.2)
    ship_velocity_clip: Tuple[float, float, float] = (15.0, 8.0, 0.15)
    max_ship_ice_force: Optional[float] = 25e6
    ice_brake_threshold: float = 0.0
    position_clip_m: float = 1e4
    # material constants that the reference keeps as class constants
    ice_density: float = ICE_DENSITY
    flexural_strength_base: float = FLEXURAL_STRENGTH_BASE
    # extension (not in the reference): contact dilation radius added to each
    # element's contact radius, dpdem/contact/detection.py:236-237 semantics.
    contact_dilation: float = 0.0


class StaticF64(NamedTuple):
    """IceStaticData (jax_ice_state.py:59-96) in binary64, without the
    all-pairs list (the oracle enumerates pairs itself)."""
    local_polys_padded: np.ndarray
    poly_mask: np.ndarray
    poly_n_verts: np.ndarray
    masses: np.ndarray
    inertia: np.ndarray
    radii: np.ndarray
    areas: np.ndarray
    thicknesses: np.ndarray
    bond_i: np.ndarray
    bond_j: np.ndarray
    rest_lengths: np.ndarray
    bond_strengths: np.ndarray
    bond_areas: np.ndarray
    sigma_bar_bond: np.ndarray
    initial_floe_of: np.ndarray
    k_n: float
    k_s: float


class ShipDynF64(NamedTuple):
    """ShipDynamicsData (jax_ice_state.py:99-114) in binary64."""
    M3inv: np.ndarray
    D3: np.ndarray
    mass: float
    Izz: float
    Kp: np.ndarray
    Kd: np.ndarray
    Ki: np.ndarray
    B_pinv: np.ndarray
    n_max: np.ndarray
    ship_hull_local: np.ndarray


def polygon_area(verts):
    """Shoelace, geometry.py:6-14 (sequential accumulation)."""
    n = len(verts)
    area = 0
    for i in range(n):
        j = (i + 1) % n
        area += verts[i, 0] * verts[j, 1]
        area -= verts[j, 0] * verts[i, 1]
    return abs(area) / 2


def polygon_centroid(verts):
    """Vertex mean, geometry.py:17-18 (not the area centroid)."""
    return np.mean(verts, axis=0)


def breaking_force(thickness, width, base):
    """properties.py:13-22: sigma_f(h) * w * h^2, sigma_f = base*sqrt(h/1)."""
    sigma_f = base * np.sqrt(thickness / 1.0)
    return sigma_f * width * thickness ** 2


def build_static_data(polygons, centroids, bonds, thicknesses, parent_ids, cfg: OracleConfig) -> StaticF64:
    """jax_ice_state.py:150-243 (binary64)."""
    n = len(polygons)
    thicknesses = np.asarray(thicknesses, dtype=np.float64)
    centroids = np.asarray(centroids, dtype=np.float64)
    areas = np.array([polygon_area(p) for p in polygons], dtype=np.float64)
    radii = np.sqrt(areas / np.pi)
    masses = areas * thicknesses * cfg.ice_density
    inertia = np.maximum(0.5 * masses * radii ** 2, 1e-10)
    local_polys = [(np.asarray(p, np.float64) - polygon_centroid(np.asarray(p, np.float64))).astype(np.float64)
                   for p in polygons]
    n_verts = np.array([len(p) for p in local_polys], dtype=np.int32)
    max_V = int(n_verts.max())
    padded = np.zeros((n, max_V, 2), dtype=np.float64)
    mask = np.zeros((n, max_V), dtype=bool)
    for i, poly in enumerate(local_polys):
        nv = len(poly)
        padded[i, :nv] = poly
        padded[i, nv:] = poly[-1]
        mask[i, :nv] = True
    nb = len(bonds)
    if nb > 0:
        bi = np.array([b[0] for b in bonds], dtype=np.int32)
        bj = np.array([b[1] for b in bonds], dtype=np.int32)
        rest = np.array([np.linalg.norm(centroids[i] - centroids[j]) for i, j in bonds], dtype=np.float64)
        bs = np.zeros(nb, dtype=np.float64)
        ba = np.zeros(nb, dtype=np.float64)
        for b, (i, j) in enumerate(bonds):
            h = min(thicknesses[i], thicknesses[j])
            w = min((radii[i] + radii[j]) / 2, cfg.max_bond_width)
            bs[b] = breaking_force(h, w, cfg.flexural_strength_base)
            ba[b] = w * h
        sigma_bar = np.maximum(bs / np.maximum(ba, 1e-10), 1e6)
    else:
        bi = np.zeros(0, np.int32)
        bj = np.zeros(0, np.int32)
        rest = np.zeros(0)
        bs = np.zeros(0)
        ba = np.zeros(0)
        sigma_bar = np.zeros(0)
    nu_r = cfg.poisson_ratio
    E = YOUNGS_MODULUS
    prefactor = E * (1.0 - nu_r) / ((1.0 + nu_r) * (1.0 - 2.0 * nu_r))
    k_n = prefactor
    k_s = prefactor * (1.0 - 2.0 * nu_r) / (2.0 * (1.0 - nu_r))
    if parent_ids is not None:
        floe = np.asarray(parent_ids, dtype=np.int32)
    else:
        floe = np.arange(n, dtype=np.int32)
    return StaticF64(padded, mask, n_verts, masses, inertia, radii, areas, thicknesses,
                     bi, bj, rest, bs, ba, sigma_bar, floe, float(k_n), float(k_s))


class Fossen:
    """FossenSupply3DOF constants, auto_icenav_ship.py:100-160."""

    def __init__(self, mass=6000.0e3, length=76.2):
        g = 9.81
        self.mass = mass
        self.length = length
        self.T_n = 1.0
        self.n_max = np.array([250, 250, 250, 250, 160, 160], dtype=float)
        K = np.diag([3.2, 3.2, 3.2, 3.2, 31.2, 31.2])
        T_cfg = np.array([[0, 0, 0, 0, 1, 1], [1, 1, 1, 1, 0, 0], [30, 22, -22, -30, -8, 8]], dtype=float)
        self.B = T_cfg @ K
        Tbis_inv = np.diag([1.0, 1.0, self.length])
        Mbis = np.array([[1.1274, 0.0, 0.0], [0.0, 1.8902, -0.0744], [0.0, -0.0744, 0.1278]], dtype=float)
        Dbis = np.array([[0.0358, 0.0, 0.0], [0.0, 0.1183, -0.0124], [0.0, -0.0041, 0.0308]], dtype=float)
        self.M3 = self.mass * (Tbis_inv @ Mbis @ Tbis_inv)
        self.M3inv = np.linalg.inv(self.M3)
        self.D3 = self.mass * math.sqrt(g / self.length) * (Tbis_inv @ Dbis @ Tbis_inv)
        self.wn = np.diag([0.1, 0.1, 0.2])
        self.zeta = np.diag([1.0 / math.sqrt(2.0)] * 3)


def hull_polygon(ship_length, ship_beam, n_pts=30):
    """62-point local hull, jax_ice_state.py:310-328 (bow tapers at +x)."""
    half_len = ship_length / 2
    half_beam = ship_beam / 2
    lx, ly = [], []
    for i in range(n_pts + 1):
        t = i / n_pts
        x_loc = half_len - t * ship_length
        x_rel = x_loc / half_len
        taper = 1.0 - 0.9 * ((x_rel - 0.6) / 0.4) ** 1.5 if x_rel > 0.6 else 1.0
        taper = max(taper, 0.02)
        lx.append(x_loc); ly.append(half_beam * taper)
    for i in range(n_pts, -1, -1):
        t = i / n_pts
        x_loc = half_len - t * ship_length
        x_rel = x_loc / half_len
        taper = 1.0 - 0.9 * ((x_rel - 0.6) / 0.4) ** 1.5 if x_rel > 0.6 else 1.0
        taper = max(taper, 0.02)
        lx.append(x_loc); ly.append(-half_beam * taper)
    return np.column_stack([lx, ly])


def build_ship_dynamics_data(ship_mass, ship_length, ship_beam, hull_local=None) -> ShipDynF64:
    """jax_ice_state.py:270-341."""
    m = ship_mass
    Izz = (1.0 / 12.0) * m * (ship_length ** 2 + ship_beam ** 2)
    s = Fossen(mass=m, length=ship_length)
    M3_diag = np.diag(np.diag(s.M3))
    D3_diag = np.diag(np.diag(s.D3))
    Kp = s.wn @ s.wn @ M3_diag
    Kd = 2.0 * s.zeta @ s.wn @ M3_diag - D3_diag
    Ki = np.zeros((3, 3))
    B_pinv = s.B.T @ np.linalg.inv(s.B @ s.B.T)
    hull = hull_polygon(ship_length, ship_beam) if hull_local is None else np.asarray(hull_local, np.float64)
    return ShipDynF64(s.M3inv, s.D3, float(s.mass), float(Izz), Kp, Kd, Ki, B_pinv, s.n_max, hull)


def time_stepping(static: StaticF64, cfg: OracleConfig, dt: float):
    """(dt_ice, n_ice_steps), jax_simulation.py:771-785."""
    nu_r = cfg.poisson_ratio
    G_shear = YOUNGS_MODULUS / (2.0 * (1.0 + nu_r))
    r_min = float(np.array(static.radii).min())
    dt_rayleigh = cfg.rayleigh_safety_factor * np.pi * r_min / (0.163 * nu_r + 0.877) * np.sqrt(cfg.ice_density / G_shear)
    dt_ice = min(dt, dt_rayleigh) if cfg.use_rayleigh_dt_cap else dt
    n_ice_steps = max(1, int(round(dt / dt_ice))) if dt_ice > 0 else 1
    return dt_ice, n_ice_steps
