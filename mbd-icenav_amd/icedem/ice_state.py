"""State pytrees and one-time builders (``medea_ice_ship/jax_ice_state.py``).

``IceState`` / ``ShipState`` / ``IceStaticData`` / ``ShipDynamicsData`` keep the
reference's field names and shapes.  Device arrays are torch tensors in HBM,
float32 (indices int32, flags bool).  Per-element 2-vectors such as
``positions`` have the reference's (n, 2) shape but structure-of-arrays
storage: they are ``(2, n)`` buffers seen through a transposed view, so the
kernels read x and y as two coalesced streams.

The builders run once on the host in binary64 NumPy, with the reference's
formulas (jax_ice_state.py:150-355), and round to float32 on upload -- what the
reference's ``jnp.array`` does in float32 mode.  The all-pairs list
(jax_ice_state.py:212-214) is not built: the uniform-grid broad phase finds the
same contact set.
"""
from __future__ import annotations

import math
from typing import List, NamedTuple, Optional, Tuple

import numpy as np
import torch

from .properties import IceProperties


class IceState(NamedTuple):
    positions: torch.Tensor          # (n, 2) view of (2, n)
    velocities: torch.Tensor         # (n, 2) view of (2, n)
    rotations: torch.Tensor          # (n,)
    angular_vels: torch.Tensor       # (n,)
    submergence: torch.Tensor        # (n,)
    bond_intact: torch.Tensor        # (n_bonds,) bool
    bond_damage: torch.Tensor        # (n_bonds,)
    bond_delta_m_0: torch.Tensor     # (n_bonds,)
    bond_delta_m_max: torch.Tensor   # (n_bonds,)
    bond_delta_s: torch.Tensor       # (n_bonds,)
    elem_active: torch.Tensor        # (n,) bool
    dpdem_floe_eligible: torch.Tensor  # (n_floes,) bool


class ShipState(NamedTuple):
    eta: torch.Tensor        # (3,)  views into one (15,) buffer
    nu: torch.Tensor         # (3,)
    n_actual: torch.Tensor   # (6,)
    e_int: torch.Tensor      # (3,)


class IceStaticData(NamedTuple):
    local_polys_padded: np.ndarray   # (n, max_V, 2) float64 (host copy)
    poly_mask: np.ndarray            # (n, max_V) bool
    poly_n_verts: np.ndarray         # (n,) int32
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


class ShipDynamicsData(NamedTuple):
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


# ── geometry helpers (geometry.py:6-18) ───────────────────────────────────────

def polygon_area(verts) -> float:
    n = len(verts)
    area = 0
    for i in range(n):
        j = (i + 1) % n
        area += verts[i, 0] * verts[j, 1]
        area -= verts[j, 0] * verts[i, 1]
    return abs(area) / 2


def polygon_centroid(verts):
    return np.mean(verts, axis=0)


def build_static_data(polygons, centroids, bonds, thicknesses, parent_ids, cfg) -> IceStaticData:
    """jax_ice_state.py:150-243 (binary64), without the all-pairs list."""
    n = len(polygons)
    thicknesses = np.asarray(thicknesses, dtype=np.float64)
    centroids = np.asarray(centroids, dtype=np.float64)
    polys = [np.asarray(p, dtype=np.float64) for p in polygons]
    areas = np.array([polygon_area(p) for p in polys], dtype=np.float64)
    radii = np.sqrt(areas / np.pi)
    masses = areas * thicknesses * getattr(cfg, "ice_density", IceProperties.ICE_DENSITY)
    inertia = np.maximum(0.5 * masses * radii ** 2, 1e-10)
    local = [(p - polygon_centroid(p)).astype(np.float64) for p in polys]
    n_verts = np.array([len(p) for p in local], dtype=np.int32)
    max_V = int(n_verts.max())
    padded = np.zeros((n, max_V, 2), dtype=np.float64)
    mask = np.zeros((n, max_V), dtype=bool)
    for i, poly in enumerate(local):
        nv = len(poly)
        padded[i, :nv] = poly
        padded[i, nv:] = poly[-1]
        mask[i, :nv] = True
    base = getattr(cfg, "flexural_strength_base", IceProperties.FLEXURAL_STRENGTH_BASE)
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
            bs[b] = IceProperties.breaking_force(h, w, base)
            ba[b] = w * h
        sigma_bar = np.maximum(bs / np.maximum(ba, 1e-10), 1e6)
    else:
        bi = np.zeros(0, np.int32); bj = np.zeros(0, np.int32)
        rest = np.zeros(0); bs = np.zeros(0); ba = np.zeros(0); sigma_bar = np.zeros(0)
    nu_r = getattr(cfg, "poisson_ratio", 0.3)
    E = IceProperties.YOUNGS_MODULUS
    prefactor = E * (1.0 - nu_r) / ((1.0 + nu_r) * (1.0 - 2.0 * nu_r))
    k_n = prefactor
    k_s = prefactor * (1.0 - 2.0 * nu_r) / (2.0 * (1.0 - nu_r))
    floe = np.asarray(parent_ids, dtype=np.int32) if parent_ids is not None else np.arange(n, dtype=np.int32)
    return IceStaticData(padded, mask, n_verts, masses, inertia, radii, areas, thicknesses, bi, bj, rest,
                         bs, ba, sigma_bar, floe, float(k_n), float(k_s))


class FossenSupply3DOF:
    """Bis-scaled supply vessel constants (auto_icenav_ship.py:100-160)."""

    def __init__(self, mass: float = 6000.0e3, length: float = 76.2):
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


def ship_hull_polygon(ship_length: float, ship_beam: float, n_pts: int = 30) -> np.ndarray:
    """Tapered hull outline in the body frame, 2 (n_pts + 1) points
    (jax_ice_state.py:310-328; n_pts = 30 gives the reference's 62 points)."""
    half_len = ship_length / 2
    half_beam = ship_beam / 2
    lx, ly = [], []
    for side, rng in ((1.0, range(n_pts + 1)), (-1.0, range(n_pts, -1, -1))):
        for i in rng:
            t = i / n_pts
            x_loc = half_len - t * ship_length
            x_rel = x_loc / half_len
            taper = 1.0 - 0.9 * ((x_rel - 0.6) / 0.4) ** 1.5 if x_rel > 0.6 else 1.0
            taper = max(taper, 0.02)
            lx.append(x_loc)
            ly.append(side * half_beam * taper)
    return np.column_stack([lx, ly])


def build_ship_dynamics_data(ship_props, cfg, ship_length, ship_beam, hull_points: Optional[int] = None,
                             ship_hull_local: Optional[np.ndarray] = None) -> ShipDynamicsData:
    """jax_ice_state.py:270-341.  ``hull_points`` (default 62) or an explicit
    ``ship_hull_local`` select the hull polygon of the hull-in-ice pass."""
    m = ship_props.mass
    Izz = (1.0 / 12.0) * m * (ship_length ** 2 + ship_beam ** 2)
    s = FossenSupply3DOF(mass=m, length=ship_length)
    M3_diag = np.diag(np.diag(s.M3))
    D3_diag = np.diag(np.diag(s.D3))
    Kp = s.wn @ s.wn @ M3_diag
    Kd = 2.0 * s.zeta @ s.wn @ M3_diag - D3_diag
    Ki = np.zeros((3, 3))
    B_pinv = s.B.T @ np.linalg.inv(s.B @ s.B.T)
    if ship_hull_local is not None:
        hull = np.asarray(ship_hull_local, dtype=np.float64)
    else:
        npts = 30 if hull_points is None else (hull_points // 2 - 1)
        hull = ship_hull_polygon(ship_length, ship_beam, npts)
    return ShipDynamicsData(s.M3inv, s.D3, float(s.mass), float(Izz), Kp, Kd, Ki, B_pinv, s.n_max, hull)


# ── device state ─────────────────────────────────────────────────────────────

def _device(device=None) -> torch.device:
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def soa_pair(values: np.ndarray | torch.Tensor | None, n: int, device) -> torch.Tensor:
    """(n, 2) float32 tensor with (2, n) storage (x stream, then y stream)."""
    buf = torch.zeros((2, n), dtype=torch.float32, device=device)
    if values is not None:
        v = torch.as_tensor(np.asarray(values, dtype=np.float32) if not torch.is_tensor(values) else values)
        buf.copy_(v.to(device=device, dtype=torch.float32).reshape(n, 2).t())
    return buf.t()


def build_initial_ice_state(n, centroids, n_bonds, initial_floe_of, device=None) -> IceState:
    """jax_ice_state.py:246-267 (float32, on the GPU)."""
    dev = _device(device)
    floe = np.asarray(initial_floe_of)
    n_floes = int(floe.max()) + 1 if n > 0 else 0
    z = lambda m: torch.zeros(m, dtype=torch.float32, device=dev)
    return IceState(
        positions=soa_pair(np.asarray(centroids, dtype=np.float64), n, dev),
        velocities=soa_pair(None, n, dev),
        rotations=z(n), angular_vels=z(n), submergence=z(n),
        bond_intact=torch.ones(n_bonds, dtype=torch.bool, device=dev),
        bond_damage=z(n_bonds), bond_delta_m_0=z(n_bonds), bond_delta_m_max=z(n_bonds), bond_delta_s=z(n_bonds),
        elem_active=torch.zeros(n, dtype=torch.bool, device=dev),
        dpdem_floe_eligible=torch.zeros(n_floes, dtype=torch.bool, device=dev),
    )


def ship_state_from_buffer(buf: torch.Tensor) -> ShipState:
    return ShipState(buf[0:3], buf[3:6], buf[6:12], buf[12:15])


def build_initial_ship_state(ship_pos, ship_heading, device=None) -> ShipState:
    """jax_ice_state.py:344-355."""
    dev = _device(device)
    buf = torch.zeros(15, dtype=torch.float32, device=dev)
    buf[0] = float(ship_pos[0]); buf[1] = float(ship_pos[1]); buf[2] = float(ship_heading)
    return ship_state_from_buffer(buf)
