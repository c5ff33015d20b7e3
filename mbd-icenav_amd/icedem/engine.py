"""Device engine: one ``libicedem`` context per scene (static data + scratch).

``DemEngine`` turns the reference's closure constants (``_make_frame_step``,
jax_simulation.py:325-361) into an ``IceDemParams`` struct, uploads
``IceStaticData`` in float32 and runs frames in place on device buffers.  The
reference-shaped API in :mod:`icedem.simulation` sits on top of it.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np
import torch

from . import _lib
from .ice_state import IceStaticData, IceState, ShipDynamicsData, ShipState, ship_state_from_buffer, soa_pair
from .properties import IceProperties

F32 = np.float32


def edge_sample_t(K: int) -> np.ndarray:
    """``jnp.linspace(1/(K+1), K/(K+1), K)`` in float32 as JAX evaluates it
    (start*(1-step) + stop*step, step = iota/(K-1), endpoint exact)."""
    start = F32(1.0 / (K + 1))
    stop = F32(K / (K + 1))
    if K == 1:
        return np.array([start], dtype=F32)
    div = K - 1
    step = (np.arange(div, dtype=F32) / F32(div)).astype(F32)
    out = start * (F32(1.0) - step) + stop * step
    return np.concatenate([out, np.array([stop], dtype=F32)]).astype(F32)


def bound_radius(local_polys_padded: np.ndarray) -> np.ndarray:
    """Per-element max |local vertex| (binary64), rounded up generously."""
    r = np.sqrt(np.sum(local_polys_padded.astype(np.float64) ** 2, axis=2)).max(axis=1)
    return (r * 1.0001 + 1e-3).astype(F32)


# Reach of the contact partner lists beyond contact (m): the lists are rebuilt
# once an element may have moved about half of it.  A wider skin means fewer
# build frames (each scans the grids for every element) but longer lists for
# the live elements, and more elements past a list's 32 partners (those scan
# the grids every frame).  On the 1.5 x 3 km field at 50 %: 1.0 m 17.0 us per
# frame (157 builds in 2000 frames), 2.0 m 15.6 (75), 3.0 m 15.0 (49), 4.0 m
# 15.0, 5.0 m 20.0 (lists overflow); with every element active 2.0 and 3.0 m
# 36.9 against 40.4 at 1.0 m, 4.0 m overflows.  2.0 m keeps a margin from
# the overflow.
DEFAULT_LIST_SKIN = 2.0


def choose_grid(centroids: np.ndarray, radii: np.ndarray, dilation: float = 0.0, cell: Optional[float] = None,
                margin: float = 60.0, max_cells: int = 1 << 22, small_quantile: float = 0.98,
                cell_factor: float = 1.0):
    """Fine uniform grid of the broad phase.

    Elements with radius <= r_s (the ``small_quantile`` radius) live in one
    cell of side ``cell`` (default cell_factor * 2 (r_s + delta)); the few
    larger ones live in a coarse grid that the library sizes.  Every element
    scans the fine cells within r + r_s + delta of its centre.  Either way
    every contacting pair is
    found (exactly the reference's contact set); positions outside the grid
    are clamped to the border cells.  Returns (x0, y0, cell, r_s, nx, ny,
    bucket_capacity, coarse_bucket_capacity).
    """
    r = np.asarray(radii, dtype=np.float64)
    r32 = r.astype(F32)
    if len(r):
        r_s = float(F32(np.quantile(r, small_quantile)))
    else:
        r_s = 1.0
    # any cell size gives the exact contact set (each element scans the cells
    # its reach box touches); 2 (r_s + delta) keeps a small element's scan to
    # at most 3x3 cells, smaller cells scan more cells holding fewer
    # non-partners.  Floor: half of that (at most 5x5 cells).
    min_cell = 2.0 * (r_s + dilation) * 1.01 + 1e-3
    cell = max(0.5 * min_cell, float(cell) if cell is not None else min_cell * cell_factor)
    c = np.asarray(centroids, dtype=np.float64)
    lo = c.min(axis=0) - margin if len(c) else np.zeros(2)
    hi = c.max(axis=0) + margin if len(c) else np.ones(2)
    while True:
        nx = max(1, int(math.ceil((hi[0] - lo[0]) / cell)))
        ny = max(1, int(math.ceil((hi[1] - lo[1]) / cell)))
        if nx * ny <= max_cells:
            break
        cell *= 1.5
    # fine-bucket capacity from the initial occupancy of the small elements
    # (2x headroom); a full bucket spills into the overflow list, so this is a
    # speed choice only
    big = r32 > F32(r_s)
    counts = np.zeros(nx * ny, dtype=np.int64)
    if len(c):
        cx = np.clip(np.floor((c[:, 0] - lo[0]) / cell), 0, nx - 1).astype(np.int64)
        cy = np.clip(np.floor((c[:, 1] - lo[1]) / cell), 0, ny - 1).astype(np.int64)
        np.add.at(counts, (cy * nx + cx)[~big], 1)
    bucket = int(max(16, 2 * int(counts.max(initial=0)) + 8))
    bucket = (bucket + 3) // 4 * 4
    # coarse-bucket capacity the same way (the library sizes the coarse grid:
    # cell >= 2 (2 r_max + 2 delta), cells over the fine grid's extent)
    bucket2 = 0
    if big.any():
        c2 = max(cell, 2.0 * (2.0 * float(r32.max()) + 2.0 * dilation) * 1.002 + 1e-3)
        nx2 = max(1, int(math.ceil(nx * cell / c2)))
        ny2 = max(1, int(math.ceil(ny * cell / c2)))
        bx = np.clip(np.floor((c[big, 0] - lo[0]) / c2), 0, nx2 - 1).astype(np.int64)
        by = np.clip(np.floor((c[big, 1] - lo[1]) / c2), 0, ny2 - 1).astype(np.int64)
        cnt2 = np.bincount(by * nx2 + bx, minlength=nx2 * ny2)
        bucket2 = int(max(8, 2 * int(cnt2.max()) + 4))
    return float(lo[0]), float(lo[1]), float(cell), r_s, nx, ny, bucket, bucket2


def make_params(static: IceStaticData, cfg, ship_dyn: ShipDynamicsData, ship_length, ship_beam, n_ice_steps,
                dt, dt_ice, use_auto_icenav, B_mat, T_n, grid=None, ship_mode: int = 0,
                ctrl_thrust_max: float = 2e6, ctrl_yaw_rate_max: float = 0.15,
                force_lanes: int = 0, list_skin: float = DEFAULT_LIST_SKIN,
                near_cand_margin: float = -1.0) -> _lib.IceDemParams:
    """Scalar parameters rounded to float32 in the reference's order of use.
    force_lanes: lanes per element in the force pass (8 or 16; 0 lets the
    library choose by element count); list_skin: reach of the contact
    partner lists beyond contact in metres (0: the grids every frame);
    near_cand_margin: margin of the next frame's near candidates (< 0: the
    library's).  None of them changes a result."""
    p = _lib.IceDemParams()
    hl = ship_length / 2.0
    hb = ship_beam / 2.0
    p.half_len = F32(hl)
    p.half_beam = F32(hb)
    K = int(getattr(cfg, "n_edge_samples", 3))
    if K > _lib.MAX_EDGE_SAMPLES:
        raise ValueError("n_edge_samples > 16")
    p.n_edge_samples = K
    if K > 0:
        t = edge_sample_t(K)
        omt = (F32(1.0) - t).astype(F32)
        for i in range(K):
            p.edge_t[i] = float(t[i])
            p.edge_omt[i] = float(omt[i])
    hull = np.asarray(ship_dyn.ship_hull_local, dtype=np.float64)
    R = max(math.hypot(hl, hb), float(np.max(np.hypot(hull[:, 0], hull[:, 1]))))
    p.ship_bound_radius = F32(R * 1.0001 + 1e-3)
    p.ship_box_half_len = F32(max(hl, float(np.max(np.abs(hull[:, 0])))) * 1.0001 + 1e-3)
    p.ship_box_half_beam = F32(max(hb, float(np.max(np.abs(hull[:, 1])))) * 1.0001 + 1e-3)
    p.w_contact_factor = F32(cfg.w_contact_radius_factor)
    p.w_contact_cap = F32(cfg.w_contact_cap)
    p.flexural_base = F32(getattr(cfg, "flexural_strength_base", IceProperties.FLEXURAL_STRENGTH_BASE))
    p.pen_crit_ref = F32(cfg.pen_crit_ref)
    p.pen_crit_thk_ref = F32(cfg.pen_crit_thickness_ref)
    p.side_push_factor = F32(cfg.side_push_factor)
    p.has_thrust_cap = 0 if cfg.max_ship_ice_force is None else 1
    p.max_ship_ice_force = F32(cfg.max_ship_ice_force if cfg.max_ship_ice_force is not None else 0.0)
    p.submergence_pen_cap = F32(cfg.submergence_pen_cap)
    p.submergence_rate = F32(cfg.submergence_rate)
    p.inv_rise_time = F32(1.0 / cfg.rise_time_const)
    p.buoyancy_coef = F32(IceProperties.WATER_DENSITY * 9.81)
    bstr32 = np.asarray(static.bond_strengths, dtype=F32)
    p.eff_strength_collision = F32(cfg.effective_strength_collision)
    p.elig_threshold = F32(bstr32.min() * F32(cfg.effective_strength_collision)) if len(bstr32) else F32(0.0)
    p.youngs_modulus = F32(IceProperties.YOUNGS_MODULUS)
    p.bond_stiffness_factor = F32(cfg.bond_stiffness_factor)
    p.hertz_factor = F32(cfg.hertz_factor)
    p.hertz_exponent = F32(cfg.hertz_exponent)
    p.ice_friction = F32(cfg.ice_ice_friction)
    p.friction_cap = F32(cfg.friction_cap)
    delta = float(getattr(cfg, "contact_dilation", 0.0))
    p.contact_dilation2 = F32(2.0 * delta)
    p.drag_ref_thickness = F32(cfg.drag_ref_thickness)
    p.drag_coef = F32(F32(0.5 * IceProperties.WATER_DENSITY) * F32(cfg.c_d_drag))
    p.linear_drag_factor = F32(cfg.linear_drag_factor)
    p.rot_drag_coef = F32(F32(F32(-0.5) * F32(getattr(cfg, "drag_coefficient_moment", 0.4)))
                          * F32(IceProperties.WATER_DENSITY))
    p.dt = F32(dt)
    p.dt_ice = F32(dt_ice)
    p.velocity_clip = F32(float(np.max(np.abs(cfg.velocity_clip))))
    p.angular_damping = F32(cfg.angular_damping)
    p.ang_cap_lo = F32(cfg.angular_cap[0])
    p.ang_cap_hi = F32(cfg.angular_cap[1])
    p.position_clip = F32(float(getattr(cfg, "position_clip_m", 1e4)))
    p.n_ice_steps = int(n_ice_steps)
    p.max_fracture_thickness = F32(getattr(cfg, "max_fracture_thickness", 15.0))
    p.k_n = F32(static.k_n)
    p.k_s = F32(static.k_s)
    p.bond_cohesion = F32(cfg.bond_cohesion)
    p.bond_friction = F32(cfg.bond_internal_friction)
    p.G_c_I = F32(cfg.bond_G_c_I)
    p.G_c_II = F32(cfg.bond_G_c_II)
    p.eta_bk = F32(cfg.bond_eta_bk)
    p.use_auto_icenav = 1 if use_auto_icenav else 0
    p.ship_mode = int(ship_mode)
    p.ctrl_thrust_max = F32(ctrl_thrust_max)
    p.ctrl_yaw_rate_max = F32(ctrl_yaw_rate_max)
    p.ship_mass = F32(ship_dyn.mass)
    p.ship_Izz = F32(ship_dyn.Izz)
    p.ice_brake_threshold = F32(float(getattr(cfg, "ice_brake_threshold", 0.0)))
    m32 = lambda a: np.asarray(a, dtype=F32)
    for name, arr in (("M3inv", ship_dyn.M3inv), ("D3", ship_dyn.D3), ("Kp", ship_dyn.Kp), ("Kd", ship_dyn.Kd),
                      ("Ki", ship_dyn.Ki)):
        getattr(p, name)[:] = [float(v) for v in m32(arr).ravel()]
    p.D3pKd[:] = [float(v) for v in (m32(ship_dyn.D3) + m32(ship_dyn.Kd)).astype(F32).ravel()]
    p.B_pinv[:] = [float(v) for v in m32(ship_dyn.B_pinv).ravel()]
    p.B[:] = [float(v) for v in m32(B_mat).ravel()]
    p.n_max[:] = [float(v) for v in m32(ship_dyn.n_max).ravel()]
    p.T_n = F32(T_n)
    vc = tuple(float(v) for v in cfg.ship_velocity_clip)
    p.vc_surge, p.vc_sway, p.vc_yaw = F32(vc[0]), F32(vc[1]), F32(vc[2])
    if grid is None:
        raise ValueError("grid (x0, y0, cell, r_s, nx, ny, cap) is required; see choose_grid")
    p.grid_x0, p.grid_y0, p.grid_cell, p.grid_rsmall = F32(grid[0]), F32(grid[1]), F32(grid[2]), F32(grid[3])
    p.grid_nx, p.grid_ny, p.grid_bucket = int(grid[4]), int(grid[5]), int(grid[6])
    p.grid_bucket2 = int(grid[7]) if len(grid) > 7 else 0
    if force_lanes not in (0, 8, 16):
        raise ValueError("force_lanes must be 0, 8 or 16")
    p.force_lanes = int(force_lanes)
    if not (list_skin >= 0.0 and math.isfinite(list_skin)):
        raise ValueError("list_skin must be finite and >= 0")
    p.list_skin = F32(list_skin)
    p.near_cand_margin = F32(near_cand_margin)
    return p


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def _stream_ptr(stream) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class DemEngine:
    """One device context: static data, parameters and per-frame scratch."""

    def __init__(self, static: IceStaticData, cfg, ship_dyn: ShipDynamicsData, ship_length, ship_beam,
                 n_ice_steps, dt, dt_ice, use_auto_icenav, B_mat, T_n, centroids=None, device=None, grid=None,
                 cell_size=None, ship_mode: int = 0, ctrl_thrust_max: float = 2e6, ctrl_yaw_rate_max: float = 0.15,
                 grid_margin: float = 60.0, small_quantile: float = 0.98, force_lanes: int = 0,
                 cell_factor: float = 1.0, list_skin: float = None, near_cand_margin: float = -1.0):
        if not torch.cuda.is_available():
            raise _lib.IceDemError("icedem needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.static = static
        self.cfg = cfg
        self.ship_dyn = ship_dyn
        self.n = int(len(static.radii))
        self.n_bonds = int(len(static.bond_i))
        self.n_floes = int(np.max(static.initial_floe_of)) + 1 if self.n else 0
        self.dt = float(dt)
        if grid is None:
            cen = np.asarray(centroids, dtype=np.float64) if centroids is not None else np.zeros((self.n, 2))
            grid = choose_grid(cen, static.radii, float(getattr(cfg, "contact_dilation", 0.0)), cell=cell_size,
                               margin=grid_margin, small_quantile=small_quantile, cell_factor=cell_factor)
        self.grid = grid
        self.params = make_params(static, cfg, ship_dyn, ship_length, ship_beam, n_ice_steps, dt, dt_ice,
                                  use_auto_icenav, B_mat, T_n, grid=grid, ship_mode=ship_mode,
                                  ctrl_thrust_max=ctrl_thrust_max, ctrl_yaw_rate_max=ctrl_yaw_rate_max,
                                  force_lanes=force_lanes,
                                  list_skin=DEFAULT_LIST_SKIN if list_skin is None else list_skin,
                                  near_cand_margin=near_cand_margin)
        lp = np.ascontiguousarray(static.local_polys_padded, dtype=F32)
        self._host = dict(
            lp=lp,
            mask=np.ascontiguousarray(static.poly_mask, dtype=np.uint8),
            mass=np.ascontiguousarray(static.masses, dtype=F32),
            inertia=np.ascontiguousarray(static.inertia, dtype=F32),
            radii=np.ascontiguousarray(static.radii, dtype=F32),
            areas=np.ascontiguousarray(static.areas, dtype=F32),
            thick=np.ascontiguousarray(static.thicknesses, dtype=F32),
            bound=bound_radius(np.asarray(static.local_polys_padded)),
            floe=np.ascontiguousarray(static.initial_floe_of, dtype=np.int32),
            bi=np.ascontiguousarray(static.bond_i, dtype=np.int32),
            bj=np.ascontiguousarray(static.bond_j, dtype=np.int32),
            rest=np.ascontiguousarray(static.rest_lengths, dtype=F32),
            bstr=np.ascontiguousarray(static.bond_strengths, dtype=F32),
            sbar=np.ascontiguousarray(static.sigma_bar_bond, dtype=F32),
            hull=np.ascontiguousarray(ship_dyn.ship_hull_local, dtype=F32),
        )
        h = self._host
        st = _lib.IceDemStatic()
        st.n, st.n_bonds, st.n_floes = self.n, self.n_bonds, self.n_floes
        st.max_v = int(lp.shape[1])
        st.n_hull = int(h["hull"].shape[0])
        for field, key in (("h_local_polys", "lp"), ("h_poly_mask", "mask"), ("h_masses", "mass"),
                           ("h_inertia", "inertia"), ("h_radii", "radii"), ("h_areas", "areas"),
                           ("h_thicknesses", "thick"), ("h_bound_radius", "bound"), ("h_floe_of", "floe"),
                           ("h_bond_i", "bi"), ("h_bond_j", "bj"), ("h_rest_lengths", "rest"),
                           ("h_bond_strengths", "bstr"), ("h_sigma_bar", "sbar"), ("h_hull_local", "hull")):
            setattr(st, field, h[key].ctypes.data if h[key].size else None)
        ctx = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.icedem_create(C.byref(st), C.byref(self.params), self.device.index or 0,
                                              C.byref(ctx)), None, "icedem_create")
        self.ctx = ctx
        self.max_v = st.max_v

    @classmethod
    def like(cls, base: "DemEngine") -> "DemEngine":
        """An engine for base's scene, parameters and device that shares its
        static tables on the GPU (icedem_create_like), with its own state
        scratch: many rollout instances of one scene."""
        eng = cls.__new__(cls)
        for k in ("lib", "device", "static", "cfg", "ship_dyn", "n", "n_bonds", "n_floes", "dt", "grid", "params",
                  "_host", "max_v"):
            setattr(eng, k, getattr(base, k))
        ctx = C.c_void_p()
        with torch.cuda.device(eng.device):
            _lib.check(eng.lib.icedem_create_like(base.ctx, C.byref(ctx)), None, "icedem_create_like")
        eng.ctx = ctx
        return eng

    def near_stats(self, st: IceState, ship_buf: torch.Tensor):
        """Elements passing the near-ship box test for the current state, and
        the 256-element blocks holding one (host-side count, for traffic
        accounting)."""
        pos = st.positions.detach().float().cpu().numpy().astype(np.float64)
        sx, sy, sh = (float(v) for v in ship_buf[:3].detach().cpu().numpy())
        b = self._host["bound"].astype(np.float64) + 1.0
        dx, dy = pos[:, 0] - sx, pos[:, 1] - sy
        u = math.cos(sh) * dx + math.sin(sh) * dy
        v = math.cos(sh) * dy - math.sin(sh) * dx
        near = (np.abs(u) <= float(self.params.ship_box_half_len) + b) & \
               (np.abs(v) <= float(self.params.ship_box_half_beam) + b)
        blocks = np.unique(np.nonzero(near)[0] // 256)
        return int(near.sum()), int(len(blocks))

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.icedem_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ── buffers ──────────────────────────────────────────────────────────────
    def check_state(self, st: IceState):
        n, nb = self.n, self.n_bonds
        for name, t, shape, dt in (("positions", st.positions, (n, 2), torch.float32),
                                   ("velocities", st.velocities, (n, 2), torch.float32)):
            if tuple(t.shape) != shape or t.dtype != dt or t.stride() != (1, n) or not t.is_cuda:
                raise ValueError(f"{name}: expected a CUDA float32 (n,2) view of (2,n) storage")
        for name, t, m, dt in (("rotations", st.rotations, n, torch.float32),
                               ("angular_vels", st.angular_vels, n, torch.float32),
                               ("submergence", st.submergence, n, torch.float32),
                               ("elem_active", st.elem_active, n, torch.bool),
                               ("bond_intact", st.bond_intact, nb, torch.bool),
                               ("bond_damage", st.bond_damage, nb, torch.float32),
                               ("bond_delta_m_0", st.bond_delta_m_0, nb, torch.float32),
                               ("bond_delta_m_max", st.bond_delta_m_max, nb, torch.float32),
                               ("bond_delta_s", st.bond_delta_s, nb, torch.float32),
                               ("dpdem_floe_eligible", st.dpdem_floe_eligible, self.n_floes, torch.bool)):
            if tuple(t.shape) != (m,) or t.dtype != dt or not t.is_contiguous() or not t.is_cuda:
                raise ValueError(f"{name}: expected a contiguous CUDA {dt} tensor of shape ({m},)")

    def state_struct(self, st: IceState, ship_buf: torch.Tensor) -> _lib.IceDemState:
        self.check_state(st)
        if ship_buf.numel() != 15 or ship_buf.dtype != torch.float32 or not ship_buf.is_contiguous():
            raise ValueError("ship state buffer must be 15 contiguous float32")
        s = _lib.IceDemState()
        s.d_pos = _ptr(st.positions); s.d_vel = _ptr(st.velocities)
        s.d_rot = _ptr(st.rotations); s.d_omega = _ptr(st.angular_vels); s.d_sub = _ptr(st.submergence)
        s.d_active = _ptr(st.elem_active)
        s.d_bond_intact = _ptr(st.bond_intact) if self.n_bonds else None
        s.d_bond_damage = _ptr(st.bond_damage) if self.n_bonds else None
        s.d_bond_dm0 = _ptr(st.bond_delta_m_0) if self.n_bonds else None
        s.d_bond_dmmax = _ptr(st.bond_delta_m_max) if self.n_bonds else None
        s.d_bond_ds = _ptr(st.bond_delta_s) if self.n_bonds else None
        s.d_floe_elig = _ptr(st.dpdem_floe_eligible) if self.n_floes else None
        s.d_ship = _ptr(ship_buf)
        return s

    def clone_state(self, st: IceState) -> IceState:
        """Fresh SoA copies (the reference's step is functional)."""
        n = self.n
        pos = soa_pair(None, n, self.device); pos.copy_(st.positions)
        vel = soa_pair(None, n, self.device); vel.copy_(st.velocities)
        c = lambda t, dt: torch.as_tensor(t).to(device=self.device, dtype=dt).clone().contiguous()
        return IceState(pos, vel, c(st.rotations, torch.float32), c(st.angular_vels, torch.float32),
                        c(st.submergence, torch.float32), c(st.bond_intact, torch.bool),
                        c(st.bond_damage, torch.float32), c(st.bond_delta_m_0, torch.float32),
                        c(st.bond_delta_m_max, torch.float32), c(st.bond_delta_s, torch.float32),
                        c(st.elem_active, torch.bool), c(st.dpdem_floe_eligible, torch.bool))

    def ship_buffer(self, ship: ShipState) -> torch.Tensor:
        parts = [torch.as_tensor(x).to(device=self.device, dtype=torch.float32).reshape(-1)
                 for x in (ship.eta, ship.nu, ship.n_actual, ship.e_int)]
        return torch.cat(parts).contiguous()

    # ── frames ───────────────────────────────────────────────────────────────
    def run_frames(self, st: IceState, ship_buf: torch.Tensor, setpoints: torch.Tensor,
                   v_water: Optional[torch.Tensor] = None, v_water_ship: Optional[torch.Tensor] = None,
                   hist: Optional[dict] = None, diag: Optional[dict] = None, use_graph: bool = True,
                   stream=None, sync: bool = True, controls: Optional[torch.Tensor] = None, hold: int = 1):
        """Run K = len(setpoints) frames in place.  v_water: (K, n, 2), (n, 2)
        (broadcast over frames) or None; any strides."""
        K = int(setpoints.shape[0])
        sp = setpoints.to(device=self.device, dtype=torch.float32).contiguous()
        if tuple(sp.shape) != (K, 6):
            raise ValueError("setpoints must be (K, 6)")
        inp = _lib.IceDemInputs()
        inp.d_setpoints = _ptr(sp)
        keep = [sp]
        if v_water is not None:
            vw = torch.as_tensor(v_water)
            if vw.device != self.device or vw.dtype != torch.float32:
                vw = vw.to(device=self.device, dtype=torch.float32)
            if vw.dim() == 2:
                if tuple(vw.shape) != (self.n, 2):
                    raise ValueError("v_water must be (n, 2) or (K, n, 2)")
                inp.vw_stride_k = 0
                inp.vw_stride_e, inp.vw_stride_c = vw.stride()
            else:
                if tuple(vw.shape) != (K, self.n, 2):
                    raise ValueError("v_water must be (K, n, 2)")
                inp.vw_stride_k, inp.vw_stride_e, inp.vw_stride_c = vw.stride()
            inp.d_v_water = _ptr(vw)
            keep.append(vw)
        if controls is not None:
            ctl = torch.as_tensor(controls).to(device=self.device, dtype=torch.float32).contiguous()
            inp.d_controls = _ptr(ctl)
            inp.n_knots = int(ctl.shape[0])
            inp.hold = int(hold)
            keep.append(ctl)
        if v_water_ship is not None:
            vws = torch.as_tensor(v_water_ship).to(device=self.device, dtype=torch.float32).contiguous()
            if tuple(vws.shape) != (K, 2):
                raise ValueError("v_water_ship must be (K, 2)")
            inp.d_v_water_ship = _ptr(vws)
            keep.append(vws)
        h = _lib.IceDemHistory()
        if hist:
            for key, field in (("positions", "d_positions"), ("rotations", "d_rotations"),
                               ("bond_intact", "d_bond_intact"), ("eta", "d_eta"),
                               ("submergence", "d_submergence"), ("J", "d_J"), ("F", "d_F")):
                if key in hist and hist[key] is not None:
                    setattr(h, field, _ptr(hist[key]))
        d = _lib.IceDemDiag()
        if diag:
            for key, field in (("pen_max", "d_pen_max"), ("force_mag", "d_force_mag"),
                               ("F_ice_on_ship", "d_F_ice_on_ship"), ("J_cumul_incr", "d_J"),
                               ("rollout", "d_rollout")):
                if key in diag and diag[key] is not None:
                    setattr(d, field, _ptr(diag[key]))
        s = self.state_struct(st, ship_buf)
        strm = _stream_ptr(stream)
        rc = self.lib.icedem_frames(self.ctx, C.byref(s), C.byref(inp), K, C.byref(h), C.byref(d),
                                    1 if use_graph else 0, strm)
        _lib.check(rc, self.ctx, "icedem_frames")
        if sync:
            _lib.check(self.lib.icedem_sync(self.ctx, strm), self.ctx, "icedem_sync")
        return keep

    def list_stats(self, active: Optional[torch.Tensor] = None) -> dict:
        """Partner-list diagnostics: frames that rebuilt the lists since
        creation, whether an element overflowed its list in the last call,
        whether the context uses lists, partners listed at the last build;
        with ``active`` (a state's elem_active) also the partners listed for
        the active elements and their number."""
        out = (C.c_int32 * 6)()
        a = None
        if active is not None:
            a = active.to(device=self.device, dtype=torch.uint8).contiguous()
            if a.numel() != self.n:
                raise ValueError(f"active has {a.numel()} flags for {self.n} elements")
        _lib.check(self.lib.icedem_list_stats(self.ctx, _ptr(a) if a is not None else None, out), self.ctx,
                   "icedem_list_stats")
        d = {"rebuilt_frames": int(out[0]), "overflowed": bool(out[1]), "lists": bool(out[2]),
             "listed_partners": int(out[3])}
        if a is not None:
            d["listed_partners_active"] = int(out[4])
            d["active_elements"] = int(out[5])
        return d

    def alloc_history(self, K: int) -> dict:
        n, nb = self.n, self.n_bonds
        dev = self.device
        pos = torch.zeros((K, 2, n), dtype=torch.float32, device=dev)
        return dict(positions=pos, rotations=torch.zeros((K, n), dtype=torch.float32, device=dev),
                    bond_intact=torch.zeros((K, nb), dtype=torch.bool, device=dev),
                    eta=torch.zeros((K, 3), dtype=torch.float32, device=dev),
                    submergence=torch.zeros((K, n), dtype=torch.float32, device=dev),
                    J=torch.zeros(K, dtype=torch.float32, device=dev),
                    F=torch.zeros((K, 2), dtype=torch.float32, device=dev))

    def alloc_diag(self) -> dict:
        dev = self.device
        return dict(pen_max=torch.zeros(self.n, dtype=torch.float32, device=dev),
                    force_mag=torch.zeros(self.n, dtype=torch.float32, device=dev),
                    F_ice_on_ship=torch.zeros(2, dtype=torch.float32, device=dev),
                    J_cumul_incr=torch.zeros(1, dtype=torch.float32, device=dev))

    # ── leaf entry points ────────────────────────────────────────────────────
    def substep_forces(self, st: IceState, ship_buf: torch.Tensor, ship_forces: torch.Tensor,
                       v_water: Optional[torch.Tensor] = None):
        n = self.n
        sf = torch.zeros((2, n), dtype=torch.float32, device=self.device)
        sf.copy_(torch.as_tensor(ship_forces).to(self.device, torch.float32).reshape(n, 2).t())
        f = torch.zeros((2, n), dtype=torch.float32, device=self.device)
        m = torch.zeros(n, dtype=torch.float32, device=self.device)
        s = self.state_struct(st, ship_buf)
        vw_p, sc, se = None, 0, 0
        vw = None
        if v_water is not None:
            vw = torch.as_tensor(v_water).to(self.device, torch.float32)
            se, sc = vw.stride()
            vw_p = _ptr(vw)
        strm = _stream_ptr(None)
        _lib.check(self.lib.icedem_substep_forces(self.ctx, C.byref(s), _ptr(sf), vw_p, sc, se, _ptr(f), _ptr(m),
                                                  strm), self.ctx, "icedem_substep_forces")
        _lib.check(self.lib.icedem_sync(self.ctx, strm), self.ctx, "icedem_sync")
        return f.t(), m

    def contact_pairs(self, positions: torch.Tensor, max_pairs: Optional[int] = None):
        n = self.n
        pos = torch.zeros((2, n), dtype=torch.float32, device=self.device)
        pos.copy_(torch.as_tensor(positions).to(self.device, torch.float32).reshape(n, 2).t())
        cap = int(max_pairs) if max_pairs is not None else 16 * n + 64
        pi = torch.zeros(cap, dtype=torch.int32, device=self.device)
        pj = torch.zeros(cap, dtype=torch.int32, device=self.device)
        cnt = C.c_int64(0)
        _lib.check(self.lib.icedem_contact_pairs(self.ctx, _ptr(pos), _ptr(pi), _ptr(pj), cap, C.byref(cnt),
                                                 _stream_ptr(None)), self.ctx, "icedem_contact_pairs")
        total = int(cnt.value)
        if total > cap:
            return self.contact_pairs(positions, max_pairs=total)
        return pi[:total], pj[:total]

    def ship_step(self, ship_buf: torch.Tensor, setpoint6, tau_ice3, v_water_ship2=None) -> torch.Tensor:
        out = ship_buf.clone()
        sp = torch.as_tensor(np.asarray(setpoint6, dtype=np.float32)).to(self.device)
        tau = torch.as_tensor(np.asarray(tau_ice3, dtype=np.float32)).to(self.device)
        vws = None
        if v_water_ship2 is not None:
            vws = torch.as_tensor(np.asarray(v_water_ship2, dtype=np.float32)).to(self.device)
        strm = _stream_ptr(None)
        _lib.check(self.lib.icedem_ship_step(self.ctx, _ptr(out), _ptr(sp), _ptr(tau),
                                             _ptr(vws) if vws is not None else None, strm),
                   self.ctx, "icedem_ship_step")
        torch.cuda.synchronize(self.device)
        return out
