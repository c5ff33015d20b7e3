"""Oracle: the DPDEM frame step, restated in NumPy (float32 or float64).

TEST INFRASTRUCTURE ONLY (the checker; see ``oracle/__init__.py``).

Follows ``medea_ice_ship/jax_simulation.py:63-268, 302-701``,
``jax_geometry.py:42-594``, ``jax_bond_fracture.py:20-208`` and
``jax_vessel_dynamics.py:54-348``.  Conventions fixed by this restatement where
the reference's XLA lowering cannot be observed (JAX is absent):

* reductions over elements (``jnp.sum`` at jax_simulation.py:462,472,476-478)
  use :func:`tree_sum` (pairwise tree over index order, zero padded to a power
  of two); reductions over a polygon's vertices (jax_geometry.py:481-483) and
  over 2/3/6-vectors accumulate sequentially in index order;
* scatter-adds (``.at[].add``) apply updates in index order (np.add.at);
* min/max/clip propagate NaN (XLA semantics) via :func:`vmax`/:func:`vmin`;
* float32 transcendentals come from :mod:`oracle.detmath`.

Pinning: ship path and connected components against the reference's own golden
tests; fracture algebra against the known answers of
``dpdem/validation/tests.py:100-146``; contact/drag/integration/hull against
closed forms and invariants (``tests/test_oracle_*.py``); bond forces,
rotational drag, the contact pre-pass, activation, eligibility and the
fracture gates against hand-worked cases that also reject planted mutations of
each method (``tests/test_oracle_pins.py``).  XLA's exact last-ulp results
remain unpinned (no JAX here).
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np

from . import detmath
from .model import (OracleConfig, StaticF64, ShipDynF64, YOUNGS_MODULUS, WATER_DENSITY)


class IceState(NamedTuple):
    """jax_ice_state.py:31-44"""
    positions: np.ndarray
    velocities: np.ndarray
    rotations: np.ndarray
    angular_vels: np.ndarray
    submergence: np.ndarray
    bond_intact: np.ndarray
    bond_damage: np.ndarray
    bond_delta_m_0: np.ndarray
    bond_delta_m_max: np.ndarray
    bond_delta_s: np.ndarray
    elem_active: np.ndarray
    dpdem_floe_eligible: np.ndarray


class ShipState(NamedTuple):
    """jax_ice_state.py:47-52"""
    eta: np.ndarray
    nu: np.ndarray
    n_actual: np.ndarray
    e_int: np.ndarray


# ── elementwise helpers with XLA semantics ────────────────────────────────────

def vmax(a, b):
    """max(a, b) that propagates NaN (XLA max)."""
    a = np.asarray(a); b = np.asarray(b)
    return np.where(np.isnan(a) | (a >= b), a, b)


def vmin(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return np.where(np.isnan(a) | (a <= b), a, b)


def vclip(x, lo, hi):
    """jnp.clip = minimum(maximum(x, lo), hi)."""
    return vmin(vmax(x, lo), hi)


def vsign(x):
    x = np.asarray(x)
    one = np.ones_like(x)
    return np.where(x > 0, one, np.where(x < 0, -one, x))


def tree_sum(a, ft):
    """Pairwise tree over index order, zero padded to a power of two."""
    a = np.asarray(a, dtype=ft).ravel()
    n = a.shape[0]
    if n == 0:
        return ft(0.0)
    m = 1 << (n - 1).bit_length()
    b = np.zeros(m, dtype=ft)
    b[:n] = a
    while b.shape[0] > 1:
        b = b[0::2] + b[1::2]
    return b[0]


def seq_sum(vals, ft):
    """Sequential accumulation from 0 in index order (short vectors)."""
    acc = ft(0.0)
    for v in vals:
        acc = acc + v
    return acc


def edge_sample_t(K, ft):
    """``jnp.linspace(1/(K+1), K/(K+1), K)`` (jax_geometry.py:190), restated as
    JAX computes it: start*(1-step) + stop*step with step = iota/(K-1), the
    endpoint appended exactly."""
    start = ft(1.0 / (K + 1))
    stop = ft(K / (K + 1))
    if K == 1:
        return np.array([start], dtype=ft)
    div = K - 1
    step = (np.arange(div, dtype=ft) / ft(div)).astype(ft)
    out = start * (ft(1.0) - step) + stop * step
    return np.concatenate([out, np.array([stop], dtype=ft)]).astype(ft)


class Oracle:
    """Restatement of ``_make_frame_step`` (jax_simulation.py:302-701)."""

    def __init__(self, static: StaticF64, cfg: OracleConfig, ship_dyn: ShipDynF64,
                 ship_length, ship_beam, n_ice_steps, dt, dt_ice, use_auto_icenav,
                 B_mat, T_n, dtype=np.float32, pair_mode="all", pair_chunk=1 << 22,
                 ship_mode=0, ctrl_thrust_max=2e6, ctrl_yaw_rate_max=0.15):
        self.ft = ft = np.dtype(dtype).type
        self.f32 = ft is np.float32
        self.cfg = cfg
        self.st64 = static
        c = lambda a: np.asarray(a, dtype=ft)
        self.local_polys = c(static.local_polys_padded)
        self.poly_mask = np.asarray(static.poly_mask, bool)
        self.masses = c(static.masses); self.inertia = c(static.inertia)
        self.radii = c(static.radii); self.areas = c(static.areas)
        self.thick = c(static.thicknesses)
        self.bond_i = np.asarray(static.bond_i, np.int64)
        self.bond_j = np.asarray(static.bond_j, np.int64)
        self.rest = c(static.rest_lengths); self.bstr = c(static.bond_strengths)
        self.sigma_bar = c(static.sigma_bar_bond)
        self.floe = np.asarray(static.initial_floe_of, np.int64)
        self.k_n = static.k_n; self.k_s = static.k_s
        self.hull_local = c(ship_dyn.ship_hull_local)
        self.M3inv = c(ship_dyn.M3inv); self.D3 = c(ship_dyn.D3)
        self.Kp = c(ship_dyn.Kp); self.Kd = c(ship_dyn.Kd); self.Ki = c(ship_dyn.Ki)
        self.B_pinv = c(ship_dyn.B_pinv); self.n_max = c(ship_dyn.n_max)
        self.B = c(B_mat); self.T_n = float(T_n)
        self.half_len = ship_length / 2.0
        self.half_beam = ship_beam / 2.0
        self.n_ice_steps = int(n_ice_steps)
        self.dt = float(dt); self.dt_ice = float(dt_ice)
        self.use_auto = bool(use_auto_icenav)
        self.pair_mode = pair_mode
        self.pair_chunk = pair_chunk
        self.t_edge = edge_sample_t(cfg.n_edge_samples, ft) if cfg.n_edge_samples > 0 else None
        # rollout extension (no reference counterpart): direct thrust control
        self.ship_mode = int(ship_mode)
        self.ctrl_thrust_max = ctrl_thrust_max
        self.ctrl_yaw_rate_max = ctrl_yaw_rate_max
        self.ship_mass = ship_dyn.mass
        self.ship_Izz = ship_dyn.Izz

    # ── transcendentals ──────────────────────────────────────────────────────
    def cos(self, x):
        return detmath.cos32(x) if self.f32 else np.cos(np.asarray(x, np.float64))

    def sin(self, x):
        return detmath.sin32(x) if self.f32 else np.sin(np.asarray(x, np.float64))

    def pow(self, x, p):
        return detmath.pow32(x, p) if self.f32 else np.power(np.asarray(x, np.float64), np.float64(p))

    def norm2(self, x, y):
        """jnp.linalg.norm over a length-2 axis: sqrt(x*x + y*y)."""
        return np.sqrt(x * x + y * y)

    def remainder(self, x, y):
        """jnp.remainder: fmod, then + y where the sign differs."""
        ft = self.ft
        r = np.fmod(ft(x), ft(y))
        if r != 0 and ((r < 0) != (ft(y) < 0)):
            r = r + ft(y)
        return ft(r)

    # ── a9: bond forces, jax_simulation.py:63-108 ────────────────────────────
    def bond_forces(self, pos, active, intact, damage):
        ft = self.ft
        n = pos.shape[0]
        forces = np.zeros((n, 2), ft)
        moments = np.zeros(n, ft)
        if self.bond_i.shape[0] == 0:
            return forces, moments
        bi, bj = self.bond_i, self.bond_j
        dx = pos[bj, 0] - pos[bi, 0]
        dy = pos[bj, 1] - pos[bi, 1]
        dist = self.norm2(dx, dy)
        dist_safe = vmax(dist, ft(1e-10))
        dirx = dx / dist_safe
        diry = dy / dist_safe
        avg_h = (self.thick[bi] + self.thick[bj]) * ft(0.5)
        stiff = ft(YOUNGS_MODULUS) * avg_h * ft(self.cfg.bond_stiffness_factor)
        stretch = dist - self.rest
        F_mag = stiff * stretch * (ft(1.0) - damage)
        Fx = dirx * F_mag
        Fy = diry * F_mag
        Fx = np.where(intact, Fx, ft(0.0))
        Fy = np.where(intact, Fy, ft(0.0))
        both_active = active[bi] & active[bj]
        both_frozen = (~active[bi]) & (~active[bj])
        ok = (both_active | both_frozen).astype(ft)
        Fx = Fx * ok
        Fy = Fy * ok
        fx = np.zeros(n, ft); fy = np.zeros(n, ft)
        np.add.at(fx, bi, Fx); np.add.at(fy, bi, Fy)
        np.add.at(fx, bj, -Fx); np.add.at(fy, bj, -Fy)
        forces[:, 0] = fx; forces[:, 1] = fy
        cross = ft(0.5) * (dx * Fy - dy * Fx)
        np.add.at(moments, bi, cross)
        np.add.at(moments, bj, -cross)
        return forces, moments

    # ── a10: ice-ice contact, jax_simulation.py:111-164 ──────────────────────
    def _pair_terms(self, pi, pj, pos, vel):
        """Per-pair contributions (contact pairs only are returned)."""
        ft = self.ft
        cfg = self.cfg
        dx = pos[pj, 0] - pos[pi, 0]
        dy = pos[pj, 1] - pos[pi, 1]
        dist = self.norm2(dx, dy)
        min_dist = self.radii[pi] + self.radii[pj]
        if cfg.contact_dilation != 0.0:
            min_dist = min_dist + ft(2.0 * cfg.contact_dilation)
        overlap = vmax(min_dist - dist, ft(0.0))
        overlap = np.where(dist > ft(1e-6), overlap, ft(0.0))
        has = overlap > ft(0.0)
        idx = np.nonzero(has)[0]
        pi, pj = pi[idx], pj[idx]
        dx, dy, dist, overlap = dx[idx], dy[idx], dist[idx], overlap[idx]
        den = vmax(dist, ft(1e-10))
        dirx = dx / den
        diry = dy / den
        f_n = ft(cfg.hertz_factor) * self.pow(overlap, cfg.hertz_exponent)
        fnx = dirx * f_n
        fny = diry * f_n
        vrx = vel[pj, 0] - vel[pi, 0]
        vry = vel[pj, 1] - vel[pi, 1]
        vn = vrx * dirx + vry * diry
        vtx = vrx - vn * dirx
        vty = vry - vn * diry
        vt = self.norm2(vtx, vty)
        vden = vmax(vt, ft(1e-10))
        thx = (-vtx) / vden
        thy = (-vty) / vden
        f_t = vmin(ft(cfg.ice_ice_friction) * f_n, ft(cfg.friction_cap))
        ftx = thx * f_t
        fty = thy * f_t
        valid = vt > ft(1e-6)
        ftx = np.where(valid, ftx, ft(0.0))
        fty = np.where(valid, fty, ft(0.0))
        fix = -fnx + ftx          # added to i
        fiy = -fny + fty
        fjx = fnx - ftx           # added to j
        fjy = fny - fty
        Fonix = -fnx - ftx
        Foniy = -fny - fty
        cross = -(dx * Foniy - dy * Fonix)
        mi = ft(0.5) * cross
        mj = ft(-0.5) * cross
        return pi, pj, fix, fiy, fjx, fjy, mi, mj

    def contact_pairs(self, pos):
        """Sorted (i<j) pairs with dist > 1e-6 and r_i + r_j (+2 delta) - dist > 0,
        by the reference's all-pairs enumeration (jax_ice_state.py:212-214)."""
        out_i, out_j = [], []
        for pi, pj in self._pair_chunks(pos):
            ft = self.ft
            dx = pos[pj, 0] - pos[pi, 0]
            dy = pos[pj, 1] - pos[pi, 1]
            dist = self.norm2(dx, dy)
            min_dist = self.radii[pi] + self.radii[pj]
            if self.cfg.contact_dilation != 0.0:
                min_dist = min_dist + ft(2.0 * self.cfg.contact_dilation)
            overlap = vmax(min_dist - dist, ft(0.0))
            overlap = np.where(dist > ft(1e-6), overlap, ft(0.0))
            k = np.nonzero(overlap > 0)[0]
            out_i.append(pi[k]); out_j.append(pj[k])
        if not out_i:
            return np.zeros(0, np.int64), np.zeros(0, np.int64)
        return np.concatenate(out_i), np.concatenate(out_j)

    def _pair_chunks(self, pos):
        n = pos.shape[0]
        if self.pair_mode == "cells":
            from .cells import candidate_pairs
            pi, pj = candidate_pairs(pos, self.radii, self.cfg.contact_dilation)
            yield pi, pj
            return
        # upper triangle in row-major order, chunked by rows
        i0 = 0
        while i0 < n - 1:
            rows, cnt = [], 0
            i1 = i0
            while i1 < n - 1 and cnt < self.pair_chunk:
                cnt += n - 1 - i1
                i1 += 1
            counts = np.arange(n - 1 - i0, n - 1 - i1, -1)
            pi = np.repeat(np.arange(i0, i1), counts)
            starts = np.repeat(np.arange(i0, i1) + 1, counts)
            offs = np.arange(pi.shape[0]) - np.repeat(np.cumsum(counts) - counts, counts)
            pj = starts + offs
            yield pi, pj
            i0 = i1

    def ice_ice_forces(self, pos, vel):
        ft = self.ft
        n = pos.shape[0]
        parts = [self._pair_terms(pi, pj, pos, vel) for pi, pj in self._pair_chunks(pos)]
        forces = np.zeros((n, 2), ft)
        moments = np.zeros(n, ft)
        if not parts:
            return forces, moments
        cat = [np.concatenate([p[k] for p in parts]) for k in range(8)]
        pi, pj, fix, fiy, fjx, fjy, mi, mj = cat
        fx = np.zeros(n, ft); fy = np.zeros(n, ft)
        np.add.at(fx, pi, fix); np.add.at(fy, pi, fiy)
        np.add.at(fx, pj, fjx); np.add.at(fy, pj, fjy)
        np.add.at(moments, pi, mi)
        np.add.at(moments, pj, mj)
        forces[:, 0] = fx; forces[:, 1] = fy
        return forces, moments

    # ── a11: drag, jax_simulation.py:167-212 ─────────────────────────────────
    def drag(self, vel, v_water):
        ft = self.ft
        cfg = self.cfg
        A_drag = self.areas * (self.thick / ft(cfg.drag_ref_thickness))
        vrx = vel[:, 0] - v_water[:, 0]
        vry = vel[:, 1] - v_water[:, 1]
        m = self.norm2(vrx, vry)
        ms = vmax(m, ft(1e-12))
        C = ft(0.5 * WATER_DENSITY) * ft(cfg.c_d_drag)      # f32(512.5) * f32(c_d)
        quad_s = C * A_drag * (ms * ms)
        qx = (vrx / ms) * quad_s
        qy = (vry / ms) * quad_s
        lin = C * A_drag * ft(cfg.linear_drag_factor)
        lx = lin * vrx
        ly = lin * vry
        big = m > ft(1e-6)
        out = np.empty((vel.shape[0], 2), ft)
        out[:, 0] = np.where(big, qx, lx)
        out[:, 1] = np.where(big, qy, ly)
        return out

    def rot_drag(self, omega, sub):
        ft = self.ft
        A_sub = self.areas * vmin(sub / vmax(self.thick, ft(1e-10)), ft(1.0))
        C = ft(-0.5) * ft(self.cfg.drag_coefficient_moment) * ft(WATER_DENSITY)
        r3 = self.radii * (self.radii * self.radii)
        return C * r3 * omega * np.abs(omega) * A_sub

    # ── a12: integration, jax_simulation.py:215-268 ──────────────────────────
    def integrate(self, st: IceState, forces, moments):
        ft = self.ft
        cfg = self.cfg
        dt = ft(self.dt_ice)
        vc = ft(float(np.max(np.abs(cfg.velocity_clip))))
        pc = ft(float(cfg.position_clip_m))
        ax = forces[:, 0] / self.masses
        ay = forces[:, 1] / self.masses
        nvx = vclip(st.velocities[:, 0] + ax * dt, -vc, vc)
        nvy = vclip(st.velocities[:, 1] + ay * dt, -vc, vc)
        npx = st.positions[:, 0] + nvx * dt
        npy = st.positions[:, 1] + nvy * dt
        alpha = moments / self.inertia
        nw = st.angular_vels * ft(cfg.angular_damping) + alpha * dt
        nw = vclip(nw, ft(cfg.angular_cap[0]), ft(cfg.angular_cap[1]))
        nr = st.rotations + nw * dt
        nvx = np.where(np.isfinite(nvx), nvx, ft(0.0))
        nvy = np.where(np.isfinite(nvy), nvy, ft(0.0))
        npx = np.where(np.isfinite(npx), npx, st.positions[:, 0])
        npy = np.where(np.isfinite(npy), npy, st.positions[:, 1])
        npx = vclip(npx, -pc, pc)
        npy = vclip(npy, -pc, pc)
        act = st.elem_active
        pos = np.empty_like(st.positions); vel = np.empty_like(st.velocities)
        vel[:, 0] = np.where(act, nvx, st.velocities[:, 0])
        vel[:, 1] = np.where(act, nvy, st.velocities[:, 1])
        pos[:, 0] = np.where(act, npx, st.positions[:, 0])
        pos[:, 1] = np.where(act, npy, st.positions[:, 1])
        nw = np.where(act, nw, st.angular_vels)
        nr = np.where(act, nr, st.rotations)
        return st._replace(positions=pos, velocities=vel, rotations=nr.astype(ft), angular_vels=nw.astype(ft))

    # ── a4: pass 1, jax_geometry.py:42-284 ───────────────────────────────────
    def world_polys(self, pos, rot):
        c = self.cos(rot); s = self.sin(rot)
        vx = self.local_polys[:, :, 0]; vy = self.local_polys[:, :, 1]
        wx = c[:, None] * vx - s[:, None] * vy + pos[:, 0:1]
        wy = s[:, None] * vx + c[:, None] * vy + pos[:, 1:2]
        return wx, wy

    def point_to_ship(self, px, py, sx, sy, sh):
        """point_to_ship_collision_batched, jax_geometry.py:42-134."""
        ft = self.ft
        hl = ft(self.half_len); hb = ft(self.half_beam)
        c_neg = self.cos(-sh); s_neg = self.sin(-sh)
        dx = px - sx; dy = py - sy
        rel_x = c_neg * dx - s_neg * dy
        rel_y = s_neg * dx + c_neg * dy
        x_rel = rel_x / hl
        in_bow = x_rel < ft(-0.6)
        t_bow = vclip(((-x_rel) - ft(0.6)) / ft(1.0 - 0.6), ft(0.0), ft(1.0))
        lhb_bow = hb * (ft(1.0) - ft(0.9) * self.pow(t_bow, 1.5))
        lhb_bow = vmax(lhb_bow, hb * ft(0.05))
        lhb = np.where(in_bow, lhb_bow, hb)
        inside = (rel_x >= -hl) & (rel_x <= hl) & (np.abs(rel_y) <= lhb)
        d_bow = rel_x + hl
        d_side = lhb - np.abs(rel_y)
        bow_angle = ft(0.3) + ft(0.5) * t_bow
        nx_bow = -self.cos(bow_angle)
        ny_bow = vsign(rel_y) * self.sin(bow_angle)
        pen_bow = vmin(d_bow, d_side) + ft(0.5)
        side = d_side <= d_bow
        nx_nb = np.where(side, ft(0.0), ft(-1.0))
        ny_nb = np.where(side, vsign(rel_y), ft(0.0))
        pen_nb = np.where(side, d_side, d_bow)
        nxb = np.where(in_bow, nx_bow, nx_nb)
        nyb = np.where(in_bow, ny_bow, ny_nb)
        pen_raw = np.where(in_bow, pen_bow, pen_nb)
        c_h = self.cos(sh); s_h = self.sin(sh)
        nxw = c_h * nxb - s_h * nyb
        nyw = s_h * nxb + c_h * nyb
        pen = np.where(inside, vmax(pen_raw, ft(0.0)), ft(0.0))
        nxo = np.where(inside, nxw, ft(0.0))
        nyo = np.where(inside, nyw, ft(0.0))
        return pen.astype(ft), nxo.astype(ft), nyo.astype(ft)

    def pass1(self, wx, wy, sx, sy, sh):
        ft = self.ft
        n, V = wx.shape
        K = self.cfg.n_edge_samples
        if K > 0:
            t = self.t_edge
            omt = (ft(1.0) - t).astype(ft)
            nxv = np.roll(wx, -1, axis=1); nyv = np.roll(wy, -1, axis=1)
            ex = wx[:, :, None] * omt[None, None, :] + nxv[:, :, None] * t[None, None, :]
            ey = wy[:, :, None] * omt[None, None, :] + nyv[:, :, None] * t[None, None, :]
            px = np.concatenate([wx[:, :, None], ex], axis=2).reshape(n, V * (1 + K))
            py = np.concatenate([wy[:, :, None], ey], axis=2).reshape(n, V * (1 + K))
            mask = np.repeat(self.poly_mask, 1 + K, axis=1)
        else:
            px, py, mask = wx, wy, self.poly_mask
        pen, nx, ny = self.point_to_ship(px, py, sx, sy, sh)
        pen = np.where(mask, pen, ft(0.0))
        best = np.argmax(pen, axis=1)
        r = np.arange(n)
        return pen[r, best], nx[r, best], ny[r, best], px[r, best], py[r, best]

    # ── a5: pass 2, jax_geometry.py:289-510 ──────────────────────────────────
    def pass2(self, wx, wy, sx, sy, sh, chunk=4096):
        ft = self.ft
        n, V = wx.shape
        c_h = self.cos(sh); s_h = self.sin(sh)
        hx = self.hull_local[:, 0] * c_h + self.hull_local[:, 1] * (-s_h) + sx
        hy = self.hull_local[:, 0] * s_h + self.hull_local[:, 1] * c_h + sy
        S = hx.shape[0]
        outs = [np.zeros(n, ft) for _ in range(5)]
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            xi = wx[a:b]; yi = wy[a:b]; m = self.poly_mask[a:b]
            xj = np.roll(xi, -1, axis=1); yj = np.roll(yi, -1, axis=1)
            n_real = vmax(seq_sum([m[:, k].astype(ft) for k in range(V)], ft), ft(1.0))
            cx = seq_sum([np.where(m[:, k], xi[:, k], ft(0.0)) for k in range(V)], ft) / n_real
            cy = seq_sum([np.where(m[:, k], yi[:, k], ft(0.0)) for k in range(V)], ft) / n_real
            P_x = hx[None, :, None]; P_y = hy[None, :, None]
            Xi = xi[:, None, :]; Xj = xj[:, None, :]; Yi = yi[:, None, :]; Yj = yj[:, None, :]
            # point_in_polygon_batched (:289-324)
            cond_y = (Yi > P_y) != (Yj > P_y)
            x_int = (Xj - Xi) * (P_y - Yi) / ((Yj - Yi) + ft(1e-10)) + Xi
            crosses = cond_y & (P_x < x_int)
            inside = (np.sum(crosses.astype(np.int32), axis=2) % 2) == 1       # (m, S)
            # distance_point_to_segment_batched (:327-355)
            vx_ = Xj - Xi; vy_ = Yj - Yi
            wx_ = P_x - Xi; wy_ = P_y - Yi
            c1 = wx_ * vx_ + wy_ * vy_
            c2 = (vx_ * vx_ + vy_ * vy_) + ft(1e-20)
            tt = vclip(c1 / c2, ft(0.0), ft(1.0))
            qx = Xi + tt * vx_; qy = Yi + tt * vy_
            ddx = P_x - qx; ddy = P_y - qy
            dist = np.sqrt((ddx * ddx + ddy * ddy) + ft(1e-20))                 # (m, S, V)
            # edge normals (:397-415)
            ex = xj - xi; ey = yj - yi
            nxe = ey; nye = -ex
            nlen = np.sqrt((nxe * nxe + nye * nye) + ft(1e-20))
            nxe = nxe / nlen; nye = nye / nlen
            midx = (xi + xj) / ft(2.0); midy = (yi + yj) / ft(2.0)
            dot = (midx - cx[:, None]) * nxe + (midy - cy[:, None]) * nye
            flip = dot < ft(0.0)
            nxe = np.where(flip, -nxe, nxe); nye = np.where(flip, -nye, nye)
            best_e = np.argmin(dist, axis=2)                                      # (m, S)
            pen_raw = np.take_along_axis(dist, best_e[:, :, None], axis=2)[:, :, 0]
            nxb = np.take_along_axis(nxe, best_e, axis=1)
            nyb = np.take_along_axis(nye, best_e, axis=1)
            pen_s = np.where(inside, pen_raw, ft(0.0))
            nx_s = np.where(inside, nxb, ft(0.0))
            ny_s = np.where(inside, nyb, ft(0.0))
            best_s = np.argmax(pen_s, axis=1)
            r = np.arange(b - a)
            pen_i = pen_s[r, best_s]
            upd = pen_i > ft(0.0)
            outs[0][a:b] = np.where(upd, pen_i, ft(0.0))
            outs[1][a:b] = np.where(upd, nx_s[r, best_s], ft(0.0))
            outs[2][a:b] = np.where(upd, ny_s[r, best_s], ft(0.0))
            outs[3][a:b] = np.where(upd, hx[best_s], ft(0.0))
            outs[4][a:b] = np.where(upd, hy[best_s], ft(0.0))
        return tuple(outs)

    def collisions_combined(self, pos, rot, sx, sy, sh):
        """compute_all_collisions_combined, jax_geometry.py:515-594."""
        ft = self.ft
        sx = ft(sx); sy = ft(sy); sh = ft(sh)
        wx, wy = self.world_polys(pos, rot)
        p1 = self.pass1(wx, wy, sx, sy, sh)
        p2 = self.pass2(wx, wy, sx, sy, sh)
        use2 = p2[0] > p1[0]
        return tuple(np.where(use2, b, a).astype(ft) for a, b in zip(p1, p2))

    # ── a13: update_bond_damage_state, jax_bond_fracture.py:29-208 ───────────
    def strength_criterion(self, sigma, tau, sigma_bar, tau_bar):
        """compute_strength_criterion, jax_bond_fracture.py:35-52."""
        ft = self.ft
        SC = ft(1e12)
        sp = vmax(vclip(sigma, -SC, SC), ft(0.0))
        tau = vclip(tau, -SC, SC)
        sb = vmax(vclip(sigma_bar, ft(1e-10), SC), ft(1e-10))
        tb = vmax(vclip(tau_bar, ft(1e-10), SC), ft(1e-10))
        q1 = sp / sb
        q2 = tau / tb
        return vclip(q1 * q1 + q2 * q2, ft(0.0), ft(1e10))

    def mohr_coulomb(self, C, mu_b, sigma):
        """compute_mohr_coulomb_strength, jax_bond_fracture.py:55-62."""
        ft = self.ft
        return vmax(ft(C) - ft(mu_b) * sigma, ft(1e-6))

    def equivalent_deformation(self, delta_n, delta_s):
        """compute_equivalent_deformation, jax_bond_fracture.py:65-74."""
        ft = self.ft
        dn2 = vclip(delta_n, ft(-1e3), ft(1e3))
        ds2 = vclip(np.abs(delta_s), ft(0.0), ft(1e3))
        dnp = vmax(dn2, ft(0.0))
        return np.sqrt(vclip(dnp * dnp + ds2 * ds2 + ft(1e-20), ft(0.0), ft(1e12)))

    def damage_variable(self, delta_m, delta_m_0, delta_m_f, delta_m_max):
        """compute_damage_variable, jax_bond_fracture.py:77-95."""
        ft = self.ft
        d0 = vclip(delta_m_0, ft(0.0), ft(1e6))
        df = vmax(vclip(delta_m_f, ft(1e-10), ft(1e6)), ft(1e-10))
        dmax = vmax(vclip(delta_m_max, ft(0.0), ft(1e6)), d0 + ft(1e-10))
        denom = dmax * ((df - d0) + ft(1e-10))
        denom = vmax(denom, ft(1e-20))
        D = (df * (dmax - d0)) / denom
        return vclip(D, ft(0.0), ft(1.0))

    def fracture_energy_ratio(self, delta_n, delta_s):
        """compute_fracture_energy_ratio, jax_bond_fracture.py:98-105."""
        ft = self.ft
        dnp = vmax(delta_n, ft(0.0))
        return (delta_s * delta_s) / (dnp * dnp + ft(1e-10))

    def bk_energy(self, G_I, G_II, G_c_I, G_c_II, eta):
        """compute_benzeggagh_kenane_energy, jax_bond_fracture.py:108-119."""
        ft = self.ft
        G_total = (G_I + G_II) + ft(1e-10)
        ratio = G_II / G_total
        return ft(G_c_I) + (ft(G_c_II) - ft(G_c_I)) * self.pow(ratio, eta)

    def critical_deformation(self, G_c, sigma_m_0):
        """compute_critical_deformation, jax_bond_fracture.py:122-128."""
        ft = self.ft
        return (ft(2.0) * G_c) / (np.abs(sigma_m_0) + ft(1e-10))

    def update_bond_damage_state(self, delta_n, delta_s, sigma_n, sigma_t, sigma_bar,
                                 C, mu_b, G_c_I, G_c_II, eta_bk,
                                 damage, dm_max, dm_0, intact):
        """update_bond_damage_state, jax_bond_fracture.py:133-208."""
        ft = self.ft
        SC = ft(1e12)
        delta_n = vclip(np.asarray(delta_n, ft), ft(-1e3), ft(1e3))
        delta_s = vclip(np.asarray(delta_s, ft), ft(-1e3), ft(1e3))
        sigma_bar = vclip(np.asarray(sigma_bar, ft), ft(1e-10), SC)
        sigma_n = np.asarray(sigma_n, ft); sigma_t = np.asarray(sigma_t, ft)
        damage = np.asarray(damage, ft); dm_max = np.asarray(dm_max, ft); dm_0 = np.asarray(dm_0, ft)
        intact = np.asarray(intact, bool)
        delta_m = self.equivalent_deformation(delta_n, delta_s)
        new_dm_max = vmax(dm_max, delta_m)
        tau_bar = self.mohr_coulomb(C, mu_b, sigma_n)
        crit = self.strength_criterion(sigma_n, sigma_t, sigma_bar, tau_bar)
        init = (crit >= ft(1.0)) & (dm_0 <= ft(0.0)) & (delta_m > ft(1e-5))
        new_dm0 = np.where(init, delta_m, dm_0)
        G_ratio = self.fracture_energy_ratio(delta_n, delta_s)
        G_I = ft(1.0) / ((ft(1.0) + G_ratio) + ft(1e-10))
        G_II = G_ratio / ((ft(1.0) + G_ratio) + ft(1e-10))
        G_c = self.bk_energy(G_I, G_II, G_c_I, G_c_II, eta_bk)
        snc = vclip(sigma_n, -SC, SC)
        stc = vclip(sigma_t, -SC, SC)
        mp = vmax(snc, ft(0.0))
        sm0_raw = np.sqrt(vclip(mp * mp + stc * stc + ft(1e-20), ft(0.0), ft(1e24)))
        sm0 = vmax(sm0_raw, ft(1e-10))
        dmf = self.critical_deformation(G_c, sm0)
        meaningful = sm0_raw >= ft(1e4)
        has_damage = (new_dm0 > ft(0.0)) & meaningful
        D = self.damage_variable(delta_m, new_dm0, dmf, new_dm_max)
        computed = np.where(has_damage, D, damage)
        nd = np.where(meaningful, vmax(computed, damage), damage)
        nd = vmin(nd, damage + ft(0.15))
        nintact = intact & (nd < ft(1.0))
        return nd.astype(ft), new_dm_max.astype(ft), new_dm0.astype(ft), nintact

    # ── a14: ship, jax_vessel_dynamics.py:54-348 ─────────────────────────────
    def _mv(self, M, v):
        ft = self.ft
        return np.array([seq_sum([M[i, k] * v[k] for k in range(M.shape[1])], ft)
                         for i in range(M.shape[0])], dtype=ft)

    def _mm(self, A, Bm):
        ft = self.ft
        out = np.zeros((A.shape[0], Bm.shape[1]), ft)
        for i in range(A.shape[0]):
            for j in range(Bm.shape[1]):
                out[i, j] = seq_sum([A[i, k] * Bm[k, j] for k in range(A.shape[1])], ft)
        return out

    def _R(self, psi):
        ft = self.ft
        c = ft(self.cos(np.asarray([psi], dtype=ft))[0]); s = ft(self.sin(np.asarray([psi], dtype=ft))[0])
        return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]], dtype=ft), c, s

    def dp_control(self, eta, nu, e_int, sp):
        """dp_control_jax, jax_vessel_dynamics.py:127-184."""
        ft = self.ft
        R, _, _ = self._R(eta[2])
        RT = R.T.copy()
        e_psi = self.remainder(eta[2] - sp[2] + ft(np.pi), ft(2.0 * np.pi)) - ft(np.pi)
        e = np.array([eta[0] - sp[0], eta[1] - sp[1], e_psi], dtype=ft)
        a = self._mv(self._mm(RT, self.Kp), e)
        b = self._mv(self.Kd, nu)
        c = self._mv(self._mm(RT, self.Ki), e_int)
        tau = ((-a) - b) - c
        e_int_new = e_int + ft(self.dt) * e
        return tau.astype(ft), e_int_new.astype(ft)

    def ship_step(self, eta, nu, n_actual, e_int, sp, spn, tau_ice, v_water_ship, control=None):
        """ship_step_jax, jax_vessel_dynamics.py:246-348.  With ship_mode 1
        (rollout extension, no reference counterpart) the thrust is
        tau = [u X_max, 0, Kd[2,2] (r_cmd r_max - r)], u and r_cmd clipped to
        [-1, 1]; the Euler step and velocity clip are the reference's."""
        ft = self.ft
        dt = ft(self.dt)
        R, c, s = self._R(eta[2])
        RT = R.T.copy()
        nu_c_body = self._mv(RT, np.array([v_water_ship[0], v_water_ship[1], 0.0], dtype=ft))
        tau_ice_body = self._mv(RT, tau_ice)
        if self.ship_mode == 1:
            u = vclip(ft(control[0]), ft(-1.0), ft(1.0))
            rc = vclip(ft(control[1]), ft(-1.0), ft(1.0))
            tau_act = np.array([u * ft(self.ctrl_thrust_max), 0.0,
                                self.Kd[2, 2] * (rc * ft(self.ctrl_yaw_rate_max) - nu[2])], dtype=ft)
            return self._euler_tail(eta, nu, tau_act, tau_ice_body, nu_c_body, c, s,
                                    np.asarray(n_actual, ft), np.asarray(e_int, ft))
        dpsi = self.remainder(spn[2] - sp[2] + ft(np.pi), ft(2.0 * np.pi)) - ft(np.pi)
        v_d = np.array([(spn[0] - sp[0]) / dt, (spn[1] - sp[1]) / dt, dpsi / dt], dtype=ft)
        nu_d = self._mv(RT, v_d)
        tau_ff = self._mv((self.D3 + self.Kd).astype(ft), nu_d)
        tau_c, e_int_new = self.dp_control(eta, nu, e_int, sp)
        if self.use_auto:
            tau3 = tau_c + tau_ff
            u = self._mv(self.B_pinv, tau3)
            n_cmd = vsign(u) * np.sqrt(np.abs(u) + ft(1e-12))
            n_cmd = vclip(n_cmd, -self.n_max, self.n_max)
            n_new = n_actual + dt * (n_cmd - n_actual) / ft(self.T_n)
            tau_act = self._mv(self.B, np.abs(n_new) * n_new)
        else:
            tau_act = tau_c + tau_ff
            n_new = n_actual
        return self._euler_tail(eta, nu, tau_act, tau_ice_body, nu_c_body, c, s, np.asarray(n_new, ft), e_int_new)

    def _euler_tail(self, eta, nu, tau_act, tau_ice_body, nu_c_body, c, s, n_new, e_int_new):
        ft = self.ft
        dt = ft(self.dt)
        tau_total = tau_act + tau_ice_body
        # ship_euler_step_jax (:54-84)
        nu_r = nu - nu_c_body
        nu_dot = self._mv(self.M3inv, tau_total - self._mv(self.D3, nu_r))
        nu_new = nu + dt * nu_dot
        vx = c * nu_new[0] - s * nu_new[1]
        vy = s * nu_new[0] + c * nu_new[1]
        psi_new = self.remainder(eta[2] + dt * nu_new[2] + ft(np.pi), ft(2.0 * np.pi)) - ft(np.pi)
        eta_new = np.array([eta[0] + dt * vx, eta[1] + dt * vy, psi_new], dtype=ft)
        vc = self.cfg.ship_velocity_clip
        nu_new = np.array([vclip(nu_new[0], ft(-vc[0]), ft(vc[0])),
                           vclip(nu_new[1], ft(-vc[1]), ft(vc[1])),
                           vclip(nu_new[2], ft(-vc[2]), ft(vc[2]))], dtype=ft)
        return eta_new, nu_new, np.asarray(n_new, ft), e_int_new

    # ── a3/a6/a7/a8/a13: frame_step, jax_simulation.py:397-699 ───────────────
    def frame_step(self, st: IceState, ship: ShipState, setpoint, setpoint_next, v_water,
                   dpdem_eligible, v_water_ship=None, control=None):
        ft = self.ft
        cfg = self.cfg
        n = st.positions.shape[0]
        if v_water_ship is None:
            v_water_ship = np.zeros(2, ft)
        setpoint = np.asarray(setpoint, ft); setpoint_next = np.asarray(setpoint_next, ft)
        v_water = np.asarray(v_water, ft); v_water_ship = np.asarray(v_water_ship, ft)
        sx, sy, sh = ship.eta[0], ship.eta[1], ship.eta[2]
        pen, nx, ny, cx, cy = self.collisions_combined(st.positions, st.rotations, sx, sy, sh)
        hc, fm, ship_on_ice, tau_ice, new_sub = self.contact_prepass(st.positions, st.submergence, pen, nx, ny,
                                                                     sx, sy)
        F_ship_x, F_ship_y = tau_ice[0], tau_ice[1]
        st = st._replace(submergence=new_sub)
        new_active, collided = self.activation(st.elem_active, hc, dpdem_eligible.shape[0])
        st = st._replace(elem_active=new_active)
        new_elig = self.eligibility(dpdem_eligible, collided, fm)
        # sub-steps (:527-532, :363-395)
        for _ in range(self.n_ice_steps):
            f_bond, m_bond = self.bond_forces(st.positions, st.elem_active, st.bond_intact, st.bond_damage)
            f_ice, m_ice = self.ice_ice_forces(st.positions, st.velocities)
            f_drag = self.drag(st.velocities, v_water)
            m_rd = self.rot_drag(st.angular_vels, st.submergence)
            forces = f_bond + f_ice - f_drag + ship_on_ice
            moments = m_bond + m_ice + m_rd
            st = self.integrate(st, forces, moments)
        # fracture (:534-653)
        if self.bond_i.shape[0] > 0:
            st = self._fracture(st, fm, new_active, new_elig)
        st = st._replace(dpdem_floe_eligible=new_elig)
        # ship (:658-689)
        if cfg.ice_brake_threshold > 0.0 and self.ship_mode == 0:
            thr_b = ft(cfg.ice_brake_threshold)
            surge = np.abs(tau_ice[0])
            bf = vclip(thr_b / vmax(surge, thr_b), ft(0.1), ft(1.0))
            eff = np.array([ship.eta[0] + bf * (setpoint[0] - ship.eta[0]),
                            ship.eta[1] + bf * (setpoint[1] - ship.eta[1]), setpoint[2]], dtype=ft)
        else:
            eff = setpoint
        eta, nu, n_act, e_int = self.ship_step(ship.eta, ship.nu, ship.n_actual, ship.e_int,
                                               eff, setpoint_next, tau_ice, v_water_ship, control)
        ship_new = ShipState(eta, nu, n_act, e_int)
        diag = {
            "pen_max": pen, "force_mag": fm.astype(ft),
            "F_ice_on_ship": np.array([F_ship_x, F_ship_y], dtype=ft),
            "J_cumul_incr": ft(np.sqrt(F_ship_x * F_ship_x + F_ship_y * F_ship_y) * ft(self.dt)),
            "nx": nx, "ny": ny, "ship_on_ice": ship_on_ice, "tau_ice": tau_ice,
        }
        return st, ship_new, diag

    # ── a6: contact pre-pass, jax_simulation.py:431-498 ──────────────────────
    def contact_prepass(self, pos, sub, pen, nx, ny, sx, sy):
        """Force magnitudes, ship forces on the ice, the ship's reaction and
        the submergence update from the hull results.  Returns (has_collision,
        force_mag (thrust-scaled), ship_on_ice (n,2: unscaled partials plus
        buoyancy), tau_ice [Fx, Fy, M] (scaled forces, unscaled moment), new
        submergence)."""
        ft = self.ft
        cfg = self.cfg
        hc = pen > ft(0.0)
        w_c = vmin(ft(cfg.w_contact_radius_factor) * self.radii, ft(cfg.w_contact_cap))
        sigma_f = ft(cfg.flexural_strength_base) * np.sqrt(self.thick)
        F_break = sigma_f * w_c * (self.thick * self.thick)
        pen_crit = ft(cfg.pen_crit_ref) * (self.thick / ft(cfg.pen_crit_thickness_ref))
        fm = F_break * vmin(pen / vmax(pen_crit, ft(1e-10)), ft(1.0))
        fm = np.where(hc, fm, ft(0.0))
        soi_x = fm * nx
        soi_y = fm * ny
        rel_y = pos[:, 1] - sy
        side_push = ft(cfg.side_push_factor) * fm * vsign(rel_y)
        soi_y = soi_y + side_push
        soi_sx, soi_sy = soi_x, soi_y
        if cfg.max_ship_ice_force is not None:
            tx = tree_sum(soi_x, ft); ty = tree_sum(soi_y, ft)
            F_tot = np.sqrt(tx * tx + ty * ty)
            cap = ft(cfg.max_ship_ice_force)
            scale = (cap / vmax(F_tot, ft(1e-6))) if F_tot > cap else ft(1.0)
            scale = ft(scale)
            soi_sx = soi_x * scale; soi_sy = soi_y * scale
            fm = fm * scale
        F_ship_x = -tree_sum(soi_sx, ft)
        F_ship_y = -tree_sum(soi_sy, ft)
        r_x = pos[:, 0] - sx
        r_y = pos[:, 1] - sy
        M_ship = tree_sum(np.where(hc, r_x * (-soi_y) - r_y * (-soi_x), ft(0.0)), ft)
        tau_ice = np.array([F_ship_x, F_ship_y, M_ship], dtype=ft)
        # submergence + buoyancy (:481-498)
        pen_c = vmin(pen, ft(cfg.submergence_pen_cap))
        new_sub = np.where(hc,
                           vmin(sub + ft(cfg.submergence_rate) * pen_c * ft(self.dt), self.thick),
                           vmax(ft(0.0), sub - ft(1.0 / cfg.rise_time_const) * sub * ft(self.dt)))
        z_sub = vmin(new_sub, self.thick)
        F_buoy = ft(WATER_DENSITY * 9.81) * self.areas * z_sub
        soi_x = soi_x + np.where(hc, F_buoy * nx, ft(0.0))
        soi_y = soi_y + np.where(hc, F_buoy * ny, ft(0.0))
        ship_on_ice = np.stack([soi_x, soi_y], axis=1).astype(ft)
        return hc, fm.astype(ft), ship_on_ice, tau_ice, new_sub.astype(ft)

    # ── a6: floe activation, jax_simulation.py:500-509 ───────────────────────
    def activation(self, active, hc, n_floes):
        """Every element of a floe with a hull contact becomes active.
        Returns (new elem_active, collided per floe)."""
        collided = np.zeros(n_floes, np.int32)
        np.maximum.at(collided, self.floe, hc.astype(np.int32))
        collided = collided.astype(bool)
        return active | collided[self.floe], collided

    # ── a7: DPDEM eligibility, jax_simulation.py:511-524 ─────────────────────
    def eligibility(self, elig, collided, fm):
        """A floe becomes eligible when collided and its largest (scaled)
        force magnitude exceeds min(bond strength) x effective strength."""
        ft = self.ft
        if self.bond_i.shape[0] == 0:
            return elig.copy()
        thr = self.bstr.min() * ft(self.cfg.effective_strength_collision)
        mx = np.zeros(elig.shape[0], ft)
        np.maximum.at(mx, self.floe, fm)
        return elig | (collided & (mx > thr))

    def _fracture(self, st: IceState, fm, new_active, new_elig):
        ft = self.ft
        cfg = self.cfg
        n = st.positions.shape[0]
        bi, bj = self.bond_i, self.bond_j
        dx = st.positions[bj, 0] - st.positions[bi, 0]
        dy = st.positions[bj, 1] - st.positions[bi, 1]
        dist = self.norm2(dx, dy)
        ds_ = vmax(dist, ft(1e-10))
        dirx = dx / ds_; diry = dy / ds_
        both = new_active[bi] & new_active[bj]
        floe_ok = new_elig[self.floe[bi]] & new_elig[self.floe[bj]]
        mft = ft(cfg.max_fracture_thickness)
        breakable = (self.thick[bi] <= mft) & (self.thick[bj] <= mft)
        near = (fm[bi] > ft(0.0)) | (fm[bj] > ft(0.0))
        broken = (~st.bond_intact).astype(np.int32)
        crack = np.zeros(n, np.int32)
        np.maximum.at(crack, bi, broken)
        np.maximum.at(crack, bj, broken)
        crack = crack.astype(bool)
        at_crack = crack[bi] | crack[bj]
        dpdem_ok = both & floe_ok & (near | at_crack) & breakable
        vrx = st.velocities[bj, 0] - st.velocities[bi, 0]
        vry = st.velocities[bj, 1] - st.velocities[bi, 1]
        vn = vrx * dirx + vry * diry
        vtx = vrx - vn * dirx; vty = vry - vn * diry
        ds_incr = self.norm2(vtx, vty) * ft(self.dt_ice)
        gate = near & both & floe_ok & breakable
        new_ds = st.bond_delta_s + np.where(gate, ds_incr, ft(0.0))
        L0 = vmax(self.rest, ft(1e-10))
        delta_n = dist - self.rest
        dn_d = np.where(near, delta_n, ft(0.0))
        sn_el = ft(self.k_n) * (dn_d / L0)
        st_el = ft(self.k_s) * (new_ds / L0)
        sigma_n = sn_el * (ft(1.0) - st.bond_damage)
        sigma_t = st_el * (ft(1.0) - st.bond_damage)
        nd, ndmx, ndm0, nint = self.update_bond_damage_state(
            delta_n, new_ds, sigma_n, sigma_t, self.sigma_bar,
            cfg.bond_cohesion, cfg.bond_internal_friction, cfg.bond_G_c_I, cfg.bond_G_c_II,
            cfg.bond_eta_bk, st.bond_damage, st.bond_delta_m_max, st.bond_delta_m_0, st.bond_intact)
        new_damage = np.where(dpdem_ok, nd, st.bond_damage)
        new_dmx = np.where(dpdem_ok, ndmx, st.bond_delta_m_max)
        new_dm0 = np.where(dpdem_ok, ndm0, st.bond_delta_m_0)
        new_int_d = np.where(dpdem_ok & breakable, nint, st.bond_intact)
        coll = (fm[bi] > ft(0.0)) | (fm[bj] > ft(0.0))
        cf = vmax(fm[bi], fm[bj])
        eff = self.bstr * ft(cfg.effective_strength_collision)
        brk = st.bond_intact & coll & (cf > eff) & breakable
        final = new_int_d & ~brk
        return st._replace(bond_damage=new_damage.astype(ft), bond_delta_m_max=new_dmx.astype(ft),
                           bond_delta_m_0=new_dm0.astype(ft), bond_delta_s=new_ds.astype(ft),
                           bond_intact=final)

    # ── a15: scan_segment, jax_simulation.py:1308-1335 ───────────────────────
    def scan_segment(self, st, ship, setpoints_K6, v_water_Kn2, v_water_ship_K2=None):
        K = setpoints_K6.shape[0]
        if v_water_ship_K2 is None:
            v_water_ship_K2 = np.zeros((K, 2), self.ft)
        hist = {k: [] for k in ("positions", "rotations", "bond_intact", "eta", "submergence", "J", "F")}
        for k in range(K):
            st, ship, d = self.frame_step(st, ship, setpoints_K6[k, :3], setpoints_K6[k, 3:],
                                          v_water_Kn2[k], st.dpdem_floe_eligible, v_water_ship_K2[k])
            hist["positions"].append(st.positions.copy()); hist["rotations"].append(st.rotations.copy())
            hist["bond_intact"].append(st.bond_intact.copy()); hist["eta"].append(ship.eta.copy())
            hist["submergence"].append(st.submergence.copy()); hist["J"].append(d["J_cumul_incr"])
            hist["F"].append(d["F_ice_on_ship"].copy())
        return st, ship, tuple(np.stack(hist[k]) for k in
                               ("positions", "rotations", "bond_intact", "eta", "submergence", "J", "F"))


def initial_ice_state(centroids, n_bonds, floe_of, ft=np.float32) -> IceState:
    """build_initial_ice_state, jax_ice_state.py:246-267."""
    n = len(centroids)
    n_floes = int(np.max(floe_of)) + 1 if n > 0 else 0
    return IceState(np.asarray(centroids, ft).copy(), np.zeros((n, 2), ft), np.zeros(n, ft), np.zeros(n, ft),
                    np.zeros(n, ft), np.ones(n_bonds, bool), np.zeros(n_bonds, ft), np.zeros(n_bonds, ft),
                    np.zeros(n_bonds, ft), np.zeros(n_bonds, ft), np.zeros(n, bool), np.zeros(n_floes, bool))


def initial_ship_state(pos, heading, ft=np.float32) -> ShipState:
    """build_initial_ship_state, jax_ice_state.py:344-355."""
    return ShipState(np.array([pos[0], pos[1], heading], ft), np.zeros(3, ft), np.zeros(6, ft), np.zeros(3, ft))


def kinetic_energy(o: Oracle, nu):
    """Ship kinetic energy 0.5 m (u^2 + v^2) + 0.5 Izz r^2 (rollout extension)."""
    ft = o.ft
    return (ft(0.5) * ft(o.ship_mass) * (nu[0] * nu[0] + nu[1] * nu[1])
            + ft(0.5) * ft(o.ship_Izz) * (nu[2] * nu[2]))


def rollout(o: Oracle, st: IceState, ship: ShipState, controls, hold, frames, v_water=None):
    """Direct-control rollout (extension): returns (state, ship, [impulse,
    KE loss, peak |F|, initial KE]) with the device's accumulation order."""
    ft = o.ft
    n = st.positions.shape[0]
    vw = np.zeros((n, 2), ft) if v_water is None else np.asarray(v_water, ft)
    controls = np.asarray(controls, ft)
    ke0 = kinetic_energy(o, ship.nu)
    imp = ft(0.0)
    peak = ft(0.0)
    zero = np.zeros(3, ft)
    for k in range(frames):
        knot = min(k // max(hold, 1), len(controls) - 1)
        st, ship, d = o.frame_step(st, ship, zero, zero, vw, st.dpdem_floe_eligible, None, controls[knot])
        F = d["F_ice_on_ship"]
        Fm = ft(np.sqrt(F[0] * F[0] + F[1] * F[1]))
        imp = ft(imp + Fm * ft(o.dt))
        peak = vmax(peak, Fm)
    ke1 = kinetic_energy(o, ship.nu)
    return st, ship, np.array([imp, ke0 - ke1, peak, ke0], dtype=ft)
