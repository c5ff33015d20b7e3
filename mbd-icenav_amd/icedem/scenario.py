"""Seeded synthetic ice scenes (host side, one-time input construction).

The reference builds scenes with shapely and an O(n^2) Python bond loop
(``medea_ice_ship/ice_field_generator.py:126-254``, ``tessellation.py:7-123``,
``bonds.py:6-16``).  This module is a shapely-free, seeded equivalent:

* convex floes: vertices on the circumcircle at jittered, sorted angles;
* non-overlap by the separating-axis test against already placed floes
  (spatial hash on circumcircles);
* tessellation: Voronoi cells of seeded interior points (scipy), clipped to
  the convex floe by Sutherland-Hodgman; cell centroid = vertex mean
  (``geometry.py:17-18``);
* proximity bonds: pairs (i<j) with ``|c_i - c_j| < 1.5 (r_i + r_j)``,
  r = sqrt(A/pi), the criterion of ``bonds.py:6-16``, found with a KD-tree and
  returned in the reference's lexicographic order.

The returned dict carries the reference driver's inputs:
``polygons, centroids, bonds, thicknesses, parent_ids``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np


def polygon_area(verts) -> float:
    """Shoelace with the reference's sequential accumulation (geometry.py:6-14)."""
    n = len(verts)
    area = 0
    for i in range(n):
        j = (i + 1) % n
        area += verts[i, 0] * verts[j, 1]
        area -= verts[j, 0] * verts[i, 1]
    return abs(area) / 2


def convex_polygon(rng, center, radius, n_verts, jitter=0.15) -> np.ndarray:
    """Convex polygon with vertices on a circle at jittered sorted angles (CCW)."""
    base = 2.0 * np.pi * np.arange(n_verts) / n_verts
    ang = base + rng.uniform(-jitter, jitter, size=n_verts)
    ang = np.sort(ang)
    pts = np.stack([center[0] + radius * np.cos(ang), center[1] + radius * np.sin(ang)], axis=1)
    return pts


def _sat_overlap(a: np.ndarray, b: np.ndarray, gap: float) -> bool:
    """Separating-axis test for two convex polygons, with a minimum gap."""
    for poly in (a, b):
        e = np.roll(poly, -1, axis=0) - poly
        normals = np.stack([e[:, 1], -e[:, 0]], axis=1)
        normals /= np.maximum(np.linalg.norm(normals, axis=1, keepdims=True), 1e-12)
        pa = a @ normals.T
        pb = b @ normals.T
        sep = (pa.min(axis=0) > pb.max(axis=0) + gap) | (pb.min(axis=0) > pa.max(axis=0) + gap)
        if np.any(sep):
            return False
    return True


def clip_convex(subject: np.ndarray, clip: np.ndarray) -> Optional[np.ndarray]:
    """Sutherland-Hodgman: subject polygon clipped by a convex CCW clip polygon."""
    out = subject
    m = len(clip)
    for k in range(m):
        if len(out) == 0:
            return None
        a = clip[k]
        b = clip[(k + 1) % m]
        ex, ey = b[0] - a[0], b[1] - a[1]
        side = (ex * (out[:, 1] - a[1]) - ey * (out[:, 0] - a[0])).tolist()   # >= 0 inside (left of edge)
        pts = out.tolist()
        res = []
        n = len(pts)
        for i in range(n):
            p, q = pts[i], pts[(i + 1) % n]
            sp, sq = side[i], side[(i + 1) % n]
            if sp >= 0:
                res.append(p)
            if (sp >= 0) != (sq >= 0):
                t = sp / (sp - sq)
                res.append([p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])])
        out = np.array(res, dtype=np.float64) if res else np.zeros((0, 2))
    if len(out) < 3:
        return None
    # drop near-duplicate consecutive vertices
    keep = np.linalg.norm(out - np.roll(out, -1, axis=0), axis=1) > 1e-7
    out = out[keep]
    return out if len(out) >= 3 else None


def tessellate_floe(rng, floe: np.ndarray, n_cells: int) -> List[np.ndarray]:
    """Voronoi cells of ~n_cells seeded interior points, clipped to the floe."""
    from scipy.spatial import Voronoi

    area = polygon_area(floe)
    lo, hi = floe.min(axis=0), floe.max(axis=0)
    spacing = 0.5 * math.sqrt(area / max(n_cells, 1))
    e = np.roll(floe, -1, axis=0) - floe
    buf = np.empty((n_cells, 2))
    npts = 0
    attempts = 0
    while npts < n_cells and attempts < 60 * n_cells:
        attempts += 1
        p = rng.uniform(lo, hi)
        side = e[:, 0] * (p[1] - floe[:, 1]) - e[:, 1] * (p[0] - floe[:, 0])
        if np.any(side < 0):
            continue
        if npts and np.min(np.sum((buf[:npts] - p) ** 2, axis=1)) < spacing * spacing:
            continue
        buf[npts] = p
        npts += 1
    if npts < 3:
        return [floe.copy()]
    P = buf[:npts].copy()
    c = floe.mean(axis=0)
    R = float(np.max(np.linalg.norm(floe - c, axis=1)))
    far = c + 20.0 * R * np.array([[1, 1], [1, -1], [-1, 1], [-1, -1], [1, 0], [-1, 0], [0, 1], [0, -1]], float)
    vor = Voronoi(np.vstack([P, far]))
    cells = []
    for i in range(len(P)):
        reg = vor.regions[vor.point_region[i]]
        if -1 in reg or len(reg) < 3:
            continue
        cell = vor.vertices[reg]
        cc = cell.mean(axis=0)
        ang = np.arctan2(cell[:, 1] - cc[1], cell[:, 0] - cc[0])
        cell = cell[np.argsort(ang)]
        clipped = clip_convex(cell, floe)
        if clipped is not None and polygon_area(clipped) > 1e-6 * area:
            cells.append(clipped)
    return cells if cells else [floe.copy()]


def proximity_bonds(centroids: np.ndarray, radii: np.ndarray, factor: float = 1.5) -> List[Tuple[int, int]]:
    """bonds.py:6-16 criterion (dist < (r_i + r_j) * factor), KD-tree accelerated."""
    from scipy.spatial import cKDTree

    n = len(centroids)
    if n < 2:
        return []
    tree = cKDTree(centroids)
    cand = tree.query_pairs(r=2.0 * float(np.max(radii)) * factor * (1 + 1e-9), output_type="ndarray")
    if len(cand) == 0:
        return []
    cand = np.sort(cand, axis=1)
    cand = cand[np.lexsort((cand[:, 1], cand[:, 0]))]
    d = centroids[cand[:, 0]] - centroids[cand[:, 1]]
    dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])
    keep = dist < (radii[cand[:, 0]] + radii[cand[:, 1]]) * factor
    return [(int(i), int(j)) for i, j in cand[keep]]


@dataclass
class Scene:
    polygons: List[np.ndarray]
    centroids: np.ndarray
    bonds: List[Tuple[int, int]]
    thicknesses: np.ndarray
    parent_ids: np.ndarray
    n_floes: int
    concentration: float
    field: Tuple[float, float, float, float]

    def as_dict(self):
        return dict(polygons=self.polygons, centroids=self.centroids, bonds=self.bonds,
                    thicknesses=self.thicknesses, parent_ids=self.parent_ids)


def place_floes(rng, x_lo, x_hi, y_lo, y_hi, concentration, max_floes, radius_fn, nverts_fn,
                jitter=0.15, gap=0.05, max_attempts=200000):
    """Random sequential placement of non-overlapping convex floes."""
    total = (x_hi - x_lo) * (y_hi - y_lo)
    target = concentration * total
    placed: List[np.ndarray] = []
    radii: List[float] = []
    centers: List[np.ndarray] = []
    grid = {}
    cell = None
    area_sum = 0.0
    attempts = 0
    while area_sum < target and len(placed) < max_floes and attempts < max_attempts:
        attempts += 1
        R = float(radius_fn(rng))
        if 2 * R >= min(x_hi - x_lo, y_hi - y_lo):
            continue
        c = np.array([rng.uniform(x_lo + R, x_hi - R), rng.uniform(y_lo + R, y_hi - R)])
        poly = convex_polygon(rng, c, R, int(nverts_fn(rng)), jitter)
        if cell is None:
            cell = 4.0 * R
        gx, gy = int(math.floor(c[0] / cell)), int(math.floor(c[1] / cell))
        reach = int(math.ceil((R + 3.0 * max(radii + [R])) / cell))
        ok = True
        for ix in range(gx - reach, gx + reach + 1):
            if not ok:
                break
            for iy in range(gy - reach, gy + reach + 1):
                for k in grid.get((ix, iy), ()):
                    if np.hypot(*(centers[k] - c)) > radii[k] + R + gap:
                        continue
                    if _sat_overlap(placed[k], poly, gap):
                        ok = False
                        break
                if not ok:
                    break
        if not ok:
            continue
        k = len(placed)
        placed.append(poly); radii.append(R); centers.append(c)
        grid.setdefault((gx, gy), []).append(k)
        area_sum += polygon_area(poly)
    # spatial order (by centre y, then x): consecutive elements are neighbours
    # in space, which keeps the GPU's per-XCD caches on one band of the field
    order = sorted(range(len(placed)), key=lambda i: (float(centers[i][1]), float(centers[i][0])))
    return [placed[i] for i in order], area_sum / total


def place_floes_sized(rng, x_lo, x_hi, y_lo, y_hi, concentration, n_floes, median_radius, sigma, nv,
                      jitter=0.3, gap=0.05, tries=300, shrink=0.93):
    """Exactly ``n_floes`` non-overlapping convex floes covering about
    ``concentration`` of the field.

    Shapes are drawn first (lognormal circumradius, 6-12 jittered vertices),
    then all are scaled by one factor so that their areas sum to the target;
    they are placed largest first by random sequential placement (large floes
    find room while the field is empty, small ones fill the gaps), and a floe
    that finds no free spot in ``tries`` draws is shrunk by ``shrink`` and
    tried again.  Returns the floes and the achieved concentration.
    """
    total = (x_hi - x_lo) * (y_hi - y_lo)
    radii = np.array([min(max(rng.lognormal(math.log(median_radius), sigma), 2.0), 6.0 * median_radius)
                      for _ in range(n_floes)])
    shapes = [convex_polygon(rng, (0.0, 0.0), R, int(rng.integers(nv[0], nv[1] + 1)), jitter) for R in radii]
    areas = np.array([polygon_area(s) for s in shapes])
    scale = math.sqrt(concentration * total / float(areas.sum()))
    shapes = [s * scale for s in shapes]
    radii = radii * scale
    order = np.argsort(-radii, kind="stable")
    cell = 2.0 * median_radius * scale
    grid = {}
    placed, centers, prad = [], [], []

    def cells_of(c, R):
        return [(ix, iy) for ix in range(int(math.floor((c[0] - R) / cell)), int(math.floor((c[0] + R) / cell)) + 1)
                for iy in range(int(math.floor((c[1] - R) / cell)), int(math.floor((c[1] + R) / cell)) + 1)]

    for k in order:
        shape, R = shapes[k], float(radii[k])
        while True:
            ok_spot = None
            for _ in range(tries):
                c = np.array([rng.uniform(x_lo + R, x_hi - R), rng.uniform(y_lo + R, y_hi - R)])
                poly = shape + c
                seen = set()
                ok = True
                for key in cells_of(c, R + gap):
                    for m in grid.get(key, ()):
                        if m in seen:
                            continue
                        seen.add(m)
                        if np.hypot(*(centers[m] - c)) > prad[m] + R + gap:
                            continue
                        if _sat_overlap(placed[m], poly, gap):
                            ok = False
                            break
                    if not ok:
                        break
                if ok:
                    ok_spot = (c, poly)
                    break
            if ok_spot is not None:
                break
            shape = shape * shrink
            R *= shrink
        c, poly = ok_spot
        m = len(placed)
        placed.append(poly); centers.append(c); prad.append(R)
        for key in cells_of(c, R):
            grid.setdefault(key, []).append(m)
    area_sum = float(sum(polygon_area(p) for p in placed))
    order = sorted(range(len(placed)), key=lambda i: (float(centers[i][1]), float(centers[i][0])))
    return [placed[i] for i in order], area_sum / total


def tessellate_all(rng, floes, cells_per_floe, floe_thickness):
    polys, cents, parents, thick = [], [], [], []
    for f, floe in enumerate(floes):
        cells = tessellate_floe(rng, floe, cells_per_floe)
        for cpoly in cells:
            polys.append(cpoly)
            cents.append(np.mean(cpoly, axis=0))
            parents.append(f)
            thick.append(floe_thickness[f])
    cents = np.array(cents)
    areas = np.array([polygon_area(p) for p in polys])
    radii = np.sqrt(areas / np.pi)
    bonds = proximity_bonds(cents, radii, 1.5)
    return polys, cents, np.array(parents, np.int32), np.array(thick, np.float64), bonds


def scene_c0(seed: int = 0) -> Scene:
    """3 convex 8-gons, circumradius U[10,25] m, in a 150 x 300 m field, ~20 cells each."""
    rng = np.random.default_rng(seed)
    x_lo, x_hi, y_lo, y_hi = -75.0, 75.0, 0.0, 300.0
    floes, conc = place_floes(rng, x_lo, x_hi, y_lo, y_hi, 1.0, 3,
                              lambda r: r.uniform(10.0, 25.0), lambda r: 8, jitter=0.15, gap=1.0)
    # keep the three floes in a column in front of the ship: shift so their
    # x-centres are near 0 is not required; the ship is aimed at the first floe
    polys, cents, parents, thick, bonds = tessellate_all(rng, floes, 20, np.ones(len(floes)))
    return Scene(polys, cents, bonds, thick, parents, len(floes), conc, (x_lo, x_hi, y_lo, y_hi))


def scene_field(seed: int, width: float, length: float, concentration: float, max_floes: int,
                median_radius: float = 8.0, sigma: float = 0.5, nv=(6, 12), cells_per_floe: int = 20,
                thickness_values=(1.0,), thickness_probs=None) -> Scene:
    rng = np.random.default_rng(seed)
    x_lo, x_hi, y_lo, y_hi = -width / 2, width / 2, 0.0, length
    floes, conc = place_floes(
        rng, x_lo, x_hi, y_lo, y_hi, concentration, max_floes,
        lambda r: min(max(r.lognormal(math.log(median_radius), sigma), 2.0), 6.0 * median_radius),
        lambda r: r.integers(nv[0], nv[1] + 1), jitter=0.3)
    th = rng.choice(np.asarray(thickness_values, float), size=len(floes), p=thickness_probs)
    polys, cents, parents, thick, bonds = tessellate_all(rng, floes, cells_per_floe, th)
    return Scene(polys, cents, bonds, thick, parents, len(floes), conc, (x_lo, x_hi, y_lo, y_hi))


def scene_c1(seed: int = 0) -> Scene:
    """300 x 750 m at 30 % (or 400 floes), lognormal radius (median 8 m, sigma 0.5)."""
    return scene_field(seed, 300.0, 750.0, 0.30, 400)


def scene_c2(seed: int = 0) -> Scene:
    """1.5 x 3 km at 50 % with exactly 5000 floes (lognormal circumradius,
    sigma 0.5, scaled to the concentration), ~20 cells each, bimodal
    thickness 0.5 / 2.0 m."""
    rng = np.random.default_rng(seed)
    x_lo, x_hi, y_lo, y_hi = -750.0, 750.0, 0.0, 3000.0
    floes, conc = place_floes_sized(rng, x_lo, x_hi, y_lo, y_hi, 0.50, 5000, 11.0, 0.5, (6, 12))
    th = rng.choice(np.array([0.5, 2.0]), size=len(floes), p=(0.5, 0.5))
    polys, cents, parents, thick, bonds = tessellate_all(rng, floes, 20, th)
    return Scene(polys, cents, bonds, thick, parents, len(floes), conc, (x_lo, x_hi, y_lo, y_hi))


def scene_patch(seed: int, n_target: int, concentration: float) -> Scene:
    """Rollout patch with ~n_target elements at the given concentration."""
    rng = np.random.default_rng(seed)
    cells_per_floe = 12
    n_floes = max(1, int(round(n_target / cells_per_floe)))
    mean_area = 2.4 * 64.0 * math.exp(2 * 0.25)
    side = math.sqrt(n_floes * mean_area / concentration)
    width = max(side, 70.0)
    length = n_floes * mean_area / concentration / width
    return scene_field(seed, width, max(length, 40.0), concentration, n_floes,
                       cells_per_floe=cells_per_floe)


def hexagon_scene():
    """The reference's shared 3-floe test scene (test_jax_full_scan.py:187-230)."""
    angles = np.linspace(0, 2 * np.pi, 7)[:-1]
    r = 3.0
    hex_pts = np.column_stack([r * np.cos(angles), r * np.sin(angles)])
    centroids = np.array([[-10.0, 0.0], [0.0, 0.0], [10.0, 0.0]])
    polygons = [hex_pts + c for c in centroids]
    return Scene(polygons, centroids, [(0, 1), (1, 2)], np.array([1.0, 1.0, 1.0]),
                 np.arange(3, dtype=np.int32), 3, 0.0, (-20.0, 20.0, -10.0, 10.0))


def straight_trajectory(x0, y0, heading, speed, dt, n_frames):
    """(n_frames + 2, 3) setpoints [x, y, psi] moving along ``heading``."""
    t = np.arange(n_frames + 2) * dt
    return np.stack([x0 + speed * t * math.cos(heading), y0 + speed * t * math.sin(heading),
                     np.full_like(t, heading)], axis=1)


def ship_start_below(scene: Scene, ship_length: float, gap: float = 1.0, target_floe: Optional[int] = None):
    """Ship centre so that its bow (+x body, heading +y) is ``gap`` m short of the
    lowest ice vertex of ``target_floe`` (default: the lowest floe)."""
    if target_floe is None:
        lows = np.full(scene.n_floes, np.inf)
        np.minimum.at(lows, np.asarray(scene.parent_ids), np.array([p[:, 1].min() for p in scene.polygons]))
        target_floe = int(np.argmin(lows))
    sel = [p for p, f in zip(scene.polygons, scene.parent_ids) if f == target_floe]
    allv = np.vstack(sel)
    x0 = float(allv[:, 0].mean())
    y0 = float(allv[:, 1].min()) - gap - ship_length / 2.0
    return x0, y0


def save_scene(path, scene: Scene) -> None:
    """Write a scene to an .npz file (vertices concatenated with per-polygon
    offsets; no pickled objects)."""
    nv = np.array([len(p) for p in scene.polygons], np.int64)
    verts = np.concatenate([np.asarray(p, np.float64) for p in scene.polygons]) if len(nv) else np.zeros((0, 2))
    np.savez(path, verts=verts, nv=nv, centroids=np.asarray(scene.centroids, np.float64),
             bonds=np.asarray(scene.bonds, np.int64).reshape(-1, 2), thicknesses=np.asarray(scene.thicknesses),
             parent_ids=np.asarray(scene.parent_ids, np.int32), n_floes=np.int64(scene.n_floes),
             concentration=np.float64(scene.concentration), field=np.asarray(scene.field, np.float64))


def load_scene(path) -> Scene:
    """Read a scene written by save_scene."""
    z = np.load(path, allow_pickle=False)
    offs = np.concatenate([[0], np.cumsum(z["nv"])])
    verts = z["verts"]
    polys = [verts[offs[i]:offs[i + 1]].copy() for i in range(len(z["nv"]))]
    bonds = [(int(i), int(j)) for i, j in z["bonds"]]
    return Scene(polys, z["centroids"], bonds, z["thicknesses"], z["parent_ids"], int(z["n_floes"]),
                 float(z["concentration"]), tuple(float(v) for v in z["field"]))
