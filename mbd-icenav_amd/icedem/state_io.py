"""Scenario inputs and simulation results on disk, without pickles.

* ``save_simulation_state`` / ``load_simulation_state``: the replay record of
  ``medea_ice_ship/state_io.py:6-43`` -- the same keys (polygons, centroids,
  bonds, thicknesses, the five histories, ship size, n_frames, dt, the output
  paths and simple_mode) -- written as one ``.npz`` of plain arrays (ragged
  polygons as concatenated vertices plus counts; the ship history as an
  (F, 3) array of x, y, heading) and read back in the reference's Python
  shapes (lists of arrays, ship history as (position, heading) pairs).
* MEDEA floe records ``{"vertices", "centre", "area"}`` (the schema that
  ``medea_ice_ship/data.py:7-21`` unpickles): ``medea_floes_from_records``
  gives ``load_medea_floes``'s (polygons, centroids, metadata);
  ``load_medea_floes`` reads records from JSON or ``.npz`` (never a pickle);
  ``scale_floes_to_world`` is ``data.py:24-47``; ``scene_from_medea`` tessellates
  and bonds the floes into the drivers' inputs.
"""
from __future__ import annotations

import json
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import scenario

STATE_KEYS = ("polygons", "centroids", "bonds", "thicknesses", "positions_history", "rotations_history",
              "bonds_history", "ship_history", "submergence_history", "ship_length", "ship_beam", "ship_draft",
              "n_frames", "dt", "output_2d", "output_3d", "simple_mode")


def _ragged(polys) -> Dict[str, np.ndarray]:
    nv = np.array([len(p) for p in polys], np.int64)
    verts = np.concatenate([np.asarray(p, np.float64).reshape(-1, 2) for p in polys]) if len(nv) else np.zeros((0, 2))
    return {"verts": verts, "nv": nv}


def _unragged(verts: np.ndarray, nv: np.ndarray) -> List[np.ndarray]:
    offs = np.concatenate([[0], np.cumsum(nv)])
    return [verts[offs[i]:offs[i + 1]].copy() for i in range(len(nv))]


def _stack(frames, dtype) -> np.ndarray:
    frames = [np.asarray(f.detach().cpu().numpy() if hasattr(f, "detach") else f, dtype=dtype) for f in frames]
    return np.stack(frames) if frames else np.zeros((0,), dtype)


def save_simulation_state(polygons, centroids, bonds, thicknesses,
                          positions_history, rotations_history, bonds_history, ship_history, submergence_history,
                          ship_length, ship_beam, ship_draft, n_frames, dt,
                          output_2d=None, output_3d=None, simple_mode=False, filepath=None):
    """Write the replay record (state_io.py:6-35) to ``filepath`` (.npz)."""
    if filepath is None:
        raise ValueError("filepath is required")
    rag = _ragged(polygons)
    ship = np.array([[float(p[0]), float(p[1]), float(h)] for p, h in ship_history], np.float64).reshape(-1, 3)
    meta = {"ship_length": float(ship_length), "ship_beam": float(ship_beam), "ship_draft": float(ship_draft),
            "n_frames": int(n_frames), "dt": float(dt), "output_2d": output_2d, "output_3d": output_3d,
            "simple_mode": bool(simple_mode)}
    np.savez(filepath, poly_verts=rag["verts"], poly_nv=rag["nv"],
             centroids=np.asarray(centroids, np.float64).reshape(-1, 2),
             bonds=np.asarray(bonds, np.int64).reshape(-1, 2), thicknesses=np.asarray(thicknesses, np.float64),
             positions_history=_stack(positions_history, np.float32),
             rotations_history=_stack(rotations_history, np.float32),
             bonds_history=_stack(bonds_history, bool), ship_history=ship,
             submergence_history=_stack(submergence_history, np.float32),
             meta=np.array(json.dumps(meta)))
    print(f"  Saved state to {filepath}", flush=True)


def load_simulation_state(filepath) -> dict:
    """Read a record written by save_simulation_state (state_io.py:38-43):
    a dict with the reference's keys and shapes."""
    z = np.load(filepath, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    state = {
        "polygons": _unragged(z["poly_verts"], z["poly_nv"]),
        "centroids": z["centroids"],
        "bonds": [(int(i), int(j)) for i, j in z["bonds"]],
        "thicknesses": z["thicknesses"],
        "positions_history": list(z["positions_history"]),
        "rotations_history": list(z["rotations_history"]),
        "bonds_history": list(z["bonds_history"]),
        "ship_history": [(row[:2].copy(), float(row[2])) for row in z["ship_history"]],
        "submergence_history": list(z["submergence_history"]),
    }
    state.update(meta)
    print(f"  Loaded state from {filepath} ({state['n_frames']} frames)", flush=True)
    return state


# ── MEDEA floe records ──────────────────────────────────────────────────────

def medea_floes_from_records(records: Sequence[dict]):
    """(polygons, centroids, metadata) of MEDEA floe records
    {"vertices": (k, 2), "centre": (2,), "area": float}, as
    load_medea_floes returns them (data.py:7-21)."""
    polygons = [np.array(r["vertices"], dtype=np.float64) for r in records]
    centroids = np.array([np.array(r["centre"], dtype=np.float64) for r in records]).reshape(-1, 2)
    areas = np.array([float(r["area"]) for r in records])
    return polygons, centroids, {"n_floes": len(polygons), "areas": areas, "source": "MEDEA"}


def save_medea_floes(records: Sequence[dict], path) -> None:
    """Write floe records as JSON (.json) or arrays (.npz)."""
    path = Path(path)
    if path.suffix == ".json":
        path.write_text(json.dumps([{"vertices": np.asarray(r["vertices"], float).tolist(),
                                     "centre": np.asarray(r["centre"], float).tolist(),
                                     "area": float(r["area"])} for r in records]))
        return
    polys, cents, meta = medea_floes_from_records(records)
    rag = _ragged(polys)
    np.savez(path, verts=rag["verts"], nv=rag["nv"], centre=cents, area=meta["areas"])


def load_medea_floes(path):
    """load_medea_floes (data.py:7-21) for records stored as JSON or .npz
    (the reference's pickles are not read: unpickling executes code)."""
    path = Path(path)
    if path.suffix == ".json":
        return medea_floes_from_records(json.loads(path.read_text()))
    if path.suffix == ".npz":
        z = np.load(path, allow_pickle=False)
        polys = _unragged(z["verts"], z["nv"])
        return medea_floes_from_records([{"vertices": p, "centre": c, "area": a}
                                         for p, c, a in zip(polys, z["centre"], z["area"])])
    raise ValueError(f"{path}: floe records are read from .json or .npz only")


def scale_floes_to_world(polygons, centroids, meters_per_pixel, center_domain=True, subsample_region=None):
    """Pixel to world coordinates (data.py:24-47): scale, optionally centre
    the vertices' bounding box on the origin, optionally keep the floes whose
    centroid lies in (x_min, x_max, y_min, y_max) (all of them if none does)."""
    scaled_polygons = [np.asarray(p, np.float64) * meters_per_pixel for p in polygons]
    scaled_centroids = np.asarray(centroids, np.float64) * meters_per_pixel
    if center_domain and scaled_polygons:
        allv = np.vstack(scaled_polygons)
        center = (allv.min(axis=0) + allv.max(axis=0)) / 2
        scaled_polygons = [p - center for p in scaled_polygons]
        scaled_centroids = scaled_centroids - center
    if subsample_region is not None:
        x_min, x_max, y_min, y_max = subsample_region
        keep = [i for i, c in enumerate(scaled_centroids) if x_min <= c[0] <= x_max and y_min <= c[1] <= y_max]
        if keep:
            scaled_polygons = [scaled_polygons[i] for i in keep]
            scaled_centroids = scaled_centroids[keep]
    return scaled_polygons, scaled_centroids


def scene_from_medea(records: Sequence[dict], meters_per_pixel: float = 1.0, thickness=1.0,
                     cells_per_floe: int = 20, seed: int = 0, center_domain: bool = True,
                     subsample_region=None, bond_factor: float = 1.5) -> scenario.Scene:
    """Drivers' inputs from MEDEA floe records: floes scaled to metres, each
    Voronoi-tessellated into about ``cells_per_floe`` convex elements (cell
    centroid = vertex mean), proximity bonds |c_i - c_j| < factor (r_i + r_j)
    (bonds.py:6-16), per-floe ``thickness`` (scalar or one per floe).  A
    floe given clockwise is reversed to the counter-clockwise order the
    tessellation clips with."""
    polys, cents, _ = medea_floes_from_records(records)
    polys, cents = scale_floes_to_world(polys, cents, meters_per_pixel, center_domain, subsample_region)
    floes = []
    for p in polys:
        p = np.asarray(p, np.float64)
        a2 = np.sum(p[:, 0] * np.roll(p[:, 1], -1) - np.roll(p[:, 0], -1) * p[:, 1])
        floes.append(p[::-1].copy() if a2 < 0 else p)
    th = np.broadcast_to(np.asarray(thickness, np.float64), (len(floes),)).copy()
    rng = np.random.default_rng(seed)
    cells, ccent, parents, thick, bonds = scenario.tessellate_all(rng, floes, cells_per_floe, th)
    if bond_factor != 1.5:
        areas = np.array([scenario.polygon_area(c) for c in cells])
        bonds = scenario.proximity_bonds(ccent, np.sqrt(areas / np.pi), bond_factor)
    allv = np.vstack(floes) if floes else np.zeros((1, 2))
    lo, hi = allv.min(axis=0), allv.max(axis=0)
    area = float((hi[0] - lo[0]) * (hi[1] - lo[1])) or 1.0
    conc = float(sum(scenario.polygon_area(f) for f in floes)) / area
    return scenario.Scene(cells, ccent, bonds, thick, parents, len(floes), conc,
                          (float(lo[0]), float(hi[0]), float(lo[1]), float(hi[1])))
