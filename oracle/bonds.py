"""Oracle: connected components of the bond graph and component COM velocity.

TEST INFRASTRUCTURE ONLY (the checker; see ``oracle/__init__.py``).

Restates ``medea_ice_ship/jax_bonds.py:37-141`` (label propagation, labels =
minimum element index after ``n_iter`` scatter-min sweeps).  Pinned by the
reference's own cases (``test_jax_full_scan.py:65-182``: chain, star,
disconnected, single node, broken bonds, 20-chain, COM velocities [2,0]/[0,2]),
restated in ``tests/test_oracle_bonds.py``, and by BFS on random graphs.
"""
from __future__ import annotations

import numpy as np


def find_connected_components(n, bond_i, bond_j, bond_intact, n_iter=64):
    labels = np.arange(n, dtype=np.int32)
    bi = np.asarray(bond_i, np.int64); bj = np.asarray(bond_j, np.int64)
    intact = np.asarray(bond_intact, bool)
    for _ in range(n_iter):
        li = labels[bi]; lj = labels[bj]
        m = np.minimum(li, lj)
        m = np.where(intact, m, n).astype(np.int32)
        new = labels.copy()
        np.minimum.at(new, bi, m)
        np.minimum.at(new, bj, m)
        labels = new
    return labels


def component_com_velocities(labels, masses, velocities, ft=np.float64):
    labels = np.asarray(labels, np.int64)
    masses = np.asarray(masses, ft); vel = np.asarray(velocities, ft)
    n = masses.shape[0]
    px = masses * vel[:, 0]; py = masses * vel[:, 1]
    tm = np.zeros(n, ft); tx = np.zeros(n, ft); ty = np.zeros(n, ft)
    np.add.at(tm, labels, masses); np.add.at(tx, labels, px); np.add.at(ty, labels, py)
    sm = np.where(np.isnan(tm) | (tm >= ft(1e-30)), tm, ft(1e-30))
    return np.stack([(tx / sm)[labels], (ty / sm)[labels]], axis=1)


def bfs_components(n, bonds, intact):
    """Breadth-first partition, bonds.py:29-52 (used to cross-check labels)."""
    adj = {i: [] for i in range(n)}
    for b, (i, j) in enumerate(bonds):
        if intact[b]:
            adj[i].append(j); adj[j].append(i)
    seen, comps = set(), []
    for s in range(n):
        if s in seen:
            continue
        comp, q = [], [s]
        while q:
            u = q.pop(0)
            if u in seen:
                continue
            seen.add(u); comp.append(u)
            q.extend(v for v in adj[u] if v not in seen)
        comps.append(comp)
    return comps
