"""Oracle: uniform-grid candidate pairs (cell-list variant of the all-pairs loop).

TEST INFRASTRUCTURE ONLY (the checker; see ``oracle/__init__.py``).

The reference evaluates the full upper triangle (jax_ice_state.py:212-214,
jax_simulation.py:127-131).  This variant returns a *superset* of the contact
pairs -- every (i<j) whose centres fall in the same or adjacent cells of a grid
of side >= 2*max(r) + 2*delta -- in the upper-triangle (row-major) order, so
the pair terms and their scatter order are identical to the all-pairs path.
Pinned by ``tests/test_oracle_contact.py`` (identical contact set and forces as
the all-pairs enumeration on random inputs).
"""
from __future__ import annotations

import numpy as np


def candidate_pairs(pos, radii, dilation=0.0):
    pos = np.asarray(pos, dtype=np.float64)
    n = pos.shape[0]
    if n < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    cell = 2.0 * float(np.max(radii)) * 1.01 + 2.0 * float(dilation) + 1e-6
    cx = np.floor(pos[:, 0] / cell).astype(np.int64)
    cy = np.floor(pos[:, 1] / cell).astype(np.int64)
    cx -= cx.min(); cy -= cy.min()
    W = int(cx.max()) + 3
    key = (cy + 1) * W + (cx + 1)
    order = np.argsort(key, kind="stable")
    skey = key[order]
    uniq, start = np.unique(skey, return_index=True)
    end = np.append(start[1:], n)
    lookup = {int(k): (int(s), int(e)) for k, s, e in zip(uniq, start, end)}
    out_i, out_j = [], []
    for k, s, e in zip(uniq, start, end):
        mine = order[s:e]
        for dyc in (-1, 0, 1):
            for dxc in (-1, 0, 1):
                nb = lookup.get(int(k) + dyc * W + dxc)
                if nb is None:
                    continue
                other = order[nb[0]:nb[1]]
                a = np.repeat(mine, other.shape[0])
                b = np.tile(other, mine.shape[0])
                keep = a < b
                out_i.append(a[keep]); out_j.append(b[keep])
    pi = np.concatenate(out_i); pj = np.concatenate(out_j)
    o = np.lexsort((pj, pi))
    return pi[o], pj[o]
