"""Hull-ice collision leaf (``medea_ice_ship/jax_geometry.py:515-594``).

``compute_all_collisions_combined`` keeps the reference's arguments and
returns ``(pen, nx, ny, cx, cy)``, each (n,), computed by the wave-per-element
hull kernel of ``libicedem``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .engine import edge_sample_t


def _dev(x, dtype=torch.float32):
    t = x if torch.is_tensor(x) else torch.as_tensor(np.asarray(x))
    return t.to(device=torch.device("cuda", torch.cuda.current_device()), dtype=dtype).contiguous()


def compute_all_collisions_combined(local_polys_padded, poly_mask, positions, rotations, ship_x, ship_y,
                                    ship_heading, half_len, half_beam, ship_hull_local, n_edge_samples=3):
    if not torch.cuda.is_available():
        raise _lib.IceDemError("icedem needs a ROCm GPU")
    lib = _lib.load()
    lp = _dev(local_polys_padded)
    n, max_v = int(lp.shape[0]), int(lp.shape[1])
    mask = _dev(poly_mask, torch.uint8)
    pos = _dev(torch.as_tensor(np.asarray(positions) if not torch.is_tensor(positions) else positions)
               .reshape(n, 2).t())
    rot = _dev(rotations)
    hull = _dev(ship_hull_local)
    K = int(n_edge_samples)
    t = edge_sample_t(K) if K > 0 else np.zeros(1, np.float32)
    omt = (np.float32(1.0) - t).astype(np.float32)
    outs = [torch.zeros(n, dtype=torch.float32, device=lp.device) for _ in range(5)]
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.icedem_collisions_combined(lp.data_ptr(), mask.data_ptr(), pos.data_ptr(), rot.data_ptr(), n, max_v,
                                        float(np.float32(ship_x)), float(np.float32(ship_y)),
                                        float(np.float32(ship_heading)), float(np.float32(half_len)),
                                        float(np.float32(half_beam)), hull.data_ptr(), int(hull.shape[0]), K,
                                        t.ctypes.data, omt.ctypes.data, *[o.data_ptr() for o in outs], stream)
    _lib.check(rc, None, "icedem_collisions_combined")
    torch.cuda.synchronize()
    return tuple(outs)
