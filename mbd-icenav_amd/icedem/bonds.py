"""Bond-graph leaves (``medea_ice_ship/jax_bonds.py:37-141``)."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def _dev(x, dtype):
    t = x if torch.is_tensor(x) else torch.as_tensor(np.asarray(x))
    return t.to(device=torch.device("cuda", torch.cuda.current_device()), dtype=dtype).contiguous()


def find_connected_components_jax(n, bond_i, bond_j, bond_intact, n_iter=64):
    """Label propagation: label = minimum element index of the component
    after ``n_iter`` sweeps (jax_bonds.py:37-89)."""
    if not torch.cuda.is_available():
        raise _lib.IceDemError("icedem needs a ROCm GPU")
    lib = _lib.load()
    bi = _dev(bond_i, torch.int32); bj = _dev(bond_j, torch.int32); it = _dev(bond_intact, torch.uint8)
    labels = torch.zeros(int(n), dtype=torch.int32, device=bi.device if bi.numel() else torch.device("cuda"))
    nb = int(bi.numel())
    rc = lib.icedem_connected_components(int(n), nb, bi.data_ptr() if nb else None, bj.data_ptr() if nb else None,
                                         it.data_ptr() if nb else None, int(n_iter), labels.data_ptr(),
                                         torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, None, "icedem_connected_components")
    return labels


def component_com_velocities_jax(labels, masses, velocities):
    """Mass-weighted component velocity broadcast to elements (jax_bonds.py:96-141)."""
    if not torch.cuda.is_available():
        raise _lib.IceDemError("icedem needs a ROCm GPU")
    lib = _lib.load()
    lab = _dev(labels, torch.int32)
    m = _dev(masses, torch.float32)
    n = int(m.numel())
    v = _dev(torch.as_tensor(np.asarray(velocities) if not torch.is_tensor(velocities) else velocities)
             .reshape(n, 2).t(), torch.float32)
    out = torch.zeros((2, n), dtype=torch.float32, device=m.device)
    rc = lib.icedem_component_com_velocities(n, lab.data_ptr(), m.data_ptr(), v.data_ptr(), out.data_ptr(),
                                             torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, None, "icedem_component_com_velocities")
    torch.cuda.synchronize()
    return out.t()


find_connected_components = find_connected_components_jax
component_com_velocities = component_com_velocities_jax
