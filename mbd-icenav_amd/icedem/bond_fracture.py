"""DPDEM bond damage leaf (``medea_ice_ship/jax_bond_fracture.py:133-208``)."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def _dev(x, dtype=torch.float32):
    t = x if torch.is_tensor(x) else torch.as_tensor(np.asarray(x))
    return t.to(device=torch.device("cuda", torch.cuda.current_device()), dtype=dtype).contiguous()


def update_bond_damage_state(delta_n, delta_s, sigma_n, sigma_t, sigma_bar, C, mu_b, G_c_I, G_c_II, eta_bk,
                             current_damage, current_delta_m_max, current_delta_m_0, intact):
    """Returns (new_damage, new_delta_m_max, new_delta_m_0, new_intact)."""
    if not torch.cuda.is_available():
        raise _lib.IceDemError("icedem needs a ROCm GPU")
    lib = _lib.load()
    ins = [_dev(a) for a in (delta_n, delta_s, sigma_n, sigma_t, sigma_bar, current_damage, current_delta_m_max,
                             current_delta_m_0)]
    it = _dev(intact, torch.uint8)
    nb = int(ins[0].numel())
    dev = ins[0].device
    outs = [torch.zeros(nb, dtype=torch.float32, device=dev) for _ in range(3)]
    oint = torch.zeros(nb, dtype=torch.uint8, device=dev)
    f = lambda v: float(np.float32(v))
    rc = lib.icedem_bond_damage(nb, *[a.data_ptr() for a in ins[:5]], f(C), f(mu_b), f(G_c_I), f(G_c_II), f(eta_bk),
                                *[a.data_ptr() for a in ins[5:]], it.data_ptr(), *[o.data_ptr() for o in outs],
                                oint.data_ptr(), torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, None, "icedem_bond_damage")
    torch.cuda.synchronize()
    return outs[0], outs[1], outs[2], oint.bool()
