"""Ocean-current inputs of the step (``medea_ice_ship/ocean_current.py:87-137``).

Only the models whose fields are static are provided (none, uniform); they
return per-element (n, 2) water velocities, and a chunk helper that broadcasts
one field over K frames without copying (stride 0), which the device step
reads directly.
"""
from __future__ import annotations

import numpy as np
import torch


class NoCurrentModel:
    is_active = False

    def sample(self, x, y):
        return np.zeros(2)

    def sample_batch(self, positions):
        return np.zeros_like(np.asarray(positions, dtype=np.float64))

    def sample_chunk(self, K, n, device=None):
        return None

    def step(self):
        pass


class UniformCurrentModel:
    """Constant (u, v) = speed (cos d, sin d) everywhere (ocean_current.py:120-137)."""
    is_active = True

    def __init__(self, speed: float, direction_rad: float):
        self.velocity = np.array([speed * np.cos(direction_rad), speed * np.sin(direction_rad)])

    def sample(self, x, y):
        return self.velocity.copy()

    def sample_batch(self, positions):
        return np.tile(self.velocity, (len(positions), 1))

    def sample_chunk(self, K, n, device=None):
        v = torch.tensor(self.velocity, dtype=torch.float32, device=device)
        return v[None, None, :].expand(K, n, 2)

    def step(self):
        pass
