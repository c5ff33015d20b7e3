"""Simulation parameters read by the frame step.

Same field names and defaults as the reference's ``SimulationConfig``
(``medea_ice_ship/config.py:46-227``) for every field the DPDEM step reads,
plus three parameters the reference keeps as class constants
(``properties.py:8-11``) and one extension:

* ``ice_density`` (IceProperties.ICE_DENSITY, 920 kg/m^3)
* ``flexural_strength_base`` (IceProperties.FLEXURAL_STRENGTH_BASE, 0.3 MPa)
* ``contact_dilation`` (extension, default 0 = reference): a dilation radius
  added to every element's contact radius, as the 3-D DEM of the reference
  does for its sphere-approximated contacts (dpdem/contact/detection.py:236-237).
"""
from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Optional, Tuple


@dataclass(frozen=True)
class SimulationConfig:
    dt: float = 0.05
    use_rayleigh_dt_cap: bool = False
    rayleigh_safety_factor: float = 0.1
    pen_crit_ref: float = 0.5
    pen_crit_thickness_ref: float = 1.0
    max_fracture_thickness: float = 15.0
    n_edge_samples: int = 4
    w_contact_radius_factor: float = 2.0
    w_contact_cap: float = 4.0
    side_push_factor: float = 0.2
    submergence_rate: float = 0.4
    submergence_pen_cap: float = 0.5
    rise_time_const: float = 0.5
    max_bond_width: float = 4.0
    bond_stiffness_factor: float = 0.001
    poisson_ratio: float = 0.3
    bond_cohesion: float = 1.8e6
    bond_internal_friction: float = 0.7
    bond_G_c_I: float = 12.0
    bond_G_c_II: float = 12.0
    bond_eta_bk: float = 1.75
    ice_ice_friction: float = 0.1
    hertz_factor: float = 2e5
    hertz_exponent: float = 1.5
    friction_cap: float = 1e5
    c_d_drag: float = 0.4
    drag_ref_thickness: float = 1.0
    linear_drag_factor: float = 0.1
    drag_coefficient_moment: float = 0.4
    velocity_clip: Tuple[float, float] = (-20.0, 20.0)
    stress_transfer: float = 0.3
    stress_diffusion_steps: int = 2
    effective_strength_collision: float = 0.1
    angular_damping: float = 0.85
    angular_cap: Tuple[float, float] = (-0.2, 0.2)
    ship_velocity_clip: Tuple[float, float, float] = (15.0, 8.0, 0.15)
    max_ship_ice_force: Optional[float] = 25e6
    ice_brake_threshold: float = 0.0
    position_clip_m: float = 1e4
    ice_density: float = 920.0
    flexural_strength_base: float = 0.3e6
    contact_dilation: float = 0.0


def get_default_simulation_config() -> SimulationConfig:
    return SimulationConfig()


def benchmark_config(**overrides) -> SimulationConfig:
    """Parameters of the benchmark scenes: contact normal stiffness 1e8 N/m
    (hertz_factor 1e8, exponent 1), friction 0.3, strength base 0.5 MPa,
    density 900 kg/m^3, frame dt 5e-3 s."""
    cfg = SimulationConfig(dt=5e-3, hertz_factor=1e8, hertz_exponent=1.0, ice_ice_friction=0.3,
                           flexural_strength_base=0.5e6, ice_density=900.0)
    return replace(cfg, **overrides) if overrides else cfg
