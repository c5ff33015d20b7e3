"""icedem -- MI355X (gfx950) DPDEM ice-ship frame step behind a C ABI.

Host-side mirror of the reference's JAX operator API
(``medea_ice_ship/jax_simulation.py`` and callees); the compute runs in
``libicedem.so`` (hand-written HIP kernels, see ``csrc/``).
"""
from .config import SimulationConfig, benchmark_config, get_default_simulation_config  # noqa: F401
from .properties import IceProperties, ShipProperties  # noqa: F401
from .ice_state import (IceState, ShipState, IceStaticData, ShipDynamicsData,  # noqa: F401
                        build_static_data, build_initial_ice_state, build_ship_dynamics_data,
                        build_initial_ship_state)

__all__ = ["SimulationConfig", "benchmark_config", "get_default_simulation_config", "IceProperties",
           "ShipProperties", "IceState", "ShipState", "IceStaticData", "ShipDynamicsData", "build_static_data",
           "build_initial_ice_state", "build_ship_dynamics_data", "build_initial_ship_state"]
