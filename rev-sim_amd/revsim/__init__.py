"""revsim — MI355X (gfx950) reversible-simulation gradients for flexible 3-site water.

Host-side mirror of rev-sim's Julia interface (water/water.jl: forward_sim, reverse_sim, train)
over the C ABI of librevsim.so (include/revsim.h).
"""
from ._abi import (RS_PRECISION_FP32, RS_PRECISION_FP64, LibraryMissing, RsConfig, config_dict,
                   load_library, make_config)
from .simulator import RevSimError, Simulator
from .system import (DEFAULT_THETA, THETA_NAMES, WORKLOADS, WaterBox, perturbed_target,
                     water_box)

__all__ = [
    "RS_PRECISION_FP32", "RS_PRECISION_FP64", "LibraryMissing", "RsConfig", "config_dict",
    "load_library", "make_config", "RevSimError", "Simulator", "DEFAULT_THETA", "THETA_NAMES",
    "WORKLOADS", "WaterBox", "perturbed_target", "water_box",
]
