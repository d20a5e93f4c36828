"""ac-fdtd time loop on MI355X: the reference API over hand-written HIP kernels.

Exports the reference's public names (``src/ac_fdtd/__init__.py:1-45``). The host physics
imports eagerly; the names that need torch and the HIP library (``AcousticFDTD``, the room
driver) are resolved on first use, so ``import ac_fdtd_amd`` loads neither torch nor the
library.
"""

from .air import AirAbsorption, RelaxationProcess
from .analytic import discrete_mode_frequency, mode_frequency, mode_initial_state
from .boundaries import (
    AbsorbingLayer,
    WallAdmittances,
    admittance_from_absorption,
    random_incidence_absorption,
    reflection_coefficient,
)
from .grid import Grid, max_stable_time_step
from .io import resample, write_wav
from .medium import AIR, Medium
from .sources import band_limited_impulse, deconvolve, gaussian_pulse

#: Names resolved on first access: (module, attribute).
_LAZY = {
    "AcousticFDTD": ("scheme", "AcousticFDTD"),
    "ImpulseResponse": ("room", "ImpulseResponse"),
    "Room": ("room", "Room"),
    "simulate_impulse_response": ("room", "simulate_impulse_response"),
    "simulate_impulse_responses": ("room", "simulate_impulse_responses"),
    "BatchedRooms": ("batch", "BatchedRooms"),
}

__all__ = [
    "AIR",
    "AbsorbingLayer",
    "AcousticFDTD",
    "AirAbsorption",
    "BatchedRooms",
    "Grid",
    "ImpulseResponse",
    "Medium",
    "RelaxationProcess",
    "Room",
    "WallAdmittances",
    "admittance_from_absorption",
    "band_limited_impulse",
    "deconvolve",
    "discrete_mode_frequency",
    "gaussian_pulse",
    "max_stable_time_step",
    "mode_frequency",
    "mode_initial_state",
    "random_incidence_absorption",
    "reflection_coefficient",
    "resample",
    "simulate_impulse_response",
    "simulate_impulse_responses",
    "write_wav",
]


def __getattr__(name: str):
    if name in _LAZY:
        import importlib

        module, attribute = _LAZY[name]
        value = getattr(importlib.import_module(f".{module}", __name__), attribute)
        globals()[name] = value
        return value
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
