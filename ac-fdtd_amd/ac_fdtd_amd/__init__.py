"""ac-fdtd time loop on MI355X: the reference API over hand-written HIP kernels."""

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
from .medium import AIR, Medium
from .sources import band_limited_impulse, deconvolve, gaussian_pulse

__all__ = [
    "AIR",
    "AbsorbingLayer",
    "AirAbsorption",
    "Grid",
    "Medium",
    "RelaxationProcess",
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
]
