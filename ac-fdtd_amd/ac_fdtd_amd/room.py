"""Room, source, receivers → impulse response, on the GPU time loop.

Follows the reference's ``src/ac_fdtd/room.py:39-176``: a rectangular room of uniform wall
absorption, a grid resolving ``max_frequency`` at 10 points per wavelength, a duration of 1.5
Sabine times by default, the band-limited impulse as excitation, and deconvolution of the
recordings. Where the reference hard-codes its NumPy solver (``room.py:149``), this runs the
HIP :class:`~ac_fdtd_amd.scheme.AcousticFDTD`, in fp64 by default (``dtype`` selects fp32).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .air import AirAbsorption
from .boundaries import WallAdmittances, admittance_from_absorption, random_incidence_absorption
from .grid import Grid
from .medium import AIR, Medium
from .metrics import air_absorption_coefficient, eyring_time, sabine_time
from .sources import band_limited_impulse, deconvolve

__all__ = ["ImpulseResponse", "Room", "simulate_impulse_response"]

DEFAULT_POINTS_PER_WAVELENGTH = 10.0
DEFAULT_DURATION_FACTOR = 1.5


@dataclass(frozen=True)
class Room:
    """Rectangular room: dimensions (m), normal-incidence absorption, air, medium."""

    dimensions: tuple[float, float, float]
    absorption: float = 0.1
    air: AirAbsorption | None = field(default_factory=AirAbsorption)
    medium: Medium = AIR

    @property
    def volume(self) -> float:
        return float(np.prod(self.dimensions))

    @property
    def surface(self) -> float:
        x, y, z = self.dimensions
        return 2.0 * (x * y + y * z + z * x)

    @property
    def wall_admittance(self) -> float:
        return admittance_from_absorption(self.absorption)

    @property
    def diffuse_absorption(self) -> float:
        """Random-incidence absorption, what Sabine and Eyring want (room.py:69-77)."""
        return random_incidence_absorption(self.wall_admittance)

    def air_coefficient(self, frequency: float) -> float:
        if self.air is None:
            return 0.0
        return air_absorption_coefficient(float(self.air.attenuation(frequency)))

    def sabine_reverberation_time(self, frequency: float = 1000.0) -> float:
        return sabine_time(self.volume, self.surface, self.diffuse_absorption,
                           self.air_coefficient(frequency))

    def eyring_reverberation_time(self, frequency: float = 1000.0) -> float:
        return eyring_time(self.volume, self.surface, self.diffuse_absorption,
                           self.air_coefficient(frequency))


@dataclass(frozen=True)
class ImpulseResponse:
    signals: np.ndarray
    recorded: np.ndarray
    excitation: np.ndarray
    sample_rate: float
    max_frequency: float
    grid: Grid


def simulate_impulse_response(
    room: Room,
    source: tuple[float, float, float],
    receivers: list[tuple[float, float, float]],
    max_frequency: float,
    duration: float | None = None,
    points_per_wavelength: float = DEFAULT_POINTS_PER_WAVELENGTH,
    courant: float = 1.0,
    dtype=np.float64,
    device="cuda",
) -> ImpulseResponse:
    """Excite the room at ``source`` and return the deconvolved responses (room.py:121-176)."""
    from .scheme import AcousticFDTD

    grid = Grid.from_max_frequency(room.dimensions, max_frequency, room.medium.sound_speed,
                                   points_per_wavelength)
    solver = AcousticFDTD(grid, medium=room.medium, courant=courant, dtype=dtype,
                          walls=WallAdmittances.from_absorption(room.absorption),
                          air_absorption=room.air, device=device)
    if duration is None:
        duration = DEFAULT_DURATION_FACTOR * room.sabine_reverberation_time()
    n_steps = int(round(duration / solver.dt))
    excitation = band_limited_impulse(n_steps, solver.dt, max_frequency)
    solver.add_volume_source(source, excitation)
    for receiver in receivers:
        solver.add_pressure_receiver(receiver)
    solver.run(n_steps)
    recorded = solver.recorded_pressure
    return ImpulseResponse(
        signals=deconvolve(recorded, excitation, solver.dt, max_frequency),
        recorded=recorded,
        excitation=excitation,
        sample_rate=solver.sample_rate,
        max_frequency=max_frequency,
        grid=grid,
    )
