"""Impulse responses of rectangular rooms on the HIP time loop.

The public names and the decisions behind them follow the reference's room driver
(``src/ac_fdtd/room.py:39-176``): the grid resolves ``max_frequency`` at 10 points per
wavelength (``Grid.from_max_frequency``), the run lasts 1.5 Sabine times unless told otherwise,
the source plays the band-limited impulse and the recordings are deconvolved by it. The
reference always runs its NumPy solver (``room.py:149``); here a run is first laid out as a
:class:`RunPlan` (grid, step count, excitation) and then executed on the GPU in the precision
asked for, one room at a time (:func:`simulate_impulse_response`) or many rooms advanced
together by the batched kernels (:func:`simulate_impulse_responses`). A single-precision run can
also state its own error: ``reference_error=True`` repeats it in fp64 and reports the receivers'
relative L2 difference.
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

__all__ = ["ImpulseResponse", "Room", "RunPlan", "simulate_impulse_response",
           "simulate_impulse_responses"]

#: Grid resolution at ``max_frequency``: the customary figure for a second-order scheme.
DEFAULT_POINTS_PER_WAVELENGTH = 10.0
#: Default run length in Sabine times: enough decay for a T30 fit with margin.
DEFAULT_DURATION_FACTOR = 1.5


@dataclass(frozen=True)
class Room:
    """A shoebox: interior ``dimensions`` in metres, one normal-incidence ``absorption`` for
    every wall, the air (``None`` for a lossless one) and the fluid."""

    dimensions: tuple[float, float, float]
    absorption: float = 0.1
    air: AirAbsorption | None = field(default_factory=AirAbsorption)
    medium: Medium = AIR

    @property
    def volume(self) -> float:
        length, width, height = self.dimensions
        return float(length * width * height)

    @property
    def surface(self) -> float:
        length, width, height = self.dimensions
        return 2.0 * (length * width + width * height + height * length)

    @property
    def wall_admittance(self) -> float:
        """Normalised admittance ``rho c Y`` that absorbs ``absorption`` at normal incidence."""
        return admittance_from_absorption(self.absorption)

    @property
    def diffuse_absorption(self) -> float:
        """The walls' absorption averaged over random incidence: the coefficient the
        statistical formulae take (larger than the normal-incidence one for these walls)."""
        return random_incidence_absorption(self.wall_admittance)

    def air_coefficient(self, frequency: float) -> float:
        """Energy attenuation of the air at ``frequency``, 1/m (0 without air losses)."""
        if self.air is None:
            return 0.0
        return air_absorption_coefficient(float(self.air.attenuation(frequency)))

    def _statistical(self, formula, frequency: float) -> float:
        return formula(self.volume, self.surface, self.diffuse_absorption,
                       self.air_coefficient(frequency))

    def sabine_reverberation_time(self, frequency: float = 1000.0) -> float:
        """Sabine's T60 at ``frequency``, seconds (an estimate to compare against)."""
        return self._statistical(sabine_time, frequency)

    def eyring_reverberation_time(self, frequency: float = 1000.0) -> float:
        """Eyring's T60 at ``frequency``, seconds."""
        return self._statistical(eyring_time, frequency)


@dataclass(frozen=True)
class ImpulseResponse:
    """Responses at the receivers, ``(receiver, sample)`` at ``sample_rate``.

    ``signals`` are deconvolved by the excitation, ``recorded`` is the raw pressure, ``grid``
    the grid that ran, ``dtype`` the precision it ran in, and ``fp64_relative_l2`` the
    receivers' relative L2 difference from an fp64 rerun when one was asked for.
    """

    signals: np.ndarray
    recorded: np.ndarray
    excitation: np.ndarray
    sample_rate: float
    max_frequency: float
    grid: Grid
    dtype: str = "float64"
    fp64_relative_l2: float | None = None


@dataclass(frozen=True)
class RunPlan:
    """How one room is simulated: its grid, time step, step count and excitation."""

    room: Room
    grid: Grid
    dt: float
    n_steps: int
    excitation: np.ndarray
    max_frequency: float
    courant: float

    @classmethod
    def for_room(cls, room: Room, max_frequency: float, duration: float | None = None,
                 points_per_wavelength: float = DEFAULT_POINTS_PER_WAVELENGTH,
                 courant: float = 1.0) -> "RunPlan":
        from .coefficients import time_step

        grid = Grid.from_max_frequency(room.dimensions, max_frequency, room.medium.sound_speed,
                                       points_per_wavelength)
        dt = time_step(grid, room.medium, courant)
        if duration is None:
            duration = DEFAULT_DURATION_FACTOR * room.sabine_reverberation_time()
        n_steps = int(round(duration / dt))
        if n_steps < 1:
            raise ValueError(f"a run of {duration} s is shorter than one time step ({dt} s)")
        excitation = band_limited_impulse(n_steps, dt, max_frequency)
        return cls(room, grid, dt, n_steps, excitation, max_frequency, courant)

    def solver(self, dtype, device):
        """An :class:`~ac_fdtd_amd.scheme.AcousticFDTD` for this room, nothing added yet."""
        from .scheme import AcousticFDTD

        solver = AcousticFDTD(self.grid, medium=self.room.medium, courant=self.courant,
                              dtype=dtype, walls=WallAdmittances.from_absorption(self.room.absorption),
                              air_absorption=self.room.air, device=device)
        if solver.dt != self.dt:
            raise RuntimeError("the solver's time step differs from the plan's")
        return solver

    def response(self, recorded: np.ndarray, dtype, error: float | None = None) -> ImpulseResponse:
        return ImpulseResponse(
            signals=deconvolve(recorded, self.excitation, self.dt, self.max_frequency),
            recorded=recorded,
            excitation=self.excitation,
            sample_rate=1.0 / self.dt,
            max_frequency=self.max_frequency,
            grid=self.grid,
            dtype=np.dtype(dtype).name,
            fp64_relative_l2=error,
        )


def _equip(solver, plan: RunPlan, source, receivers):
    solver.add_volume_source(source, plan.excitation)
    for point in receivers:
        solver.add_pressure_receiver(point)
    return solver


def _run_alone(plan: RunPlan, source, receivers, dtype, device) -> np.ndarray:
    solver = _equip(plan.solver(dtype, device), plan, source, receivers)
    solver.run(plan.n_steps)
    return solver.recorded_pressure


def _relative_l2(test: np.ndarray, reference: np.ndarray) -> float:
    scale = float(np.linalg.norm(reference))
    return float(np.linalg.norm(test - reference)) / scale if scale > 0.0 else float("nan")


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
    reference_error: bool = False,
) -> ImpulseResponse:
    """Excite ``room`` at ``source`` and return the responses at ``receivers`` (positions in
    metres). ``max_frequency`` sets the grid and through it the cost, which grows as its
    fourth power; ``duration`` defaults to 1.5 Sabine times. ``dtype`` float64 (the default)
    or float32; with ``reference_error`` an fp32 run is repeated in fp64 to state its error."""
    plan = RunPlan.for_room(room, max_frequency, duration, points_per_wavelength, courant)
    recorded = _run_alone(plan, source, receivers, dtype, device)
    error = None
    if reference_error and np.dtype(dtype) != np.float64:
        error = _relative_l2(recorded, _run_alone(plan, source, receivers, np.float64, device))
    return plan.response(recorded, dtype, error)


def simulate_impulse_responses(
    jobs,
    max_frequency: float,
    duration: float | None = None,
    points_per_wavelength: float = DEFAULT_POINTS_PER_WAVELENGTH,
    courant: float = 1.0,
    dtype=np.float32,
    device="cuda",
    reference_error: bool = False,
) -> list[ImpulseResponse]:
    """Many rooms at once: ``jobs`` is a list of ``(room, source, receivers)``. All rooms
    advance together through the batched kernels (:class:`~ac_fdtd_amd.batch.BatchedRooms`)
    for the longest of their step counts; each response keeps its own room's steps, which are
    the same bits a run of that room alone gives. fp32 by default; ``reference_error`` reruns
    each room alone in fp64 and states its error."""
    from .batch import BatchedRooms

    plans = [RunPlan.for_room(room, max_frequency, duration, points_per_wavelength, courant)
             for room, _, _ in jobs]
    solvers = [_equip(plan.solver(dtype, device), plan, source, receivers)
               for plan, (_, source, receivers) in zip(plans, jobs)]
    batch = BatchedRooms(solvers)
    try:
        batch.run(max(plan.n_steps for plan in plans))
    finally:
        batch.close()
    responses = []
    for plan, solver, (_, source, receivers) in zip(plans, solvers, jobs):
        recorded = solver.recorded_pressure[:, : plan.n_steps]
        error = None
        if reference_error and np.dtype(dtype) != np.float64:
            error = _relative_l2(recorded, _run_alone(plan, source, receivers, np.float64, device))
        responses.append(plan.response(recorded, dtype, error))
    return responses
