"""Per-solver scalars and tables of the time step, rounded to the working precision.

Follows the reference constructor ``src/ac_fdtd/scheme.py:107-212``. Every number is computed
in fp64 exactly as the reference computes it and then rounded once to the working precision,
with one deliberate exception kept from the reference: the relaxation target scale
``dt * sum(gain)`` is summed over the *already rounded* gains (``scheme.py:189-191``), which for
fp32 makes it an fp32 sum. The HIP kernels receive these values through the descriptor as
doubles that are exactly representable in the working precision, so kernel and oracle start
from the same bits.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .air import AirAbsorption, RelaxationProcess
from .boundaries import AbsorbingLayer, WallAdmittances, layer_factors, wall_coefficient_table
from .grid import Grid, max_stable_time_step
from .medium import Medium

__all__ = ["MAX_RELAXATION", "SchemeCoefficients", "build_coefficients"]

#: Relaxation states the kernels carry at most: O2, N2, the classical surrogate, one spare.
MAX_RELAXATION = 4


@dataclass
class SchemeCoefficients:
    dtype: np.dtype
    dt: float
    velocity: float  # dt / (rho dx)
    pressure: float  # rho c^2 dt / dx
    source: float  # rho c^2 dt
    relax_decay: list[float] = field(default_factory=list)  # exp(-dt / tau)
    relax_mean: list[float] = field(default_factory=list)  # dt * (tau / dt) (1 - decay)
    relax_gain: list[float] = field(default_factory=list)  # rho c_relaxed^2 Delta / dx
    relax_scale: float = 0.0  # dt * sum(gain)
    wall_is_wall: np.ndarray | None = None  # (64,) bool, by wall code
    wall_from_updated: np.ndarray | None = None  # (64,) working precision
    wall_from_previous: np.ndarray | None = None  # (64,) working precision
    cell_profiles: list[np.ndarray] | None = None  # 3 x (n_a,), sponge at cell centres
    face_profiles: list[np.ndarray] | None = None  # 3 x (n_a + 1,), sponge on faces
    sponge_strength: float | None = None  # sigma_max of the layer, 1/s
    relaxation: tuple[RelaxationProcess, ...] = ()  # the processes the gains came from

    @property
    def n_relax(self) -> int:
        return len(self.relax_gain)

    @property
    def has_walls(self) -> bool:
        return bool(self.wall_is_wall is not None and self.wall_is_wall.any())

    @property
    def has_layer(self) -> bool:
        return self.cell_profiles is not None


def time_step(grid: Grid, medium: Medium, courant: float) -> float:
    """``courant`` times the 3D Courant limit ``dx / (c sqrt 3)`` (scheme.py:127)."""
    if not 0.0 < courant <= 1.0:
        raise ValueError(f"the Courant number must lie in (0, 1]; {courant} was given")
    return courant * max_stable_time_step(grid.dx, medium.sound_speed)


def _round(dtype: np.dtype, value) -> float:
    return float(dtype.type(value))


def build_coefficients(
    grid: Grid,
    medium: Medium,
    courant: float,
    dtype,
    walls: WallAdmittances | None,
    absorbing_layer: AbsorbingLayer | None,
    air_absorption: AirAbsorption | None,
    sponge_strength: float | None = None,
    relaxation_strengths=None,
    wall_table_always: bool = False,
) -> SchemeCoefficients:
    """Coefficients of the time step. ``sponge_strength`` (sigma_max) and
    ``relaxation_strengths`` (one Delta per process) override the values the layer and the air
    model would give, for differentiating with respect to them; ``wall_table_always`` builds
    the wall table even for rigid walls (needed by the adjoint, which differentiates every
    face)."""
    dt = time_step(grid, medium, courant)
    dtype = np.dtype(dtype)
    if dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError(f"dtype must be float32 or float64, got {dtype}")
    walls = walls or WallAdmittances()
    c, rho, dx = medium.sound_speed, medium.density, grid.dx

    out = SchemeCoefficients(
        dtype=dtype,
        dt=dt,
        velocity=_round(dtype, dt / (rho * dx)),
        pressure=_round(dtype, rho * c**2 * dt / dx),
        source=_round(dtype, rho * c**2 * dt),
    )

    if air_absorption is not None:
        processes = air_absorption.processes(c, dt)
        if relaxation_strengths is not None:
            strengths = [float(x) for x in relaxation_strengths]
            if len(strengths) != len(processes):
                raise ValueError(f"expected {len(processes)} relaxation strengths")
            processes = tuple(
                RelaxationProcess(strength=x, relaxation_time=q.relaxation_time)
                for x, q in zip(strengths, processes)
            )
        out.relaxation = tuple(processes)
        if len(processes) > MAX_RELAXATION:
            raise ValueError(f"at most {MAX_RELAXATION} relaxation processes are supported")
        relaxed_c2 = c**2 / (1.0 + sum(process.strength for process in processes))
        rounded_gains = []
        for process in processes:
            decay = math.exp(-dt / process.relaxation_time)
            mean_weight = (process.relaxation_time / dt) * (1.0 - decay)
            gain = rho * relaxed_c2 * process.strength / dx
            out.relax_decay.append(_round(dtype, decay))
            out.relax_mean.append(_round(dtype, dt * mean_weight))
            out.relax_gain.append(_round(dtype, gain))
            rounded_gains.append(dtype.type(gain))
        # Summed over the rounded gains, in the working precision (see module docstring).
        out.relax_scale = _round(dtype, dt * sum(rounded_gains))

    if not walls.is_rigid or wall_table_always:
        is_wall, updated, previous = wall_coefficient_table(grid, medium, dt, walls)
        out.wall_is_wall = is_wall
        out.wall_from_updated = updated.astype(dtype)
        out.wall_from_previous = previous.astype(dtype)

    if absorbing_layer is not None:
        if sponge_strength is None:
            sponge_strength = absorbing_layer.max_damping(dx, c)
        out.sponge_strength = float(sponge_strength)
        pressure, velocity = layer_factors(grid, medium, dt, absorbing_layer, sponge_strength)
        out.cell_profiles = [pressure[a].astype(dtype) for a in range(3)]
        out.face_profiles = [velocity[a][a].astype(dtype) for a in range(3)]
    return out
