"""Closed forms for the rigid box: modal frequencies and an exact single-mode initial state.

Follows the reference's ``src/ac_fdtd/analytic.py:42-125``. The cosine mode shapes are exact
eigenvectors of the discrete rigid-wall operator, so a run started from
:func:`mode_initial_state` stays ``cos(2 pi f_d t) * Phi`` at :func:`discrete_mode_frequency`
to round-off — the strongest whole-scheme check available.
"""

from __future__ import annotations

import math

import numpy as np

from .grid import Grid
from .medium import Medium

__all__ = [
    "FOURTH_ORDER",
    "discrete_mode_frequency",
    "fourth_order_difference",
    "fourth_order_mode_frequency",
    "fourth_order_mode_initial_state",
    "mode_frequency",
    "mode_initial_state",
    "mode_pressure_field",
    "wavenumbers",
]


def mode_frequency(lengths, mode, sound_speed: float) -> float:
    """``(c/2) sqrt(sum (n_a / L_a)^2)`` (analytic.py:42-57)."""
    return 0.5 * sound_speed * math.sqrt(sum((n / length) ** 2 for n, length in zip(mode, lengths)))


def wavenumbers(lengths, mode) -> tuple[float, float, float]:
    return tuple(n * math.pi / length for n, length in zip(mode, lengths))


def discrete_mode_frequency(grid: Grid, mode, medium: Medium, dt: float) -> float:
    """``asin((c dt / dx) sqrt(sum sin^2(k dx / 2))) / (pi dt)`` (analytic.py:66-87)."""
    dx = grid.dx
    spatial = math.sqrt(sum(math.sin(k * dx / 2.0) ** 2 for k in wavenumbers(grid.lengths, mode)))
    argument = medium.sound_speed * dt / dx * spatial
    if argument > 1.0:
        raise ValueError(
            f"the discrete dispersion relation has no real frequency for mode {mode} "
            f"(sin(pi f dt) would be {argument:.6f}): the time step exceeds the stability limit "
            "of this grid"
        )
    return math.asin(argument) / (math.pi * dt)


def mode_pressure_field(grid: Grid, mode) -> np.ndarray:
    """Unit-amplitude product of cosines at cell centres."""
    field = np.ones(grid.shape)
    for axis, (n, length) in enumerate(zip(mode, grid.lengths)):
        shape = [1, 1, 1]
        shape[axis] = -1
        field = field * np.cos(n * math.pi * grid.cell_centres(axis) / length).reshape(shape)
    return field


def mode_initial_state(grid: Grid, mode, medium: Medium, dt: float, amplitude: float = 1.0):
    """``(p^0, v^{-1/2})`` with ``v^{-1/2} = (dt / (2 rho)) grad p^0`` (analytic.py:99-125)."""
    pressure = amplitude * mode_pressure_field(grid, mode)
    scale = dt / (2.0 * medium.density * grid.dx)
    velocity = []
    for axis in range(3):
        component = np.zeros(grid.velocity_shape(axis))
        inner = tuple(slice(1, -1) if a == axis else slice(None) for a in range(3))
        hi = tuple(slice(1, None) if a == axis else slice(None) for a in range(3))
        lo = tuple(slice(None, -1) if a == axis else slice(None) for a in range(3))
        component[inner] = scale * (pressure[hi] - pressure[lo])
        velocity.append(component)
    return pressure, velocity


#: Weights ``(c1, c2)`` of the fourth-order staggered difference
#: ``c1 (u[i] - u[i-1]) - c2 (u[i+1] - u[i-2])``.
FOURTH_ORDER = (9.0 / 8.0, 1.0 / 24.0)


def fourth_order_mode_frequency(grid: Grid, mode, medium: Medium, dt: float,
                                coefficients=FOURTH_ORDER) -> float:
    """Frequency of a rigid-box mode under the fourth-order staggered scheme (whose image walls
    keep the cosine modes exact): ``sin(pi f dt) = (c dt / dx) sqrt(sum_a s_a^2)`` with
    ``s_a = c1 sin(k_a dx / 2) - c2 sin(3 k_a dx / 2)``, the stencil's symbol."""
    c1, c2 = coefficients
    dx = grid.dx
    symbol = math.sqrt(sum((c1 * math.sin(k * dx / 2.0) - c2 * math.sin(1.5 * k * dx)) ** 2
                           for k in wavenumbers(grid.lengths, mode)))
    argument = medium.sound_speed * dt / dx * symbol
    if argument > 1.0:
        raise ValueError(
            f"the fourth-order dispersion relation has no real frequency for mode {mode} "
            f"(sin(pi f dt) would be {argument:.6f}): the time step exceeds the stability limit")
    return math.asin(argument) / (math.pi * dt)


def fourth_order_difference(field: np.ndarray, axis: int, coefficients=FOURTH_ORDER) -> np.ndarray:
    """The fourth-order difference of a cell field on the interior faces of ``axis``, the field
    mirrored evenly across the walls (``u[-1] = u[0]``, ``u[n] = u[n-1]``)."""
    c1, c2 = coefficients
    n = field.shape[axis]
    faces = np.arange(1, n)

    def image(index):
        index = np.where(index < 0, -1 - index, index)
        return np.clip(np.where(index >= n, 2 * n - 1 - index, index), 0, n - 1)

    near = np.take(field, faces, axis=axis) - np.take(field, faces - 1, axis=axis)
    far = np.take(field, image(faces + 1), axis=axis) - np.take(field, image(faces - 2), axis=axis)
    return c1 * near - c2 * far


def fourth_order_mode_initial_state(grid: Grid, mode, medium: Medium, dt: float,
                                    amplitude: float = 1.0, coefficients=FOURTH_ORDER):
    """``(p^0, v^{-1/2})`` of an exact mode under the fourth-order scheme:
    ``v^{-1/2} = (dt / (2 rho dx)) D p^0`` with the fourth-order difference ``D``."""
    pressure = amplitude * mode_pressure_field(grid, mode)
    scale = dt / (2.0 * medium.density * grid.dx)
    velocity = []
    for axis in range(3):
        component = np.zeros(grid.velocity_shape(axis))
        inner = tuple(slice(1, -1) if a == axis else slice(None) for a in range(3))
        component[inner] = scale * fourth_order_difference(pressure, axis, coefficients)
        velocity.append(component)
    return pressure, velocity
