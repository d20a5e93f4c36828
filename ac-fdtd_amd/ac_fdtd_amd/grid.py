"""Uniform Cartesian grid of cubic cells and the Courant time step it implies.

Follows the reference's ``src/ac_fdtd/grid.py:37-163``. Layout convention, shared by every
backend and by the HIP kernels: fields are C-ordered ``(nx, ny, nz)`` with ``z`` unit-stride;
pressure sits at cell centres, the velocity component normal to an axis sits on the faces of
that axis and therefore carries one extra plane along it.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

__all__ = ["Grid", "max_stable_time_step"]

# A coordinate within this fraction of a cell below a cell boundary is counted in the upper
# cell, so that 1.2 / 0.01 = 119.99999999999999 lands in cell 120 (grid.py:34, :131-147).
_SNAP = 1e-6


def max_stable_time_step(dx: float, sound_speed: float) -> float:
    """Courant limit ``dx / (c * sqrt(3))`` of the 3D staggered leapfrog (grid.py:37-45)."""
    return dx / (sound_speed * math.sqrt(3))


@dataclass(frozen=True)
class Grid:
    """``nx * ny * nz`` cubic cells of side ``dx`` (grid.py:48-163).

    Cell ``(i, j, k)`` is centred at ``((i + 0.5) dx, (j + 0.5) dx, (k + 0.5) dx)``.
    """

    dx: float
    nx: int
    ny: int
    nz: int
    requested_lengths: tuple[float, float, float] | None = None

    def __post_init__(self) -> None:
        if not self.dx > 0.0:
            raise ValueError(f"the cell size dx must be positive; {self.dx} was given")
        if min(self.nx, self.ny, self.nz) < 1:
            raise ValueError(f"every axis needs at least one cell; the shape was {self.shape}")

    @classmethod
    def from_lengths(cls, lengths: tuple[float, float, float], dx: float) -> Grid:
        """Round every side to a whole number of cells, at least one (grid.py:70-74)."""
        nx, ny, nz = (max(1, round(side / dx)) for side in lengths)
        return cls(dx=dx, nx=nx, ny=ny, nz=nz, requested_lengths=tuple(lengths))

    @classmethod
    def from_max_frequency(
        cls,
        lengths: tuple[float, float, float],
        max_frequency: float,
        sound_speed: float,
        points_per_wavelength: float = 10.0,
    ) -> Grid:
        """Spacing ``c / (ppw * f_max)``, then :meth:`from_lengths` (grid.py:76-96)."""
        if not max_frequency > 0.0:
            raise ValueError(f"the band limit must be a positive frequency; {max_frequency} Hz was given")
        if not points_per_wavelength > 0.0:
            raise ValueError(f"the resolution must be a positive number of points per wavelength; {points_per_wavelength} was given")
        return cls.from_lengths(lengths, sound_speed / (points_per_wavelength * max_frequency))

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.nx, self.ny, self.nz)

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def lengths(self) -> tuple[float, float, float]:
        """Room actually simulated, ``n * dx`` per axis."""
        return (self.nx * self.dx, self.ny * self.dx, self.nz * self.dx)

    @property
    def length_error(self) -> tuple[float, float, float]:
        """Simulated minus requested length per axis; zeros if nothing was requested."""
        if self.requested_lengths is None:
            return (0.0, 0.0, 0.0)
        return tuple(got - asked for got, asked in zip(self.lengths, self.requested_lengths))

    def velocity_shape(self, axis: int) -> tuple[int, int, int]:
        """Shape of the face-normal velocity component of ``axis``."""
        shape = [self.nx, self.ny, self.nz]
        shape[axis] += 1
        return tuple(shape)

    def cell_index(self, point: tuple[float, float, float]) -> tuple[int, int, int]:
        """Cell containing ``point``, clipped into the grid (grid.py:131-147)."""
        index = []
        for coordinate, count in zip(point, self.shape):
            cell = math.floor(coordinate / self.dx + _SNAP)
            index.append(int(min(max(cell, 0), count - 1)))
        return tuple(index)

    def flat_index(self, point: tuple[float, float, float]) -> int:
        """Row-major offset of :meth:`cell_index` in the pressure field."""
        i, j, k = self.cell_index(point)
        return (i * self.ny + j) * self.nz + k

    def cell_centres(self, axis: int) -> np.ndarray:
        return (np.arange(self.shape[axis]) + 0.5) * self.dx

    def memory_bytes(self, itemsize: int = 8, extra_fields: int = 0) -> int:
        """Bytes of pressure, three velocities and ``extra_fields`` cell-sized arrays."""
        faces = sum(math.prod(self.velocity_shape(axis)) for axis in range(3))
        return (self.n_cells * (1 + extra_fields) + faces) * itemsize
