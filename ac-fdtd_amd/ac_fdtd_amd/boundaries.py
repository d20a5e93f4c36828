"""Admittance walls and the graded sponge layer: host-side coefficient builders.

Follows the reference's ``src/ac_fdtd/boundaries.py:49-318``.

Walls
-----
A locally reacting wall of normalised admittance ``xi`` enters only the pressure update of the
cells that touch it, through the time-centred closed form

    p_new = p_tmp * 1/(1+a)  -  p_old * a/(1+a),      a = sum over touched faces of c dt xi / (2 dx)

(``boundaries.py:5-28``, ``scheme.py:328-335``). Edges and corners sum two or three faces, in
the fixed face order :data:`FACES`. The reference stores ``a`` as a sparse list of flat
indices; the kernels instead classify each axis index as interior / low wall / high wall /
both (a one-cell axis) and look the pair ``(1/(1+a), a/(1+a))`` up in a 64-entry table, which
holds the same fp64-rounded numbers without an O(N^2) index array.

Sponge
------
``exp(-sigma(x) dt)`` with ``sigma`` graded as ``sigma_max * inside^order`` from the inner face
of the layer, evaluated at each field's own positions (cell centres, or faces along a velocity
component's own axis) — ``boundaries.py:238-283``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .grid import Grid
from .medium import Medium

__all__ = [
    "FACES",
    "AbsorbingLayer",
    "WallAdmittances",
    "admittance_from_absorption",
    "axis_wall_codes",
    "cell_wall_codes",
    "edge_slabs",
    "layer_factors",
    "layer_weights",
    "random_incidence_absorption",
    "reflection_coefficient",
    "wall_coefficient_sums",
    "wall_coefficient_table",
    "wall_update_coefficients",
]

#: Face order used everywhere: (axis, side), side 0 = low, 1 = high (boundaries.py:73).
FACES = ((0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1))

#: Per-axis index class used by the wall table: bit 0 = touches the low face, bit 1 = high.
WALL_CODE_LOW = 1
WALL_CODE_HIGH = 2


def admittance_from_absorption(absorption: float) -> float:
    """``xi = (1 - R) / (1 + R)`` with ``R = sqrt(1 - alpha)`` (boundaries.py:76-87)."""
    if not 0.0 <= absorption <= 1.0:
        raise ValueError(f"an absorption coefficient lies between 0 and 1; {absorption} was given")
    reflection = math.sqrt(1.0 - absorption)
    return (1.0 - reflection) / (1.0 + reflection)


def reflection_coefficient(normalised_admittance: float, angle: float = 0.0) -> float:
    """Plane-wave pressure reflection ``(cos t - xi) / (cos t + xi)`` (boundaries.py:90-95)."""
    cosine = math.cos(angle)
    return (cosine - normalised_admittance) / (cosine + normalised_admittance)


def random_incidence_absorption(normalised_admittance: float, n_angles: int = 1024) -> float:
    """Paris-formula average of ``1 - R(t)^2`` weighted by ``sin 2t`` (boundaries.py:98-117)."""
    angles = (np.arange(n_angles) + 0.5) * (math.pi / 2.0) / n_angles
    cosines = np.cos(angles)
    reflection = (cosines - normalised_admittance) / (cosines + normalised_admittance)
    return float(np.trapezoid((1.0 - reflection**2) * np.sin(2.0 * angles), angles))


@dataclass(frozen=True)
class WallAdmittances:
    """Normalised admittance of the six walls, ``0`` rigid, ``1`` matched (boundaries.py:120-164)."""

    x_min: float = 0.0
    x_max: float = 0.0
    y_min: float = 0.0
    y_max: float = 0.0
    z_min: float = 0.0
    z_max: float = 0.0

    def __post_init__(self) -> None:
        for face, value in zip(FACES, self.values):
            if value < 0.0:
                raise ValueError(
                    f"the admittance of wall {face} is {value} < 0: such a wall would feed "
                    "energy into the room (it generates sound) and the run would diverge"
                )

    @property
    def values(self) -> tuple[float, float, float, float, float, float]:
        """Admittances in :data:`FACES` order."""
        return (self.x_min, self.x_max, self.y_min, self.y_max, self.z_min, self.z_max)

    @property
    def is_rigid(self) -> bool:
        return not any(self.values)

    @classmethod
    def uniform(cls, normalised_admittance: float) -> WallAdmittances:
        return cls(*(normalised_admittance,) * 6)

    @classmethod
    def from_absorption(cls, absorption: float) -> WallAdmittances:
        return cls.uniform(admittance_from_absorption(absorption))

    @classmethod
    def matched(cls) -> WallAdmittances:
        return cls.uniform(1.0)


def _face_coefficients(grid: Grid, medium: Medium, dt: float, walls: WallAdmittances):
    """``a`` contributed by each face, ``(c dt / (2 dx)) * xi`` (boundaries.py:180, :200)."""
    scale = medium.sound_speed * dt / (2.0 * grid.dx)
    return [scale * xi if xi != 0.0 else 0.0 for xi in walls.values]


def axis_wall_codes(count: int) -> np.ndarray:
    """Wall class of every index along one axis of ``count`` cells (see module docstring)."""
    codes = np.zeros(count, dtype=np.int64)
    codes[0] |= WALL_CODE_LOW
    codes[count - 1] |= WALL_CODE_HIGH
    return codes


def _summed_coefficient(face_a, cx: int, cy: int, cz: int) -> float:
    """Sum of the touched faces' ``a`` in :data:`FACES` order, from 0.0 (np.add.at order)."""
    total = 0.0
    for axis, code in enumerate((cx, cy, cz)):
        for side, bit in ((0, WALL_CODE_LOW), (1, WALL_CODE_HIGH)):
            value = face_a[2 * axis + side]
            if code & bit and value != 0.0:
                total = total + value
    return total


def wall_coefficient_table(
    grid: Grid, medium: Medium, dt: float, walls: WallAdmittances
) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """64-entry wall table indexed by ``cx + 4 cy + 16 cz``.

    Returns ``(is_wall, from_updated, from_previous)``: whether any touched face absorbs, and
    ``1/(1+a)`` and ``a/(1+a)`` in fp64 (rounded to the working precision by the caller, as the
    reference does with ``.astype(dtype)`` at ``scheme.py:156-157``).
    """
    face_a = _face_coefficients(grid, medium, dt, walls)
    is_wall = np.zeros(64, dtype=bool)
    updated = np.ones(64)
    previous = np.zeros(64)
    for code in range(64):
        cx, cy, cz = code & 3, (code >> 2) & 3, (code >> 4) & 3
        touched = [
            face_a[2 * axis + side]
            for axis, c in enumerate((cx, cy, cz))
            for side, bit in ((0, WALL_CODE_LOW), (1, WALL_CODE_HIGH))
            if c & bit
        ]
        if not any(value != 0.0 for value in touched):
            continue
        a = _summed_coefficient(face_a, cx, cy, cz)
        is_wall[code] = True
        updated[code] = 1.0 / (1.0 + a)
        previous[code] = a / (1.0 + a)
    return is_wall, updated, previous


def wall_coefficient_sums(
    grid: Grid, medium: Medium, dt: float, walls: WallAdmittances
) -> tuple[np.ndarray, np.ndarray]:
    """``(is_wall, a)`` per wall code ``cx + 4 cy + 16 cz``: whether any touched face absorbs,
    and the summed ``a`` of the touched faces (0 for interior and rigid codes), fp64."""
    face_a = _face_coefficients(grid, medium, dt, walls)
    is_wall = np.zeros(64, dtype=bool)
    total = np.zeros(64)
    for code in range(64):
        cx, cy, cz = code & 3, (code >> 2) & 3, (code >> 4) & 3
        total[code] = _summed_coefficient(face_a, cx, cy, cz)
        is_wall[code] = total[code] != 0.0
    return is_wall, total


def cell_wall_codes(grid: Grid) -> np.ndarray:
    """Wall code ``cx + 4 cy + 16 cz`` of every cell, shape ``grid.shape``."""
    cx, cy, cz = (axis_wall_codes(n) for n in grid.shape)
    return cx[:, None, None] + 4 * cy[None, :, None] + 16 * cz[None, None, :]


def wall_update_coefficients(
    grid: Grid, medium: Medium, dt: float, walls: WallAdmittances
) -> tuple[np.ndarray, np.ndarray]:
    """Sorted flat indices of the absorbing boundary cells and their summed ``a``.

    Same result as the reference's sparse builder (``boundaries.py:167-207``); meant for small
    grids (the CPU oracle), since it classifies every cell.
    """
    face_a = _face_coefficients(grid, medium, dt, walls)
    if not any(face_a):
        return np.empty(0, dtype=np.int64), np.empty(0)
    cx = axis_wall_codes(grid.nx)[:, None, None]
    cy = axis_wall_codes(grid.ny)[None, :, None]
    cz = axis_wall_codes(grid.nz)[None, None, :]
    codes = (cx + 4 * cy + 16 * cz).ravel()
    is_wall, _, _ = wall_coefficient_table(grid, medium, dt, walls)
    indices = np.flatnonzero(is_wall[codes]).astype(np.int64)
    per_code = np.array(
        [_summed_coefficient(face_a, c & 3, (c >> 2) & 3, (c >> 4) & 3) for c in range(64)]
    )
    return indices, per_code[codes[indices]]


@dataclass(frozen=True)
class AbsorbingLayer:
    """Graded sponge lining ``faces`` (boundaries.py:210-241).

    ``thickness`` in cells; ``target_reflection`` sets ``sigma_max`` through the standard
    polynomial-grading estimate; ``order`` is the grading exponent.
    """

    thickness: int = 16
    target_reflection: float = 1e-5
    order: float = 3.0
    faces: tuple[tuple[int, int], ...] = FACES

    def __post_init__(self) -> None:
        if self.thickness < 1:
            raise ValueError(f"the layer needs a thickness of one cell or more; {self.thickness} was given")
        if not 0.0 < self.target_reflection < 1.0:
            raise ValueError(f"the layer's target reflection must lie strictly between 0 and 1; {self.target_reflection} was given")

    def max_damping(self, dx: float, sound_speed: float) -> float:
        """``sigma_max = -(order + 1) c ln(R) / (2 d)`` with ``d = thickness * dx``."""
        depth = self.thickness * dx
        return -(self.order + 1.0) * sound_speed * math.log(self.target_reflection) / (2.0 * depth)


def layer_weights(grid: Grid, layer: AbsorbingLayer) -> tuple[list[np.ndarray], list[np.ndarray]]:
    """Grading ``inside^order`` per axis at cell centres and on faces (0 outside the layer).

    The sponge factor at a position is ``exp(-sigma_max * weight * dt)``; the weights are the
    part that does not depend on the damping strength, so gradients with respect to
    ``sigma_max`` are ``-weight * dt`` times the factor.
    """
    depth = layer.thickness * grid.dx

    def weight(positions: np.ndarray, axis: int) -> np.ndarray:
        extent = grid.lengths[axis]
        distance = np.full(positions.shape, np.inf)
        if (axis, 0) in layer.faces:
            distance = np.minimum(distance, positions)
        if (axis, 1) in layer.faces:
            distance = np.minimum(distance, extent - positions)
        inside = np.clip(1.0 - distance / depth, 0.0, 1.0)
        return inside**layer.order

    cells = [weight(grid.cell_centres(axis), axis) for axis in range(3)]
    faces = [weight(np.arange(grid.shape[axis] + 1) * grid.dx, axis) for axis in range(3)]
    return cells, faces


def layer_factors(
    grid: Grid,
    medium: Medium,
    dt: float,
    layer: AbsorbingLayer,
    sigma_max: float | None = None,
) -> tuple[list[np.ndarray], list[list[np.ndarray]]]:
    """Per-axis 1D damping factors: ``(pressure[axis], velocity[component][axis])``.

    ``velocity[c][a]`` is sampled on faces when ``a == c`` and at cell centres otherwise, so it
    equals ``pressure[a]`` for ``a != c``. All factors are fp64 (boundaries.py:244-283).
    ``sigma_max`` overrides the layer's own damping strength.
    """
    if sigma_max is None:
        sigma_max = layer.max_damping(grid.dx, medium.sound_speed)
    cells, faces = layer_weights(grid, layer)
    pressure = [np.exp(-sigma_max * w * dt) for w in cells]
    velocity = []
    for component in range(3):
        velocity.append(
            [
                np.exp(-sigma_max * faces[axis] * dt) if axis == component else pressure[axis]
                for axis in range(3)
            ]
        )
    return pressure, velocity


def edge_slabs(factor: np.ndarray, axis: int) -> list[tuple[tuple[slice, ...], np.ndarray]]:
    """The leading and trailing runs of ``factor < 1`` as broadcastable slabs (boundaries.py:286-318).

    Used by the CPU oracle, which multiplies only the damped slabs as the reference does.
    """
    damped = factor < 1.0
    if not damped.any():
        return []
    shape = [1, 1, 1]
    shape[axis] = -1

    def slab(span: slice) -> tuple[slice, ...]:
        return tuple(span if a == axis else slice(None) for a in range(3))

    if damped.all():
        return [(slab(slice(None)), factor.reshape(shape))]
    n = factor.size
    lead = int(np.argmin(damped)) if damped[0] else 0
    trail = int(np.argmin(damped[::-1])) if damped[-1] else 0
    out = []
    if lead:
        out.append((slab(slice(0, lead)), factor[:lead].reshape(shape)))
    if trail:
        out.append((slab(slice(n - trail, None)), factor[n - trail :].reshape(shape)))
    return out
