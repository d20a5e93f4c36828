"""CPU oracle of the first-order scheme with frequency-dependent (IIR) walls, for checking only.

TEST INFRASTRUCTURE: only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline may import this module; the product never does.

The reference plans this boundary (``PLAN.md:21-24``) and does not implement it. Each wall's
normalised admittance is a conductance plus mass-spring-damper branches, each branch turned into
a biquad by the bilinear transform (:mod:`ac_fdtd_amd.impedance`). A step is the reference's
(:class:`oracle.scheme.OracleFDTD`, ``src/ac_fdtd/scheme.py:285-363``) with the walls' closed
form taken at each wall's instantaneous admittance ``g + sum_b b0_b``, after which every
frequency-dependent face cell updates its branch states from the mean pressure of the step,

    x = (f_u (p_tmp + p_old)) / 2 = (p^{n+1} + p^n) / 2       (f_u = 1 / (1 + a) of the cell)
    y_b = b0_b x + s1_b,  s1_b <- s2_b - a1_b y_b,  s2_b <- b2_b x - a2_b y_b     (per branch)
    H = sum_b s1_b,

and the branches' memory for the next step, ``H / (rho c)`` as an outward velocity, is written
on the wall face (minus for the low faces, whose outward normal points down the axis), where the
next pressure update's divergence reads it. So the outward wall velocity at ``n + 1/2`` is
``(g x + sum_b y_b) / (rho c)``, the discrete admittance applied to the mean pressure.

Pinned (``tests/test_impedance.py``) by: a wall of branches whose memory is never read (all
``b2, a1, a2 = 0``) reducing bit for bit to the first-order oracle with real admittances;
passivity (the acoustic energy over interior faces never exceeds its start); normal-incidence
reflection in a duct against ``(1 - xi) / (1 + xi)`` of the wall's discrete admittance per
frequency band; and a planted sign error on the wall velocity failing the passivity check.
"""

from __future__ import annotations

import numpy as np

from ac_fdtd_amd.boundaries import FACES, WallAdmittances
from ac_fdtd_amd.grid import Grid
from ac_fdtd_amd.impedance import FrequencyDependentWalls
from ac_fdtd_amd.medium import AIR, Medium
from oracle.scheme import OracleFDTD

__all__ = ["OracleImpedanceFDTD"]


def _plane(grid: Grid, axis: int, side: int) -> np.ndarray:
    """Flat cell indices of a face plane, in the boundary tape's order for that plane."""
    shape = grid.shape
    ranges = [np.arange(n) for n in shape]
    ranges[axis] = np.array([0 if side == 0 else shape[axis] - 1])
    mesh = np.meshgrid(*ranges, indexing="ij")
    return ((mesh[0] * shape[1] + mesh[1]) * shape[2] + mesh[2]).ravel()


class OracleImpedanceFDTD(OracleFDTD):
    """The first-order oracle with :class:`FrequencyDependentWalls`."""

    def __init__(self, grid: Grid, medium: Medium = AIR, courant: float = 1.0, dtype=np.float64,
                 walls: FrequencyDependentWalls | None = None, air_absorption=None,
                 wall_sign: float = 1.0):
        walls = walls or FrequencyDependentWalls()
        probe = OracleFDTD(Grid(dx=grid.dx, nx=1, ny=1, nz=1), medium, courant)
        self.iir_walls = walls
        super().__init__(grid, medium, courant, dtype,
                         walls=WallAdmittances(*walls.instantaneous(probe.dt)),
                         air_absorption=air_absorption)
        t = self.dtype.type
        self.table = walls.branch_table(self.dt).astype(self.dtype)
        self.mask = walls.face_mask()
        self.velocity_scale = t(1.0 / (medium.density * medium.sound_speed))
        self.wall_sign = wall_sign  # 1; -1 plants the sign error the passivity test must catch
        self.faces = []  # (face, axis, side, positions in wall_idx, states (cells, B, 2))
        for face, (axis, side) in enumerate(FACES):
            if not (self.mask >> face) & 1:
                continue
            cells = _plane(grid, axis, side)
            where = np.searchsorted(self.wall_idx, cells)
            states = np.zeros((cells.size, self.table.shape[1], 2), dtype=self.dtype)
            self.faces.append((face, axis, side, where, states))

    def step(self):
        super().step()
        half = self.dtype.type(0.5)
        for face, axis, side, where, states in self.faces:
            x = (self.wall_fu[where] * self.boundary_sum[where]) * half
            memory = np.zeros_like(x)
            for b in range(self.table.shape[1]):
                b0, b2, a1, a2 = self.table[face, b]
                y = b0 * x + states[:, b, 0]
                first = states[:, b, 1] - a1 * y
                second = b2 * x - a2 * y
                states[:, b, 0], states[:, b, 1] = first, second
                memory = memory + first
            velocity = memory * self.velocity_scale
            if side == 0:
                velocity = -velocity
            if self.wall_sign != 1.0:
                velocity = velocity * self.dtype.type(self.wall_sign)
            wall = tuple((0 if side == 0 else self.grid.shape[a]) if a == axis else slice(None)
                         for a in range(3))
            self.velocity[axis][wall] = velocity.reshape(self.velocity[axis][wall].shape)

    def acoustic_energy(self) -> float:
        """The reference's energy (``scheme.py:365-389``) over interior faces only: the wall
        faces hold the walls' memory, not acoustic velocity."""
        m, dx = self.medium, self.grid.dx
        total = self.energy()
        for axis in range(3):
            for end in (0, -1):
                wall = tuple(end if a == axis else slice(None) for a in range(3))
                values = self.velocity[axis][wall].astype(np.float64)
                total -= dx**3 * 0.5 * m.density * float(np.dot(values.ravel(), values.ravel()))
        return total
