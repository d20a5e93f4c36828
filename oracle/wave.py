"""CPU oracle of the pressure-only (second-order wave equation) scheme, for checking only.

TEST INFRASTRUCTURE: only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline may import this module; the product never does.

The reference plans this variant and does not implement it (``PLAN.md:52-56``, M8 at
``PLAN.md:116``): "the second-order wave equation on pressure alone carries two fields instead
of four". Eliminating the velocity from the reference's staggered leapfrog
(``src/ac_fdtd/scheme.py:285-363``) gives it exactly. With ``K = rho c^2 dt / dx``,
``k = dt / (rho dx)`` and ``lambda^2 = K k = (c dt / dx)^2``:

* interior: ``p^{n+1} = 2 p^n - p^{n-1} + lambda^2 L p^n``, ``L`` the 7-point Laplacian over the
  cell's interior faces (a wall face carries no flux: the rigid wall is structural, as the
  velocity update never writes wall faces);
* a cell touching absorbing faces with summed ``a`` (the reference's time-centred closed form,
  ``boundaries.py:167-207``): subtracting the closed form at steps n and n - 1,
  ``(1 + a) p^{n+1} = 2 p^n - (1 - a) p^{n-1} + lambda^2 L p^n``;
* a soft source adding ``S^n = rho c^2 dt s[n]`` after step n (``scheme.py:360-363``) enters as
  ``S^n - S^{n-1}`` after the division.

So started from ``(p^1, p^0)`` of a first-order run, this scheme reproduces that run's
pressures up to rounding (no sponge, no air: those attach to the first-order fields).

Arithmetic of one step, in the working precision, which the HIP kernel reproduces bit for bit:
``L = (Lx + Ly) + Lz`` with ``La = F[hi] - F[lo]`` and face fluxes ``F = p[i+1] - p[i]`` (0 on
wall faces); ``p_new = ((2 p - q p_prev) + lambda^2 L) f`` with ``f = 1 / (1 + a)`` and
``q = 1 - a`` per cell (1 and 1 inside); then, per source in insertion order,
``p_new[cell] += src s[n] - src s[n - 1]`` (terms beyond the signal are 0).

Discrete invariant (rigid walls): ``H^{n+1/2} = |p^{n+1} - p^n|^2 + lambda^2 <G p^{n+1}, G p^n>``
with ``G`` the undivided face differences; it is conserved exactly (summation by parts of
``L = -G^T G``), and absorbing walls only decrease it.

How this oracle is pinned (``tests/test_wave.py``): against the first-order oracle
``oracle/scheme.py`` (itself bit-for-bit the reference) on runs with walls and sources; the
exact single mode at the first-order scheme's discrete modal frequency; conservation of
``H``; and a planted sign error failing those checks.
"""

from __future__ import annotations

import numpy as np

from ac_fdtd_amd.boundaries import WallAdmittances, cell_wall_codes, wall_coefficient_sums
from ac_fdtd_amd.grid import Grid, max_stable_time_step
from ac_fdtd_amd.medium import AIR, Medium

__all__ = ["OracleWaveFDTD"]


def _hi(axis):
    return tuple(slice(1, None) if a == axis else slice(None) for a in range(3))


def _lo(axis):
    return tuple(slice(None, -1) if a == axis else slice(None) for a in range(3))


def _inner(axis):
    return tuple(slice(1, -1) if a == axis else slice(None) for a in range(3))


def face_fluxes(p: np.ndarray, axis: int) -> np.ndarray:
    """``p[i+1] - p[i]`` on interior faces of ``axis``, 0 on its two wall planes (the shape of
    the first-order scheme's velocity component along ``axis``)."""
    shape = list(p.shape)
    shape[axis] += 1
    flux = np.zeros(shape, dtype=p.dtype)
    flux[_inner(axis)] = p[_hi(axis)] - p[_lo(axis)]
    return flux


def laplacian(p: np.ndarray) -> np.ndarray:
    """``(Lx + Ly) + Lz``, ``La = F[hi] - F[lo]``: the undivided Laplacian with no flux
    through the walls."""
    parts = []
    for axis in range(3):
        flux = face_fluxes(p, axis)
        parts.append(flux[_hi(axis)] - flux[_lo(axis)])
    return (parts[0] + parts[1]) + parts[2]


class OracleWaveFDTD:
    """The pressure-only scheme on NumPy arrays: state ``(p^n, p^{n-1})`` as ``p`` and
    ``p_previous``; the same source / receiver / run API as the first-order oracle."""

    def __init__(self, grid: Grid, medium: Medium = AIR, courant: float = 1.0, dtype=np.float64,
                 walls: WallAdmittances | None = None):
        if not 0.0 < courant <= 1.0:
            raise ValueError(f"the Courant number must lie in (0, 1]; {courant} was given")
        self.grid, self.medium = grid, medium
        self.dtype = np.dtype(dtype)
        t = self.dtype.type
        self.walls = walls or WallAdmittances()
        self.dt = courant * max_stable_time_step(grid.dx, medium.sound_speed)
        c, rho, dx, dt = medium.sound_speed, medium.density, grid.dx, self.dt
        self.courant2 = t((c * dt / dx) ** 2)
        self.k_src = t(rho * c**2 * dt)
        _, a = wall_coefficient_sums(grid, medium, dt, self.walls)
        codes = cell_wall_codes(grid)
        self.f = (1.0 / (1.0 + a)).astype(self.dtype)[codes]
        self.q = (1.0 - a).astype(self.dtype)[codes]
        self.p = np.zeros(grid.shape, dtype=self.dtype)
        self.p_previous = np.zeros(grid.shape, dtype=self.dtype)
        self.step_index = 0
        self.sources = []
        self.receivers = []

    @property
    def time(self):
        return self.step_index * self.dt

    def add_volume_source(self, point, signal):
        self.sources.append((self.grid.cell_index(point), np.asarray(signal, dtype=self.dtype)))

    def add_pressure_receiver(self, point):
        self.receivers.append((self.grid.cell_index(point), []))
        return len(self.receivers) - 1

    @property
    def recorded_pressure(self):
        if not self.receivers:
            return np.zeros((0, 0))
        return np.array([samples for _, samples in self.receivers])

    def run(self, n_steps):
        for _ in range(n_steps):
            self.step()

    def _injection(self, signal, n):
        zero = self.dtype.type(0)
        now = self.k_src * signal[n] if n < len(signal) else zero
        before = self.k_src * signal[n - 1] if 1 <= n <= len(signal) else zero
        return now - before

    def step(self):
        two = self.dtype.type(2)
        update = (self.p * two - self.q * self.p_previous) + self.courant2 * laplacian(self.p)
        new = update * self.f
        for index, signal in self.sources:
            new[index] += self._injection(signal, self.step_index)
        self.p_previous, self.p = self.p, new
        self.step_index += 1
        for index, samples in self.receivers:
            samples.append(self.p[index])

    def invariant(self, dtype=np.float64) -> float:
        """``|p^n - p^{n-1}|^2 + lambda^2 <G p^n, G p^{n-1}>`` in fp64 (the discrete energy of the
        last half step, conserved with rigid walls)."""
        now = self.p.astype(dtype)
        before = self.p_previous.astype(dtype)
        coupling = sum(float(np.sum(face_fluxes(now, a) * face_fluxes(before, a))) for a in range(3))
        return float(np.sum((now - before) ** 2)) + float(self.courant2) * coupling
