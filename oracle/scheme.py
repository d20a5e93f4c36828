"""CPU oracle: NumPy restatement of the reference time loop, for checking only.

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker or the timed CPU baseline; the
product (``ac_fdtd_amd``) never calls it and fails loudly without its HIP library.

What it restates, in the reference's exact operation order (``src/ac_fdtd/scheme.py``):

* constructor coefficients, ``scheme.py:127-212``, including the sparse wall list built by
  concatenating face planes in FACES order and summing with ``np.add.at``
  (``boundaries.py:167-207``);
* ``step`` = velocity update (``:307-311``) -> velocity sponge, x then y then z slabs
  (``:288-290``, ``:298-301``) -> pressure update with the unscaled divergence
  ``((dvx + dvy) + dvz)`` (``:313-326``) and the relaxation states (``:337-358``) -> wall
  closed form (``:328-335``) -> pressure sponge (``:292``) -> source injection with the
  pre-increment step index (``:360-363``) -> ``step_index += 1`` -> receiver samples
  (``:294-296``);
* ``energy`` for the ``(p^n, v^{n-1/2})`` pairing (``:365-389``).

Host physics it borrows from the product package (``Grid``, ``Medium``, ``WallAdmittances``,
``AbsorbingLayer``, ``AirAbsorption.processes``, ``layer_factors``, ``edge_slabs``) is pinned
separately against golden vectors written from the reference and against closed forms.

How it is pinned: against the reference's own C kernel compiled from its sources into
``oracle/_ref`` (``tests/test_oracle.py``, 1e-13 relative on the five parity configurations),
against the golden vectors in ``tests/golden`` written by importing the reference, and against
the invariants the reference tests (energy conservation, exact single mode, passivity,
reflection, ISO duct attenuation, bitwise zero-admittance / no-air equivalence).
"""

from __future__ import annotations

import math

import numpy as np

from ac_fdtd_amd.boundaries import FACES, WallAdmittances, edge_slabs, layer_factors
from ac_fdtd_amd.grid import Grid, max_stable_time_step
from ac_fdtd_amd.medium import AIR, Medium

__all__ = ["OracleFDTD", "sparse_wall_coefficients"]


def _hi(axis):
    return tuple(slice(1, None) if a == axis else slice(None) for a in range(3))


def _lo(axis):
    return tuple(slice(None, -1) if a == axis else slice(None) for a in range(3))


def _inner(axis):
    return tuple(slice(1, -1) if a == axis else slice(None) for a in range(3))


def sparse_wall_coefficients(grid: Grid, medium: Medium, dt: float, walls: WallAdmittances):
    """Sorted unique flat indices of absorbing boundary cells and their summed ``a``.

    Restates ``boundaries.py:167-207``: one plane per absorbing face, concatenated in FACES
    order, deduplicated with ``np.unique`` and summed with ``np.add.at``.
    """
    shape = grid.shape
    strides = np.array([shape[1] * shape[2], shape[2], 1])
    scale = medium.sound_speed * dt / (2.0 * grid.dx)
    planes, values = [], []
    for (axis, side), xi in zip(FACES, walls.values):
        if xi == 0.0:
            continue
        ranges = [np.arange(n) for n in shape]
        ranges[axis] = np.array([0 if side == 0 else shape[axis] - 1])
        mesh = np.meshgrid(*ranges, indexing="ij")
        flat = sum(m.ravel() * s for m, s in zip(mesh, strides)).astype(np.int64)
        planes.append(flat)
        values.append(np.full(flat.size, scale * xi))
    if not planes:
        return np.empty(0, dtype=np.int64), np.empty(0)
    flat = np.concatenate(planes)
    unique, inverse = np.unique(flat, return_inverse=True)
    summed = np.zeros(unique.size)
    np.add.at(summed, inverse, np.concatenate(values))
    return unique, summed


class OracleFDTD:
    """The reference scheme on NumPy arrays; same constructor and run API as the product."""

    def __init__(
        self,
        grid: Grid,
        medium: Medium = AIR,
        courant: float = 1.0,
        dtype=np.float64,
        walls=None,
        absorbing_layer=None,
        air_absorption=None,
    ):
        if not 0.0 < courant <= 1.0:
            raise ValueError(f"the Courant number must lie in (0, 1]; {courant} was given")
        self.grid, self.medium, self.courant = grid, medium, courant
        self.dtype = np.dtype(dtype)
        self.walls = walls or WallAdmittances()
        self.dt = courant * max_stable_time_step(grid.dx, medium.sound_speed)
        t = self.dtype.type
        self.p = np.zeros(grid.shape, dtype=self.dtype)
        self.velocity = [np.zeros(grid.velocity_shape(a), dtype=self.dtype) for a in range(3)]
        self.step_index = 0

        c, rho, dx, dt = medium.sound_speed, medium.density, grid.dx, self.dt
        self.k_vel = t(dt / (rho * dx))
        self.k_pres = t(rho * c**2 * dt / dx)
        self.k_src = t(rho * c**2 * dt)

        idx, a = sparse_wall_coefficients(grid, medium, dt, self.walls)
        self.wall_idx = idx
        self.wall_fu = (1.0 / (1.0 + a)).astype(self.dtype)
        self.wall_fp = (a / (1.0 + a)).astype(self.dtype)

        self.relax = []  # (state, decay, mean, gain)
        if air_absorption is not None:
            procs = air_absorption.processes(c, dt)
            c2 = c**2 / (1.0 + sum(q.strength for q in procs))
            for q in procs:
                decay = math.exp(-dt / q.relaxation_time)
                mean = (q.relaxation_time / dt) * (1.0 - decay)
                self.relax.append(
                    (
                        np.zeros(grid.shape, dtype=self.dtype),
                        t(decay),
                        t(dt * mean),
                        t(rho * c2 * q.strength / dx),
                    )
                )
        self.relax_scale = t(dt * sum(g for _, _, _, g in self.relax))

        self.p_slabs, self.v_slabs = [], [[], [], []]
        if absorbing_layer is not None:
            pf, vf = layer_factors(grid, medium, dt, absorbing_layer)
            self.p_slabs = [s for a in range(3) for s in edge_slabs(pf[a].astype(self.dtype), a)]
            for comp in range(3):
                self.v_slabs[comp] = [
                    s for a in range(3) for s in edge_slabs(vf[comp][a].astype(self.dtype), a)
                ]

        self.sources = []  # (index tuple, signal)
        self.receivers = []  # (index tuple, list of samples)

    # -- API -----------------------------------------------------------------------------
    @property
    def time(self):
        return self.step_index * self.dt

    @property
    def sample_rate(self):
        return 1.0 / self.dt

    @property
    def relaxation_states(self):
        return [state for state, _, _, _ in self.relax]

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

    # -- one step --------------------------------------------------------------------------
    def _face_difference(self, axis):
        """``p[i+1] - p[i]`` on the interior faces of ``axis``."""
        return self.p[_hi(axis)] - self.p[_lo(axis)]

    def _divergence(self):
        """``((dvx + dvy) + dvz)`` over every cell, wall faces included (scheme.py:317-322)."""
        v = self.velocity
        div = v[0][_hi(0)] - v[0][_lo(0)]
        div += v[1][_hi(1)] - v[1][_lo(1)]
        div += v[2][_hi(2)] - v[2][_lo(2)]
        return div

    def step(self):
        for a in range(3):
            self.velocity[a][_inner(a)] -= self.k_vel * self._face_difference(a)
        for a in range(3):
            for where, f in self.v_slabs[a]:
                self.velocity[a][where] *= f
        self._pressure()
        for where, f in self.p_slabs:
            self.p[where] *= f
        for index, signal in self.sources:
            if self.step_index < len(signal):
                self.p[index] += self.k_src * signal[self.step_index]
        self.step_index += 1
        for index, samples in self.receivers:
            samples.append(self.p[index])

    def _pressure(self):
        flat = self.p.reshape(-1)
        before = flat[self.wall_idx].copy() if self.wall_idx.size else None
        div = self._divergence()
        for state, decay, mean, gain in self.relax:
            target = div * gain
            offset = state - target
            np.multiply(offset, decay, out=state)
            state += target
            self.p += offset * mean
        if self.relax:
            self.p += div * self.relax_scale
        self.p -= div * self.k_pres
        if before is not None:
            # p_tmp + p_old on the wall cells, before the closed form (the boundary tape).
            self.boundary_sum = flat[self.wall_idx] + before
            flat[self.wall_idx] = flat[self.wall_idx] * self.wall_fu - before * self.wall_fp

    # -- diagnostics ------------------------------------------------------------------------
    def energy(self):
        m, dx = self.medium, self.grid.dx
        pot = float(np.dot(self.p.ravel(), self.p.ravel())) / (2.0 * m.density * m.sound_speed**2)
        kin = coup = 0.0
        for a in range(3):
            comp = self.velocity[a]
            kin += 0.5 * m.density * float(np.dot(comp.ravel(), comp.ravel()))
            grad = self._face_difference(a) / dx
            coup += float(np.sum(comp[_inner(a)] * grad))
        return dx**3 * (pot + kin - 0.5 * self.dt * coup)
