"""CPU oracle of the first-order scheme with a split-field perfectly matched layer, for checking.

TEST INFRASTRUCTURE: only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline may import this module; the product never does.

The reference keeps "a true PML in reserve" (``PLAN.md:21-24``) and explains that its sponge is
not one (``src/ac_fdtd/boundaries.py:42-46``). Berenger's split-field layer for the acoustic
system splits the pressure into ``p = p_x + p_y + p_z``, each part driven by its own axis,

    dp_a/dt + sigma_a p_a = -rho c^2 dv_a/da,      dv_a/dt + sigma_a v_a = -(1/rho) dp/da,

which is reflectionless at every angle and frequency in the continuum. ``sigma_a`` is graded
like the sponge (``sigma_max (depth/d)^order``, ``sigma_max = -(order + 1) c ln R / (2 d)``,
``boundaries.py:210-241``) but damps each field only along its own axis. Discretised with the
damping centred in time: ``A = (1 - sigma dt/2) / (1 + sigma dt/2)``, ``B = 1 / (1 + sigma dt/2)``.

One step, in the working precision:

* every interior face: ``v_a = A_f v_a - B_f (vel (p[i] - p[i-1]))`` (A_f = B_f = 1 outside the
  layer, where this is the reference's update bit for bit);
* cells where no axis damps (``A_c = 1`` on all three): the reference's pressure update
  ``p = p - pres ((Dx + Dy) + Dz)``, ``Da = v_a[hi] - v_a[lo]``;
* layer cells: ``p_x = p - p_y - p_z`` (only p_y and p_z are stored, so ``p`` stays the public
  field and may be edited), then ``p_a = A_c,a p_a - B_c,a (pres Da)`` for each axis and
  ``p = (p_x + p_y) + p_z``;
* then sources and receivers as the reference (``scheme.py:360-363``, ``:294-296``).

Walls are rigid (the layer ends on them); no sponge, no air. Pinned (``tests/test_pml.py``)
by: no layer faces being bit for bit the first-order oracle; a layer with ``sigma = 0`` agreeing
with it to rounding; reflection at normal and oblique incidence measured against a much larger
domain, far below the sponge's of the same thickness; and a planted mistake (the velocity left
undamped, so the layer is no longer matched) reflecting far more.
"""

from __future__ import annotations

import numpy as np

from ac_fdtd_amd.boundaries import AbsorbingLayer, layer_weights
from ac_fdtd_amd.grid import Grid
from ac_fdtd_amd.medium import AIR, Medium
from oracle.scheme import OracleFDTD, _hi, _inner, _lo

__all__ = ["OraclePMLFDTD", "pml_coefficients"]


def pml_coefficients(grid: Grid, medium: Medium, dt: float, layer: AbsorbingLayer,
                     sigma_max: float | None = None):
    """Per axis ``(A, B)`` at cell centres and on faces (fp64):
    ``A = (1 - s) / (1 + s)``, ``B = 1 / (1 + s)``, ``s = sigma dt / 2``."""
    if sigma_max is None:
        sigma_max = layer.max_damping(grid.dx, medium.sound_speed)
    cells, faces = layer_weights(grid, layer)

    def pair(weight):
        half = 0.5 * sigma_max * weight * dt
        return (1.0 - half) / (1.0 + half), 1.0 / (1.0 + half)

    return [pair(w) for w in cells], [pair(w) for w in faces]


class OraclePMLFDTD(OracleFDTD):
    """First-order scheme with a split-field PML on ``layer.faces`` (rigid walls behind)."""

    def __init__(self, grid: Grid, medium: Medium = AIR, courant: float = 1.0, dtype=np.float64,
                 layer: AbsorbingLayer | None = None, sigma_max: float | None = None,
                 damp_velocity: bool = True):
        super().__init__(grid, medium, courant, dtype)
        layer = layer or AbsorbingLayer()
        cells, faces = pml_coefficients(grid, medium, self.dt, layer, sigma_max)
        if not damp_velocity:  # planted mistake for the tests: a layer that is not matched
            faces = [(np.ones_like(f[0]), np.ones_like(f[1])) for f in faces]
        t = self.dtype
        shape = [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]
        self.cell_a = [c[0].astype(t).reshape(shape[a]) for a, c in enumerate(cells)]
        self.cell_b = [c[1].astype(t).reshape(shape[a]) for a, c in enumerate(cells)]
        self.face_a = [f[0].astype(t).reshape(shape[a]) for a, f in enumerate(faces)]
        self.face_b = [f[1].astype(t).reshape(shape[a]) for a, f in enumerate(faces)]
        one = t.type(1)
        self.layer_cells = ((self.cell_a[0] != one) | (self.cell_a[1] != one)
                            | (self.cell_a[2] != one))
        self.p_y = np.zeros(grid.shape, dtype=t)
        self.p_z = np.zeros(grid.shape, dtype=t)

    def step(self):
        for a in range(3):
            inner = _inner(a)
            f_a, f_b = self.face_a[a][inner], self.face_b[a][inner]
            self.velocity[a][inner] = (f_a * self.velocity[a][inner]
                                       - f_b * (self.k_vel * self._face_difference(a)))
        v = self.velocity
        parts = [v[a][_hi(a)] - v[a][_lo(a)] for a in range(3)]
        plain = self.p - ((parts[0] + parts[1]) + parts[2]) * self.k_pres
        p_x = (self.p - self.p_y) - self.p_z
        split = [p_x, self.p_y, self.p_z]
        new = [self.cell_a[a] * split[a] - self.cell_b[a] * (self.k_pres * parts[a])
               for a in range(3)]
        layered = (new[0] + new[1]) + new[2]
        mask = self.layer_cells
        self.p = np.where(mask, layered, plain)
        self.p_y = np.where(mask, new[1], self.p_y)
        self.p_z = np.where(mask, new[2], self.p_z)
        for index, signal in self.sources:
            if self.step_index < len(signal):
                self.p[index] += self.k_src * signal[self.step_index]
        self.step_index += 1
        for index, samples in self.receivers:
            samples.append(self.p[index])
