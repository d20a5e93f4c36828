"""CPU oracle of the fourth-order-in-space staggered scheme, for checking only.

TEST INFRASTRUCTURE: only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline may import this module; the product never does.

The reference lists "a 4th-order-in-space variant as an option" (``PLAN.md:17``, milestone M8
at ``PLAN.md:116``) and does not implement it. This restates the reference's first-order step
(``src/ac_fdtd/scheme.py:285-363``: velocity update, sponge, pressure update with relaxation,
sponge, sources, receivers; through :class:`oracle.scheme.OracleFDTD`) with both spatial
differences replaced by the standard fourth-order staggered one,

    D u (at the point between u[i-1] and u[i]) = c1 (u[i] - u[i-1]) - c2 (u[i+1] - u[i-2]),
    c1 = 9/8, c2 = 1/24,

for the pressure gradient on faces and the velocity divergence in cells. The rigid walls are
imposed by images, the fields being extended past a wall face as if the room were mirrored in
it: pressure even (``p[-1] = p[0]``, ``p[n] = p[n-1]``), normal velocity odd (``v[-1] = -v[1]``,
``v[n+1] = -v[n-1]``, with ``v[0]`` and ``v[n]`` the wall faces). With these images the
gradient is minus the transpose of the divergence, so the discrete energy (the reference's
``energy`` with the fourth-order gradient in its coupling term) is conserved, and the rigid
box's cosine modes stay exact eigenmodes, at
``sin(pi f dt) = (c dt / dx) sqrt(sum_a (c1 sin(k_a dx / 2) - c2 sin(3 k_a dx / 2))^2)``.
Stability needs ``c dt / dx <= 1 / (sqrt(3) (c1 + c2)) = 6 / (7 sqrt 3)``, so the Courant
number is a fraction of that limit. Walls are rigid; the sponge and air absorption are those of
the first-order scheme, unchanged.

How it is pinned (``tests/test_fourth.py``): with ``(c1, c2) = (1, 0)`` it is bit for bit the
first-order oracle (so the sponge, air, source and receiver parts are the reference's); energy
conservation to 1e-11 in a rigid box; exact single modes at the frequency above (a wrong c1,
c2 or image breaks it); and convergence of a mode's frequency to the continuum one at fourth
order in dx.
"""

from __future__ import annotations

import numpy as np

from ac_fdtd_amd.grid import Grid
from ac_fdtd_amd.medium import AIR, Medium
from oracle.scheme import OracleFDTD

__all__ = ["FOURTH_ORDER", "OracleFourthFDTD", "STABILITY_SCALE"]

FOURTH_ORDER = (9.0 / 8.0, 1.0 / 24.0)
#: The fourth-order Courant limit over the second-order one: 1 / (c1 + c2).
STABILITY_SCALE = 6.0 / 7.0


def _cell_image(index: np.ndarray, n: int) -> np.ndarray:
    """Cell indices mirrored back into ``[0, n)`` across the wall faces."""
    index = np.where(index < 0, -1 - index, index)
    return np.where(index >= n, 2 * n - 1 - index, index)


def _face_image(index: np.ndarray, n: int) -> tuple[np.ndarray, np.ndarray]:
    """Face indices mirrored back into ``[0, n]`` and the sign of the odd extension."""
    sign = np.where((index < 0) | (index > n), -1.0, 1.0)
    index = np.where(index < 0, -index, index)
    return np.where(index > n, 2 * n - index, index), sign


class OracleFourthFDTD(OracleFDTD):
    """The fourth-order staggered scheme on NumPy arrays (rigid walls; optional sponge and air).

    ``courant`` is the fraction of the fourth-order stability limit; ``coefficients`` the
    difference weights ``(c1, c2)`` (``(1, 0)`` is the second-order scheme).
    """

    def __init__(self, grid: Grid, medium: Medium = AIR, courant: float = 1.0, dtype=np.float64,
                 absorbing_layer=None, air_absorption=None, coefficients=FOURTH_ORDER):
        c1, c2 = coefficients
        scale = 1.0 / (c1 + c2)
        if not 0.0 < courant <= 1.0:
            raise ValueError(f"the Courant number must lie in (0, 1]; {courant} was given")
        super().__init__(grid, medium, courant * scale, dtype, walls=None,
                         absorbing_layer=absorbing_layer, air_absorption=air_absorption)
        t = self.dtype.type
        self.c1, self.c2 = t(c1), t(c2)

    def _face_difference(self, axis):
        n = self.grid.shape[axis]
        faces = np.arange(1, n)
        near = np.take(self.p, faces, axis=axis) - np.take(self.p, faces - 1, axis=axis)
        far = (np.take(self.p, _cell_image(faces + 1, n), axis=axis)
               - np.take(self.p, _cell_image(faces - 2, n), axis=axis))
        return self.c1 * near - self.c2 * far

    def _axis_divergence(self, axis):
        v = self.velocity[axis]
        n = self.grid.shape[axis]
        cells = np.arange(n)
        near = np.take(v, cells + 1, axis=axis) - np.take(v, cells, axis=axis)
        upper, upper_sign = _face_image(cells + 2, n)
        lower, lower_sign = _face_image(cells - 1, n)
        shape = [1, 1, 1]
        shape[axis] = n
        far_upper = np.take(v, upper, axis=axis) * upper_sign.astype(self.dtype).reshape(shape)
        far_lower = np.take(v, lower, axis=axis) * lower_sign.astype(self.dtype).reshape(shape)
        return self.c1 * near - self.c2 * (far_upper - far_lower)

    def _divergence(self):
        div = self._axis_divergence(0)
        div += self._axis_divergence(1)
        div += self._axis_divergence(2)
        return div
