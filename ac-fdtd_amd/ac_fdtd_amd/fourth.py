"""``FourthOrderFDTD``: the staggered scheme with fourth-order spatial differences, on one GPU.

The reference lists "a 4th-order-in-space variant as an option" (``PLAN.md:17``; milestone M8,
``PLAN.md:116``) without implementing it. This is the reference's step
(``src/ac_fdtd/scheme.py:285-363``) with the pressure gradient and the velocity divergence
taken by the standard fourth-order staggered difference
``c1 (u[i] - u[i-1]) - c2 (u[i+1] - u[i-2])`` (``c1 = 9/8``, ``c2 = 1/24``), the rigid walls
imposed by images (pressure even, normal velocity odd across each wall face), and the sponge,
air absorption, sources and receivers unchanged. Dispersion error falls as ``dx^4`` instead of
``dx^2``, so a target accuracy needs fewer points per wavelength; the price is a Courant limit
``6/7`` of the second-order one (``courant`` is the fraction of the fourth-order limit) and
twice the stencil reads, which the kernels take from L1/L2 (HBM bytes per cell-update are those
of the two-sweep step). Walls are rigid: a locally reacting wall has no closed form under the
wider stencil, so admittances are refused.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .analytic import FOURTH_ORDER
from .boundaries import WallAdmittances
from .medium import AIR
from .scheme import AcousticFDTD

__all__ = ["FOURTH_ORDER", "FourthOrderFDTD"]


def _cell_images(index: np.ndarray, n: int) -> np.ndarray:
    index = np.where(index < 0, -1 - index, index)
    return np.clip(np.where(index >= n, 2 * n - 1 - index, index), 0, n - 1)


class FourthOrderFDTD(AcousticFDTD):
    """The fourth-order staggered scheme with the reference solver's API.

    ``courant`` is the fraction of the fourth-order stability limit ``c dt / dx <= 1 /
    (sqrt(3) (c1 + c2))``; ``coefficients`` are ``(c1, c2)`` (``(1, 0)`` runs the second-order
    arithmetic through the same kernels).
    """

    def __init__(self, grid, medium=AIR, courant: float = 1.0, dtype=np.float64,
                 walls: WallAdmittances | None = None, absorbing_layer=None, air_absorption=None,
                 device="cuda", tile_j: int = 0, chunk_x: int = 0, coefficients=FOURTH_ORDER):
        if walls is not None and not walls.is_rigid:
            raise ValueError("the fourth-order scheme takes rigid walls only: a locally reacting "
                             "wall has no closed form under its wider stencil")
        if not 0.0 < courant <= 1.0:
            raise ValueError(f"the Courant number must lie in (0, 1]; {courant} was given")
        c1, c2 = (float(c) for c in coefficients)
        super().__init__(grid, medium, courant * (1.0 / (c1 + c2)), dtype, None, absorbing_layer,
                         air_absorption, device, fused=False, tile_j=tile_j, chunk_x=chunk_x,
                         time_block=1)
        self.courant = courant
        self.difference = (c1, c2)
        self._rounded = tuple(float(self.np_dtype.type(c)) for c in (c1, c2))

    def two_step_applies(self) -> bool:
        return False

    def capture(self, n_steps: int, record: bool = True):
        raise NotImplementedError("graph capture covers the second-order solver")

    def timed_steps(self, n_steps: int):
        raise NotImplementedError("per-kernel timing covers the second-order solver")

    def run(self, n_steps: int, record: bool = True) -> None:
        """Advance ``n_steps`` (a velocity and a pressure sweep each) on the current stream."""
        n_steps = int(n_steps)
        if n_steps <= 0:
            return
        recording = None
        if self._receivers and record:
            recording = torch.zeros((len(self._receivers), n_steps), **self._opts)
        desc = self._descriptor(recording, tape=False)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        with torch.cuda.device(self.device):
            status = self._lib.acfdtd_hip_fourth_run(ctypes.byref(desc), *self._rounded,
                                                     self.step_index, n_steps, 0,
                                                     ctypes.c_void_p(stream))
        _native.check(status, "acfdtd_hip_fourth_run")
        self.step_index += n_steps
        if recording is not None:
            self._runs.append(recording)

    def _face_difference(self, field: torch.Tensor, axis: int) -> torch.Tensor:
        c1, c2 = self.difference
        n = field.shape[axis]
        faces = np.arange(1, n)

        def take(index):
            return torch.index_select(field, axis, torch.as_tensor(index, device=field.device))

        return c1 * (take(faces) - take(faces - 1)) - c2 * (
            take(_cell_images(faces + 1, n)) - take(_cell_images(faces - 2, n)))

    def energy(self) -> float:
        """The reference's discrete energy (``scheme.py:365-389``) with the fourth-order
        gradient in its coupling term, which makes it the scheme's conserved quantity in a rigid
        box; reduced in fp64 on the device."""
        m, dx = self.medium, self.grid.dx
        p = self.p.double()
        potential = float(torch.sum(p * p)) / (2.0 * m.density * m.sound_speed**2)
        kinetic = coupling = 0.0
        for axis in range(3):
            v = self.velocity[axis].double()
            kinetic += 0.5 * m.density * float(torch.sum(v * v))
            if p.shape[axis] > 1:
                inner = v.narrow(axis, 1, p.shape[axis] - 1)
                coupling += float(torch.sum(inner * self._face_difference(p, axis) / dx))
        return dx**3 * (potential + kinetic - 0.5 * self.dt * coupling)
