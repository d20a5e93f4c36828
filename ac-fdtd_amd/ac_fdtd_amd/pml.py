"""``PMLFDTD``: the first-order solver with a split-field perfectly matched layer, on one GPU.

The reference lines free-field boundaries with a graded sponge and keeps "a true PML in
reserve for when the layer's measured -30 dB is not enough" (``PLAN.md:21-24``;
``src/ac_fdtd/boundaries.py:42-46``: the sponge damps both equations by the same ``sigma`` and
is only approximately reflectionless at grazing incidence). This is Berenger's split-field
layer: inside it the pressure is carried as ``p = p_x + p_y + p_z``, each part and each
velocity component damped only along its own axis, which is reflectionless at every angle in
the continuum (the arithmetic is stated in ``include/acfdtd_hip.h``, ``acfdtd_hip_pml_desc``).
The layer's thickness, grading order, target reflection and faces come from an
:class:`~ac_fdtd_amd.boundaries.AbsorbingLayer`, as for the sponge, so the two are compared
like for like. Walls are rigid behind the layer; the sponge and air absorption are not combined
with it. Two extra cell fields (``p_y``, ``p_z``) are allocated; they carry traffic only in the
layer.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .boundaries import AbsorbingLayer, layer_weights
from .medium import AIR
from .scheme import AcousticFDTD

__all__ = ["PMLFDTD", "pml_coefficients"]


def pml_coefficients(grid, medium, dt: float, layer: AbsorbingLayer,
                     sigma_max: float | None = None):
    """Per axis ``(A, B)`` at cell centres and on faces (fp64): ``A = (1 - s) / (1 + s)``,
    ``B = 1 / (1 + s)`` with ``s = sigma dt / 2`` and ``sigma = sigma_max (depth / d)^order``."""
    if sigma_max is None:
        sigma_max = layer.max_damping(grid.dx, medium.sound_speed)
    cells, faces = layer_weights(grid, layer)

    def pair(weight):
        half = 0.5 * sigma_max * weight * dt
        return (1.0 - half) / (1.0 + half), 1.0 / (1.0 + half)

    return [pair(w) for w in cells], [pair(w) for w in faces]


class PMLFDTD(AcousticFDTD):
    """The reference solver's API with a split-field PML on ``layer.faces``. ``fused``
    (default) advances the steps in fused pairs, one sweep per step between two sets of fields
    with lanes of 4 (fp32) or 2 (fp64) z cells where nz allows; False, the two in-place sweeps
    per step. The same results either way."""

    def __init__(self, grid, medium=AIR, courant: float = 1.0, dtype=np.float64,
                 layer: AbsorbingLayer | None = None, walls=None, absorbing_layer=None,
                 air_absorption=None, device="cuda", tile_j: int = 0, chunk_x: int = 0,
                 sigma_max: float | None = None, fused: bool = True):
        if (walls is not None and not walls.is_rigid) or absorbing_layer is not None \
                or air_absorption is not None:
            raise ValueError("the PML runs with rigid walls behind it and without the sponge "
                             "or air absorption")
        super().__init__(grid, medium, courant, dtype, None, None, None, device, fused=fused,
                         tile_j=tile_j, chunk_x=chunk_x, time_block=1)
        self.layer = layer or AbsorbingLayer()
        cells, faces = pml_coefficients(grid, medium, self.dt, self.layer, sigma_max)
        opts = self._opts
        self._profiles = {
            name: [torch.as_tensor(pair[index].astype(self.np_dtype), device=self.device)
                   for pair in source]
            for name, source, index in (("cell_a", cells, 0), ("cell_b", cells, 1),
                                        ("face_a", faces, 0), ("face_b", faces, 1))
        }
        self.p_y = torch.zeros(grid.shape, **opts)
        self.p_z = torch.zeros(grid.shape, **opts)
        self._pml = _native.PMLDesc()
        self._pml.abi_version = _native.ABI_VERSION
        self._pml.struct_size = ctypes.sizeof(_native.PMLDesc)
        for name, tensors in self._profiles.items():
            for axis, tensor in enumerate(tensors):
                getattr(self._pml, name)[axis] = tensor.data_ptr()
        self._pml.p_y = self.p_y.data_ptr()
        self._pml.p_z = self.p_z.data_ptr()

    def two_step_applies(self) -> bool:
        return False

    def capture(self, n_steps: int, record: bool = True):
        raise NotImplementedError("graph capture covers the solver without a PML")

    def timed_steps(self, n_steps: int):
        raise NotImplementedError("per-kernel timing covers the solver without a PML")

    def run(self, n_steps: int, record: bool = True) -> None:
        n_steps = int(n_steps)
        if n_steps <= 0:
            return
        recording = None
        if self._receivers and record:
            recording = torch.zeros((len(self._receivers), n_steps), **self._opts)
        desc = self._descriptor(recording, tape=False)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        with torch.cuda.device(self.device):
            status = self._lib.acfdtd_hip_pml_run(ctypes.byref(desc), ctypes.byref(self._pml),
                                                  self.step_index, n_steps, 0,
                                                  ctypes.c_void_p(stream))
        _native.check(status, "acfdtd_hip_pml_run")
        self.step_index += n_steps
        if recording is not None:
            self._runs.append(recording)
