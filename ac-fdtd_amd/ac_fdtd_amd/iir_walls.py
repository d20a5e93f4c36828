"""``IIRWallFDTD``: the first-order solver with frequency-dependent (IIR) walls, on one GPU.

The walls are :class:`~ac_fdtd_amd.impedance.FrequencyDependentWalls`: per face a real
admittance, or a passive conductance-plus-branches model (given or fitted to octave-band
absorption), the boundary the reference plans at ``PLAN.md:21-24``. Each step runs the
first-order sweeps with the walls' instantaneous admittance in the wall closed form
(``src/ac_fdtd/boundaries.py:167-207``) and the boundary tape on, then one small kernel updates
every frequency-dependent face cell's branch filters and writes their memory as the wall
face's velocity (``include/acfdtd_hip.h``, ``acfdtd_hip_iir_run``). Air absorption and any
number of sources and receivers work as in :class:`~ac_fdtd_amd.scheme.AcousticFDTD`; the
sponge does not (it would damp the wall faces' memory), and steps run one per sweep.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .boundaries import WallAdmittances
from .coefficients import time_step
from .impedance import FrequencyDependentWalls
from .medium import AIR
from .scheme import AcousticFDTD

__all__ = ["IIRWallFDTD"]


class IIRWallFDTD(AcousticFDTD):
    """The reference solver's API with frequency-dependent walls (``walls``). ``fused``
    (default) advances the steps in fused pairs, one sweep per step between two sets of fields;
    False, the two in-place sweeps per step. The same results either way."""

    def __init__(self, grid, medium=AIR, courant: float = 1.0, dtype=np.float64,
                 walls: FrequencyDependentWalls | None = None, absorbing_layer=None,
                 air_absorption=None, device="cuda", tile_j: int = 0, chunk_x: int = 0,
                 fused: bool = True):
        if absorbing_layer is not None:
            raise ValueError("frequency-dependent walls run without a sponge, which would damp "
                             "the walls' memory held on the wall faces")
        self.iir_walls = walls or FrequencyDependentWalls()
        dt = time_step(grid, medium, courant)
        if self.iir_walls.branch_count() > _native.MAX_BRANCHES:
            raise ValueError(f"at most {_native.MAX_BRANCHES} branches per wall")
        super().__init__(grid, medium, courant, dtype,
                         WallAdmittances(*self.iir_walls.instantaneous(dt)), None,
                         air_absorption, device, fused=fused, tile_j=tile_j, chunk_x=chunk_x,
                         time_block=1)
        self._branches = self.iir_walls.branch_count()
        self._face_mask = self.iir_walls.face_mask() if self._branches else 0
        stride = self.wall_tape_stride()
        self._scratch_tape = torch.zeros((1, stride), **self._opts)
        self._states = torch.zeros((stride, max(self._branches, 1), 2), **self._opts)
        table = self.iir_walls.branch_table(self.dt).astype(self.np_dtype).astype(np.float64)
        self._iir = _native.IIRDesc()
        self._iir.abi_version = _native.ABI_VERSION
        self._iir.struct_size = ctypes.sizeof(_native.IIRDesc)
        self._iir.n_branches = self._branches
        self._iir.face_mask = self._face_mask
        self._iir.velocity_scale = float(self.np_dtype.type(1.0 / (medium.density
                                                                   * medium.sound_speed)))
        coef = np.zeros((6, _native.MAX_BRANCHES, 4))
        coef[:, : self._branches] = table
        for n, value in enumerate(coef.ravel()):
            self._iir.coef[n] = value
        self._iir.state = self._states.data_ptr()

    def two_step_applies(self) -> bool:
        return False

    def capture(self, n_steps: int, record: bool = True):
        raise NotImplementedError("graph capture covers the solver without IIR walls")

    def timed_steps(self, n_steps: int):
        raise NotImplementedError("per-kernel timing covers the solver without IIR walls")

    def attach_wall_tape(self, tape, step0: int = 0) -> None:
        raise NotImplementedError("the IIR walls use the boundary tape themselves")

    def run(self, n_steps: int, record: bool = True) -> None:
        n_steps = int(n_steps)
        if n_steps <= 0:
            return
        recording = None
        if self._receivers and record:
            recording = torch.zeros((len(self._receivers), n_steps), **self._opts)
        desc = self._descriptor(recording, tape=False)
        desc.wall_tape = self._scratch_tape.data_ptr()
        desc.wall_tape_stride = self._scratch_tape.shape[1]
        desc.wall_tape_step0 = self.step_index
        desc.wall_tape_rows = 1
        stream = torch.cuda.current_stream(self.device).cuda_stream
        with torch.cuda.device(self.device):
            status = self._lib.acfdtd_hip_iir_run(ctypes.byref(desc), ctypes.byref(self._iir),
                                                  self.step_index, n_steps, 0,
                                                  ctypes.c_void_p(stream))
        _native.check(status, "acfdtd_hip_iir_run")
        self.step_index += n_steps
        if recording is not None:
            self._runs.append(recording)

    def acoustic_energy(self) -> float:
        """:meth:`energy` over interior faces only: the wall faces hold the walls' memory, not
        acoustic velocity (fp64)."""
        m, dx = self.medium, self.grid.dx
        total = self.energy()
        for axis in range(3):
            v = self.velocity[axis].double()
            for end in (0, v.shape[axis] - 1):
                wall = v.select(axis, end)
                total -= dx**3 * 0.5 * m.density * float(torch.sum(wall * wall))
        return total
