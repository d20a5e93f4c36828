"""``PressureOnlyFDTD``: the second-order wave equation on pressure alone, on one GPU.

The reference plans this variant and does not implement it (``PLAN.md:52-56``: "the
second-order wave equation on pressure alone carries two fields instead of four, which on a
bandwidth-bound problem is close to a factor two in wall time"; milestone M8,
``PLAN.md:116``). It is the reference's staggered step (``src/ac_fdtd/scheme.py:285-363``)
with the velocity eliminated: for rigid and locally reacting walls and soft sources it gives
the same pressures as :class:`~ac_fdtd_amd.scheme.AcousticFDTD` up to rounding, from two cell
fields instead of four (one sweep per step reads ``p^n`` and ``p^{n-1}`` and writes ``p^{n+1}``
over ``p^{n-1}``: 12 B per cell-update in fp32, against the first-order fused step's 32 B
without air). The sponge and air absorption attach to the velocity and relaxation fields of
the first-order system (``PLAN.md:55-56``), so this solver refuses them.

State: ``p`` holds ``p^n`` and ``p_previous`` ``p^{n-1}`` (torch tensors on the GPU); the two
swap roles every step, so the attributes are re-pointed after a run. With ``time_block=2`` (the
default in fp32) the kernel advances two steps per sweep through a scratch pair of fields, reading
``p^n`` and ``p^{n-1}`` and writing ``p^{n+2}`` and ``p^{n+1}``: 8 B per cell-update in fp32
instead of 12, the same bits as one step per sweep. A first-order run hands
over through :meth:`PressureOnlyFDTD.set_state` with ``p^{n}`` and ``p^{n-1}``.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .boundaries import WallAdmittances, wall_coefficient_sums
from .coefficients import time_step
from .grid import Grid
from .medium import AIR, Medium

__all__ = ["PressureOnlyFDTD"]

_TORCH = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}


class PressureOnlyFDTD:
    """The pressure-only scheme with the reference solver's API (``add_volume_source``,
    ``add_pressure_receiver``, ``run``, ``step``, ``recorded_pressure``, ``energy``).

    ``energy()`` is the scheme's conserved quantity, not joules: the pressure-only state has no
    velocity to form the kinetic energy from. It is
    ``dx^3 / (2 rho c^2 lambda^2) (|p^n - p^{n-1}|^2 + lambda^2 <G p^n, G p^{n-1}>)`` with ``G``
    the differences across interior faces and ``lambda = c dt / dx``; exactly conserved in a
    rigid box, non-increasing with absorbing walls.
    """

    def __init__(self, grid: Grid, medium: Medium = AIR, courant: float = 1.0, dtype=np.float64,
                 walls: WallAdmittances | None = None, absorbing_layer=None, air_absorption=None,
                 device="cuda", tile_j: int = 0, chunk_x: int = 0,
                 time_block: int | None = None) -> None:
        if absorbing_layer is not None or air_absorption is not None:
            raise ValueError(
                "the pressure-only scheme has no velocity or relaxation fields for a sponge or "
                "air absorption to act on; use AcousticFDTD for those")
        from .scheme import _require_gpu

        self.device = _require_gpu(device)
        self._lib = _native.library()
        self.grid, self.medium, self.courant = grid, medium, courant
        self.walls = walls or WallAdmittances()
        self.np_dtype = np.dtype(dtype)
        if self.np_dtype not in _TORCH:
            raise ValueError(f"dtype must be float32 or float64, got {self.np_dtype}")
        self.dtype = _TORCH[self.np_dtype]
        self.dt = time_step(grid, medium, courant)
        c, rho, dx = medium.sound_speed, medium.density, grid.dx
        self.courant2 = float(self.np_dtype.type((c * self.dt / dx) ** 2))
        self.source_coefficient = float(self.np_dtype.type(rho * c**2 * self.dt))
        if time_block is None:
            # Measured on MI355X (walls): in fp32 two steps per sweep is 15 % faster than one at
            # 256^3 (both fields fit the 256 MB last-level cache) and 35 % at 512^3; in fp64 its
            # lane groups of two cells spend more on halo work than it saves (240 against 305
            # Gcell/s at 256^3).
            time_block = 2 if self.np_dtype == np.float32 else 1
        if time_block not in (1, 2):
            raise ValueError(f"time_block must be 1 or 2, got {time_block}")
        self.tile_j, self.chunk_x, self.time_block = tile_j, chunk_x, time_block
        self.step_index = 0
        opts = {"device": self.device, "dtype": self.dtype}
        self.p = torch.zeros(grid.shape, **opts)
        self.p_previous = torch.zeros(grid.shape, **opts)
        # Scratch pair of the two-step sweep (each four steps pass through it and come back).
        self._swap = (torch.empty(grid.shape, **opts), torch.empty(grid.shape, **opts)) \
            if time_block == 2 else None
        is_wall, total = wall_coefficient_sums(grid, medium, self.dt, self.walls)
        self._wall_mask = int(sum(1 << code for code in range(64) if is_wall[code]))
        self._wall_table = None
        if self._wall_mask:
            table = np.concatenate([1.0 / (1.0 + total), 1.0 - total]).astype(self.np_dtype)
            self._wall_table = torch.as_tensor(table, device=self.device)
        self._sources: list[tuple[int, int, np.ndarray]] = []
        self._receivers: list[tuple[int, int]] = []
        self._receiver_first_run: list[int] = []
        self._runs: list[torch.Tensor] = []
        self._tables = None

    # -- reference API -------------------------------------------------------------------------
    @property
    def time(self) -> float:
        return self.step_index * self.dt

    @property
    def sample_rate(self) -> float:
        return 1.0 / self.dt

    def _cell_of(self, point) -> tuple[int, int]:
        i, j, k = self.grid.cell_index(point)
        return i, j * self.grid.nz + k

    def add_volume_source(self, point, signal) -> None:
        """Soft source; sample ``n`` adds ``rho c^2 dt s[n]`` to ``p^{n+1}`` as in the
        first-order scheme, which this scheme carries as ``S^n - S^{n-1}``."""
        i, jk = self._cell_of(point)
        self._sources.append((i, jk, np.asarray(signal, dtype=self.np_dtype).ravel()))
        self._tables = None

    def add_pressure_receiver(self, point) -> int:
        self._receivers.append(self._cell_of(point))
        self._receiver_first_run.append(len(self._runs))
        self._tables = None
        return len(self._receivers) - 1

    @property
    def recorded_pressure(self) -> np.ndarray:
        if not self._receivers:
            return np.zeros((0, 0))
        rows = []
        for channel, first in enumerate(self._receiver_first_run):
            parts = [run[channel] for run in self._runs[first:]]
            rows.append(torch.cat(parts) if parts else torch.zeros(0, device=self.device,
                                                                   dtype=self.dtype))
        if len({row.numel() for row in rows}) != 1:
            raise ValueError("receivers added part-way through the run have shorter rows")
        return torch.stack(rows).cpu().numpy()

    def probe_pressure(self, point) -> float:
        return float(self.p[self.grid.cell_index(point)])

    def set_state(self, p_now, p_before, step_index: int = 0) -> None:
        """Start from ``p^n = p_now`` and ``p^{n-1} = p_before`` at step ``step_index`` (e.g.
        two consecutive pressures of a first-order run)."""
        self.p.copy_(torch.as_tensor(p_now))
        self.p_previous.copy_(torch.as_tensor(p_before))
        self.step_index = int(step_index)

    def step(self) -> None:
        self.run(1)

    def run(self, n_steps: int) -> None:
        """Advance ``n_steps`` on the current stream (asynchronous)."""
        n_steps = int(n_steps)
        if n_steps <= 0:
            return
        recording = None
        if self._receivers:
            recording = torch.zeros((len(self._receivers), n_steps), device=self.device,
                                    dtype=self.dtype)
        desc = self._descriptor(recording)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        with torch.cuda.device(self.device):
            status = self._lib.acfdtd_hip_wave_run(ctypes.byref(desc), self.step_index, n_steps,
                                                   0, ctypes.c_void_p(stream))
        _native.check(status, "acfdtd_hip_wave_run")
        self.step_index += n_steps
        if n_steps % 2:
            self.p, self.p_previous = self.p_previous, self.p
        if recording is not None:
            self._runs.append(recording)

    def synchronize(self) -> None:
        torch.cuda.synchronize(self.device)

    def energy(self) -> float:
        """The discrete invariant in J-like units (see the class docstring), reduced in fp64."""
        now, before = self.p.double(), self.p_previous.double()
        coupling = 0.0
        for axis in range(3):
            n = now.shape[axis]
            coupling += float(torch.sum(torch.diff(now, dim=axis) * torch.diff(before, dim=axis))) \
                if n > 1 else 0.0
        lam2 = self.courant2
        raw = float(torch.sum((now - before) ** 2)) + lam2 * coupling
        m = self.medium
        return self.grid.dx**3 * raw / (2.0 * m.density * m.sound_speed**2 * lam2)

    # -- plumbing -------------------------------------------------------------------------------
    def _event_tables(self):
        if self._tables is None:
            cells = lengths = signal = receivers = None
            if self._sources:
                width = max(max(s.size for _, _, s in self._sources), 1)
                table = np.zeros((len(self._sources), width), dtype=self.np_dtype)
                for n, (_, _, s) in enumerate(self._sources):
                    table[n, : s.size] = s
                cells = torch.as_tensor(np.array([(i, jk) for i, jk, _ in self._sources],
                                                 dtype=np.int32).ravel(), device=self.device)
                lengths = torch.as_tensor(np.array([s.size for _, _, s in self._sources],
                                                   dtype=np.int64), device=self.device)
                signal = torch.as_tensor(table, device=self.device)
            if self._receivers:
                receivers = torch.as_tensor(np.array(self._receivers, dtype=np.int32).ravel(),
                                            device=self.device)
            self._tables = (cells, lengths, signal, receivers)
        return self._tables

    def _descriptor(self, recording) -> _native.WaveDesc:
        g = self.grid
        d = _native.WaveDesc()
        d.abi_version = _native.ABI_VERSION
        d.struct_size = ctypes.sizeof(_native.WaveDesc)
        d.nx, d.ny, d.nz = g.nx, g.ny, g.nz
        d.real_bytes = self.np_dtype.itemsize
        d.tile_j, d.chunk_x = self.tile_j, self.chunk_x
        if self._swap is not None:
            d.time_block = 2
            d.p_swap, d.p_prev_swap = (t.data_ptr() for t in self._swap)
        d.courant2 = self.courant2
        d.src_coef = self.source_coefficient
        d.wall_mask = self._wall_mask
        d.wall_table = self._wall_table.data_ptr() if self._wall_table is not None else None
        for name, tensor in (("p", self.p), ("p_prev", self.p_previous)):
            if tuple(tensor.shape) != g.shape or tensor.dtype != self.dtype \
                    or not tensor.is_contiguous() or tensor.device != self.device:
                raise ValueError(f"{name} must be a contiguous {self.dtype} tensor of shape "
                                 f"{g.shape} on {self.device}")
            setattr(d, name, tensor.data_ptr())
        cells, lengths, signal, receivers = self._event_tables()
        if cells is not None:
            d.n_sources = len(self._sources)
            d.source_cell, d.source_length = cells.data_ptr(), lengths.data_ptr()
            d.source_signal, d.source_stride = signal.data_ptr(), signal.shape[1]
        if recording is not None:
            d.n_receivers = len(self._receivers)
            d.receiver_cell = receivers.data_ptr()
            d.recording = recording.data_ptr()
            d.recording_stride = recording.shape[1]
        return d
