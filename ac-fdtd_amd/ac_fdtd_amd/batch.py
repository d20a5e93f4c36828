"""Many rooms, one launch per step: :class:`BatchedRooms` over a list of solvers.

Each room is an ordinary :class:`~ac_fdtd_amd.scheme.AcousticFDTD` (its own grid, walls, air,
sources, receivers and state tensors); the batch advances them together through the library's
batch API (``acfdtd_hip_batch_*`` in ``include/acfdtd_hip.h``), one fused launch per step over
a tile table that covers every room. Results are bit-identical to running each room alone.
There is no reference counterpart: the reference runs one room at a time.
"""

from __future__ import annotations

import ctypes

import torch

from . import _native
from .scheme import AcousticFDTD

__all__ = ["BatchedRooms", "lane_width"]


def lane_width(solver: AcousticFDTD) -> int:
    """z cells per lane the fused kernels use for this room (4 / 2 / 1 as nz allows)."""
    widest = 4 if solver.np_dtype.itemsize == 4 else 2
    width = widest
    while width > 1 and solver.grid.nz % width:
        width //= 2
    return width


class _Group:
    """One library batch handle: rooms that share a lane width."""

    def __init__(self, solvers: list[AcousticFDTD], lib) -> None:
        self.solvers = solvers
        self.device = solvers[0].device
        self._lib = lib
        self._handle = None
        descs = self._descriptors(None)
        size = lib.acfdtd_hip_batch_workspace_bytes(descs, len(solvers))
        if size == 0:
            raise RuntimeError(
                "these rooms cannot be batched: " + lib.acfdtd_hip_last_error().decode(errors="replace")
            )
        self._workspace = torch.empty(size, dtype=torch.uint8, device=self.device)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            status = lib.acfdtd_hip_batch_create(
                descs, len(solvers), ctypes.c_void_p(self._workspace.data_ptr()), size,
                ctypes.c_void_p(_stream(self.device)), ctypes.byref(handle))
        _native.check(status, "acfdtd_hip_batch_create")
        self._handle = handle

    @property
    def two_step(self) -> bool:
        """Whether this group advances two steps per sweep."""
        return bool(self._lib.acfdtd_hip_batch_two_step(self._handle))

    def _descriptors(self, recordings):
        array = (_native.Desc * len(self.solvers))()
        for n, solver in enumerate(self.solvers):
            array[n] = solver._descriptor(None if recordings is None else recordings[n], tape=False)
        return array

    def run(self, step_index: int, n_steps: int) -> list:
        recordings = [
            torch.zeros((len(s._receivers), n_steps), **s._opts) if s._receivers else None
            for s in self.solvers
        ]
        descs = self._descriptors(recordings)
        stream = ctypes.c_void_p(_stream(self.device))
        with torch.cuda.device(self.device):
            _native.check(self._lib.acfdtd_hip_batch_update(self._handle, descs,
                                                            len(self.solvers), stream),
                          "acfdtd_hip_batch_update")
            _native.check(self._lib.acfdtd_hip_batch_run(self._handle, step_index, n_steps, 0,
                                                         stream), "acfdtd_hip_batch_run")
        return recordings

    def close(self) -> None:
        if self._handle is not None:
            self._lib.acfdtd_hip_batch_destroy(self._handle)
            self._handle = None


def _stream(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _bind(lib) -> None:
    lib.acfdtd_hip_batch_workspace_bytes.restype = ctypes.c_size_t
    lib.acfdtd_hip_batch_workspace_bytes.argtypes = [ctypes.POINTER(_native.Desc), ctypes.c_int32]
    lib.acfdtd_hip_batch_create.restype = ctypes.c_int
    lib.acfdtd_hip_batch_create.argtypes = [
        ctypes.POINTER(_native.Desc), ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t,
        ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    lib.acfdtd_hip_batch_update.restype = ctypes.c_int
    lib.acfdtd_hip_batch_update.argtypes = [
        ctypes.c_void_p, ctypes.POINTER(_native.Desc), ctypes.c_int32, ctypes.c_void_p]
    lib.acfdtd_hip_batch_run.restype = ctypes.c_int
    lib.acfdtd_hip_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                         ctypes.c_int64, ctypes.c_void_p]
    lib.acfdtd_hip_batch_two_step.restype = ctypes.c_int
    lib.acfdtd_hip_batch_two_step.argtypes = [ctypes.c_void_p]
    lib.acfdtd_hip_batch_destroy.restype = None
    lib.acfdtd_hip_batch_destroy.argtypes = [ctypes.c_void_p]


class BatchedRooms:
    """Advance ``solvers`` together. They must share device, precision, relaxation count,
    sponge presence, compensation and tiling, use the fused path, and be at the same step.

    Rooms are grouped by lane width (how many z cells each lane takes, set by nz): one launch
    per group and step, so uneven rooms keep the widest accesses their rows allow.
    """

    def __init__(self, solvers: list[AcousticFDTD]) -> None:
        if not solvers:
            raise ValueError("a batch needs at least one room")
        first = solvers[0]
        for solver in solvers:
            if solver.device != first.device or solver.dtype != first.dtype:
                raise ValueError("rooms of a batch must share device and dtype")
            if not solver.fused:
                raise ValueError("batched rooms use the fused path (fused=True)")
            if solver.step_index != first.step_index:
                raise ValueError("rooms of a batch must be at the same step")
        self.solvers = list(solvers)
        self._lib = _native.library()
        _bind(self._lib)
        by_width: dict[int, list[AcousticFDTD]] = {}
        for solver in solvers:
            by_width.setdefault(lane_width(solver), []).append(solver)
        self._groups = []
        try:
            for group in by_width.values():
                self._groups.append(_Group(group, self._lib))
        except Exception:
            self.close()
            raise

    @property
    def step_index(self) -> int:
        return self.solvers[0].step_index

    @property
    def two_step(self) -> list[bool]:
        """Per lane-width group, whether it advances two steps per sweep."""
        return [group.two_step for group in self._groups]

    @property
    def n_cells(self) -> int:
        return sum(solver.grid.n_cells for solver in self.solvers)

    def run(self, n_steps: int) -> None:
        """Advance every room ``n_steps``; receivers record as with ``AcousticFDTD.run``."""
        n_steps = int(n_steps)
        if n_steps <= 0:
            return
        if not self._groups:
            raise RuntimeError("the batch has been closed")
        step = self.step_index
        for group in self._groups:
            recordings = group.run(step, n_steps)
            for solver, recording in zip(group.solvers, recordings):
                solver.step_index += n_steps
                if recording is not None:
                    solver._runs.append(recording)

    def close(self) -> None:
        for group in getattr(self, "_groups", []):
            group.close()
        self._groups = []

    def __del__(self):
        self.close()
