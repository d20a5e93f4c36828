"""Many rooms, one launch per step: :class:`BatchedRooms` over a list of solvers.

Each room is an ordinary :class:`~ac_fdtd_amd.scheme.AcousticFDTD` (its own grid, walls, air,
sources, receivers and state tensors); the batch advances them together through the library's
batch API (``acfdtd_hip_batch_*`` in ``include/acfdtd_hip.h``), one launch per two steps over a
tile table that covers every room. Every room runs on the widest lanes of its precision (4 z
cells fp32, 2 fp64): a room whose nz that width does not divide is advanced in rows padded to a
multiple of it (the descriptor's ``z_pitch``), in buffers the batch owns, and its state is copied
into them at the start of :meth:`BatchedRooms.run` and back at the end, so the room's own tensors
keep the reference layout. Results are bit-identical to running each room alone.
:meth:`BatchedRooms.capture` records a step loop into a HIP graph. There is no reference
counterpart: the reference runs one room at a time.
"""

from __future__ import annotations

import ctypes

import torch

from . import _native
from .scheme import AcousticFDTD, PaddedRows, padded_row

__all__ = ["BatchedRooms", "CapturedBatch", "lane_width", "padded_row"]


def lane_width(solver: AcousticFDTD) -> int:
    """z cells per lane the single-room fused kernels use for this room (4 / 2 / 1 as nz
    allows)."""
    widest = 4 if solver.np_dtype.itemsize == 4 else 2
    width = widest
    while width > 1 and solver.grid.nz % width:
        width //= 2
    return width


class _Group:
    """One library batch handle over ``solvers``; rooms whose nz the lane width does not
    divide advance in padded rows (:class:`~ac_fdtd_amd.scheme.PaddedRows`)."""

    def __init__(self, solvers: list[AcousticFDTD], lib) -> None:
        self.solvers = solvers
        self._padded = [PaddedRows(s) if padded_row(s) != s.grid.nz else None for s in solvers]
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
        for n, (solver, padded) in enumerate(zip(self.solvers, self._padded)):
            recording = None if recordings is None else recordings[n]
            array[n] = (padded.descriptor(recording) if padded is not None
                        else solver._descriptor(recording, tape=False))
        return array

    def load(self) -> None:
        """Copy the padded rooms' state into their padded rows (stream-ordered)."""
        for padded in self._padded:
            if padded is not None:
                padded.load()

    def store(self) -> None:
        """Copy the padded rooms' state back into their own tensors (stream-ordered)."""
        for padded in self._padded:
            if padded is not None:
                padded.store()

    def upload(self, recordings) -> None:
        """Point the device room table at the rooms' current buffers and ``recordings``
        (synchronises the stream)."""
        descs = self._descriptors(recordings)
        stream = ctypes.c_void_p(_stream(self.device))
        with torch.cuda.device(self.device):
            _native.check(self._lib.acfdtd_hip_batch_update(self._handle, descs,
                                                            len(self.solvers), stream),
                          "acfdtd_hip_batch_update")

    def launch(self, step_index: int, n_steps: int) -> None:
        stream = ctypes.c_void_p(_stream(self.device))
        with torch.cuda.device(self.device):
            _native.check(self._lib.acfdtd_hip_batch_run(self._handle, step_index, n_steps, 0,
                                                         stream), "acfdtd_hip_batch_run")

    def run(self, step_index: int, n_steps: int) -> list:
        recordings = [
            torch.zeros((len(s._receivers), n_steps), **s._opts) if s._receivers else None
            for s in self.solvers
        ]
        self.upload(recordings)
        with torch.cuda.device(self.device):
            self.load()
            self.launch(step_index, n_steps)
            self.store()
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


def _kernel_key(solver: AcousticFDTD) -> tuple:
    """What rooms of one launch must share: the kernel's template (relaxation count, sponge,
    compensation) and the workgroup shape."""
    return (solver.coefficients.n_relax, bool(solver.coefficients.has_layer), solver.compensated,
            solver.tile_j, solver._psi_swap is not None)


class BatchedRooms:
    """Advance ``solvers`` together. They must share device and precision, use the fused path,
    and be at the same step.

    Every room takes the widest lanes of the precision (rooms whose nz it does not divide in
    padded rows, see :func:`padded_row`). Rooms that share the kernel's configuration
    (relaxation count, sponge presence, compensation, tiling) advance in one launch per two
    steps, one launch per configuration: the 64 rooms of one air model and wall model, whatever
    their sizes, in one.
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
        self._tables = None  # the CapturedBatch whose buffers the room tables point at, if any
        self._lib = _native.library()
        _bind(self._lib)
        by_kernel: dict[tuple, list[AcousticFDTD]] = {}
        for solver in self.solvers:
            by_kernel.setdefault(_kernel_key(solver), []).append(solver)
        self._groups = []
        try:
            for group in by_kernel.values():
                self._groups.append(_Group(group, self._lib))
        except Exception:
            self.close()
            raise

    @property
    def step_index(self) -> int:
        return self.solvers[0].step_index

    @property
    def two_step(self) -> list[bool]:
        """Per launch group (one per kernel configuration), whether it advances two steps per
        sweep."""
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
        self._tables = None
        for group in self._groups:
            recordings = group.run(step, n_steps)
            for solver, recording in zip(group.solvers, recordings):
                solver.step_index += n_steps
                if recording is not None:
                    solver._runs.append(recording)

    def capture(self, n_steps: int) -> "CapturedBatch":
        """Capture steps ``step_index`` .. ``step_index + n_steps - 1`` of :meth:`run` into one
        HIP graph (:class:`CapturedBatch`) without running them."""
        return CapturedBatch(self, n_steps)

    def close(self) -> None:
        for group in getattr(self, "_groups", []):
            group.close()
        self._groups = []

    def __del__(self):
        self.close()


class CapturedBatch:
    """A step loop of :class:`BatchedRooms` captured into one HIP graph.

    The graph holds every launch of ``n_steps`` steps of all groups from the batch's step at
    capture time, with the copies of padded rooms' state into their padded rows and back. The launches read each group's device room table (the rooms' buffers and
    where their receivers record); the table is pointed at this capture's recordings before
    the capture and again before a replay that follows an eager :meth:`BatchedRooms.run`
    (which re-points it, and synchronises the stream to do so). :meth:`replay` runs the same
    steps on the rooms' current fields and writes :attr:`recordings` (one
    ``(channel, n_steps)`` tensor or None per room, in the batch's order); it changes neither
    the rooms' ``step_index`` nor their recordings.
    """

    def __init__(self, batch: BatchedRooms, n_steps: int) -> None:
        n_steps = int(n_steps)
        if n_steps <= 0:
            raise ValueError("capture needs at least one step")
        if not batch._groups:
            raise RuntimeError("the batch has been closed")
        self.batch = batch
        self.n_steps = n_steps
        self.step0 = batch.step_index
        self.recordings = [
            torch.zeros((len(s._receivers), n_steps), **s._opts) if s._receivers else None
            for s in batch.solvers
        ]
        index = {id(s): n for n, s in enumerate(batch.solvers)}
        self._group_recordings = [[self.recordings[index[id(s)]] for s in group.solvers]
                                  for group in batch._groups]
        self._point_tables()
        device = batch.solvers[0].device
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.device(device):
            capture_stream = torch.cuda.Stream(device)
            capture_stream.wait_stream(torch.cuda.current_stream(device))
            with torch.cuda.graph(self.graph, stream=capture_stream):
                for group in batch._groups:
                    group.load()  # padded rooms: their state into the padded rows, and back
                    group.launch(self.step0, n_steps)
                    group.store()
            torch.cuda.current_stream(device).wait_stream(capture_stream)

    def _point_tables(self) -> None:
        for group, recordings in zip(self.batch._groups, self._group_recordings):
            group.upload(recordings)
        self.batch._tables = self

    def replay(self) -> None:
        """Run the captured steps on the current stream (asynchronous unless the room tables
        had to be re-pointed first)."""
        if not self.batch._groups:
            raise RuntimeError("the batch has been closed")
        if self.batch._tables is not self:
            self._point_tables()
        with torch.cuda.device(self.batch.solvers[0].device):
            self.graph.replay()
