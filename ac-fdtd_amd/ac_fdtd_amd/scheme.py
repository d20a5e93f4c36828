"""``AcousticFDTD`` on MI355X: the reference solver API over the HIP time loop.

Drop-in for the reference's duck-typed solver classes (``src/ac_fdtd/scheme.py:82-389``,
``c_backend.py:160-382``, ``torch_backend.py:74-339``): same constructor arguments, the same
``(p^n, v^{n-1/2})`` state on public tensors, ``add_volume_source`` / ``add_pressure_receiver``
/ ``run`` / ``step`` / ``recorded_pressure`` / ``energy``. The state lives in device memory as
torch tensors; :meth:`run` hands them to the library, which advances ``n_steps`` with two fused
kernels per step on the current HIP stream. There is no CPU path: without a GPU or without the
built library the constructor raises.

Precision: ``float64`` runs the plain kernels, which reproduce the NumPy reference bit for bit;
``float32`` runs the same kernels in single precision, and ``compensated=True`` switches the
fp32 pressure increment to error-free summation (see ``csrc/sweeps.hip``).

Kernels: with ``fused=True`` (the default) the solver keeps a second set of ``p`` and velocity
buffers and the library advances steps in pairs with one fused sweep per step (A -> B -> A),
moving 8 + 2 n_relax instead of 12 + 2 n_relax field elements per cell-update; ``fused=False``
uses the two in-place sweeps. Both give the same bits, and the state always ends in the
public tensors.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .air import AirAbsorption
from .boundaries import AbsorbingLayer, WallAdmittances
from .coefficients import build_coefficients
from .grid import Grid
from .medium import AIR, Medium

__all__ = ["AcousticFDTD", "CapturedSteps", "PaddedRows", "padded_row", "synchronize"]

_TORCH_DTYPES = {np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}


def _numpy_dtype(dtype) -> np.dtype:
    if isinstance(dtype, torch.dtype):
        dtype = {torch.float32: np.float32, torch.float64: np.float64}.get(dtype, dtype)
    return np.dtype(dtype)


def synchronize() -> None:
    torch.cuda.synchronize()


def _require_gpu(device) -> torch.device:
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError(f"the HIP backend runs on a GPU device ('cuda' on ROCm), got {device}")
    if not torch.cuda.is_available():
        raise RuntimeError("no ROCm GPU is visible: the HIP backend has no CPU fallback")
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    return device


class AcousticFDTD:
    """The staggered pressure-velocity leapfrog in a rectangular room, on one GPU.

    Args:
        grid, medium, courant, walls, absorbing_layer, air_absorption: as the reference.
        dtype: ``float64`` (default, the reference's precision) or ``float32``; numpy or torch.
        device: a ROCm GPU, ``"cuda"`` / ``"cuda:N"``.
        compensated: fp32 only — error-free summation of the pressure increment.
        fused: one sweep per step through a second set of buffers (default), or two in place.
        tile_j, chunk_x: kernel tiling (rows per workgroup, planes per march); 0 = default.
        autotune: time the candidate tilings of the fused step on scratch fields of this
            configuration and keep the fastest (overrides ``tile_j`` / ``chunk_x``; see
            :mod:`ac_fdtd_amd.autotune`).
        time_block: steps per sweep over memory, 2 (default) or 1. With 2 (fused path only)
            the run advances two steps per sweep wherever it can (a second set of relaxation
            states is allocated for it); results are bit-identical to single steps. With
            ``autotune`` the faster of the two is kept.
        staged: two-step sweeps request each plane's inputs one plane ahead, straight into
            LDS, so that they travel while the previous plane's second step runs (same
            results; needs 16-byte aligned fields, which torch allocations are). With
            ``autotune`` it is chosen by timing.
    """

    def __init__(
        self,
        grid: Grid,
        medium: Medium = AIR,
        courant: float = 1.0,
        dtype=np.float64,
        walls: WallAdmittances | None = None,
        absorbing_layer: AbsorbingLayer | None = None,
        air_absorption: AirAbsorption | None = None,
        device="cuda",
        compensated: bool = False,
        fused: bool = True,
        tile_j: int = 0,
        chunk_x: int = 0,
        sponge_strength: float | None = None,
        relaxation_strengths=None,
        wall_table_always: bool = False,
        autotune: bool = False,
        time_block: int = 2,
        staged: bool = False,
    ) -> None:
        if time_block not in (1, 2):
            raise ValueError(f"time_block must be 1 or 2, got {time_block}")
        self.device = _require_gpu(device)
        self._lib = _native.library()
        self.grid = grid
        self.medium = medium
        self.courant = courant
        self.walls = walls or WallAdmittances()
        self.absorbing_layer = absorbing_layer
        self.air_absorption = air_absorption
        self.np_dtype = _numpy_dtype(dtype)
        self.coefficients = build_coefficients(
            grid,
            medium,
            courant,
            self.np_dtype,
            self.walls,
            absorbing_layer,
            air_absorption,
            sponge_strength=sponge_strength,
            relaxation_strengths=relaxation_strengths,
            wall_table_always=wall_table_always,
        )
        if compensated and self.np_dtype != np.float32:
            raise ValueError("compensated summation is an fp32 mode")
        self.compensated = compensated
        self.tile_j, self.chunk_x = tile_j, chunk_x
        # Two-step sweeps request each plane's inputs a plane ahead, into LDS (same results).
        self.staged = bool(staged)
        self.dtype = _TORCH_DTYPES[self.np_dtype]
        self.dt = self.coefficients.dt
        self.step_index = 0

        opts = {"device": self.device, "dtype": self.dtype}
        self.p = torch.zeros(grid.shape, **opts)
        self.velocity = [torch.zeros(grid.velocity_shape(a), **opts) for a in range(3)]
        k = self.coefficients
        self.relaxation_states = torch.zeros((k.n_relax, *grid.shape), **opts)
        self.fused = fused
        self.time_block = time_block if fused else 1
        self._swap = None
        self._psi_swap = None
        if fused:
            self._swap = [torch.zeros(grid.shape, **opts)] + [
                torch.zeros(grid.velocity_shape(a), **opts) for a in range(3)
            ]
            if self.time_block == 2 and k.n_relax:
                self._psi_swap = torch.zeros((k.n_relax, *grid.shape), **opts)

        self._wall_table = None
        self._wall_mask = 0
        if k.wall_is_wall is not None:
            table = np.concatenate([k.wall_from_updated, k.wall_from_previous])
            self._wall_table = torch.as_tensor(table, **opts)
            self._wall_mask = int(sum(1 << c for c in range(64) if k.wall_is_wall[c]))
        self._wall_tape: tuple[torch.Tensor, int] | None = None
        self._div_tape: tuple[torch.Tensor, int] | None = None
        self._layer_tape: tuple[torch.Tensor, int] | None = None
        self._layer_extent = None
        # A grid whose nz the widest lane does not divide runs in padded rows (PaddedRows,
        # allocated at the first run long enough to use them); False keeps the reference layout.
        self.padded_rows = True
        self._padded = None
        self._cell = self._face = None
        if k.has_layer:
            self._cell = [torch.as_tensor(x, **opts) for x in k.cell_profiles]
            self._face = [torch.as_tensor(x, **opts) for x in k.face_profiles]

        self._sources: list[tuple[int, int, np.ndarray]] = []
        self._source_tensors = None
        self._receivers: list[tuple[int, int]] = []
        self._receiver_first_run: list[int] = []  # index into _runs of each receiver's first run
        self._receiver_tensor = None
        self._runs: list[torch.Tensor] = []  # per run(): (receivers at that time, n_steps)
        self._energy_workspace = None
        if autotune:
            from .autotune import forward_plan

            self.time_block, self.tile_j, self.chunk_x, self.staged = forward_plan(self)

    # -- reference API ----------------------------------------------------------------------
    @property
    def vx(self) -> torch.Tensor:
        return self.velocity[0]

    @property
    def vy(self) -> torch.Tensor:
        return self.velocity[1]

    @property
    def vz(self) -> torch.Tensor:
        return self.velocity[2]

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
        """Soft volume source; sample ``n`` acts on the step from ``p^n`` (scheme.py:231-244)."""
        i, jk = self._cell_of(point)
        samples = np.asarray(signal, dtype=self.np_dtype).ravel()
        self._sources.append((i, jk, samples))
        self._source_tensors = None

    def add_pressure_receiver(self, point) -> int:
        """Record ``p`` at ``point`` after every later step; returns the channel (:246-255)."""
        self._receivers.append(self._cell_of(point))
        self._receiver_first_run.append(len(self._runs))
        self._receiver_tensor = None
        return len(self._receivers) - 1

    @property
    def recorded_pressure(self) -> np.ndarray:
        """``(channel, sample)`` on the host. Like the reference, receivers added part-way
        through have shorter rows, and a ragged set cannot form one array (ValueError)."""
        if not self._receivers:
            return np.zeros((0, 0))
        rows = []
        for channel, first in enumerate(self._receiver_first_run):
            parts = [run[channel] for run in self._runs[first:]]
            rows.append(torch.cat(parts) if parts else torch.zeros(0, **self._opts))
        lengths = {row.numel() for row in rows}
        if len(lengths) != 1:
            raise ValueError(
                "receivers were added part-way through the run and have different lengths "
                f"{sorted(lengths)}; read them one at a time with recorded_channel()"
            )
        return torch.stack(rows).cpu().numpy()

    def recorded_channel(self, channel: int) -> np.ndarray:
        first = self._receiver_first_run[channel]
        parts = [run[channel] for run in self._runs[first:]]
        return torch.cat(parts).cpu().numpy() if parts else np.zeros(0, dtype=self.np_dtype)

    def probe_pressure(self, point) -> float:
        return float(self.p[self.grid.cell_index(point)])

    def probe_velocity(self, point, axis: int) -> float:
        """Lower face of the cell containing ``point``, at ``time - dt/2`` (scheme.py:269-283)."""
        return float(self.velocity[axis][self.grid.cell_index(point)])

    def step(self) -> None:
        self.run(1)

    def run(self, n_steps: int, record: bool = True) -> None:
        """Advance ``n_steps`` on the current stream (asynchronous, like any GPU launch).

        ``record=False`` skips the receivers (and an attached boundary tape) for these steps,
        as a recomputation from a checkpoint does; an attached divergence tape is written
        either way."""
        launch = self.prepare_run(n_steps, record)
        if launch is not None:
            launch()

    def prepare_run(self, n_steps: int, record: bool = True):
        """:meth:`run` in two halves: this one allocates the recording and builds the
        descriptor; the returned callable (None for ``n_steps <= 0``) launches the steps on the
        current stream and does run's bookkeeping. Call it once, before any other step of this
        solver. For timing the steps alone."""
        if type(self).run is not AcousticFDTD.run:
            raise NotImplementedError(f"{type(self).__name__} advances its own scheme: use run()")
        n_steps = int(n_steps)
        if n_steps <= 0:
            return None
        recording = None
        if self._receivers and record:
            recording = torch.zeros((len(self._receivers), n_steps), **self._opts)
        padded = self._padded_rows(n_steps, record)
        if padded is not None:
            desc = padded.descriptor(recording)
        else:
            desc = self._descriptor(recording, tape=record, div=True)
        step = self.step_index

        def launch() -> None:
            if self.step_index != step:
                raise RuntimeError("the solver advanced since this run was prepared")
            stream = torch.cuda.current_stream(self.device).cuda_stream
            with torch.cuda.device(self.device):
                if padded is not None:
                    padded.load()
                status = self._lib.acfdtd_hip_run(
                    ctypes.byref(desc), step, n_steps, 0, ctypes.c_void_p(stream))
                _native.check(status, "acfdtd_hip_run")
                if padded is not None:
                    padded.store()
            self.step_index += n_steps
            if recording is not None:
                self._runs.append(recording)

        return launch

    #: Runs shorter than this stay in the reference layout: the padded rows' copies in and out
    #: would cost more than the wider lanes save.
    PADDED_MIN_STEPS = 8

    def _padded_rows(self, n_steps: int, record: bool):
        """The padded rows this run takes (allocated on first use), or None: only the fused path
        of a grid whose nz the widest lane does not divide, with no tape to write (tapes keep the
        reference layout), for at least PADDED_MIN_STEPS steps."""
        if (not self.padded_rows or self._swap is None or n_steps < self.PADDED_MIN_STEPS
                or padded_row(self) == self.grid.nz or self._div_tape is not None
                or self._layer_tape is not None
                or (record and self._wall_tape is not None)):
            return None
        if self._padded is None:
            self._padded = PaddedRows(self)
        return self._padded

    def capture(self, n_steps: int, record: bool = True) -> "CapturedSteps":
        """Capture steps ``step_index`` .. ``step_index + n_steps - 1`` of :meth:`run` into one
        HIP graph (:class:`CapturedSteps`) without running them. A replay runs those steps
        again on whatever the fields hold, e.g. after :meth:`load_state` in an optimisation
        loop that restarts every pass from the same state."""
        return CapturedSteps(self, n_steps, record)

    def timed_steps(self, n_steps: int) -> dict[str, float]:
        """Advance ``n_steps`` like :meth:`run`, one kernel per library call, with HIP events
        around every kernel on the launch stream; returns the mean duration in ms of each
        kernel kind used (``"fused2"`` for a two-step sweep, ``"fused"`` for the fused step,
        ``"velocity"`` and ``"pressure"`` for the two sweeps). The state advances exactly as
        :meth:`run` would."""
        n_steps = int(n_steps)
        if n_steps <= 0:
            return {}
        if self._wall_tape is not None:
            # Checked before the first launch, so a short tape leaves the state untouched.
            rows, step0 = self._wall_tape[0].shape[0], self._wall_tape[1]
            if self.step_index < step0 or self.step_index + n_steps - step0 > rows:
                raise ValueError(
                    f"the attached wall tape holds steps {step0} to {step0 + rows - 1}; "
                    f"steps {self.step_index} to {self.step_index + n_steps - 1} do not fit")
        recording = None
        if self._receivers:
            recording = torch.zeros((len(self._receivers), n_steps), **self._opts)
        desc = self._descriptor(recording)
        stream = torch.cuda.current_stream(self.device)
        handle = ctypes.c_void_p(stream.cuda_stream)
        ref = ctypes.byref(desc)
        launches = []  # (kind, call)
        n = 0
        if self.two_step_applies():
            quads = n_steps - n_steps % 4
            for q in range(0, quads, 4):
                for half in (0, 1):
                    launches.append(("fused2", lambda s=q + 2 * half, r=half: _native.check(
                        self._lib.acfdtd_hip_fused2_step(ref, self.step_index + s, s, r, handle),
                        "acfdtd_hip_fused2_step")))
            n = quads
        if self.fused:
            pairs = n + (n_steps - n) - (n_steps - n) % 2
            for q in range(n, pairs, 2):
                for half in (0, 1):
                    launches.append(("fused", lambda s=q + half, r=half: _native.check(
                        self._lib.acfdtd_hip_fused_step(ref, self.step_index + s, s, r, handle),
                        "acfdtd_hip_fused_step")))
            n = pairs
        for s in range(n, n_steps):
            launches.append(("velocity", lambda: _native.check(
                self._lib.acfdtd_hip_velocity_sweep(ref, handle), "acfdtd_hip_velocity_sweep")))
            launches.append(("pressure", lambda s=s: _native.check(
                self._lib.acfdtd_hip_pressure_sweep(ref, self.step_index + s, s, handle),
                "acfdtd_hip_pressure_sweep")))
        events = [torch.cuda.Event(enable_timing=True) for _ in range(len(launches) + 1)]
        with torch.cuda.device(self.device):
            events[0].record(stream)
            for e, (_, call) in enumerate(launches):
                call()
                events[e + 1].record(stream)
        events[-1].synchronize()
        self.step_index += n_steps
        if recording is not None:
            self._runs.append(recording)
        totals: dict[str, list[float]] = {}
        for e, (kind, _) in enumerate(launches):
            totals.setdefault(kind, []).append(events[e].elapsed_time(events[e + 1]))
        return {kind: sum(v) / len(v) for kind, v in totals.items()}

    def two_step_applies(self) -> bool:
        """Whether :meth:`run` advances two steps per sweep: ``time_block=2`` on the fused path,
        with at most 64 sources and 256 receivers, whose tables the sweep keeps in LDS (the
        library's own test, restated). Any nz: lanes hold 4, 2 or 1 cells (fp32), 2 or 1
        (fp64), the widest that divides nz."""
        if self.time_block != 2 or not self.fused:
            return False
        if self.coefficients.n_relax and self._psi_swap is None:
            return False
        return len(self._sources) <= 64 and len(self._receivers) <= 256

    def adjoint_two_step_applies(self) -> bool:
        """Whether the adjoint of this solver's run can advance two steps per sweep (for wall and
        signal gradients): as :meth:`two_step_applies`, with 16-byte lanes, so nz a multiple of
        4 (fp32) or 2 (fp64)."""
        vec = 4 if self.np_dtype.itemsize == 4 else 2
        return self.two_step_applies() and self.grid.nz % vec == 0

    def synchronize(self) -> None:
        torch.cuda.synchronize(self.device)

    def energy(self) -> float:
        """Acoustic energy of ``(p^n, v^{n-1/2})`` in J, reduced in fp64 on the device.

        ``dx^3 (sum p^2 / (2 rho c^2) + rho/2 sum v^2 - dt/2 sum v . grad p)``
        (scheme.py:365-389).
        """
        desc = self._descriptor(None)
        size = self._lib.acfdtd_hip_energy_workspace_bytes(ctypes.byref(desc))
        if self._energy_workspace is None or self._energy_workspace.numel() < size // 8 + 3:
            self._energy_workspace = torch.empty(size // 8 + 3, dtype=torch.float64,
                                                 device=self.device)
        out = self._energy_workspace[-3:]
        stream = torch.cuda.current_stream(self.device).cuda_stream
        with torch.cuda.device(self.device):
            status = self._lib.acfdtd_hip_energy(
                ctypes.byref(desc),
                ctypes.c_void_p(self._energy_workspace.data_ptr()),
                ctypes.c_void_p(out.data_ptr()),
                ctypes.c_void_p(stream),
            )
        _native.check(status, "acfdtd_hip_energy")
        pp, vv, cp = out.cpu().tolist()
        m, dx = self.medium, self.grid.dx
        potential = pp / (2.0 * m.density * m.sound_speed**2)
        kinetic = 0.5 * m.density * vv
        return dx**3 * (potential + kinetic - 0.5 * self.dt * (cp / dx))

    # -- plumbing ---------------------------------------------------------------------------
    @property
    def _opts(self):
        return {"device": self.device, "dtype": self.dtype}

    def _checked(self, tensor: torch.Tensor, shape, name: str) -> int:
        if tuple(tensor.shape) != tuple(shape) or tensor.dtype != self.dtype:
            raise ValueError(f"{name} must be a {self.dtype} tensor of shape {tuple(shape)}")
        if tensor.device != self.device or not tensor.is_contiguous():
            raise ValueError(f"{name} must be contiguous on {self.device}")
        return tensor.data_ptr()

    def _source_arrays(self):
        if self._source_tensors is None and self._sources:
            length = max(s.size for _, _, s in self._sources)
            signal = np.zeros((len(self._sources), max(length, 1)), dtype=self.np_dtype)
            for n, (_, _, s) in enumerate(self._sources):
                signal[n, : s.size] = s
            cells = np.array([(i, jk) for i, jk, _ in self._sources], dtype=np.int32).ravel()
            lengths = np.array([s.size for _, _, s in self._sources], dtype=np.int64)
            self._source_tensors = (
                torch.as_tensor(cells, device=self.device),
                torch.as_tensor(lengths, device=self.device),
                torch.as_tensor(signal, device=self.device),
            )
        return self._source_tensors

    # -- state, tapes and recordings for the adjoint ------------------------------------------
    def save_state(self, with_relaxation: bool = True) -> list[torch.Tensor]:
        """Copies of p, vx, vy, vz (and the relaxation states), e.g. a checkpoint."""
        fields = [self.p, *self.velocity]
        if with_relaxation:
            fields.append(self.relaxation_states)
        return [f.clone() for f in fields]

    def load_state(self, state: list[torch.Tensor], step_index: int) -> None:
        for target, source in zip([self.p, *self.velocity, self.relaxation_states], state):
            target.copy_(source)
        self.step_index = int(step_index)

    def attach_wall_tape(self, tape: torch.Tensor | None, step0: int = 0) -> None:
        """Record ``p_tmp + p_old`` on the boundary into ``tape`` (rows of
        :meth:`wall_tape_stride` elements, one per step from ``step0``); None detaches."""
        if tape is not None:
            if tape.dim() != 2 or tape.shape[1] < self.wall_tape_stride() or not tape.is_contiguous():
                raise ValueError("wall tape must be a contiguous (rows, wall_tape_stride) tensor")
            if tape.dtype != self.dtype or tape.device != self.device:
                raise ValueError(f"wall tape must be {self.dtype} on {self.device}")
        self._wall_tape = None if tape is None else (tape, int(step0))

    def attach_div_tape(self, tape: torch.Tensor | None, step0: int = 0) -> None:
        """Write every cell's velocity divergence of each step ``n`` (the one its pressure
        update uses) to row ``n - step0`` of ``tape``, a contiguous ``(rows, nx * ny * nz)``
        tensor of the working precision; None detaches. The air-strength gradients read it
        (:mod:`ac_fdtd_amd.autograd`). Needs air absorption."""
        if tape is not None:
            if not self.coefficients.n_relax:
                raise ValueError("a divergence tape needs air absorption (relaxation states)")
            if tape.dim() != 2 or tape.shape[1] != self.grid.n_cells or not tape.is_contiguous():
                raise ValueError("divergence tape must be a contiguous (rows, n_cells) tensor")
            if tape.dtype != self.dtype or tape.device != self.device:
                raise ValueError(f"divergence tape must be {self.dtype} on {self.device}")
        self._div_tape = None if tape is None else (tape, int(step0))

    def layer_extent(self) -> tuple[tuple[int, int, int], tuple[int, int, int]]:
        """Cells of the absorbing layer at the low and high end of each axis (where the grading
        weights are nonzero), the frame a layer tape holds."""
        if self.absorbing_layer is None:
            raise ValueError("the solver has no absorbing layer")
        if self._layer_extent is None:
            from .boundaries import layer_weights

            cells, faces = layer_weights(self.grid, self.absorbing_layer)
            lo, hi = [], []
            for wc, wf in zip(cells, faces):
                # A cell is in the layer when its own weight or that of a face it owns (its lower
                # face; the last cell also the far face) is nonzero.
                need = np.asarray(wc) > 0
                n = need.size
                face_on = np.asarray(wf) > 0
                need[: n] |= face_on[:n]
                need[n - 1] |= bool(face_on[n])
                lead = n if need.all() else int(np.argmin(need))
                trail = n if need.all() else int(np.argmin(need[::-1]))
                if not need.all() and need[lead : n - trail].any():
                    raise ValueError("the absorbing layer is not at the ends of an axis")
                lo.append(lead)
                hi.append(trail)
            self._layer_extent = (tuple(lo), tuple(hi))
        return self._layer_extent

    def layer_tape_stride(self) -> int:
        """Elements per row of a layer tape (:meth:`attach_layer_tape`)."""
        d = _native.Desc()
        d.nx, d.ny, d.nz = self.grid.shape
        lo, hi = self.layer_extent()
        for a in range(3):
            d.layer_lo[a], d.layer_hi[a] = lo[a], hi[a]
        return int(self._lib.acfdtd_hip_layer_tape_stride(ctypes.byref(d)))

    def attach_layer_tape(self, tape: torch.Tensor | None, step0: int = 0) -> None:
        """Write the fields of the absorbing layer's cells after each step ``n`` (pressure,
        lower face velocities, and the far faces) to row ``n - step0`` of ``tape``, a contiguous
        16-byte aligned ``(rows, layer_tape_stride())`` tensor of the working precision; None
        detaches. The sponge-strength gradient reads it (:mod:`ac_fdtd_amd.autograd`)."""
        if tape is not None:
            if self.absorbing_layer is None:
                raise ValueError("a layer tape needs an absorbing layer")
            if tape.dim() != 2 or tape.shape[1] != self.layer_tape_stride() or not tape.is_contiguous():
                raise ValueError("layer tape must be a contiguous (rows, layer_tape_stride) tensor")
            if tape.dtype != self.dtype or tape.device != self.device:
                raise ValueError(f"layer tape must be {self.dtype} on {self.device}")
            if tape.data_ptr() % 16:
                raise ValueError("layer tape must be 16-byte aligned")
        self._layer_tape = None if tape is None else (tape, int(step0))

    def wall_tape_stride(self) -> int:
        return int(self._lib.acfdtd_hip_wall_tape_stride(*self.grid.shape))

    def recording_tensor(self) -> torch.Tensor:
        """All samples so far as one ``(channel, sample)`` device tensor (equal lengths)."""
        if not self._receivers:
            return torch.zeros((0, 0), **self._opts)
        if len(set(self._receiver_first_run)) > 1:
            raise ValueError("receivers were added part-way through the run")
        return torch.cat(self._runs, dim=1) if self._runs else torch.zeros(
            (len(self._receivers), 0), **self._opts
        )

    def _descriptor(self, recording: torch.Tensor | None, tape: bool = True,
                    div: bool | None = None) -> _native.Desc:
        """The C-ABI descriptor of the current state; ``tape`` includes the boundary tape,
        ``div`` (default: as ``tape``) the divergence tape."""
        g, k = self.grid, self.coefficients
        d = _native.Desc()
        d.abi_version = _native.ABI_VERSION
        d.struct_size = ctypes.sizeof(_native.Desc)
        d.nx, d.ny, d.nz = g.nx, g.ny, g.nz
        d.real_bytes = self.np_dtype.itemsize
        d.n_relax = k.n_relax
        d.has_layer = int(k.has_layer)
        d.flags = (_native.COMPENSATED if self.compensated else 0) | (
            _native.STAGED if self.staged else 0)
        d.tile_j, d.chunk_x = self.tile_j, self.chunk_x
        d.time_block = self.time_block
        d.vel_coef, d.pres_coef, d.src_coef = k.velocity, k.pressure, k.source
        d.relax_scale = k.relax_scale
        for r in range(k.n_relax):
            d.relax_decay[r] = k.relax_decay[r]
            d.relax_mean[r] = k.relax_mean[r]
            d.relax_gain[r] = k.relax_gain[r]
        d.wall_mask = self._wall_mask
        d.wall_table = self._wall_table.data_ptr() if self._wall_table is not None else None
        d.p = self._checked(self.p, g.shape, "p")
        for a, name in enumerate(("vx", "vy", "vz")):
            setattr(d, name, self._checked(self.velocity[a], g.velocity_shape(a), name))
        if k.n_relax:
            d.psi = self._checked(self.relaxation_states, (k.n_relax, *g.shape), "psi")
        if k.has_layer:
            for a in range(3):
                d.cell[a] = self._cell[a].data_ptr()
                d.face[a] = self._face[a].data_ptr()
        sources = self._source_arrays()
        if sources is not None:
            cells, lengths, signal = sources
            d.n_sources = len(self._sources)
            d.source_cell = cells.data_ptr()
            d.source_length = lengths.data_ptr()
            d.source_signal = signal.data_ptr()
            d.source_stride = signal.shape[1]
        if recording is not None:
            if self._receiver_tensor is None:
                cells = np.array(self._receivers, dtype=np.int32).ravel()
                self._receiver_tensor = torch.as_tensor(cells, device=self.device)
            d.n_receivers = len(self._receivers)
            d.receiver_cell = self._receiver_tensor.data_ptr()
            d.recording = recording.data_ptr()
            d.recording_stride = recording.shape[1]
        if tape and self._wall_tape is not None:
            buffer, step0 = self._wall_tape
            d.wall_tape = buffer.data_ptr()
            d.wall_tape_stride = buffer.shape[1]
            d.wall_tape_step0 = step0
            d.wall_tape_rows = buffer.shape[0]
        if (tape if div is None else div) and self._div_tape is not None:
            buffer, step0 = self._div_tape
            d.div_tape = buffer.data_ptr()
            d.div_tape_step0 = step0
            d.div_tape_rows = buffer.shape[0]
        if (tape if div is None else div) and self._layer_tape is not None:
            buffer, step0 = self._layer_tape
            d.layer_tape = buffer.data_ptr()
            d.layer_tape_step0 = step0
            d.layer_tape_rows = buffer.shape[0]
            d.layer_tape_stride = buffer.shape[1]
            lo, hi = self.layer_extent()
            for a in range(3):
                d.layer_lo[a], d.layer_hi[a] = lo[a], hi[a]
        if self._swap is not None:
            d.p_swap = self._checked(self._swap[0], g.shape, "p swap")
            for a, name in enumerate(("vx_swap", "vy_swap", "vz_swap")):
                setattr(d, name, self._checked(self._swap[a + 1], g.velocity_shape(a), name))
        if self._psi_swap is not None:
            d.psi_swap = self._checked(self._psi_swap, (k.n_relax, *g.shape), "psi swap")
        return d


def padded_row(solver) -> int:
    """Row length of ``solver``'s fields in padded rows: nz rounded up to the widest lane of its
    precision (4 cells fp32, 2 fp64); nz itself when that width divides it."""
    widest = 4 if solver.np_dtype.itemsize == 4 else 2
    return -(-solver.grid.nz // widest) * widest


class PaddedRows:
    """A solver's state, swap set and cell tables in rows of ``padded_row`` elements (every
    field, vz included): the layout ``acfdtd_hip_run`` and the batch API take as the
    descriptor's ``z_pitch``, so that a room whose nz is not a multiple of the widest lane (4
    cells fp32, 2 fp64) advances on full-width lanes. The elements past the room's own in each
    row are padding, which the kernels never let reach a cell. :meth:`load` and :meth:`store`
    copy the state in and out (stream-ordered device copies)."""

    def __init__(self, solver) -> None:
        self.solver = solver
        g = solver.grid
        self.zp = padded_row(solver)
        opts = solver._opts
        zp = self.zp

        def fields():
            return [torch.zeros((g.nx, g.ny, zp), **opts),
                    torch.zeros((g.nx + 1, g.ny, zp), **opts),
                    torch.zeros((g.nx, g.ny + 1, zp), **opts),
                    torch.zeros((g.nx, g.ny, zp), **opts)]

        n_relax = solver.coefficients.n_relax
        self.state = fields()
        self.psi = torch.zeros((n_relax, g.nx, g.ny, zp), **opts) if n_relax else None
        self.swap = fields() if solver._swap is not None else None
        self.psi_swap = (torch.zeros((n_relax, g.nx, g.ny, zp), **opts)
                         if solver._psi_swap is not None else None)
        self._tables: dict[tuple, torch.Tensor] = {}

    def _cells(self, cells) -> torch.Tensor:
        """(i, j * nz + k) pairs re-addressed to rows of zp: (i, j * zp + k), on the device
        (kept, so that a table a captured graph's room table points at stays alive)."""
        key = tuple(cells)
        if key not in self._tables:
            nz = self.solver.grid.nz
            table = np.array([(i, (jk // nz) * self.zp + jk % nz) for i, jk in cells],
                             dtype=np.int32).ravel()
            self._tables[key] = torch.as_tensor(table, device=self.solver.device)
        return self._tables[key]

    def _views(self):
        """(padded view, room tensor) pairs of the state that crosses a run's boundaries."""
        s, nz = self.solver, self.solver.grid.nz
        pairs = [(self.state[0][..., :nz], s.p)]
        pairs += [(self.state[a + 1][..., : nz + (a == 2)], s.velocity[a]) for a in range(3)]
        if self.psi is not None:
            pairs.append((self.psi[..., :nz], s.relaxation_states))
        return pairs

    def load(self) -> None:
        """Copy the room's state into the padded rows."""
        for padded, room in self._views():
            padded.copy_(room)

    def store(self) -> None:
        """Copy the padded rows' state back into the room's tensors."""
        for padded, room in self._views():
            room.copy_(padded)

    def descriptor(self, recording) -> _native.Desc:
        s = self.solver
        d = s._descriptor(recording, tape=False)
        d.z_pitch = self.zp
        d.p, d.vx, d.vy, d.vz = (t.data_ptr() for t in self.state)
        if self.psi is not None:
            d.psi = self.psi.data_ptr()
        if self.swap is not None:
            d.p_swap, d.vx_swap, d.vy_swap, d.vz_swap = (t.data_ptr() for t in self.swap)
        if self.psi_swap is not None:
            d.psi_swap = self.psi_swap.data_ptr()
        if s._sources:
            d.source_cell = self._cells([(i, jk) for i, jk, _ in s._sources]).data_ptr()
        if recording is not None and s._receivers:
            d.receiver_cell = self._cells(s._receivers).data_ptr()
        return d


class CapturedSteps:
    """A step loop of :class:`AcousticFDTD` captured into one HIP graph.

    The graph holds every launch :meth:`AcousticFDTD.run` makes for ``n_steps`` steps from the
    solver's ``step_index`` at capture time, with the field, swap, relaxation, source and
    receiver buffers of that time: :meth:`replay` runs the same steps (the same source samples
    and recording columns) on the fields' current contents and writes the receivers' samples
    into :attr:`recording` (``(channel, n_steps)``). It changes neither the solver's
    ``step_index`` nor its recordings, and must not outlive the buffers it captured (a later
    :meth:`AcousticFDTD.add_volume_source` or receiver replaces the source and receiver
    tables). A boundary tape is not captured. The graph is submitted as one unit; the kernels
    are those of :meth:`AcousticFDTD.run`, so a replay computes the same bits.
    """

    def __init__(self, solver: AcousticFDTD, n_steps: int, record: bool = True) -> None:
        n_steps = int(n_steps)
        if n_steps <= 0:
            raise ValueError("capture needs at least one step")
        self.solver = solver
        self.n_steps = n_steps
        self.step0 = solver.step_index
        self.recording = None
        if solver._receivers and record:
            self.recording = torch.zeros((len(solver._receivers), n_steps), **solver._opts)
        # Device tables (sources, receivers) are uploaded here, before the capture: a capture
        # records no host-to-device copies.
        solver._source_arrays()
        self._desc = solver._descriptor(self.recording, tape=False)
        self.graph = torch.cuda.CUDAGraph()
        lib = solver._lib
        with torch.cuda.device(solver.device):
            capture_stream = torch.cuda.Stream(solver.device)
            capture_stream.wait_stream(torch.cuda.current_stream(solver.device))
            with torch.cuda.graph(self.graph, stream=capture_stream):
                handle = ctypes.c_void_p(torch.cuda.current_stream(solver.device).cuda_stream)
                status = lib.acfdtd_hip_run(ctypes.byref(self._desc), self.step0, n_steps, 0,
                                            handle)
            torch.cuda.current_stream(solver.device).wait_stream(capture_stream)
        _native.check(status, "acfdtd_hip_run (captured)")

    def replay(self) -> None:
        """Run the captured steps on the current stream (asynchronous)."""
        with torch.cuda.device(self.solver.device):
            self.graph.replay()
