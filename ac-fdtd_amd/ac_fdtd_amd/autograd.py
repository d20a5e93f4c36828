"""Gradients of receiver recordings with respect to the room's parameters, on the GPU.

:class:`RoomResponse` maps parameters to the ``(receiver, step)`` pressure recording of one
room through the HIP time loop, as a differentiable torch operation. Differentiable inputs:

* ``admittance`` — the six normalised wall admittances ``xi`` (FACES order);
* ``signal`` — the source signals, ``(n_sources, n_samples)``;
* ``sponge_strength`` — ``sigma_max`` of the absorbing layer, 1/s;
* ``air_strength`` — the relaxation strengths ``Delta`` of the air model, one per process.

The backward pass is the hand-written adjoint of ``csrc/adjoint.hip``: the transposed,
time-reversed leapfrog, two kernels per step like the forward. The adjoint recurrence needs no
forward state (the step is linear in it); the parameter gradients do:

* walls need ``p_tmp + p_old`` on the boundary only, which the forward kernels write to a tape
  of ``2 (ny nz + nx nz + nx ny)`` values per step, so wall and source gradients cost one
  forward pass plus one adjoint pass;
* the air strengths need every cell's velocity divergence at every step (one value per cell
  and step). The forward kernels write it to a divergence tape when ``n_steps`` rows of it fit
  in ``history_bytes``, and the adjoint reads it, two steps per sweep like the wall gradient:
  one forward and one adjoint pass again. Beyond that budget the forward pass keeps a
  checkpoint of the full state every ``checkpoint_every`` steps, and the backward pass
  recomputes each segment's divergences (two steps per sweep) into a ring of
  ``checkpoint_every`` rows before running the adjoint through it;
* the sponge strength needs the damped forward fields of the layer's cells at every step. When
  ``n_steps`` rows of them (a layer tape, about 16 B per layer cell and step in fp32), plus the
  divergence tape if the air gradients are asked for too, fit in ``history_bytes`` and the
  adjoint runs two steps per sweep, the forward kernels write them and the two-step adjoint
  reads them: one forward and one adjoint pass again (the last one to four adjoint steps, which
  run one per sweep, read states unpacked from the tape). Otherwise the forward pass keeps
  checkpoints every ``checkpoint_every`` steps, and each segment's fields are recomputed into
  a ring of states before the adjoint (one step per sweep) runs through it; requested together
  with the sponge, the air gradients then take the same path.

There is no reference counterpart: the reference's PyTorch backend uses ``out=`` operations and
host-float coefficients and cannot be differentiated (``src/ac_fdtd/torch_backend.py:149-195``,
``:278-309``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _native
from .air import AirAbsorption
from .boundaries import AbsorbingLayer, WallAdmittances, layer_weights
from .coefficients import build_coefficients
from .grid import Grid
from .medium import AIR, Medium
from .scheme import AcousticFDTD

__all__ = ["RoomResponse", "GradientReport"]


@dataclass
class GradientReport:
    """What the last backward pass did: adjoint steps, recomputed steps, state checkpoints,
    whether the adjoint ran two steps per sweep."""

    adjoint_steps: int = 0
    adjoint_two_step: bool = False
    recomputed_steps: int = 0
    checkpoints: int = 0
    checkpoint_bytes: int = 0
    tape_bytes: int = 0
    div_tape_bytes: int = 0  # divergence rows held for the air gradients (tape or ring)
    layer_tape_bytes: int = 0  # layer-cell fields held for the sponge gradient


class RoomResponse:
    """A differentiable room: parameters in, receiver recording out.

    Args:
        grid, medium, courant, walls, absorbing_layer, air_absorption: as
            :class:`~ac_fdtd_amd.scheme.AcousticFDTD`; these fix everything that is not passed
            to :meth:`__call__`.
        n_steps: steps per evaluation.
        sources: ``[(point, signal), ...]``; the signals are the defaults for ``signal``.
        receivers: receiver points.
        checkpoint_every: state checkpoint interval K, used beyond ``history_bytes``: for the
            sponge gradient (memory: ``ceil(n_steps / K)`` full states plus ``K + 1`` field
            sets) and for the air gradients (``ceil(n_steps / K)`` full states plus ``K``
            divergence rows).
        history_bytes: device memory the forward pass may keep for the volumetric gradients
            (default 64 GiB, of the MI355X's 288 GB): a divergence tape of ``n_steps * n_cells``
            values for the air strengths, and with the sponge strength a layer tape of
            ``n_steps`` rows (``solver.layer_tape_stride()`` values each) as well, when they fit;
            else checkpoints and recomputation.
        autotune: pick the tilings of the forward and of the adjoint step by timing them
            (:mod:`ac_fdtd_amd.autotune`); gradients then differ from an untuned run only in
            the summation order of their fp64 partials.
        time_block: 2 (default) advances the forward pass, and the adjoint pass when only
            wall and signal gradients are asked for, two steps per sweep where the grid allows
            (same results; the adjoint's fp64 partials summed in another order); 1 one step.
    """

    def __init__(
        self,
        grid: Grid,
        n_steps: int,
        sources,
        receivers,
        medium: Medium = AIR,
        courant: float = 1.0,
        dtype=np.float32,
        walls: WallAdmittances | None = None,
        absorbing_layer: AbsorbingLayer | None = None,
        air_absorption: AirAbsorption | None = None,
        device="cuda",
        checkpoint_every: int = 100,
        history_bytes: int = 64 << 30,
        fused: bool = True,
        compensated: bool = False,
        autotune: bool = False,
        time_block: int = 2,
    ) -> None:
        if n_steps < 1:
            raise ValueError("n_steps must be >= 1")
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.grid, self.medium, self.courant = grid, medium, courant
        self.n_steps = int(n_steps)
        self.dtype = np.dtype(dtype) if not isinstance(dtype, torch.dtype) else (
            np.dtype(np.float32) if dtype == torch.float32 else np.dtype(np.float64)
        )
        self.walls = walls or WallAdmittances()
        self.absorbing_layer = absorbing_layer
        self.air_absorption = air_absorption
        self.device = device
        self.checkpoint_every = int(checkpoint_every)
        self.history_bytes = int(history_bytes)
        self.fused, self.compensated = fused, compensated
        self.autotune = autotune
        if time_block not in (1, 2):
            raise ValueError("time_block must be 1 or 2")
        self.time_block = int(time_block)
        self.source_points = [point for point, _ in sources]
        self.signals = [np.asarray(signal, dtype=np.float64) for _, signal in sources]
        self.receivers = list(receivers)
        base = build_coefficients(grid, medium, courant, np.float64, self.walls,
                                  absorbing_layer, air_absorption)
        self.dt = base.dt
        self.base_sponge_strength = base.sponge_strength
        self.base_air_strengths = [q.strength for q in base.relaxation]
        self.report = GradientReport()

    # -- defaults for the parameters, as tensors ---------------------------------------------
    def default_admittance(self, **options) -> torch.Tensor:
        return torch.tensor(self.walls.values, dtype=torch.float64, **options)

    def default_signal(self, **options) -> torch.Tensor:
        length = max((s.size for s in self.signals), default=0)
        out = torch.zeros((len(self.signals), length), dtype=torch.float64, **options)
        for n, s in enumerate(self.signals):
            out[n, : s.size] = torch.as_tensor(s)
        return out

    def default_sponge_strength(self, **options) -> torch.Tensor:
        if self.base_sponge_strength is None:
            raise ValueError("the room has no absorbing layer")
        return torch.tensor(self.base_sponge_strength, dtype=torch.float64, **options)

    def default_air_strength(self, **options) -> torch.Tensor:
        if not self.base_air_strengths:
            raise ValueError("the room has no air absorption")
        return torch.tensor(self.base_air_strengths, dtype=torch.float64, **options)

    def __call__(self, admittance=None, signal=None, sponge_strength=None, air_strength=None):
        """The ``(n_receivers, n_steps)`` recording, differentiable in every tensor given."""
        return _RoomResponseFunction.apply(self, admittance, signal, sponge_strength, air_strength)

    # -- building blocks ------------------------------------------------------------------------
    def solver(self, admittance=None, signal=None, sponge_strength=None, air_strength=None,
               wall_table_always: bool = False) -> AcousticFDTD:
        """A forward solver with these parameter values (plain floats / arrays or tensors)."""
        xi = self.walls.values if admittance is None else tuple(
            float(x) for x in torch.as_tensor(admittance).detach().cpu().double().tolist()
        )
        sigma = None if sponge_strength is None else float(torch.as_tensor(sponge_strength))
        delta = None if air_strength is None else [
            float(x) for x in torch.as_tensor(air_strength).detach().cpu().double().tolist()
        ]
        solver = AcousticFDTD(
            self.grid,
            medium=self.medium,
            courant=self.courant,
            dtype=self.dtype,
            walls=WallAdmittances(*xi),
            absorbing_layer=self.absorbing_layer,
            air_absorption=self.air_absorption,
            device=self.device,
            compensated=self.compensated,
            fused=self.fused,
            sponge_strength=sigma,
            relaxation_strengths=delta,
            wall_table_always=wall_table_always,
            autotune=self.autotune and not wall_table_always,
            time_block=self.time_block,
        )
        if self.autotune and wall_table_always:
            # The wall gradient's forward pass writes the boundary tape: tune with one.
            from .autotune import forward_plan

            solver.time_block, solver.tile_j, solver.chunk_x, solver.staged = forward_plan(
                solver, tape=True)
        signals = self.signals if signal is None else list(
            torch.as_tensor(signal).detach().cpu().double().numpy()
        )
        if len(signals) != len(self.source_points):
            raise ValueError(f"expected {len(self.source_points)} source signals")
        for point, s in zip(self.source_points, signals):
            solver.add_volume_source(point, s)
        for point in self.receivers:
            solver.add_pressure_receiver(point)
        return solver


class _RoomResponseFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, room: RoomResponse, admittance, signal, sponge_strength, air_strength):
        need = ctx.needs_input_grad
        want_walls = admittance is not None and need[1]
        want_signal = need[2]
        want_sponge = sponge_strength is not None and need[3]
        want_air = air_strength is not None and need[4]
        if want_sponge and room.absorbing_layer is None:
            raise ValueError("sponge_strength given but the room has no absorbing layer")
        if want_air and room.air_absorption is None:
            raise ValueError("air_strength given but the room has no air absorption")
        solver = room.solver(admittance, signal, sponge_strength, air_strength,
                             wall_table_always=want_walls)
        report = GradientReport()
        if want_walls:
            tape = torch.empty((room.n_steps, solver.wall_tape_stride()), dtype=solver.dtype,
                               device=solver.device)
            solver.attach_wall_tape(tape, 0)
            report.tape_bytes = tape.numel() * tape.element_size()
        checkpoints = []
        div_tape = None
        tape_bytes = room.n_steps * room.grid.n_cells * solver.np_dtype.itemsize
        layer_bytes = None
        if (want_sponge and room.fused and room.time_block == 2
                and solver.adjoint_two_step_applies()):
            try:
                layer_bytes = room.n_steps * solver.layer_tape_stride() * solver.np_dtype.itemsize
            except ValueError:
                layer_bytes = None  # a layer not at the ends of its axes: the ring path
        layered = layer_bytes is not None and (
            layer_bytes + (tape_bytes if want_air else 0) <= room.history_bytes)
        if layered:
            # The layer cells' fields (and the divergences) of every step, for the two-step
            # adjoint to read: no recomputation.
            layer_tape = torch.empty((room.n_steps, solver.layer_tape_stride()),
                                     dtype=solver.dtype, device=solver.device)
            solver.attach_layer_tape(layer_tape, 0)
            report.layer_tape_bytes = layer_bytes
            if want_air:
                div_tape = torch.empty((room.n_steps, room.grid.n_cells), dtype=solver.dtype,
                                       device=solver.device)
                solver.attach_div_tape(div_tape, 0)
                report.div_tape_bytes = tape_bytes
            solver.run(room.n_steps)
        elif want_air and not want_sponge and tape_bytes <= room.history_bytes:
            # Every step's divergences, for the adjoint to read: no recomputation.
            div_tape = torch.empty((room.n_steps, room.grid.n_cells), dtype=solver.dtype,
                                   device=solver.device)
            solver.attach_div_tape(div_tape, 0)
            solver.run(room.n_steps)
            report.div_tape_bytes = tape_bytes
        elif want_sponge or want_air:
            K = room.checkpoint_every
            for start in range(0, room.n_steps, K):
                checkpoints.append(solver.save_state())
                solver.run(min(K, room.n_steps - start))
            report.checkpoints = len(checkpoints)
            report.checkpoint_bytes = sum(t.numel() * t.element_size() for c in checkpoints for t in c)
        else:
            solver.run(room.n_steps)
        recording = solver.recording_tensor()
        ctx.room, ctx.solver, ctx.checkpoints, ctx.report = room, solver, checkpoints, report
        ctx.want = (want_walls, want_signal, want_sponge, want_air)
        ctx.layered = layered
        ctx.inputs = (admittance, signal, sponge_strength, air_strength)
        return recording

    @staticmethod
    def backward(ctx, grad_recording):
        room, solver = ctx.room, ctx.solver
        want_walls, want_signal, want_sponge, want_air = ctx.want
        grads = _run_adjoint(room, solver, grad_recording, ctx.checkpoints, ctx.want, ctx.report,
                             layered=ctx.layered)
        room.report = ctx.report
        admittance, signal, sponge_strength, air_strength = ctx.inputs

        def like(value, reference):
            if value is None or reference is None:
                return None
            return torch.as_tensor(value).to(dtype=reference.dtype, device=reference.device)

        out_signal = None
        if want_signal and signal is not None:
            g = grads["signal"][:, : signal.shape[1]]
            out_signal = like(g, signal)
        ctx.solver = ctx.checkpoints = None
        return (
            None,
            like(grads.get("admittance"), admittance) if want_walls else None,
            out_signal,
            like(grads.get("sponge_strength"), sponge_strength) if want_sponge else None,
            like(grads.get("air_strength"), air_strength) if want_air else None,
        )


def _run_adjoint(room, solver: AcousticFDTD, grad_recording, checkpoints, want, report,
                 layered: bool = False):
    """Transpose every step of the forward run, newest first; return the parameter gradients.
    ``layered``: the forward pass wrote a layer tape of every step (the sponge gradient)."""
    want_walls, want_signal, want_sponge, want_air = want
    lib = solver._lib
    g, k = solver.grid, solver.coefficients
    opts = {"device": solver.device, "dtype": solver.dtype}
    n_steps = room.n_steps
    lam = [torch.zeros(g.shape, **opts)] + [torch.zeros(g.velocity_shape(a), **opts) for a in range(3)]
    lpsi = torch.zeros((k.n_relax, *g.shape), **opts)
    ldiv = torch.zeros(g.shape, **opts)
    G = grad_recording.to(**opts).contiguous()
    recording = solver.recording_tensor()
    # Air gradients from divergences (a tape of every step, or a ring per recomputed segment)
    # unless the sponge gradient's forward states are recomputed anyway.
    divergences = want_air and (layered or not want_sponge)
    fwd = solver._descriptor(recording)
    two_step = (room.fused and room.time_block == 2 and (layered or not want_sponge)
                and solver.adjoint_two_step_applies())
    if room.autotune:
        from .autotune import adjoint_tiling

        if two_step:
            fwd.tile_j, fwd.chunk_x = adjoint_tiling(solver, two_step=True, air=divergences,
                                                     sponge=want_sponge and layered)
        else:
            fwd.tile_j, fwd.chunk_x = adjoint_tiling(solver, fwdgrad=want_sponge or want_air)
    n_partials = lib.acfdtd_hip_adjoint_partials_count(ctypes.byref(fwd))
    partials = torch.zeros(n_partials, dtype=torch.float64, device=solver.device)
    sources = solver._source_arrays()
    grad_signal = None
    if want_signal and sources is not None:
        grad_signal = torch.zeros(sources[2].shape, **opts)

    adj = _native.AdjointDesc()
    adj.abi_version = _native.ABI_VERSION
    adj.struct_size = ctypes.sizeof(_native.AdjointDesc)
    adj.grads = (
        (_native.GRAD_WALLS if want_walls else 0)
        | (_native.GRAD_SPONGE if want_sponge else 0)
        | (_native.GRAD_AIR if want_air else 0)
    )
    adj.lp, adj.lvx, adj.lvy, adj.lvz = (t.data_ptr() for t in lam)
    adj.lpsi = lpsi.data_ptr() if k.n_relax else None
    adj.ldiv = ldiv.data_ptr()
    adj.grad_recording = G.data_ptr() if G.numel() else None
    adj.grad_recording_stride = G.shape[1] if G.dim() == 2 else 0
    adj.grad_recording_step0 = 0
    adj.grad_signal = grad_signal.data_ptr() if grad_signal is not None else None
    adj.dt = solver.dt
    adj.partials = partials.data_ptr()
    weights = None
    if want_sponge:
        cells, faces = layer_weights(g, room.absorbing_layer)
        weights = [torch.as_tensor(w, **opts) for w in (*cells, *faces)]
        for a in range(3):
            adj.weight_cell[a] = weights[a].data_ptr()
            adj.weight_face[a] = weights[3 + a].data_ptr()

    if room.fused:
        # Second set for the fused adjoint steps (acfdtd_hip_adjoint_run ping-pongs ldiv and
        # the adjoint velocities between the two).
        swaps = [torch.zeros(g.shape, **opts)] + [torch.zeros(g.velocity_shape(a), **opts)
                                                  for a in range(3)]
        adj.ldiv_swap, adj.lvx_swap, adj.lvy_swap, adj.lvz_swap = (t.data_ptr() for t in swaps)
        if two_step:
            # Two adjoint steps per sweep also ping-pong lp and lpsi.
            swaps += [torch.zeros(g.shape, **opts), torch.zeros((k.n_relax, *g.shape), **opts)]
            adj.lp_swap = swaps[4].data_ptr()
            adj.lpsi_swap = swaps[5].data_ptr() if k.n_relax else None
            adj.time_block = 2
    div_ring = None
    if checkpoints and divergences:
        # One ring of K divergence rows, reused by every segment: row q = step start + q.
        div_ring = torch.empty((room.checkpoint_every, g.n_cells), **opts)
        report.div_tape_bytes = div_ring.numel() * div_ring.element_size()
        fwd.div_tape, fwd.div_tape_step0, fwd.div_tape_rows = (div_ring.data_ptr(), 0,
                                                               room.checkpoint_every)
    report.adjoint_two_step = bool(lib.acfdtd_hip_adjoint_two_step(ctypes.byref(fwd), ctypes.byref(adj)))

    stream = torch.cuda.current_stream(solver.device).cuda_stream
    handle = ctypes.c_void_p(stream)

    def use_ring(ring, step0, count):
        """Forward states for the adjoint: slot q of each ring field holds step step0 + q."""
        if ring is None:
            adj.fwd_p = adj.fwd_vx = adj.fwd_vy = adj.fwd_vz = None
            adj.fwd_step0 = adj.fwd_count = 0
        else:
            adj.fwd_p, adj.fwd_vx, adj.fwd_vy, adj.fwd_vz = (t.data_ptr() for t in ring)
            adj.fwd_step0, adj.fwd_count = step0, count

    def last_local(ring, slot):
        """The local part of the last forward step, before any adjoint step."""
        if ring is None:
            adj.fwd_p = adj.fwd_vx = adj.fwd_vy = adj.fwd_vz = None
        else:
            adj.fwd_p, adj.fwd_vx, adj.fwd_vy, adj.fwd_vz = (t[slot].data_ptr() for t in ring)
        _native.check(lib.acfdtd_hip_adjoint_cell_sweep(ctypes.byref(fwd), ctypes.byref(adj),
                                                        n_steps - 1, 0, 1, handle),
                      "acfdtd_hip_adjoint_cell_sweep")

    def adjoint_steps(step_hi, step_lo):
        _native.check(lib.acfdtd_hip_adjoint_run(ctypes.byref(fwd), ctypes.byref(adj), step_hi,
                                                 step_lo, handle), "acfdtd_hip_adjoint_run")

    def unpacked(steps):
        """Forward states after the given steps, unpacked from the layer tape into slots."""
        states = [torch.zeros((len(steps), *g.shape), **opts)] + [
            torch.zeros((len(steps), *g.velocity_shape(a)), **opts) for a in range(3)]
        for q, n in enumerate(steps):
            if n is None:
                continue
            _native.check(lib.acfdtd_hip_layer_tape_unpack(
                ctypes.byref(fwd), n, *(ctypes.c_void_p(t[q].data_ptr()) for t in states), handle),
                "acfdtd_hip_layer_tape_unpack")
        return states

    with torch.cuda.device(solver.device):
        if layered:
            # Adjoint steps n_steps - 1 .. r in groups of four through the two-step sweep, which
            # reads the layer tape; the last r (1 to 4) one per sweep through states unpacked from
            # it (slot q: the state after step q - 1; slot 0, the initial state, is never read).
            quads = (n_steps - 1) // 4
            r = n_steps - 4 * quads
            last_local(unpacked([n_steps - 1]), 0)
            use_ring(None, 0, 0)
            if quads:
                adjoint_steps(n_steps - 1, r)
            ring = unpacked([None] + list(range(r)))
            use_ring(ring, 0, r + 1)
            adjoint_steps(r - 1, 0)
            del ring
        elif div_ring is not None:
            K = room.checkpoint_every
            ring = div_ring
            use_ring(None, 0, 0)
            last = len(checkpoints) - 1
            for seg in range(last, -1, -1):
                start, stop = seg * K, min(seg * K + K, n_steps)
                solver.load_state(checkpoints[seg], start)
                solver.attach_div_tape(ring, start)
                solver.run(stop - start, record=False)
                report.recomputed_steps += stop - start
                fwd.div_tape, fwd.div_tape_step0, fwd.div_tape_rows = ring.data_ptr(), start, K
                # Adjoint step m reads the divergences of step m - 1: this segment's rows serve
                # steps start + 1 .. stop (and the last step's local part).
                if seg == last:
                    last_local(None, 0)
                hi = n_steps - 1 if seg == last else stop
                lo = start + 1 if start > 0 else 0
                if hi >= lo:
                    adjoint_steps(hi, lo)
            solver.attach_div_tape(None)
            del ring, div_ring
        elif checkpoints:
            K = room.checkpoint_every
            # One ring of K + 1 states, reused by every segment: slot q = step start + q.
            ring = [torch.empty((K + 1, *g.shape), **opts)] + [
                torch.empty((K + 1, *g.velocity_shape(a)), **opts) for a in range(3)
            ]
            report.checkpoint_bytes += sum(t.numel() * t.element_size() for t in ring)
            state_desc = solver._descriptor(None, tape=False)
            for seg in range(len(checkpoints) - 1, -1, -1):
                start, stop = seg * K, min(seg * K + K, n_steps)
                solver.load_state(checkpoints[seg], start)
                for field, saved in zip(ring, checkpoints[seg][:4]):
                    field[0].copy_(saved)
                _native.check(lib.acfdtd_hip_run_states(
                    ctypes.byref(state_desc), start, stop - start,
                    *(ctypes.c_void_p(t.data_ptr()) for t in ring), handle),
                    "acfdtd_hip_run_states")
                report.recomputed_steps += stop - start
                if stop == n_steps:
                    last_local(ring, n_steps - start)
                use_ring(ring, start, stop - start + 1)
                adjoint_steps(stop - 1, start)
            del ring
        else:
            use_ring(None, 0, 0)
            last_local(None, 0)
            adjoint_steps(n_steps - 1, 0)
    report.adjoint_steps = n_steps

    sums = partials.view(-1, _native.GRAD_SLOTS).sum(dim=0).cpu().numpy()
    out = {"adjoint_state": lam}
    c, dx, dt, rho = solver.medium.sound_speed, g.dx, solver.dt, solver.medium.density
    if want_walls:
        out["admittance"] = sums[0:6] * (c * dt / (2.0 * dx))
    if want_sponge:
        out["sponge_strength"] = np.float64(sums[6])
    if want_air:
        delta = np.array([q.strength for q in k.relaxation])
        total = 1.0 + delta.sum()
        # gain_r = rho c^2 Delta_r / ((1 + sum Delta) dx);  relax_scale = dt sum_r gain_r
        jac = (rho * c**2 / dx) * (np.eye(delta.size) / total - delta[:, None] / total**2)
        d_gain = sums[8 : 8 + delta.size] + sums[7] * dt
        out["air_strength"] = jac.T @ d_gain
    if grad_signal is not None:
        out["signal"] = grad_signal
    return out
