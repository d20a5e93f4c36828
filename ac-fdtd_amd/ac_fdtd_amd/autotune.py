"""Tile shapes of the fused kernels, picked by timing them on the GPU.

A fused launch cuts each x-plane into workgroups of ``64 x tile_j`` lanes and gives every
workgroup ``chunk_x`` consecutive planes to march. Which shape is fastest depends on the grid
(small grids need many short marches to fill 256 CUs; long marches re-read fewer halo planes),
on the precision and on the terms the step carries, and it differs between the forward step
and the adjoint step, so it is measured: every candidate shape runs a few steps on scratch
fields that share the solver's descriptor (coefficients, walls, sponge, relaxation) and is
timed with HIP events on the launch stream; the fastest wins. Where the solver may advance
two steps per sweep (``time_block=2``), the best two-step shape and the best one-step shape are
timed against each other too, and the faster plan is kept.

Results are kept per configuration for the life of the process and, when the environment
variable ``ACFDTD_TUNE_CACHE`` names a JSON file, across processes. A tiling never changes the
fields a step computes (each cell's arithmetic is the same whatever the tile), only the order
in which the adjoint's fp64 gradient partials are summed.
"""

from __future__ import annotations

import ctypes
import json
import os
from pathlib import Path

import torch

from . import _native

__all__ = ["CANDIDATES", "CANDIDATES_TWO_STEP", "adjoint_tiling", "forward_plan", "forward_tiling"]

#: (tile_j, chunk_x) shapes tried for one step per sweep.
CANDIDATES = tuple((tile_j, chunk) for tile_j in (1, 2, 4, 8) for chunk in (1, 2, 3, 4, 6, 8))
#: (tile_j, chunk_x) shapes tried for two steps per sweep (2 tile_j rows per block): taller
#: blocks spread the halo rows thinner, longer marches the halo planes; short marches keep
#: small grids' workgroups many and their per-plane latency chain short.
CANDIDATES_TWO_STEP = tuple((tile_j, chunk) for tile_j in (2, 3, 4, 6, 8)
                            for chunk in (1, 2, 3, 4, 6, 8, 12, 16, 24, 32))

_cache: dict[str, tuple[int, ...]] = {}
_loaded = False


def _cache_file() -> Path | None:
    name = os.environ.get("ACFDTD_TUNE_CACHE")
    return Path(name) if name else None


def _lookup(key: str):
    global _loaded
    if not _loaded:
        _loaded = True
        path = _cache_file()
        if path is not None and path.exists():
            try:
                stored = json.loads(path.read_text())
                _cache.update({k: tuple(v) for k, v in stored.items()})
            except (OSError, ValueError):
                pass
    return _cache.get(key)


def _store(key: str, value: tuple[int, ...]) -> None:
    _cache[key] = value
    path = _cache_file()
    if path is not None:
        try:
            stored = json.loads(path.read_text()) if path.exists() else {}
        except (OSError, ValueError):
            stored = {}
        stored[key] = list(value)
        path.write_text(json.dumps(stored, indent=1, sort_keys=True))


def _key(kind: str, solver, tape: bool = False) -> str:
    g, k = solver.grid, solver.coefficients
    props = torch.cuda.get_device_properties(solver.device)
    # The architecture and CU count identify the device type; the marketing name is not
    # reliable (it comes from a table some installations lack).
    name = f"{getattr(props, 'gcnArchName', 'gpu').split(':')[0]}x{props.multi_processor_count}"
    return (f"{kind}:{name}:{g.nx}x{g.ny}x{g.nz}:f{solver.np_dtype.itemsize * 8}:r{k.n_relax}"
            f":w{int(solver._wall_mask != 0)}:l{int(k.has_layer)}:c{int(solver.compensated)}"
            + (":t1" if tape else ""))


def _steps_per_sample(solver) -> int:
    """Enough steps for about 5 ms per timing, at least 8 (four two-step launches: fewer made
    the timings of large grids too noisy to rank close candidates), at most 60, a multiple of 4
    (so a two-step timing runs two-step sweeps only)."""
    bytes_per_cell = solver.np_dtype.itemsize * (8 + 2 * solver.coefficients.n_relax)
    seconds = solver.grid.n_cells * bytes_per_cell / 4e12
    steps = int(min(60, max(8, 5e-3 / max(seconds, 1e-9))))
    return steps + (-steps) % 4


class _Scratch:
    """Scratch forward fields (both sets) with a copy of the solver's descriptor on them; no
    sources or receivers, and a scratch boundary tape of ``tape_rows`` rows when that is
    nonzero and the solver has walls."""

    def __init__(self, solver, tape_rows: int = 0) -> None:
        g, k = solver.grid, solver.coefficients
        opts = {"device": solver.device, "dtype": solver.dtype}
        gen = torch.Generator(device=solver.device).manual_seed(0)

        def field(shape):
            return torch.randn(shape, generator=gen, **opts) * 1e-3

        self.sets = [[field(g.shape)] + [field(g.velocity_shape(a)) for a in range(3)]
                     for _ in range(2)]
        self.psi = torch.zeros((k.n_relax, *g.shape), **opts)
        self.psi_swap = torch.zeros((k.n_relax, *g.shape), **opts) if k.n_relax else None
        d = solver._descriptor(None, tape=False)
        d.n_sources = d.n_receivers = 0
        d.source_cell = d.source_length = d.source_signal = None
        d.receiver_cell = d.recording = None
        d.p, d.vx, d.vy, d.vz = (t.data_ptr() for t in self.sets[0])
        d.p_swap, d.vx_swap, d.vy_swap, d.vz_swap = (t.data_ptr() for t in self.sets[1])
        d.psi = self.psi.data_ptr() if k.n_relax else None
        d.psi_swap = self.psi_swap.data_ptr() if k.n_relax else None
        self.tape = None
        if tape_rows > 0 and solver._wall_table is not None:
            self.tape = torch.zeros((tape_rows, solver.wall_tape_stride()), **opts)
            d.wall_tape = self.tape.data_ptr()
            d.wall_tape_stride, d.wall_tape_step0 = self.tape.shape[1], 0
            d.wall_tape_rows = tape_rows
        self.desc = d


def _time(solver, launch, shape, steps: int, repeats: int = 2) -> float:
    """Mean ms per step of ``launch(steps)`` with ``shape`` = (tile_j, chunk_x); best of
    ``repeats`` timings after one warm-up."""
    stream = torch.cuda.current_stream(solver.device)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    launch(min(steps, 4), shape)
    best = float("inf")
    for _ in range(repeats):
        start.record(stream)
        launch(steps, shape)
        stop.record(stream)
        stop.synchronize()
        best = min(best, start.elapsed_time(stop) / steps)
    return best


def _warm_up(solver, launch, shape, steps: int, seconds: float = 0.05) -> None:
    """Run ``launch`` for about ``seconds`` before timing anything, so that the first
    candidates are not timed while the device is still raising its clocks."""
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(torch.cuda.current_stream(solver.device))
    launch(steps, shape)
    stop.record(torch.cuda.current_stream(solver.device))
    stop.synchronize()
    per_call = max(start.elapsed_time(stop) * 1e-3, 1e-6)
    for _ in range(min(200, int(seconds / per_call))):
        launch(steps, shape)
    torch.cuda.synchronize(solver.device)


def _pick(solver, kind, launch, candidates, tape: bool = False):
    key = _key(kind, solver, tape)
    found = _lookup(key)
    if found is not None:
        return found
    steps = _steps_per_sample(solver)
    with torch.cuda.device(solver.device):
        _warm_up(solver, launch, candidates[0], steps)
        timings = {shape: _time(solver, launch, shape, steps) for shape in candidates}
    best = min(timings, key=timings.get)
    _store(key, best)
    return best


def forward_tiling(solver, time_block: int = 1, candidates=None, tape: bool = False,
                   staged: bool = False) -> tuple[int, int]:
    """Fastest (tile_j, chunk_x) of the fused forward sweep advancing ``time_block`` steps, for
    ``solver``'s configuration (writing a boundary tape when ``tape``; the staged two-step
    sweep when ``staged``)."""
    if not solver.fused:
        return solver.tile_j, solver.chunk_x
    if candidates is None:
        candidates = CANDIDATES_TWO_STEP if time_block == 2 else CANDIDATES
    scratch = _Scratch(solver, _steps_per_sample(solver) if tape else 0)
    scratch.desc.time_block = time_block
    scratch.desc.flags = (scratch.desc.flags & ~_native.STAGED) | (_native.STAGED if staged else 0)
    lib = solver._lib
    handle = ctypes.c_void_p(torch.cuda.current_stream(solver.device).cuda_stream)

    def launch(steps, shape):
        scratch.desc.tile_j, scratch.desc.chunk_x = shape
        _native.check(lib.acfdtd_hip_run(ctypes.byref(scratch.desc), 0, steps, 0, handle),
                      "acfdtd_hip_run")

    kind = ("forward2s" if staged else "forward2") if time_block == 2 else "forward"
    return _pick(solver, kind, launch, candidates, tape)


def _staged_possible(solver) -> bool:
    """The staged sweep needs 16-byte lanes: 4 fp32 or 2 fp64 cells, so nz a multiple of that."""
    return solver.grid.nz % (16 // solver.np_dtype.itemsize) == 0


def forward_plan(solver, tape: bool = False) -> tuple[int, int, int, bool]:
    """(time_block, tile_j, chunk_x, staged) for ``solver``: the fastest one-step tiling and,
    when the solver may advance two steps per sweep, the fastest two-step tiling, plain and
    staged (each plane's inputs requested a plane ahead into LDS), whichever runs more steps
    per second. With ``tape`` every candidate writes a boundary tape, as a forward pass of the
    wall gradient does (the tape's stores change which plan is fastest)."""
    single = forward_tiling(solver, 1, tape=tape)
    if not solver.two_step_applies():
        return 1, single[0], single[1], False
    double = forward_tiling(solver, 2, tape=tape)
    # The staged sweep's best shape differs from the plain one's (at 512^3 fp64: 8 rows and
    # 32-plane marches staged, 71 Gcell/s, against 65 staged at the plain sweep's best shape).
    staged = forward_tiling(solver, 2, tape=tape, staged=True) if _staged_possible(solver) else None
    key = _key("plan", solver, tape)
    found = _lookup(key)
    if found is not None:
        return (*found, False)[:4] if len(found) == 3 else tuple(found)
    handle = ctypes.c_void_p(torch.cuda.current_stream(solver.device).cuda_stream)
    launches = {}
    steps = 4 * _steps_per_sample(solver)
    plans = [((1, False), single), ((2, False), double)]
    if staged is not None:
        plans.append(((2, True), staged))
    for plan, shape in plans:
        scratch = _Scratch(solver, steps if tape else 0)
        scratch.desc.time_block = plan[0]
        scratch.desc.tile_j, scratch.desc.chunk_x = shape
        scratch.desc.flags = (scratch.desc.flags & ~_native.STAGED) | (_native.STAGED if plan[1] else 0)

        def launch(steps, _shape, desc=scratch.desc, keep=scratch):
            _native.check(solver._lib.acfdtd_hip_run(ctypes.byref(desc), 0, steps, 0, handle),
                          "acfdtd_hip_run")

        launches[plan] = (launch, shape)
    per_step = {plan: float("inf") for plan in launches}
    with torch.cuda.device(solver.device):
        _warm_up(solver, *launches[(1, False)], steps)
        # Interleaved rounds, the best of each: a drift of the clocks favours none.
        for _ in range(3):
            for plan, (launch, shape) in launches.items():
                per_step[plan] = min(per_step[plan], _time(solver, launch, shape, steps, 1))
    best = min(per_step, key=per_step.get)
    shape = launches[best][1]
    result = (best[0], shape[0], shape[1], best[1])
    _store(key, result)
    return result


def adjoint_tiling(solver, fwdgrad: bool = False, candidates=None,
                   two_step: bool = False, air: bool = False,
                   sponge: bool = False) -> tuple[int, int]:
    """Fastest (tile_j, chunk_x) of the fused adjoint step: with the wall gradient when the
    solver has a wall table, and with the sponge and air gradients too when ``fwdgrad``; of the
    two-step adjoint sweep when ``two_step`` (wall gradient, with ``air`` the air gradients
    from a divergence tape, with ``sponge`` the sponge gradient from a layer tape)."""
    if not solver.fused:
        return solver.tile_j, solver.chunk_x
    if candidates is None:
        candidates = CANDIDATES_TWO_STEP if two_step else CANDIDATES
    g, k = solver.grid, solver.coefficients
    lib = solver._lib
    opts = {"device": solver.device, "dtype": solver.dtype}
    scratch = _Scratch(solver)
    steps = _steps_per_sample(solver)
    fwd = scratch.desc
    walls = solver._wall_table is not None
    tape = None
    if walls:
        tape = torch.zeros((steps + 1, solver.wall_tape_stride()), **opts)
        fwd.wall_tape = tape.data_ptr()
        fwd.wall_tape_stride, fwd.wall_tape_step0, fwd.wall_tape_rows = tape.shape[1], 0, tape.shape[0]
    # Nonzero adjoint fields, so that boundary waves reduce their wall partials as in a run.
    gen = torch.Generator(device=solver.device).manual_seed(1)
    lam = [[torch.randn(g.shape, generator=gen, **opts)] +
           [torch.randn(g.velocity_shape(a), generator=gen, **opts) for a in range(3)]
           for _ in range(2)]
    lpsi = torch.zeros((k.n_relax, *g.shape), **opts)
    adj = _native.AdjointDesc()
    adj.abi_version = _native.ABI_VERSION
    adj.struct_size = ctypes.sizeof(_native.AdjointDesc)
    adj.grads = _native.GRAD_WALLS if walls else 0
    adj.lp, adj.lvx, adj.lvy, adj.lvz = (t.data_ptr() for t in lam[0])
    ldiv = [torch.randn(g.shape, generator=gen, **opts) for _ in range(2)]
    adj.ldiv, adj.ldiv_swap = ldiv[0].data_ptr(), ldiv[1].data_ptr()
    adj.lvx_swap, adj.lvy_swap, adj.lvz_swap = (t.data_ptr() for t in lam[1][1:])
    adj.lpsi = lpsi.data_ptr() if k.n_relax else None
    adj.dt = solver.dt
    second = None
    if two_step:
        second = [torch.randn(g.shape, generator=gen, **opts), torch.zeros((k.n_relax, *g.shape), **opts)]
        adj.lp_swap = second[0].data_ptr()
        adj.lpsi_swap = second[1].data_ptr() if k.n_relax else None
        adj.time_block = 2
    ring = weights = divergences = None
    if two_step and air and k.n_relax:
        # Adjoint steps n .. 1 read the divergences of steps n - 1 .. 0.
        divergences = torch.zeros((steps + 2, g.n_cells), **opts)
        fwd.div_tape, fwd.div_tape_step0, fwd.div_tape_rows = (divergences.data_ptr(), 0,
                                                               divergences.shape[0])
        adj.grads |= _native.GRAD_AIR
    layer_tape = None
    if two_step and sponge and k.has_layer:
        # Adjoint steps n .. 1 read layer tape rows n .. 0.
        from .boundaries import layer_weights

        layer_tape = torch.zeros((steps + 2, solver.layer_tape_stride()), **opts)
        fwd.layer_tape, fwd.layer_tape_step0 = layer_tape.data_ptr(), 0
        fwd.layer_tape_rows, fwd.layer_tape_stride = layer_tape.shape
        lo, hi = solver.layer_extent()
        for a in range(3):
            fwd.layer_lo[a], fwd.layer_hi[a] = lo[a], hi[a]
        adj.grads |= _native.GRAD_SPONGE
        cells, faces = layer_weights(g, solver.absorbing_layer)
        weights = [torch.as_tensor(w, **opts) for w in (*cells, *faces)]
        for a in range(3):
            adj.weight_cell[a] = weights[a].data_ptr()
            adj.weight_face[a] = weights[3 + a].data_ptr()
    if fwdgrad and not two_step and (k.has_layer or k.n_relax):
        # Adjoint steps n .. 1 read forward states 1 .. n + 1.
        ring = [torch.zeros((steps + 2, *g.shape), **opts)] + [
            torch.zeros((steps + 2, *g.velocity_shape(a)), **opts) for a in range(3)]
        adj.fwd_p, adj.fwd_vx, adj.fwd_vy, adj.fwd_vz = (t.data_ptr() for t in ring)
        adj.fwd_step0, adj.fwd_count = 0, steps + 2
        if k.n_relax:
            adj.grads |= _native.GRAD_AIR
        if k.has_layer:
            adj.grads |= _native.GRAD_SPONGE
            weights = [torch.zeros(g.shape[a] + extra, **opts) for extra in (0, 1) for a in range(3)]
            for a in range(3):
                adj.weight_cell[a] = weights[a].data_ptr()
                adj.weight_face[a] = weights[3 + a].data_ptr()
    size = 0
    for shape in candidates:
        fwd.tile_j, fwd.chunk_x = shape
        size = max(size, lib.acfdtd_hip_adjoint_partials_count(ctypes.byref(fwd)))
    partials = torch.zeros(max(size, 1), dtype=torch.float64, device=solver.device)
    adj.partials = partials.data_ptr()
    handle = ctypes.c_void_p(torch.cuda.current_stream(solver.device).cuda_stream)

    def launch(n, shape):
        # Adjoint steps n .. 1: local parts of steps n - 1 .. 0 (tape rows 0 .. n - 1).
        fwd.tile_j, fwd.chunk_x = shape
        _native.check(lib.acfdtd_hip_adjoint_run(ctypes.byref(fwd), ctypes.byref(adj), n, 1,
                                                 handle), "acfdtd_hip_adjoint_run")

    kind = ("adjoint2" + ("-sponge" if layer_tape is not None else "")
            + ("-air" if divergences is not None else "")) if two_step else (
        "adjoint-fwdgrad" if ring is not None else "adjoint")
    return _pick(solver, kind, launch, candidates)
