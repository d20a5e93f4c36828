"""Throughput of the fused HIP time loop, in Gcell-updates/s, on one MI355X.

Workload (the throughput configuration): a cubic grid of ``--size``^3 cells at dx = 1 cm,
walls of normal-incidence absorption 0.15 on all six faces, a 16-cell sponge on all faces, air
absorption at 20 C / 50 % RH (three relaxation states), a Gaussian-derivative pulse of seeded
width 4-8 cells at a seeded interior cell, 8 receivers at seeded interior cells, ``--steps``
steps. All state is resident in HBM before timing; the timed region is one call of the library
advancing ``--steps`` steps, bracketed by HIP events and synchronisations.

Printed: one JSON line. ``value`` is fp32 Gcell-updates/s at ``--size``^3 (two steps per
sweep where the autotuner finds that faster, tile shapes tuned). ``roofline`` is for the
dominant kernel: ``achieved`` counts the bytes that kernel must move per launch (each field
element it reads or writes counted once: 28 B per cell-update for the two-step sweep in fp32
with three relaxation states, two cell-updates per launch) over the launch's mean duration from
HIP events on the launch stream; ``traffic`` is the HBM bytes per launch rocprofv3 measured for
the same kernel and workload (committed profile), when there is one. Under ``extra``: the same
launch time on the byte count of the reference's two-sweep step (72 B per cell-update, a
"reference-step equivalent" rate that exceeds the HBM peak when a kernel advances two steps
per pass), the run at one step per sweep, the fp64 plain kernel and the compensated fp32
kernel, the error of fp32 against fp64 on runs of ``--accuracy-steps`` steps whatever
``--steps`` is (receiver relative L2, T20 per octave band), rigid-box energy drift in fp64 and
fp32 over as many steps, a 384^3 fp32 run, 64 uneven rooms advanced as one batch (each checked
against its own fp64 single-room run), the cost of a wall-admittance gradient against a forward
pass at 128^3, a 512^3 room with its T20 per octave band against fp64 and its wall gradient,
and the stream ceiling. ``cpu_baseline`` is the reference's own compiled C loop
(``oracle/_ref``, all host cores this process may use) on a bounded sample of the same
workload.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _path in (ROOT / "ac-fdtd_amd", ROOT):
    if str(_path) not in sys.path:
        sys.path.insert(0, str(_path))

import numpy as np  # noqa: E402

METRIC = "Gcell-updates/s at 256³ (fp64 plain vs fp32/batched, %HBM); grad/fwd time"
HBM_PEAK_GBS = 8000.0
DX = 0.01
SPONGE = 16
MARGIN = 24  # source and receivers at least this many cells from every face


def log(message: str) -> None:
    print(f"[bench] {message}", file=sys.stderr, flush=True)


def workload(size: int, steps: int, seed: int = 2024):
    """Seeded source/receiver cells and pulse of the throughput workload."""
    from ac_fdtd_amd import AIR, AbsorbingLayer, AirAbsorption, Grid, WallAdmittances

    grid = Grid(dx=DX, nx=size, ny=size, nz=size)
    rng = np.random.default_rng(seed)
    width_cells = rng.uniform(4.0, 8.0)

    def point():
        cell = rng.integers(MARGIN, size - MARGIN, size=3)
        return tuple((cell + 0.5) * DX)

    source = point()
    receivers = [point() for _ in range(8)]
    config = {
        "walls": WallAdmittances.from_absorption(0.15),
        "absorbing_layer": AbsorbingLayer(thickness=SPONGE),
        "air_absorption": AirAbsorption(temperature=20.0, relative_humidity=50.0),
    }
    width = width_cells * DX / AIR.sound_speed
    return grid, config, source, receivers, width


def make_solver(cls, size, steps, dtype, **kwargs):
    from ac_fdtd_amd.sources import gaussian_derivative_pulse

    grid, config, source, receivers, width = workload(size, steps)
    solver = cls(grid, dtype=dtype, **config, **kwargs)
    solver.add_volume_source(source, gaussian_derivative_pulse(steps + 16, solver.dt, width))
    for r in receivers:
        solver.add_pressure_receiver(r)
    return solver


def update_bytes(n_relax: int, itemsize: int) -> int:
    """Algorithmic bytes per cell-update, the roofline basis: each field element of the
    reference's two-sweep step counted once (velocity sweep: read p, read and write vx, vy, vz;
    pressure sweep: read vx, vy, vz, read and write p and each relaxation state), 12 + 2 n_relax
    elements."""
    return (12 + 2 * n_relax) * itemsize


def launch_bytes(kind: str, n_relax: int, itemsize: int) -> int:
    """Algorithmic bytes per cell of one launch of ``kind`` on that basis: the two-step sweep
    does two cell-updates, the fused step one, the velocity and pressure sweeps their parts."""
    return {
        "fused2": 2 * update_bytes(n_relax, itemsize),
        "fused": update_bytes(n_relax, itemsize),
        "velocity": 7 * itemsize,
        "pressure": (5 + 2 * n_relax) * itemsize,
    }[kind]


def launch_minimum_bytes(kind: str, n_relax: int, itemsize: int) -> float:
    """Bytes per cell a launch of ``kind`` must move: :func:`minimum_bytes` times the
    cell-updates it performs (two for the two-step sweep)."""
    return minimum_bytes(kind, n_relax, itemsize) * (2 if kind == "fused2" else 1)


def minimum_bytes(kind: str, n_relax: int, itemsize: int) -> float:
    """Bytes per cell-update a launch of ``kind`` must move itself: the fused step reads p, v
    from one set, writes them to the other and reads and writes the relaxation states
    (8 + 2 n_relax elements); the two-step sweep does so once for two updates."""
    fused = (8 + 2 * n_relax) * itemsize
    return {"fused2": fused / 2, "fused": fused, "velocity": 7 * itemsize,
            "pressure": (5 + 2 * n_relax) * itemsize}[kind]

PROFILES = ROOT / "profiles"


def measured_traffic(size: int, kernel: str):
    """HBM bytes per launch of ``kernel`` measured by rocprofv3 (FETCH_SIZE and WRITE_SIZE,
    calibrated on a copy of known size) for this workload, from the committed profiles; None
    when no profile covers this kernel and grid."""
    for path in sorted(PROFILES.glob("*.json")):
        record = json.loads(path.read_text())
        if record.get("kind") == kernel and record.get("grid") == [size] * 3:
            return record["traffic_bytes_per_launch"]
    return None

def gpu_throughput(size, steps, warmup, dtype, compensated=False, autotune=True, time_block=2):
    import torch

    from ac_fdtd_amd.scheme import AcousticFDTD

    solver = make_solver(AcousticFDTD, size, steps + warmup + 40, dtype, compensated=compensated,
                         autotune=autotune, time_block=time_block)
    solver.run(warmup)
    # The recording buffer and the descriptor are prepared before the timed region, which
    # holds the steps' launches alone.
    launch = solver.prepare_run(steps)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    wall = time.perf_counter()
    start.record()
    launch()
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall
    ms = start.elapsed_time(stop)
    cells = solver.grid.n_cells
    kernels = solver.timed_steps(20)
    recorded = solver.recorded_pressure
    if not np.all(np.isfinite(recorded)):
        raise RuntimeError("non-finite receiver samples")
    return {
        "gcells": cells * steps / (ms * 1e-3) / 1e9,
        "ms_per_step": ms / steps,
        "wall_s": wall,
        "kernels": kernels,
        "n_relax": solver.coefficients.n_relax,
        "cells": cells,
        "recorded": recorded.astype(np.float64),
        "sample_rate": solver.sample_rate,
        "tiling": [solver.tile_j, solver.chunk_x],
        "time_block": solver.time_block,
        "staged": solver.staged,
    }


def relative_l2(test: np.ndarray, reference: np.ndarray) -> float:
    return float(np.linalg.norm(test - reference) / np.linalg.norm(reference))


OCTAVES = (500.0, 1000.0, 2000.0, 4000.0)


def error_against(test: dict, reference: dict) -> dict:
    """Receiver relative L2 of ``test`` against ``reference`` and the largest relative T20
    deviation per octave band over the receivers (bands whose decay is too short for a T20
    in either run are left out)."""
    from ac_fdtd_amd.metrics import octave_band_reverberation

    a, b, rate = test["recorded"], reference["recorded"], reference["sample_rate"]
    worst = {}
    for channel in range(b.shape[0]):
        t_test = octave_band_reverberation(a[channel], rate, OCTAVES)
        t_ref = octave_band_reverberation(b[channel], rate, OCTAVES)
        for band in OCTAVES:
            if np.isfinite(t_test[band]) and np.isfinite(t_ref[band]):
                deviation = abs(t_test[band] / t_ref[band] - 1.0)
                worst[band] = max(worst.get(band, 0.0), deviation)
    return {
        "receiver_rel_l2": relative_l2(a, b),
        "t20_max_rel_deviation_by_octave": {f"{int(k)}": float(v) for k, v in worst.items()},
    }


def rigid_energy_drift(size, steps, dtype=np.float64):
    """Rigid box from a seeded random pressure: max |E/E0 - 1| over the run (energy reduced in
    fp64 on the device for either precision)."""
    import torch

    from ac_fdtd_amd import Grid
    from ac_fdtd_amd.scheme import AcousticFDTD

    solver = AcousticFDTD(Grid(dx=DX, nx=size, ny=size, nz=size), dtype=dtype)
    solver.p.copy_(torch.as_tensor(np.random.default_rng(0).standard_normal(solver.grid.shape)))
    initial = solver.energy()
    drift = 0.0
    chunk = max(1, steps // 8)
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        solver.run(n)
        done += n
        drift = max(drift, abs(solver.energy() / initial - 1.0))
    return drift


def batched_rooms(n_rooms, steps, warmup, check=True, seed=3):
    """Rooms of sides U[3, 10] m at 1 kHz (the spacing ``Grid.from_max_frequency`` gives:
    c / (10 x 1 kHz) = 34.3 mm), six per-wall absorptions U[0.05, 0.5], air at T U[15, 25] C and
    RH U[30, 70] %, one source and 4 receivers at least 0.5 m from every wall, the repo's
    band-limited impulse at 1 kHz; advanced together in fp32, then each room alone in fp64 for
    the error."""
    import torch

    from ac_fdtd_amd import AIR, AirAbsorption, Grid, WallAdmittances, admittance_from_absorption
    from ac_fdtd_amd.batch import BatchedRooms
    from ac_fdtd_amd.scheme import AcousticFDTD
    from ac_fdtd_amd.sources import band_limited_impulse

    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n_rooms):
        sides = rng.uniform(3.0, 10.0, 3)
        walls = [admittance_from_absorption(a) for a in rng.uniform(0.05, 0.5, 6)]
        air = (rng.uniform(15.0, 25.0), rng.uniform(30.0, 70.0))
        points = [tuple(0.5 + rng.uniform(0.0, 1.0, 3) * (sides - 1.0)) for _ in range(5)]
        specs.append((tuple(sides), walls, air, points))

    def room(spec, dtype):
        sides, walls, air, points = spec
        grid = Grid.from_max_frequency(sides, 1000.0, AIR.sound_speed)
        solver = AcousticFDTD(grid, dtype=dtype, walls=WallAdmittances(*walls),
                              air_absorption=AirAbsorption(*air))
        solver.add_volume_source(points[0], band_limited_impulse(warmup + steps, solver.dt, 1000.0))
        for point in points[1:]:
            solver.add_pressure_receiver(point)
        return solver

    rooms = [room(spec, np.float32) for spec in specs]
    cells = sum(r.grid.n_cells for r in rooms)
    shapes = [r.grid.shape for r in rooms]
    batch = BatchedRooms(rooms)
    batch.run(warmup)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    batch.run(steps)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    groups = len(batch._groups)
    padded = sum(1 for g in batch._groups for r in g._padded if r is not None)
    batch.close()
    recorded = [r.recorded_pressure.astype(np.float64) for r in rooms]
    out = {
        "rooms": n_rooms,
        "cells": cells,
        "cells_per_axis_range": [int(min(min(s) for s in shapes)), int(max(max(s) for s in shapes))],
        "steps": steps,
        "launch_groups": groups,
        "rooms_in_padded_rows": padded,
        "gcells": round(cells * steps / (ms * 1e-3) / 1e9, 3),
        "ms_per_step": round(ms / steps, 4),
    }
    del rooms, batch
    torch.cuda.empty_cache()
    if check:
        errors = []
        for spec, mine in zip(specs, recorded):
            reference = room(spec, np.float64)
            reference.run(warmup + steps)
            errors.append(relative_l2(mine, reference.recorded_pressure))
            del reference
        torch.cuda.empty_cache()
        out["receiver_rel_l2_vs_fp64_max"] = float(max(errors))
        out["receiver_rel_l2_vs_fp64_median"] = float(np.median(errors))
    return out


def gradient_cost(size, steps, seed=4):
    """Wall-admittance gradient of an L2 receiver loss at ``size``^3, fp32: a hidden target set
    of six absorptions U[0.05, 0.5], the model at uniform 0.2, 4 receivers, the band-limited
    impulse at 1 kHz on a 34.3 mm grid. Returns seconds for a forward pass and for forward +
    backward, with walls only (boundary tape, no checkpoints) and with walls and the air
    strengths: from a divergence tape of every step (the default within 64 GiB), and from
    checkpoints every 100 steps with each segment's divergences recomputed (as beyond it); and,
    in a room lined with the 16-cell sponge on all faces, with walls, sponge and air strengths:
    from a layer tape and a divergence tape of every step (the default within 64 GiB), and from
    checkpoints every 100 steps with each segment's states recomputed into a ring (as beyond
    it)."""
    import torch

    from ac_fdtd_amd import (AIR, AbsorbingLayer, AirAbsorption, Grid, WallAdmittances,
                             admittance_from_absorption)
    from ac_fdtd_amd.autograd import RoomResponse
    from ac_fdtd_amd.sources import band_limited_impulse

    rng = np.random.default_rng(seed)
    dx = AIR.sound_speed / (10.0 * 1000.0)
    grid = Grid(dx=dx, nx=size, ny=size, nz=size)
    lengths = np.array(grid.lengths)
    hidden = [admittance_from_absorption(a) for a in rng.uniform(0.05, 0.5, 6)]
    points = [tuple((0.15 + 0.7 * rng.uniform(0.0, 1.0, 3)) * lengths) for _ in range(5)]
    start = admittance_from_absorption(0.2)

    def model(air, history_bytes=64 << 30, sponge=False):
        from ac_fdtd_amd.grid import max_stable_time_step

        dt = max_stable_time_step(dx, AIR.sound_speed)
        return RoomResponse(
            grid, steps, sources=[(points[0], band_limited_impulse(steps, dt, 1000.0))],
            receivers=points[1:], dtype=np.float32,
            walls=WallAdmittances.from_absorption(0.2),
            air_absorption=AirAbsorption(20.0, 50.0) if air else None, checkpoint_every=100,
            absorbing_layer=AbsorbingLayer(thickness=16) if sponge else None,
            history_bytes=history_bytes, autotune=True,
        )

    def seconds(call):
        torch.cuda.synchronize()
        began = time.perf_counter()
        call()
        torch.cuda.synchronize()
        return time.perf_counter() - began

    out = {"grid": [size] * 3, "steps": steps}
    # walls_and_air is RoomResponse's default for this size (its divergence tape fits the
    # 64 GiB history budget); walls_and_air_checkpointed forces the path beyond the budget.
    cases = (("walls_tape", False, 64 << 30, False), ("walls_and_air", True, 64 << 30, False),
             ("walls_and_air_checkpointed", True, 0, False),
             ("walls_sponge_and_air", True, 64 << 30, True),
             ("walls_sponge_and_air_checkpointed", True, 0, True))
    for key, air, history, sponge in cases:
        room = model(air, history, sponge)
        target = room(torch.tensor(hidden, dtype=torch.float64)).detach()
        xi = torch.full((6,), start, dtype=torch.float64, requires_grad=True)
        strengths = room.default_air_strength().requires_grad_(True) if air else None
        sigma = room.default_sponge_strength().requires_grad_(True) if sponge else None

        def forward():
            with torch.no_grad():
                room(xi.detach(), air_strength=None if strengths is None else strengths.detach(),
                     sponge_strength=None if sigma is None else sigma.detach())

        def forward_backward():
            loss = ((room(xi, air_strength=strengths, sponge_strength=sigma) - target) ** 2).sum()
            loss.backward()

        forward()
        forward_backward()
        t_fwd = min(seconds(forward) for _ in range(3))
        t_grad = min(seconds(forward_backward) for _ in range(3))
        out[key] = {
            "forward_s": round(t_fwd, 4),
            "forward_backward_s": round(t_grad, 4),
            "grad_over_forward": round(t_grad / t_fwd, 3),
            "recomputed_steps": room.report.recomputed_steps,
            "checkpoints": room.report.checkpoints,
            "divergence_bytes": room.report.div_tape_bytes,
            "layer_tape_bytes": room.report.layer_tape_bytes,
            "adjoint_two_steps_per_sweep": room.report.adjoint_two_step,
            "path": ("boundary tape" if not air else
                     "layer tape and divergence tape of every step"
                     if sponge and room.report.checkpoints == 0 else
                     "checkpoints every 100 steps, states recomputed per segment into a ring"
                     if sponge else "divergence tape of every step"
                     if room.report.checkpoints == 0 else
                     "checkpoints every 100 steps, divergences recomputed per segment"),
        }
        if not np.all(np.isfinite(xi.grad.cpu().numpy())):
            raise RuntimeError("non-finite admittance gradient")
    return out


LARGE_OCTAVES = (250.0, 500.0, 1000.0, 2000.0, 4000.0)


def large_room_workload(size, steps, seed=5):
    """The large-room configuration: ``size``^3 at dx = 1 cm, the 16-cell sponge on the two x
    faces (rigid behind it), walls of absorption 0.1 on the four others, air at 20 C / 30 % RH,
    a seeded Gaussian-derivative pulse of ``steps`` samples, 32 seeded receivers."""
    from ac_fdtd_amd import (AIR, AbsorbingLayer, AirAbsorption, Grid, WallAdmittances,
                             admittance_from_absorption)
    from ac_fdtd_amd.grid import max_stable_time_step
    from ac_fdtd_amd.sources import gaussian_derivative_pulse

    grid = Grid(dx=DX, nx=size, ny=size, nz=size)
    rng = np.random.default_rng(seed)
    width = rng.uniform(4.0, 8.0) * DX / AIR.sound_speed

    def point():
        return tuple((rng.integers(MARGIN, size - MARGIN, size=3) + 0.5) * DX)

    source = point()
    receivers = [point() for _ in range(32)]
    xi = admittance_from_absorption(0.1)
    config = {
        "walls": WallAdmittances(0.0, 0.0, xi, xi, xi, xi),
        "absorbing_layer": AbsorbingLayer(thickness=SPONGE, faces=((0, 0), (0, 1))),
        "air_absorption": AirAbsorption(temperature=20.0, relative_humidity=30.0),
    }
    dt = max_stable_time_step(DX, AIR.sound_speed)
    return grid, config, source, receivers, gaussian_derivative_pulse(steps, dt, width), dt


def large_room(size, steps, seed=5):
    """The large-room configuration (:func:`large_room_workload`) for ``steps`` steps: fp32
    through the differentiable room (boundary tape attached, so the forward pass is the
    gradient's) and its wall-admittance gradient, then fp64 for the T20 deviation per octave
    band."""
    import torch

    from ac_fdtd_amd.autograd import RoomResponse
    from ac_fdtd_amd.metrics import octave_band_reverberation
    from ac_fdtd_amd.scheme import AcousticFDTD

    grid, config, source, receivers, signal, dt = large_room_workload(size, steps, seed)

    def seconds(call):
        torch.cuda.synchronize()
        began = time.perf_counter()
        out = call()
        torch.cuda.synchronize()
        return time.perf_counter() - began, out

    room = RoomResponse(grid, steps, [(source, signal)], receivers, dtype=np.float32,
                        autotune=True, **config)
    # Tune the forward and adjoint tiles now, outside the timed passes.
    from ac_fdtd_amd.autotune import adjoint_tiling

    adjoint_tiling(room.solver(wall_table_always=True))
    torch.cuda.empty_cache()
    admittance = room.default_admittance(device="cuda").requires_grad_()
    t_forward, recording = seconds(lambda: room(admittance))
    weights = torch.as_tensor(np.random.default_rng(seed + 1).standard_normal(recording.shape),
                              dtype=recording.dtype, device=recording.device)
    t_backward, (gradient,) = seconds(lambda: torch.autograd.grad((recording * weights).sum(),
                                                                  [admittance]))
    report = room.report
    f32 = recording.detach().double().cpu().numpy()
    del recording, weights
    torch.cuda.empty_cache()

    solver = AcousticFDTD(grid, dtype=np.float64, autotune=True, **config)
    solver.add_volume_source(source, signal)
    for r in receivers:
        solver.add_pressure_receiver(r)
    t_f64, _ = seconds(lambda: solver.run(steps))
    f64 = solver.recorded_pressure
    del solver
    torch.cuda.empty_cache()

    rate = 1.0 / dt
    deviation, reference = {}, {}
    for channel in range(f64.shape[0]):
        t_a = octave_band_reverberation(f32[channel], rate, LARGE_OCTAVES)
        t_b = octave_band_reverberation(f64[channel], rate, LARGE_OCTAVES)
        for band in LARGE_OCTAVES:
            if np.isfinite(t_a[band]) and np.isfinite(t_b[band]):
                deviation[band] = max(deviation.get(band, 0.0), abs(t_a[band] / t_b[band] - 1.0))
                reference.setdefault(band, []).append(t_b[band])
    cells = grid.n_cells
    return {
        "grid": [size] * 3,
        "steps": steps,
        "receivers": len(receivers),
        "fp32_forward_with_tape_gcells": round(cells * steps / t_forward / 1e9, 3),
        "fp64_forward_gcells": round(cells * steps / t_f64 / 1e9, 3),
        "wall_gradient_backward_s": round(t_backward, 3),
        "wall_gradient_over_forward": round((t_forward + t_backward) / t_forward, 3),
        "wall_gradient": [float(g) for g in gradient.cpu()],
        "hbm_gradient_storage_bytes": report.tape_bytes + report.checkpoint_bytes,
        "receiver_rel_l2_fp32_vs_fp64": relative_l2(f32, f64),
        "t20_fp64_median_s_by_octave": {f"{int(b)}": float(np.median(v))
                                        for b, v in reference.items()},
        "t20_max_rel_deviation_by_octave": {f"{int(b)}": float(v) for b, v in deviation.items()},
    }

def large_room_air_gradient(size, steps, seed=5):
    """The large-room configuration for ``steps`` steps, fp32: seconds of a forward pass and of
    forward + backward for the gradient with respect to the six wall admittances and the three
    air strengths. ``steps`` rows of divergences do not fit the 64 GiB history budget at 512^3
    (537 MB per step), so this is the checkpointed path: the state every 100 steps, each
    segment's divergences recomputed into a 100-row ring, the adjoint two steps per sweep."""
    import torch

    from ac_fdtd_amd.autograd import RoomResponse

    grid, config, source, receivers, signal, _ = large_room_workload(size, steps, seed)
    room = RoomResponse(grid, steps, [(source, signal[:steps])], receivers, dtype=np.float32,
                        checkpoint_every=100, **config)
    admittance = room.default_admittance(device="cuda").requires_grad_()
    strengths = room.default_air_strength(device="cuda").requires_grad_()

    def seconds(call):
        torch.cuda.synchronize()
        began = time.perf_counter()
        call()
        torch.cuda.synchronize()
        return time.perf_counter() - began

    def forward():
        with torch.no_grad():
            room(admittance.detach(), air_strength=strengths.detach())

    weights = None

    def forward_backward():
        nonlocal weights
        recording = room(admittance, air_strength=strengths)
        if weights is None:
            weights = torch.as_tensor(
                np.random.default_rng(seed + 1).standard_normal(recording.shape),
                dtype=recording.dtype, device=recording.device)
        torch.autograd.grad((recording * weights).sum(), [admittance, strengths])

    forward()
    t_forward = seconds(forward)
    t_both = seconds(forward_backward)
    report = room.report
    out = {
        "grid": [size] * 3,
        "steps": steps,
        "forward_s": round(t_forward, 3),
        "forward_backward_s": round(t_both, 3),
        "grad_over_forward": round(t_both / t_forward, 3),
        "recomputed_steps": report.recomputed_steps,
        "checkpoints": report.checkpoints,
        "hbm_bytes": report.checkpoint_bytes + report.div_tape_bytes + report.tape_bytes,
        "adjoint_two_steps_per_sweep": report.adjoint_two_step,
    }
    del room
    torch.cuda.empty_cache()
    return out


def timed_run(solver, steps: int, warmup: int) -> float:
    """Milliseconds per step of ``solver.run(steps)`` after ``warmup`` steps, from HIP events
    on the current stream."""
    import torch

    from ac_fdtd_amd.scheme import AcousticFDTD

    solver.run(warmup)
    if isinstance(solver, AcousticFDTD) and type(solver).run is AcousticFDTD.run:
        launch = solver.prepare_run(steps)  # the recording and descriptor outside the timing
    else:
        def launch():
            solver.run(steps)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    launch()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def scheme_variants(size: int, steps: int, warmup: int) -> dict:
    """The throughput workload's grid, source and receivers with walls of absorption 0.15 and
    neither sponge nor air (what the pressure-only scheme supports): the pressure-only scheme
    against the first-order scheme on the same configuration. The pressure-only kernel moves 3
    elements per cell-update with one step per sweep and 2 with two (halo re-reads not
    counted); fp32 runs both, fp64 one step per sweep (its default)."""
    import torch

    from ac_fdtd_amd import Grid
    from ac_fdtd_amd.scheme import AcousticFDTD
    from ac_fdtd_amd.sources import gaussian_derivative_pulse
    from ac_fdtd_amd.wave import PressureOnlyFDTD

    grid, config, source, receivers, width = workload(size, steps)
    walls = {"walls": config["walls"]}
    out = {"grid": [size] * 3, "steps": steps, "configuration": "walls alpha 0.15, no sponge, no air"}
    cells = grid.n_cells
    recordings = {}
    for name, cls, dtype, kwargs in (
        ("pressure_only_fp32", PressureOnlyFDTD, np.float32, {"time_block": 2}),
        ("pressure_only_fp32_one_step_per_sweep", PressureOnlyFDTD, np.float32, {"time_block": 1}),
        ("pressure_only_fp64", PressureOnlyFDTD, np.float64, {"time_block": 1}),
        ("first_order_fp32", AcousticFDTD, np.float32, {"autotune": True}),
    ):
        solver = cls(grid, dtype=dtype, **walls, **kwargs)
        solver.add_volume_source(source, gaussian_derivative_pulse(steps + warmup + 8, solver.dt,
                                                                   width))
        for r in receivers:
            solver.add_pressure_receiver(r)
        ms = timed_run(solver, steps, warmup)
        entry = {"gcells": round(cells / (ms * 1e-3) / 1e9, 3), "ms_per_step": round(ms, 5)}
        if cls is PressureOnlyFDTD:
            moved = (2 if kwargs["time_block"] == 2 else 3) * np.dtype(dtype).itemsize
            entry["bytes_per_cell_update"] = moved
            entry["hbm_fraction"] = round(moved * cells / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        recordings[name] = solver.recorded_pressure.astype(np.float64)
        out[name] = entry
        del solver
        torch.cuda.empty_cache()
    # Beyond the 256 MB last-level cache: two fp32 fields of 512^3 are 1.07 GB.
    big = Grid(dx=DX, nx=2 * size, ny=2 * size, nz=2 * size)
    for name, time_block in (("pressure_only_fp32_double_size", 2),
                             ("pressure_only_fp32_double_size_one_step_per_sweep", 1)):
        solver = PressureOnlyFDTD(big, dtype=np.float32, time_block=time_block, **walls)
        solver.add_volume_source(source, gaussian_derivative_pulse(steps + warmup + 8, solver.dt,
                                                                   width))
        ms = timed_run(solver, max(steps // 4, 20), warmup)
        moved = 8 if time_block == 2 else 12
        out[name] = {
            "grid": [2 * size] * 3, "gcells": round(big.n_cells / (ms * 1e-3) / 1e9, 3),
            "ms_per_step": round(ms, 5), "bytes_per_cell_update": moved,
            "hbm_fraction": round(moved * big.n_cells / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        del solver
        torch.cuda.empty_cache()
    # Fourth order in space: rigid walls (images), with the throughput workload's sponge and air.
    from ac_fdtd_amd.fourth import FourthOrderFDTD

    layered = {"absorbing_layer": config["absorbing_layer"],
               "air_absorption": config["air_absorption"]}
    for name, cls, kwargs in (("fourth_order_fp32", FourthOrderFDTD, {}),
                              ("second_order_two_sweeps_fp32", AcousticFDTD,
                               {"fused": False, "time_block": 1})):
        solver = cls(grid, dtype=np.float32, **layered, **kwargs)
        solver.add_volume_source(source, gaussian_derivative_pulse(steps + warmup + 8, solver.dt,
                                                                   width))
        for r in receivers:
            solver.add_pressure_receiver(r)
        ms = timed_run(solver, steps, warmup)
        n_relax = solver.coefficients.n_relax
        out[name] = {"configuration": "rigid walls, 16-cell sponge, air 20 C / 50 %",
                     "gcells": round(cells / (ms * 1e-3) / 1e9, 3), "ms_per_step": round(ms, 5),
                     "hbm_fraction": round(update_bytes(n_relax, 4) * cells / (ms * 1e-3) / 1e9
                                           / HBM_PEAK_GBS, 4)}
        del solver
        torch.cuda.empty_cache()
    out["pressure_only_over_first_order_speed"] = round(
        out["pressure_only_fp32"]["gcells"] / out["first_order_fp32"]["gcells"], 3)
    out["pressure_only_fp32_vs_fp64_receiver_rel_l2"] = relative_l2(
        recordings["pressure_only_fp32"], recordings["pressure_only_fp64"])
    return out


def boundary_variants(size: int, steps: int, warmup: int) -> dict:
    """The two boundaries the reference plans beyond its sponge and real admittance: the
    split-field PML's reflection against the sponge's (a 96^3 room lined with 12 cells on all
    faces, three receivers from the centre to a corner, residual against a free domain), and
    the fp32 throughput at ``size``^3 of the PML (16-cell layer) and of walls with
    frequency-dependent admittance fitted to a porous absorber's octave table."""
    import torch

    from ac_fdtd_amd import AbsorbingLayer, Grid
    from ac_fdtd_amd.iir_walls import IIRWallFDTD
    from ac_fdtd_amd.impedance import FrequencyDependentWalls
    from ac_fdtd_amd.pml import PMLFDTD
    from ac_fdtd_amd.scheme import AcousticFDTD
    from ac_fdtd_amd.sources import gaussian_derivative_pulse

    out = {}
    n, depth, run_steps, dx = 96, 12, 160, 0.01
    pulse = np.exp(-((np.arange(run_steps) - 30.0) / 8.0) ** 2)
    shift = run_steps * dx
    points = [(0.5 * n, 0.5 * n, 0.5 * n), (depth + 4, 0.5 * n, 0.5 * n),
              (depth + 4, depth + 4, depth + 4)]

    def record(solver, offset):
        solver.add_volume_source(tuple(c * dx + offset for c in (0.5 * n, 0.4 * n, 0.45 * n)),
                                 pulse)
        for point in points:
            solver.add_pressure_receiver(tuple(c * dx + offset for c in point))
        solver.run(run_steps)
        return solver.recorded_pressure

    side = n + 2 * run_steps
    free = record(AcousticFDTD(Grid(dx=dx, nx=side, ny=side, nz=side), time_block=1), shift)
    torch.cuda.empty_cache()
    grid = Grid(dx=dx, nx=n, ny=n, nz=n)
    layer = AbsorbingLayer(thickness=depth)
    for name, solver in (("pml", PMLFDTD(grid, layer=layer)),
                         ("sponge", AcousticFDTD(grid, absorbing_layer=layer))):
        recorded = record(solver, 0.0)
        out[f"{name}_residual_db_centre_face_corner"] = [
            round(20 * math.log10(np.linalg.norm(recorded[r] - free[r]) / np.linalg.norm(free[r])), 1)
            for r in range(3)]
    del free
    torch.cuda.empty_cache()

    big, config, source, receivers, width = workload(size, steps)
    porous = {125: 0.1, 250: 0.2, 500: 0.45, 1000: 0.7, 2000: 0.8, 4000: 0.85}
    for name, solver in (
        ("pml_fp32", PMLFDTD(big, dtype=np.float32, layer=AbsorbingLayer(thickness=SPONGE))),
        ("iir_walls_fp32", IIRWallFDTD(big, dtype=np.float32,
                                       walls=FrequencyDependentWalls.uniform(porous),
                                       air_absorption=config["air_absorption"])),
    ):
        solver.add_volume_source(source, gaussian_derivative_pulse(steps + warmup + 8, solver.dt,
                                                                   width))
        for r in receivers:
            solver.add_pressure_receiver(r)
        ms = timed_run(solver, steps, warmup)
        out[name] = {"grid": [size] * 3, "gcells": round(big.n_cells / (ms * 1e-3) / 1e9, 3),
                     "ms_per_step": round(ms, 5)}
        del solver
        torch.cuda.empty_cache()
    return out


def stream_ceiling(gib: float = 1.0, copies: int = 20) -> float:
    """GB/s (read + write) of the library's 16-byte-per-lane copy over ``gib`` GiB, four times
    the 256 MB MALL: the HBM bandwidth a plain streaming kernel reaches here."""
    import ctypes

    import torch

    from ac_fdtd_amd import _native

    lib = _native.library()
    n_bytes = int(gib * 2**30)
    src = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    handle = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def copy():
        _native.check(lib.acfdtd_hip_stream_copy(ctypes.c_void_p(src.data_ptr()),
                                                 ctypes.c_void_p(dst.data_ptr()), n_bytes,
                                                 handle), "acfdtd_hip_stream_copy")

    for _ in range(3):
        copy()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(copies):
        copy()
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / copies
    return 2.0 * n_bytes / (ms * 1e-3) / 1e9


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as info:
            for line in info:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def numpy_baseline(size, seconds):
    """The oracle's NumPy restatement of the reference loop (one thread, as NumPy runs), fp32,
    same workload, bounded sample."""
    from oracle.scheme import OracleFDTD

    solver = make_solver(OracleFDTD, size, 4000, np.float32)
    solver.run(1)
    started = time.perf_counter()
    solver.run(1)
    per_step = time.perf_counter() - started
    steps = max(2, min(200, int(seconds / per_step)))
    started = time.perf_counter()
    solver.run(steps)
    elapsed = time.perf_counter() - started
    return {
        "value": round(solver.grid.n_cells * steps / elapsed / 1e9, 5),
        "unit": "Gcell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"NumPy restatement of the reference step (oracle/scheme.py), fp32, same "
        f"{size}^3 workload, {steps} steps after 2 warm-up steps, {elapsed:.1f} s",
    }

def cpu_baseline(size, seconds):
    """The reference C loop (oracle/_ref) on the host, fp32, same workload, bounded sample."""
    from oracle import ref_c

    if not ref_c.is_available(np.float32):
        return None
    threads = ref_c.host_thread_count()
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    solver = make_solver(ref_c.RefCFDTD, size, 4000, np.float32, threads=threads)
    solver.run(2)
    started = time.perf_counter()
    solver.run(2)
    per_step = (time.perf_counter() - started) / 2
    steps = max(3, min(4000, int(seconds / per_step)))
    started = time.perf_counter()
    solver.run(steps)
    elapsed = time.perf_counter() - started
    return {
        "value": round(solver.grid.n_cells * steps / elapsed / 1e9, 4),
        "unit": "Gcell-updates/s",
        "cores": solver.threads,
        "kind": "reference",
        "sample": f"reference C loop (acfdtd.c, fp32, -O3 -ffp-contract=off, pthreads) on the "
        f"same {size}^3 walls+sponge+air workload, {steps} steps after 4 warm-up steps, "
        f"{elapsed:.1f} s, on {cpu_model()}",
    }


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=4000)
    parser.add_argument("--warmup", type=int, default=3)
    parser.add_argument("--size", type=int, default=256)
    parser.add_argument("--cpu-seconds", type=float, default=15.0)
    parser.add_argument("--skip-cpu", action="store_true")
    parser.add_argument("--skip-extra", action="store_true")
    parser.add_argument("--batch-rooms", type=int, default=64)
    parser.add_argument("--large-steps", type=int, default=8000,
                        help="steps of the 512^3 room (0 skips it)")
    parser.add_argument("--batch-steps", type=int, default=3000)
    parser.add_argument("--steps-variants", type=int, default=400,
                        help="steps of each scheme-variant run (pressure-only, fourth order)")
    parser.add_argument("--accuracy-steps", type=int, default=4000,
                        help="steps of the fp32-vs-fp64 and energy-drift runs (any --steps)")
    parser.add_argument("--skip-check", action="store_true",
                        help="skip the fp64 single-room runs that check the batched rooms")
    parser.add_argument(
        "--mode",
        choices=("bench", "copy", "gradient", "batch", "variants", "large-gradient"),
        default="bench",
        help="copy: five plain 512 MiB device-to-device copies, the known-size reference a "
        "profiler's FETCH_SIZE / WRITE_SIZE counters are scaled by; gradient: only the "
        "wall-gradient cost at 128^3 (for profiling the adjoint kernels); batch: only the "
        "batched rooms (for profiling the batch kernels)",
    )
    args = parser.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(
            f"bench.py measures one GPU; --gpus {args.gpus} / WORLD_SIZE {world} is not "
            "supported (rooms shard by slab with one-plane halo exchanges, which this build "
            "does not implement)",
            file=sys.stderr,
        )
        return 2

    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 2

    if args.mode == "copy":
        source = torch.ones(1 << 27, dtype=torch.float32, device="cuda")  # 512 MiB
        target = torch.empty_like(source)
        for _ in range(5):
            target.copy_(source)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": source.numel() * 4, "copies": 5}), flush=True)
        return 0
    if args.mode == "large-gradient":
        print(json.dumps(large_room_air_gradient(512, 1000)), flush=True)
        return 0
    if args.mode == "gradient":
        print(json.dumps(gradient_cost(128, 2000)), flush=True)
        return 0
    if args.mode == "variants":
        print(json.dumps({"schemes": scheme_variants(args.size, args.steps_variants, args.warmup),
                          "boundaries": boundary_variants(args.size, args.steps_variants,
                                                          args.warmup)}), flush=True)
        return 0
    if args.mode == "batch":
        print(json.dumps(batched_rooms(args.batch_rooms, args.batch_steps, args.warmup,
                                       check=not args.skip_check)), flush=True)
        return 0

    log(f"fp32 {args.size}^3, {args.steps} steps")
    f32 = gpu_throughput(args.size, args.steps, args.warmup, np.float32)
    log(f"fp32: {f32['gcells']:.2f} Gcell/s, kernels {f32['kernels']}")
    # The dominant kernel: the one with the largest mean launch time. Its roofline counts the
    # bytes that kernel itself must move per launch (each field element it reads or writes
    # once, e.g. 28 B per cell-update for the two-step sweep in fp32 with three relaxation
    # states, two cell-updates per launch), so the fraction is one of HBM bandwidth.
    dominant = max(f32["kernels"], key=f32["kernels"].get)
    sweep_bytes = launch_minimum_bytes(dominant, f32["n_relax"], 4)
    achieved = sweep_bytes * f32["cells"] / (f32["kernels"][dominant] * 1e-3) / 1e9
    traffic = measured_traffic(args.size, dominant)

    extra = {
        "time_block": f32["time_block"],
        "staged": f32["staged"],
        "tiling_tile_j_chunk_x": f32["tiling"],
        "kernel_ms": {k: round(v, 4) for k, v in f32["kernels"].items()},
        "kernel_minimum_bytes_per_cell_update": minimum_bytes(dominant, f32["n_relax"], 4),
        # The same launch time on the bytes of the reference's own step structure (two sweeps,
        # every field across HBM every step: 72 B per cell-update in fp32 with air). Above 1
        # for kernels that advance several steps per pass over memory: not an HBM fraction.
        "reference_step_equivalent": {
            "bytes_per_cell_update": update_bytes(f32["n_relax"], 4),
            "gbs": round(launch_bytes(dominant, f32["n_relax"], 4) * f32["cells"]
                         / (f32["kernels"][dominant] * 1e-3) / 1e9, 1),
        },
    }
    if traffic is not None:
        extra["measured_traffic_over_minimum"] = round(traffic / (sweep_bytes * f32["cells"]), 4)
        extra["measured_traffic_gbs"] = round(traffic / (f32["kernels"][dominant] * 1e-3) / 1e9, 1)
    if not args.skip_extra:
        log("fp32 one step per sweep")
        one = gpu_throughput(args.size, args.steps, args.warmup, np.float32, time_block=1)
        kind = max(one["kernels"], key=one["kernels"].get)
        extra["fp32_one_step_per_sweep"] = {
            "gcells": round(one["gcells"], 3),
            "tiling_tile_j_chunk_x": one["tiling"],
            "kernel_ms": {k: round(v, 4) for k, v in one["kernels"].items()},
            "roofline_frac": round(launch_minimum_bytes(kind, one["n_relax"], 4) * one["cells"]
                                   / (one["kernels"][kind] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "bitwise_equal_recordings": bool(np.array_equal(one["recorded"], f32["recorded"])),
        }
        del one
        # Accuracy runs: a fixed number of steps whatever --steps is, long enough for every
        # receiver to record the direct sound and the first reflections.
        log(f"accuracy: fp32, fp32 compensated, fp64 over {args.accuracy_steps} steps")
        acc32 = (f32 if args.steps == args.accuracy_steps else
                 gpu_throughput(args.size, args.accuracy_steps, args.warmup, np.float32))
        comp = gpu_throughput(args.size, args.accuracy_steps, args.warmup, np.float32,
                              compensated=True)
        extra["fp32_compensated_gcells"] = round(comp["gcells"], 3)
        log("fp64 plain")
        f64 = gpu_throughput(args.size, args.accuracy_steps, args.warmup, np.float64)
        extra["fp64_plain_gcells"] = round(f64["gcells"], 3)
        extra["fp32_over_fp64_speed"] = round(f32["gcells"] / f64["gcells"], 3)
        extra["fp64_time_block"] = f64["time_block"]
        kind64 = max(f64["kernels"], key=f64["kernels"].get)
        extra["fp64_dominant_kernel_hbm_fraction"] = round(
            launch_minimum_bytes(kind64, f64["n_relax"], 8) * f64["cells"]
            / (f64["kernels"][kind64] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        extra["accuracy_steps"] = args.accuracy_steps
        extra["fp32_error_vs_fp64"] = error_against(acc32, f64)
        extra["fp32_compensated_error_vs_fp64"] = error_against(comp, f64)
        for key in ("fp32_error_vs_fp64", "fp32_compensated_error_vs_fp64"):
            if not math.isfinite(extra[key]["receiver_rel_l2"]):
                raise RuntimeError(f"{key}: non-finite receiver error")
        del comp, f64, acc32
        log("fp32 384^3")
        big = gpu_throughput(384, args.steps, args.warmup, np.float32)
        extra["fp32_384_gcells"] = round(big["gcells"], 3)
        del big
        log(f"rigid-box energy drift, fp64 and fp32, {args.accuracy_steps} steps")
        extra["fp64_rigid_energy_drift_256"] = rigid_energy_drift(args.size, args.accuracy_steps)
        extra["fp32_rigid_energy_drift_256"] = rigid_energy_drift(args.size, args.accuracy_steps,
                                                                  np.float32)
        torch.cuda.empty_cache()
        if args.batch_rooms > 0:
            log(f"{args.batch_rooms} batched rooms, {args.batch_steps} steps")
            extra["batched_rooms_fp32"] = batched_rooms(args.batch_rooms, args.batch_steps,
                                                        args.warmup, check=not args.skip_check)
            log(f"batched: {extra['batched_rooms_fp32']}")
        log("scheme variants")
        extra["scheme_variants"] = scheme_variants(args.size, args.steps_variants, args.warmup)
        log(f"variants: {extra['scheme_variants']}")
        log("boundary variants")
        extra["boundary_variants"] = boundary_variants(args.size, args.steps_variants, args.warmup)
        log(f"boundaries: {extra['boundary_variants']}")
        log("gradient cost")
        extra["wall_gradient_cost_fp32"] = gradient_cost(128, 2000)
        log(f"gradient: {extra['wall_gradient_cost_fp32']}")
        torch.cuda.empty_cache()
        if args.large_steps > 0:
            log(f"large room 512^3, {args.large_steps} steps")
            extra["large_room_512"] = large_room(512, args.large_steps)
            log(f"large room: {extra['large_room_512']}")
            log("large room 512^3: wall and air gradients, 1000 steps")
            extra["large_room_512"]["walls_and_air_gradient"] = large_room_air_gradient(512, 1000)
            log(f"large room air gradient: {extra['large_room_512']['walls_and_air_gradient']}")

    log("stream ceiling")
    ceiling = stream_ceiling()
    extra["stream_copy_ceiling_gbs"] = round(ceiling, 1)
    extra["dominant_kernel_fraction_of_stream_ceiling"] = round(achieved / ceiling, 4)

    baseline = None
    if not args.skip_cpu:
        log("cpu baseline")
        baseline = cpu_baseline(args.size, args.cpu_seconds)
        log("numpy baseline")
        extra["numpy_baseline"] = numpy_baseline(args.size, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(f32["gcells"], 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"throughput {args.size}^3, dx 1 cm, walls alpha 0.15, 16-cell sponge on "
            "all faces, air 20 C / 50 % RH, seeded Gaussian-derivative pulse, 8 receivers",
            "grid": [args.size] * 3,
            "relaxation_states": f32["n_relax"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": dominant,
            "algorithmic_bytes_per_launch": sweep_bytes * f32["cells"],
        },
        "cpu_baseline": baseline,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
