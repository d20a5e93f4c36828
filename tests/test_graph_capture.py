"""Step loops captured into a HIP graph replay the eager run bit for bit.

A capture records every launch of ``run`` (two-step sweeps, one-step pairs, a last single
step) with its step indices; each replay starts from a state restored eagerly and must give
the fields, relaxation states and receiver samples of an eager run of the same steps from that
state. Replays alternate with eager work and synchronisations, and every one is checked.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from cases import PARITY

from ac_fdtd_amd import Grid

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]


def _solver(shape, dtype, time_block):
    from ac_fdtd_amd.scheme import AcousticFDTD

    rng = np.random.default_rng(23)
    grid = Grid(dx=0.03, nx=shape[0], ny=shape[1], nz=shape[2])
    solver = AcousticFDTD(grid, dtype=dtype, time_block=time_block, **PARITY["all"])
    solver.p.copy_(torch.as_tensor(rng.standard_normal(grid.shape)))
    lengths = np.array(grid.lengths)
    solver.add_volume_source(tuple(0.4 * lengths), rng.standard_normal(40))
    for f in (0.2, 0.55, 0.8):
        solver.add_pressure_receiver(tuple(f * lengths))
    return solver


@pytest.mark.parametrize("n_steps", [12, 15])
@pytest.mark.parametrize("time_block", [1, 2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_captured_steps_replay_the_eager_run(dtype, time_block, n_steps):
    eager = _solver((14, 11, 16), dtype, time_block)
    graph = _solver((14, 11, 16), dtype, time_block)
    eager.run(5)
    graph.run(5)
    start = graph.save_state()
    captured = graph.capture(n_steps)
    assert graph.step_index == 5 and captured.step0 == 5
    eager.run(n_steps)
    want = eager.save_state()
    want_recording = eager.recording_tensor()[:, 5:].clone()
    for replay in range(3):
        graph.load_state(start, 5)
        graph.p.add_(0.0)  # eager work on the same stream between replays
        captured.replay()
        torch.cuda.synchronize()
        for got, ref in zip(graph.save_state(), want):
            assert torch.equal(got, ref), f"replay {replay}"
        assert torch.equal(captured.recording, want_recording), f"replay {replay}"
    assert graph.step_index == 5


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_captured_batch_replays_the_eager_batch(dtype):
    """A batch step loop in a graph: replays from restored states give the eager batch's
    fields and samples, also after an eager run has re-pointed the room tables."""
    from ac_fdtd_amd.batch import BatchedRooms

    shapes = [(14, 11, 16), (9, 13, 10), (12, 7, 9)]  # fp32: the last two in padded rows
    eager = [_solver(shape, dtype, 2) for shape in shapes]
    graph = [_solver(shape, dtype, 2) for shape in shapes]
    eager_batch, graph_batch = BatchedRooms(eager), BatchedRooms(graph)
    eager_batch.run(3)
    graph_batch.run(3)
    starts = [s.save_state() for s in graph]
    captured = graph_batch.capture(14)
    eager_batch.run(14)
    want = [s.save_state() for s in eager]
    want_recordings = [s.recording_tensor()[:, 3:].clone() for s in eager]
    for replay in range(3):
        for solver, start in zip(graph, starts):
            solver.load_state(start, 3)
        if replay == 1:
            graph_batch.run(2)  # an eager run re-points the room tables
            for solver, start in zip(graph, starts):
                solver.load_state(start, 3)
        captured.replay()
        torch.cuda.synchronize()
        for solver, ref, ref_rec, rec in zip(graph, want, want_recordings, captured.recordings):
            for got, exp in zip(solver.save_state(), ref):
                assert torch.equal(got, exp), f"replay {replay}"
            assert torch.equal(rec, ref_rec), f"replay {replay}"
    eager_batch.close()
    graph_batch.close()
