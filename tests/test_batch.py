"""Batched rooms: one launch per step over many uneven rooms, bit-identical to each room alone.

The rooms differ in size (including nz that are multiples of 4, of 2 only, and odd: those the
widest lane does not divide run in padded rows, all rooms in one launch), walls, air, sources
and receivers; every room's fields and recordings must equal those of the same room run alone,
for even and odd step counts (the odd one through the in-place sweeps on padded rows), fp64 and
fp32.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from ac_fdtd_amd import AbsorbingLayer, AirAbsorption, Grid, WallAdmittances

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

SHAPES = [(23, 17, 12), (9, 31, 20), (16, 8, 36), (30, 21, 9), (5, 5, 5), (12, 40, 44)]


def _room(shape, seed, dtype, layer):
    from ac_fdtd_amd.scheme import AcousticFDTD

    rng = np.random.default_rng(seed)
    grid = Grid(dx=0.03, nx=shape[0], ny=shape[1], nz=shape[2])
    solver = AcousticFDTD(
        grid,
        dtype=dtype,
        walls=WallAdmittances(*rng.uniform(0.0, 0.5, 6)),
        air_absorption=AirAbsorption(rng.uniform(15, 25), rng.uniform(30, 70)),
        absorbing_layer=AbsorbingLayer(thickness=2) if layer else None,
    )
    lengths = np.array(grid.lengths)
    solver.p.copy_(torch.as_tensor(rng.standard_normal(grid.shape)))
    solver.add_volume_source(tuple(rng.uniform(0.2, 0.8) * lengths), rng.standard_normal(40))
    for _ in range(2):
        solver.add_pressure_receiver(tuple(rng.uniform(0.1, 0.9) * lengths))
    return solver


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("layer", [False, True])
@pytest.mark.parametrize("shapes", [SHAPES, [s for s in SHAPES if s[2] % 4 == 0]],
                         ids=["mixed-nz", "nz-multiple-of-4"])
def test_batch_equals_rooms_alone(dtype, layer, shapes):
    from ac_fdtd_amd.batch import BatchedRooms

    alone = [_room(shape, n, dtype, layer) for n, shape in enumerate(shapes)]
    batched = [_room(shape, n, dtype, layer) for n, shape in enumerate(shapes)]
    batch = BatchedRooms(batched)
    assert batch.two_step == [True]  # one launch group; two steps per sweep for any nz
    for n_steps in (6, 3, 10):
        batch.run(n_steps)
        for solver in alone:
            solver.run(n_steps)
    batch.close()
    for a, b in zip(alone, batched):
        assert b.step_index == a.step_index == 19
        np.testing.assert_array_equal(b.p.cpu().numpy(), a.p.cpu().numpy())
        for axis in range(3):
            np.testing.assert_array_equal(b.velocity[axis].cpu().numpy(),
                                          a.velocity[axis].cpu().numpy())
        np.testing.assert_array_equal(b.recorded_pressure, a.recorded_pressure)


def test_batch_refuses_mixed_precisions_and_groups_mixed_kernels():
    """Rooms of different precisions are refused; rooms whose kernels differ (here sponge
    presence) advance in one launch group each, still each as if alone."""
    from ac_fdtd_amd.batch import BatchedRooms

    a = _room((8, 8, 8), 0, np.float64, False)
    b = _room((8, 8, 8), 1, np.float32, False)
    with pytest.raises(ValueError, match="dtype"):
        BatchedRooms([a, b])
    c = _room((9, 8, 7), 2, np.float64, True)
    alone = [_room((8, 8, 8), 0, np.float64, False), _room((9, 8, 7), 2, np.float64, True)]
    batch = BatchedRooms([a, c])
    assert batch.two_step == [True, True]
    batch.run(9)
    batch.close()
    for solver, reference in zip((a, c), alone):
        reference.run(9)
        np.testing.assert_array_equal(solver.p.cpu().numpy(), reference.p.cpu().numpy())
        np.testing.assert_array_equal(solver.recorded_pressure, reference.recorded_pressure)
