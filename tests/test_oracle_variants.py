"""Every kernel variant the bench and the batch run, against the CPU oracle at the same dtype.

The oracle (``oracle/scheme.py``) is bit-for-bit the reference NumPy scheme in fp64 and fp32
(``tests/test_oracle.py``, golden vectors), so each HIP path here must equal it bit for bit:
fields, relaxation states and recordings. Grids are chosen for the lane width the kernels pick
from nz: nz = 16 gives 4-cell fp32 lanes (the 16-byte lanes of the bench's headline sweep,
with its register prefetch of the next plane) and 2-cell fp64 lanes (with the fp64 register
prefetch when there is no sponge), nz = 14 gives 2-cell lanes in both precisions and nz = 15
single-cell lanes. Each runs one step per sweep, two steps per sweep, and two steps per sweep
staged through LDS, on the reference's five parity configurations (rigid, walls, sponge, air,
all of them), plus the bench's own tile shape (6 rows of 64 lane groups, 24-plane marches) on
a slab of the 256^3 workload's rows. Batched rooms (the batch's two-step kernel on uneven rooms,
those whose nz the widest lane does not divide in padded rows) are compared room by room with
their own oracles.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from cases import PARITY, PARITY_RECEIVER, PARITY_SOURCE, parity_signal

from ac_fdtd_amd import AbsorbingLayer, AirAbsorption, Grid, WallAdmittances
from oracle.scheme import OracleFDTD

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

STEPS = 150  # a multiple of four plus two: two-step quads, then a fused pair
MODES = {
    "one-step": {"time_block": 1},
    "two-step": {"time_block": 2},
    "two-step-staged": {"time_block": 2, "staged": True},
}
NZ = {16: "nz16", 14: "nz14", 15: "nz15"}


def _grid(nz: int, nx: int = 20, ny: int = 17) -> Grid:
    return Grid(dx=0.03, nx=nx, ny=ny, nz=nz)


def _prepare(solver, grid, seed):
    rng = np.random.default_rng(seed)
    initial = rng.standard_normal(grid.shape)
    velocity = [1e-3 * rng.standard_normal(grid.velocity_shape(a)) for a in range(3)]
    if isinstance(solver.p, torch.Tensor):
        solver.p.copy_(torch.as_tensor(initial))
        for a in range(3):
            solver.velocity[a].copy_(torch.as_tensor(velocity[a]))
    else:
        solver.p[:] = initial
        for a in range(3):
            solver.velocity[a][:] = velocity[a]
    solver.add_volume_source(PARITY_SOURCE, parity_signal())
    solver.add_volume_source((0.45, 0.1, 0.3), parity_signal()[::3])
    for point in (PARITY_RECEIVER, PARITY_SOURCE, (0.05, 0.45, 0.4)):
        solver.add_pressure_receiver(point)
    return solver


def _assert_equal_to_oracle(candidate, reference):
    np.testing.assert_array_equal(candidate.p.cpu().numpy(), reference.p)
    for axis in range(3):
        np.testing.assert_array_equal(candidate.velocity[axis].cpu().numpy(),
                                      reference.velocity[axis])
    for n, state in enumerate(reference.relaxation_states):
        np.testing.assert_array_equal(candidate.relaxation_states[n].cpu().numpy(), state)
    np.testing.assert_array_equal(candidate.recorded_pressure, reference.recorded_pressure)


def _pair(grid, dtype, config, seed=3, **kwargs):
    from ac_fdtd_amd.scheme import AcousticFDTD

    reference = _prepare(OracleFDTD(grid, dtype=dtype, **config), grid, seed)
    candidate = _prepare(AcousticFDTD(grid, dtype=dtype, **config, **kwargs), grid, seed)
    return candidate, reference


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", list(PARITY))
@pytest.mark.parametrize("nz", list(NZ), ids=list(NZ.values()))
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
def test_every_sweep_is_bitwise_the_oracle(dtype, nz, name, mode):
    grid = _grid(nz)
    candidate, reference = _pair(grid, dtype, PARITY[name], **MODES[mode])
    assert candidate.two_step_applies() == (MODES[mode]["time_block"] == 2)
    for n_steps in (STEPS - 7, 7):  # the second run starts on an odd step
        candidate.run(n_steps)
        reference.run(n_steps)
    _assert_equal_to_oracle(candidate, reference)


@pytest.mark.parametrize("mode", ["two-step", "two-step-staged"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
def test_the_bench_tile_shape_is_bitwise_the_oracle(dtype, mode):
    """The tuned shape of the 256^3 bench (tile_j 3: 6 own rows and 2 halo rows of 64 lane
    groups of 4 cells, 24-plane marches) on 56 planes of 20 rows of 256 cells: several x
    chunks and y blocks, so workgroup edges, halo rows and halo planes all fall inside the
    grid; walls, the sponge and air all on."""
    grid = _grid(256, nx=56, ny=20)
    config = {
        "walls": WallAdmittances.from_absorption(0.15),
        "absorbing_layer": AbsorbingLayer(thickness=5),
        "air_absorption": AirAbsorption(temperature=20.0, relative_humidity=50.0),
    }
    candidate, reference = _pair(grid, dtype, config, seed=8, tile_j=3, chunk_x=24,
                                 **MODES[mode])
    candidate.run(30)
    reference.run(30)
    _assert_equal_to_oracle(candidate, reference)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
def test_walls_and_air_without_a_sponge_is_bitwise_the_oracle(dtype):
    """The fp64 register prefetch runs only without the sponge (2-cell lanes); fp32 4-cell lanes
    prefetch either way. Row segments (nz = 512 is 128 lane groups, longer than a block)."""
    grid = _grid(512, nx=12, ny=9)
    config = {"walls": WallAdmittances(0.1, 0.3, 0.0, 0.2, 0.05, 0.4),
              "air_absorption": AirAbsorption(18.0, 40.0)}
    for mode in ("two-step", "two-step-staged"):
        candidate, reference = _pair(grid, dtype, config, seed=9, **MODES[mode])
        candidate.run(22)
        reference.run(22)
        _assert_equal_to_oracle(candidate, reference)


@pytest.mark.parametrize("compensated", [False, True])
def test_fp32_four_cell_lanes_against_fp64(compensated):
    """Not bitwise (the compensated increment rounds once instead of R + 2 times): fp32 on
    4-cell lanes stays within 1e-4 of peak of the fp64 oracle over 400 steps, the reference's
    own fp32 tolerance (test_c_backend.py:101-109)."""
    from ac_fdtd_amd.scheme import AcousticFDTD

    grid = _grid(16)
    reference = _prepare(OracleFDTD(grid, **PARITY["all"]), grid, 4)
    single = _prepare(AcousticFDTD(grid, dtype=np.float32, compensated=compensated,
                                   **PARITY["all"]), grid, 4)
    reference.run(400)
    single.run(400)
    recorded = reference.recorded_pressure
    error = np.max(np.abs(recorded - single.recorded_pressure)) / np.max(np.abs(recorded))
    assert error < 1e-4


# Uneven rooms: nz multiples of 4, of 2 only and odd (fp32: the last two in padded rows).
BATCH_SHAPES = [(23, 17, 12), (9, 31, 22), (16, 8, 36), (30, 21, 9), (12, 14, 44), (7, 19, 15)]


def _batch_room(cls, shape, seed, dtype, layer):
    rng = np.random.default_rng(seed)
    grid = Grid(dx=0.03, nx=shape[0], ny=shape[1], nz=shape[2])
    solver = cls(
        grid,
        dtype=dtype,
        walls=WallAdmittances(*rng.uniform(0.0, 0.5, 6)),
        air_absorption=AirAbsorption(rng.uniform(15, 25), rng.uniform(30, 70)),
        absorbing_layer=AbsorbingLayer(thickness=2) if layer else None,
    )
    lengths = np.array(grid.lengths)
    initial = rng.standard_normal(grid.shape)
    if isinstance(solver.p, torch.Tensor):
        solver.p.copy_(torch.as_tensor(initial))
    else:
        solver.p[:] = initial
    solver.add_volume_source(tuple(rng.uniform(0.2, 0.8) * lengths), rng.standard_normal(40))
    for _ in range(2):
        solver.add_pressure_receiver(tuple(rng.uniform(0.1, 0.9) * lengths))
    return solver


@pytest.mark.parametrize("layer", [False, True], ids=["walls-air", "walls-air-sponge"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
def test_batched_rooms_are_bitwise_their_own_oracles(dtype, layer):
    from ac_fdtd_amd.batch import BatchedRooms
    from ac_fdtd_amd.scheme import AcousticFDTD

    rooms = [_batch_room(AcousticFDTD, s, n, dtype, layer) for n, s in enumerate(BATCH_SHAPES)]
    oracles = [_batch_room(OracleFDTD, s, n, dtype, layer) for n, s in enumerate(BATCH_SHAPES)]
    batch = BatchedRooms(rooms)
    assert all(batch.two_step)
    for n_steps in (9, 12):
        batch.run(n_steps)
        for oracle in oracles:
            oracle.run(n_steps)
    batch.close()
    for room, oracle in zip(rooms, oracles):
        _assert_equal_to_oracle(room, oracle)
