"""Two steps per sweep (time_block=2) against one step per sweep, bit for bit.

The two-step sweep recomputes each cell's first step in halo lanes and halo planes and keeps
the fields of the intermediate step in registers and LDS; every cell's arithmetic is the
single-step sweep's, so fields, relaxation states and recordings must be identical. The cases
put workgroup edges inside the grid (rows of lane groups split across workgroups, several x
chunks), run step counts with every remainder modulo four (the remainder runs as single
steps), and cover every combination of walls, sponge, air and compensation in fp32 and fp64.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from cases import PARITY

from ac_fdtd_amd import AbsorbingLayer, AirAbsorption, Grid, WallAdmittances

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

# (shape, tile_j, chunk_x): blocks of 2 tile_j rows; whole rows of nz / V lane groups when
# (2 tile_j + 2) rows fit in 512 lanes, else row segments with a halo lane group at each end
# (nz = 512 and the 37-row case, and nz = 87 with 1-cell lanes). V is the widest of 4, 2, 1
# (fp32) or 2, 1 (fp64) that divides nz.
CASES = [
    ((9, 11, 12), 4, 8),
    ((13, 40, 64), 1, 3),
    ((13, 40, 64), 2, 1),
    ((21, 17, 32), 8, 5),
    ((1, 6, 8), 1, 1),
    ((34, 30, 128), 2, 4),
    ((9, 13, 512), 2, 4),
    ((11, 37, 128), 8, 4),
    ((10, 9, 10), 2, 3),   # fp32 lanes of 2 cells
    ((12, 11, 15), 1, 4),  # lanes of 1 cell
    ((9, 13, 87), 2, 4),   # 1-cell lanes, rows of 87 lanes in segments
    ((7, 10, 160), 0, 0),  # default shape: 6 whole rows (fp32), 4 whole rows (fp64)
    ((6, 19, 87), 0, 0),   # default shape: 8 rows in segments
]


def _solver(shape, dtype, config, time_block, tile_j, chunk_x, compensated=False, staged=False):
    from ac_fdtd_amd.scheme import AcousticFDTD

    rng = np.random.default_rng(11)
    grid = Grid(dx=0.03, nx=shape[0], ny=shape[1], nz=shape[2])
    solver = AcousticFDTD(grid, dtype=dtype, time_block=time_block, tile_j=tile_j,
                          chunk_x=chunk_x, compensated=compensated, staged=staged, **config)
    solver.p.copy_(torch.as_tensor(rng.standard_normal(grid.shape)))
    for a in range(3):
        solver.velocity[a].copy_(torch.as_tensor(1e-3 * rng.standard_normal(grid.velocity_shape(a))))
    lengths = np.array(grid.lengths)
    solver.add_volume_source(tuple(0.37 * lengths), rng.standard_normal(23))
    solver.add_volume_source(tuple(0.71 * lengths), rng.standard_normal(9))
    for f in (0.13, 0.5, 0.86):
        solver.add_pressure_receiver(tuple(f * lengths))
    return solver


def _assert_same(a, b):
    assert torch.equal(a.p, b.p)
    for axis in range(3):
        assert torch.equal(a.velocity[axis], b.velocity[axis])
    assert torch.equal(a.relaxation_states, b.relaxation_states)
    np.testing.assert_array_equal(a.recorded_pressure, b.recorded_pressure)


CONFIGS = {
    "rigid": {},
    "walls": {"walls": WallAdmittances(0.1, 0.3, 0.0, 0.2, 0.05, 0.4)},
    "layer": {"absorbing_layer": AbsorbingLayer(thickness=3)},
    "air": {"air_absorption": AirAbsorption(18.0, 40.0)},
    "all": PARITY["all"],
}


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("config", list(CONFIGS))
@pytest.mark.parametrize("case", range(len(CASES)))
def test_two_step_sweep_is_bitwise_single_steps(dtype, config, case):
    shape, tile_j, chunk_x = CASES[case]
    one = _solver(shape, dtype, CONFIGS[config], 1, tile_j, chunk_x)
    two = _solver(shape, dtype, CONFIGS[config], 2, tile_j, chunk_x)
    for n_steps in (8, 5, 4, 7, 6):
        one.run(n_steps)
        two.run(n_steps)
    _assert_same(one, two)
    assert one.step_index == two.step_index == 30


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("config", list(CONFIGS))
@pytest.mark.parametrize("case", range(len(CASES)))
def test_staged_two_step_sweep_is_bitwise_single_steps(dtype, config, case):
    """The staged sweep (each plane's inputs requested a plane ahead, into LDS) computes the
    same bits as single steps, with a boundary tape too."""
    shape, tile_j, chunk_x = CASES[case]
    one = _solver(shape, dtype, CONFIGS[config], 1, tile_j, chunk_x)
    two = _solver(shape, dtype, CONFIGS[config], 2, tile_j, chunk_x, staged=True)
    tapes = None
    if config == "all":
        tapes = [torch.full((30, s.wall_tape_stride()), -7.0, dtype=s.dtype, device=s.device)
                 for s in (one, two)]
        one.attach_wall_tape(tapes[0], 0)
        two.attach_wall_tape(tapes[1], 0)
    for n_steps in (8, 5, 4, 7, 6):
        one.run(n_steps)
        two.run(n_steps)
    _assert_same(one, two)
    if tapes is not None:
        assert torch.equal(tapes[0], tapes[1])


def test_two_step_compensated_fp32():
    shape, tile_j, chunk_x = CASES[3]
    one = _solver(shape, np.float32, CONFIGS["all"], 1, tile_j, chunk_x, compensated=True)
    two = _solver(shape, np.float32, CONFIGS["all"], 2, tile_j, chunk_x, compensated=True)
    one.run(13)
    two.run(13)
    _assert_same(one, two)


def test_two_step_narrow_lanes_and_the_tape():
    """nz not a multiple of 4 runs two steps per sweep on narrower lanes; with a boundary tape
    attached the two-step sweep writes the same tape rows as single steps."""
    shape = (10, 9, 10)  # fp32 lanes of 2 cells
    one = _solver(shape, np.float32, CONFIGS["all"], 1, 0, 0)
    two = _solver(shape, np.float32, CONFIGS["all"], 2, 0, 0)
    one.run(9)
    two.run(9)
    _assert_same(one, two)
    shape = (10, 9, 12)
    one = _solver(shape, np.float32, CONFIGS["all"], 1, 0, 0)
    two = _solver(shape, np.float32, CONFIGS["all"], 2, 0, 0)
    tapes = [torch.zeros((9, s.wall_tape_stride()), dtype=s.dtype, device=s.device) for s in (one, two)]
    one.attach_wall_tape(tapes[0], 0)
    two.attach_wall_tape(tapes[1], 0)
    one.run(9)
    two.run(9)
    _assert_same(one, two)
    assert torch.equal(tapes[0], tapes[1])


def test_more_sources_than_the_sweep_holds_run_single_steps():
    """The two-step sweep keeps the source and receiver tables in LDS (64 and 256 entries);
    beyond them the solver runs single steps, with the same results."""
    shape, tile_j, chunk_x = CASES[5]
    one = _solver(shape, np.float32, CONFIGS["all"], 1, tile_j, chunk_x)
    two = _solver(shape, np.float32, CONFIGS["all"], 2, tile_j, chunk_x)
    assert two.two_step_applies()
    rng = np.random.default_rng(5)
    lengths = np.array(one.grid.lengths)
    for _ in range(63):
        where, signal = tuple(rng.uniform(0.05, 0.95, 3) * lengths), rng.standard_normal(12)
        one.add_volume_source(where, signal)
        two.add_volume_source(where, signal)
    assert not two.two_step_applies()
    one.run(11)
    two.run(11)
    _assert_same(one, two)


@pytest.mark.parametrize("pad", [0, 1], ids=["aligned", "unaligned"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_two_step_boundary_tape_is_bitwise_single_steps(dtype, case, pad):
    """Owner lanes write the tape rows of both steps of a sweep; halo lanes and halo planes
    write nothing. With the library's stride the x and y face rows go out as vector stores;
    one element more makes rows unaligned and takes the per-cell stores."""
    shape, tile_j, chunk_x = CASES[case]
    one = _solver(shape, dtype, CONFIGS["all"], 1, tile_j, chunk_x)
    two = _solver(shape, dtype, CONFIGS["all"], 2, tile_j, chunk_x)
    tapes = [torch.full((17, s.wall_tape_stride() + pad), -7.0, dtype=s.dtype, device=s.device)
             for s in (one, two)]
    one.attach_wall_tape(tapes[0], 0)
    two.attach_wall_tape(tapes[1], 0)
    one.run(17)
    two.run(17)
    _assert_same(one, two)
    assert torch.equal(tapes[0], tapes[1])
    nx, ny, nz = shape
    stride = 2 * (ny * nz + nx * nz + nx * ny)  # the row's written length; the stride pads it
    assert one.wall_tape_stride() >= stride and one.wall_tape_stride() % 4 == 0
    assert not torch.any(tapes[1][:, :stride] == -7.0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("config", ["walls", "all"])
def test_balanced_chunks_are_bitwise_single_steps(dtype, config):
    """A single room whose two-step tiles overrun a whole round of workgroups runs in balanced x
    chunks (row blocks in nchunk or nchunk + 1 chunks of 2 to 3 planes here, chunk edges at
    uneven planes): the same bits as single steps, boundary tape included. 62 row blocks of 2
    rows (one wavefront per workgroup) in 34 chunks of 6 planes are 2108 tiles, more than the
    8 one-wavefront workgroups per CU that run at once."""
    shape, tile_j, chunk_x = (200, 124, 32), 1, 6
    units = torch.cuda.get_device_properties(0).multi_processor_count * 8
    tiles = -(-shape[1] // 2) * -(-shape[0] // chunk_x)
    assert tiles > units, "the case must overrun one round to exercise the balanced chunks"
    one = _solver(shape, dtype, CONFIGS[config], 1, tile_j, chunk_x)
    two = _solver(shape, dtype, CONFIGS[config], 2, tile_j, chunk_x)
    tapes = None
    if config == "all":
        tapes = [torch.full((30, s.wall_tape_stride()), -7.0, dtype=s.dtype, device=s.device)
                 for s in (one, two)]
        one.attach_wall_tape(tapes[0], 0)
        two.attach_wall_tape(tapes[1], 0)
    for n_steps in (8, 5, 4, 7, 6):
        one.run(n_steps)
        two.run(n_steps)
    _assert_same(one, two)
    if tapes is not None:
        assert torch.equal(tapes[0], tapes[1])
