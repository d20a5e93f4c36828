"""The pressure-only (second-order wave equation) scheme: oracle pins on the CPU, the HIP kernel
against the oracle on the GPU.

The reference has no implementation of this variant (``PLAN.md:52-56``, ``:116``), so its
oracle (``oracle/wave.py``) is pinned by what the mathematics fixes: it reproduces the
first-order oracle (itself bit-for-bit the reference) with walls and sources, because it is that
scheme with the velocity eliminated; an exact single mode of the rigid box oscillates at the
first-order scheme's discrete modal frequency; its discrete invariant is conserved in a rigid
box and decreases with absorbing walls; and a planted sign error fails those checks.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from cases import PARITY_RECEIVER, PARITY_SOURCE, parity_signal

from ac_fdtd_amd import AIR, Grid, WallAdmittances, discrete_mode_frequency, mode_initial_state
from ac_fdtd_amd.analytic import mode_pressure_field
from oracle import wave as wave_oracle
from oracle.scheme import OracleFDTD
from oracle.wave import OracleWaveFDTD

WALLS = WallAdmittances(0.3, 0.1, 0.0, 0.5, 0.2, 0.05)


def _handover(grid, walls, steps, dtype=np.float64, seed=0):
    """A first-order oracle run and the pressure-only oracle started from its p^1, p^0."""
    rng = np.random.default_rng(seed)
    p0 = rng.standard_normal(grid.shape)
    first = OracleFDTD(grid, walls=walls, dtype=dtype)
    first.p[:] = p0
    second = OracleWaveFDTD(grid, walls=walls, dtype=dtype)
    for solver in (first, second):
        solver.add_volume_source(PARITY_SOURCE, parity_signal())
        solver.add_pressure_receiver(PARITY_RECEIVER)
    first.run(1)
    second.p[:] = first.p
    second.p_previous[:] = p0
    second.step_index = 1
    worst = 0.0
    for _ in range(steps):
        first.step()
        second.step()
        worst = max(worst, float(np.max(np.abs(first.p - second.p)) / np.max(np.abs(first.p))))
    return worst


@pytest.mark.parametrize("walls", [WallAdmittances(), WALLS], ids=["rigid", "walls"])
def test_oracle_is_the_first_order_scheme_without_velocity(walls):
    grid = Grid(dx=0.03, nx=20, ny=17, nz=15)
    assert _handover(grid, walls, 300) < 1e-11


@pytest.mark.parametrize("mode", [(1, 0, 0), (2, 1, 1), (3, 2, 1)])
def test_oracle_single_mode_at_the_discrete_frequency(mode):
    grid = Grid(dx=0.04, nx=12, ny=10, nz=9)
    solver = OracleWaveFDTD(grid)
    frequency = discrete_mode_frequency(grid, mode, AIR, solver.dt)
    shape = mode_pressure_field(grid, mode)
    solver.p[:] = shape
    solver.p_previous[:] = math.cos(2 * math.pi * frequency * solver.dt) * shape
    worst = 0.0
    for _ in range(600):
        solver.step()
        expected = math.cos(2 * math.pi * frequency * solver.time) * shape
        worst = max(worst, float(np.max(np.abs(solver.p - expected))))
    assert worst < 1e-11


def test_oracle_invariant_is_conserved_and_walls_only_remove_it():
    grid = Grid(dx=0.03, nx=14, ny=12, nz=11)
    rng = np.random.default_rng(2)
    for walls, rigid in ((WallAdmittances(), True), (WALLS, False)):
        solver = OracleWaveFDTD(grid, walls=walls)
        solver.p[:] = rng.standard_normal(grid.shape)
        solver.p_previous[:] = solver.p
        values = []
        for _ in range(300):
            solver.step()
            values.append(solver.invariant())
        if rigid:
            assert (max(values) - min(values)) / values[0] < 1e-12
        else:
            assert np.all(np.diff(values) <= 1e-12 * values[0])
            assert values[-1] < 0.5 * values[0]


def test_a_planted_sign_error_fails_the_pins(monkeypatch):
    """The Laplacian's sign flipped in the oracle: the hand-over drifts at once."""
    original = wave_oracle.laplacian
    monkeypatch.setattr(wave_oracle, "laplacian", lambda p: -original(p))
    grid = Grid(dx=0.03, nx=20, ny=17, nz=15)
    assert _handover(grid, WALLS, 20) > 1e-3


def test_the_wave_descriptor_is_checked_before_any_launch():
    from ac_fdtd_amd import _native
    from ac_fdtd_amd.build import LIBRARY

    if not LIBRARY.exists():
        pytest.skip("HIP library not built")
    lib = _native.library()

    def desc(**overrides):
        d = _native.WaveDesc()
        d.abi_version = _native.ABI_VERSION
        d.struct_size = ctypes.sizeof(_native.WaveDesc)
        d.nx = d.ny = d.nz = 4
        d.real_bytes = 8
        d.p, d.p_prev = 4096, 8192
        for key, value in overrides.items():
            setattr(d, key, value)
        return d

    for overrides, message in (
        ({"struct_size": 4}, "bytes"),
        ({"nx": 0}, "cell counts"),
        ({"real_bytes": 2}, "real_bytes"),
        ({"p_prev": 4096}, "distinct"),
        ({"p": 4100}, "aligned"),
        ({"wall_mask": 3}, "wall table"),
        ({"n_sources": 1}, "source arrays"),
        ({"tile_j": 9}, "tile_j"),
        ({"time_block": 3}, "time_block"),
        ({"time_block": 2, "p_swap": 12288}, "both swap buffers"),
        ({"time_block": 2, "p_swap": 12288, "p_prev_swap": 4096}, "distinct"),
        ({"time_block": 2, "p_swap": 12288, "p_prev_swap": 16388}, "aligned"),
    ):
        d = desc(**overrides)
        code = -2 if message == "bytes" else -1
        assert lib.acfdtd_hip_wave_run(ctypes.byref(d), 0, 1, 0, None) == code, message
        assert message in lib.acfdtd_hip_last_error().decode()
    recording = desc(n_receivers=1, receiver_cell=4096, recording=4096, recording_stride=3)
    assert lib.acfdtd_hip_wave_run(ctypes.byref(recording), 0, 4, 0, None) == -1
    assert lib.acfdtd_hip_wave_run(ctypes.byref(desc()), 0, 0, 0, None) == 0


# ------------------------------------------------------------------------------------ GPU --

def _pair(grid, dtype, walls, seed=3, **kwargs):
    import torch

    from ac_fdtd_amd.wave import PressureOnlyFDTD

    rng = np.random.default_rng(seed)
    now, before = rng.standard_normal(grid.shape), rng.standard_normal(grid.shape)
    oracle = OracleWaveFDTD(grid, walls=walls, dtype=dtype)
    oracle.p[:], oracle.p_previous[:] = now, before
    solver = PressureOnlyFDTD(grid, walls=walls, dtype=dtype, **kwargs)
    solver.set_state(torch.as_tensor(now), torch.as_tensor(before))
    lengths = np.array(grid.lengths)
    for s in (oracle, solver):
        s.add_volume_source(tuple(0.3 * lengths), parity_signal())
        s.add_volume_source(tuple(0.3 * lengths), parity_signal()[::2])
        s.add_volume_source(tuple(0.7 * lengths), parity_signal()[::5])
        for f in (0.11, 0.3, 0.83):
            s.add_pressure_receiver(tuple(f * lengths))
    return solver, oracle


def _assert_same(solver, oracle):
    np.testing.assert_array_equal(solver.p.cpu().numpy(), oracle.p)
    np.testing.assert_array_equal(solver.p_previous.cpu().numpy(), oracle.p_previous)
    np.testing.assert_array_equal(solver.recorded_pressure, oracle.recorded_pressure)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
@pytest.mark.parametrize("time_block", [1, 2])
@pytest.mark.parametrize("walls", [WallAdmittances(), WALLS], ids=["rigid", "walls"])
@pytest.mark.parametrize("nz", [16, 14, 15])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
def test_hip_wave_is_bitwise_the_oracle(dtype, nz, walls, time_block):
    """Odd and even step counts swap the fields' roles; with two steps per sweep, 37 steps are
    nine pairs of sweeps and one single step, 40 steps ten pairs."""
    grid = Grid(dx=0.03, nx=20, ny=17, nz=nz)
    solver, oracle = _pair(grid, dtype, walls, time_block=time_block)
    for n_steps in (37, 1, 40, 2, 3):
        solver.run(n_steps)
        oracle.run(n_steps)
        _assert_same(solver, oracle)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
@pytest.mark.parametrize("time_block", [1, 2])
@pytest.mark.parametrize("tile_j,chunk_x", [(1, 1), (2, 5), (8, 3), (4, 0)])
def test_hip_wave_tiles_do_not_change_the_answer(tile_j, chunk_x, time_block):
    """Several workgroups per plane and several marches along x, so workgroup edges fall inside
    the grid (rows of 64 lane groups of 4 cells: one step per sweep splits them between
    wavefronts; two steps per sweep own 2 or 4 whole rows with tile_j 1 or 2, and three row
    segments of 22 lane groups with tile_j 4 or 8)."""
    grid = Grid(dx=0.02, nx=23, ny=19, nz=256)
    solver, oracle = _pair(grid, np.float32, WALLS, seed=5, tile_j=tile_j, chunk_x=chunk_x,
                           time_block=time_block)
    solver.run(14)
    oracle.run(14)
    _assert_same(solver, oracle)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
@pytest.mark.parametrize("shape,dtype,tile_j,chunk_x", [
    ((10, 5, 130), np.float64, 8, 4),   # 65 lane groups of 2: three segments, 16 rows
    ((7, 11, 61), np.float32, 8, 3),    # 61 single cells: three segments
    ((9, 7, 12), np.float32, 0, 0),     # 3 groups of 4, all rows in one workgroup
    ((5, 1, 4), np.float32, 0, 0),      # one row of one group
    ((1, 6, 7), np.float64, 1, 1),      # one plane
    ((17, 40, 16), np.float64, 0, 5),   # 8 groups: 62 rows per workgroup, ragged last block
], ids=["seg-fp64", "seg-odd", "one-block", "one-row", "one-plane", "ragged"])
def test_hip_wave_two_step_sweep_shapes(shape, dtype, tile_j, chunk_x):
    """Two steps per sweep on whole rows and row segments, single rows and planes."""
    grid = Grid(dx=0.03, nx=shape[0], ny=shape[1], nz=shape[2])
    solver, oracle = _pair(grid, dtype, WALLS, seed=7, tile_j=tile_j, chunk_x=chunk_x,
                           time_block=2)
    for n_steps in (8, 6):
        solver.run(n_steps)
        oracle.run(n_steps)
    _assert_same(solver, oracle)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_hip_wave_reproduces_the_first_order_kernel():
    """Hand-over from the first-order HIP solver (fp64): the pressures agree to rounding."""
    import torch

    from ac_fdtd_amd.scheme import AcousticFDTD
    from ac_fdtd_amd.wave import PressureOnlyFDTD

    grid = Grid(dx=0.03, nx=40, ny=33, nz=28)
    first = AcousticFDTD(grid, walls=WALLS)
    first.p.copy_(torch.as_tensor(np.random.default_rng(1).standard_normal(grid.shape)))
    p0 = first.p.clone()
    first.run(1)
    second = PressureOnlyFDTD(grid, walls=WALLS)
    second.set_state(first.p, p0, 1)
    first.run(400)
    second.run(400)
    scale = float(torch.max(torch.abs(first.p)))
    assert float(torch.max(torch.abs(first.p - second.p))) < 1e-11 * scale


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_hip_wave_single_mode_at_128():
    import torch

    from ac_fdtd_amd.wave import PressureOnlyFDTD

    grid = Grid(dx=0.04, nx=128, ny=128, nz=128)
    mode = (3, 2, 5)
    solver = PressureOnlyFDTD(grid)
    frequency = discrete_mode_frequency(grid, mode, AIR, solver.dt)
    shape = torch.as_tensor(mode_pressure_field(grid, mode), device=solver.device)
    solver.set_state(shape, math.cos(2 * math.pi * frequency * solver.dt) * shape)
    energy = solver.energy()
    worst = 0.0
    for _ in range(12):
        solver.run(50)
        expected = math.cos(2 * math.pi * frequency * solver.time) * shape
        worst = max(worst, float(torch.max(torch.abs(solver.p - expected))))
    assert worst < 1e-11
    assert solver.energy() == pytest.approx(energy, rel=1e-12)
    # The first-order solver's exact mode has the same frequency: same closed form.
    p, _ = mode_initial_state(grid, mode, AIR, solver.dt)
    assert np.allclose(p, shape.cpu().numpy())


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_hip_wave_refuses_the_first_order_only_physics():
    from ac_fdtd_amd import AbsorbingLayer, AirAbsorption
    from ac_fdtd_amd.wave import PressureOnlyFDTD

    grid = Grid(dx=0.03, nx=8, ny=8, nz=8)
    for kwargs in ({"absorbing_layer": AbsorbingLayer(thickness=2)},
                   {"air_absorption": AirAbsorption()}):
        with pytest.raises(ValueError, match="pressure-only"):
            PressureOnlyFDTD(grid, **kwargs)
