"""The fourth-order-in-space staggered scheme: oracle pins on the CPU, the HIP sweeps against
the oracle on the GPU.

The reference lists this variant (``PLAN.md:17``, ``:116``) and has no implementation, so the
oracle (``oracle/fourth.py``) is pinned by: reducing bit for bit to the first-order oracle (the
reference's scheme) with weights (1, 0), sponge, air, sources and receivers included; energy
conservation in a rigid box; exact single modes at the closed-form fourth-order frequency
(which a wrong weight or wall image breaks); and the fourth-order convergence of that frequency.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from cases import PARITY, PARITY_RECEIVER, PARITY_SOURCE, parity_signal

from ac_fdtd_amd import AIR, AbsorbingLayer, AirAbsorption, Grid, mode_frequency
from ac_fdtd_amd.analytic import (
    fourth_order_mode_frequency,
    fourth_order_mode_initial_state,
    mode_pressure_field,
)
from oracle.fourth import OracleFourthFDTD
from oracle.scheme import OracleFDTD

CONFIGS = {
    "rigid": {},
    "layer": {"absorbing_layer": AbsorbingLayer(thickness=3)},
    "air": {"air_absorption": AirAbsorption(18.0, 40.0)},
    "layer-air": {"absorbing_layer": PARITY["all"]["absorbing_layer"],
                  "air_absorption": PARITY["all"]["air_absorption"]},
}


def _prepare(solver, grid, seed=0):
    rng = np.random.default_rng(seed)
    initial = rng.standard_normal(grid.shape)
    velocity = [1e-3 * rng.standard_normal(grid.velocity_shape(a)) for a in range(3)]
    if hasattr(solver.p, "copy_"):
        import torch

        solver.p.copy_(torch.as_tensor(initial))
        for a in range(3):
            solver.velocity[a].copy_(torch.as_tensor(velocity[a]))
    else:
        solver.p[:] = initial
        for a in range(3):
            solver.velocity[a][:] = velocity[a]
    solver.add_volume_source(PARITY_SOURCE, parity_signal())
    solver.add_volume_source(PARITY_SOURCE, parity_signal()[::3])
    solver.add_pressure_receiver(PARITY_RECEIVER)
    solver.add_pressure_receiver(PARITY_SOURCE)
    return solver


@pytest.mark.parametrize("name", list(CONFIGS))
def test_weights_one_zero_are_bitwise_the_first_order_oracle(name):
    grid = Grid(dx=0.03, nx=20, ny=17, nz=15)
    first = _prepare(OracleFDTD(grid, **CONFIGS[name]), grid)
    fourth = _prepare(OracleFourthFDTD(grid, coefficients=(1.0, 0.0), **CONFIGS[name]), grid)
    first.run(60)
    fourth.run(60)
    np.testing.assert_array_equal(fourth.p, first.p)
    for a in range(3):
        np.testing.assert_array_equal(fourth.velocity[a], first.velocity[a])
    np.testing.assert_array_equal(fourth.recorded_pressure, first.recorded_pressure)
    assert fourth.energy() == first.energy()


@pytest.mark.parametrize("courant", [1.0, 0.6])
def test_oracle_conserves_energy_in_a_rigid_box(courant):
    grid = Grid(dx=0.03, nx=14, ny=11, nz=9)
    solver = OracleFourthFDTD(grid, courant=courant)
    solver.p[:] = np.random.default_rng(1).standard_normal(grid.shape)
    initial = solver.energy()
    drift = 0.0
    for _ in range(30):
        solver.run(20)
        drift = max(drift, abs(solver.energy() / initial - 1.0))
    assert drift < 1e-11


def _mode_error(solver_cls, grid, mode, steps, **kwargs):
    solver = solver_cls(grid, **kwargs)
    frequency = fourth_order_mode_frequency(grid, mode, AIR, solver.dt)
    pressure, velocity = fourth_order_mode_initial_state(grid, mode, AIR, solver.dt)
    shape = mode_pressure_field(grid, mode)
    if hasattr(solver.p, "copy_"):
        import torch

        solver.p.copy_(torch.as_tensor(pressure))
        for a in range(3):
            solver.velocity[a].copy_(torch.as_tensor(velocity[a]))
        expected_shape = torch.as_tensor(shape, device=solver.device)
    else:
        solver.p[:] = pressure
        for a in range(3):
            solver.velocity[a][:] = velocity[a]
        expected_shape = shape
    worst = 0.0
    chunk = 50
    for _ in range(steps // chunk):
        solver.run(chunk)
        expected = math.cos(2 * math.pi * frequency * solver.time) * expected_shape
        difference = solver.p - expected
        worst = max(worst, float(abs(difference).max()))
    return worst


@pytest.mark.parametrize("mode", [(1, 0, 0), (0, 2, 1), (3, 2, 1), (5, 4, 3)])
def test_oracle_single_mode_at_the_fourth_order_frequency(mode):
    grid = Grid(dx=0.04, nx=12, ny=10, nz=9)
    assert _mode_error(OracleFourthFDTD, grid, mode, 600) < 1e-11


def test_a_wrong_weight_breaks_the_mode():
    """A planted mistake: the oracle run with c2 of the wrong sign drifts off the mode."""
    grid = Grid(dx=0.04, nx=12, ny=10, nz=9)
    solver = OracleFourthFDTD(grid, coefficients=(9.0 / 8.0, -1.0 / 24.0), courant=0.8)
    frequency = fourth_order_mode_frequency(grid, (3, 2, 1), AIR, solver.dt)
    pressure, velocity = fourth_order_mode_initial_state(grid, (3, 2, 1), AIR, solver.dt)
    solver.p[:] = pressure
    for a in range(3):
        solver.velocity[a][:] = velocity[a]
    solver.run(200)
    expected = math.cos(2 * math.pi * frequency * solver.time) * pressure
    assert np.max(np.abs(solver.p - expected)) > 1e-3


def test_the_spatial_error_falls_at_fourth_order():
    """With the time step made negligible (Courant 0.002), a mode's frequency error against
    the continuum falls 16-fold per halving of dx (the second-order stencil: 4-fold)."""
    mode, lengths = (2, 1, 1), (1.0, 0.8, 0.6)
    exact = mode_frequency(lengths, mode, AIR.sound_speed)
    errors, errors2 = [], []
    for dx in (0.1, 0.05, 0.025):
        grid = Grid.from_lengths(lengths, dx)
        dt = 0.002 * dx / (AIR.sound_speed * math.sqrt(3.0))
        errors.append(abs(fourth_order_mode_frequency(grid, mode, AIR, dt) / exact - 1.0))
        errors2.append(abs(fourth_order_mode_frequency(grid, mode, AIR, dt, (1.0, 0.0)) / exact - 1.0))
    for coarse, fine in zip(errors, errors[1:]):
        assert 14.0 < coarse / fine < 17.0
    for coarse, fine in zip(errors2, errors2[1:]):
        assert 3.5 < coarse / fine < 4.5


def test_admittance_walls_are_refused_on_the_host():
    from ac_fdtd_amd import WallAdmittances
    from ac_fdtd_amd.fourth import FourthOrderFDTD

    with pytest.raises(ValueError, match="rigid walls only"):
        FourthOrderFDTD(Grid(dx=0.1, nx=3, ny=3, nz=3), walls=WallAdmittances.uniform(0.2),
                        device="cuda")


# ------------------------------------------------------------------------------------ GPU --


def _hip(grid, **kwargs):
    from ac_fdtd_amd.fourth import FourthOrderFDTD

    return FourthOrderFDTD(grid, **kwargs)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
@pytest.mark.parametrize("shape,tile_j,chunk_x", [
    ((20, 17, 15), 0, 0), ((13, 40, 64), 2, 3), ((9, 6, 130), 1, 1), ((3, 2, 70), 4, 2),
    ((1, 5, 9), 0, 0),
])
@pytest.mark.parametrize("name", list(CONFIGS))
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
def test_hip_fourth_order_is_bitwise_the_oracle(dtype, name, shape, tile_j, chunk_x):
    grid = Grid(dx=0.03, nx=shape[0], ny=shape[1], nz=shape[2])
    oracle = _prepare(OracleFourthFDTD(grid, dtype=dtype, **CONFIGS[name]), grid, 4)
    solver = _prepare(_hip(grid, dtype=dtype, tile_j=tile_j, chunk_x=chunk_x, **CONFIGS[name]),
                      grid, 4)
    assert solver.dt == oracle.dt
    for n in (17, 12):
        oracle.run(n)
        solver.run(n)
    np.testing.assert_array_equal(solver.p.cpu().numpy(), oracle.p)
    for a in range(3):
        np.testing.assert_array_equal(solver.velocity[a].cpu().numpy(), oracle.velocity[a])
    for n, state in enumerate(oracle.relaxation_states):
        np.testing.assert_array_equal(solver.relaxation_states[n].cpu().numpy(), state)
    np.testing.assert_array_equal(solver.recorded_pressure, oracle.recorded_pressure)
    assert solver.energy() == pytest.approx(oracle.energy(), rel=1e-12)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_hip_fourth_order_mode_and_energy_at_128():
    grid = Grid(dx=0.04, nx=128, ny=128, nz=128)
    assert _mode_error(_hip, grid, (3, 2, 5), 600) < 1e-11
    import torch

    solver = _hip(grid)
    solver.p.copy_(torch.as_tensor(np.random.default_rng(2).standard_normal(grid.shape)))
    initial = solver.energy()
    drift = 0.0
    for _ in range(8):
        solver.run(250)
        drift = max(drift, abs(solver.energy() / initial - 1.0))
    assert drift < 1e-11
