"""Frequency-dependent (IIR) walls: the wall model and its fit, the oracle's pins, and the HIP
wall update against the oracle.

The reference plans these walls (``PLAN.md:21-24``) without implementing them, so the oracle
(``oracle/impedance.py``) is pinned by what the physics fixes: every fitted wall is positive
real (passive) in continuous and discrete time; the wall's filter recursion equals an
independent IIR filter (``scipy.signal.lfilter``) on the same mean pressures; the acoustic
energy never exceeds its start, and a planted sign error on the wall velocity breaks that;
and a pulse down a duct reflects with ``|(1 - xi) / (1 + xi)|`` of the discrete admittance,
band by band (the reference's normal-incidence test, ``tests/test_boundaries.py:113-144``,
for a frequency-dependent wall).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from cases import PARITY_RECEIVER, PARITY_SOURCE, parity_signal

from ac_fdtd_amd import AirAbsorption, Grid
from ac_fdtd_amd.impedance import FrequencyDependentWalls, fit_wall
from oracle.impedance import OracleImpedanceFDTD

TABLES = [
    {125: 0.1, 250: 0.2, 500: 0.45, 1000: 0.7, 2000: 0.8, 4000: 0.85},  # porous absorber
    {125: 0.6, 250: 0.3, 500: 0.15, 1000: 0.1, 2000: 0.1, 4000: 0.1},  # panel absorber
    {125: 0.02, 250: 0.03, 500: 0.03, 1000: 0.04, 2000: 0.05, 4000: 0.07},  # plaster
]
DT = 0.01 / (343.0 * math.sqrt(3.0))


@pytest.fixture(scope="module")
def porous():
    return fit_wall(TABLES[0])


@pytest.mark.parametrize("table", range(len(TABLES)))
def test_fit_meets_the_octave_table_and_stays_passive(table):
    wall = fit_wall(TABLES[table])
    centres = np.array(sorted(TABLES[table]))
    np.testing.assert_allclose(wall.absorption(centres), [TABLES[table][c] for c in centres],
                               atol=0.03)
    frequencies = np.geomspace(1.0, 0.49 / DT, 4000)
    assert np.all(wall.admittance(frequencies).real >= 0.0)
    assert np.all(wall.discrete_admittance(frequencies, DT).real >= -1e-12)
    # The bilinear transform only warps frequency: close to the continuous wall in the audio band
    # (frequency warping grows as (2 pi f dt)^2 / 12: 0.05 % at 4 kHz on a 1 cm grid).
    audio = np.geomspace(50.0, 4000.0, 200)
    exact = wall.admittance(audio)
    assert np.all(np.abs(wall.discrete_admittance(audio, DT) - exact) <= 0.03 * np.abs(exact) + 0.01)


def test_wall_recursion_is_an_iir_filter_of_the_mean_pressure(porous):
    """Record the mean pressure x of one wall cell and the velocity the update writes; the
    memory written after step n plus the instantaneous part of step n + 1 equals the sum of the
    branch biquads run by scipy.signal.lfilter on x."""
    from scipy.signal import lfilter

    grid = Grid(dx=0.03, nx=12, ny=10, nz=9)
    solver = OracleImpedanceFDTD(grid, walls=FrequencyDependentWalls(x_max=porous))
    solver.p[:] = np.random.default_rng(0).standard_normal(grid.shape)
    _, _, _, where, _ = solver.faces[0]
    cell = 17
    x, written = [], []
    rho_c = solver.medium.density * solver.medium.sound_speed
    for _ in range(120):
        solver.step()
        x.append(float((solver.wall_fu[where] * solver.boundary_sum[where])[cell] * 0.5))
        written.append(float(solver.velocity[0][-1].ravel()[cell]) * rho_c)
    x = np.array(x)
    table = porous.branch_count() if hasattr(porous, "branch_count") else len(porous.branches)
    assert table == len(porous.branches)
    total = np.zeros_like(x)
    instantaneous = 0.0
    for branch in porous.branches:
        b0, b2, a1, a2 = branch.biquad(solver.dt)
        total += lfilter([b0, 0.0, b2], [1.0, a1, a2], x)
        instantaneous += b0
    np.testing.assert_allclose(np.array(written[:-1]) + instantaneous * x[1:], total[1:],
                               rtol=1e-9, atol=1e-12 * np.max(np.abs(total)))


def _resonant_wall():
    """One lightly damped branch (500 Hz, Q = 5): a wall whose memory carries most of its
    response, so a wrong sign on that memory feeds energy in."""
    from ac_fdtd_amd.impedance import FrequencyDependentWall, ImpedanceBranch

    omega, resistance, q = 2.0 * math.pi * 500.0, 0.2, 5.0
    return FrequencyDependentWall(0.0, (ImpedanceBranch(resistance * q / omega, resistance,
                                                        resistance * q * omega),))


@pytest.mark.parametrize("case", ["fitted", "resonant", "resonant-planted-sign-error"])
def test_energy_never_exceeds_its_start(case):
    grid = Grid(dx=0.03, nx=16, ny=13, nz=11)
    if case == "fitted":
        walls = FrequencyDependentWalls(TABLES[0], 0.3, TABLES[1], 0.0, TABLES[2], 0.1)
    else:
        walls = FrequencyDependentWalls.uniform(_resonant_wall())
    sign = -1.0 if "planted" in case else 1.0
    solver = OracleImpedanceFDTD(grid, walls=walls, air_absorption=AirAbsorption(),
                                 wall_sign=sign)
    solver.p[:] = np.random.default_rng(1).standard_normal(grid.shape)
    start = solver.acoustic_energy()
    worst = 0.0
    for _ in range(300):
        solver.step()
        worst = max(worst, solver.acoustic_energy() / start)
    if sign == 1.0:
        assert worst <= 1.0 + 1e-12
    else:
        assert worst > 10.0


def _duct_reflection(wall_model):
    grid = Grid(dx=0.01, nx=600, ny=1, nz=1)

    def run(wall):
        solver = OracleImpedanceFDTD(grid, walls=FrequencyDependentWalls(x_min=1.0, x_max=wall))
        index = np.arange(2600)
        solver.add_volume_source((1.0, 0.005, 0.005), np.exp(-(((index - 120.0) / 30.0) ** 2)))
        solver.add_pressure_receiver((2.0, 0.005, 0.005))
        solver.run(2600)
        return solver.recorded_pressure[0], solver.dt

    rigid, dt = run(0.0)
    measured, _ = run(wall_model)
    window = slice(1300, 2600)
    frequencies = np.fft.rfftfreq(1300, dt)
    ratio = np.abs(np.fft.rfft(measured[window]) / np.fft.rfft(rigid[window]))
    return frequencies, ratio, dt


def test_normal_incidence_reflection_follows_the_discrete_admittance(porous):
    """Reflected spectrum over the rigid wall's (|R| = 1 exactly in the scheme), per octave
    band, against |(1 - xi_d) / (1 + xi_d)|: within 0.02 where the grid resolves the band
    (at least 50 points per wavelength)."""
    frequencies, ratio, dt = _duct_reflection(porous)
    xi = porous.discrete_admittance(frequencies, dt)
    theory = np.abs((1.0 - xi) / (1.0 + xi))
    for centre in (125.0, 250.0, 500.0):
        band = (frequencies > centre / math.sqrt(2.0)) & (frequencies < centre * math.sqrt(2.0))
        assert np.mean(ratio[band]) == pytest.approx(np.mean(theory[band]), abs=0.02), centre
    # The wall absorbs much more at 500 Hz than at 125 Hz, as its table says.
    low = (frequencies > 100.0) & (frequencies < 150.0)
    high = (frequencies > 400.0) & (frequencies < 600.0)
    assert np.mean(ratio[high]) < np.mean(ratio[low]) - 0.1


# ------------------------------------------------------------------------------------ GPU --


def _prepare(solver, grid, seed=0):
    rng = np.random.default_rng(seed)
    initial = rng.standard_normal(grid.shape)
    if hasattr(solver.p, "copy_"):
        import torch

        solver.p.copy_(torch.as_tensor(initial))
    else:
        solver.p[:] = initial
    solver.add_volume_source(PARITY_SOURCE, parity_signal())
    solver.add_pressure_receiver(PARITY_RECEIVER)
    solver.add_pressure_receiver((0.05, 0.05, 0.05))
    return solver


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
@pytest.mark.parametrize("air", [False, True], ids=["no-air", "air"])
@pytest.mark.parametrize("nz", [16, 15])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "two-sweep"])
def test_hip_iir_walls_are_bitwise_the_oracle(dtype, nz, air, fused):
    from ac_fdtd_amd.iir_walls import IIRWallFDTD

    grid = Grid(dx=0.03, nx=20, ny=17, nz=nz)
    walls = FrequencyDependentWalls(TABLES[0], 0.3, TABLES[1], 0.0, TABLES[2], TABLES[0])
    extra = {"air_absorption": AirAbsorption(18.0, 40.0)} if air else {}
    oracle = _prepare(OracleImpedanceFDTD(grid, dtype=dtype, walls=walls, **extra), grid)
    solver = _prepare(IIRWallFDTD(grid, dtype=dtype, walls=walls, fused=fused, **extra), grid)
    for n in (23, 40):
        oracle.run(n)
        solver.run(n)
    np.testing.assert_array_equal(solver.p.cpu().numpy(), oracle.p)
    for a in range(3):
        np.testing.assert_array_equal(solver.velocity[a].cpu().numpy(), oracle.velocity[a])
    np.testing.assert_array_equal(solver.recorded_pressure, oracle.recorded_pressure)
    assert solver.acoustic_energy() == pytest.approx(oracle.acoustic_energy(), rel=1e-10)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_hip_iir_walls_stay_passive_at_96():
    import torch

    from ac_fdtd_amd.iir_walls import IIRWallFDTD

    grid = Grid(dx=0.03, nx=96, ny=96, nz=96)
    solver = IIRWallFDTD(grid, walls=FrequencyDependentWalls.uniform(TABLES[1]))
    solver.p.copy_(torch.as_tensor(np.random.default_rng(3).standard_normal(grid.shape)))
    start = solver.acoustic_energy()
    energies = []
    for _ in range(10):
        solver.run(100)
        energies.append(solver.acoustic_energy() / start)
    assert max(energies) <= 1.0 + 1e-12
    assert energies[-1] < 0.5
