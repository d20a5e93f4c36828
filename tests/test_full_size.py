"""The reference's physics checks and the workloads' accuracy targets at full size, on the GPU.

* Rigid box at 256^3 in fp64, 4000 steps from a seeded random pressure: the discrete energy
  (reduced in fp64 on the device) stays within 1e-11 of its start, the reference's tolerance
  (``tests/test_scheme.py:28-42``), through the two-step sweep the bench times.
* An exact single mode of the rigid box at 128^3 stays ``cos(2 pi f_d t) Phi`` to 1e-11 at the
  discrete modal frequency (``tests/test_scheme.py:75-98``, closed form ``analytic.py:66-125``).
* The throughput workload (256^3, walls 0.15, 16-cell sponge, air 20 C / 50 %): fp32 and
  compensated fp32 within 1e-5 receiver relative L2 of fp64 over 2000 steps.
* The batched workload: the 64 seeded rooms (sides U[3, 10] m at 1 kHz) advanced together in
  fp32, each within 1e-5 of its own fp64 single-room run after 1000 steps (every receiver has
  the direct sound by then).
* The large room: 512^3 with the sponge on two faces, 32 receivers, fp32 within 1e-5 of fp64
  over 1500 steps.
* The reference's normal-incidence reflection and ISO duct attenuation checks, on the HIP
  kernel (the CPU oracle passes them in ``tests/test_oracle.py``).

The fp64 single-room kernel these compare against is bit-for-bit the oracle
(``tests/test_gpu_sweeps.py``, ``tests/test_oracle_variants.py``).
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import bench
from ac_fdtd_amd import AIR, Grid, discrete_mode_frequency, mode_initial_state
from ac_fdtd_amd.analytic import mode_pressure_field

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]


def _relative_l2(test, reference):
    error = float(np.linalg.norm(test - reference) / np.linalg.norm(reference))
    assert math.isfinite(error)
    return error


def test_rigid_box_256_fp64_conserves_energy_over_4000_steps():
    drift = bench.rigid_energy_drift(256, 4000, np.float64)
    print(f"256^3 fp64 rigid box, 4000 steps: max |E/E0 - 1| = {drift:.3e}", flush=True)
    assert drift < 1e-11


@pytest.mark.parametrize("mode", [(1, 0, 0), (0, 0, 3), (2, 1, 1), (5, 3, 2)])
def test_single_mode_at_128_stays_exactly_sinusoidal(mode):
    from ac_fdtd_amd.scheme import AcousticFDTD

    grid = Grid(dx=0.04, nx=128, ny=128, nz=128)
    solver = AcousticFDTD(grid)
    assert solver.two_step_applies()
    frequency = discrete_mode_frequency(grid, mode, AIR, solver.dt)
    pressure, velocity = mode_initial_state(grid, mode, AIR, solver.dt)
    solver.p.copy_(torch.as_tensor(pressure))
    for axis in range(3):
        solver.velocity[axis].copy_(torch.as_tensor(velocity[axis]))
    shape = torch.as_tensor(mode_pressure_field(grid, mode), device=solver.device)
    worst = 0.0
    for _ in range(12):
        solver.run(50)
        expected = math.cos(2.0 * math.pi * frequency * solver.time) * shape
        worst = max(worst, float(torch.max(torch.abs(solver.p - expected))))
    print(f"mode {mode} at 128^3, 600 steps: max |p - p_exact| = {worst:.3e}", flush=True)
    assert worst < 1e-11


def _throughput_recording(dtype, steps, **kwargs):
    from ac_fdtd_amd.scheme import AcousticFDTD

    solver = bench.make_solver(AcousticFDTD, 256, steps, dtype, **kwargs)
    solver.run(steps)
    recorded = solver.recorded_pressure.astype(np.float64)
    del solver
    torch.cuda.empty_cache()
    return recorded


def test_throughput_workload_fp32_within_1e5_of_fp64():
    steps = 2000
    reference = _throughput_recording(np.float64, steps)
    for compensated in (False, True):
        single = _throughput_recording(np.float32, steps, compensated=compensated)
        error = _relative_l2(single, reference)
        print(f"256^3 walls+sponge+air, {steps} steps, fp32 (compensated={compensated}): "
              f"receiver relative L2 {error:.3e}", flush=True)
        assert error <= 1e-5


def test_batched_64_rooms_within_1e5_of_fp64_single_rooms():
    result = bench.batched_rooms(64, 1000, 0, check=True)
    print(f"64 batched rooms: {result}", flush=True)
    assert result["rooms"] == 64
    assert math.isfinite(result["receiver_rel_l2_vs_fp64_max"])
    assert result["receiver_rel_l2_vs_fp64_max"] <= 1e-5


def test_large_room_512_fp32_within_1e5_of_fp64():
    from ac_fdtd_amd.scheme import AcousticFDTD

    steps = 1500
    grid, config, source, receivers, signal, _ = bench.large_room_workload(512, steps)
    recordings = {}
    for dtype in (np.float32, np.float64):
        solver = AcousticFDTD(grid, dtype=dtype, **config)
        solver.add_volume_source(source, signal)
        for point in receivers:
            solver.add_pressure_receiver(point)
        solver.run(steps)
        recordings[dtype] = solver.recorded_pressure.astype(np.float64)
        del solver
        torch.cuda.empty_cache()
    error = _relative_l2(recordings[np.float32], recordings[np.float64])
    print(f"512^3 large room, {steps} steps, 32 receivers: fp32 relative L2 {error:.3e}",
          flush=True)
    assert error <= 1e-5


def _hip_factory(grid, **kwargs):
    from ac_fdtd_amd.scheme import AcousticFDTD

    return AcousticFDTD(grid, **kwargs)


@pytest.mark.parametrize(("xi", "expected"), [(0.0, 1.0), (0.25, 0.6), (0.5, 1 / 3), (1.0, 0.0)])
def test_normal_incidence_reflection_on_the_hip_kernel(xi, expected):
    """The reference's reflection check (``tests/test_boundaries.py:113-144``) on the HIP time
    loop: a pulse down a one-cell duct off a wall of admittance xi reflects (1 - xi)/(1 + xi)
    within 0.03."""
    from test_oracle import _reflection

    assert _reflection(_hip_factory, xi) == pytest.approx(expected, abs=0.03)


@pytest.mark.parametrize("frequency", [2000.0, 4000.0])
def test_iso_duct_attenuation_on_the_hip_kernel(frequency):
    """The reference's ISO 9613-1 check (``tests/test_air.py:91-124``) on the HIP time loop: a
    pulse down a 3000-cell duct, attenuation between two stations within 5 % of the standard."""
    from test_oracle import _iso_duct

    measured, expected = _iso_duct(_hip_factory, frequency)
    assert measured == pytest.approx(expected, rel=0.05)
