"""Room driver and acoustics metrics (reference ``tests/test_room.py``).

CPU: the metrics on signals with known answers. GPU: the room driver end to end on the HIP time
loop, bit-for-bit against the same room on the CPU oracle, and its decay against Sabine/Eyring.
"""

from __future__ import annotations

import numpy as np
import pytest

from ac_fdtd_amd import admittance_from_absorption, random_incidence_absorption
from ac_fdtd_amd.metrics import (
    clarity,
    early_decay_time,
    eyring_time,
    octave_band,
    octave_band_reverberation,
    reverberation_time,
    sabine_time,
    schroeder_decay,
    schroeder_frequency,
)
from ac_fdtd_amd.room import Room, simulate_impulse_response

RATE = 20000.0


def _decaying_noise(t60: float, duration: float = 2.0, seed: int = 0) -> np.ndarray:
    t = np.arange(int(duration * RATE)) / RATE
    return np.random.default_rng(seed).standard_normal(t.size) * 10.0 ** (-3.0 * t / t60)


@pytest.mark.parametrize("t60", [0.3, 0.8, 1.6])
def test_reverberation_time_recovers_a_known_decay(t60):
    response = _decaying_noise(t60)
    assert reverberation_time(response, RATE, 20.0) == pytest.approx(t60, rel=0.02)
    assert reverberation_time(response, RATE, 30.0) == pytest.approx(t60, rel=0.02)


def test_early_decay_and_schroeder_curve():
    response = _decaying_noise(0.8)
    assert early_decay_time(response, RATE) == pytest.approx(0.8, rel=0.05)
    decay = schroeder_decay(_decaying_noise(0.5))
    assert decay[0] == pytest.approx(0.0)
    assert np.all(np.diff(decay) <= 1e-12)


def test_clarity_is_higher_for_a_drier_room():
    assert clarity(_decaying_noise(0.3), RATE) > clarity(_decaying_noise(1.5), RATE)


def test_octave_band_keeps_its_band_and_rejects_the_rest():
    t = np.arange(int(0.5 * RATE)) / RATE
    in_band, out_band = np.sin(2 * np.pi * 500.0 * t), np.sin(2 * np.pi * 4000.0 * t)
    kept = octave_band(in_band, RATE, 500.0)[1000:-1000]
    rejected = octave_band(out_band, RATE, 500.0)[1000:-1000]
    assert np.std(kept) == pytest.approx(np.std(in_band), rel=0.05)
    assert np.std(rejected) < 1e-4 * np.std(out_band)


def test_octave_band_reverberation_of_a_band_limited_decay():
    bands = octave_band_reverberation(_decaying_noise(0.6), RATE, (250.0, 1000.0, 4000.0))
    for value in bands.values():
        assert value == pytest.approx(0.6, rel=0.1)


def test_statistical_formulae():
    for absorption in (0.1, 0.4, 0.8):
        assert eyring_time(100.0, 130.0, absorption) < sabine_time(100.0, 130.0, absorption)
    assert eyring_time(100.0, 130.0, 0.999) < 0.05
    for normal in (0.05, 0.15, 0.3):
        assert 1.3 < random_incidence_absorption(admittance_from_absorption(normal)) / normal < 2.0
    assert 200.0 < schroeder_frequency(20.0, 0.3) < 400.0


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_a_small_room_decays_at_roughly_the_predicted_rate():
    """End to end on the GPU (reference test_room.py:146-161): T20 between 0.6 x Eyring and
    1.6 x Sabine."""
    room = Room(dimensions=(2.0, 1.7, 1.5), absorption=0.3, air=None)
    response = simulate_impulse_response(
        room, source=(0.6, 0.5, 0.4), receivers=[(1.4, 1.1, 1.0)], max_frequency=500.0,
        duration=4.0 * room.eyring_reverberation_time(),
    )
    assert np.all(np.isfinite(response.signals))
    measured = reverberation_time(response.recorded[0], response.sample_rate, 20.0)
    assert 0.6 * room.eyring_reverberation_time() < measured < 1.6 * room.sabine_reverberation_time()


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_room_driver_is_bitwise_the_oracle():
    """The same room through the GPU driver and through the CPU oracle: identical recordings."""
    from ac_fdtd_amd import AirAbsorption, Grid, WallAdmittances, band_limited_impulse
    from oracle.scheme import OracleFDTD

    room = Room(dimensions=(1.2, 0.9, 0.8), absorption=0.2, air=AirAbsorption(22.0, 40.0))
    response = simulate_impulse_response(room, (0.3, 0.3, 0.3), [(0.9, 0.6, 0.5), (0.2, 0.7, 0.6)],
                                         max_frequency=800.0, duration=0.02)
    grid = Grid.from_max_frequency(room.dimensions, 800.0, room.medium.sound_speed)
    oracle = OracleFDTD(grid, walls=WallAdmittances.from_absorption(0.2), air_absorption=room.air)
    n_steps = response.recorded.shape[1]
    oracle.add_volume_source((0.3, 0.3, 0.3), band_limited_impulse(n_steps, oracle.dt, 800.0))
    for point in [(0.9, 0.6, 0.5), (0.2, 0.7, 0.6)]:
        oracle.add_pressure_receiver(point)
    oracle.run(n_steps)
    np.testing.assert_array_equal(response.recorded, oracle.recorded_pressure)
