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
from ac_fdtd_amd.io import resample, resampling_ratio, write_wav
from ac_fdtd_amd.room import Room, RunPlan, simulate_impulse_response, simulate_impulse_responses

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


def test_resampling_keeps_a_tone_and_its_duration():
    """reference tests/test_room.py:126-134, plus the tone's frequency from its spectrum."""
    source_rate = 59400.0
    tone = np.sin(2 * np.pi * 1000.0 * np.arange(int(0.2 * source_rate)) / source_rate)
    out = resample(tone, source_rate, 48000.0)
    assert out.size == pytest.approx(0.2 * 48000.0, rel=0.01)
    middle = out[1000:-1000]
    assert np.max(np.abs(middle)) == pytest.approx(1.0, rel=0.01)
    spectrum = np.abs(np.fft.rfft(middle * np.hanning(middle.size)))
    peak = np.argmax(spectrum) * 48000.0 / middle.size
    assert peak == pytest.approx(1000.0, abs=48000.0 / middle.size)
    ratio = resampling_ratio(17320.508075688773, 48000.0)
    assert ratio.denominator <= 1000
    assert float(ratio) == pytest.approx(48000.0 / 17320.508075688773, rel=1e-4)
    with pytest.raises(ValueError, match="positive"):
        resample(tone, 0.0, 48000.0)


def test_wav_files_read_back(tmp_path):
    """reference tests/test_room.py:137-143: two channels at 48 kHz, peak 1; and unnormalised
    frames at the grid's own rate read back bit for bit as float32."""
    from scipy.io import wavfile

    signals = np.random.default_rng(0).standard_normal((2, 6000))
    path = write_wav(tmp_path / "sub" / "ir.wav", signals, 59400.0, target_rate=48000.0)
    rate, data = wavfile.read(path)
    assert rate == 48000 and data.dtype == np.float32 and data.shape[1] == 2
    assert np.max(np.abs(data)) == pytest.approx(1.0, rel=1e-6)
    raw = write_wav(tmp_path / "raw.wav", signals[0], 17320.5, target_rate=None, normalise=False)
    rate, data = wavfile.read(raw)
    assert rate == 17320 and data.shape == (6000,)
    np.testing.assert_array_equal(data, signals[0].astype(np.float32))


def test_package_import_needs_neither_torch_nor_the_library():
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, 'ac-fdtd_amd'); import ac_fdtd_amd as a; "
            "[getattr(a, n) for n in a.__all__ if n not in ('AcousticFDTD', 'BatchedRooms')]; "
            "assert 'torch' not in sys.modules; print('ok')")
    root = __import__("pathlib").Path(__file__).resolve().parents[1]
    out = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"


def test_run_plan_follows_the_room():
    """The grid resolves max_frequency at 10 points per wavelength, the default duration is
    1.5 Sabine times, the excitation is the band-limited impulse of that many steps."""
    room = Room(dimensions=(3.0, 2.5, 2.2), absorption=0.25)
    plan = RunPlan.for_room(room, 700.0)
    assert plan.grid.dx == pytest.approx(room.medium.sound_speed / 7000.0)
    assert plan.n_steps == int(round(1.5 * room.sabine_reverberation_time() / plan.dt))
    assert plan.excitation.shape == (plan.n_steps,)
    assert plan.dt == pytest.approx(plan.grid.dx / (room.medium.sound_speed * np.sqrt(3.0)))
    with pytest.raises(ValueError, match="time step"):
        RunPlan.for_room(room, 700.0, duration=1e-7)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_batched_room_driver_is_bitwise_rooms_alone_and_states_its_error():
    """Three rooms of different sizes and durations advanced together in fp32: each response
    equals its room run alone, and the fp32 error against an fp64 rerun is stated and small."""
    jobs = [
        (Room((1.1, 0.9, 0.8), absorption=0.2), (0.3, 0.3, 0.3), [(0.8, 0.6, 0.5)]),
        (Room((1.6, 0.7, 1.0), absorption=0.4, air=None), (0.4, 0.2, 0.3),
         [(1.2, 0.5, 0.7), (0.3, 0.5, 0.8)]),
        (Room((0.9, 1.3, 0.6), absorption=0.1), (0.2, 0.9, 0.3), [(0.6, 0.4, 0.4)]),
    ]
    batched = simulate_impulse_responses(jobs, 900.0, duration=0.03, reference_error=True)
    for (room, source, receivers), response in zip(jobs, batched):
        alone = simulate_impulse_response(room, source, receivers, 900.0, duration=0.03,
                                          dtype=np.float32)
        assert response.dtype == "float32"
        np.testing.assert_array_equal(response.recorded, alone.recorded)
        np.testing.assert_array_equal(response.signals, alone.signals)
        assert 0.0 < response.fp64_relative_l2 < 1e-5


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
