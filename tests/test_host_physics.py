"""Host-side physics: grid, walls, sponge, air, signals, closed forms (CPU).

Two kinds of pin: the reference's own closed-form tests re-expressed (``tests/test_grid.py``,
``test_boundaries.py:27-110``, ``test_air.py:15-67``, ``test_scheme.py:101-115``), and
bit-for-bit agreement with vectors written from the reference implementation
(``tests/golden/make_golden.py``).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from ac_fdtd_amd import (
    AIR,
    AbsorbingLayer,
    AirAbsorption,
    Grid,
    Medium,
    WallAdmittances,
    admittance_from_absorption,
    band_limited_impulse,
    deconvolve,
    discrete_mode_frequency,
    gaussian_pulse,
    max_stable_time_step,
    mode_frequency,
    random_incidence_absorption,
    reflection_coefficient,
)
from ac_fdtd_amd.boundaries import (
    axis_wall_codes,
    layer_factors,
    wall_coefficient_table,
    wall_update_coefficients,
)
from ac_fdtd_amd.coefficients import build_coefficients
from ac_fdtd_amd.sources import gaussian_derivative_pulse

FREQUENCIES = np.array([125.0, 250.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0])


# -- grid (reference tests/test_grid.py) ------------------------------------------------------
def test_lengths_report_the_room_actually_simulated():
    grid = Grid.from_lengths((1.75, 1.0, 1.0), dx=0.1)
    assert grid.shape == (18, 10, 10)
    assert grid.lengths[0] == pytest.approx(1.8)
    assert grid.length_error[0] == pytest.approx(0.05)
    assert grid.length_error[1] == pytest.approx(0.0)


def test_spacing_and_time_step():
    grid = Grid.from_max_frequency((3.0, 3.0, 3.0), 1000.0, 343.0, 10.0)
    assert grid.dx == pytest.approx(0.0343)
    dt = max_stable_time_step(0.05, AIR.sound_speed)
    assert AIR.sound_speed * dt * math.sqrt(3) / 0.05 == pytest.approx(1.0)


def test_cell_index_clips_and_snaps():
    grid = Grid.from_lengths((1.0, 1.0, 1.0), dx=0.1)
    assert grid.cell_index((0.0, 0.0, 0.0)) == (0, 0, 0)
    assert grid.cell_index((0.55, 0.05, 0.05)) == (5, 0, 0)
    assert grid.cell_index((1.0, 1.0, 1.0)) == (9, 9, 9)
    fine = Grid(dx=0.01, nx=200, ny=10, nz=10)
    assert fine.cell_index((1.2, 0.0, 0.0))[0] == 120  # 1.2 / 0.01 = 119.99999999999999


def test_memory_estimate_counts_the_velocity_planes():
    grid = Grid.from_lengths((1.0, 1.0, 1.0), dx=0.1)
    faces = 3 * 11 * 10 * 10
    assert grid.memory_bytes(itemsize=4) == (1000 + faces) * 4
    assert grid.memory_bytes(itemsize=4, extra_fields=2) == (3000 + faces) * 4


@pytest.mark.parametrize("bad", [0.0, -1.0])
def test_rejects_nonsense_spacing(bad):
    with pytest.raises(ValueError):
        Grid(dx=bad, nx=1, ny=1, nz=1)


# -- walls (reference tests/test_boundaries.py) ------------------------------------------------
def test_absorption_and_admittance_are_inverse():
    for absorption in (0.0, 0.1, 0.5, 0.9, 1.0):
        xi = admittance_from_absorption(absorption)
        assert 1.0 - reflection_coefficient(xi) ** 2 == pytest.approx(absorption)
    assert reflection_coefficient(0.0) == pytest.approx(1.0)
    assert reflection_coefficient(1.0) == pytest.approx(0.0)


def test_random_incidence_absorption_exceeds_normal_for_a_live_wall():
    # Paris formula: a wall quoted at 0.15 at normal incidence absorbs about 0.25 diffusely.
    xi = admittance_from_absorption(0.15)
    assert random_incidence_absorption(xi) == pytest.approx(0.25, abs=0.02)
    # A matched wall: 1 - ((cos t - 1)/(cos t + 1))^2 averaged with sin 2t; closed form
    # 16 ln 2 - 32/3 + ... is not needed, the numeric value 0.9109 is (brute-force quadrature).
    angles = np.linspace(0.0, math.pi / 2.0, 200001)
    r = (np.cos(angles) - 1.0) / (np.cos(angles) + 1.0)
    brute = np.trapezoid((1.0 - r**2) * np.sin(2.0 * angles), angles)
    assert random_incidence_absorption(1.0) == pytest.approx(brute, rel=1e-5)


def test_negative_admittance_is_refused():
    with pytest.raises(ValueError, match="generates sound"):
        WallAdmittances(x_min=-0.1)


def test_corner_cells_sum_every_face_they_touch():
    grid = Grid.from_lengths((0.3, 0.3, 0.3), 0.1)
    indices, coefficient = wall_update_coefficients(grid, AIR, 1e-5, WallAdmittances.uniform(1.0))
    assert indices.size == 26
    assert np.array_equal(indices, np.unique(indices))
    counts = np.round(coefficient / coefficient.min()).astype(int)
    assert list(np.bincount(counts)[1:]) == [6, 12, 8]


def test_wall_table_agrees_with_the_sparse_list():
    grid = Grid(dx=0.02, nx=5, ny=1, nz=4)
    walls = WallAdmittances(0.3, 0.0, 0.2, 0.4, 0.0, 0.7)
    dt = max_stable_time_step(grid.dx, AIR.sound_speed)
    indices, a = wall_update_coefficients(grid, AIR, dt, walls)
    is_wall, updated, previous = wall_coefficient_table(grid, AIR, dt, walls)
    cx, cy, cz = (axis_wall_codes(n) for n in grid.shape)
    codes = (cx[:, None, None] + 4 * cy[None, :, None] + 16 * cz[None, None, :]).ravel()
    assert np.array_equal(np.flatnonzero(is_wall[codes]), indices)
    assert np.array_equal(updated[codes[indices]], 1.0 / (1.0 + a))
    assert np.array_equal(previous[codes[indices]], a / (1.0 + a))


def test_walls_match_the_reference_vectors(golden):
    grid = Grid.from_lengths((0.6, 0.5, 0.45), 0.03)
    dt = max_stable_time_step(grid.dx, AIR.sound_speed)
    walls = WallAdmittances(x_min=0.2, x_max=0.05, y_min=0.3, z_max=0.7)
    indices, coefficient = wall_update_coefficients(grid, AIR, dt, walls)
    assert np.array_equal(indices, golden["walls_indices"])
    assert np.array_equal(coefficient, golden["walls_coefficient"])
    duct = Grid(dx=0.01, nx=7, ny=1, nz=3)
    indices, coefficient = wall_update_coefficients(duct, AIR, 1e-5, WallAdmittances.uniform(0.4))
    assert np.array_equal(indices, golden["walls_duct_indices"])
    assert np.array_equal(coefficient, golden["walls_duct_coefficient"])


# -- sponge -------------------------------------------------------------------------------------
def test_the_layer_leaves_the_interior_alone():
    grid = Grid.from_lengths((1.0, 1.0, 1.0), 0.05)
    pressure, velocity = layer_factors(grid, AIR, 1e-5, AbsorbingLayer(thickness=4))
    for axis in range(3):
        assert np.all(pressure[axis][4:-4] == 1.0)
        assert np.any(pressure[axis][:4] < 1.0)
        assert np.all(velocity[axis][axis][5:-5] == 1.0)
        for other in range(3):
            if other != axis:
                assert np.array_equal(velocity[axis][other], pressure[other])


def test_a_layer_can_line_a_subset_of_the_walls():
    grid = Grid.from_lengths((1.0, 1.0, 1.0), 0.05)
    pressure, _ = layer_factors(grid, AIR, 1e-5, AbsorbingLayer(thickness=4, faces=((0, 1),)))
    assert np.all(pressure[0][:-4] == 1.0)
    assert np.any(pressure[0][-4:] < 1.0)
    assert np.all(pressure[1] == 1.0) and np.all(pressure[2] == 1.0)


def test_the_layer_grading_integrates_to_the_target_reflection():
    """Round trip through the layer: exp(-2/c * integral sigma) = R. Sampling the profile on
    the faces of a fine grid and multiplying the per-step factors over one crossing at speed
    c must reproduce sqrt(R) per pass, independently of the formula's shape."""
    n, thickness, target = 4000, 2000, 1e-3
    grid = Grid(dx=1e-3, nx=n, ny=1, nz=1)
    dt = grid.dx / AIR.sound_speed  # one cell per step
    layer = AbsorbingLayer(thickness=thickness, target_reflection=target, faces=((0, 0),))
    _, velocity = layer_factors(grid, AIR, dt, layer)
    one_way = np.prod(velocity[0][0][: thickness + 1] ** 0.5 * velocity[0][0][1 : thickness + 2] ** 0.5)
    assert one_way**2 == pytest.approx(target, rel=1e-2)


def test_layer_matches_the_reference_vectors(golden):
    grid = Grid.from_lengths((0.6, 0.5, 0.45), 0.03)
    dt = max_stable_time_step(grid.dx, AIR.sound_speed)
    layers = {
        "all": AbsorbingLayer(thickness=4),
        "subset": AbsorbingLayer(
            thickness=3, target_reflection=1e-3, order=2.0, faces=((0, 1), (2, 0))
        ),
    }
    for tag, layer in layers.items():
        pressure, velocity = layer_factors(grid, AIR, dt, layer)
        for axis in range(3):
            assert np.array_equal(pressure[axis], golden[f"layer_{tag}_p{axis}"])
            assert np.array_equal(velocity[axis][axis], golden[f"layer_{tag}_v{axis}"])


# -- air (reference tests/test_air.py) ---------------------------------------------------------
def test_relaxation_frequencies_land_where_they_should():
    air = AirAbsorption(temperature=20.0, relative_humidity=50.0)
    assert 100.0 < air.nitrogen_relaxation_frequency < 1000.0
    assert 10e3 < air.oxygen_relaxation_frequency < 100e3


def test_attenuation_matches_published_iso_9613_1_values():
    """ISO 9613-1 Table 1 at 20 C, 50 % RH, 101.325 kPa (dB/km): 125 Hz 0.44, 250 Hz 1.31,
    500 Hz 2.73, 1 kHz 4.66, 2 kHz 9.86, 4 kHz 29.7, 8 kHz 105 (published to 3 figures)."""
    table = np.array([0.44, 1.31, 2.73, 4.66, 9.86, 29.7, 105.0])
    computed = 1000.0 * AirAbsorption(20.0, 50.0).attenuation(FREQUENCIES)
    np.testing.assert_allclose(computed, table, rtol=0.02)


def test_attenuation_rises_with_frequency_and_falls_with_humidity():
    dry = AirAbsorption(temperature=20.0, relative_humidity=10.0).attenuation(FREQUENCIES)
    damp = AirAbsorption(temperature=20.0, relative_humidity=80.0).attenuation(FREQUENCIES)
    assert np.all(np.diff(dry) > 0.0) and np.all(np.diff(damp) > 0.0)
    assert dry[-1] > 2.0 * damp[-1]


def test_relaxation_dominates_the_audio_band():
    full = AirAbsorption(20.0, 50.0).attenuation(FREQUENCIES)
    classical = AirAbsorption(20.0, 50.0, relaxation=False).attenuation(FREQUENCIES)
    assert np.all(full / classical > 5.0)


def test_the_implemented_medium_reproduces_the_standard():
    air = AirAbsorption(temperature=20.0, relative_humidity=50.0)
    f = np.geomspace(50.0, 16000.0, 60)
    model = air.model_attenuation(f, AIR.sound_speed, 3.4e-6)
    assert np.max(np.abs(model / air.attenuation(f) - 1.0)) < 0.01
    assert 1e-4 < air.total_strength(AIR.sound_speed, 3.4e-6) < 1e-2


def test_air_matches_the_reference_vectors(golden):
    for temperature, humidity in ((20.0, 50.0), (15.0, 70.0), (25.0, 30.0), (20.0, 30.0)):
        key = f"air_{temperature:g}_{humidity:g}"
        air = AirAbsorption(temperature=temperature, relative_humidity=humidity)
        processes = air.processes(AIR.sound_speed, float(golden[key + "_dt"]))
        got = np.array([(q.strength, q.relaxation_time) for q in processes])
        assert np.array_equal(got, golden[key + "_processes"])
        assert np.array_equal(air.attenuation(golden["air_frequencies"]), golden[key + "_attenuation"])
        h = [air.molar_water_concentration, air.oxygen_relaxation_frequency,
             air.nitrogen_relaxation_frequency]
        assert np.array_equal(np.array(h), golden[key + "_h"])


def test_medium_matches_the_reference_vectors(golden):
    for temperature in (15.0, 20.0, 25.0):
        medium = Medium.air(temperature)
        assert np.array_equal(golden[f"medium_{temperature:g}"], [medium.sound_speed, medium.density])
    # Ideal gas and c ~ sqrt(T): 343.2 m/s and 1.204 kg/m^3 at 20 C.
    assert AIR.sound_speed == pytest.approx(343.21, abs=0.01)
    assert AIR.density == pytest.approx(1.2041, abs=1e-4)


# -- signals ------------------------------------------------------------------------------------
def test_signals_match_the_reference_vectors(golden):
    grid = Grid.from_max_frequency((5.0, 4.0, 3.0), 1000.0, AIR.sound_speed)
    dt = max_stable_time_step(grid.dx, AIR.sound_speed)
    assert np.array_equal(band_limited_impulse(1000, dt, 1000.0), golden["band_limited_impulse"])
    assert np.array_equal(gaussian_pulse(300, dt, 1000.0), golden["gaussian_pulse"])
    assert np.array_equal(golden["room_grid"], [grid.dx, *grid.shape])
    cells = [grid.cell_index((1.0, 1.0, 1.2)), grid.cell_index((3.5, 2.5, 1.4))]
    assert np.array_equal(golden["room_cells"], cells)


def test_band_limited_impulse_is_flat_in_band_and_empty_above():
    dt = 1.0 / 48000.0
    signal = band_limited_impulse(8192, dt, 4000.0)
    spectrum = np.abs(np.fft.rfft(signal))
    f = np.fft.rfftfreq(signal.size, dt)
    assert np.allclose(spectrum[(f > 100) & (f < 3000)], 1.0, atol=0.02)
    assert np.max(spectrum[f > 6000]) < 1e-3
    assert signal.sum() == pytest.approx(1.0, rel=1e-3)  # DC gain of the windowed sinc


def test_gaussian_pulse_is_40_db_down_at_the_stated_band_edge():
    dt = 1e-5
    pulse = gaussian_pulse(20000, dt, 1000.0)
    spectrum = np.abs(np.fft.rfft(pulse))
    f = np.fft.rfftfreq(pulse.size, dt)
    level = 20 * np.log10(np.interp(1000.0, f, spectrum) / spectrum[0])
    assert level == pytest.approx(-40.0, abs=0.1)


def test_gaussian_derivative_pulse_is_the_derivative():
    dt, width = 1e-5, 3e-4
    t = (np.arange(400) + 0.5) * dt
    pulse = gaussian_derivative_pulse(400, dt, width)
    gauss = np.exp(-(((t - 4 * width) / width) ** 2))
    np.testing.assert_allclose(pulse, np.gradient(gauss, dt) / 2.0, atol=2e-3 * np.abs(pulse).max())
    assert abs(pulse.sum()) < 1e-6 * np.abs(pulse).sum()


def test_deconvolution_recovers_a_delayed_impulse():
    dt, n = 1e-4, 2048
    excitation = band_limited_impulse(n, dt, 2000.0)
    response = np.zeros(n)
    response[300] = 1.0
    recorded = np.real(np.fft.irfft(np.fft.rfft(response) * np.fft.rfft(excitation), n=n))
    recovered = deconvolve(recorded, excitation, dt, 2000.0)
    assert int(np.argmax(recovered)) == 300


# -- closed forms (reference tests/test_scheme.py:101-115, test_dispersion.py) ---------------------
def test_discrete_frequency_approaches_the_analytic_one_at_second_order():
    mode, lengths = (2, 1, 1), (1.6, 1.2, 1.0)
    errors = []
    for dx in (0.05, 0.025, 0.0125):
        grid = Grid.from_lengths(lengths, dx)
        dt = max_stable_time_step(grid.dx, AIR.sound_speed)
        exact = mode_frequency(grid.lengths, mode, AIR.sound_speed)
        errors.append(abs(discrete_mode_frequency(grid, mode, AIR, dt) - exact) / exact)
    ratios = [errors[i] / errors[i + 1] for i in range(2)]
    assert all(3.5 < r < 4.5 for r in ratios)


def test_body_diagonal_mode_is_exact_at_the_courant_limit():
    """On the body diagonal at Courant 1 the discrete dispersion is exact: a mode with equal
    wavenumber on all three axes has sin(w dt/2) = sqrt(3) (c dt/dx) sin(k dx/2), which at
    c dt/dx = 1/sqrt(3) and k dx = pi gives w dt = pi: the Nyquist frequency, exactly.
    asin is evaluated at 1 - O(1e-16) there, where it amplifies round-off to sqrt(2e-16)."""
    grid = Grid(dx=0.05, nx=8, ny=8, nz=8)
    dt = max_stable_time_step(grid.dx, AIR.sound_speed)
    f = discrete_mode_frequency(grid, (8, 8, 8), AIR, dt)
    assert f == pytest.approx(0.5 / dt, rel=3e-8)


# -- coefficients ---------------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_coefficients_agree_with_the_oracle(dtype):
    from oracle.scheme import OracleFDTD

    grid = Grid.from_lengths((0.6, 0.5, 0.45), 0.03)
    config = {
        "walls": WallAdmittances(x_min=0.2, z_max=0.7),
        "absorbing_layer": AbsorbingLayer(thickness=3),
        "air_absorption": AirAbsorption(temperature=15.0, relative_humidity=70.0),
    }
    k = build_coefficients(grid, AIR, 1.0, dtype, config["walls"], config["absorbing_layer"],
                           config["air_absorption"])
    o = OracleFDTD(grid, dtype=dtype, **config)
    assert (k.velocity, k.pressure, k.source) == (o.k_vel, o.k_pres, o.k_src)
    assert k.relax_scale == o.relax_scale
    assert [(d, m, g) for d, m, g in zip(k.relax_decay, k.relax_mean, k.relax_gain)] == [
        (d, m, g) for _, d, m, g in o.relax
    ]
    assert k.n_relax == 3


def test_courant_outside_the_unit_interval_is_refused():
    grid = Grid(dx=0.1, nx=2, ny=2, nz=2)
    for bad in (0.0, 1.01, -0.5):
        with pytest.raises(ValueError, match="(?i)courant"):
            build_coefficients(grid, AIR, bad, np.float64, None, None, None)
