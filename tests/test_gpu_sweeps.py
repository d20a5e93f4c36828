"""The HIP time loop against the CPU oracle and the reference's golden vectors, on the GPU.

fp64: bit-for-bit equal to the oracle (which is itself bit-for-bit the reference NumPy scheme,
``tests/test_oracle.py``) on every field, and to the reference's recorded vectors; energy to
1e-11 relative, the reference's own tolerance (``test_c_backend.py:51-65``). fp32: bit-for-bit
equal to the reference NumPy scheme run in fp32, and within 1e-4 of peak of fp64 over 400
steps (``test_c_backend.py:101-109``).
"""

from __future__ import annotations

import numpy as np
import pytest

from cases import (
    PARITY,
    PARITY_GRID,
    PARITY_RECEIVER,
    PARITY_SOURCE,
    PARITY_STEPS,
    parity_signal,
    prepare,
    to_numpy,
)

from ac_fdtd_amd import (
    AIR,
    AbsorbingLayer,
    AirAbsorption,
    Grid,
    WallAdmittances,
    band_limited_impulse,
)
from oracle.scheme import OracleFDTD

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]


def hip(grid=PARITY_GRID, **kwargs):
    from ac_fdtd_amd.scheme import AcousticFDTD

    return AcousticFDTD(grid, **kwargs)


def assert_same_state(candidate, reference):
    np.testing.assert_array_equal(to_numpy(candidate.p), reference.p)
    for axis in range(3):
        np.testing.assert_array_equal(to_numpy(candidate.velocity[axis]), reference.velocity[axis])
    for n, state in enumerate(reference.relaxation_states):
        np.testing.assert_array_equal(to_numpy(candidate.relaxation_states[n]), state)


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "two-sweep"])
@pytest.mark.parametrize("name", list(PARITY))
def test_fp64_is_bitwise_the_oracle(name, fused):
    reference = prepare(OracleFDTD(PARITY_GRID, **PARITY[name]))
    candidate = prepare(hip(fused=fused, **PARITY[name]))
    reference.run(PARITY_STEPS)
    candidate.run(PARITY_STEPS)
    assert_same_state(candidate, reference)
    np.testing.assert_array_equal(candidate.recorded_pressure, reference.recorded_pressure)
    assert candidate.energy() == pytest.approx(reference.energy(), rel=1e-11)


@pytest.mark.parametrize("name", list(PARITY))
def test_fp64_matches_the_reference_vectors(name, golden):
    candidate = prepare(hip(**PARITY[name]))
    candidate.run(PARITY_STEPS)
    np.testing.assert_array_equal(to_numpy(candidate.p), golden[f"parity_{name}_p"])
    np.testing.assert_array_equal(candidate.recorded_pressure, golden[f"parity_{name}_recording"])
    assert candidate.energy() == pytest.approx(float(golden[f"parity_{name}_energy"]), rel=1e-11)
    if name == "all":
        for axis in range(3):
            np.testing.assert_array_equal(
                to_numpy(candidate.velocity[axis]), golden[f"parity_all_v{axis}"]
            )


def test_fp32_is_bitwise_the_reference_in_fp32(golden):
    candidate = prepare(hip(dtype=np.float32, **PARITY["all"]))
    candidate.run(PARITY_STEPS)
    np.testing.assert_array_equal(to_numpy(candidate.p), golden["parity_all_f32_p"])
    np.testing.assert_array_equal(candidate.recorded_pressure, golden["parity_all_f32_recording"])


@pytest.mark.parametrize("compensated", [False, True])
def test_fp32_stays_close_to_fp64(compensated):
    reference = prepare(OracleFDTD(PARITY_GRID, **PARITY["all"]))
    single = prepare(hip(dtype=np.float32, compensated=compensated, **PARITY["all"]))
    reference.run(400)
    single.run(400)
    recorded = reference.recorded_pressure[0]
    error = np.max(np.abs(recorded - single.recorded_pressure[0])) / np.max(np.abs(recorded))
    assert error < 1e-4


def test_room_receiver_signal_matches_the_reference(golden):
    """The 5 x 4 x 3 m room: absorption 0.15, 1 kHz grid (146 x 117 x 87 cells), air at
    20 C / 50 %, band-limited impulse at (1.0, 1.0, 1.2) m, receiver at (3.5, 2.5, 1.4) m,
    1000 steps. Target 1e-13 relative; the fp64 kernel is in fact bitwise."""
    grid = Grid.from_max_frequency((5.0, 4.0, 3.0), 1000.0, AIR.sound_speed)
    assert grid.shape == (146, 117, 87)
    solver = hip(
        grid,
        walls=WallAdmittances.from_absorption(0.15),
        air_absorption=AirAbsorption(temperature=20.0, relative_humidity=50.0),
    )
    solver.add_volume_source((1.0, 1.0, 1.2), band_limited_impulse(1000, solver.dt, 1000.0))
    solver.add_pressure_receiver((3.5, 2.5, 1.4))
    solver.run(1000)
    recorded = solver.recorded_pressure
    expected = golden["room_recording"]
    relative = np.max(np.abs(recorded - expected)) / np.max(np.abs(expected))
    assert relative <= 1e-13
    np.testing.assert_array_equal(recorded, expected)
    assert solver.energy() == pytest.approx(float(golden["room_energy"]), rel=1e-11)


def test_running_in_pieces_is_bitwise_running_at_once():
    whole = prepare(hip(**PARITY["all"]))
    pieces = prepare(hip(**PARITY["all"]))
    whole.run(120)
    for n in (30, 1, 29, 60):
        pieces.run(n)
    np.testing.assert_array_equal(to_numpy(pieces.p), to_numpy(whole.p))
    np.testing.assert_array_equal(pieces.recorded_pressure, whole.recorded_pressure)


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["fp32", "fp64"])
@pytest.mark.parametrize("nz", [13, 14, 15, 17])
def test_padded_rows_are_the_reference_layout_bit_for_bit(nz, dtype):
    """A grid whose nz the widest lane does not divide runs in padded rows (full-width lanes,
    state copied in and out): fields, relaxation states and recordings equal the same run in
    the reference layout and the oracle's, for run lengths with every remainder."""
    grid = Grid(dx=0.03, nx=16, ny=12, nz=nz)
    config = dict(PARITY["all"])
    padded = prepare(hip(grid, dtype=dtype, **config), grid)
    plain = prepare(hip(grid, dtype=dtype, **config), grid)
    oracle = prepare(OracleFDTD(grid, dtype=dtype, **config), grid)
    plain.padded_rows = False
    for n in (9, 10, 11, 12):
        padded.run(n)
        plain.run(n)
        oracle.run(n)
    widest = 4 if np.dtype(dtype).itemsize == 4 else 2
    assert (padded._padded is not None) == (nz % widest != 0) and plain._padded is None
    assert_same_state(padded, oracle)
    for a, b in zip(padded.save_state(), plain.save_state()):
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    np.testing.assert_array_equal(padded.recorded_pressure, oracle.recorded_pressure)
    np.testing.assert_array_equal(padded.recorded_pressure, plain.recorded_pressure)


def test_a_prepared_run_is_run():
    """prepare_run + its launch (what the bench times) is run() bit for bit; a launch after the
    solver moved on, and the subclasses that advance their own scheme, are refused."""
    from ac_fdtd_amd.fourth import FourthOrderFDTD

    direct = prepare(hip(**PARITY["all"]))
    staged = prepare(hip(**PARITY["all"]))
    direct.run(37)
    launch = staged.prepare_run(37)
    launch()
    assert staged.step_index == direct.step_index == 37
    np.testing.assert_array_equal(to_numpy(staged.p), to_numpy(direct.p))
    np.testing.assert_array_equal(staged.recorded_pressure, direct.recorded_pressure)
    stale = staged.prepare_run(5)
    staged.run(1)
    with pytest.raises(RuntimeError, match="advanced"):
        stale()
    assert staged.prepare_run(0) is None
    with pytest.raises(NotImplementedError):
        FourthOrderFDTD(PARITY_GRID).prepare_run(4)


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "two-sweep"])
@pytest.mark.parametrize("tile_j,chunk_x", [(1, 1), (2, 3), (8, 7), (4, 64)])
def test_tiling_does_not_change_the_answer(tile_j, chunk_x, fused):
    """The reference's thread-count invariance (test_c_backend.py:68-82), for tile shapes."""
    base = prepare(hip(**PARITY["all"]))
    tiled = prepare(hip(tile_j=tile_j, chunk_x=chunk_x, fused=fused, **PARITY["all"]))
    base.run(60)
    tiled.run(60)
    np.testing.assert_array_equal(to_numpy(tiled.p), to_numpy(base.p))
    for axis in range(3):
        np.testing.assert_array_equal(to_numpy(tiled.velocity[axis]), to_numpy(base.velocity[axis]))


@pytest.mark.parametrize("fused", [True, False], ids=["fused", "two-sweep"])
@pytest.mark.parametrize(
    "shape", [(1, 1, 1), (7, 1, 1), (1, 9, 1), (3, 1, 70), (5, 3, 130), (2, 66, 65), (33, 5, 64)]
)
def test_odd_shapes_match_the_oracle(shape, fused):
    """Ducts, single cells, one-cell axes (both walls on one cell), z rows that are not a
    multiple of the 64-lane wavefront, with walls, a layer and air all on."""
    grid = Grid(dx=0.02, nx=shape[0], ny=shape[1], nz=shape[2])
    config = {
        "walls": WallAdmittances(0.3, 0.1, 0.2, 0.0, 0.5, 0.4),
        "absorbing_layer": AbsorbingLayer(thickness=2),
        "air_absorption": AirAbsorption(),
    }
    reference = OracleFDTD(grid, **config)
    candidate = hip(grid, fused=fused, **config)
    for solver in (reference, candidate):
        prepare(solver, grid, seed=7)
    reference.run(41)
    candidate.run(41)
    assert_same_state(candidate, reference)
    np.testing.assert_array_equal(candidate.recorded_pressure, reference.recorded_pressure)


def test_several_sources_and_receivers_in_insertion_order():
    """Two sources in one cell of different lengths, one elsewhere; three receivers, one of
    them on a source cell (scheme.py:360-363 adds sources in insertion order)."""
    config = PARITY["all"]
    reference = OracleFDTD(PARITY_GRID, **config)
    candidate = hip(**config)
    rng = np.random.default_rng(11)
    signals = [rng.standard_normal(40), rng.standard_normal(90), rng.standard_normal(7)]
    points = [PARITY_SOURCE, PARITY_SOURCE, (0.5, 0.1, 0.4)]
    for solver in (reference, candidate):
        for point, signal in zip(points, signals):
            solver.add_volume_source(point, signal)
        for point in (PARITY_RECEIVER, PARITY_SOURCE, (0.0, 0.49, 0.44)):
            solver.add_pressure_receiver(point)
        solver.run(100)
    assert_same_state(candidate, reference)
    np.testing.assert_array_equal(candidate.recorded_pressure, reference.recorded_pressure)


def test_a_receiver_added_part_way_records_from_then_on():
    reference = OracleFDTD(PARITY_GRID, **PARITY["walls"])
    candidate = hip(**PARITY["walls"])
    for solver in (reference, candidate):
        prepare(solver)
        solver.run(20)
        solver.add_pressure_receiver((0.1, 0.1, 0.1))
        solver.run(15)
    expected = [np.array(samples) for _, samples in reference.receivers]
    assert [e.size for e in expected] == [35, 15]
    for channel, samples in enumerate(expected):
        np.testing.assert_array_equal(candidate.recorded_channel(channel), samples)
    with pytest.raises(ValueError):
        candidate.recorded_pressure  # noqa: B018  (ragged, as in the reference)


def test_user_edits_between_runs_are_seen_by_both_kernel_paths():
    """State written by the caller between runs (as the reference tests do through
    ``solver.p[:] = ...``), including non-zero wall faces, is what the next run starts from."""
    import torch

    config = PARITY["all"]
    reference = prepare(OracleFDTD(PARITY_GRID, **config))
    candidate = prepare(hip(**config))
    rng = np.random.default_rng(5)
    for n_steps in (4, 3):
        reference.run(n_steps)
        candidate.run(n_steps)
        for axis in range(3):
            noise = rng.standard_normal(reference.velocity[axis].shape)
            reference.velocity[axis][:] = noise
            candidate.velocity[axis].copy_(torch.as_tensor(noise))
    reference.run(6)
    candidate.run(6)
    assert_same_state(candidate, reference)


def test_receivers_record_the_same_samples_as_probing_by_hand():
    solver = hip()
    solver.add_volume_source(PARITY_SOURCE, np.ones(50))
    solver.add_pressure_receiver(PARITY_RECEIVER)
    solver.run(20)
    solver.run(20)
    recorded = solver.recorded_pressure
    assert recorded.shape == (1, 40)
    assert recorded[0, -1] == solver.probe_pressure(PARITY_RECEIVER)


def test_source_signal_longer_and_shorter_than_the_run():
    reference = OracleFDTD(PARITY_GRID)
    candidate = hip()
    for solver in (reference, candidate):
        solver.add_volume_source(PARITY_SOURCE, parity_signal()[:13])
        solver.add_pressure_receiver(PARITY_RECEIVER)
        solver.run(25)
    np.testing.assert_array_equal(candidate.recorded_pressure, reference.recorded_pressure)
