"""The tile autotuner: cache behaviour (CPU), and on the GPU that a tuned solver computes the
same fields as an untuned one (bit for bit: tiles never change a cell's arithmetic) and the
same gradients (to the round-off of the fp64 partial sums)."""

from __future__ import annotations

import json

import numpy as np
import pytest
import torch

from cases import PARITY

from ac_fdtd_amd import Grid, autotune


def test_cache_round_trips_through_the_file(tmp_path, monkeypatch):
    path = tmp_path / "tiles.json"
    monkeypatch.setenv("ACFDTD_TUNE_CACHE", str(path))
    monkeypatch.setattr(autotune, "_cache", {})
    monkeypatch.setattr(autotune, "_loaded", False)
    autotune._store("forward:x", (2, 3))
    assert json.loads(path.read_text()) == {"forward:x": [2, 3]}
    monkeypatch.setattr(autotune, "_cache", {})
    monkeypatch.setattr(autotune, "_loaded", False)
    assert autotune._lookup("forward:x") == (2, 3)
    assert autotune._lookup("forward:y") is None


def test_candidates_are_valid_tilings():
    for tile_j, chunk in autotune.CANDIDATES:
        assert tile_j in (1, 2, 4, 8) and chunk >= 1


def _solver(autotuned: bool, dtype):
    from ac_fdtd_amd.scheme import AcousticFDTD

    grid = Grid(dx=0.03, nx=40, ny=28, nz=36)
    solver = AcousticFDTD(grid, dtype=dtype, autotune=autotuned, **PARITY["all"])
    rng = np.random.default_rng(0)
    solver.p.copy_(torch.as_tensor(rng.standard_normal(grid.shape)))
    solver.add_volume_source((0.3, 0.2, 0.5), rng.standard_normal(50))
    solver.add_pressure_receiver((0.9, 0.6, 0.4))
    return solver


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tuned_solver_is_bitwise_the_default(dtype):
    tuned, plain = _solver(True, dtype), _solver(False, dtype)
    assert (tuned.tile_j, tuned.chunk_x) in autotune.CANDIDATES + autotune.CANDIDATES_TWO_STEP
    assert tuned.time_block in (1, 2)
    tuned.run(41)
    plain.run(41)
    assert torch.equal(tuned.p, plain.p)
    for a in range(3):
        assert torch.equal(tuned.velocity[a], plain.velocity[a])
    np.testing.assert_array_equal(tuned.recorded_pressure, plain.recorded_pressure)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
@pytest.mark.parametrize("air", [False, True], ids=["walls", "walls+air"])
def test_tuned_gradients_match_untuned(air):
    from ac_fdtd_amd.autograd import RoomResponse

    grid = Grid(dx=0.04, nx=24, ny=20, nz=16)
    rng = np.random.default_rng(1)
    sources = [((0.3, 0.3, 0.2), rng.standard_normal(60))]
    receivers = [(0.7, 0.5, 0.4), (0.2, 0.6, 0.5)]
    weights = torch.as_tensor(rng.standard_normal((2, 90))).cuda()
    grads = []
    for tuned in (False, True):
        room = RoomResponse(grid, 90, sources, receivers, dtype=np.float64, checkpoint_every=25,
                            autotune=tuned, **PARITY["all"])
        xi = room.default_admittance(device="cuda").requires_grad_()
        inputs = [xi]
        strengths = None
        if air:
            strengths = room.default_air_strength(device="cuda").requires_grad_()
            inputs.append(strengths)
        out = room(admittance=xi, air_strength=strengths)
        grads.append(torch.autograd.grad((out * weights).sum(), inputs))
    for a, b in zip(*grads):
        np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=1e-11)
