"""The energy entry point that allocates nothing (``acfdtd_hip_energy_noalloc``) against the
one that takes a caller workspace (``acfdtd_hip_energy``), through the C-ABI.

Two calls in a row, the second on a smaller grid (the library's scratch keeps the first call's
partials, which the second must not see): bit for bit where the descriptor's tiling fits the
scratch, within 1e-13 of the largest term where the library re-tiles (longer marches, fp64 sums in another
order). Also: the second result follows a change of state, and fp32 fields reduce alike.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from cases import PARITY

from ac_fdtd_amd import Grid

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]


def _solver(shape, dtype, seed, **kwargs):
    from ac_fdtd_amd.scheme import AcousticFDTD

    rng = np.random.default_rng(seed)
    grid = Grid(dx=0.03, nx=shape[0], ny=shape[1], nz=shape[2])
    solver = AcousticFDTD(grid, dtype=dtype, **PARITY["all"], **kwargs)
    solver.p.copy_(torch.as_tensor(rng.standard_normal(grid.shape)))
    for a in range(3):
        solver.velocity[a].copy_(torch.as_tensor(rng.standard_normal(grid.velocity_shape(a))))
    return solver


def _both(solver):
    from ac_fdtd_amd import _native

    lib = _native.library()
    desc = solver._descriptor(None)
    stream = ctypes.c_void_p(torch.cuda.current_stream(solver.device).cuda_stream)
    size = lib.acfdtd_hip_energy_workspace_bytes(ctypes.byref(desc))
    assert size > 0
    workspace = torch.empty(size // 8 + 1, dtype=torch.float64, device=solver.device)
    with_workspace = torch.full((3,), np.nan, dtype=torch.float64, device=solver.device)
    without = torch.full((3,), np.nan, dtype=torch.float64, device=solver.device)
    with torch.cuda.device(solver.device):
        _native.check(lib.acfdtd_hip_energy(ctypes.byref(desc), ctypes.c_void_p(workspace.data_ptr()),
                                            ctypes.c_void_p(with_workspace.data_ptr()), stream),
                      "acfdtd_hip_energy")
        _native.check(lib.acfdtd_hip_energy_noalloc(ctypes.byref(desc),
                                                    ctypes.c_void_p(without.data_ptr()), stream),
                      "acfdtd_hip_energy_noalloc")
    torch.cuda.synchronize()
    return with_workspace.cpu().numpy(), without.cpu().numpy(), size // 24


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["fp64", "fp32"])
def test_no_alloc_energy_equals_the_workspace_energy(dtype):
    from ac_fdtd_amd import _native

    for shape, seed in (((40, 33, 64), 1), ((9, 7, 13), 2)):  # the second call is smaller
        solver = _solver(shape, dtype, seed)
        expected, got, tiles = _both(solver)
        assert tiles <= _native.ENERGY_SCRATCH_TILES
        assert np.all(np.isfinite(expected))
        np.testing.assert_array_equal(got, expected)
        solver.run(5)  # a new state: the next reduction follows it
        expected_after, got_after, _ = _both(solver)
        assert not np.array_equal(expected_after, expected)
        np.testing.assert_array_equal(got_after, expected_after)


def test_no_alloc_energy_retiles_when_the_tiling_exceeds_the_scratch():
    from ac_fdtd_amd import _native

    big = _solver((700, 50, 64), np.float64, 3, tile_j=1, chunk_x=1)
    expected, got, tiles = _both(big)
    assert tiles > _native.ENERGY_SCRATCH_TILES
    # Relative to the largest term: the coupling sum cancels, its own relative error is larger.
    np.testing.assert_allclose(got, expected, rtol=0, atol=1e-13 * np.max(np.abs(expected)))
    small = _solver((11, 5, 8), np.float64, 4)
    expected, got, _ = _both(small)
    np.testing.assert_array_equal(got, expected)
