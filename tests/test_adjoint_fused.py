"""The fused adjoint step against the two-sweep adjoint, bit for bit.

``acfdtd_hip_adjoint_run`` advances the adjoint either as one fused sweep per step (pairs over
two buffer sets, when the swap set is given) or as the face and cell sweeps. The per-cell
arithmetic of the two is the same, so the adjoint fields and the source-signal gradient must be
identical; the fp64 gradient partials are summed in another order and agree to round-off.
Cases cover 4-, 2- and 1-wide lanes (nz a multiple of 4, of 2, odd), odd segment lengths (so
each call ends with a two-sweep step), the boundary-tape path alone and the recompute path with
every gradient, fp64 and fp32.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from cases import PARITY

from ac_fdtd_amd import Grid

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

SHAPES = [(20, 14, 12), (17, 9, 10), (9, 11, 7)]


def _adjoint(fused: bool, dtype, shape, n_steps: int, every: int, want, time_block=1,
             tiling=None):
    from ac_fdtd_amd.autograd import GradientReport, RoomResponse, _run_adjoint

    rng = np.random.default_rng(5)
    grid = Grid(dx=0.04, nx=shape[0], ny=shape[1], nz=shape[2])
    lengths = np.array(grid.lengths)
    sources = [(tuple(0.3 * lengths + 0.01), rng.standard_normal(n_steps - 7))]
    receivers = [tuple(f * lengths) for f in (0.2, 0.55, 0.8)]
    room = RoomResponse(grid, n_steps, sources, receivers, dtype=dtype, checkpoint_every=every,
                        fused=fused, time_block=time_block, **PARITY["all"])
    solver = room.solver(wall_table_always=True)
    if tiling is not None:
        solver.tile_j, solver.chunk_x = tiling
    tape = torch.empty((n_steps, solver.wall_tape_stride()), dtype=solver.dtype,
                       device=solver.device)
    solver.attach_wall_tape(tape, 0)
    checkpoints = []
    if want[2] or want[3]:
        for start in range(0, n_steps, every):
            checkpoints.append(solver.save_state())
            solver.run(min(every, n_steps - start))
    else:
        solver.run(n_steps)
    G = torch.as_tensor(rng.standard_normal((len(receivers), n_steps)), dtype=solver.dtype,
                        device=solver.device)
    report = GradientReport()
    out = _run_adjoint(room, solver, G, checkpoints, want, report)
    torch.cuda.synchronize()
    return out, report


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", SHAPES, ids=["nz12", "nz10", "nz7"])
@pytest.mark.parametrize("mode", ["tape", "recompute"])
def test_fused_adjoint_is_bitwise_the_two_sweep_adjoint(dtype, shape, mode):
    want = (True, True, False, False) if mode == "tape" else (True, True, True, True)
    n_steps, every = 61, 17
    fused, _ = _adjoint(True, dtype, shape, n_steps, every, want)
    plain, report = _adjoint(False, dtype, shape, n_steps, every, want)
    if mode == "recompute":
        assert report.recomputed_steps == n_steps
    for a, b in zip(fused["adjoint_state"], plain["adjoint_state"]):
        assert torch.equal(a, b)
        assert a.abs().max().item() > 0
    assert torch.equal(fused["signal"], plain["signal"])
    keys = ["admittance"] + (["sponge_strength", "air_strength"] if mode == "recompute" else [])
    for key in keys:
        got, ref = np.atleast_1d(fused[key]), np.atleast_1d(plain[key])
        assert np.all(np.isfinite(ref)) and np.abs(ref).max() > 0
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


# (shape, tile_j, chunk_x): workgroup edges inside the grid along x and y; the last case
# splits rows of 32 fp32 lane groups into segments (10 rows of 32 do not fit 512 lanes).
TWO_STEP_CASES = [((20, 14, 12), 1, 3), ((13, 22, 16), 2, 5), ((9, 30, 64), 3, 4),
                  ((11, 37, 128), 8, 4),
                  # 62 row blocks of 2 rows in 34 chunks overrun a round of one-wavefront
                  # workgroups on 256 CUs: balanced chunks of 2 to 3 planes
                  ((200, 124, 32), 1, 6)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", range(len(TWO_STEP_CASES)))
@pytest.mark.parametrize("n_steps", [61, 62])
def test_two_step_adjoint_is_bitwise_the_single_step_adjoint(dtype, case, n_steps):
    """Two adjoint steps per sweep (wall and signal gradients) against one per sweep: the
    adjoint fields and the signal gradient bit for bit, the wall partials to round-off. 61
    steps leave one step after the last four-step group, 62 leave two."""
    shape, tile_j, chunk_x = TWO_STEP_CASES[case]
    want = (True, True, False, False)
    two, report = _adjoint(True, dtype, shape, n_steps, 17, want, time_block=2,
                           tiling=(tile_j, chunk_x))
    assert report.adjoint_two_step
    one, report = _adjoint(True, dtype, shape, n_steps, 17, want, time_block=1,
                           tiling=(tile_j, chunk_x))
    assert not report.adjoint_two_step
    for a, b in zip(two["adjoint_state"], one["adjoint_state"]):
        assert torch.equal(a, b)
        assert a.abs().max().item() > 0
    assert torch.equal(two["signal"], one["signal"])
    got, ref = np.atleast_1d(two["admittance"]), np.atleast_1d(one["admittance"])
    assert np.all(np.isfinite(ref)) and np.abs(ref).max() > 0
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())
