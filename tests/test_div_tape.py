"""The divergence tape and the air-strength gradients that read it.

The forward kernels write every cell's velocity divergence of each step to the tape (the value
its pressure update uses); the adjoint reads it in place of recomputed forward velocities.

* The tape's rows are the oracle's divergences bit for bit (``OracleFDTD._divergence``, the
  ``((dvx + dvy) + dvz)`` its pressure update of that step used, in the working precision),
  on grids whose nz selects each kernel family (two steps per sweep on the
  widest lanes, one step per sweep on narrower ones, the in-place sweeps for an odd last step),
  and writing it leaves the run's fields and recordings unchanged.
* The air gradients from the whole-run tape, from tapes recomputed per checkpoint segment, and
  from recomputed forward states (the sponge's path beyond the history budget) are the same
  sums in other orders.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from cases import PARITY

from ac_fdtd_amd import Grid, WallAdmittances
from oracle.scheme import OracleFDTD

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]


def _divergence(oracle: OracleFDTD) -> np.ndarray:
    """The divergence the oracle's last step used: that of its current velocities."""
    return oracle._divergence().ravel()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(14, 11, 16), (12, 9, 14), (10, 9, 7)],
                         ids=["nz16", "nz14", "nz7"])
@pytest.mark.parametrize("staged", [False, True], ids=["plain", "staged"])
def test_divergence_tape_rows_are_the_oracles(dtype, shape, staged):
    from ac_fdtd_amd.scheme import AcousticFDTD

    grid = Grid(dx=0.04, nx=shape[0], ny=shape[1], nz=shape[2])
    config = dict(PARITY["all"])
    rng = np.random.default_rng(5)
    initial = rng.standard_normal(grid.shape)
    signal = rng.standard_normal(30)
    n_steps = 11  # two sweeps of two steps, a pair of single steps, an odd in-place step

    def setup(solver):
        if hasattr(solver.p, "copy_"):
            solver.p.copy_(torch.as_tensor(initial))
        else:
            solver.p[:] = initial
        solver.add_volume_source((0.2, 0.18, 0.15), signal)
        solver.add_pressure_receiver((0.3, 0.2, 0.2))
        return solver

    taped = setup(AcousticFDTD(grid, dtype=dtype, staged=staged, **config))
    plain = setup(AcousticFDTD(grid, dtype=dtype, staged=staged, **config))
    oracle = setup(OracleFDTD(grid, dtype=dtype, **config))
    tape = torch.full((n_steps + 2, grid.n_cells), float("nan"), dtype=taped.dtype,
                      device=taped.device)
    taped.attach_div_tape(tape, 3)  # rows for steps 3 .. 15
    oracle.run(3)
    state = (oracle.p, *oracle.velocity, np.stack(oracle.relaxation_states))
    taped.load_state([torch.as_tensor(x) for x in state], 3)
    plain.load_state(taped.save_state(), 3)
    taped.run(n_steps)
    plain.run(n_steps)
    rows = tape.cpu().numpy()
    for n in range(n_steps):
        oracle.step()
        np.testing.assert_array_equal(rows[n], _divergence(oracle), err_msg=f"step {3 + n}")
    assert np.isnan(rows[n_steps:]).all()  # rows past the run untouched
    for a, b in zip(taped.save_state(), plain.save_state()):
        assert torch.equal(a, b)
    np.testing.assert_array_equal(taped.recorded_pressure, plain.recorded_pressure)


def _room(dtype, **kwargs):
    from ac_fdtd_amd.autograd import RoomResponse

    grid = Grid(dx=0.04, nx=20, ny=14, nz=12)
    config = dict(PARITY["all"])
    config["walls"] = WallAdmittances(0.2, 0.05, 0.1, 0.3, 0.15, 0.25)
    rng = np.random.default_rng(3)
    sources = [((0.31, 0.27, 0.22), rng.standard_normal(45))]
    receivers = [(0.12, 0.41, 0.3), (0.5, 0.2, 0.11)]
    return RoomResponse(grid, 90, sources, receivers, dtype=dtype, checkpoint_every=17,
                        **config, **kwargs)


def _assert_close(got, want, rtol):
    got = got.detach().double().cpu().reshape(-1)
    want = want.detach().double().cpu().reshape(-1)
    scale = want.abs().max().item()
    assert scale > 0
    assert (got - want).abs().max().item() <= rtol * scale, (got, want)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_air_gradient_paths_agree(dtype):
    """Whole-run tape, per-segment tapes and recomputed states give the same air and wall
    gradients; the tape paths run two adjoint steps per sweep."""
    weights = torch.as_tensor(np.random.default_rng(2).standard_normal((2, 90))).cuda()
    n_cells = 20 * 14 * 12
    itemsize = np.dtype(dtype).itemsize

    def grads(room, with_sponge=False):
        params = [room.default_admittance(device="cuda").requires_grad_(),
                  room.default_air_strength(device="cuda").requires_grad_()]
        kwargs = {"admittance": params[0], "air_strength": params[1]}
        if with_sponge:
            params.append(room.default_sponge_strength(device="cuda").requires_grad_())
            kwargs["sponge_strength"] = params[2]
        rec = room(**kwargs)
        return torch.autograd.grad((rec * weights.to(rec.dtype)).sum(), params)

    room = _room(dtype)
    g_tape = grads(room)
    report = room.report
    assert report.recomputed_steps == 0 and report.checkpoints == 0
    assert report.div_tape_bytes == 90 * n_cells * itemsize
    assert report.adjoint_two_step

    room = _room(dtype, history_bytes=0)
    g_seg = grads(room)
    report = room.report
    assert report.recomputed_steps == 90 and report.checkpoints == 6
    assert report.div_tape_bytes == 17 * n_cells * itemsize
    assert report.adjoint_two_step

    room = _room(dtype, history_bytes=0)
    g_states = grads(room, with_sponge=True)
    assert room.report.recomputed_steps == 90 and not room.report.adjoint_two_step

    for other in (g_seg, g_states):
        _assert_close(other[0], g_tape[0], 1e-11)
        _assert_close(other[1], g_tape[1], 1e-11)
