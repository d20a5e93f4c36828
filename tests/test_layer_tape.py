"""The layer tape and the sponge-strength gradient that reads it.

The forward kernels write, after every step, the fields of the absorbing layer's cells (the
pressure and each cell's lower face velocities, and the far faces of the grid); the two-step
adjoint reads them for the sponge-strength gradient instead of forward states recomputed from
checkpoints.

* Unpacked into full fields, each row is the oracle's state after that step, bit for bit, at
  every cell whose grading weight is nonzero and on every face whose summed face weight is
  nonzero (what the gradient multiplies), and zero or the oracle's value elsewhere, on grids
  whose nz selects each kernel family (two steps per sweep on the widest lanes; one step per
  sweep, copied out after each step, on narrower ones); writing it leaves the run's fields and
  recordings unchanged.
* The sponge, air and wall gradients from the layer tape (two adjoint steps per sweep, the last
  one to four steps from unpacked states) and from recomputed forward states (one step per
  sweep) are the same sums in other orders. Against autograd through the torch restatement the
  same path is checked in ``test_gradients.py``.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from cases import PARITY

from ac_fdtd_amd import AbsorbingLayer, Grid, WallAdmittances
from ac_fdtd_amd.boundaries import layer_weights
from oracle.scheme import OracleFDTD

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

LAYERS = {
    "all-faces": AbsorbingLayer(thickness=3),
    "two-faces": AbsorbingLayer(thickness=4, faces=((0, 0), (2, 1))),
}


def _unpack(solver, step):
    """Row ``step`` of the solver's layer tape as full p, vx, vy, vz arrays."""
    from ac_fdtd_amd import _native

    g = solver.grid
    opts = {"dtype": solver.dtype, "device": solver.device}
    fields = [torch.full(g.shape, float("nan"), **opts)] + [
        torch.full(g.velocity_shape(a), float("nan"), **opts) for a in range(3)]
    desc = solver._descriptor(None)
    stream = torch.cuda.current_stream(solver.device).cuda_stream
    _native.check(solver._lib.acfdtd_hip_layer_tape_unpack(
        ctypes.byref(desc), step, *(ctypes.c_void_p(t.data_ptr()) for t in fields),
        ctypes.c_void_p(stream)), "acfdtd_hip_layer_tape_unpack")
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in fields]


def _needed(grid, layer):
    """Where the sponge gradient reads forward values: cells with a nonzero summed cell weight,
    faces with a nonzero summed face weight."""
    cells, faces = layer_weights(grid, layer)
    wx, wy, wz = (np.asarray(w) for w in cells)
    fx, fy, fz = (np.asarray(w) for w in faces)
    p = (wx[:, None, None] + wy[None, :, None] + wz[None, None, :]) > 0
    vx = (fx[:, None, None] + wy[None, :, None] + wz[None, None, :]) > 0
    vy = (wx[:, None, None] + fy[None, :, None] + wz[None, None, :]) > 0
    vz = (wx[:, None, None] + wy[None, :, None] + fz[None, None, :]) > 0
    return [p, vx, vy, vz]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(14, 11, 16), (12, 9, 14), (10, 9, 7)],
                         ids=["nz16", "nz14", "nz7"])
@pytest.mark.parametrize("layer", list(LAYERS), ids=list(LAYERS))
def test_layer_tape_rows_are_the_oracles(dtype, shape, layer):
    from ac_fdtd_amd.scheme import AcousticFDTD

    grid = Grid(dx=0.04, nx=shape[0], ny=shape[1], nz=shape[2])
    config = dict(PARITY["all"])
    config["absorbing_layer"] = LAYERS[layer]
    rng = np.random.default_rng(5)
    initial = rng.standard_normal(grid.shape)
    signal = rng.standard_normal(30)
    n_steps = 11  # two sweeps of two steps, then single steps

    def setup(solver):
        if hasattr(solver.p, "copy_"):
            solver.p.copy_(torch.as_tensor(initial))
        else:
            solver.p[:] = initial
        solver.add_volume_source((0.2, 0.18, 0.15), signal)
        solver.add_pressure_receiver((0.3, 0.2, 0.2))
        return solver

    taped = setup(AcousticFDTD(grid, dtype=dtype, **config))
    plain = setup(AcousticFDTD(grid, dtype=dtype, **config))
    oracle = setup(OracleFDTD(grid, dtype=dtype, **config))
    tape = torch.full((n_steps + 2, taped.layer_tape_stride()), float("nan"), dtype=taped.dtype,
                      device=taped.device)
    taped.attach_layer_tape(tape, 3)  # rows for steps 3 .. 15
    oracle.run(3)
    state = (oracle.p, *oracle.velocity, np.stack(oracle.relaxation_states))
    taped.load_state([torch.as_tensor(x) for x in state], 3)
    plain.load_state(taped.save_state(), 3)
    taped.run(n_steps)
    plain.run(n_steps)
    needed = _needed(grid, LAYERS[layer])
    for n in range(n_steps):
        oracle.step()
        rows = _unpack(taped, 3 + n)
        for name, got, want, need in zip(("p", "vx", "vy", "vz"), rows,
                                         (oracle.p, *oracle.velocity), needed):
            msg = f"step {3 + n} {name}"
            np.testing.assert_array_equal(got[need], want[need], err_msg=msg)
            assert np.all((got == 0) | (got == want)), msg
        # The far faces are always held.
        np.testing.assert_array_equal(rows[1][-1], oracle.velocity[0][-1])
        np.testing.assert_array_equal(rows[2][:, -1], oracle.velocity[1][:, -1])
        np.testing.assert_array_equal(rows[3][..., -1], oracle.velocity[2][..., -1])
    assert torch.isnan(tape[n_steps:]).all()  # rows past the run untouched
    for a, b in zip(taped.save_state(), plain.save_state()):
        assert torch.equal(a, b)
    np.testing.assert_array_equal(taped.recorded_pressure, plain.recorded_pressure)


def _room(dtype, config, n_steps, **kwargs):
    from ac_fdtd_amd.autograd import RoomResponse

    grid = Grid(dx=0.04, nx=20, ny=14, nz=12)
    rng = np.random.default_rng(3)
    sources = [((0.31, 0.27, 0.22), rng.standard_normal(45))]
    receivers = [(0.12, 0.41, 0.3), (0.5, 0.2, 0.11)]
    return RoomResponse(grid, n_steps, sources, receivers, dtype=dtype, checkpoint_every=17,
                        **config, **kwargs)


def _assert_close(got, want, rtol):
    got = got.detach().double().cpu().reshape(-1)
    want = want.detach().double().cpu().reshape(-1)
    scale = want.abs().max().item()
    assert scale > 0
    assert (got - want).abs().max().item() <= rtol * scale, (got, want)


CONFIGS = {
    "all": dict(PARITY["all"]),
    "walls+layer": {"walls": WallAdmittances(0.1, 0.0, 0.3, 0.2, 0.0, 0.05),
                    "absorbing_layer": AbsorbingLayer(thickness=3, faces=((0, 1), (1, 0), (2, 0)))},
}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("config", list(CONFIGS), ids=list(CONFIGS))
@pytest.mark.parametrize("n_steps", [89, 90, 92])  # 1, 2 and 4 steps after the groups of four
def test_sponge_gradient_paths_agree(dtype, config, n_steps):
    """Layer tape (two adjoint steps per sweep) and recomputed states (one per sweep) give the
    same sponge, wall and air gradients and the same recording."""
    weights = torch.as_tensor(np.random.default_rng(2).standard_normal((2, n_steps))).cuda()
    cfg = CONFIGS[config]
    itemsize = np.dtype(dtype).itemsize

    def grads(room):
        params = {"admittance": room.default_admittance(device="cuda").requires_grad_(),
                  "sponge_strength": room.default_sponge_strength(device="cuda").requires_grad_()}
        if cfg.get("air_absorption") is not None:
            params["air_strength"] = room.default_air_strength(device="cuda").requires_grad_()
        rec = room(**params)
        return rec.detach(), torch.autograd.grad((rec * weights.to(rec.dtype)).sum(),
                                                 list(params.values()))

    room = _room(dtype, cfg, n_steps)
    rec_tape, g_tape = grads(room)
    report = room.report
    assert report.recomputed_steps == 0 and report.checkpoints == 0
    assert report.adjoint_two_step
    solver = room.solver()
    assert report.layer_tape_bytes == n_steps * solver.layer_tape_stride() * itemsize

    room = _room(dtype, cfg, n_steps, history_bytes=0)
    rec_states, g_states = grads(room)
    assert room.report.recomputed_steps == n_steps and not room.report.adjoint_two_step

    assert torch.equal(rec_tape, rec_states)
    for got, want in zip(g_tape, g_states):
        _assert_close(got, want, 1e-11)
