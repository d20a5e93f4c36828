"""Gradients of a receiver-signal loss through the hand-written adjoint.

Oracles:

* ``oracle/torch_scheme.py``, a differentiable restatement of the reference scheme, pinned
  bit-for-bit to the NumPy oracle at default parameters (CPU tests below), differentiated by
  ``torch.autograd`` in fp64;
* central finite differences of the HIP fp64 forward pass.

Tolerances (fp64): adjoint vs autograd 1e-9 relative (both exact up to round-off); adjoint vs
finite differences 1e-6 relative. The differences are central differences at steps h and h/2
(h = 3e-2 of the parameter) combined by Richardson extrapolation, (4 D(h/2) - D(h)) / 3, which
cancels the O(h^2) truncation term; much smaller steps are limited by round-off, because a
wall's effect on the loss is small next to the loss itself.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from cases import PARITY, PARITY_GRID, PARITY_RECEIVER, PARITY_SOURCE, parity_initial, parity_signal

from ac_fdtd_amd import AIR, AbsorbingLayer, AirAbsorption, Grid, WallAdmittances
from ac_fdtd_amd.coefficients import build_coefficients
from oracle.scheme import OracleFDTD
from oracle.torch_scheme import TorchScheme

F64 = torch.float64


def richardson(loss, base: torch.Tensor, index: int, rel: float = 3e-2) -> float:
    """d loss / d base[index] by extrapolated central differences (see module docstring)."""
    h = rel * (abs(float(base.detach().reshape(-1)[index])) or 1e-2)

    def central(step):
        plus, minus = base.detach().clone(), base.detach().clone()
        plus.reshape(-1)[index] += step
        minus.reshape(-1)[index] -= step
        return (float(loss(plus)) - float(loss(minus))) / (2 * step)

    return (4.0 * central(h / 2) - central(h)) / 3.0


def _defaults(grid, config):
    k = build_coefficients(grid, AIR, 1.0, np.float64, config.get("walls"),
                           config.get("absorbing_layer"), config.get("air_absorption"))
    xi = torch.tensor((config.get("walls") or WallAdmittances()).values, dtype=F64)
    sigma = torch.tensor(k.sponge_strength, dtype=F64) if k.sponge_strength else None
    air = torch.tensor([q.strength for q in k.relaxation], dtype=F64) if k.relaxation else None
    return xi, sigma, air


# -- CPU: pin the differentiable oracle --------------------------------------------------------
@pytest.mark.parametrize("name", list(PARITY))
def test_torch_restatement_is_bitwise_the_oracle(name):
    config = PARITY[name]
    reference = OracleFDTD(PARITY_GRID, **config)
    reference.p[:] = parity_initial()
    reference.add_volume_source(PARITY_SOURCE, parity_signal())
    reference.add_pressure_receiver(PARITY_RECEIVER)
    reference.run(120)
    scheme = TorchScheme(PARITY_GRID, AIR, absorbing_layer=config.get("absorbing_layer"),
                         air_absorption=config.get("air_absorption"),
                         sources=[PARITY_SOURCE], receivers=[PARITY_RECEIVER])
    xi, sigma, air = _defaults(PARITY_GRID, config)
    signal = torch.tensor(parity_signal(), dtype=F64)[None]
    recording = scheme.simulate(120, xi, signal, sigma, air, initial_p=parity_initial())
    assert np.array_equal(recording.numpy(), reference.recorded_pressure)


def test_torch_restatement_gradient_matches_its_own_differences():
    """Autograd on the restatement agrees with central differences of the same restatement:
    the oracle's derivative is the derivative of the function it computes."""
    grid = Grid(dx=0.05, nx=9, ny=7, nz=6)
    config = PARITY["all"]
    scheme = TorchScheme(grid, AIR, absorbing_layer=config["absorbing_layer"],
                         air_absorption=config["air_absorption"],
                         sources=[(0.2, 0.15, 0.1)], receivers=[(0.3, 0.2, 0.2)])
    xi, sigma, air = _defaults(grid, config)
    signal = torch.tensor(parity_signal()[:60], dtype=F64)[None]
    weights = torch.as_tensor(np.random.default_rng(1).standard_normal((1, 60)))

    def loss(x, s, a):
        return (scheme.simulate(60, x, signal, s, a) * weights).sum()

    params = [xi.clone().requires_grad_(), sigma.clone().requires_grad_(),
              air.clone().requires_grad_()]
    grads = torch.autograd.grad(loss(*params), params)
    for which, (base, grad) in enumerate(zip(params, grads)):
        for e in range(base.numel()):

            def partial(value, which=which):
                args = [p.detach() for p in params]
                args[which] = value
                return loss(*args)

            fd = richardson(partial, base, e)
            assert float(grad.reshape(-1)[e]) == pytest.approx(fd, rel=1e-6)


# -- GPU: the adjoint ---------------------------------------------------------------------------
def _room(grid, config, n_steps, dtype=np.float64, checkpoint_every=40, **kwargs):
    from ac_fdtd_amd.autograd import RoomResponse

    rng = np.random.default_rng(3)
    sources = [((0.31, 0.27, 0.22), rng.standard_normal(n_steps // 2))]
    receivers = [(0.12, 0.41, 0.3), (0.5, 0.2, 0.11)]
    return RoomResponse(grid, n_steps, sources, receivers, dtype=dtype,
                        checkpoint_every=checkpoint_every, **config, **kwargs), sources, receivers


def _oracle_grads(room, config, sources, receivers, weights, device="cuda"):
    scheme = TorchScheme(room.grid, AIR, absorbing_layer=config.get("absorbing_layer"),
                         air_absorption=config.get("air_absorption"),
                         sources=[p for p, _ in sources], receivers=receivers, device=device)
    params = {
        "admittance": room.default_admittance(device=device).requires_grad_(),
        "signal": room.default_signal(device=device).requires_grad_(),
    }
    if config.get("absorbing_layer") is not None:
        params["sponge_strength"] = room.default_sponge_strength(device=device).requires_grad_()
    if config.get("air_absorption") is not None:
        params["air_strength"] = room.default_air_strength(device=device).requires_grad_()
    rec = scheme.simulate(room.n_steps, params["admittance"], params["signal"],
                          params.get("sponge_strength"), params.get("air_strength"))
    loss = (rec * weights.to(device)).sum()
    grads = torch.autograd.grad(loss, list(params.values()))
    return rec.detach(), dict(zip(params, grads))


def _adjoint_grads(room, config, weights):
    params = {
        "admittance": room.default_admittance(device="cuda").requires_grad_(),
        "signal": room.default_signal(device="cuda").requires_grad_(),
    }
    if config.get("absorbing_layer") is not None:
        params["sponge_strength"] = room.default_sponge_strength(device="cuda").requires_grad_()
    if config.get("air_absorption") is not None:
        params["air_strength"] = room.default_air_strength(device="cuda").requires_grad_()
    rec = room(**params)
    loss = (rec * weights.to(rec.device, rec.dtype)).sum()
    grads = torch.autograd.grad(loss, list(params.values()))
    return rec.detach(), dict(zip(params, grads))


def _assert_close(got, want, rtol):
    got = got.detach().double().cpu().reshape(-1)
    want = want.detach().double().cpu().reshape(-1)
    scale = want.abs().max().item()
    assert scale > 0
    assert (got - want).abs().max().item() <= rtol * scale, (got, want)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
@pytest.mark.parametrize("name", ["walls", "all", "air", "layer"])
@pytest.mark.parametrize("fused", [True, False], ids=["fused", "two-sweep"])
@pytest.mark.parametrize("nz", [11, 12])
def test_adjoint_matches_autograd(name, fused, nz):
    """Every parameter gradient against torch.autograd through the restatement, fp64. With
    nz = 12 the fused path runs two steps per sweep, forward (with the divergence tape when
    the air strengths are asked for without the sponge) and adjoint; with nz = 11 one."""
    grid = Grid(dx=0.04, nx=16, ny=13, nz=nz)
    config = dict(PARITY[name])
    if name == "layer":
        config["walls"] = WallAdmittances(0.1, 0.0, 0.3, 0.2, 0.0, 0.05)
    room, sources, receivers = _room(grid, config, 120, fused=fused)
    weights = torch.as_tensor(np.random.default_rng(9).standard_normal((2, 120)))
    rec_o, g_o = _oracle_grads(room, config, sources, receivers, weights)
    rec_a, g_a = _adjoint_grads(room, config, weights)
    _assert_close(rec_a, rec_o, 1e-12)
    for key in g_o:
        _assert_close(g_a[key], g_o[key], 1e-9)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_adjoint_at_64_cubed_matches_autograd_and_central_differences():
    """64^3 fp64: walls + sponge + air, wall admittances checked against autograd through the
    restatement and against central differences of the HIP forward pass (1e-6)."""
    grid = Grid(dx=0.02, nx=64, ny=64, nz=64)
    config = {
        "walls": WallAdmittances(0.2, 0.1, 0.3, 0.05, 0.15, 0.25),
        "absorbing_layer": AbsorbingLayer(thickness=6, faces=((0, 1), (2, 0))),
        "air_absorption": AirAbsorption(),
    }
    n_steps = 160
    room, sources, receivers = _room(grid, config, n_steps)
    weights = torch.as_tensor(np.random.default_rng(4).standard_normal((2, n_steps)))
    _, g_o = _oracle_grads(room, config, sources, receivers, weights)
    _, g_a = _adjoint_grads(room, config, weights)
    for key in g_o:
        _assert_close(g_a[key], g_o[key], 1e-9)

    # Central differences of an L2 misfit against the response of hidden admittances.
    xi = room.default_admittance(device="cuda")
    with torch.no_grad():
        target = room(admittance=torch.tensor([0.3, 0.05, 0.2, 0.1, 0.25, 0.4], dtype=F64))

    def loss(values):
        with torch.no_grad():
            return float(((room(admittance=values) - target) ** 2).sum())

    probe = xi.clone().requires_grad_()
    (g_a["admittance"],) = torch.autograd.grad(((room(admittance=probe) - target) ** 2).sum(), [probe])

    for f in range(6):
        fd = richardson(loss, xi, f)
        assert float(g_a["admittance"][f]) == pytest.approx(fd, rel=1e-6)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_checkpointed_and_taped_wall_gradients_agree():
    """The wall gradient from the boundary tape alone and from the recompute path (requested
    together with the air strengths) are the same adjoint: equal to round-off."""
    grid = Grid(dx=0.04, nx=20, ny=14, nz=12)
    config = dict(PARITY["all"])
    room, _, _ = _room(grid, config, 90, checkpoint_every=17)
    weights = torch.as_tensor(np.random.default_rng(2).standard_normal((2, 90))).cuda()
    xi = room.default_admittance(device="cuda").requires_grad_()
    (g_tape,) = torch.autograd.grad((room(admittance=xi) * weights).sum(), [xi])
    assert room.report.recomputed_steps == 0
    xi2 = room.default_admittance(device="cuda").requires_grad_()
    air = room.default_air_strength(device="cuda").requires_grad_()
    room.history_bytes = 0  # the air gradients from checkpoints and recomputed segments
    g_both = torch.autograd.grad((room(admittance=xi2, air_strength=air) * weights).sum(), [xi2, air])
    assert room.report.recomputed_steps == 90 and room.report.checkpoints == 6
    _assert_close(g_both[0], g_tape, 1e-12)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_fp32_adjoint_tracks_fp64():
    grid = Grid(dx=0.04, nx=24, ny=18, nz=16)
    config = dict(PARITY["all"])
    weights = torch.as_tensor(np.random.default_rng(8).standard_normal((2, 200)))
    grads = {}
    for dtype in (np.float64, np.float32):
        room, _, _ = _room(grid, config, 200, dtype=dtype)
        _, grads[dtype] = _adjoint_grads(room, config, weights)
    for key in grads[np.float64]:
        _assert_close(grads[np.float32][key], grads[np.float64][key], 2e-3)


@pytest.mark.gpu
@pytest.mark.usefixtures("require_gpu")
def test_gradient_descent_recovers_hidden_wall_admittances():
    """The use the adjoint is for: fit the six wall admittances of a room to a target response
    simulated from hidden values, starting from uniform walls. 80 Adam steps on the L2 misfit
    of six receivers (fp64) bring every admittance within 0.01 of its hidden value."""
    from ac_fdtd_amd.autograd import RoomResponse

    grid = Grid(dx=0.03, nx=24, ny=20, nz=16)
    rng = np.random.default_rng(3)
    sources = [((0.31, 0.27, 0.22), rng.standard_normal(60))]
    receivers = [(0.1, 0.1, 0.1), (0.6, 0.5, 0.4), (0.1, 0.5, 0.3), (0.6, 0.1, 0.2),
                 (0.35, 0.3, 0.05), (0.35, 0.3, 0.42)]
    room = RoomResponse(grid, 240, sources, receivers, dtype=np.float64,
                        walls=WallAdmittances.uniform(0.2))
    hidden = torch.tensor([0.3, 0.05, 0.2, 0.1, 0.25, 0.4], dtype=F64, device="cuda")
    with torch.no_grad():
        target = room(admittance=hidden)
    xi = room.default_admittance(device="cuda").clone().requires_grad_()
    optimizer = torch.optim.Adam([xi], lr=0.02)
    losses = []
    for _ in range(80):
        optimizer.zero_grad()
        loss = ((room(admittance=xi) - target) ** 2).sum()
        loss.backward()
        optimizer.step()
        with torch.no_grad():
            xi.clamp_(min=0.0)
        losses.append(float(loss.detach()))
    assert losses[-1] < 1e-3 * losses[0]
    assert float((xi.detach() - hidden).abs().max()) < 0.01
