"""Differentiable CPU/GPU restatement of the reference scheme in torch, for checking gradients.

TEST INFRASTRUCTURE (same rule as ``oracle/scheme.py``): only tests, ``smoke()`` and the CPU
baseline of ``bench.py`` may use it, as the checker.

The reference's own torch backend cannot be differentiated: it updates with ``out=`` ops and
takes its coefficients as host floats (``src/ac_fdtd/torch_backend.py:149-195``, ``:278-309``).
This module restates one step of ``src/ac_fdtd/scheme.py:285-363`` with out-of-place torch
operations, with the wall admittances, the sponge strength ``sigma_max``, the relaxation
strengths and the source signals as tensors, so ``torch.autograd`` differentiates it. The
wall closed form is applied on every boundary cell (with ``a = 0`` on rigid faces it is the
identity), which is what makes ``dL/dxi`` defined on rigid faces too.

Pinned against the NumPy oracle: run in fp64 with the parameters at their defaults it gives
the same receiver signal to 1e-13 relative (``tests/test_gradients.py``).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from ac_fdtd_amd.boundaries import FACES, layer_weights
from ac_fdtd_amd.grid import max_stable_time_step

__all__ = ["TorchScheme"]


class TorchScheme:
    """One room; :meth:`simulate` returns the ``(receiver, step)`` recording."""

    def __init__(self, grid, medium, courant=1.0, absorbing_layer=None, air_absorption=None,
                 sources=(), receivers=(), device="cpu", dtype=torch.float64):
        self.grid, self.medium = grid, medium
        self.layer, self.air = absorbing_layer, air_absorption
        self.dt = courant * max_stable_time_step(grid.dx, medium.sound_speed)
        self.opts = {"device": device, "dtype": dtype}
        self.sources = [grid.cell_index(p) for p in sources]
        self.receivers = [grid.cell_index(p) for p in receivers]
        c, rho, dx, dt = medium.sound_speed, medium.density, grid.dx, self.dt
        self.k_vel = dt / (rho * dx)
        self.k_pres = rho * c**2 * dt / dx
        self.k_src = rho * c**2 * dt
        self.wall_scale = c * dt / (2.0 * dx)
        # Boundary-plane masks in FACES order.
        nx, ny, nz = grid.shape
        self.face_masks = []
        for axis, side in FACES:
            mask = torch.zeros(grid.shape, **self.opts)
            index = [slice(None)] * 3
            index[axis] = 0 if side == 0 else grid.shape[axis] - 1
            mask[tuple(index)] = 1.0
            self.face_masks.append(mask)
        self.processes = air_absorption.processes(c, dt) if air_absorption is not None else ()
        if absorbing_layer is not None:
            cells, faces = layer_weights(grid, absorbing_layer)
            self.w_cell = [torch.as_tensor(w, **self.opts) for w in cells]
            self.w_face = [torch.as_tensor(w, **self.opts) for w in faces]

    @staticmethod
    def _axis(values, axis):
        shape = [1, 1, 1]
        shape[axis] = -1
        return values.reshape(shape)

    def simulate(self, n_steps, admittance, signal, sponge_strength=None, air_strength=None,
                 initial_p=None):
        g, dt = self.grid, self.dt
        opts = self.opts
        p = torch.zeros(g.shape, **opts) if initial_p is None else torch.as_tensor(initial_p, **opts)
        v = [torch.zeros(g.velocity_shape(a), **opts) for a in range(3)]

        a_cell = sum(self.wall_scale * admittance[f] * self.face_masks[f] for f in range(6))
        fu = 1.0 / (1.0 + a_cell)
        fp = a_cell / (1.0 + a_cell)

        damp_p = damp_v = None
        if self.layer is not None:
            sigma = sponge_strength
            fc = [torch.exp(-sigma * self.w_cell[a] * dt) for a in range(3)]
            ff = [torch.exp(-sigma * self.w_face[a] * dt) for a in range(3)]
            damp_p = [self._axis(fc[a], a) for a in range(3)]
            damp_v = [[self._axis(ff[a] if a == comp else fc[a], a) for a in range(3)]
                      for comp in range(3)]

        relax = []
        if self.processes:
            strengths = air_strength
            total = 1.0 + strengths.sum()
            c, rho, dx = self.medium.sound_speed, self.medium.density, g.dx
            for r, q in enumerate(self.processes):
                decay = math.exp(-dt / q.relaxation_time)
                mean = dt * ((q.relaxation_time / dt) * (1.0 - decay))
                gain = rho * (c**2 / total) * strengths[r] / dx
                relax.append((decay, mean, gain))
            scale = dt * sum(gain for _, _, gain in relax)
            psi = [torch.zeros(g.shape, **opts) for _ in relax]

        rows = []
        for n in range(n_steps):
            new_v = []
            for a in range(3):
                hi = [slice(None)] * 3
                lo = [slice(None)] * 3
                hi[a], lo[a] = slice(1, None), slice(None, -1)
                grad = p[tuple(hi)] - p[tuple(lo)]
                pad = [0, 0, 0, 0, 0, 0]
                pad[2 * (2 - a)] = pad[2 * (2 - a) + 1] = 1
                update = torch.nn.functional.pad(grad, pad)
                w = v[a] - self.k_vel * update
                if damp_v is not None:
                    w = w * damp_v[a][0] * damp_v[a][1] * damp_v[a][2]
                new_v.append(w)
            v = new_v
            div = (v[0][1:] - v[0][:-1]) + (v[1][:, 1:] - v[1][:, :-1]) + (v[2][:, :, 1:] - v[2][:, :, :-1])
            p_tmp = p
            if relax:
                new_psi = []
                for (decay, mean, gain), state in zip(relax, psi):
                    target = div * gain
                    offset = state - target
                    new_psi.append(offset * decay + target)
                    p_tmp = p_tmp + offset * mean
                psi = new_psi
                p_tmp = p_tmp + div * scale
            p_tmp = p_tmp - div * self.k_pres
            p_new = p_tmp * fu - p * fp
            if damp_p is not None:
                p_new = p_new * damp_p[0] * damp_p[1] * damp_p[2]
            for s, index in enumerate(self.sources):
                if n < signal.shape[1]:
                    bump = torch.zeros(g.shape, **opts)
                    bump[index] = 1.0
                    p_new = p_new + bump * (self.k_src * signal[s, n])
            p = p_new
            rows.append(torch.stack([p[index] for index in self.receivers]))
        return torch.stack(rows, dim=1)
