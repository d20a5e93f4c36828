"""Configurations shared by the CPU and GPU tests.

``PARITY`` are the reference's five backend-parity configurations
(``tests/test_c_backend.py:15-29`` and ``tests/test_torch_backend.py:23-37``): a 20 x 17 x 15
grid, a seeded standard-normal initial pressure, a Gaussian source at (0.2, 0.2, 0.2) m and a
receiver at (0.4, 0.3, 0.3) m, run for 150 steps.
"""

from __future__ import annotations

import numpy as np

from ac_fdtd_amd import AbsorbingLayer, AirAbsorption, Grid, WallAdmittances

PARITY_GRID = Grid.from_lengths((0.6, 0.5, 0.45), 0.03)
PARITY_SOURCE = (0.2, 0.2, 0.2)
PARITY_RECEIVER = (0.4, 0.3, 0.3)
PARITY_STEPS = 150

PARITY = {
    "rigid": {},
    "walls": {"walls": WallAdmittances.uniform(0.3)},
    "layer": {"absorbing_layer": AbsorbingLayer(thickness=4)},
    "air": {"air_absorption": AirAbsorption()},
    "all": {
        "walls": WallAdmittances(x_min=0.2, z_max=0.7),
        "absorbing_layer": AbsorbingLayer(thickness=3),
        "air_absorption": AirAbsorption(temperature=15.0, relative_humidity=70.0),
    },
}


def parity_signal() -> np.ndarray:
    return np.exp(-(((np.arange(200) - 30) / 8.0) ** 2))


def parity_initial(grid=PARITY_GRID, seed=0) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal(grid.shape)


def prepare(solver, grid=PARITY_GRID, seed=0):
    """Seeded initial pressure, the Gaussian source and the receiver, on any backend."""
    initial = parity_initial(grid, seed)
    if hasattr(solver.p, "copy_"):
        import torch

        solver.p.copy_(torch.as_tensor(initial, dtype=solver.p.dtype))
    else:
        solver.p[:] = initial
    solver.add_volume_source(PARITY_SOURCE, parity_signal())
    solver.add_pressure_receiver(PARITY_RECEIVER)
    return solver


def to_numpy(x):
    return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)
