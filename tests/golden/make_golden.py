"""Write the golden vectors in ``reference_vectors.npz`` from the reference implementation.

Runs only where the reference tree exists (``REFERENCE`` env var, default ``/root/reference``),
never on the GPU box and never from the test suite: the suite reads the committed ``.npz``.

What it records, and why these need the reference itself:

* host coefficient builders whose exact fp64 values the reference's own tests pin only to
  1-5 %: ``Medium.air``, ``AirAbsorption.processes`` / ``attenuation``, ``layer_factors``,
  ``wall_update_coefficients``, ``band_limited_impulse``, ``gaussian_pulse``,
  ``Grid.from_max_frequency`` / ``cell_index``;
* the NumPy scheme's outputs on the reference's five backend-parity configurations
  (``tests/test_c_backend.py:15-29``: 20x17x15 cells, seeded initial pressure, Gaussian source,
  150 steps) in fp64 and, for the combined one, fp32 — the exact association the reference
  compiled C does not share;
* the receiver signal of the 5 x 4 x 3 m room (absorption 0.15, 1 kHz grid, air at 20 C / 50 %,
  band-limited impulse at (1.0, 1.0, 1.2) m, receiver at (3.5, 2.5, 1.4) m, 1000 steps).

Usage: ``python tests/golden/make_golden.py``
"""

from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

REFERENCE = Path(os.environ.get("REFERENCE", "/root/reference"))
sys.path.insert(0, str(REFERENCE / "src"))

from ac_fdtd.air import AirAbsorption  # noqa: E402
from ac_fdtd.boundaries import (  # noqa: E402
    AbsorbingLayer,
    WallAdmittances,
    layer_factors,
    wall_update_coefficients,
)
from ac_fdtd.grid import Grid  # noqa: E402
from ac_fdtd.medium import AIR, Medium  # noqa: E402
from ac_fdtd.scheme import AcousticFDTD  # noqa: E402
from ac_fdtd.sources import band_limited_impulse, gaussian_pulse  # noqa: E402

OUT = Path(__file__).resolve().parent / "reference_vectors.npz"

PARITY_GRID = Grid.from_lengths((0.6, 0.5, 0.45), 0.03)
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


def parity_run(name, dtype=np.float64):
    solver = AcousticFDTD(PARITY_GRID, dtype=dtype, **PARITY[name])
    solver.p[:] = np.random.default_rng(0).standard_normal(PARITY_GRID.shape)
    solver.add_volume_source((0.2, 0.2, 0.2), np.exp(-(((np.arange(200) - 30) / 8.0) ** 2)))
    solver.add_pressure_receiver((0.4, 0.3, 0.3))
    solver.run(150)
    return solver


def main() -> None:
    out: dict[str, np.ndarray] = {}

    # Medium
    for temperature in (15.0, 20.0, 25.0):
        medium = Medium.air(temperature)
        out[f"medium_{temperature:g}"] = np.array([medium.sound_speed, medium.density])

    # Air absorption
    frequencies = np.array([63.0, 125.0, 250.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0])
    out["air_frequencies"] = frequencies
    for temperature, humidity, dt in ((20.0, 50.0, 5.773502691896258e-05), (15.0, 70.0, 3.4e-6),
                                      (25.0, 30.0, 1.6823e-05), (20.0, 30.0, 1e-5)):
        air = AirAbsorption(temperature=temperature, relative_humidity=humidity)
        key = f"air_{temperature:g}_{humidity:g}"
        out[key + "_processes"] = np.array(
            [(q.strength, q.relaxation_time) for q in air.processes(AIR.sound_speed, dt)]
        )
        out[key + "_dt"] = np.array(dt)
        out[key + "_attenuation"] = air.attenuation(frequencies)
        out[key + "_h"] = np.array(
            [air.molar_water_concentration, air.oxygen_relaxation_frequency,
             air.nitrogen_relaxation_frequency]
        )

    # Sponge profiles on a grid that is not a cube, layer on a subset of faces and on all faces
    grid = Grid.from_lengths((0.6, 0.5, 0.45), 0.03)
    dt = grid.dx / (AIR.sound_speed * np.sqrt(3))
    for tag, layer in (("all", AbsorbingLayer(thickness=4)),
                       ("subset", AbsorbingLayer(thickness=3, target_reflection=1e-3, order=2.0,
                                                 faces=((0, 1), (2, 0))))):
        pressure, velocity = layer_factors(grid, AIR, dt, layer)
        for axis in range(3):
            out[f"layer_{tag}_p{axis}"] = pressure[axis]
            out[f"layer_{tag}_v{axis}"] = velocity[axis][axis]

    # Walls
    indices, coefficient = wall_update_coefficients(
        grid, AIR, dt, WallAdmittances(x_min=0.2, x_max=0.05, y_min=0.3, z_max=0.7)
    )
    out["walls_indices"] = indices
    out["walls_coefficient"] = coefficient
    duct = Grid(dx=0.01, nx=7, ny=1, nz=3)
    indices, coefficient = wall_update_coefficients(duct, AIR, 1e-5, WallAdmittances.uniform(0.4))
    out["walls_duct_indices"] = indices
    out["walls_duct_coefficient"] = coefficient

    # Signals
    room_grid = Grid.from_max_frequency((5.0, 4.0, 3.0), 1000.0, AIR.sound_speed)
    room_dt = room_grid.dx / (AIR.sound_speed * np.sqrt(3))
    out["band_limited_impulse"] = band_limited_impulse(1000, room_dt, 1000.0)
    out["gaussian_pulse"] = gaussian_pulse(300, room_dt, 1000.0)

    # Grid
    out["room_grid"] = np.array([room_grid.dx, room_grid.nx, room_grid.ny, room_grid.nz])
    out["room_cells"] = np.array(
        [room_grid.cell_index((1.0, 1.0, 1.2)), room_grid.cell_index((3.5, 2.5, 1.4))]
    )

    # Parity configurations
    for name in PARITY:
        solver = parity_run(name)
        out[f"parity_{name}_p"] = solver.p
        out[f"parity_{name}_recording"] = solver.recorded_pressure
        out[f"parity_{name}_energy"] = np.array(solver.energy())
        if name == "all":
            for axis in range(3):
                out[f"parity_{name}_v{axis}"] = solver.velocity[axis]
    single = parity_run("all", np.float32)
    out["parity_all_f32_recording"] = single.recorded_pressure
    out["parity_all_f32_p"] = single.p

    # The 5 x 4 x 3 m room
    solver = AcousticFDTD(
        room_grid,
        walls=WallAdmittances.from_absorption(0.15),
        air_absorption=AirAbsorption(temperature=20.0, relative_humidity=50.0),
    )
    solver.add_volume_source((1.0, 1.0, 1.2), band_limited_impulse(1000, solver.dt, 1000.0))
    solver.add_pressure_receiver((3.5, 2.5, 1.4))
    solver.run(1000)
    out["room_recording"] = solver.recorded_pressure
    out["room_energy"] = np.array(solver.energy())

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
