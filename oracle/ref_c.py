"""Driver for the reference's own compiled C time loop (``oracle/_ref``), for checking only.

TEST INFRASTRUCTURE, under the same rule as ``oracle/scheme.py``: tests, ``smoke()`` and the
CPU-baseline leg of ``bench.py`` only.

``oracle/Makefile`` compiles ``src/ac_fdtd/c/acfdtd.c`` from the reference tree with the
reference's flags. This module is the binding a caller needs: a ctypes mirror of the struct at
``acfdtd.c:56-97`` (field for field, as the reference's own mirror at ``c_backend.py:116-157``
must be) and a restatement of the coefficient assembly of ``c_backend.py:190-237`` and
``:302-356`` — which differs from the NumPy scheme in association (the mean weight is
``(dt * (tau / dt)) * (1 - decay)``, the target scale sums fp64 gains, the velocity sponge is
``(v - c dp) * ((fx fy) fz)``), so it agrees with the oracle to a few ulp rather than bitwise.
"""

from __future__ import annotations

import ctypes
import math
import os
from pathlib import Path

import numpy as np

from ac_fdtd_amd.boundaries import WallAdmittances, layer_factors
from ac_fdtd_amd.grid import Grid, max_stable_time_step
from ac_fdtd_amd.medium import AIR, Medium

from .scheme import sparse_wall_coefficients

__all__ = ["RefCFDTD", "is_available", "library_path"]

_DIR = Path(__file__).resolve().parent / "_ref"
_MAX_RELAX = 4
_libs: dict[str, ctypes.CDLL] = {}


def library_path(dtype) -> Path:
    return _DIR / f"acfdtd_{'f64' if np.dtype(dtype) == np.float64 else 'f32'}.so"


def is_available(dtype=np.float64) -> bool:
    return library_path(dtype).exists()


def _lib(dtype) -> ctypes.CDLL:
    key = np.dtype(dtype).name
    if key not in _libs:
        path = library_path(dtype)
        if not path.exists():
            raise RuntimeError(f"{path} is not built: run `make -C oracle` first")
        lib = ctypes.CDLL(str(path))
        lib.acfdtd_struct_size.restype = ctypes.c_size_t
        lib.acfdtd_is_double.restype = ctypes.c_int
        lib.acfdtd_set_threads.restype = ctypes.c_int
        lib.acfdtd_set_threads.argtypes = [ctypes.c_int]
        lib.acfdtd_thread_count.restype = ctypes.c_int
        lib.acfdtd_run.restype = None
        lib.acfdtd_run.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        _libs[key] = lib
    return _libs[key]


def _struct(real):
    ptr = ctypes.POINTER(real)
    iptr = ctypes.POINTER(ctypes.c_int64)

    class State(ctypes.Structure):
        _fields_ = [
            ("nx", ctypes.c_int32),
            ("ny", ctypes.c_int32),
            ("nz", ctypes.c_int32),
            ("n_relax", ctypes.c_int32),
            ("n_wall", ctypes.c_int64),
            ("n_sources", ctypes.c_int64),
            ("n_receivers", ctypes.c_int64),
            ("n_steps", ctypes.c_int64),
            ("source_length", ctypes.c_int64),
            ("vel_coef", real),
            ("pres_coef", real),
            ("src_coef", real),
            ("relax_scale", real),
            ("p", ptr),
            ("vx", ptr),
            ("vy", ptr),
            ("vz", ptr),
            ("psi", ptr),
            ("relax_decay", real * _MAX_RELAX),
            ("relax_mean", real * _MAX_RELAX),
            ("relax_gain", real * _MAX_RELAX),
            ("cell", ptr * 3),
            ("face", ptr * 3),
            ("has_layer", ctypes.c_int32),
            ("wall_index", iptr),
            ("wall_from_updated", ptr),
            ("wall_from_previous", ptr),
            ("wall_scratch", ptr),
            ("source_index", iptr),
            ("source_signal", ptr),
            ("receiver_index", iptr),
            ("recording", ptr),
        ]

    return State


class RefCFDTD:
    """The reference C loop behind the same constructor / run API as the oracle."""

    def __init__(
        self,
        grid: Grid,
        medium: Medium = AIR,
        courant: float = 1.0,
        dtype=np.float64,
        walls=None,
        absorbing_layer=None,
        air_absorption=None,
        threads: int | None = None,
    ):
        self.grid, self.medium = grid, medium
        self.dtype = np.dtype(dtype)
        self.real = ctypes.c_double if self.dtype == np.float64 else ctypes.c_float
        self.lib = _lib(self.dtype)
        self.State = _struct(self.real)
        if ctypes.sizeof(self.State) != self.lib.acfdtd_struct_size():
            raise RuntimeError("ctypes mirror does not match the reference struct")
        if threads is not None:
            self.lib.acfdtd_set_threads(int(threads))
        self.dt = courant * max_stable_time_step(grid.dx, medium.sound_speed)
        self.step_index = 0
        self.p = np.zeros(grid.shape, dtype=self.dtype)
        self.velocity = [np.zeros(grid.velocity_shape(a), dtype=self.dtype) for a in range(3)]

        procs = air_absorption.processes(medium.sound_speed, self.dt) if air_absorption else ()
        self.psi = np.zeros((len(procs), *grid.shape), dtype=self.dtype)
        c2 = medium.sound_speed**2 / (1.0 + sum(q.strength for q in procs))
        self.decay, self.mean, self.gain = (np.zeros(_MAX_RELAX) for _ in range(3))
        for n, q in enumerate(procs):
            d = math.exp(-self.dt / q.relaxation_time)
            self.decay[n] = d
            self.mean[n] = self.dt * (q.relaxation_time / self.dt) * (1.0 - d)
            self.gain[n] = medium.density * c2 * q.strength / grid.dx
        self.relax_scale = self.dt * float(self.gain.sum())

        idx, a = sparse_wall_coefficients(grid, medium, self.dt, walls or WallAdmittances())
        self.wall_index = np.ascontiguousarray(idx, dtype=np.int64)
        self.wall_fu = (1.0 / (1.0 + a)).astype(self.dtype)
        self.wall_fp = (a / (1.0 + a)).astype(self.dtype)
        self.wall_scratch = np.zeros(idx.size, dtype=self.dtype)

        self.has_layer = absorbing_layer is not None
        if self.has_layer:
            pf, vf = layer_factors(grid, medium, self.dt, absorbing_layer)
            self.cell = [pf[a].astype(self.dtype) for a in range(3)]
            self.face = [vf[a][a].astype(self.dtype) for a in range(3)]
        else:
            self.cell = [np.ones(grid.shape[a], dtype=self.dtype) for a in range(3)]
            self.face = [np.ones(grid.shape[a] + 1, dtype=self.dtype) for a in range(3)]

        self.source_index = np.zeros(0, dtype=np.int64)
        self.source_signal = np.zeros((0, 0), dtype=self.dtype)
        self.receiver_index = np.zeros(0, dtype=np.int64)
        self.recordings: list[np.ndarray] = []

    @property
    def threads(self) -> int:
        return int(self.lib.acfdtd_thread_count())

    def set_threads(self, n: int) -> int:
        return int(self.lib.acfdtd_set_threads(int(n)))

    def add_volume_source(self, point, signal):
        signal = np.asarray(signal, dtype=self.dtype)
        length = max(self.source_signal.shape[1], signal.size)
        padded = np.zeros((self.source_signal.shape[0] + 1, length), dtype=self.dtype)
        padded[:-1, : self.source_signal.shape[1]] = self.source_signal
        padded[-1, : signal.size] = signal
        self.source_index = np.append(self.source_index, self.grid.flat_index(point))
        self.source_signal = padded

    def add_pressure_receiver(self, point):
        self.receiver_index = np.append(self.receiver_index, self.grid.flat_index(point))
        return self.receiver_index.size - 1

    @property
    def recorded_pressure(self):
        if not self.recordings:
            return np.zeros((self.receiver_index.size, 0))
        return np.concatenate(self.recordings, axis=1)

    def _ptr(self, array):
        return array.ctypes.data_as(ctypes.POINTER(self.real))

    def run(self, n_steps: int):
        if n_steps <= 0:
            return
        rec = np.zeros((self.receiver_index.size, n_steps), dtype=self.dtype)
        arr3 = ctypes.POINTER(self.real) * 3
        reals4 = self.real * _MAX_RELAX
        iptr = ctypes.POINTER(ctypes.c_int64)
        m, g = self.medium, self.grid
        state = self.State(
            nx=g.nx,
            ny=g.ny,
            nz=g.nz,
            n_relax=self.psi.shape[0],
            n_wall=self.wall_index.size,
            n_sources=self.source_index.size,
            n_receivers=self.receiver_index.size,
            n_steps=n_steps,
            source_length=self.source_signal.shape[1],
            vel_coef=self.dt / (m.density * g.dx),
            pres_coef=m.density * m.sound_speed**2 * self.dt / g.dx,
            src_coef=m.density * m.sound_speed**2 * self.dt,
            relax_scale=self.relax_scale,
            p=self._ptr(self.p),
            vx=self._ptr(self.velocity[0]),
            vy=self._ptr(self.velocity[1]),
            vz=self._ptr(self.velocity[2]),
            psi=self._ptr(self.psi),
            relax_decay=reals4(*self.decay),
            relax_mean=reals4(*self.mean),
            relax_gain=reals4(*self.gain),
            cell=arr3(*[self._ptr(x) for x in self.cell]),
            face=arr3(*[self._ptr(x) for x in self.face]),
            has_layer=int(self.has_layer),
            wall_index=self.wall_index.ctypes.data_as(iptr),
            wall_from_updated=self._ptr(self.wall_fu),
            wall_from_previous=self._ptr(self.wall_fp),
            wall_scratch=self._ptr(self.wall_scratch),
            source_index=self.source_index.ctypes.data_as(iptr),
            source_signal=self._ptr(self.source_signal),
            receiver_index=self.receiver_index.ctypes.data_as(iptr),
            recording=self._ptr(rec),
        )
        self.lib.acfdtd_run(ctypes.byref(state), self.step_index)
        self.step_index += n_steps
        if self.receiver_index.size:
            self.recordings.append(rec)


def host_thread_count() -> int:
    """Cores this process may run on (the reference C defaults to all online processors)."""
    return len(os.sched_getaffinity(0))
