"""The C-ABI library on the CPU: it loads, exports what ``include/*.h`` declares, checks its
descriptor and reports errors — all without touching a GPU (validation precedes any HIP call).
Also the Python layer's argument errors (reference ``scheme.py:117-118``,
``c_backend.py:195-200``)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from ac_fdtd_amd import _native
from ac_fdtd_amd.build import LIBRARY

INCLUDE = Path(__file__).resolve().parents[1] / "include"
EINVAL, EABI = -1, -2


def declared_functions() -> list[str]:
    names = []
    for header in INCLUDE.glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names += re.findall(r"\b(acfdtd_\w+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not LIBRARY.exists():
        pytest.skip("HIP library not built")
    return _native.library()


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert "acfdtd_hip_run" in names and len(names) >= 9
    for name in names:
        assert hasattr(lib, name), name


def test_struct_size_and_abi_version(lib):
    assert lib.acfdtd_hip_struct_size() == ctypes.sizeof(_native.Desc)
    assert lib.acfdtd_hip_abi_version() == _native.ABI_VERSION
    assert lib.acfdtd_hip_supports_real(4) and lib.acfdtd_hip_supports_real(8)
    assert not lib.acfdtd_hip_supports_real(2)


def _desc(**overrides):
    d = _native.Desc()
    d.abi_version = _native.ABI_VERSION
    d.struct_size = ctypes.sizeof(_native.Desc)
    d.nx = d.ny = d.nz = 4
    d.real_bytes = 8
    fake = 4096  # never dereferenced: every case fails validation first
    d.p = d.vx = d.vy = d.vz = fake
    for key, value in overrides.items():
        setattr(d, key, value)
    return d


@pytest.mark.parametrize(
    ("overrides", "code", "message"),
    [
        ({"abi_version": 99}, EABI, "ABI version"),
        ({"struct_size": 8}, EABI, "bytes"),
        ({"nz": 0}, EINVAL, "cell counts"),
        ({"real_bytes": 2}, EINVAL, "real_bytes"),
        ({"n_relax": 5}, EINVAL, "n_relax"),
        ({"n_relax": 2}, EINVAL, "relaxation states"),
        ({"flags": 1}, EINVAL, "fp32 mode"),
        ({"flags": 8}, EINVAL, "unknown flags"),
        ({"tile_j": 9}, EINVAL, "tile_j"),
        ({"tile_j": -1}, EINVAL, "tile_j"),
        ({"time_block": 3}, EINVAL, "time_block"),
        ({"p": None}, EINVAL, "field pointer"),
        ({"wall_mask": 5}, EINVAL, "wall table"),
        ({"has_layer": 1}, EINVAL, "sponge profile"),
        ({"n_sources": 1}, EINVAL, "source arrays"),
        ({"n_receivers": 2}, EINVAL, "receiver arrays"),
        ({"reserved": 1}, EINVAL, "reserved"),
    ],
)
def test_malformed_descriptors_are_refused(lib, overrides, code, message):
    desc = _desc(**overrides)
    assert lib.acfdtd_hip_run(ctypes.byref(desc), 0, 1, 0, None) == code
    assert message in lib.acfdtd_hip_last_error().decode()
    assert lib.acfdtd_hip_energy(ctypes.byref(desc), None, None, None) == code


def test_null_descriptor_and_bad_run_arguments(lib):
    assert lib.acfdtd_hip_run(None, 0, 1, 0, None) == EINVAL
    desc = _desc()
    assert lib.acfdtd_hip_run(ctypes.byref(desc), 0, -1, 0, None) == EINVAL
    assert lib.acfdtd_hip_run(ctypes.byref(desc), -3, 1, 0, None) == EINVAL
    recording = _desc(n_receivers=1, receiver_cell=4096, recording=4096, recording_stride=10)
    assert lib.acfdtd_hip_run(ctypes.byref(recording), 0, 11, 0, None) == EINVAL
    assert "too short" in lib.acfdtd_hip_last_error().decode()
    assert lib.acfdtd_hip_pressure_sweep(ctypes.byref(recording), 0, 10, None) == EINVAL
    assert lib.acfdtd_hip_energy(ctypes.byref(desc), None, None, None) == EINVAL
    # Zero steps is a valid no-op and enqueues nothing.
    assert lib.acfdtd_hip_run(ctypes.byref(desc), 0, 0, 0, None) == 0


def test_short_wall_tapes_are_refused_before_any_launch(lib):
    """A tape must hold a row for every step a call writes: the two-step sweep writes rows for
    ``step`` and ``step + 1``, the fused step and the pressure sweep one row, the run all of
    them."""
    tape = {"wall_tape": 4096, "wall_tape_stride": lib.acfdtd_hip_wall_tape_stride(4, 4, 4),
            "wall_tape_step0": 10, "wall_tape_rows": 4}
    swaps = {"p_swap": 4096, "vx_swap": 4096, "vy_swap": 4096, "vz_swap": 4096, "time_block": 2}
    desc = _desc(**tape, **swaps)
    for step in (9, 13, 14):  # before the tape, its last row alone, and past its end
        assert lib.acfdtd_hip_fused2_step(ctypes.byref(desc), step, 0, 0, None) == EINVAL
        assert "wall tape" in lib.acfdtd_hip_last_error().decode()
    for step in (9, 14):
        assert lib.acfdtd_hip_fused_step(ctypes.byref(desc), step, 0, 0, None) == EINVAL
        assert lib.acfdtd_hip_pressure_sweep(ctypes.byref(desc), step, 0, None) == EINVAL
    assert lib.acfdtd_hip_run(ctypes.byref(desc), 11, 4, 0, None) == EINVAL
    assert "wall tape" in lib.acfdtd_hip_last_error().decode()


def test_padded_rows_are_checked_by_the_batch_api(lib):
    """z_pitch (padded rows) is a layout of acfdtd_hip_run and the batch API alone: the other
    entry points refuse it; it must exceed nz and takes no wall tape; a valid one is accepted,
    and a padded room takes lanes as wide as the pitch allows."""
    from ac_fdtd_amd.batch import _bind

    _bind(lib)
    padded = _desc(nz=5, z_pitch=8)
    assert lib.acfdtd_hip_energy(ctypes.byref(padded), None, None, None) == EINVAL
    assert "acfdtd_hip_run and the batch API only" in lib.acfdtd_hip_last_error().decode()
    assert lib.acfdtd_hip_pressure_sweep(ctypes.byref(padded), 0, 0, None) == EINVAL
    assert lib.acfdtd_hip_fused2_step(ctypes.byref(padded), 0, 0, 0, None) == EINVAL
    assert lib.acfdtd_hip_run(ctypes.byref(_desc(nz=5, z_pitch=5)), 0, 1, 0, None) == EINVAL
    assert "greater than nz" in lib.acfdtd_hip_last_error().decode()
    swaps = {"p_swap": 4096, "vx_swap": 4096, "vy_swap": 4096, "vz_swap": 4096}
    rooms = (_native.Desc * 2)(_desc(nz=4, **swaps), _desc(nz=5, z_pitch=8, **swaps))
    assert lib.acfdtd_hip_batch_workspace_bytes(rooms, 2) > 0
    for bad, message in (({"z_pitch": 5}, "greater than nz"), ({"z_pitch": 3}, "greater than nz"),
                         ({"z_pitch": 8, "wall_tape": 4096, "wall_tape_stride": 1000,
                           "wall_tape_rows": 1}, "no wall tape")):
        rooms[1] = _desc(nz=5, **swaps, **bad)
        assert lib.acfdtd_hip_batch_workspace_bytes(rooms, 2) == 0
        assert message in lib.acfdtd_hip_last_error().decode()


def test_divergence_tapes_are_checked_before_any_launch(lib):
    """A divergence tape needs relaxation states, must hold a row for every step a call
    advances, and is refused with padded rows and by the batch API."""
    from ac_fdtd_amd.batch import _bind

    _bind(lib)
    air = {"n_relax": 1, "psi": 4096}
    tape = {"div_tape": 4096, "div_tape_step0": 10, "div_tape_rows": 4}
    assert lib.acfdtd_hip_run(ctypes.byref(_desc(**tape)), 10, 1, 0, None) == EINVAL
    assert "relaxation states" in lib.acfdtd_hip_last_error().decode()
    desc = _desc(**air, **tape)
    for offset, n in ((9, 1), (10, 5), (13, 2)):
        assert lib.acfdtd_hip_run(ctypes.byref(desc), offset, n, 0, None) == EINVAL
        assert "divergence tape" in lib.acfdtd_hip_last_error().decode()
    assert lib.acfdtd_hip_pressure_sweep(ctypes.byref(desc), 14, 0, None) == EINVAL
    assert lib.acfdtd_hip_fused2_step(ctypes.byref(desc), 13, 0, 0, None) == EINVAL
    assert "divergence tape" in lib.acfdtd_hip_last_error().decode()
    swaps = {"p_swap": 4096, "vx_swap": 4096, "vy_swap": 4096, "vz_swap": 4096}
    rooms = (_native.Desc * 1)(_desc(nz=5, z_pitch=8, **swaps, **air, **tape))
    assert lib.acfdtd_hip_batch_workspace_bytes(rooms, 1) == 0
    assert "padded rows take no divergence tape" in lib.acfdtd_hip_last_error().decode()
    rooms[0] = _desc(**swaps, **air, **tape)
    assert lib.acfdtd_hip_batch_workspace_bytes(rooms, 1) == 0
    assert "divergence tapes" in lib.acfdtd_hip_last_error().decode()


def _frame_cells(shape, lo, hi):
    """Cells of a layer tape's frame, counted by brute force over a mask: whole planes and rows
    at the layer's x and y ends, z cells within lo[2] and hi[2] rounded up to whole groups of
    4, everything when an axis's ends meet."""
    nx, ny, nz = shape
    zl, zh = -(-lo[2] // 4) * 4, -(-hi[2] // 4) * 4
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    in_x = (i < lo[0]) | (i >= nx - hi[0]) if lo[0] + hi[0] < nx else np.ones_like(i, bool)
    if zl + zh >= nz:
        in_yz = np.ones_like(i, bool)
    else:
        in_y = (j < lo[1]) | (j >= ny - hi[1]) if lo[1] + hi[1] < ny else np.ones_like(i, bool)
        in_yz = in_y | (k < zl) | (k >= nz - zh)
    return int((in_x | in_yz).sum())


@pytest.mark.parametrize(("shape", "lo", "hi"), [
    ((20, 14, 12), (3, 3, 3), (3, 3, 3)),
    ((16, 16, 16), (4, 0, 0), (0, 0, 4)),
    ((9, 7, 6), (2, 1, 0), (1, 0, 1)),
    ((10, 9, 8), (0, 0, 3), (0, 0, 3)),     # z ends meet after rounding: whole planes
    ((6, 5, 12), (4, 3, 0), (3, 2, 0)),     # x and y ends meet
])
def test_layer_tape_stride_counts_the_frame(lib, shape, lo, hi):
    d = _desc(nx=shape[0], ny=shape[1], nz=shape[2])
    for a in range(3):
        d.layer_lo[a], d.layer_hi[a] = lo[a], hi[a]
    nx, ny, nz = shape
    far = ny * nz + nx * nz + nx * ny
    want = -(-(4 * _frame_cells(shape, lo, hi) + far) // 4) * 4
    assert lib.acfdtd_hip_layer_tape_stride(ctypes.byref(d)) == want
    assert lib.acfdtd_hip_layer_tape_stride(None) == -1


def test_layer_tapes_are_checked_before_any_launch(lib):
    """A layer tape needs the absorbing layer, alignment, a long enough stride and a row for
    every step acfdtd_hip_run advances; the entry points that would not write it refuse it."""
    from ac_fdtd_amd.batch import _bind

    _bind(lib)
    layer = {"has_layer": 1, "cell": (ctypes.c_void_p * 3)(4096, 4096, 4096),
             "face": (ctypes.c_void_p * 3)(4096, 4096, 4096)}
    stride = lib.acfdtd_hip_layer_tape_stride(ctypes.byref(_desc()))
    tape = {"layer_tape": 4096, "layer_tape_step0": 10, "layer_tape_rows": 4,
            "layer_tape_stride": stride}
    assert lib.acfdtd_hip_run(ctypes.byref(_desc(**tape)), 10, 1, 0, None) == EINVAL
    assert "absorbing layer" in lib.acfdtd_hip_last_error().decode()
    for bad in ({"layer_tape": 4100}, {"layer_tape_stride": stride - 4},
                {"layer_tape_stride": stride + 2}):
        desc = _desc(**layer, **{**tape, **bad})
        assert lib.acfdtd_hip_run(ctypes.byref(desc), 10, 1, 0, None) == EINVAL
        assert "layer tape" in lib.acfdtd_hip_last_error().decode()
    desc = _desc(**layer, **tape)
    for offset, n in ((9, 1), (10, 5), (13, 2)):
        assert lib.acfdtd_hip_run(ctypes.byref(desc), offset, n, 0, None) == EINVAL
        assert "layer tape does not cover" in lib.acfdtd_hip_last_error().decode()
    swaps = {"p_swap": 4096, "vx_swap": 4096, "vy_swap": 4096, "vz_swap": 4096, "time_block": 2}
    desc = _desc(**layer, **tape, **swaps)
    assert lib.acfdtd_hip_velocity_sweep(ctypes.byref(desc), None) == EINVAL
    assert lib.acfdtd_hip_pressure_sweep(ctypes.byref(desc), 10, 0, None) == EINVAL
    assert lib.acfdtd_hip_fused_step(ctypes.byref(desc), 10, 0, 0, None) == EINVAL
    assert lib.acfdtd_hip_fused2_step(ctypes.byref(desc), 10, 0, 0, None) == EINVAL
    assert "acfdtd_hip_run only" in lib.acfdtd_hip_last_error().decode()
    rooms = (_native.Desc * 1)(desc)
    assert lib.acfdtd_hip_batch_workspace_bytes(rooms, 1) == 0
    assert "layer tapes" in lib.acfdtd_hip_last_error().decode()
    fields = [ctypes.c_void_p(4096)] * 4
    assert lib.acfdtd_hip_layer_tape_unpack(ctypes.byref(desc), 14, *fields, None) == EINVAL
    assert lib.acfdtd_hip_layer_tape_unpack(ctypes.byref(_desc(**layer)), 0, *fields, None) == EINVAL
    assert "no layer tape" in lib.acfdtd_hip_last_error().decode()


def test_energy_without_allocation_refuses_grids_beyond_its_scratch(lib):
    """ceil(ny / 8) * ceil(nz / 64) workgroup columns must fit the fixed scratch; refused before
    any HIP call."""
    columns = -(-2400 // 8) * -(-12800 // 64)
    assert columns > _native.ENERGY_SCRATCH_TILES
    desc = _desc(nx=2, ny=2400, nz=12800)
    out = ctypes.c_double * 3
    assert lib.acfdtd_hip_energy_noalloc(ctypes.byref(desc), out(), None) == EINVAL
    assert "energy scratch" in lib.acfdtd_hip_last_error().decode()
    assert lib.acfdtd_hip_energy_noalloc(ctypes.byref(_desc()), None, None) == EINVAL
    assert lib.acfdtd_hip_energy_noalloc(None, out(), None) == EINVAL


def test_energy_workspace_size_follows_the_tiling(lib):
    small = lib.acfdtd_hip_energy_workspace_bytes(ctypes.byref(_desc()))
    large = lib.acfdtd_hip_energy_workspace_bytes(ctypes.byref(_desc(nx=200, ny=130, nz=70)))
    assert small > 0 and small % 24 == 0 and large > small
    assert lib.acfdtd_hip_energy_workspace_bytes(ctypes.byref(_desc(nx=0))) == 0


def test_the_python_solver_refuses_to_run_without_a_gpu():
    import torch

    from ac_fdtd_amd import Grid
    from ac_fdtd_amd.scheme import AcousticFDTD

    grid = Grid(dx=0.1, nx=3, ny=3, nz=3)
    with pytest.raises(ValueError, match="GPU device"):
        AcousticFDTD(grid, device="cpu")
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            AcousticFDTD(grid)


def test_host_side_argument_errors():
    from ac_fdtd_amd import Grid
    from ac_fdtd_amd.coefficients import build_coefficients
    from ac_fdtd_amd.medium import AIR

    grid = Grid(dx=0.1, nx=3, ny=3, nz=3)
    with pytest.raises(ValueError, match="dtype"):
        build_coefficients(grid, AIR, 1.0, np.float16, None, None, None)
