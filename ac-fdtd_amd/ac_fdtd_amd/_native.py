"""ctypes binding of ``include/acfdtd_hip.h``: the descriptor mirror and the loaded library.

The reference binds its C loop the same way (``src/ac_fdtd/c_backend.py:83-157``): a ctypes
``Structure`` mirrored field for field and checked against the library's own ``sizeof``.

torch is imported before the library is opened on purpose: torch ships its own HIP runtime
(soname ``libamdhip64.so.7``), and a library opened afterwards binds to that same copy, so
device pointers and streams from torch tensors are valid in the library's calls.
"""

from __future__ import annotations

import ctypes

import torch  # noqa: F401  (must precede the dlopen below; see module docstring)

from .build import LIBRARY

__all__ = [
    "ABI_VERSION",
    "AdjointDesc",
    "COMPENSATED",
    "Desc",
    "STAGED",
    "WaveDesc",
    "GRAD_AIR",
    "IIRDesc",
    "GRAD_SLOTS",
    "GRAD_SPONGE",
    "GRAD_WALLS",
    "MAX_RELAXATION",
    "PMLDesc",
    "check",
    "library",
]

ABI_VERSION = 1
MAX_RELAXATION = 4
#: Workgroups whose energy partials fit the library's fixed scratch (acfdtd_hip_energy_noalloc).
ENERGY_SCRATCH_TILES = 32768
WALL_CODES = 64
COMPENSATED = 1
STAGED = 2

_c = ctypes
_ptr = _c.c_void_p


class Desc(_c.Structure):
    """Mirror of ``acfdtd_hip_desc``."""

    _fields_ = [
        ("abi_version", _c.c_uint32),
        ("struct_size", _c.c_uint32),
        ("nx", _c.c_int32),
        ("ny", _c.c_int32),
        ("nz", _c.c_int32),
        ("real_bytes", _c.c_int32),
        ("n_relax", _c.c_int32),
        ("has_layer", _c.c_int32),
        ("n_sources", _c.c_int32),
        ("n_receivers", _c.c_int32),
        ("flags", _c.c_uint32),
        ("tile_j", _c.c_int32),
        ("chunk_x", _c.c_int32),
        ("time_block", _c.c_int32),
        ("vel_coef", _c.c_double),
        ("pres_coef", _c.c_double),
        ("src_coef", _c.c_double),
        ("relax_scale", _c.c_double),
        ("relax_decay", _c.c_double * MAX_RELAXATION),
        ("relax_mean", _c.c_double * MAX_RELAXATION),
        ("relax_gain", _c.c_double * MAX_RELAXATION),
        ("wall_mask", _c.c_uint64),
        ("wall_table", _ptr),
        ("p", _ptr),
        ("vx", _ptr),
        ("vy", _ptr),
        ("vz", _ptr),
        ("psi", _ptr),
        ("cell", _ptr * 3),
        ("face", _ptr * 3),
        ("source_cell", _ptr),
        ("source_length", _ptr),
        ("source_signal", _ptr),
        ("source_stride", _c.c_int64),
        ("receiver_cell", _ptr),
        ("recording", _ptr),
        ("recording_stride", _c.c_int64),
        ("p_swap", _ptr),
        ("vx_swap", _ptr),
        ("vy_swap", _ptr),
        ("vz_swap", _ptr),
        ("wall_tape", _ptr),
        ("wall_tape_stride", _c.c_int64),
        ("wall_tape_step0", _c.c_int64),
        ("wall_tape_rows", _c.c_int64),
        ("psi_swap", _ptr),
        ("z_pitch", _c.c_int32),
        ("reserved", _c.c_int32),
        ("div_tape", _ptr),
        ("div_tape_step0", _c.c_int64),
        ("div_tape_rows", _c.c_int64),
        ("layer_tape", _ptr),
        ("layer_tape_step0", _c.c_int64),
        ("layer_tape_rows", _c.c_int64),
        ("layer_tape_stride", _c.c_int64),
        ("layer_lo", _c.c_int32 * 3),
        ("layer_hi", _c.c_int32 * 3),
    ]


GRAD_WALLS = 1
GRAD_SPONGE = 2
GRAD_AIR = 4
GRAD_SLOTS = 12  # per-wavefront gradient accumulators: 6 walls, sponge, scale, 4 gains


class AdjointDesc(_c.Structure):
    """Mirror of ``acfdtd_hip_adjoint_desc``."""

    _fields_ = [
        ("abi_version", _c.c_uint32),
        ("struct_size", _c.c_uint32),
        ("grads", _c.c_uint32),
        ("time_block", _c.c_int32),
        ("lp", _ptr),
        ("lvx", _ptr),
        ("lvy", _ptr),
        ("lvz", _ptr),
        ("lpsi", _ptr),
        ("ldiv", _ptr),
        ("grad_recording", _ptr),
        ("grad_recording_stride", _c.c_int64),
        ("grad_recording_step0", _c.c_int64),
        ("grad_signal", _ptr),
        ("fwd_p", _ptr),
        ("fwd_vx", _ptr),
        ("fwd_vy", _ptr),
        ("fwd_vz", _ptr),
        ("weight_cell", _ptr * 3),
        ("weight_face", _ptr * 3),
        ("dt", _c.c_double),
        ("partials", _ptr),
        ("ldiv_swap", _ptr),
        ("lvx_swap", _ptr),
        ("lvy_swap", _ptr),
        ("lvz_swap", _ptr),
        ("fwd_step0", _c.c_int64),
        ("fwd_count", _c.c_int64),
        ("lp_swap", _ptr),
        ("lpsi_swap", _ptr),
    ]


class WaveDesc(_c.Structure):
    """Mirror of ``acfdtd_hip_wave_desc`` (the pressure-only scheme)."""

    _fields_ = [
        ("abi_version", _c.c_uint32),
        ("struct_size", _c.c_uint32),
        ("nx", _c.c_int32),
        ("ny", _c.c_int32),
        ("nz", _c.c_int32),
        ("real_bytes", _c.c_int32),
        ("n_sources", _c.c_int32),
        ("n_receivers", _c.c_int32),
        ("tile_j", _c.c_int32),
        ("chunk_x", _c.c_int32),
        ("courant2", _c.c_double),
        ("src_coef", _c.c_double),
        ("wall_mask", _c.c_uint64),
        ("wall_table", _ptr),
        ("p", _ptr),
        ("p_prev", _ptr),
        ("source_cell", _ptr),
        ("source_length", _ptr),
        ("source_signal", _ptr),
        ("source_stride", _c.c_int64),
        ("receiver_cell", _ptr),
        ("recording", _ptr),
        ("recording_stride", _c.c_int64),
        ("time_block", _c.c_int32),
        ("reserved", _c.c_int32),
        ("p_swap", _ptr),
        ("p_prev_swap", _ptr),
    ]


MAX_BRANCHES = 8


class IIRDesc(_c.Structure):
    """Mirror of ``acfdtd_hip_iir_desc`` (frequency-dependent walls)."""

    _fields_ = [
        ("abi_version", _c.c_uint32),
        ("struct_size", _c.c_uint32),
        ("n_branches", _c.c_int32),
        ("face_mask", _c.c_uint32),
        ("velocity_scale", _c.c_double),
        ("coef", _c.c_double * (6 * MAX_BRANCHES * 4)),
        ("state", _ptr),
    ]


class PMLDesc(_c.Structure):
    """Mirror of ``acfdtd_hip_pml_desc`` (split-field perfectly matched layer)."""

    _fields_ = [
        ("abi_version", _c.c_uint32),
        ("struct_size", _c.c_uint32),
        ("cell_a", _ptr * 3),
        ("cell_b", _ptr * 3),
        ("face_a", _ptr * 3),
        ("face_b", _ptr * 3),
        ("p_y", _ptr),
        ("p_z", _ptr),
    ]


_library: ctypes.CDLL | None = None


def library() -> ctypes.CDLL:
    """The loaded HIP library; raises if it was never built or does not match this binding."""
    global _library
    if _library is not None:
        return _library
    if not LIBRARY.exists():
        raise RuntimeError(
            f"{LIBRARY} is not built; run `python -m ac_fdtd_amd.build` "
            "(or __graft_entry__.build()) first"
        )
    lib = ctypes.CDLL(str(LIBRARY))
    lib.acfdtd_hip_struct_size.restype = _c.c_size_t
    lib.acfdtd_hip_abi_version.restype = _c.c_int
    lib.acfdtd_hip_supports_real.restype = _c.c_int
    lib.acfdtd_hip_supports_real.argtypes = [_c.c_int]
    lib.acfdtd_hip_last_error.restype = _c.c_char_p
    lib.acfdtd_hip_run.restype = _c.c_int
    lib.acfdtd_hip_run.argtypes = [_c.POINTER(Desc), _c.c_int64, _c.c_int64, _c.c_int64, _ptr]
    lib.acfdtd_hip_velocity_sweep.restype = _c.c_int
    lib.acfdtd_hip_velocity_sweep.argtypes = [_c.POINTER(Desc), _ptr]
    lib.acfdtd_hip_pressure_sweep.restype = _c.c_int
    lib.acfdtd_hip_pressure_sweep.argtypes = [_c.POINTER(Desc), _c.c_int64, _c.c_int64, _ptr]
    lib.acfdtd_hip_fused_step.restype = _c.c_int
    lib.acfdtd_hip_fused_step.argtypes = [_c.POINTER(Desc), _c.c_int64, _c.c_int64, _c.c_int, _ptr]
    lib.acfdtd_hip_fused2_step.restype = _c.c_int
    lib.acfdtd_hip_fused2_step.argtypes = [_c.POINTER(Desc), _c.c_int64, _c.c_int64, _c.c_int, _ptr]
    lib.acfdtd_hip_wall_tape_stride.restype = _c.c_int64
    lib.acfdtd_hip_wall_tape_stride.argtypes = [_c.c_int32, _c.c_int32, _c.c_int32]
    lib.acfdtd_hip_layer_tape_stride.restype = _c.c_int64
    lib.acfdtd_hip_layer_tape_stride.argtypes = [_c.POINTER(Desc)]
    lib.acfdtd_hip_layer_tape_unpack.restype = _c.c_int
    lib.acfdtd_hip_layer_tape_unpack.argtypes = [_c.POINTER(Desc), _c.c_int64, _ptr, _ptr, _ptr,
                                                 _ptr, _ptr]
    lib.acfdtd_hip_adjoint_partials_count.restype = _c.c_size_t
    lib.acfdtd_hip_adjoint_partials_count.argtypes = [_c.POINTER(Desc)]
    lib.acfdtd_hip_adjoint_face_sweep.restype = _c.c_int
    lib.acfdtd_hip_adjoint_face_sweep.argtypes = [_c.POINTER(Desc), _c.POINTER(AdjointDesc), _ptr]
    lib.acfdtd_hip_adjoint_cell_sweep.restype = _c.c_int
    lib.acfdtd_hip_adjoint_cell_sweep.argtypes = [
        _c.POINTER(Desc), _c.POINTER(AdjointDesc), _c.c_int64, _c.c_int, _c.c_int, _ptr
    ]
    lib.acfdtd_hip_adjoint_run.restype = _c.c_int
    lib.acfdtd_hip_adjoint_run.argtypes = [
        _c.POINTER(Desc), _c.POINTER(AdjointDesc), _c.c_int64, _c.c_int64, _ptr
    ]
    lib.acfdtd_hip_adjoint_two_step.restype = _c.c_int
    lib.acfdtd_hip_adjoint_two_step.argtypes = [_c.POINTER(Desc), _c.POINTER(AdjointDesc)]
    lib.acfdtd_hip_run_states.restype = _c.c_int
    lib.acfdtd_hip_run_states.argtypes = [
        _c.POINTER(Desc), _c.c_int64, _c.c_int64, _ptr, _ptr, _ptr, _ptr, _ptr
    ]
    lib.acfdtd_hip_stream_copy.restype = _c.c_int
    lib.acfdtd_hip_stream_copy.argtypes = [_ptr, _ptr, _c.c_size_t, _ptr]
    lib.acfdtd_hip_energy_workspace_bytes.restype = _c.c_size_t
    lib.acfdtd_hip_energy_workspace_bytes.argtypes = [_c.POINTER(Desc)]
    lib.acfdtd_hip_energy.restype = _c.c_int
    lib.acfdtd_hip_energy.argtypes = [_c.POINTER(Desc), _ptr, _ptr, _ptr]
    lib.acfdtd_hip_energy_noalloc.restype = _c.c_int
    lib.acfdtd_hip_energy_noalloc.argtypes = [_c.POINTER(Desc), _ptr, _ptr]
    lib.acfdtd_hip_pml_run.restype = _c.c_int
    lib.acfdtd_hip_pml_run.argtypes = [_c.POINTER(Desc), _c.POINTER(PMLDesc), _c.c_int64,
                                       _c.c_int64, _c.c_int64, _ptr]
    lib.acfdtd_hip_iir_run.restype = _c.c_int
    lib.acfdtd_hip_iir_run.argtypes = [_c.POINTER(Desc), _c.POINTER(IIRDesc), _c.c_int64,
                                       _c.c_int64, _c.c_int64, _ptr]
    lib.acfdtd_hip_fourth_run.restype = _c.c_int
    lib.acfdtd_hip_fourth_run.argtypes = [_c.POINTER(Desc), _c.c_double, _c.c_double, _c.c_int64,
                                          _c.c_int64, _c.c_int64, _ptr]
    lib.acfdtd_hip_wave_struct_size.restype = _c.c_size_t
    lib.acfdtd_hip_wave_run.restype = _c.c_int
    lib.acfdtd_hip_wave_run.argtypes = [_c.POINTER(WaveDesc), _c.c_int64, _c.c_int64, _c.c_int64,
                                        _ptr]
    if lib.acfdtd_hip_abi_version() != ABI_VERSION:
        raise RuntimeError(
            f"library ABI {lib.acfdtd_hip_abi_version()} does not match binding {ABI_VERSION}"
        )
    if lib.acfdtd_hip_struct_size() != ctypes.sizeof(Desc):
        raise RuntimeError(
            "the ctypes mirror of acfdtd_hip_desc no longer matches the library: "
            f"{ctypes.sizeof(Desc)} bytes here, {lib.acfdtd_hip_struct_size()} there"
        )
    for name, mirror in (("acfdtd_hip_wave_struct_size", WaveDesc),):
        if getattr(lib, name)() != ctypes.sizeof(mirror):
            raise RuntimeError(f"the ctypes mirror {mirror.__name__} no longer matches the library")
    _library = lib
    return lib


def check(status: int, what: str) -> None:
    """Raise ``RuntimeError`` with the library's message for a non-zero status."""
    if status != 0:
        message = library().acfdtd_hip_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({status}): {message}")
