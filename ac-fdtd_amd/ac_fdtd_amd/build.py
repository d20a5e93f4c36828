"""Compile the gfx950 HIP library in-tree: ``python -m ac_fdtd_amd.build``.

The reference builds its C loop on first use (``src/ac_fdtd/c_backend.py:55-79``). Here the
build is explicit and ahead of time, because the library is cross-compiled for gfx950 on a
machine that may have no GPU and is then shipped with the tree; the loader never compiles.

Flags:

* ``-ffp-contract=off`` keeps every ``a * b + c`` as two roundings, which is what lets the fp64
  kernels reproduce the NumPy oracle bit for bit (the reference uses the same flag for the
  same reason, ``c_backend.py:20-24``). Fused multiply-adds the fp32 compensated path wants
  are written explicitly.
* ``-target-feature -packed-fp32-ops`` keeps the compiler from forming packed fp32
  instructions on its own.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE_DIR = Path(__file__).resolve().parent
SOURCE_DIR = PACKAGE_DIR.parent / "csrc"
INCLUDE_DIR = PACKAGE_DIR.parent.parent / "include"
#: The library the loader opens; ACFDTD_HIP_LIBRARY points it at another build of the same
#: sources (tile experiments), the default is the in-tree build.
LIBRARY = Path(os.environ.get("ACFDTD_HIP_LIBRARY", PACKAGE_DIR / "lib" / "libacfdtd_hip.so"))
SOURCES = (
    "sweeps.hip",
    "fused.hip",
    "batch.hip",
    "adjoint.hip",
    "adjoint_fused.hip",
    "adjoint_fused64.hip",
    "fused2.hip",
    "fused2_64.hip",
    "adjoint2.hip",
    "adjoint2_64.hip",
    "adjoint2s.hip",
    "adjoint2s_64.hip",
    "fused2s.hip",
    "fused2s_64.hip",
    "fused2d.hip",
    "fused2d_64.hip",
    "fused2p.hip",
    "fused2p_64.hip",
    "fused2_v2.hip",
    "fused2_v1.hip",
    "fused2_64_v1.hip",
    "batch2.hip",
    "batch2_v2.hip",
    "batch2_v1.hip",
    "batch2_64.hip",
    "batch2_64_v1.hip",
    "wave.hip",
    "fourth.hip",
    "iir.hip",
    "pml.hip",
)
ARCH = os.environ.get("ACFDTD_HIP_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    found = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(found).exists():
        raise RuntimeError("hipcc not found: the ROCm toolchain is needed to build the library")
    return found


def compile_command(source: Path, output: Path) -> list[str]:
    return [hipcc(), *FLAGS, "-c", f"-I{INCLUDE_DIR}", f"-I{SOURCE_DIR}", str(source), "-o", str(output)]


def link_command(objects: list[Path], output: Path) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(output)]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed ({result.returncode}):\n{result.stderr[-4000:]}")


def _stale(source: Path, obj: Path) -> bool:
    """Whether ``obj`` is older than its source or than any header of the library."""
    if not obj.exists():
        return True
    built = obj.stat().st_mtime
    headers = [*SOURCE_DIR.glob("*.hpp"), *INCLUDE_DIR.glob("*.h")]
    return any(path.stat().st_mtime > built for path in (source, *headers))


def build_library(verbose: bool = False, incremental: bool = False) -> Path:
    """Compile every source (in parallel) and link :data:`LIBRARY`, replacing it atomically.

    ``incremental`` keeps objects in ``_build/`` beside the sources and recompiles only those
    older than their source or a header (a development convenience; the default rebuilds
    everything)."""
    LIBRARY.parent.mkdir(parents=True, exist_ok=True)
    with tempfile.TemporaryDirectory(prefix="acfdtd_build_") as scratch:
        folder = SOURCE_DIR.parent / "_build" if incremental else Path(scratch)
        folder.mkdir(exist_ok=True)
        objects = [folder / (Path(name).stem + ".o") for name in SOURCES]
        todo = [(SOURCE_DIR / name, obj) for name, obj in zip(SOURCES, objects)
                if not incremental or _stale(SOURCE_DIR / name, obj)]
        jobs = min(max(1, len(todo)), max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            for future in [pool.submit(_run, compile_command(src, obj), verbose) for src, obj in todo]:
                future.result()
        staging = LIBRARY.with_name(LIBRARY.name + f".tmp{os.getpid()}")
        try:
            _run(link_command(objects, staging), verbose)
        except RuntimeError:
            staging.unlink(missing_ok=True)
            raise
        os.replace(staging, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
