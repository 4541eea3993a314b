"""Compile the HIP engine into ``vibeqc_amd/libvibeqc_amd.so`` for gfx950.

Every translation unit is compiled by ``hipcc --offload-arch=gfx950`` in
parallel, then linked against rocBLAS and rocSOLVER.  Packed-FP32 VALU
instructions are disabled for all device code.

Usage: python -m vibeqc_amd.build  (or vibeqc_amd.build.build())
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
CSRC = PACKAGE / "csrc"
BUILD = PACKAGE / "_build"
LIBRARY = PACKAGE / "libvibeqc_amd.so"
CLASS_UNITS = tuple(f"integrals_c{a}{b}.hip" for a in range(4) for b in range(a + 1))
GRAD_UNITS = tuple(f"integrals_g{a}{b}.hip" for a in range(4) for b in range(a + 1))
# Largest translation units first so the parallel build finishes soonest.
SOURCES = (
    tuple(reversed(GRAD_UNITS))
    + tuple(reversed(CLASS_UNITS))
    + ("integrals.hip", "integrals_grad.hip", "jk.hip", "noalloc.hip", "plan.hip", "uhf.hip", "gradient.hip", "scf.hip", "basis.cpp", "batch.cpp", "capi.cpp", "tuning.cpp")
)
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

COMMON_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
    "-Wall",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the engine is built with ROCm's hipcc")
    return exe


def _compile(src: str) -> Path:
    source = CSRC / src
    obj = BUILD / (source.stem + ".o")
    cmd = [_hipcc(), *COMMON_FLAGS]
    if source.suffix == ".hip":
        cmd += ["-x", "hip"]
    cmd += ["-c", str(source), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{proc.stdout}\n{proc.stderr}")
    return obj


def _stale(src: str) -> bool:
    """True when the object of ``src`` is older than the source or any header."""
    obj = BUILD / (Path(src).stem + ".o")
    if not obj.exists():
        return True
    built = obj.stat().st_mtime
    deps = [CSRC / src, *CSRC.glob("*.hpp"), *(ROOT / "include").rglob("*.h")]
    return any(d.stat().st_mtime > built for d in deps)


def build(jobs: int | None = None, changed_only: bool = False) -> Path:
    """Compile the sources and link the shared library; returns its path.
    With ``changed_only`` the objects that are newer than their source and
    every header are reused."""
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, min(16, os.cpu_count() or 1)))
    todo = [s for s in SOURCES if not changed_only or _stale(s)]
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        list(pool.map(_compile, todo))
    objects = [BUILD / (Path(s).stem + ".o") for s in SOURCES]
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [
        _hipcc(),
        "--offload-arch=gfx950",
        "-shared",
        "-o",
        str(tmp),
        *map(str, objects),
        f"-L{ROCM / 'lib'}",
        "-lrocsolver",
        "-lrocblas",
        f"-Wl,-rpath,{ROCM / 'lib'}",
    ]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if a != "--changed"]
    path = build(int(args[0]) if args else None, changed_only="--changed" in sys.argv[1:])
    print(path)
