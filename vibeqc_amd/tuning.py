"""Registry of measured launch shapes for the occupied-factor J/K build.

``FockPlan.autotune`` times alternative U i-tile widths and SYRK split-K
counts on a plan's own tensors and records the fastest here, keyed by device
architecture, precision, AO count, occupied columns and padded auxiliary
count; every later build of that shape, on any plan in the process (batches
included), uses it.  The registry starts empty, so the default rules apply
until a caller tunes or loads a profile.  With ``VIBEQC_TUNING_FILE`` set the
engine reads that file on first use and rewrites it after every record.

The reference tunes its kernels when it generates them and keeps user-local
schedules per device (``generativeqc.h:1191-1212``, ``profiles.py``); these
kernels are hand-written, so what is tuned here is their launch shape.
"""

from __future__ import annotations

import ctypes as ct
import os
from dataclasses import dataclass

from . import _native

PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2}


@dataclass(frozen=True)
class TunedShape:
    arch: str
    precision: str
    nbf: int
    occupied: int
    ldq: int
    u_tile: int
    syrk_splits: int


def _precision_code(precision: str) -> int:
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}") from None


def _check(status: int, what: str) -> None:
    if status != 0:
        raise _native.NativeError(status, what)


def record(arch: str, precision: str, nbf: int, occupied: int, ldq: int, u_tile: int, syrk_splits: int,
           ms: float = 0.0) -> None:
    """Record a shape (0 for either field keeps that field's default rule)."""
    lib = _native.load()
    _check(
        lib.vibeqc_tuning_record(arch.encode(), _precision_code(precision), nbf, occupied, ldq, u_tile, syrk_splits, ms),
        "invalid tuning entry",
    )


def lookup(arch: str, precision: str, nbf: int, occupied: int, ldq: int):
    """The recorded (u_tile, syrk_splits) for this key, or None."""
    lib = _native.load()
    shape = (ct.c_int32 * 2)()
    found = ct.c_int32(0)
    _check(
        lib.vibeqc_tuning_lookup(arch.encode(), _precision_code(precision), nbf, occupied, ldq, shape, ct.byref(found)),
        "invalid tuning key",
    )
    return (shape[0], shape[1]) if found.value else None


def count() -> int:
    lib = _native.load()
    n = ct.c_uint64(0)
    _check(lib.vibeqc_tuning_count(ct.byref(n)), "tuning registry unavailable")
    return n.value


def load(path: str | os.PathLike) -> bool:
    """Merge a tuning file into the registry; False when it does not exist."""
    lib = _native.load()
    found = ct.c_int32(0)
    _check(lib.vibeqc_tuning_load(os.fsencode(path), ct.byref(found)), f"malformed tuning file {path}")
    return bool(found.value)


def save(path: str | os.PathLike) -> None:
    lib = _native.load()
    _check(lib.vibeqc_tuning_save(os.fsencode(path)), f"cannot write tuning file {path}")


def clear() -> None:
    _check(_native.load().vibeqc_tuning_clear(), "tuning registry unavailable")
