"""Persistent ragged batch of DF-RHF systems (host mirror of
python/generativeqc/batch.py:123-900 for the RHF density-fitted path).

``Calculator.prepare_batch(systems)`` returns a :class:`PreparedBatch`;
``execute()`` solves every system (energies, and forces unless
``properties=("energy",)``) and returns a :class:`BatchResult` in input order.
Systems of any size share one call: the engine buckets them by shape and runs
them concurrently on the HIP device.  Each item carries its own status, so one
failure does not abort the batch (``strict=True`` raises instead).
"""

from __future__ import annotations

import ctypes as ct
import threading
from dataclasses import dataclass

import numpy as np

from . import _native
from .fock import NativeContext

_STATUS_TEXT = {
    0: "success",
    1: "invalid argument",
    2: "ABI mismatch",
    3: "not implemented",
    4: "not converged",
    5: "numerical failure",
    6: "device error",
    7: "out of memory",
    8: "internal error",
    9: "precision unavailable",
}


@dataclass(frozen=True)
class BatchItemResult:
    index: int
    status: int
    status_message: str
    energy: float
    forces: np.ndarray | None
    converged: bool
    iterations: int
    energy_change: float
    density_rms: float
    executed_backend: str
    bucket_id: int
    warm_start_used: bool
    warm_start_fallback: bool
    fock_builds: int | None = None
    # "cold", "cold_fallback", "in_process_warm" or "persistent_restart"
    # (a seed restored from a checkpoint), as in the reference's batch items.
    restart_origin: str = "cold"

    @property
    def succeeded(self) -> bool:
        return self.status == 0


@dataclass(frozen=True)
class BatchResult:
    """Input-ordered results for a ragged batch; forces are never padded."""

    items: tuple

    @property
    def energies(self) -> np.ndarray:
        return np.asarray([it.energy if it.succeeded else np.nan for it in self.items], dtype=np.float64)

    @property
    def failure_indices(self) -> tuple:
        return tuple(it.index for it in self.items if not it.succeeded)

    @property
    def succeeded(self) -> bool:
        return not self.failure_indices

    def raise_for_failures(self) -> None:
        failures = [f"{it.index}: {it.status_message}" for it in self.items if not it.succeeded]
        if failures:
            raise RuntimeError("batched item failures: " + "; ".join(failures))


class PreparedBatch:
    """A prepared native batch; reuse it with new coordinates of the same atoms."""

    def __init__(self, calculator, systems, *, charges=None, multiplicities=None, warm_start: bool = True):
        from .calculator import _normalise_atoms

        systems = [_normalise_atoms(s) for s in systems]
        if not systems:
            raise ValueError("a batch needs at least one system")
        charges = list(charges) if charges is not None else [0] * len(systems)
        multiplicities = list(multiplicities) if multiplicities is not None else [1] * len(systems)
        if len(charges) != len(systems) or len(multiplicities) != len(systems):
            raise ValueError("charges/multiplicities must match the systems")
        self._lock = threading.RLock()
        self._natoms = [len(s) for s in systems]
        self._calculator = calculator
        self._systems = systems
        self._charges = [int(q) for q in charges]
        self._multiplicities = [int(m) for m in multiplicities]
        self._last_status = [-1] * len(systems)
        self._restart_indices: set = set()
        self._ctx = NativeContext(calculator.device_id)
        lib = self._lib = self._ctx.lib
        handles = []
        try:
            orb, aux = [], []
            for atoms, q, m in zip(systems, charges, multiplicities):
                o, a = calculator._bases(atoms, q, m)
                orb.append(self._ctx.system(o))
                aux.append(self._ctx.system(a))
            handles = orb + aux
            n = len(systems)
            orb_arr = (ct.c_void_p * n)(*[h.value for h in orb])
            aux_arr = (ct.c_void_p * n)(*[h.value for h in aux])
            desc = calculator._method_descriptor()
            self._handle = ct.c_void_p()
            status = lib.vibeqc_batch_prepare(
                self._ctx.handle,
                orb_arr,
                aux_arr,
                n,
                ct.byref(desc),
                _native.BATCH_WARM_STARTS if warm_start else 0,
                ct.byref(self._handle),
            )
            _native.check(status, lib.generativeqc_context_get_last_detail(self._ctx.handle))
        finally:
            for h in handles:
                lib.generativeqc_system_destroy(h)

    def __len__(self) -> int:
        return len(self._natoms)

    def execute(self, coordinates=None, *, strict: bool = False, properties=None) -> BatchResult:
        """Solve every system; ``coordinates`` optionally replaces each
        geometry ((natoms, 3) Bohr arrays, or None to keep one)."""
        if properties is None:
            requested = frozenset({"energy", "forces"})
        else:
            if isinstance(properties, (str, bytes)):
                raise TypeError("properties must be an iterable of property names")
            requested = frozenset(properties)
        if "energy" not in requested:
            raise ValueError("properties must include 'energy'")
        unknown = requested - {"energy", "forces"}
        if unknown:
            raise ValueError(f"unsupported properties: {', '.join(sorted(map(repr, unknown)))}")
        want_forces = "forces" in requested
        n = len(self._natoms)
        inputs = None
        keep = []
        if coordinates is not None:
            if len(coordinates) != n:
                raise ValueError("one coordinate array (or None) per system is required")
            inputs = (_native.BatchInputC * n)()
            for i, c in enumerate(coordinates):
                inputs[i].struct_size = ct.sizeof(_native.BatchInputC)
                inputs[i].abi_version = _native.ABI_VERSION
                if c is None:
                    continue
                arr = np.ascontiguousarray(c, dtype=np.float64)
                if arr.shape != (self._natoms[i], 3):
                    raise ValueError(f"coordinates of system {i} must have shape ({self._natoms[i]}, 3)")
                keep.append(arr)
                inputs[i].coordinates = _native.pointer(arr)
                inputs[i].coordinate_count = arr.size
        forces = [np.empty((k, 3)) if want_forces else None for k in self._natoms]
        results = (_native.BatchItemResultC * n)()
        for i in range(n):
            results[i].struct_size = ct.sizeof(_native.BatchItemResultC)
            results[i].abi_version = _native.ABI_VERSION
            results[i].forces = _native.pointer(forces[i])
            results[i].force_count = 0 if forces[i] is None else forces[i].size
        with self._lock:
            if not self._handle:
                raise RuntimeError("batch is closed")
            status = self._lib.generativeqc_batch_execute(self._handle, inputs, 0 if inputs is None else n, results, n)
            _native.check(status, "batch execution failed")
            builds = ct.c_uint64()
            items = []
            for i in range(n):
                r = results[i]
                self._lib.generativeqc_batch_get_last_fock_builds(self._handle, i, ct.byref(builds))
                ok = r.status == 0
                self._last_status[i] = r.status
                if r.warm_start_fallback:
                    origin = "cold_fallback"
                elif r.warm_start_used:
                    origin = "persistent_restart" if i in self._restart_indices else "in_process_warm"
                else:
                    origin = "cold"
                if ok:
                    self._restart_indices.discard(i)  # the retained seed is now this run's
                f = forces[i]
                if f is not None and ok:
                    f.setflags(write=False)
                items.append(
                    BatchItemResult(
                        index=i,
                        status=r.status,
                        status_message=_STATUS_TEXT.get(r.status, "unknown status"),
                        energy=r.energy if ok else float("nan"),
                        forces=f if ok else None,
                        converged=bool(r.converged),
                        iterations=r.iterations,
                        energy_change=r.energy_change,
                        density_rms=r.density_rms,
                        executed_backend="hip",
                        bucket_id=r.bucket_id,
                        warm_start_used=bool(r.warm_start_used),
                        warm_start_fallback=bool(r.warm_start_fallback),
                        fock_builds=int(builds.value),
                        restart_origin=origin,
                    )
                )
        result = BatchResult(tuple(items))
        if strict:
            result.raise_for_failures()
        return result

    def _ensure_open(self) -> None:
        if not getattr(self, "_handle", None):
            raise RuntimeError("batch is closed")

    def save_checkpoint(self, path, *, max_bytes: int = 256 << 20) -> dict:
        """Atomically persist the retained converged densities with their
        identities and source diagnostics (batch.py:1312-1322)."""
        from .checkpoint import save_checkpoint

        with self._lock:
            return save_checkpoint(self, path, max_bytes=max_bytes)

    def load_checkpoint(self, path, *, allow_warm: bool = False, strict: bool = True, max_bytes: int = 256 << 20) -> dict:
        """Restore compatible seeds for the next execution (batch.py:1324-1343):
        exact restarts by default, changed geometry/controls with allow_warm."""
        from .checkpoint import load_checkpoint

        with self._lock:
            return load_checkpoint(self, path, allow_warm=allow_warm, strict=strict, max_bytes=max_bytes)

    def clear_warm_starts(self) -> None:
        with self._lock:
            self._ensure_open()
            _native.check(self._lib.generativeqc_batch_clear_warm_starts(self._handle), "clearing warm starts failed")
            self._restart_indices.clear()

    def close(self) -> None:
        with self._lock:
            if getattr(self, "_handle", None):
                self._lib.generativeqc_batch_destroy(self._handle)
                self._handle = ct.c_void_p()
            self._ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["BatchItemResult", "BatchResult", "PreparedBatch"]
