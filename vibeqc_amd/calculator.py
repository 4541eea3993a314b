"""HF ``Calculator`` entry point (host mirror of python/generativeqc/calculator.py:146-200).

The density-fitted Hartree-Fock paths are provided: ``method="rhf"``
(closed shell) and ``method="uhf"`` (any multiplicity; also in ragged batches), with ``density_fitting`` in ("hip", "cuda", "auto", True).  Coordinates are in
Bohr; results carry the energy in Hartree and convergence diagnostics.
``precision`` selects the J/K arithmetic: "fp64" (the reference's DF
precision), "fp32", "bf16x3", or "mixed" (= bf16x3 iterations followed by FP64
refinement to the requested tolerances).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .basis import Atom, AOBasis, CanonicalBasis, NUMBERS, atoms_from, shells_for
from .fock import FockBuildSpec, FockPlan

_PRECISIONS = {"fp64": "fp64", "fp32": "fp32", "bf16x3": "bf16x3", "mixed": "bf16x3"}


@dataclass(frozen=True)
class Result:
    energy: float
    forces: np.ndarray | None
    iterations: int
    converged: bool
    energy_change: float
    density_rms: float
    fock_builds: int
    density: np.ndarray
    nbf: int
    naux: int


def _normalise_atoms(atoms) -> tuple:
    out = []
    for a in atoms:
        if isinstance(a, Atom):
            out.append(a)
            continue
        z, pos = a
        if isinstance(z, str):
            z = NUMBERS[z.capitalize()]
        out.append(Atom(int(z), tuple(float(v) for v in pos)))
    return tuple(out)


class Calculator:
    def __init__(
        self,
        method: str = "rhf",
        basis="cc-pvdz",
        device: str = "hip",
        device_id: int = 0,
        basis_representation: str | None = None,
        *,
        density_fitting="hip",
        auxiliary_basis="cc-pvdz-jkfit",
        density_fitting_relative_threshold: float = 1.0e-10,
        density_fitting_memory_budget_bytes: int = 0,
        max_iterations: int = 100,
        energy_tolerance: float = 1.0e-10,
        density_tolerance: float = 1.0e-8,
        diis_history: int = 8,
        screening_tolerance: float | None = None,
        precision: str = "fp64",
    ):
        if str(method).lower() not in ("rhf", "uhf"):
            raise NotImplementedError("the MI355X engine provides density-fitted RHF and UHF only")
        if density_fitting not in ("hip", "cuda", "auto", True):
            raise NotImplementedError("only density-fitted J/K is provided (density_fitting='hip')")
        if device not in ("hip", "cuda", "gpu"):
            raise NotImplementedError("the engine runs on the HIP device only")
        if precision not in _PRECISIONS:
            raise ValueError("precision must be fp64, fp32, bf16x3 or mixed")
        if not (0.0 < density_fitting_relative_threshold < 1.0):
            raise ValueError("metric threshold must lie strictly between zero and one")
        for v in (energy_tolerance, density_tolerance):
            if not (math.isfinite(v) and v > 0):
                raise ValueError("tolerances must be finite and positive")
        self.method = str(method).lower()
        self.basis = basis
        self.auxiliary_basis = auxiliary_basis
        self.device_id = device_id
        if basis_representation is None and isinstance(basis, CanonicalBasis):
            basis_representation = basis.representation  # a canonical record states its own
        self.representation = "cartesian" if basis_representation == "cartesian" else "real_spherical"
        self.metric_threshold = density_fitting_relative_threshold
        self.budget = density_fitting_memory_budget_bytes
        self.max_iterations = max_iterations
        self.energy_tolerance = energy_tolerance
        self.density_tolerance = density_tolerance
        self.diis_history = diis_history
        self.screening_tolerance = 1e-12 if screening_tolerance is None else screening_tolerance
        self.precision = precision

    def _bases(self, atoms, charge, multiplicity):
        orbital = AOBasis(
            atoms,
            self.basis if isinstance(self.basis, tuple) else shells_for(self.basis, atoms),
            self.representation,
            charge,
            multiplicity,
        )
        aux_source = self.auxiliary_basis
        aux = AOBasis(
            atoms,
            aux_source if isinstance(aux_source, tuple) else shells_for(aux_source, atoms),
            self.representation,
            charge,
            multiplicity,
        )
        return orbital, aux

    def _method_descriptor(self, auxiliary=None, *, single=False):
        """generativeqc_method_descriptor of this calculator (generativeqc.h:723-769).
        ``auxiliary`` is an optional native system handle carrying the
        auxiliary shells; a single calculation (``single``) also accepts fp32,
        selected after preparation."""
        from . import _native

        if self.precision not in ("fp64", "mixed", "bf16x3") and not (single and self.precision == "fp32"):
            raise NotImplementedError("batches run in fp64 or mixed (bf16x3 + FP64 refinement) precision")
        return _native.descriptor(
            _native.MethodDescriptorC,
            method=_native.METHOD_UHF if self.method == "uhf" else _native.METHOD_RHF,
            max_iterations=self.max_iterations,
            diis_history=self.diis_history,
            energy_tolerance=self.energy_tolerance,
            density_tolerance=self.density_tolerance,
            screening_tolerance=self.screening_tolerance,
            density_fitting_mode=_native.DENSITY_FITTING_CUDA,
            density_fitting_auxiliary_basis=auxiliary,
            density_fitting_relative_threshold=self.metric_threshold,
            density_fitting_memory_budget_bytes=self.budget,
            precision_mode=_native.PRECISION_MODE["auto" if self.precision in ("mixed", "bf16x3") else "fp64"],
        )

    def prepare_batch(self, systems, *, charges=None, multiplicities=None, warm_start: bool = True):
        """Prepare a persistent ragged batch (calculator.py:1931-1963)."""
        from .batch import PreparedBatch

        return PreparedBatch(self, systems, charges=charges, multiplicities=multiplicities, warm_start=warm_start)

    def batch_singlepoint(self, systems, *, charges=None, multiplicities=None, strict: bool = False, properties=None):
        """One-shot ragged batch with per-system status (calculator.py:1965-1980)."""
        with self.prepare_batch(systems, charges=charges, multiplicities=multiplicities, warm_start=False) as batch:
            return batch.execute(strict=strict, properties=properties)

    def prepare(self, atoms, *, charge: int = 0, multiplicity: int = 1) -> FockPlan:
        """Build the device plan (integrals, metric factor, B) for one geometry."""
        atoms = _normalise_atoms(atoms)
        orbital, aux = self._bases(atoms, charge, multiplicity)
        return FockPlan(
            orbital,
            FockBuildSpec.hf("unrestricted" if self.method == "uhf" else "restricted"),
            auxiliary=aux,
            device_id=self.device_id,
            screening_tolerance=self.screening_tolerance,
            metric_relative_threshold=self.metric_threshold,
            device_budget_bytes=self.budget,
            precision=_PRECISIONS[self.precision],
        )

    def singlepoint(self, atoms, *, charge: int = 0, multiplicity: int = 1, properties=None) -> Result:
        """Energy (Hartree) and, unless ``properties=("energy",)``, analytic
        forces -dE/dR (natom, 3) in Hartree/bohr (calculator.py:2128-2200:
        the default requests every supported property)."""
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
            raise ValueError(f"method {self.method!r} does not support properties: {', '.join(sorted(map(repr, unknown)))}")
        return self._calculation(_normalise_atoms(atoms), charge, multiplicity, "forces" in requested)

    def _calculation(self, atoms, charge, multiplicity, want_forces) -> Result:
        """One prepared native calculation: generativeqc_calculation_prepare /
        _execute / _destroy, bound as the reference's Calculator binds them
        (calculator.py:2291-2346)."""
        import ctypes as ct

        from . import _native
        from .fock import NativeContext

        orbital, aux = self._bases(atoms, charge, multiplicity)
        ctx = NativeContext(self.device_id)
        lib = ctx.lib
        handles, calc = [], ct.c_void_p()
        try:
            handles = [ctx.system(orbital), ctx.system(aux)]
            desc = self._method_descriptor(handles[1], single=True)
            status = lib.generativeqc_calculation_prepare(ctx.handle, handles[0], ct.byref(desc), ct.byref(calc))
            _native.check(status, lib.generativeqc_context_get_last_detail(ctx.handle))
            if self.precision == "fp32":
                _native.check(lib.vibeqc_calculation_set_precision(calc, _native.PRECISION["fp32"]),
                              lib.generativeqc_context_get_last_detail(ctx.handle))
            forces = np.empty((len(atoms), 3)) if want_forces else None
            result = _native.descriptor(
                _native.ResultDescriptorC,
                forces=_native.pointer(forces),
                force_count=0 if forces is None else forces.size,
                executed_backend=_native.BACKEND_HIP,
            )
            status = lib.generativeqc_calculation_execute(calc, ct.byref(result))
            _native.check(status, lib.generativeqc_context_get_last_detail(ctx.handle))
            counts = (ct.c_uint64 * 4)()
            _native.check(lib.vibeqc_calculation_last_state(calc, None, 0, counts), "calculation state query failed")
            density = np.empty(counts[0])
            _native.check(lib.vibeqc_calculation_last_state(calc, _native.pointer(density), counts[0], counts),
                          "calculation state copy failed")
            n = int(counts[2])
            density = density.reshape((2, n, n) if self.method == "uhf" else (n, n))
            density.setflags(write=False)
            if forces is not None:
                forces.setflags(write=False)
            return Result(result.energy, forces, result.iterations, bool(result.converged), result.energy_change,
                          result.density_rms, int(counts[1]), density, n, int(counts[3]))
        finally:
            if calc:
                lib.generativeqc_calculation_destroy(calc)
            for h in handles:
                lib.generativeqc_system_destroy(h)
            ctx.close()


__all__ = ["Calculator", "Result", "atoms_from"]
