"""Prepared DF-JK Fock operator (host mirror of the reference's ``FockPlan``).

Same public shape as python/generativeqc/fock.py:33-792: ``FockTerm``,
``FockBuildSpec.hf(...)``, ``FockPlan(basis, spec, auxiliary=...)`` with
``evaluate(D)`` -> ``FockEvaluation`` and ``solve()`` -> ``FockScfResult``.
Execution always happens on the HIP device through the C ABI; there is no CPU
path.  Extensions: ``precision=`` ("fp64", "fp32", "bf16x3"),
``evaluate_occupied(C)`` and ``benchmark_occupied(C, repeat)``.
"""

from __future__ import annotations

import ctypes as ct
import math
import threading
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .basis import AOBasis

_OPERATORS = {"full_range": 0, "short_range": 1, "long_range": 2}
_APPROXIMATIONS = {"exact": 0, "density_fitted": 1}
_SPINS = {"restricted": 0, "unrestricted": 1}


@dataclass(frozen=True)
class FockTerm:
    """One operator/approximation and its signed Fock coefficient (fock.py:33-53)."""

    present: bool = True
    coefficient: float = 1.0
    operator: str = "full_range"
    omega: float = 0.0
    approximation: str = "density_fitted"

    def __post_init__(self):
        if type(self.present) is not bool:
            raise TypeError("Fock presence must be bool")
        if self.operator not in _OPERATORS or self.approximation not in _APPROXIMATIONS:
            raise ValueError("unknown Fock operator/approximation")
        if not math.isfinite(self.coefficient) or not math.isfinite(self.omega) or self.omega < 0:
            raise ValueError("Fock coefficients and nonnegative omega must be finite")


@dataclass(frozen=True)
class FockBuildSpec:
    """J/K semantics (fock.py:56-101).  derivative_order 1 (the default, as in
    the reference) lets evaluate() return the fixed-density two-electron
    gradient and solve() the complete forces."""

    spin: str = "restricted"
    derivative_order: int = 1
    coulomb: FockTerm = field(default_factory=FockTerm)
    exchange: FockTerm = field(default_factory=lambda: FockTerm(coefficient=-0.5))

    def __post_init__(self):
        if self.spin not in _SPINS or self.derivative_order not in (0, 1):
            raise ValueError("unsupported Fock spin/derivative order")

    @classmethod
    def hf(cls, spin="restricted", *, coulomb="density_fitted", exchange="density_fitted", derivative_order=1):
        return cls(
            spin,
            derivative_order,
            FockTerm(approximation=coulomb),
            FockTerm(coefficient=-1.0 if spin == "unrestricted" else -0.5, approximation=exchange),
        )


@dataclass(frozen=True, eq=False)
class FockEvaluation:
    """Fixed-density energies (Hartree), raw J/K and the assembled Fock matrix."""

    energy_one_electron: float
    energy_two_electron: float
    nuclear_repulsion: float
    fock: np.ndarray
    coulomb: np.ndarray | None
    exchange: np.ndarray | None
    gradient: np.ndarray | None = None

    @property
    def energy(self) -> float:
        return self.energy_one_electron + self.energy_two_electron + self.nuclear_repulsion


@dataclass(frozen=True, eq=False)
class FockScfResult:
    energy: float
    density: np.ndarray
    forces: np.ndarray | None
    iterations: int
    energy_change: float
    density_rms: float
    initial_density_used: bool
    fock_builds: int


def _term(t: FockTerm):
    return _native.FockTermC(
        present=1 if t.present else 0,
        coefficient=t.coefficient,
        op=_OPERATORS[t.operator],
        omega=t.omega,
        approximation=_APPROXIMATIONS[t.approximation],
    )


def _immutable(a):
    a = np.asarray(a)
    a.setflags(write=False)
    return a


class NativeContext:
    """Owns one generativeqc_context on a HIP device."""

    def __init__(self, device_id: int = 0):
        self.lib = _native.load()
        self.handle = ct.c_void_p()
        desc = _native.descriptor(_native.ContextDescriptor, device_id=device_id, backend=_native.BACKEND_HIP)
        status = self.lib.generativeqc_context_create(ct.byref(desc), ct.byref(self.handle))
        _native.check(status, "HIP context creation failed (is a gfx950 device visible?)")

    def system(self, basis: AOBasis):
        atoms = (_native.Atom * basis.natom)(
            *[_native.Atom(a.atomic_number, *a.position) for a in basis.atoms]
        )
        prims = []
        shells = []
        for s in basis.shells:
            shells.append(_native.Shell(s.atom_index, s.angular_momentum, len(prims), len(s.primitives)))
            prims.extend(_native.Primitive(p.exponent, p.coefficient) for p in s.primitives)
        shell_arr = (_native.Shell * len(shells))(*shells)
        prim_arr = (_native.Primitive * len(prims))(*prims)
        desc = _native.descriptor(
            _native.SystemDescriptor,
            atoms=atoms,
            atom_count=basis.natom,
            shells=shell_arr,
            shell_count=len(shells),
            primitives=prim_arr,
            primitive_count=len(prims),
            charge=basis.charge,
            multiplicity=basis.multiplicity,
            basis_representation=_native.BASIS_SPHERICAL
            if basis.representation == "real_spherical"
            else _native.BASIS_CARTESIAN,
        )
        handle = ct.c_void_p()
        status = self.lib.generativeqc_system_create(self.handle, ct.byref(desc), ct.byref(handle))
        _native.check(status, self.lib.generativeqc_context_get_last_detail(self.handle))
        return handle

    def close(self):
        if self.handle:
            self.lib.generativeqc_context_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FockPlan:
    """Own device-resident DF J/K sources for one orbital (+ auxiliary) basis."""

    def __init__(
        self,
        basis: AOBasis,
        spec: FockBuildSpec | None = None,
        *,
        auxiliary: AOBasis | None = None,
        device: str = "hip",
        device_id: int = 0,
        screening_tolerance: float = 1e-12,
        metric_relative_threshold: float = 1e-10,
        device_budget_bytes: int = 0,
        precision: str = "fp64",
    ):
        if device not in ("hip", "cuda", "gpu"):
            raise ValueError("the MI355X engine executes on the HIP device only (device='hip')")
        if not math.isfinite(screening_tolerance) or screening_tolerance < 0:
            raise ValueError("screening tolerance must be nonnegative and finite")
        if not math.isfinite(metric_relative_threshold) or not 0 < metric_relative_threshold < 1:
            raise ValueError("metric threshold must lie strictly between zero and one")
        self._lock = threading.RLock()
        self.basis = basis
        self.spec = FockBuildSpec.hf() if spec is None else spec
        self._ctx = NativeContext(device_id)
        lib = self._lib = self._ctx.lib
        self._handle = ct.c_void_p()
        system = self._ctx.system(basis)
        aux = self._ctx.system(auxiliary) if auxiliary is not None else ct.c_void_p()
        try:
            spec_c = _native.descriptor(
                _native.FockSpecC,
                spec_version=1,
                spin=_SPINS[self.spec.spin],
                derivative_order=self.spec.derivative_order,
                coulomb=_term(self.spec.coulomb),
                exchange=_term(self.spec.exchange),
            )
            controls = _native.descriptor(
                _native.FockControlsC,
                screening_tolerance=screening_tolerance,
                metric_relative_threshold=metric_relative_threshold,
                device_budget_bytes=device_budget_bytes,
            )
            status = lib.generativeqc_fock_plan_create(
                self._ctx.handle, system, aux if auxiliary is not None else None, ct.byref(spec_c), ct.byref(controls), ct.byref(self._handle)
            )
            _native.check(status, lib.generativeqc_context_get_last_detail(self._ctx.handle))
        finally:
            lib.generativeqc_system_destroy(system)
            if auxiliary is not None:
                lib.generativeqc_system_destroy(aux)
        self.set_precision(precision)

    # -- helpers -------------------------------------------------------------
    def _ensure_open(self):
        if not self._handle:
            raise RuntimeError("Fock plan is closed")

    def _check(self, status):
        _native.check(status, self._lib.generativeqc_fock_plan_last_error(self._handle))

    @property
    def nao(self) -> int:
        return self.basis.nao

    def shape(self) -> dict:
        out = (ct.c_uint64 * 6)()
        self._check(self._lib.vibeqc_fock_plan_shape(self._handle, out))
        return {
            "nbf": out[0],
            "naux": out[1],
            "npair": out[2],
            "metric_rank": out[3],
            "cholesky": bool(out[4]),
            "device_bytes": out[5],
        }

    def streaming(self) -> dict:
        """Whether the fitted tensor is regenerated per build to honour
        ``device_budget_bytes`` (``streamed``), with the auxiliary block width
        and the three-centre row tiles of each block."""
        out = (ct.c_uint64 * 3)()
        self._check(self._lib.vibeqc_fock_plan_streaming(self._handle, out))
        return {"streamed": bool(out[0]), "auxiliary_block": out[1], "row_tiles": out[2]}

    def last_solve_stats(self) -> dict:
        """Statistics of the last solve(): iterations, iterations with
        reduced-precision Fock builds, new densities from purification,
        FP64-stage builds of a mixed solve done incrementally, and BF16X3
        builds with single bf16 products."""
        out = (ct.c_uint64 * 5)()
        with self._lock:
            self._check(self._lib.vibeqc_fock_plan_last_solve_stats(self._handle, out, 5))
        return {"iterations": out[0], "reduced_iterations": out[1], "purified_densities": out[2],
                "incremental_builds": out[3], "single_term_builds": out[4]}

    def one_electron(self):
        """(S, Hcore) as nbf x nbf arrays."""
        n = self.nao
        S = np.empty((n, n))
        H = np.empty((n, n))
        with self._lock:
            self._ensure_open()
            self._check(self._lib.vibeqc_fock_plan_one_electron(self._handle, _native.pointer(S), _native.pointer(H), n * n))
        return S, H

    @property
    def diagnostics(self) -> dict:
        with self._lock:
            self._ensure_open()
            d = _native.descriptor(_native.FockDiagnosticC)
            self._check(self._lib.generativeqc_fock_plan_diagnostic(self._handle, ct.byref(d)))
            return {
                "backend": "hip",
                "nbf": d.nbf,
                "coordinate_count": d.coordinate_count,
                "device_bytes": d.device_bytes,
                "auxiliary_rank": d.auxiliary_rank,
                "auxiliary_functions": d.auxiliary_tile,
                "screening_tolerance": d.screening_tolerance,
                "metric_relative_threshold": d.metric_relative_threshold,
                "df_value_backend": d.df_value_backend.decode(),
                "df_value_mapping": d.df_value_mapping.decode(),
            }

    def set_precision(self, precision: str) -> None:
        if precision not in _native.PRECISION:
            raise ValueError("precision must be fp64, fp32 or bf16x3")
        with self._lock:
            self._ensure_open()
            self._check(self._lib.vibeqc_fock_plan_set_precision(self._handle, _native.PRECISION[precision]))
            self.precision = precision

    # -- operations ----------------------------------------------------------
    def evaluate(self, density, *, derivative: bool = False) -> FockEvaluation:
        """Raw J/K, Fock (with Hcore) and energies at a fixed density:
        [AO, AO] for restricted spin, [2, AO, AO] (alpha, beta) for unrestricted.

        With ``derivative=True`` (plan created with derivative_order 1) the
        result's ``gradient`` holds the fixed-density two-electron energy
        derivative dE2/dR, flat (3 natom), Hartree/bohr (fock.py:610-672)."""
        if type(derivative) is not bool:
            raise TypeError("derivative must be bool")
        n = self.nao
        separate = self.spec.spin == "unrestricted"
        d = np.array(density, dtype=np.float64, order="C", copy=True)
        if d.shape != ((2, n, n) if separate else (n, n)) or not np.isfinite(d).all():
            raise ValueError("Fock density shape/spin/finiteness mismatch")
        with self._lock:
            self._ensure_open()
            j = np.empty((n, n)) if self.spec.coulomb.present else None
            k = np.empty(d.shape) if self.spec.exchange.present else None
            f = np.empty(d.shape)
            g = np.empty(self.basis.natom * 3) if derivative else None
            out = _native.descriptor(
                _native.FockResultC,
                matrix_count=n * n,
                gradient_count=0 if g is None else g.size,
                coulomb=_native.pointer(j),
                exchange_alpha=_native.pointer(k[0] if separate and k is not None else k),
                exchange_beta=_native.pointer(k[1] if separate and k is not None else None),
                fock_alpha=_native.pointer(f[0] if separate else f),
                fock_beta=_native.pointer(f[1] if separate else None),
                gradient=_native.pointer(g),
            )
            status = self._lib.generativeqc_fock_plan_evaluate(
                self._handle,
                _native.pointer(d[0] if separate else d),
                n * n,
                _native.pointer(d[1] if separate else None),
                n * n if separate else 0,
                ct.byref(out),
            )
            self._check(status)
            return FockEvaluation(
                out.energy_one_electron,
                out.energy_two_electron,
                out.nuclear_repulsion,
                _immutable(f),
                _immutable(j) if j is not None else None,
                _immutable(k) if k is not None else None,
                _immutable(g) if g is not None else None,
            )

    def evaluate_occupied(self, occupied):
        """Raw (J, K) for D = 2 C C^T with C nbf x nocc."""
        n = self.nao
        c = np.array(occupied, dtype=np.float64, order="C", copy=True)
        if c.ndim != 2 or c.shape[0] != n:
            raise ValueError("occupied factor must be nbf x nocc")
        j = np.empty((n, n))
        k = np.empty((n, n))
        with self._lock:
            self._ensure_open()
            self._check(
                self._lib.vibeqc_fock_plan_evaluate_occupied(
                    self._handle, _native.pointer(c), n, c.shape[1], _native.pointer(j), _native.pointer(k)
                )
            )
        return j, k

    def benchmark_occupied(self, occupied, repeat: int = 10) -> dict:
        """Device-timed Fock builds from a resident occupied factor (ms per phase)."""
        n = self.nao
        c = np.array(occupied, dtype=np.float64, order="C", copy=True)
        ms = (ct.c_double * 5)()
        with self._lock:
            self._ensure_open()
            self._check(
                self._lib.vibeqc_fock_plan_benchmark_occupied(
                    self._handle, _native.pointer(c), n, c.shape[1], int(repeat), ms
                )
            )
        return {"exchange_u_ms": ms[0], "exchange_syrk_ms": ms[1], "coulomb_ms": ms[2], "build_ms": ms[3], "wall_ms": ms[4]}

    def autotune(self, occupied, repeat: int = 3) -> dict:
        """Time the occupied J/K build of ``occupied`` (AO x nocc) at the
        plan's precision for the default launch shape and its alternatives
        (U i-tile width, SYRK split-K count), and record the fastest in the
        process-wide tuning registry (``vibeqc_amd.tuning``), which later
        builds of this shape use.  Returns the chosen ``u_tile`` and
        ``syrk_splits`` with the best and default build times (ms)."""
        n = self.nao
        c = np.array(occupied, dtype=np.float64, order="C", copy=True)
        if c.ndim != 2 or c.shape[0] != n or not 0 < c.shape[1] <= n:
            raise ValueError("occupied factor must be (AO, nocc) with 0 < nocc <= AO")
        shape = (ct.c_int32 * 2)()
        ms = (ct.c_double * 2)()
        with self._lock:
            self._ensure_open()
            self._check(
                self._lib.vibeqc_fock_plan_autotune(self._handle, _native.pointer(c), n, c.shape[1], int(repeat), shape, ms)
            )
        return {"u_tile": shape[0], "syrk_splits": shape[1], "best_ms": ms[0], "default_ms": ms[1]}

    def launch_shape(self, nocc: int) -> dict:
        """The U i-tile width and SYRK split count that builds with ``nocc``
        occupied columns use at the current precision."""
        shape = (ct.c_int32 * 2)()
        with self._lock:
            self._ensure_open()
            self._check(self._lib.vibeqc_fock_plan_launch_shape(self._handle, int(nocc), shape))
        return {"u_tile": shape[0], "syrk_splits": shape[1]}

    def solve(
        self,
        *,
        initial_density=None,
        compute_forces: bool = False,
        max_iterations: int = 100,
        diis_history: int = 8,
        energy_tolerance: float = 1e-10,
        density_tolerance: float = 1e-8,
    ) -> FockScfResult:
        """Device-resident RHF (restricted spec) or UHF (unrestricted spec) SCF
        for the declared J/K energy (fock.py:674-776); UHF densities are
        [2, AO, AO] (alpha, beta).

        ``compute_forces`` returns the complete forces -dE/dR, (natom, 3),
        Hartree/bohr, at the converged density."""
        if type(compute_forces) is not bool:
            raise TypeError("compute_forces must be bool")
        n = self.nao
        shape = (2, n, n) if self.spec.spin == "unrestricted" else (n, n)
        seed = None
        if initial_density is not None:
            seed = np.array(initial_density, dtype=np.float64, order="C", copy=True)
            if seed.shape != shape or not np.isfinite(seed).all():
                raise ValueError("initial density shape/spin/finiteness mismatch")
        density = np.empty(shape)
        controls = _native.descriptor(
            _native.ScfControlsC,
            max_iterations=max_iterations,
            diis_history=diis_history,
            energy_tolerance=energy_tolerance,
            density_tolerance=density_tolerance,
        )
        forces = np.empty((self.basis.natom, 3)) if compute_forces else None
        out = _native.descriptor(
            _native.ScfResultC,
            density=_native.pointer(density),
            forces=_native.pointer(forces),
            density_count=density.size,
            force_count=0 if forces is None else forces.size,
        )
        with self._lock:
            self._ensure_open()
            status = self._lib.generativeqc_fock_plan_solve(
                self._handle, ct.byref(controls), _native.pointer(seed), 0 if seed is None else seed.size, ct.byref(out)
            )
            self._check(status)
        return FockScfResult(
            out.energy,
            _immutable(density),
            _immutable(forces) if forces is not None else None,
            out.iterations,
            out.energy_change,
            out.density_rms,
            bool(out.initial_density_used),
            out.fock_builds,
        )

    def close(self) -> None:
        with self._lock:
            if self._handle:
                self._lib.generativeqc_fock_plan_destroy(self._handle)
                self._handle = ct.c_void_p()
            self._ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *_):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
