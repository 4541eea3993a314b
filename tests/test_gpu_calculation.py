"""The native calculation ABI (generativeqc_calculation_prepare / _execute /
_destroy, reference generativeqc.h:1335-1347) bound through ctypes the way
the reference's Calculator binds it (python/generativeqc/calculator.py:
2291-2346): a context, an orbital system and an auxiliary system, a method
descriptor, a caller-owned result descriptor with optional force storage.

Energies against the reference's GPU4PySCF cc-pVDZ / cc-pVDZ-JKFIT water
fixture (1e-8 Eh, forces 1e-7 Eh/bohr); status semantics of
c_api_calculation.cpp:24-131 (ABI mismatch, short force buffer, energy-only
request, NOT_CONVERGED with diagnostics, exact J/K refused); UHF through the
same entry points against the prepared-plan UHF solve (1e-10 Eh)."""

from __future__ import annotations

import ctypes as ct

import numpy as np
import pytest

from _cases import ccpvdz_case

pytestmark = pytest.mark.gpu


class _Native:
    def __init__(self, case):
        from vibeqc_amd import _native
        from vibeqc_amd.fock import NativeContext

        self.N = _native
        self.ctx = NativeContext(0)
        self.lib = self.ctx.lib
        self.system = self.ctx.system(case.orbital())
        self.aux = self.ctx.system(case.auxiliary())
        self.calc = ct.c_void_p()

    def descriptor(self, **overrides):
        N = self.N
        values = dict(
            method=N.METHOD_RHF,
            max_iterations=100,
            diis_history=8,
            energy_tolerance=1e-10,
            density_tolerance=1e-8,
            screening_tolerance=1e-12,
            density_fitting_mode=N.DENSITY_FITTING_CUDA,
            density_fitting_auxiliary_basis=self.aux,
            density_fitting_relative_threshold=1e-10,
            density_fitting_memory_budget_bytes=0,
            precision_mode=N.PRECISION_MODE["fp64"],
        )
        values.update(overrides)
        return N.descriptor(N.MethodDescriptorC, **values)

    def prepare(self, desc):
        if self.calc:
            self.lib.generativeqc_calculation_destroy(self.calc)
            self.calc = ct.c_void_p()
        return self.lib.generativeqc_calculation_prepare(self.ctx.handle, self.system, ct.byref(desc), ct.byref(self.calc))

    def result(self, forces=None, count=None):
        N = self.N
        return N.descriptor(
            N.ResultDescriptorC,
            forces=N.pointer(forces),
            force_count=(0 if forces is None else forces.size) if count is None else count,
            executed_backend=N.BACKEND_CPU_REFERENCE,
        )

    def detail(self):
        return self.lib.generativeqc_context_get_last_detail(self.ctx.handle)

    def close(self):
        if self.calc:
            self.lib.generativeqc_calculation_destroy(self.calc)
        self.lib.generativeqc_system_destroy(self.system)
        self.lib.generativeqc_system_destroy(self.aux)
        self.ctx.close()


@pytest.fixture
def water():
    case, meta = ccpvdz_case("ccpvdz_water")
    native = _Native(case)
    yield native, meta
    native.close()


def test_energy_and_forces_match_reference_fixture(water):
    native, meta = water
    assert native.prepare(native.descriptor()) == 0
    forces = np.full((3, 3), np.nan)
    res = native.result(forces)
    assert native.lib.generativeqc_calculation_execute(native.calc, ct.byref(res)) == 0
    assert res.converged == 1 and res.executed_backend == native.N.BACKEND_HIP
    assert res.iterations > 1 and res.density_rms < 1e-8 and res.energy_change < 1e-10
    ref = np.array(meta["forces_hartree_per_bohr"])
    print(f"calculation ABI water: dE {res.energy - meta['energy_hartree']:.2e}, max dF {np.abs(forces - ref).max():.2e}")
    assert abs(res.energy - meta["energy_hartree"]) < 1e-8
    assert np.abs(forces - ref).max() < 1e-7
    # Energy only: NULL forces with count 0; a second execution is a fresh SCF.
    res2 = native.result()
    assert native.lib.generativeqc_calculation_execute(native.calc, ct.byref(res2)) == 0
    assert res2.energy == res.energy and res2.iterations == res.iterations


def test_status_semantics(water):
    native, _ = water
    N, lib = native.N, native.lib
    bad = native.descriptor()
    bad.struct_size = 8
    calc = ct.c_void_p(123)
    assert lib.generativeqc_calculation_prepare(native.ctx.handle, native.system, ct.byref(bad), ct.byref(calc)) == 2
    assert not calc  # nulled on failure
    assert native.prepare(native.descriptor(density_fitting_mode=N.DENSITY_FITTING_NONE)) == 3
    assert not native.calc
    assert native.prepare(native.descriptor(density_fitting_relative_threshold=2.0)) == 1
    assert native.prepare(native.descriptor()) == 0
    short = np.zeros(5)
    assert lib.generativeqc_calculation_execute(native.calc, ct.byref(native.result(short, count=5))) == 1
    assert lib.generativeqc_calculation_execute(native.calc, ct.byref(native.result(None, count=9))) == 1
    mismatch = native.result()
    mismatch.abi_version = 7
    assert lib.generativeqc_calculation_execute(native.calc, ct.byref(mismatch)) == 2
    assert lib.generativeqc_calculation_execute(native.calc, None) == 1
    # One iteration cannot converge: NOT_CONVERGED with the diagnostics written.
    assert native.prepare(native.descriptor(max_iterations=1)) == 0
    forces = np.full((3, 3), 7.0)
    res = native.result(forces)
    assert lib.generativeqc_calculation_execute(native.calc, ct.byref(res)) == 4
    assert res.converged == 0 and res.iterations == 1 and np.isfinite(res.energy)
    assert b"converge" in native.detail()
    assert np.all(forces == 7.0)  # forces are written only on success


def test_mixed_precision_and_calculator_route(water):
    from vibeqc_amd import Calculator

    native, meta = water
    assert native.prepare(native.descriptor(precision_mode=native.N.PRECISION_MODE["auto"])) == 0
    res = native.result()
    assert native.lib.generativeqc_calculation_execute(native.calc, ct.byref(res)) == 0
    assert abs(res.energy - meta["energy_hartree"]) < 1e-8
    # Calculator.singlepoint runs through the same entry points.
    case, _ = ccpvdz_case("ccpvdz_water")
    r = Calculator(precision="mixed").singlepoint(case.atoms)
    assert abs(r.energy - res.energy) < 1e-9
    assert r.density.shape == (24, 24) and r.nbf == 24 and r.naux == 116 and r.fock_builds > r.iterations


def test_uhf_through_the_calculation_abi():
    from vibeqc_amd import Calculator
    from vibeqc_amd.basis import Atom

    atoms = (Atom(8, (0.0, 0.0, 0.0)), Atom(1, (0.0, 0.0, 1.83)))
    calc = Calculator(method="uhf", energy_tolerance=1e-11, density_tolerance=1e-9)
    r = calc.singlepoint(atoms, multiplicity=2)
    with calc.prepare(atoms, multiplicity=2) as plan:
        ref = plan.solve(compute_forces=True, energy_tolerance=1e-11, density_tolerance=1e-9)
    print(f"OH UHF: calculation {r.energy:.10f} plan {ref.energy:.10f}")
    assert r.density.shape == (2, 19, 19)
    assert abs(r.energy - ref.energy) < 1e-10
    assert np.abs(np.asarray(r.forces) - np.asarray(ref.forces)).max() < 1e-8
