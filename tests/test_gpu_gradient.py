"""Analytic DF-RHF nuclear gradients on the HIP device, against the oracle,
the reference's GPU4PySCF force fixtures and finite differences.

Tolerances:
  * fixed-density two-electron gradient vs the oracle (same density, FP64 both
    sides, different summation order): 1e-9 relative to the largest component;
  * converged forces vs GPU4PySCF cc-pVDZ/cc-pVDZ-JKFIT forces: 1e-7 Eh/bohr
    (the reference's own forces agree with them to ~2e-9; SCF convergence to
    1e-10 Eh / 1e-8 RMS density bounds ours);
  * (H2O)8 forces vs central differences of the energy (h = 1e-3 bohr): 1e-5
    Eh/bohr (O(h^2) truncation); mixed-precision forces vs FP64: 1e-6 Eh/bohr.
"""

from __future__ import annotations

import numpy as np
import pytest

from _cases import ccpvdz_case, molecule_case, posthf_case

pytestmark = pytest.mark.gpu


def _plan(case, **kw):
    from vibeqc_amd import FockBuildSpec, FockPlan

    return FockPlan(case.orbital(), FockBuildSpec.hf(), auxiliary=case.auxiliary(), **kw)


@pytest.mark.parametrize("name", ["ccpvdz_water", "ccpvdz_water4"])
def test_forces_match_reference_forces(name):
    case, meta = ccpvdz_case(name)
    with _plan(case) as plan:
        r = plan.solve(compute_forces=True)
    ref = np.array(meta["forces_hartree_per_bohr"])
    err = np.abs(r.forces - ref).max()
    print(f"{name}: max |F - F_ref| = {err:.3e} Eh/bohr, sum F = {np.abs(r.forces.sum(0)).max():.2e}")
    assert r.forces.shape == ref.shape
    assert err < 1e-7
    assert np.abs(r.forces.sum(axis=0)).max() < 1e-9


@pytest.mark.parametrize("name,kind", [("water", "posthf"), ("ccpvdz_water", "ccpvdz")])
@pytest.mark.parametrize("indefinite", [False, True])
def test_fixed_density_gradient_matches_oracle(oracle, name, kind, indefinite):
    """evaluate(derivative=True) returns dE2/dR at the given density; the
    posthf water case is Cartesian, the cc-pVDZ one real-spherical."""
    case = posthf_case(name)[0] if kind == "posthf" else ccpvdz_case(name)[0]
    sh, ax = case.oracle_shells(oracle), case.oracle_shells(oracle, True)
    problem = case.oracle_problem(oracle)
    n = problem.S.shape[0]
    rng = np.random.default_rng(5)
    if indefinite:
        D = rng.normal(size=(n, n))
        D = 0.05 * (D + D.T)
    else:
        D = oracle.df_rhf(problem).density
    ref = oracle.df_two_electron_gradient(case.xyz, len(case.atoms), sh, case.spherical, ax, case.spherical,
                                          problem.A, problem.M, D)
    with _plan(case) as plan:
        ev = plan.evaluate(D, derivative=True)
    g = np.asarray(ev.gradient).reshape(-1, 3)
    err = np.abs(g - ref).max()
    print(f"{name} indefinite={indefinite}: max |g - g_oracle| = {err:.3e} (max |g| {np.abs(ref).max():.3e})")
    assert err < 1e-9 * max(1.0, np.abs(ref).max())


def test_forces_are_deterministic():
    case, _ = ccpvdz_case("ccpvdz_water4")
    with _plan(case) as plan:
        f1 = plan.solve(compute_forces=True).forces
        f2 = plan.solve(compute_forces=True).forces
    assert np.array_equal(np.asarray(f1), np.asarray(f2))


def test_water8_forces_central_differences_and_mixed_precision():
    """(H2O)8, 8 A cube, N(0, 0.05 A) displacements, cc-pVDZ/cc-pVDZ-JKFIT:
    FP64 forces against central differences of the energy on a sample of
    coordinates, and mixed-precision (BF16X3 then FP64 refinement) forces
    against FP64."""
    from vibeqc_amd.geometry import water_cluster

    atoms = water_cluster(8, 8.0, seed=3, displacement_angstrom=0.05)
    case = molecule_case("water8", atoms)
    tight = dict(energy_tolerance=1e-11, density_tolerance=1e-9)
    with _plan(case) as plan:
        r64 = plan.solve(compute_forces=True, **tight)
    with _plan(case, precision="bf16x3") as plan:
        rmx = plan.solve(compute_forces=True, **tight)
    dmix = np.abs(np.asarray(rmx.forces) - np.asarray(r64.forces)).max()
    print(f"(H2O)8: E64 {r64.energy:.10f}, Emixed - E64 = {rmx.energy - r64.energy:.2e}, max |dF| mixed = {dmix:.2e}")
    assert abs(rmx.energy - r64.energy) < 1e-6
    assert dmix < 1e-6

    xyz = case.xyz
    h = 1e-3

    def energy(shifted):
        moved = tuple(type(a)(a.atomic_number, tuple(map(float, p))) for a, p in zip(case.atoms, shifted))
        with _plan(molecule_case("water8", moved)) as plan:
            return plan.solve(**tight).energy

    for a, x in [(0, 0), (1, 1), (2, 2), (9, 0), (13, 1), (23, 2)]:
        xp, xm = xyz.copy(), xyz.copy()
        xp[a, x] += h
        xm[a, x] -= h
        fd = -(energy(xp) - energy(xm)) / (2 * h)
        print(f"atom {a} axis {x}: analytic {r64.forces[a, x]: .8f} central difference {fd: .8f}")
        assert abs(r64.forces[a, x] - fd) < 1e-5


def test_torch_energy_backward_is_minus_forces():
    import torch

    from vibeqc_amd import Calculator
    from vibeqc_amd.torch import energy

    case, meta = ccpvdz_case("ccpvdz_water")
    coords = torch.tensor(case.xyz, dtype=torch.float64, requires_grad=True)
    calc = Calculator(method="rhf", basis="cc-pvdz", auxiliary_basis="cc-pvdz-jkfit")
    e = energy(coords, case.z, calc)
    e.backward()
    ref = np.array(meta["forces_hartree_per_bohr"])
    assert abs(e.item() - meta["energy_hartree"]) < 1e-8
    assert np.abs(-coords.grad.numpy() - ref).max() < 1e-7


def test_torch_density_energy_backward_is_fock():
    import torch

    from vibeqc_amd.torch import density_energy

    case, _ = ccpvdz_case("ccpvdz_water")
    with _plan(case) as plan:
        r = plan.solve()
        D = torch.tensor(np.asarray(r.density), requires_grad=True)
        e = density_energy(plan, D)
        e.backward()
        F = plan.evaluate(np.asarray(r.density)).fock
    assert abs(e.item() - r.energy) < 1e-8
    assert np.abs(D.grad.numpy() - F).max() < 1e-12
