"""Unrestricted (UHF) DF-JK on the HIP device against the UHF oracle: raw
J/K/Fock/energies and the fixed-density gradient of unrestricted plans, the
converged UHF energy, densities and forces (OH radical, cc-pVDZ), the HF+
exact-integral fixture, the one-electron and closed-shell limits, and the
mixed-precision stage."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

from _cases import Case, molecule_case
from test_oracle_uhf import OH, TIGHT, _electrons, _hf_plus_case, _radical

pytestmark = pytest.mark.gpu


def _plan(case, **kw):
    from vibeqc_amd.fock import FockBuildSpec, FockPlan

    return FockPlan(case.orbital(), FockBuildSpec.hf("unrestricted"), auxiliary=case.auxiliary(), **kw)


def _spin_densities(n, na, nb, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in (na, nb):
        c = rng.standard_normal((n, k)) * 0.3
        out.append(c @ c.T)
    return np.stack(out)


def test_unrestricted_evaluate_matches_oracle(oracle):
    case = _radical(np.array([a.position for a in OH]))
    p = case.oracle_problem(oracle)
    with _plan(case) as plan:
        n = plan.nao
        D = _spin_densities(n, 5, 4, seed=3)
        ev = plan.evaluate(D, derivative=True)
    J, Ka, Kb = oracle.uhf_jk(p.B, D[0], D[1])
    print(f"dJ {np.abs(ev.coulomb - J).max():.2e} dKa {np.abs(ev.exchange[0] - Ka).max():.2e} "
          f"dKb {np.abs(ev.exchange[1] - Kb).max():.2e}")
    assert np.abs(ev.coulomb - J).max() < 1e-10
    assert np.abs(ev.exchange[0] - Ka).max() < 1e-10
    assert np.abs(ev.exchange[1] - Kb).max() < 1e-10
    assert np.abs(ev.fock[0] - (p.H + J - Ka)).max() < 1e-10
    assert np.abs(ev.fock[1] - (p.H + J - Kb)).max() < 1e-10
    e2 = 0.5 * np.sum((D[0] + D[1]) * J) - 0.5 * np.sum(D[0] * Ka) - 0.5 * np.sum(D[1] * Kb)
    assert abs(ev.energy_two_electron - e2) < 1e-10
    assert abs(ev.energy_one_electron - np.sum((D[0] + D[1]) * p.H)) < 1e-10
    natom = len(case.atoms)
    shells, aux = case.oracle_shells(oracle), case.oracle_shells(oracle, True)
    g = oracle.df_two_electron_gradient(case.xyz, natom, shells, True, aux, True, p.A, p.M, D[0] + D[1], 1.0, 0.0)
    for s in range(2):
        g += oracle.df_two_electron_gradient(case.xyz, natom, shells, True, aux, True, p.A, p.M, D[s], 0.0, -1.0)
    print(f"dgrad {np.abs(ev.gradient.reshape(-1, 3) - g).max():.2e}")
    assert np.abs(ev.gradient.reshape(-1, 3) - g).max() < 1e-9


def test_uhf_radical_energy_density_and_forces(oracle):
    from vibeqc_amd import Calculator

    case = _radical(np.array([a.position for a in OH]))
    p = case.oracle_problem(oracle)
    na, nb = oracle.spin_occupations(_electrons(case), 2)
    ref = oracle.df_uhf(p, na, nb, **TIGHT)
    g = oracle.df_uhf_gradient(
        case.z, case.xyz, case.oracle_shells(oracle), True, case.oracle_shells(oracle, True), True, p, ref, (na, nb)
    )
    calc = Calculator(method="uhf", energy_tolerance=1e-12, density_tolerance=1e-10)
    r = calc.singlepoint(case.atoms, multiplicity=2)
    print(f"E {r.energy:.12f} oracle {ref.energy:.12f} dE {r.energy - ref.energy:.2e} it={r.iterations} "
          f"dF {np.abs(np.asarray(r.forces) + g).max():.2e}")
    assert r.converged
    assert abs(r.energy - ref.energy) < 1e-9
    # The radical's pi hole may sit in any combination of the degenerate pi
    # pair, so the densities are checked through their invariants.
    D = np.asarray(r.density)
    assert D.shape == (2, p.S.shape[0], p.S.shape[0])
    for s, k in enumerate((na, nb)):
        assert abs(np.trace(D[s] @ p.S) - k) < 1e-9
        assert np.abs(D[s] @ p.S @ D[s] - D[s]).max() < 1e-8
    assert np.abs(np.asarray(r.forces) + g).max() < 1e-7


def test_uhf_hf_plus_fixture(oracle):
    from vibeqc_amd import Calculator

    case, data = _hf_plus_case()
    p = case.oracle_problem(oracle)
    na, nb = oracle.spin_occupations(_electrons(case), case.multiplicity)
    ref = oracle.df_uhf(p, na, nb, **TIGHT)
    calc = Calculator(
        method="uhf", basis=case.shells, auxiliary_basis=case.aux_shells, basis_representation="cartesian",
        energy_tolerance=1e-12, density_tolerance=1e-10,
    )
    r = calc.singlepoint(case.atoms, charge=1, multiplicity=2, properties=("energy",))
    print(f"E {r.energy:.12f} oracle {ref.energy:.12f} exact {data['energy']:.12f}")
    assert abs(r.energy - ref.energy) < 1e-8
    assert abs(r.energy - data["energy"]) < 1e-6
    # Seeded with the fixture's densities the solve stays in the fixture's
    # orientation of the degenerate pi hole.
    with calc.prepare(case.atoms, charge=1, multiplicity=2) as plan:
        seeded = plan.solve(initial_density=np.array(data["density"]), energy_tolerance=1e-12, density_tolerance=1e-10)
    assert abs(seeded.energy - ref.energy) < 1e-8
    assert np.abs(np.asarray(seeded.density) - np.array(data["density"])).max() < 1e-5


@pytest.mark.parametrize("charge", [0, 1])
def test_one_electron_limit(oracle, charge):
    from vibeqc_amd import Calculator
    from vibeqc_amd.basis import Atom

    atoms = (Atom(1, (0.0, 0.0, 0.0)),) if charge == 0 else (Atom(1, (0.0, 0.0, -0.7)), Atom(1, (0.0, 0.0, 0.7)))
    case = molecule_case("one", atoms)
    case.charge = charge
    p = case.oracle_problem(oracle)
    e0 = scipy.linalg.eigh(p.H, p.S, eigvals_only=True)[0] + p.enuc
    r = Calculator(method="uhf").singlepoint(atoms, charge=charge, multiplicity=2)
    assert abs(r.energy - e0) < 1e-9
    assert np.abs(np.asarray(r.density)[1]).max() == 0.0


def test_singlet_uhf_equals_rhf():
    from vibeqc_amd import Calculator
    from vibeqc_amd.geometry import water_cluster

    atoms = water_cluster(2, 3.5, seed=11)
    rhf = Calculator().singlepoint(atoms)
    uhf = Calculator(method="uhf").singlepoint(atoms)
    assert abs(uhf.energy - rhf.energy) < 1e-9
    assert np.abs(np.asarray(uhf.forces) - np.asarray(rhf.forces)).max() < 1e-7
    for s in range(2):
        assert np.abs(np.asarray(uhf.density)[s] - 0.5 * np.asarray(rhf.density)).max() < 1e-6


def test_mixed_precision_uhf_and_seeded_restart():
    from vibeqc_amd.basis import Atom
    from vibeqc_amd import Calculator
    from vibeqc_amd.geometry import water_cluster

    # (H2O)2 + OH: a doublet with 67 basis functions.
    atoms = tuple(water_cluster(2, 3.5, seed=11)) + (Atom(8, (7.0, 0.0, 0.0)), Atom(1, (7.0, 0.0, 1.83)))
    e64 = Calculator(method="uhf").singlepoint(atoms, multiplicity=2, properties=("energy",))
    mixed = Calculator(method="uhf", precision="mixed").singlepoint(atoms, multiplicity=2, properties=("energy",))
    print(f"fp64 {e64.energy:.12f} it={e64.iterations} mixed {mixed.energy:.12f} it={mixed.iterations}")
    assert abs(mixed.energy - e64.energy) < 1e-6
    calc = Calculator(method="uhf")
    with calc.prepare(atoms, multiplicity=2) as plan:
        again = plan.solve(initial_density=np.asarray(e64.density))
    assert abs(again.energy - e64.energy) < 1e-9
    assert again.iterations <= 3


def test_uhf_batch_matches_single_plans_and_checkpoints(tmp_path):
    from vibeqc_amd import Calculator
    from vibeqc_amd.basis import Atom
    from vibeqc_amd.checkpoint import inspect_checkpoint
    from vibeqc_amd.geometry import water_cluster

    systems = [
        OH,
        tuple(water_cluster(1, 3.0, seed=5)),
        (Atom(1, (0.0, 0.0, 0.0)),),
        (Atom(8, (0.0, 0.0, 0.0)), Atom(1, (0.0, 0.2, 1.9))),
    ]
    mult = [2, 1, 2, 2]
    calc = Calculator(method="uhf")
    batch = calc.batch_singlepoint(systems, multiplicities=mult)
    assert batch.succeeded
    for item, atoms, m in zip(batch.items, systems, mult):
        single = calc.singlepoint(atoms, multiplicity=m)
        print(f"item {item.index}: E={item.energy:.10f} dE={item.energy - single.energy:.2e} it={item.iterations}")
        assert abs(item.energy - single.energy) < 1e-9
        assert np.abs(np.asarray(item.forces) - np.asarray(single.forces)).max() < 1e-8
    path = tmp_path / "uhf.vqcp"
    with calc.prepare_batch(systems, multiplicities=mult) as prepared:
        first = prepared.execute(properties=("energy",), strict=True)
        prepared.save_checkpoint(path)
    doc = inspect_checkpoint(path)
    n = [it["seed"]["nbf"] for it in doc["items"]]
    assert [b["shape"] for b in doc["blobs"] if b["name"].startswith("density")] == [[2, k, k] for k in n]
    with calc.prepare_batch(systems, multiplicities=mult) as prepared:
        prepared.load_checkpoint(path)
        restored = prepared.execute(properties=("energy",), strict=True)
    for a, b in zip(first.items, restored.items):
        assert b.restart_origin == "persistent_restart"
        assert abs(a.energy - b.energy) < 1e-9
        assert b.iterations <= 3


def test_near_degenerate_radical_state_follows_the_seed(oracle):
    """(H2O)2 ... OH: the OH pi pair is nearly degenerate, so the core guess's
    frontier rotation acts inside an almost arbitrary eigenbasis and cold
    starts may reach different UHF stationary points (the oracle's numpy and
    the device's rocSOLVER eigenvectors differ there).  Seeded with the
    oracle's converged densities, the device SCF stays at the oracle's state."""
    from vibeqc_amd import Calculator
    from vibeqc_amd.basis import Atom
    from vibeqc_amd.geometry import water_cluster

    atoms = tuple(water_cluster(2, 3.5, seed=11)) + (Atom(8, (7.0, 0.0, 0.0)), Atom(1, (7.0, 0.0, 1.83)))
    case = molecule_case("cluster", atoms)
    case.multiplicity = 2
    p = case.oracle_problem(oracle)
    na, nb = oracle.spin_occupations(_electrons(case), 2)
    ref = oracle.df_uhf(p, na, nb, **TIGHT)
    calc = Calculator(method="uhf", energy_tolerance=1e-12, density_tolerance=1e-10)
    with calc.prepare(atoms, multiplicity=2) as plan:
        r = plan.solve(initial_density=ref.density, energy_tolerance=1e-12, density_tolerance=1e-10)
    print(f"seeded E {r.energy:.12f} oracle {ref.energy:.12f} it={r.iterations}")
    assert abs(r.energy - ref.energy) < 1e-9
    assert np.abs(np.asarray(r.density) - ref.density).max() < 1e-6


@pytest.mark.parametrize(
    "name,atoms,charge,multiplicity",
    [
        ("o2_triplet", ((8, (0.0, 0.0, -1.14)), (8, (0.1, 0.05, 1.14))), 0, 3),
        ("water_cation", ((8, (0.0, 0.0, 0.0)), (1, (0.0, 1.43, 1.11)), (1, (0.0, -1.43, 1.11))), 1, 2),
    ],
)
def test_other_spin_states_match_oracle(oracle, name, atoms, charge, multiplicity):
    """A triplet (n_alpha - n_beta = 2) and a cation doublet from cold starts:
    energy to 1e-9 Eh and forces to 1e-7 Eh/bohr of the UHF oracle."""
    from vibeqc_amd import Calculator
    from vibeqc_amd.basis import Atom

    atoms = tuple(Atom(z, xyz) for z, xyz in atoms)
    case = molecule_case(name, atoms)
    case.charge, case.multiplicity = charge, multiplicity
    p = case.oracle_problem(oracle)
    na, nb = oracle.spin_occupations(_electrons(case), multiplicity)
    ref = oracle.df_uhf(p, na, nb, **TIGHT)
    g = oracle.df_uhf_gradient(
        case.z, case.xyz, case.oracle_shells(oracle), True, case.oracle_shells(oracle, True), True, p, ref, (na, nb)
    )
    r = Calculator(method="uhf", energy_tolerance=1e-12, density_tolerance=1e-10).singlepoint(
        atoms, charge=charge, multiplicity=multiplicity
    )
    print(f"{name}: E {r.energy:.12f} oracle {ref.energy:.12f} dF {np.abs(np.asarray(r.forces) + g).max():.2e}")
    assert abs(r.energy - ref.energy) < 1e-9
    assert np.abs(np.asarray(r.forces) + g).max() < 1e-7


def test_mixed_precision_uhf_forces():
    """BF16X3 iterations + FP64 refinement: energy within 1 uEh and forces
    within 1e-6 Eh/bohr of the FP64 UHF (the forces come from FP64 builds)."""
    from vibeqc_amd import Calculator

    atoms = ((8, (0.0, 0.0, -1.14)), (8, (0.1, 0.05, 1.14)))
    e64 = Calculator(method="uhf").singlepoint(atoms, multiplicity=3)
    mixed = Calculator(method="uhf", precision="mixed").singlepoint(atoms, multiplicity=3)
    print(f"dE {mixed.energy - e64.energy:.2e} dF {np.abs(np.asarray(mixed.forces) - np.asarray(e64.forces)).max():.2e}")
    assert abs(mixed.energy - e64.energy) < 1e-6
    assert np.abs(np.asarray(mixed.forces) - np.asarray(e64.forces)).max() < 1e-6


def test_uhf_batch_with_forces_eight_radicals():
    """Eight OH doublets (def2-SVP / cc-pVDZ-JKFIT, d shells on O) as one UHF
    batch with forces, executed three times in one process: every execution
    succeeds and each member's energy and forces equal the single-plan ones
    (1e-9 Eh, 1e-8 Eh/bohr); repeated executions agree bitwise."""
    from vibeqc_amd import Calculator
    from vibeqc_amd.basis import Atom

    systems = [(Atom(8, (0.0, 0.0, 0.0)), Atom(1, (0.0, 0.03 * k, 1.76 + 0.02 * k))) for k in range(8)]
    calc = Calculator(method="uhf", basis="def2-svp")
    runs = []
    with calc.prepare_batch(systems, multiplicities=[2] * 8, warm_start=False) as batch:
        for _ in range(3):
            res = batch.execute(strict=True)
            runs.append((np.asarray(res.energies), np.stack([np.asarray(it.forces) for it in res.items])))
    for e, f in runs[1:]:
        assert np.array_equal(e, runs[0][0]) and np.array_equal(f, runs[0][1])
    for k, atoms in enumerate(systems):
        single = calc.singlepoint(atoms, multiplicity=2)
        de = runs[0][0][k] - single.energy
        df = np.abs(runs[0][1][k] - np.asarray(single.forces)).max()
        print(f"OH {k}: E {single.energy:.10f} dE {de:.1e} dF {df:.1e}")
        assert abs(de) < 1e-9 and df < 1e-8
