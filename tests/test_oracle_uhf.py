"""The UHF oracle pinned without retyping its formulas:

* one electron (H atom, H2+): J and K of a rank-one density cancel exactly in
  DF, so the UHF energy is the lowest generalized eigenvalue of (H, S) plus
  E_nuc;
* a closed-shell singlet: UHF reduces to the (golden-pinned) RHF oracle, with
  each spin density half the RHF density;
* HF+ (doublet): the reference's exact-integral PySCF UHF fixture.  The
  one-electron matrices match to 1e-10.  With a large even-tempered fitting
  set (DF error about 5e-7 Eh here) the DF energy at the fixture's densities
  and the converged DF-UHF energy, densities and orbital energies agree with
  the exact ones, and the core guess (with its beta frontier rotation)
  reaches the same state as the fixture's densities;
* the analytic DF-UHF gradient of the OH radical matches central finite
  differences of the oracle's converged energy.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

from _cases import Case, molecule_case
from conftest import golden_json
from vibeqc_amd.basis import Atom, Primitive, Shell, shells_for

TIGHT = dict(energy_tolerance=1e-12, density_tolerance=1e-10, max_iterations=200)


def even_tempered(atom: int, lmax: int, lo: float, hi: float, count: int):
    """Uncontracted fitting shells: count - 2l exponents from lo to hi / 4^l per l."""
    return [
        Shell(atom, l, (Primitive(float(e), 1.0),))
        for l in range(lmax + 1)
        for e in np.geomspace(lo, hi / 4**l, max(2, count - 2 * l))
    ]


def _hf_plus_case():
    meta = golden_json("uhf_hf_plus.json")
    inp = meta["inputs"]
    atoms = tuple(Atom(int(z), tuple(map(float, p))) for z, p in zip(inp["atomic_numbers"], inp["coordinates"]))
    shells = tuple(
        Shell(r["atom_index"], r["angular_momentum"], tuple(Primitive(float(e), float(c)) for e, c in r["primitives"]))
        for r in inp["shells"]
    )
    aux = tuple(even_tempered(0, 3, 0.05, 3000.0, 18) + even_tempered(1, 3, 0.05, 600.0, 14))
    case = Case("hf_plus", atoms, shells, aux, "cartesian", inp["charge"], inp["multiplicity"])
    return case, meta["data"]


def _electrons(case):
    return sum(case.z) - case.charge


@pytest.mark.parametrize(
    "atoms,charge",
    [
        ((Atom(1, (0.0, 0.0, 0.0)),), 0),
        ((Atom(1, (0.0, 0.0, -0.7)), Atom(1, (0.0, 0.0, 0.7))), 1),
    ],
)
def test_one_electron_uhf_is_the_core_eigenvalue(oracle, atoms, charge):
    case = molecule_case("one_electron", atoms)
    case.charge = charge
    case.multiplicity = 2
    p = case.oracle_problem(oracle)
    na, nb = oracle.spin_occupations(_electrons(case), 2)
    assert (na, nb) == (1, 0)
    r = oracle.df_uhf(p, na, nb, **TIGHT)
    assert r.converged
    e0 = scipy.linalg.eigh(p.H, p.S, eigvals_only=True)[0]
    assert abs(r.energy - (e0 + p.enuc)) < 1e-10
    assert np.abs(r.density[1]).max() == 0.0


def test_singlet_uhf_equals_rhf(oracle):
    from vibeqc_amd.geometry import water_cluster

    case = molecule_case("water", water_cluster(1, 3.0, seed=7))
    p = case.oracle_problem(oracle)
    rhf = oracle.df_rhf(p, **TIGHT)
    uhf = oracle.df_uhf(p, p.nocc, p.nocc, **TIGHT)
    assert rhf.converged and uhf.converged
    assert abs(uhf.energy - rhf.energy) < 1e-9
    for s in range(2):
        assert np.abs(uhf.density[s] - 0.5 * rhf.density).max() < 1e-7


def test_hf_plus_against_exact_uhf_fixture(oracle):
    case, data = _hf_plus_case()
    p = case.oracle_problem(oracle)
    assert np.abs(p.S - np.array(data["overlap"])).max() < 1e-10
    assert np.abs(p.H - np.array(data["hcore"])).max() < 1e-10
    assert abs(p.enuc - data["nuclear_repulsion"]) < 1e-12
    na, nb = oracle.spin_occupations(_electrons(case), case.multiplicity)
    occ = np.array(data["mo_occ"])
    assert (na, nb) == (int(occ[0].sum()), int(occ[1].sum())) == (5, 4)
    D = np.array(data["density"])
    # DF energy at the exact solution's densities.
    J, Ka, Kb = oracle.uhf_jk(p.B, D[0], D[1])
    e_at = oracle.uhf_energy(D[0], D[1], p.H, p.H + J - Ka, p.H + J - Kb) + p.enuc
    assert abs(e_at - data["energy"]) < 1e-6
    seeded = oracle.df_uhf(p, na, nb, initial_density=D, **TIGHT)
    cold = oracle.df_uhf(p, na, nb, **TIGHT)
    assert seeded.converged and cold.converged
    assert abs(seeded.energy - data["energy"]) < 1e-6
    assert seeded.energy <= e_at + 1e-10  # variational in the DF model
    assert abs(cold.energy - seeded.energy) < 1e-9
    assert np.abs(seeded.density - D).max() < 1e-5
    assert np.abs(seeded.orbital_energies - np.array(data["mo_energy"])).max() < 1e-4


OH = (Atom(8, (0.0, 0.0, 0.0)), Atom(1, (0.3, 0.1, 1.83)))


def _radical(xyz):
    atoms = tuple(Atom(a.atomic_number, tuple(map(float, x))) for a, x in zip(OH, xyz))
    case = molecule_case("oh", atoms)
    case.multiplicity = 2
    return case


def test_uhf_gradient_matches_finite_differences(oracle):
    base = _radical(np.array([a.position for a in OH]))
    na, nb = oracle.spin_occupations(_electrons(base), 2)

    def energy(xyz):
        c = _radical(xyz)
        p = c.oracle_problem(oracle)
        return oracle.df_uhf(p, na, nb, **TIGHT), p, c

    r, p, c = energy(base.xyz)
    assert r.converged and p.rank == p.M.shape[0]
    g = oracle.df_uhf_gradient(c.z, c.xyz, c.oracle_shells(oracle), True, c.oracle_shells(oracle, True), True, p, r, (na, nb))
    assert np.abs(g.sum(axis=0)).max() < 1e-8  # translational invariance
    h = 1e-4
    for atom, axis in ((0, 2), (1, 0)):
        plus, minus = base.xyz.copy(), base.xyz.copy()
        plus[atom, axis] += h
        minus[atom, axis] -= h
        fd = (energy(plus)[0].energy - energy(minus)[0].energy) / (2 * h)
        assert abs(g[atom, axis] - fd) < 2e-7, (atom, axis, g[atom, axis], fd)
