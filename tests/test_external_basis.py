"""One-electron matrices against the reference's external-basis oracle
(tests/data/external_basis/oracles.npz: water STO-3G overlap, kinetic and
core Hamiltonian from PySCF/libcint, Cartesian and spherical, at the case
geometry and at a changed geometry).  The CPU test checks the oracle; the
GPU test checks the device one-electron kernels through the plan."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import golden_json
from vibeqc_amd.basis import NUMBERS, Atom, shells_for

CASES = [(name, geom) for name in ("water_cartesian", "water_spherical") for geom in ("atoms", "changed_atoms")]


def _case(name, geom):
    rec = golden_json("external_water_sto3g.json")[name]
    atoms = tuple(Atom(NUMBERS[sym], tuple(map(float, xyz))) for sym, xyz in rec[geom])
    tag = "" if geom == "atoms" else "changed_"
    ref = {k: np.array(rec[tag + k]) for k in ("overlap", "kinetic", "hcore")}
    return atoms, rec["representation"] == "spherical", ref


@pytest.mark.parametrize("name,geom", CASES)
def test_oracle_one_electron_matches_external_oracle(oracle, name, geom):
    atoms, spherical, ref = _case(name, geom)
    shells = shells_for("sto-3g", atoms)
    oshells = [oracle.Shell(s.atom_index, s.l, tuple(p.exponent for p in s.primitives), tuple(p.coefficient for p in s.primitives)) for s in shells]
    z = [a.atomic_number for a in atoms]
    S, T, V = oracle.one_electron(z, np.array([a.position for a in atoms]), oshells, spherical)
    assert np.abs(S - ref["overlap"]).max() < 1e-10
    assert np.abs(T - ref["kinetic"]).max() < 1e-10
    assert np.abs(T + V - ref["hcore"]).max() < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name,geom", CASES)
def test_device_one_electron_matches_external_oracle(name, geom):
    from vibeqc_amd.basis import AOBasis
    from vibeqc_amd.fock import FockBuildSpec, FockPlan

    atoms, spherical, ref = _case(name, geom)
    rep = "real_spherical" if spherical else "cartesian"
    orb = AOBasis(atoms, shells_for("sto-3g", atoms), rep, 0, 1)
    with FockPlan(orb, FockBuildSpec.hf(), auxiliary=orb) as plan:
        S, H = plan.one_electron()
    assert np.abs(np.asarray(S) - ref["overlap"]).max() < 1e-10
    assert np.abs(np.asarray(H) - ref["hcore"]).max() < 1e-10
