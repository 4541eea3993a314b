"""GPU parity: the HIP engine (through the C ABI) against the FP64 oracle and the
reference's golden vectors.

Tolerances:
  * FP64 J/K/Fock: 1e-10 absolute against the oracle and the PySCF fixtures;
  * FP64 SCF energy: 1e-9 Eh against the fixtures (the target is 1e-8 Eh);
  * FP32 J/K: 1e-5 x max|X| (B rounded once to FP32, FP32 MFMA accumulation);
  * BF16X3 J/K: 2e-4 x max|K| and 1e-5 x max|J| (hi+lo bfloat16 operands,
    about 2^-16 relative per product; J accumulates in FP64);
  * mixed-precision SCF (bf16x3 iterations + FP64 refinement): 1e-6 Eh
    (1 micro-Hartree) against the FP64 SCF.
"""

from __future__ import annotations

import numpy as np
import pytest

from _cases import ccpvdz_case, molecule_case, posthf_case
from vibeqc_amd import FockPlan, _native
from vibeqc_amd.geometry import water_cluster

pytestmark = pytest.mark.gpu


def _plan(case, **kw):
    return FockPlan(case.orbital(), auxiliary=case.auxiliary(), **kw)


@pytest.mark.parametrize("name", ["h2", "lih", "water"])
def test_posthf_fixture_fock_and_energy(oracle, name):
    case, meta, ref = posthf_case(name)
    P = case.oracle_problem(oracle)
    C, occ = ref["df_C"], ref["df_occ"]
    D = (C * occ) @ C.T
    J, K = oracle.jk(P.B, D)
    with _plan(case) as plan:
        ev = plan.evaluate(D)
        assert np.abs(ev.coulomb - J).max() < 1e-10
        assert np.abs(ev.exchange - K).max() < 1e-10
        assert np.abs(ev.fock - ref["df_F"]).max() < 1e-10
        e = ev.energy_one_electron + ev.energy_two_electron + ev.nuclear_repulsion
        assert abs(e - meta["df_hf_energy"]) < 1e-9
        r = plan.solve(energy_tolerance=1e-12, density_tolerance=1e-10)
        assert abs(r.energy - meta["df_hf_energy"]) < 1e-9


def test_orbital_f_shells_against_fixture_and_oracle(oracle):
    """HeH+ with an orbital f shell fitted by itself (PySCF posthf f_heh): J/K
    against the oracle (1e-10), the fixture's Fock matrix (1e-10) and DF-RHF
    energy (1e-9 Eh), and forces against the oracle's analytic gradient, which
    central differences pin (1e-8 Eh/bohr)."""
    case, meta, ref = posthf_case("f_heh")
    assert max(s.l for s in case.shells) == 3
    P = case.oracle_problem(oracle)
    C, occ = ref["df_C"], ref["df_occ"]
    D = (C * occ) @ C.T
    J, K = oracle.jk(P.B, D)
    tight = dict(energy_tolerance=1e-12, density_tolerance=1e-10)
    with _plan(case) as plan:
        ev = plan.evaluate(D)
        r = plan.solve(compute_forces=True, **tight)
    assert np.abs(ev.coulomb - J).max() < 1e-10 and np.abs(ev.exchange - K).max() < 1e-10
    assert np.abs(ev.fock - ref["df_F"]).max() < 1e-10
    assert abs(r.energy - meta["df_hf_energy"]) < 1e-9
    sh, ax = case.oracle_shells(oracle), case.oracle_shells(oracle, True)
    g = oracle.df_rhf_gradient(case.z, case.xyz, sh, case.spherical, ax, case.spherical, P, oracle.df_rhf(P, **tight))
    print(f"f_heh: dE {r.energy - meta['df_hf_energy']:.2e}, max |F + g_oracle| {np.abs(np.asarray(r.forces) + g).max():.2e}")
    assert np.abs(np.asarray(r.forces) + g).max() < 1e-8


def test_orbital_g_shells_report_not_implemented():
    from vibeqc_amd.basis import AOBasis, Primitive, Shell

    case, _, _ = posthf_case("f_heh")
    shells = case.shells + (Shell(0, 4, (Primitive(1.2, 1.0),)),)
    with pytest.raises(_native.NativeError) as err:
        FockPlan(AOBasis(case.atoms, shells, case.representation, case.charge, case.multiplicity),
                 auxiliary=case.auxiliary())
    assert err.value.status == 3


@pytest.mark.parametrize("name", ["ccpvdz_water", "ccpvdz_water4"])
def test_ccpvdz_jkfit_energy(oracle, name):
    case, meta = ccpvdz_case(name)
    with _plan(case) as plan:
        shape = plan.shape()
        assert shape["nbf"] == meta["aos"] and shape["naux"] == meta["auxiliary_count"]
        assert shape["metric_rank"] == meta["stock_effective_rank"]
        r = plan.solve(energy_tolerance=1e-12, density_tolerance=1e-10)
    assert abs(r.energy - meta["energy_hartree"]) < 1e-8
    ref = oracle.df_rhf(case.oracle_problem(oracle), energy_tolerance=1e-12, density_tolerance=1e-10)
    assert abs(r.energy - ref.energy) < 1e-9
    assert np.abs(np.asarray(r.density) - ref.density).max() < 1e-6


@pytest.fixture(scope="module")
def water4(oracle):
    case = molecule_case("water4", water_cluster(4, 6.0, seed=2024))
    P = case.oracle_problem(oracle)
    rng = np.random.default_rng(11)
    C = rng.normal(size=(P.S.shape[0], 20)) * 0.3
    return case, P, C


def test_jk_random_density_fp64(oracle, water4):
    case, P, C = water4
    D = 2 * C @ C.T
    J, K = oracle.jk(P.B, D)
    with _plan(case) as plan:
        ev = plan.evaluate(D)
        j2, k2 = plan.evaluate_occupied(C)
    assert np.abs(ev.coulomb - J).max() < 1e-10 * max(1, np.abs(J).max())
    assert np.abs(ev.exchange - K).max() < 1e-10 * max(1, np.abs(K).max())
    assert np.abs(j2 - J).max() < 1e-10 * max(1, np.abs(J).max())
    assert np.abs(k2 - K).max() < 1e-10 * max(1, np.abs(K).max())


def test_jk_indefinite_density_fp64(oracle, water4):
    case, P, C = water4
    rng = np.random.default_rng(5)
    X = rng.normal(size=P.S.shape) * 0.05
    D = X + X.T  # indefinite: K gets signed factor columns
    J, K = oracle.jk(P.B, D)
    with _plan(case) as plan:
        ev = plan.evaluate(D)
    assert np.abs(ev.coulomb - J).max() < 1e-10 * max(1, np.abs(J).max())
    assert np.abs(ev.exchange - K).max() < 1e-10 * max(1, np.abs(K).max())


@pytest.mark.parametrize("precision,tol_j,tol_k", [("fp32", 1e-5, 1e-5), ("bf16x3", 1e-5, 2e-4)])
def test_jk_reduced_precision(oracle, water4, precision, tol_j, tol_k):
    case, P, C = water4
    D = 2 * C @ C.T
    J, K = oracle.jk(P.B, D)
    with _plan(case, precision=precision) as plan:
        j, k = plan.evaluate_occupied(C)
    ej = np.abs(j - J).max() / np.abs(J).max()
    ek = np.abs(k - K).max() / np.abs(K).max()
    print(f"{precision}: rel J err {ej:.3e}, rel K err {ek:.3e}")
    assert ej < tol_j and ek < tol_k


def test_mixed_precision_scf_within_one_microhartree(water4):
    case, _, _ = water4
    with _plan(case) as plan:
        ref = plan.solve(energy_tolerance=1e-11, density_tolerance=1e-9)
    for precision in ("bf16x3", "fp32"):
        with _plan(case, precision=precision) as plan:
            r = plan.solve(energy_tolerance=1e-11, density_tolerance=1e-9)
        print(f"{precision}: dE = {r.energy - ref.energy:.3e} Eh, builds {r.fock_builds} vs {ref.fock_builds}")
        assert abs(r.energy - ref.energy) < 1e-6


def test_nonsymmetric_density(oracle, water4):
    """A nonsymmetric density (as the reference's raw contractions accept,
    fock_build.hpp:173): J depends on its symmetric part, K = sum_Q B_Q D B_Q
    keeps its orientation — both against the oracle (1e-10 relative), also for
    an unrestricted plan's spin densities.  The derivative still needs a
    symmetric density (INVALID_ARGUMENT, outputs untouched)."""
    from vibeqc_amd import FockBuildSpec

    case, P, C = water4
    n = P.S.shape[0]
    rng = np.random.default_rng(21)
    D = 2 * C @ C.T + 0.05 * rng.normal(size=(n, n))
    J, K = oracle.jk(P.B, D)
    assert np.abs(K - K.T).max() > 1e-3  # the antisymmetric part reaches K
    with _plan(case) as plan:
        ev = plan.evaluate(D)
    print(f"nonsymmetric D: rel dJ {np.abs(ev.coulomb - J).max() / np.abs(J).max():.1e}, "
          f"rel dK {np.abs(ev.exchange - K).max() / np.abs(K).max():.1e}")
    assert np.abs(ev.coulomb - J).max() < 1e-10 * np.abs(J).max()
    assert np.abs(ev.exchange - K).max() < 1e-10 * np.abs(K).max()
    Db = 0.5 * D.T + 0.02 * rng.normal(size=(n, n))
    with FockPlan(case.orbital(), FockBuildSpec.hf("unrestricted"), auxiliary=case.auxiliary()) as plan:
        ev = plan.evaluate(np.stack([0.5 * D, Db]))
    for dev, dens in ((ev.exchange[0], 0.5 * D), (ev.exchange[1], Db)):
        Ks = oracle.jk(P.B, dens)[1]
        assert np.abs(dev - Ks).max() < 1e-10 * np.abs(Ks).max()
    with _plan(case) as plan:  # FockBuildSpec.hf(): derivative order 1
        with pytest.raises(_native.NativeError) as err:
            plan.evaluate(D, derivative=True)
        assert err.value.status == 1


def test_not_converged_status(water4):
    case, _, _ = water4
    with _plan(case) as plan:
        with pytest.raises(_native.NotConvergedError):
            plan.solve(max_iterations=1)
        # The plan stays usable after a failure.
        r = plan.solve()
        assert np.isfinite(r.energy)


def test_cartesian_f_shells_and_open_shell_f(oracle):
    """The f_heh orbital and fitting sets in the Cartesian representation
    (10-component f): J/K of a random density against the oracle (1e-10), the
    closed-shell SCF energy (1e-9 Eh) and forces (1e-8 Eh/bohr) against the
    oracle's; and neutral HeH (doublet, UHF) with the same basis against the
    UHF oracle's energy (1e-9 Eh) and gradient (1e-8 Eh/bohr)."""
    from dataclasses import replace

    from vibeqc_amd import Calculator, FockPlan

    case, _, _ = posthf_case("f_heh")
    cart = replace(case, representation="cartesian")
    P = cart.oracle_problem(oracle)
    n = P.S.shape[0]
    assert n == 1 + 10 + 1
    rng = np.random.default_rng(12)
    C = rng.normal(size=(n, 2)) * 0.4
    D = 2 * C @ C.T
    J, K = oracle.jk(P.B, D)
    tight = dict(energy_tolerance=1e-12, density_tolerance=1e-10)
    with FockPlan(cart.orbital(), auxiliary=cart.auxiliary()) as plan:
        ev = plan.evaluate(D)
        r = plan.solve(compute_forces=True, **tight)
    assert np.abs(ev.coulomb - J).max() < 1e-10 and np.abs(ev.exchange - K).max() < 1e-10
    sh, ax = cart.oracle_shells(oracle), cart.oracle_shells(oracle, True)
    ref = oracle.df_rhf(P, **tight)
    g = oracle.df_rhf_gradient(cart.z, cart.xyz, sh, False, ax, False, P, ref)
    print(f"Cartesian f_heh: dE {r.energy - ref.energy:.2e}, max dF {np.abs(np.asarray(r.forces) + g).max():.2e}")
    assert abs(r.energy - ref.energy) < 1e-9
    assert np.abs(np.asarray(r.forces) + g).max() < 1e-8

    # Neutral HeH doublet (three electrons), UHF with the spherical f basis.
    open_shell = replace(case, charge=0, multiplicity=2)
    Po = open_shell.oracle_problem(oracle)
    na, nb = oracle.spin_occupations(3, 2)
    uref = oracle.df_uhf(Po, na, nb, **tight)
    sho, axo = open_shell.oracle_shells(oracle), open_shell.oracle_shells(oracle, True)
    gu = oracle.df_uhf_gradient(open_shell.z, open_shell.xyz, sho, True, axo, True, Po, uref, (na, nb))
    res = Calculator(method="uhf", basis=case.shells, auxiliary_basis=case.aux_shells,
                     basis_representation="real_spherical", energy_tolerance=1e-12,
                     density_tolerance=1e-10).singlepoint(case.atoms, charge=0, multiplicity=2)
    print(f"HeH doublet UHF: dE {res.energy - uref.energy:.2e}, max dF {np.abs(np.asarray(res.forces) + gu).max():.2e}")
    assert abs(res.energy - uref.energy) < 1e-9
    assert np.abs(np.asarray(res.forces) + gu).max() < 1e-8
