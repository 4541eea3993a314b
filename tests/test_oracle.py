"""Pins the FP64 CPU oracle against the reference's golden vectors and against
what the mathematics fixes (closed forms, invariants, brute force).  No GPU."""

from __future__ import annotations

import math

import numpy as np
import pytest

from _cases import CCPVDZ, POSTHF, ccpvdz_case, posthf_case
from conftest import golden_json


def _known_shells(O, records):
    return [O.Shell(a, l, tuple(p[0] for p in prims), tuple(p[1] for p in prims)) for a, l, prims in records]


def test_boys_closed_form_and_recurrence(oracle):
    # F0(T) = 1/2 sqrt(pi/T) erf(sqrt T); F_m(0) = 1/(2m+1); upward identity.
    for t in (1e-6, 0.3, 2.0, 7.5, 15.9, 16.1, 30.0, 80.0):
        f = oracle.boys(8, t)
        assert abs(f[0] - 0.5 * math.sqrt(math.pi / t) * math.erf(math.sqrt(t))) < 1e-14 * max(1.0, f[0]) + 1e-15
        for m in range(1, 9):
            assert abs(f[m] - ((2 * m - 1) * f[m - 1] - math.exp(-t)) / (2 * t)) < 1e-12 * f[m - 1] + 1e-16 or t < 1e-3
    f0 = oracle.boys(6, 0.0 + 1e-300)
    for m in range(7):
        assert abs(f0[m] - 1.0 / (2 * m + 1)) < 1e-15


def test_libcint_cartesian_known_answers(oracle):
    ka = golden_json("libcint_known_answers.json")["cartesian_h2"]
    z = [a[0] for a in ka["atoms_bohr"]]
    xyz = np.array([a[1] for a in ka["atoms_bohr"]])
    orb = _known_shells(oracle, ka["orbital_shells"])
    aux = _known_shells(oracle, ka["auxiliary_shells"])
    A = oracle.three_center(xyz, orb, False, aux, False)
    M = oracle.metric(xyz, aux, False)
    assert A.shape == (8, 8, 8) and M.shape == (8, 8)
    assert abs(A.sum() - ka["sum_three_center"]) < ka["sum_three_center_tol"]
    assert abs(M.sum() - ka["sum_metric"]) < ka["sum_metric_tol"]
    for (i, j, p), v in ka["three_center"]:
        assert abs(A[i, j, p] - v) < ka["value_tol"]
    for (p, q), v in ka["metric"]:
        assert abs(M[p, q] - v) < ka["value_tol"]


def test_libcint_spherical_d_known_answers(oracle):
    ka = golden_json("libcint_known_answers.json")["spherical_d"]
    xyz = np.array([a[1] for a in ka["atoms_bohr"]])
    orb = _known_shells(oracle, ka["orbital_shells"])
    aux = _known_shells(oracle, ka["auxiliary_shells"])
    A = oracle.three_center(xyz, orb, True, aux, True)
    M = oracle.metric(xyz, aux, True)
    assert abs(M.sum() - ka["sum_metric"]) < ka["sum_tol"]
    assert abs(A.sum() - ka["sum_three_center"]) < ka["sum_tol"]
    for (p, q), v in ka["metric"]:
        assert abs(M[p, q] - v) < ka["value_tol"]
    for (i, j, p), v in ka["three_center"]:
        assert abs(A[i, j, p] - v) < ka["value_tol"]


def test_ss_s_closed_form(oracle):
    # (s_a s_b | s_c) for unit-normalised primitives has a closed form.
    xyz = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 1.1], [1.5, 0.2, -0.7]])
    a, b, g = 1.3, 0.7, 0.45
    orb = [oracle.Shell(0, 0, (a,), (1.0,)), oracle.Shell(1, 0, (b,), (1.0,))]
    aux = [oracle.Shell(2, 0, (g,), (1.0,))]
    A = oracle.three_center(xyz, orb, False, aux, False)
    p = a + b
    P = (a * xyz[0] + b * xyz[1]) / p
    rab = np.sum((xyz[0] - xyz[1]) ** 2)
    T = p * g / (p + g) * np.sum((P - xyz[2]) ** 2)
    F0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
    norm = lambda e: (2 * e / math.pi) ** 0.75
    ref = norm(a) * norm(b) * norm(g) * 2 * math.pi**2.5 / (p * g * math.sqrt(p + g)) * math.exp(-a * b / p * rab) * F0
    assert abs(A[0, 1, 0] - ref) < 1e-14
    assert abs(A[1, 0, 0] - ref) < 1e-14


def test_overlap_diagonal_is_one(oracle):
    case, _ = ccpvdz_case("ccpvdz_water")
    S, _, _ = oracle.one_electron(case.z, case.xyz, case.oracle_shells(oracle), True)
    assert np.allclose(np.diag(S), 1.0, atol=1e-13)


@pytest.mark.parametrize("name", POSTHF)
def test_posthf_fixture_integrals(oracle, name):
    case, meta, ref = posthf_case(name)
    P = case.oracle_problem(oracle)
    assert np.abs(P.A - ref["raw_three_center"]).max() < 1e-12
    assert np.abs(P.M - ref["metric"]).max() < 2e-12
    assert np.abs(P.S - ref["df_S"]).max() < 1e-13
    assert np.abs(P.H - ref["df_h"]).max() < 1e-12


@pytest.mark.parametrize("name", POSTHF)
def test_posthf_fixture_fock_and_energy(oracle, name):
    case, meta, ref = posthf_case(name)
    P = case.oracle_problem(oracle)
    C, occ = ref["df_C"], ref["df_occ"]
    D = (C * occ) @ C.T
    J, K = oracle.jk(P.B, D)
    assert np.abs(oracle.fock(P.H, J, K) - ref["df_F"]).max() < 1e-11
    r = oracle.df_rhf(P, energy_tolerance=1e-12, density_tolerance=1e-10)
    assert r.converged
    assert abs(r.energy - meta["df_hf_energy"]) < 1e-10


def test_metric_inverse_sqrt_identity_and_rank(oracle):
    case, _, ref = posthf_case("lih")
    X, rank, cond = oracle.metric_inverse_sqrt(ref["metric"])
    assert rank == ref["metric"].shape[0]
    assert np.abs(X @ ref["metric"] @ X - np.eye(rank)).max() < 1e-10
    Xd, rank_d, cond_d = oracle.metric_inverse_sqrt(np.array([[2.0, 2.0], [2.0, 2.0]]))
    assert rank_d == 1 and cond_d == 1.0


def test_jk_against_brute_force_and_occupied(oracle):
    case, _, ref = posthf_case("water")
    P = case.oracle_problem(oracle)
    rng = np.random.default_rng(7)
    C = rng.normal(size=(P.S.shape[0], 3))
    D = 2 * C @ C.T
    J, K = oracle.jk(P.B, D)
    Jb, Kb = oracle.brute_force_jk(P.B, D)
    Jo, Ko = oracle.jk_occupied(P.B, math.sqrt(2) * C)
    assert np.abs(J - Jb).max() < 1e-12 and np.abs(K - Kb).max() < 1e-12
    assert np.abs(J - Jo).max() < 1e-12 and np.abs(K - Ko).max() < 1e-12


def test_energy_variation_matches_fock(oracle):
    # dE/dD . Delta = sum F . Delta for E = tr D H + 1/2 tr D (J - K/2) (test_fock.py:42-83).
    case, _, _ = posthf_case("water")
    P = case.oracle_problem(oracle)
    rng = np.random.default_rng(3)
    D = rng.normal(size=P.S.shape)
    D = D + D.T
    delta = rng.normal(size=P.S.shape)
    delta = delta + delta.T

    def energy(Dm):
        J, K = oracle.jk(P.B, Dm)
        return float(np.sum(Dm * P.H) + oracle.two_electron_energy(Dm, J, K))

    h = 1e-5
    fd = (energy(D + h * delta) - energy(D - h * delta)) / (2 * h)
    J, K = oracle.jk(P.B, D)
    assert abs(fd - np.sum(oracle.fock(P.H, J, K) * delta)) < 1e-7 * max(1.0, abs(fd))


@pytest.mark.parametrize("name", CCPVDZ[:2])
def test_ccpvdz_jkfit_energy(oracle, name):
    case, meta = ccpvdz_case(name)
    P = case.oracle_problem(oracle)
    assert P.S.shape[0] == meta["aos"] and P.M.shape[0] == meta["auxiliary_count"]
    assert P.rank == meta["stock_effective_rank"]
    assert abs(P.condition - meta["metric_condition"]) < 1e-6 * meta["metric_condition"]
    r = oracle.df_rhf(P, energy_tolerance=1e-12, density_tolerance=1e-10)
    assert r.converged
    assert abs(r.energy - meta["energy_hartree"]) < 1e-9


def test_planted_mistake_is_caught(oracle):
    # A dropped exchange factor (K instead of K/2) must break the fixture check.
    case, meta, ref = posthf_case("water")
    P = case.oracle_problem(oracle)
    C, occ = ref["df_C"], ref["df_occ"]
    D = (C * occ) @ C.T
    J, K = oracle.jk(P.B, D)
    assert np.abs(P.H + J - K - ref["df_F"]).max() > 1e-3


# ---------------------------------------------------------------------------
# Nuclear gradient (density_fitting.cpp:898-917; hf_gradient.cpp:8-18)
# ---------------------------------------------------------------------------


def _oracle_gradient(O, case, problem=None, **scf_kw):
    sh, ax = case.oracle_shells(O), case.oracle_shells(O, True)
    problem = problem or O.build_problem(case.z, case.xyz, sh, case.spherical, ax, case.spherical, charge=case.charge)
    scf = O.df_rhf(problem, **scf_kw)
    return O.df_rhf_gradient(case.z, case.xyz, sh, case.spherical, ax, case.spherical, problem, scf), problem, scf


@pytest.mark.parametrize("name", ["ccpvdz_water", "ccpvdz_water4"])
def test_df_rhf_gradient_matches_gpu4pyscf_forces(oracle, name):
    """GPU4PySCF 1.8.1 cc-pVDZ/cc-pVDZ-JKFIT forces (issue409 reference JSONs);
    the reference's own forces agree with them to ~2e-9 Eh/bohr."""
    case, meta = ccpvdz_case(name)
    g, _, _ = _oracle_gradient(oracle, case)
    forces = np.array(meta["forces_hartree_per_bohr"])
    assert np.abs(-g - forces).max() < 1e-8
    assert np.abs(g.sum(axis=0)).max() < 1e-10  # translational invariance


@pytest.mark.parametrize("name", ["lih", "f_heh"])
def test_df_rhf_gradient_central_differences(oracle, name):
    """Analytic gradient vs central differences of the converged DF-RHF energy
    (h = 1e-4 bohr); f_heh carries an auxiliary f shell."""
    case, _, _ = posthf_case(name)
    tight = dict(energy_tolerance=1e-13, density_tolerance=1e-11)
    g, _, _ = _oracle_gradient(oracle, case, **tight)
    sh, ax = case.oracle_shells(oracle), case.oracle_shells(oracle, True)

    def energy(xyz):
        p = oracle.build_problem(case.z, xyz, sh, case.spherical, ax, case.spherical, charge=case.charge)
        return oracle.df_rhf(p, **tight).energy

    h = 1e-4
    for a in range(len(case.atoms)):
        for x in range(3):
            xp, xm = case.xyz.copy(), case.xyz.copy()
            xp[a, x] += h
            xm[a, x] -= h
            fd = (energy(xp) - energy(xm)) / (2 * h)
            assert abs(g[a, x] - fd) < 1e-7, (a, x, g[a, x], fd)


def test_gradient_planted_mistake_is_caught(oracle):
    """Dropping the metric response W_M must fail the fixture check."""
    case, meta = ccpvdz_case("ccpvdz_water")
    g, problem, scf = _oracle_gradient(oracle, case)
    sh, ax = case.oracle_shells(oracle), case.oracle_shells(oracle, True)
    Co = scf.coefficients[:, : problem.nocc]
    _, WM = oracle.df_weights(problem.A, problem.M, 2.0 * Co @ Co.T, Co)
    broken = g - oracle.grad_metric(case.xyz, len(case.atoms), ax, case.spherical, WM)
    forces = np.array(meta["forces_hartree_per_bohr"])
    assert np.abs(-broken - forces).max() > 1e-4


def test_dense_weights_equal_occupied_weights(oracle):
    """For D = 2 C C^T the factor-free weights equal the occupied-factor ones."""
    case, _ = ccpvdz_case("ccpvdz_water")
    _, problem, scf = _oracle_gradient(oracle, case)
    Co = scf.coefficients[:, : problem.nocc]
    D = 2.0 * Co @ Co.T
    wa1, wm1 = oracle.df_weights(problem.A, problem.M, D, Co)
    wa2, wm2 = oracle.df_weights_dense(problem.A, problem.M, D)
    # Different contraction orders through M^-1 agree to roundoff.
    assert np.abs(wa1 - wa2).max() < 1e-10 * np.abs(wa1).max()
    assert np.abs(wm1 - wm2).max() < 1e-10 * np.abs(wm1).max()


def test_fixed_density_two_electron_gradient_central_differences(oracle):
    """dE2/dR at a fixed indefinite AO density (the basis moves with the atoms)
    against central differences of 1/2 sum D (J - K/2)."""
    case, _, _ = posthf_case("lih")
    sh, ax = case.oracle_shells(oracle), case.oracle_shells(oracle, True)
    n = sum(2 * s.l + 1 if case.spherical else (s.l + 1) * (s.l + 2) // 2 for s in sh)
    rng = np.random.default_rng(11)
    D = rng.normal(size=(n, n))
    D = D + D.T

    def e2(xyz):
        A = oracle.three_center(xyz, sh, case.spherical, ax, case.spherical)
        M = oracle.metric(xyz, ax, case.spherical)
        X, _, _ = oracle.metric_inverse_sqrt(M)
        J, K = oracle.jk(oracle.df_tensor(A, X), D)
        return oracle.two_electron_energy(D, J, K), A, M

    _, A, M = e2(case.xyz)
    g = oracle.df_two_electron_gradient(case.xyz, len(case.atoms), sh, case.spherical, ax, case.spherical, A, M, D)
    h = 1e-4
    for a in range(len(case.atoms)):
        for x in range(3):
            xp, xm = case.xyz.copy(), case.xyz.copy()
            xp[a, x] += h
            xm[a, x] -= h
            fd = (e2(xp)[0] - e2(xm)[0]) / (2 * h)
            assert abs(g[a, x] - fd) < 1e-7 * max(1.0, abs(fd)), (a, x, g[a, x], fd)
