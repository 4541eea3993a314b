"""Metric factorisation routes on the HIP device against the FP64 oracle.

  * Rank-deficient metric: the water cc-pVDZ-JKFIT auxiliary set with every
    shell listed twice has a metric of exact rank naux / 2 (the duplicated
    functions span the same space, the situation of the reference's rank-one
    {2, 2, 2, 2} metric, tests/native/test_density_fitting.cpp:574-597).  The
    Cholesky pivots vanish, so the plan takes the eigen-cutoff route
    (density_fitting_metric.cpp:101-152): rank naux / 2 as in the oracle, J/K
    equal to the oracle's on the same duplicated set (1e-9 relative) and to
    the plan without duplicates (the fitted spaces are identical, 1e-8), and
    the same SCF energy.
  * NULL auxiliary: the orbital basis fits itself (fock_prepared.cpp:170);
    the reference's self-fitted PySCF DF-RHF energies (posthf water, LiH,
    H2; 1e-9 Eh) and Fock matrices (1e-10).
"""

from __future__ import annotations

import numpy as np
import pytest

from _cases import ccpvdz_case, posthf_case

pytestmark = pytest.mark.gpu


def test_rank_deficient_metric_takes_the_eigen_cutoff_route(oracle):
    from vibeqc_amd import FockPlan
    from vibeqc_amd.basis import AOBasis

    case, _ = ccpvdz_case("ccpvdz_water")
    doubled = tuple(s for s in case.aux_shells for _ in range(2))
    aux2 = AOBasis(case.atoms, doubled, case.representation)
    P = oracle.build_problem(case.z, case.xyz, case.oracle_shells(oracle), True,
                             [oracle.Shell(s.atom_index, s.l, tuple(p.exponent for p in s.primitives),
                                           tuple(p.coefficient for p in s.primitives)) for s in doubled], True)
    rng = np.random.default_rng(3)
    C = rng.normal(size=(P.S.shape[0], 5)) * 0.3
    D = 2 * C @ C.T
    J, K = oracle.jk(P.B, D)
    with FockPlan(case.orbital(), auxiliary=aux2) as plan:
        shape = plan.shape()
        ev = plan.evaluate(D)
        e2 = plan.solve().energy
    with FockPlan(case.orbital(), auxiliary=case.auxiliary()) as plan:
        ref = plan.evaluate(D)
        e1 = plan.solve().energy
    print(f"duplicated aux: naux {shape['naux']}, rank {shape['metric_rank']} (oracle {P.rank}), "
          f"cholesky {shape['cholesky']}, dE vs single aux {e2 - e1:.2e}")
    assert shape["naux"] == 2 * 116 and shape["metric_rank"] == P.rank == 116 and not shape["cholesky"]
    for dev, orc in ((ev.coulomb, J), (ev.exchange, K)):
        assert np.abs(dev - orc).max() < 1e-9 * np.abs(orc).max()
    for dev, single in ((ev.coulomb, ref.coulomb), (ev.exchange, ref.exchange)):
        assert np.abs(dev - single).max() < 1e-8 * np.abs(single).max()
    assert abs(e2 - e1) < 1e-8


@pytest.mark.parametrize("name", ["water", "lih", "h2"])
def test_null_auxiliary_fits_with_the_orbital_basis(name):
    from vibeqc_amd import FockPlan

    case, meta, ref = posthf_case(name)
    C, occ = ref["df_C"], ref["df_occ"]
    D = (C * occ) @ C.T
    with FockPlan(case.orbital()) as plan:  # no auxiliary basis: the orbital basis fits itself
        assert plan.shape()["naux"] == plan.shape()["nbf"]
        ev = plan.evaluate(D)
        r = plan.solve(energy_tolerance=1e-12, density_tolerance=1e-10)
    assert np.abs(ev.fock - ref["df_F"]).max() < 1e-10
    assert abs(r.energy - meta["df_hf_energy"]) < 1e-9
