"""Device-memory budget of a prepared plan (device_budget_bytes; the
reference's budgeted DF planner, src/scf/density_fitting.cpp:1041-1218).

cc-pVDZ / cc-pVDZ-JKFIT (H2O)4 (96 AO, 464 auxiliary functions) holds about
52 MB resident.  Under budgets of 16, 24 and 48 MB the plan keeps only the
metric factor and regenerates B in auxiliary blocks (several blocks and
three-centre row tiles under 16 MB) for every Fock build.  Streamed J/K (positive and
indefinite densities, occupied factors) agree with the resident plan's to
1e-10 relative (the block path maps the three-centre rows with an explicit
L^-1 where the resident path solves with L), the streamed SCF energy to
1e-10 Eh, and the plan stays inside its budget.  The eigen-cutoff metric
(the auxiliary set listed twice, 928 functions of rank 464) streams with the
symmetric X under 40 MB.  An 8 MB budget
cannot hold the metric and one block: OUT_OF_MEMORY.  Forces need the
resident tensor: NOT_IMPLEMENTED.
"""

from __future__ import annotations

import numpy as np
import pytest

from _cases import ccpvdz_case

pytestmark = pytest.mark.gpu

MB = 1 << 20
TIGHT = dict(energy_tolerance=1e-11, density_tolerance=1e-9)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _densities(n, seed=7):
    rng = np.random.default_rng(seed)
    C = rng.normal(size=(n, 20)) * 0.3
    E = rng.normal(size=(n, 6)) * 0.2
    return C, 2 * C @ C.T, 2 * C @ C.T - E @ E.T


@pytest.fixture(scope="module")
def resident():
    from vibeqc_amd import FockPlan

    case, _ = ccpvdz_case("ccpvdz_water4")
    with FockPlan(case.orbital(), auxiliary=case.auxiliary()) as plan:
        assert not plan.streaming()["streamed"]
        C, D, Dind = _densities(plan.nao)
        out = dict(ev=plan.evaluate(D), ind=plan.evaluate(Dind), occ=plan.evaluate_occupied(C),
                   scf=plan.solve(**TIGHT))
    return case, C, D, Dind, out


@pytest.mark.parametrize("budget_mb", [16, 24, 48])
def test_streamed_plan_matches_resident(resident, budget_mb):
    from vibeqc_amd import FockPlan, _native

    case, C, D, Dind, ref = resident
    with FockPlan(case.orbital(), auxiliary=case.auxiliary(), device_budget_bytes=budget_mb * MB) as plan:
        info = plan.streaming()
        assert info["streamed"] and info["auxiliary_block"] % 128 == 0 and info["row_tiles"] >= 1
        if budget_mb == 16:
            assert info["auxiliary_block"] < plan.shape()["naux"] and info["row_tiles"] > 1
        ev, ind, occ = plan.evaluate(D), plan.evaluate(Dind), plan.evaluate_occupied(C)
        scf = plan.solve(**TIGHT)
        held = plan.shape()["device_bytes"]
        with pytest.raises(_native.NativeError) as err:
            plan.solve(compute_forces=True, **TIGHT)
        assert err.value.status == 3
    errs = [_rel(ev.coulomb, ref["ev"].coulomb), _rel(ev.exchange, ref["ev"].exchange),
            _rel(ind.coulomb, ref["ind"].coulomb), _rel(ind.exchange, ref["ind"].exchange),
            _rel(occ[0], ref["occ"][0]), _rel(occ[1], ref["occ"][1])]
    print(f"budget {budget_mb} MB: {info}, held {held / MB:.1f} MB, max rel J/K {max(errs):.1e}, "
          f"dE {scf.energy - ref['scf'].energy:.1e} ({scf.iterations} vs {ref['scf'].iterations} it)")
    assert max(errs) < 1e-10
    assert abs(scf.energy - ref["scf"].energy) < 1e-10
    assert held <= budget_mb * MB


def test_streamed_eigen_cutoff_metric():
    from vibeqc_amd import FockPlan
    from vibeqc_amd.basis import AOBasis

    case, _ = ccpvdz_case("ccpvdz_water4")
    aux2 = AOBasis(case.atoms, tuple(s for s in case.aux_shells for _ in range(2)), case.representation)
    _, D, Dind = _densities(96, seed=5)
    with FockPlan(case.orbital(), auxiliary=aux2) as plan:
        ref = plan.evaluate(Dind)
    with FockPlan(case.orbital(), auxiliary=aux2, device_budget_bytes=40 * MB) as plan:
        info, shape = plan.streaming(), plan.shape()
        ev = plan.evaluate(Dind)
    print(f"duplicated aux streamed: {info}, rank {shape['metric_rank']}")
    assert info["streamed"] and info["auxiliary_block"] < shape["naux"] and not shape["cholesky"]
    assert _rel(ev.coulomb, ref.coulomb) < 1e-10 and _rel(ev.exchange, ref.exchange) < 1e-10


def test_budget_below_one_block_is_out_of_memory():
    from vibeqc_amd import FockPlan, _native

    case, _ = ccpvdz_case("ccpvdz_water4")
    with pytest.raises(_native.NativeError) as err:
        FockPlan(case.orbital(), auxiliary=case.auxiliary(), device_budget_bytes=8 * MB)
    assert err.value.status == 7


def test_calculator_energy_under_budget(resident):
    from vibeqc_amd import Calculator

    case, _, _, _, ref = resident
    r = Calculator(density_fitting_memory_budget_bytes=24 * MB, energy_tolerance=1e-11,
                   density_tolerance=1e-9).singlepoint(case.atoms, properties=("energy",))
    assert abs(r.energy - ref["scf"].energy) < 1e-10


def test_streamed_unrestricted_plan():
    """(H2O)4+ doublet: the UHF SCF of a plan streamed under 24 MB equals the
    resident plan's (1e-10 Eh), and so do its unrestricted J/K (1e-10
    relative)."""
    from dataclasses import replace

    from vibeqc_amd import FockBuildSpec, FockPlan

    case, _ = ccpvdz_case("ccpvdz_water4")
    cation = replace(case, charge=1, multiplicity=2)
    spec = FockBuildSpec.hf("unrestricted")
    rng = np.random.default_rng(2)
    Ca, Cb = rng.normal(size=(96, 20)) * 0.3, rng.normal(size=(96, 19)) * 0.3
    D = np.stack([Ca @ Ca.T, Cb @ Cb.T])
    out = {}
    for budget in (0, 24 * MB):
        with FockPlan(cation.orbital(), spec, auxiliary=cation.auxiliary(), device_budget_bytes=budget) as plan:
            assert plan.streaming()["streamed"] == (budget > 0)
            ev = plan.evaluate(D)
            r = plan.solve(**TIGHT)
        out[budget] = (ev, r.energy)
    (e0, E0), (e1, E1) = out[0], out[24 * MB]
    print(f"(H2O)4+ UHF streamed - resident: {E1 - E0:.1e} Eh")
    assert abs(E1 - E0) < 1e-10
    for a, b in ((e1.coulomb, e0.coulomb), (e1.exchange[0], e0.exchange[0]), (e1.exchange[1], e0.exchange[1])):
        assert _rel(a, b) < 1e-10
