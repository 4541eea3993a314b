"""Ragged batches on the HIP device: every item equals the single-system
plan's result (same kernels; energies to 1e-9 Eh, forces to 1e-8 Eh/bohr),
items carry their own status, warm starts reuse densities, and the shape
buckets group equal shapes."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _clusters():
    from vibeqc_amd.geometry import water_cluster

    sizes = [1, 3, 2, 1, 3]
    return [water_cluster(n, 3.0 * n ** (1.0 / 3.0), seed=100 + k) for k, n in enumerate(sizes)], sizes


def test_batch_matches_single_plans():
    from vibeqc_amd import Calculator

    systems, sizes = _clusters()
    calc = Calculator(method="rhf", basis="cc-pvdz", auxiliary_basis="cc-pvdz-jkfit")
    batch = calc.batch_singlepoint(systems)
    assert batch.succeeded
    buckets = {}
    for item, n, atoms in zip(batch.items, sizes, systems):
        single = calc.singlepoint(atoms)
        print(f"item {item.index}: n={n} E={item.energy:.10f} dE={item.energy - single.energy:.2e} it={item.iterations}")
        assert abs(item.energy - single.energy) < 1e-9
        assert np.abs(np.asarray(item.forces) - np.asarray(single.forces)).max() < 1e-8
        buckets.setdefault(n, set()).add(item.bucket_id)
    assert all(len(v) == 1 for v in buckets.values())
    assert len({next(iter(v)) for v in buckets.values()}) == len(buckets)


def test_batch_mixed_precision_within_one_microhartree():
    from vibeqc_amd import Calculator

    systems, _ = _clusters()
    e64 = Calculator().batch_singlepoint(systems, properties=("energy",)).energies
    mixed = Calculator(precision="mixed").batch_singlepoint(systems, properties=("energy",))
    assert mixed.succeeded
    assert all(it.forces is None for it in mixed.items)
    assert np.abs(mixed.energies - e64).max() < 1e-6


def test_prepared_batch_warm_start_and_new_coordinates():
    from vibeqc_amd import Calculator

    systems, _ = _clusters()
    calc = Calculator()
    with calc.prepare_batch(systems, warm_start=True) as batch:
        first = batch.execute(properties=("energy",))
        again = batch.execute(properties=("energy",))
        assert all(it.warm_start_used for it in again.items)
        assert np.abs(again.energies - first.energies).max() < 1e-9
        assert max(it.iterations for it in again.items) < max(it.iterations for it in first.items)
        moved = [np.array([a.position for a in s]) + 0.01 for s in systems]
        shifted = batch.execute(moved, properties=("energy",))
    for a, b, c in zip(first.items, again.items, shifted.items):
        print(f"item {a.index}: first {a.energy:.12f} it={a.iterations} again {b.energy:.12f} it={b.iterations} "
              f"shifted {c.energy:.12f} it={c.iterations} warm={c.warm_start_used} fb={c.warm_start_fallback}")
    # A rigid translation leaves the energy unchanged.
    assert np.abs(shifted.energies - first.energies).max() < 1e-8


def test_batch_item_failures_are_isolated():
    """An open-shell member fails with NOT_IMPLEMENTED (status 3), an
    unconverged run with NOT_CONVERGED (status 4); the other members succeed."""
    from vibeqc_amd import Calculator

    systems, _ = _clusters()
    calc = Calculator()
    mixed = calc.batch_singlepoint(systems[:3], charges=[0, 1, 0], multiplicities=[1, 2, 1], properties=("energy",))
    assert mixed.failure_indices == (1,)
    assert mixed.items[1].status == 3
    assert mixed.items[0].succeeded and mixed.items[2].succeeded
    with pytest.raises(RuntimeError):
        calc.batch_singlepoint(systems[:3], charges=[0, 1, 0], multiplicities=[1, 2, 1], strict=True)
    short = Calculator(max_iterations=2).batch_singlepoint(systems[:2], properties=("energy",))
    assert all(it.status == 4 for it in short.items)


def test_torch_batched_energy_backward():
    import torch

    from vibeqc_amd import Calculator
    from vibeqc_amd.torch import batched_energy

    systems, _ = _clusters()
    calc = Calculator()
    coords = [torch.tensor([a.position for a in s], dtype=torch.float64, requires_grad=True) for s in systems[:3]]
    zs = [[a.atomic_number for a in s] for s in systems[:3]]
    e = batched_energy(coords, zs, calc)
    e.sum().backward()
    for c, atoms in zip(coords, systems[:3]):
        ref = calc.singlepoint(atoms)
        assert np.abs(-c.grad.numpy() - np.asarray(ref.forces)).max() < 1e-8


def test_batch_is_reproducible(monkeypatch):
    """Energies and forces of repeated executions are bitwise identical, with
    one worker and with concurrent workers (their SCFs overlap; setups and
    forces run one system at a time)."""
    from vibeqc_amd import Calculator
    from vibeqc_amd.geometry import water_cluster

    rng = np.random.default_rng(9)
    sizes = rng.integers(1, 6, size=48)
    systems = [water_cluster(int(n), 3.0 * n ** (1.0 / 3.0), seed=300 + k) for k, n in enumerate(sizes)]

    def run():
        res = batch.execute()
        return np.asarray(res.energies), np.concatenate([np.asarray(it.forces).ravel() for it in res.items])

    with Calculator().prepare_batch(systems, warm_start=False) as batch:
        monkeypatch.setenv("VIBEQC_BATCH_WORKERS", "1")
        serial = [run() for _ in range(2)]
        monkeypatch.delenv("VIBEQC_BATCH_WORKERS")
        concurrent = [run() for _ in range(3)]
    for e, f in serial[1:] + concurrent:
        assert np.array_equal(e, serial[0][0]) and np.array_equal(f, serial[0][1])
