"""Checkpoint restart through the batch's warm-state ABI on the HIP device:
a fresh batch restored from a file restarts from the saved densities
(restart_origin "persistent_restart", same energies, few iterations); changed
geometry needs allow_warm; another basis is rejected without touching the
neighbours' seeds; clearing forgets every seed."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _systems():
    from vibeqc_amd.geometry import water_cluster

    return [water_cluster(n, 3.0 * n ** (1.0 / 3.0), seed=40 + n) for n in (1, 2, 3)]


def test_save_and_restore_in_a_fresh_batch(tmp_path):
    from vibeqc_amd import Calculator
    from vibeqc_amd.checkpoint import CheckpointError, inspect_checkpoint

    systems = _systems()
    calc = Calculator()
    path = tmp_path / "hf.vqcp"
    with calc.prepare_batch(systems) as batch:
        cold = batch.execute(properties=("energy",), strict=True)
        report = batch.save_checkpoint(path)
    assert report["saved_indices"] == [0, 1, 2]
    assert all(it.restart_origin == "cold" for it in cold.items)
    doc = inspect_checkpoint(path)
    assert [it["status"] for it in doc["items"]] == [0, 0, 0]

    with calc.prepare_batch(systems) as batch:
        rep = batch.load_checkpoint(path)
        assert [r["compatibility"] for r in rep["items"]] == ["exact_restart"] * 3
        warm = batch.execute(properties=("energy",), strict=True)
        again = batch.execute(properties=("energy",), strict=True)
    for c, w in zip(cold.items, warm.items):
        print(f"item {c.index}: cold it={c.iterations} restored it={w.iterations} dE={w.energy - c.energy:.2e}")
        assert w.restart_origin == "persistent_restart" and w.warm_start_used
        assert abs(w.energy - c.energy) < 1e-9
        assert w.iterations <= 3
    assert all(it.restart_origin == "in_process_warm" for it in again.items)

    moved = [[type(a)(a.atomic_number, tuple(x + 0.05 for x in a.position)) for a in s] for s in systems]
    with calc.prepare_batch(moved) as batch:
        with pytest.raises(CheckpointError, match="geometry"):
            batch.load_checkpoint(path)
        rep = batch.load_checkpoint(path, allow_warm=True)
        assert [r["compatibility"] for r in rep["items"]] == ["warm_start_compatible"] * 3
        shifted = batch.execute(properties=("energy",), strict=True)
    with Calculator().prepare_batch(moved, warm_start=False) as batch:
        reference = batch.execute(properties=("energy",), strict=True)
    assert all(it.restart_origin == "persistent_restart" for it in shifted.items)
    assert np.abs(shifted.energies - reference.energies).max() < 1e-8


def test_incompatible_basis_and_clear(tmp_path):
    from vibeqc_amd import Calculator
    from vibeqc_amd.checkpoint import CheckpointError

    systems = _systems()[:2]
    path = tmp_path / "hf.vqcp"
    with Calculator().prepare_batch(systems) as batch:
        batch.execute(properties=("energy",), strict=True)
        batch.save_checkpoint(path)
    other = Calculator(basis="def2-svp", auxiliary_basis="cc-pvdz-jkfit")
    with other.prepare_batch(systems) as batch:
        with pytest.raises(CheckpointError, match="transport"):
            batch.load_checkpoint(path)
        rep = batch.load_checkpoint(path, strict=False)
        assert [r["compatibility"] for r in rep["items"]] == ["transport_required"] * 2
        assert all(not r["restored_fields"] for r in rep["items"])
        result = batch.execute(properties=("energy",), strict=True)
        assert all(it.restart_origin == "cold" for it in result.items)
    with Calculator().prepare_batch(systems) as batch:
        batch.load_checkpoint(path)
        batch.clear_warm_starts()
        result = batch.execute(properties=("energy",), strict=True)
        assert all(it.restart_origin == "cold" for it in result.items)


def test_empty_and_refused_restores(tmp_path):
    """A batch that has not executed saves slots without seeds (status -1);
    restoring into a batch without warm starts is refused before any change."""
    from vibeqc_amd import Calculator
    from vibeqc_amd.checkpoint import CheckpointError, inspect_checkpoint

    systems = _systems()[:2]
    empty = tmp_path / "empty.vqcp"
    full = tmp_path / "full.vqcp"
    calc = Calculator()
    with calc.prepare_batch(systems) as batch:
        report = batch.save_checkpoint(empty)
        assert report["saved_indices"] == []
        batch.execute(properties=("energy",), strict=True)
        batch.save_checkpoint(full)
    doc = inspect_checkpoint(empty)
    assert all(it["seed"] is None and it["status"] == -1 for it in doc["items"]) and doc["blobs"] == []
    with calc.prepare_batch(systems) as batch:
        rep = batch.load_checkpoint(empty, strict=False)
        assert all(r["compatibility"] == "incompatible" for r in rep["items"])
        with pytest.raises(CheckpointError, match="missing reusable state"):
            batch.load_checkpoint(empty)
    with calc.prepare_batch(systems, warm_start=False) as batch:
        with pytest.raises(CheckpointError, match="native"):
            batch.load_checkpoint(full)
        result = batch.execute(properties=("energy",), strict=True)
        assert all(it.restart_origin == "cold" for it in result.items)
