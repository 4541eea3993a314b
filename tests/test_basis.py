"""Canonical basis records (the reference's checked interchange format,
python/generativeqc/basis.py:283-373), pinned by the reference's own
checksummed cc-pVDZ / cc-pVDZ-JKFIT records."""

from __future__ import annotations

import json
from pathlib import Path

import pytest

from vibeqc_amd.basis import Atom, canonical_hash, load_basis, load_canonical_basis, shells_for, write_canonical_basis

GOLDEN = Path(__file__).resolve().parent / "golden"
WATER = (Atom(8, (0.0, 0.0, 0.0)), Atom(1, (0.0, 1.43, 1.11)), Atom(1, (0.0, -1.43, 1.11)))


@pytest.mark.parametrize("name", ["cc-pvdz", "cc-pvdz-jkfit"])
def test_reference_checksums_reproduce(name):
    payload = json.loads((GOLDEN / f"canonical_{name}.json").read_text())
    expected = payload.pop("checksum")
    assert canonical_hash(payload) == expected


@pytest.mark.parametrize("name", ["cc-pvdz", "cc-pvdz-jkfit"])
def test_canonical_records_expand_like_builtin_data(name):
    record = load_canonical_basis(GOLDEN / f"canonical_{name}.json")
    assert record.representation == "spherical"
    assert record.shells_for(WATER) == shells_for(name, WATER)
    assert shells_for(record, WATER) == shells_for(name, WATER)


def test_tampering_is_rejected(tmp_path):
    payload = json.loads((GOLDEN / "canonical_cc-pvdz.json").read_text())
    bad = json.loads(json.dumps(payload))
    bad["elements"][0]["shells"][0]["coefficients"][0][0] = "2E0"
    (tmp_path / "coef.json").write_text(json.dumps(bad))
    with pytest.raises(ValueError, match="checksum"):
        load_canonical_basis(tmp_path / "coef.json")
    extra = json.loads(json.dumps(payload))
    extra["comment"] = "x"
    extra.pop("checksum")
    extra["checksum"] = canonical_hash(extra)
    (tmp_path / "extra.json").write_text(json.dumps(extra))
    with pytest.raises(ValueError, match="unknown fields"):
        load_canonical_basis(tmp_path / "extra.json")
    (tmp_path / "dup.json").write_text('{"schema": "generativeqc.basis", "schema": "generativeqc.basis"}')
    with pytest.raises(ValueError, match="duplicate"):
        load_canonical_basis(tmp_path / "dup.json")


def test_write_and_read_round_trip(tmp_path):
    path = tmp_path / "sto3g.json"
    checksum = write_canonical_basis(path, "sto-3g", load_basis("sto-3g"), representation="cartesian")
    record = load_canonical_basis(path)
    assert record.checksum == checksum and record.representation == "cartesian"
    assert record.shells_for(WATER) == shells_for("sto-3g", WATER)


def test_calculator_accepts_canonical_records():
    from vibeqc_amd import Calculator

    orb = load_canonical_basis(GOLDEN / "canonical_cc-pvdz.json")
    aux = load_canonical_basis(GOLDEN / "canonical_cc-pvdz-jkfit.json")
    by_record = Calculator(basis=orb, auxiliary_basis=aux)
    by_name = Calculator(basis="cc-pvdz", auxiliary_basis="cc-pvdz-jkfit")
    assert by_record.representation == by_name.representation == "real_spherical"
    (o1, a1), (o2, a2) = by_record._bases(WATER, 0, 1), by_name._bases(WATER, 0, 1)
    assert o1.nao == o2.nao == 24 and a1.nao == a2.nao and o1.shells == o2.shells and a1.shells == a2.shells
