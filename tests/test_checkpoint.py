"""Checkpoint files (python/generativeqc/checkpoint.py format) without a GPU:
the container round-trips, every corruption the reference's readers reject is
rejected before anything is loaded (tests/python/test_checkpoint.py and
test_endpoint_checkpoint_read_failures.py cover the same cases there), and
the geometry identity is the reference's formula."""

from __future__ import annotations

import hashlib
import json

import numpy as np
import pytest

from vibeqc_amd import Calculator
from vibeqc_amd.basis import Atom, canonical_hash
from vibeqc_amd.checkpoint import (
    HEADER,
    MAGIC,
    CheckpointError,
    build_checkpoint,
    geometry_hash,
    inspect_checkpoint,
    parse_manifest,
    write_checkpoint,
)

H2 = (Atom(1, (0.0, 0.0, -0.7)), Atom(1, (0.0, 0.0, 0.7)))
WATER = (Atom(8, (0.0, 0.0, 0.0)), Atom(1, (0.0, 1.43, 1.11)), Atom(1, (0.0, -1.43, 1.11)))


def _seed(n, natom, energy=-1.0):
    rng = np.random.default_rng(n)
    c = rng.standard_normal((n, 2))
    return {
        "density": c @ c.T,
        "coordinates": rng.standard_normal((natom, 3)),
        "energy": energy,
        "energy_change": 1e-12,
        "density_rms": 1e-9,
        "iterations": 7,
    }


def _file(tmp_path, seeds=(True, False)):
    calc = Calculator()
    systems = [H2, WATER]
    n = [10, 24]
    data = build_checkpoint(
        calc, systems, [0, 0], [1, 1], [0, -1], [_seed(n[i], len(systems[i])) if s else None for i, s in enumerate(seeds)]
    )
    path = tmp_path / "hf.vqcp"
    write_checkpoint(path, data)
    return path, data


def _split(data):
    magic, length, digest = HEADER.unpack(data[: HEADER.size])
    raw = data[HEADER.size : HEADER.size + length]
    return json.loads(raw), data[HEADER.size + length :]


def _rewrite(path, doc, payload, *, fix_header=True):
    raw = json.dumps(doc, sort_keys=True, separators=(",", ":")).encode()
    digest = hashlib.sha256(raw).digest() if fix_header else b"\0" * 32
    path.write_bytes(HEADER.pack(MAGIC, len(raw), digest) + raw + payload)


def test_round_trip_and_layout(tmp_path):
    path, data = _file(tmp_path)
    doc = inspect_checkpoint(path)
    assert doc["schema"] == "generativeqc.hf_checkpoint" and doc["schema_version"] == 1
    assert [it["index"] for it in doc["items"]] == [0, 1]
    assert doc["items"][1]["seed"] is None and doc["items"][1]["model"] is None
    seed = doc["items"][0]["seed"]
    assert seed["nbf"] == 10 and seed["converged"] is True
    assert [b["name"] for b in doc["blobs"]] == ["density_0", "coordinates_0"]
    assert doc["blobs"][0]["shape"] == [1, 10, 10] and doc["blobs"][1]["shape"] == [2, 3]
    _, payload = _split(data)
    dens = np.frombuffer(payload[: 800], dtype="<f8").reshape(10, 10)
    assert np.array_equal(dens, _seed(10, 2)["density"])
    model = doc["items"][0]["model"]
    assert model["identity"] == canonical_hash({k: v for k, v in model.items() if k != "identity"})


def test_geometry_identity_is_the_reference_formula():
    xyz = [(0.0, -0.0, 1.5), (0.25, 2.0, -3.0)]
    expected = hashlib.sha256(
        json.dumps([[8, ["0x0.0p+0", "0x0.0p+0", (1.5).hex()]], [1, [(0.25).hex(), (2.0).hex(), (-3.0).hex()]]],
                   sort_keys=True, separators=(",", ":")).encode()
    ).hexdigest()
    assert geometry_hash([8, 1], xyz) == expected


@pytest.mark.parametrize(
    "mutate,message",
    [
        (lambda d: d.update(schema="other"), "schema"),
        (lambda d: d.update(array_format="float32"), "array format"),
        (lambda d: d["items"][0].update(index=1), "ordering"),
        (lambda d: d["items"][0].update(extra=1), "fields"),
        (lambda d: d["items"][0]["seed"].update(converged=False), "converged"),
        (lambda d: d["items"][0]["model"].update(charge=1), "identity checksum"),
        (lambda d: d["items"][0].update(provider="other"), "provider"),
        (lambda d: d["blobs"][0].update(offset=8), "offset"),
        (lambda d: d["blobs"][1].update(shape=[3, 2]), "dimensions"),
        (lambda d: d["items"][1].update(provider="x"), "no-state"),
    ],
)
def test_semantic_corruption_is_rejected(tmp_path, mutate, message):
    path, data = _file(tmp_path)
    doc, payload = _split(data)
    mutate(doc)
    _rewrite(path, doc, payload)
    with pytest.raises(CheckpointError, match=message):
        inspect_checkpoint(path)


def test_byte_corruption_is_rejected(tmp_path):
    path, data = _file(tmp_path)
    doc, payload = _split(data)
    bad = bytearray(data)
    bad[-1] ^= 1  # blob byte
    path.write_bytes(bytes(bad))
    with pytest.raises(CheckpointError, match="blob checksum"):
        inspect_checkpoint(path)
    _rewrite(path, doc, payload, fix_header=False)
    with pytest.raises(CheckpointError, match="manifest checksum"):
        inspect_checkpoint(path)
    path.write_bytes(data[:-8])
    with pytest.raises(CheckpointError, match="truncated|missing"):
        inspect_checkpoint(path)
    path.write_bytes(data + b"\0" * 8)
    with pytest.raises(CheckpointError, match="trailing"):
        inspect_checkpoint(path)
    path.write_bytes(b"NOTACKPT" + data[8:])
    with pytest.raises(CheckpointError, match="header"):
        inspect_checkpoint(path)
    with pytest.raises(MemoryError):
        inspect_checkpoint(tmp_path / "hf.vqcp", max_bytes=HEADER.size + 10)


def test_duplicate_keys_and_nonfinite_json_are_rejected(tmp_path):
    path, data = _file(tmp_path)
    doc, payload = _split(data)
    raw = json.dumps(doc, sort_keys=True, separators=(",", ":"))
    dup = raw.replace('"schema_version":1', '"schema_version":1,"schema_version":1', 1).encode()
    path.write_bytes(HEADER.pack(MAGIC, len(dup), hashlib.sha256(dup).digest()) + dup + payload)
    with pytest.raises(CheckpointError, match="duplicate"):
        inspect_checkpoint(path)
    nan = raw.replace('"energy_change":1e-12', '"energy_change":NaN', 1).encode()
    with pytest.raises(CheckpointError, match="nonfinite"):
        parse_manifest(nan, len(payload), 1 << 20)


def test_nonfinite_seed_is_refused_on_save(tmp_path):
    seed = _seed(10, 2)
    seed["density"][0, 0] = np.inf
    with pytest.raises(CheckpointError):
        build_checkpoint(Calculator(), [H2], [0], [1], [0], [seed])
