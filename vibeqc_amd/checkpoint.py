"""Portable HF checkpoints of a prepared batch's retained densities
(the reference's file format, python/generativeqc/checkpoint.py:1-8,
docs/user/checkpoint.md).

A file is a fixed header ``<8sQ32s>`` (magic ``VQHFCP01``, manifest length,
SHA-256 of the manifest), a UTF-8 JSON manifest (sorted keys, compact
separators, no NaN/Inf) and the numeric blobs: contiguous little-endian
binary64 arrays in C order, each with its own SHA-256 in the manifest.  The
manifest (schema ``generativeqc.hf_checkpoint``, version 1) lists one item per
batch slot in input order with its ordered nuclei, scientific model identity,
basis metadata, solver controls, source backend, last status and, when a
converged seed is retained, the seed's diagnostics and blob names
(``density_<i>`` of shape [spins, n, n], ``coordinates_<i>`` of shape [natom, 3]).

Identities follow the reference's definitions where they are formulas over
the inputs (the geometry hash over ordered nuclei and hex-spelled
coordinates; the model identity as the canonical hash of the resolved model);
the basis hash is this package's canonical hash of the expanded shells.
Loading classifies each slot against the target batch (exact restart, warm
compatible, transport required, incompatible), imports only what is allowed,
and never bypasses the target's own SCF: a restored density is a seed.
"""

from __future__ import annotations

import ctypes as ct
import hashlib
import json
import math
import os
import struct
import tempfile
import time
from pathlib import Path

import numpy as np

from . import _native
from .basis import canonical_hash

MAGIC = b"VQHFCP01"
HEADER = struct.Struct("<8sQ32s")
ARRAY_FORMAT = "ieee754-binary64-little-endian-c-order"
SCHEMA = "generativeqc.hf_checkpoint"
MAX_MANIFEST = 4 << 20
MAX_ITEMS = 10000
PROVIDER = "generativeqc.hf.ri-pseudoinverse/v1"
HAMILTONIAN = "all-electron-nonrelativistic-coulomb"
ITEM_KEYS = frozenset("index atomic_numbers request_identity basis_metadata model provider controls backend status seed".split())
SEED_KEYS = frozenset("density coordinates nbf energy energy_change density_rms iterations converged".split())
BLOB_KEYS = frozenset("name shape offset bytes sha256".split())
CONTROL_KEYS = frozenset(
    "max_iterations energy_tolerance density_tolerance diis_history screening_tolerance "
    "target_accuracy precision_policy runtime_policy".split()
)
MODEL_KEYS = frozenset(
    "method geometry_hash basis_hash electron_count multiplicity charge representation hamiltonian "
    "approximation auxiliary_basis_hash metric_relative_threshold schema_version identity".split()
)


class CheckpointError(ValueError):
    """A checkpoint is malformed or incompatible; no seed has been applied."""


# -- identities ----------------------------------------------------------------


def _hex(x: float) -> str:
    return float(x).hex() if x else "0x0.0p+0"


def geometry_hash(numbers, coordinates) -> str:
    """Canonical hash of ordered (Z, hex-spelled xyz) pairs."""
    return canonical_hash([[int(z), [_hex(v) for v in xyz]] for z, xyz in zip(numbers, coordinates)])


def _basis_metadata(basis, role: str) -> dict:
    payload = {
        "representation": "spherical" if basis.representation == "real_spherical" else "cartesian",
        "elements": {"atomic_numbers": [a.atomic_number for a in basis.atoms]},
        "shells": [[s.atom_index, s.l, [[_hex(p.exponent), _hex(p.coefficient)] for p in s.primitives]] for s in basis.shells],
        "normalization": "normalized-primitives; native-normalized-contractions",
    }
    identity = canonical_hash(payload)
    return {
        "role": role,
        "mathematical_identity": identity,
        "representation": payload["representation"],
        "electrons": {
            "atomic_numbers": [a.atomic_number for a in basis.atoms],
            "ionic_charge": basis.charge,
            "electron_count": basis.electrons,
            "multiplicity": basis.multiplicity,
        },
        "shell_count": len(basis.shells),
        "primitive_count": sum(len(s.primitives) for s in basis.shells),
    }


def resolved_model(calculator, atoms, charge: int, multiplicity: int) -> tuple[dict, dict]:
    """(model, basis_metadata) of one system under ``calculator``."""
    orbital, aux = calculator._bases(atoms, charge, multiplicity)
    meta = {"orbital": _basis_metadata(orbital, "orbital"), "auxiliary": _basis_metadata(aux, "auxiliary")}
    meta["model_identity"] = canonical_hash(
        {"calculator": _signature(calculator), "basis": meta, "charge": int(charge), "multiplicity": int(multiplicity)}
    )
    model = {
        "method": calculator.method,
        "geometry_hash": geometry_hash([a.atomic_number for a in atoms], [a.position for a in atoms]),
        "basis_hash": meta["orbital"]["mathematical_identity"],
        "electron_count": orbital.electrons,
        "multiplicity": int(multiplicity),
        "charge": int(charge),
        "representation": calculator.representation,
        "hamiltonian": HAMILTONIAN,
        "approximation": "density_fitting",
        "auxiliary_basis_hash": meta["auxiliary"]["mathematical_identity"],
        "metric_relative_threshold": float(calculator.metric_threshold),
        "schema_version": 1,
    }
    model["identity"] = canonical_hash(model)
    return model, meta


def _signature(calculator) -> dict:
    return {
        "method": calculator.method,
        "representation": calculator.representation,
        "metric_relative_threshold": float(calculator.metric_threshold),
    }


def controls(calculator) -> dict:
    return {
        "max_iterations": int(calculator.max_iterations),
        "energy_tolerance": float(calculator.energy_tolerance),
        "density_tolerance": float(calculator.density_tolerance),
        "diis_history": int(calculator.diis_history),
        "screening_tolerance": float(calculator.screening_tolerance),
        "target_accuracy": None,
        "precision_policy": calculator.precision,
        "runtime_policy": {"VIBEQC_BATCH_WORKERS": os.environ.get("VIBEQC_BATCH_WORKERS")},
    }


# -- reading and validation ----------------------------------------------------


def _manifest_bytes(doc) -> bytes:
    return json.dumps(doc, sort_keys=True, separators=(",", ":"), allow_nan=False).encode("utf-8")


def _unique(pairs):
    out = {}
    for k, v in pairs:
        if k in out:
            raise CheckpointError(f"duplicate manifest key: {k}")
        out[k] = v
    return out


def _reject_constant(name):
    raise CheckpointError(f"nonfinite JSON: {name}")


def _int(value, name, lo, hi):
    if type(value) is not int or not lo <= value <= hi:
        raise CheckpointError(f"invalid {name}")
    return value


def _fields(value, keys, name):
    if not isinstance(value, dict) or set(value) != keys:
        raise CheckpointError(f"unknown or missing required {name} fields")


def _is_hash(v) -> bool:
    return isinstance(v, str) and len(v) == 64 and all(c in "0123456789abcdef" for c in v)


def _number(v, name, *, nonnegative=False):
    if type(v) not in (int, float) or not math.isfinite(v) or (nonnegative and v < 0):
        raise CheckpointError(f"invalid {name}")
    return float(v)


def parse_manifest(raw: bytes, payload_bytes: int, max_bytes: int) -> dict:
    """Validate a manifest against the schema and the payload length."""
    try:
        doc = json.loads(raw, object_pairs_hook=_unique, parse_constant=_reject_constant)
    except (UnicodeError, ValueError, RecursionError) as error:
        if isinstance(error, CheckpointError):
            raise
        raise CheckpointError(f"invalid checkpoint manifest: {error}") from error
    _fields(doc, {"schema", "schema_version", "array_format", "items", "blobs"}, "manifest")
    if doc["schema"] != SCHEMA or type(doc["schema_version"]) is not int or doc["schema_version"] != 1:
        raise CheckpointError("unsupported required checkpoint schema")
    if doc["array_format"] != ARRAY_FORMAT:
        raise CheckpointError("unsupported dtype/endian/array format")
    items, blobs = doc["items"], doc["blobs"]
    if not isinstance(items, list) or not 1 <= len(items) <= MAX_ITEMS:
        raise CheckpointError("invalid checkpoint item count")
    if not isinstance(blobs, list) or len(blobs) > 2 * len(items):
        raise CheckpointError("invalid checkpoint blob count")
    expected = {}
    for index, item in enumerate(items):
        _fields(item, ITEM_KEYS, "item")
        if type(item["index"]) is not int or item["index"] != index:
            raise CheckpointError("checkpoint item ordering mismatch")
        numbers = item["atomic_numbers"]
        if not isinstance(numbers, list) or not numbers or len(numbers) > max_bytes // 24:
            raise CheckpointError("invalid ordered nuclei")
        for z in numbers:
            _int(z, "atomic number", 1, 118)
        if not _is_hash(item["request_identity"]) or not isinstance(item["basis_metadata"], dict):
            raise CheckpointError("missing prepared identity/provenance")
        if item["basis_metadata"].get("model_identity") != item["request_identity"]:
            raise CheckpointError("prepared identity/provenance mismatch")
        _int(item["status"], "item status", -1, 2**31 - 1)
        if item["backend"] not in ("cpu", "cuda"):
            raise CheckpointError("invalid solver provenance")
        c = item["controls"]
        _fields(c, CONTROL_KEYS, "solver controls")
        for name in ("max_iterations", "diis_history"):
            _int(c[name], name, 0, 2**31 - 1)
        for name in ("energy_tolerance", "density_tolerance", "screening_tolerance"):
            _number(c[name], name, nonnegative=True)
        seed = item["seed"]
        if seed is None:
            if item["model"] is not None or item["provider"] is not None:
                raise CheckpointError("no-state item carries reusable identity")
            continue
        model = item["model"]
        _fields(model, MODEL_KEYS, "model")
        body = {k: v for k, v in model.items() if k != "identity"}
        if canonical_hash(body) != model["identity"]:
            raise CheckpointError("resolved-model identity checksum mismatch")
        if item["basis_metadata"]["orbital"]["mathematical_identity"] != model["basis_hash"]:
            raise CheckpointError("orbital basis provenance/model mismatch")
        if sum(numbers) - model["charge"] != model["electron_count"]:
            raise CheckpointError("impossible electron count for ordered nuclei")
        if item["provider"] != PROVIDER:
            raise CheckpointError("unsupported scientific provider identity")
        _fields(seed, SEED_KEYS, "seed")
        n = _int(seed["nbf"], "AO dimension", 1, math.isqrt(max_bytes // 8))
        _int(seed["iterations"], "source iterations", 0, 2**31 - 1)
        if seed["converged"] is not True:
            raise CheckpointError("schema 1 requires a converged source seed")
        for name in ("energy", "energy_change", "density_rms"):
            _number(seed[name], "source diagnostics")
        if seed["density_rms"] < 0:
            raise CheckpointError("negative source residual")
        spins = 2 if model["method"] == "uhf" else 1
        for name, shape in (("density", [spins, n, n]), ("coordinates", [len(numbers), 3])):
            if seed[name] != f"{name}_{index}":
                raise CheckpointError("unexpected blob name")
            expected[seed[name]] = shape
    offset = 0
    seen = set()
    for blob in blobs:
        _fields(blob, BLOB_KEYS, "blob")
        name, shape = blob["name"], blob["shape"]
        if not isinstance(shape, list) or any(type(d) is not int for d in shape):
            raise CheckpointError("noninteger blob dimensions")
        if not isinstance(name, str) or name in seen or name not in expected or shape != expected[name]:
            raise CheckpointError("inconsistent or duplicate blob dimensions")
        size = math.prod(shape) * 8
        if size > max_bytes or type(blob["offset"]) is not int or blob["offset"] != offset or blob["bytes"] != size:
            raise CheckpointError("inconsistent blob byte count/offset")
        if not _is_hash(blob["sha256"]):
            raise CheckpointError("invalid blob checksum")
        seen.add(name)
        offset += size
    if seen != set(expected) or offset != payload_bytes:
        raise CheckpointError("truncated, trailing or missing checkpoint blobs")
    return doc


def _read(path, max_bytes: int, retain: bool = True):
    _int(max_bytes, "checkpoint byte limit", HEADER.size, 2**63 - 1)
    arrays = {}
    with open(path, "rb") as f:
        size = os.fstat(f.fileno()).st_size
        if size > max_bytes:
            raise MemoryError("checkpoint file exceeds max_bytes")
        header = f.read(HEADER.size)
        if len(header) != HEADER.size:
            raise CheckpointError("truncated checkpoint header")
        magic, length, digest = HEADER.unpack(header)
        if magic != MAGIC or length > min(MAX_MANIFEST, size - HEADER.size):
            raise CheckpointError("invalid checkpoint header/schema or manifest length")
        raw = f.read(length)
        if hashlib.sha256(raw).digest() != digest:
            raise CheckpointError("manifest checksum mismatch")
        doc = parse_manifest(raw, size - HEADER.size - length, max_bytes)
        for blob in doc["blobs"]:
            h = hashlib.sha256()
            data = bytearray() if retain else None
            left = blob["bytes"]
            while left:
                chunk = f.read(min(left, 1 << 20))
                if not chunk:
                    raise CheckpointError("truncated checkpoint blob")
                h.update(chunk)
                if retain:
                    data += chunk
                left -= len(chunk)
            if h.hexdigest() != blob["sha256"]:
                raise CheckpointError(f"blob checksum mismatch: {blob['name']}")
            if retain:
                array = np.frombuffer(bytes(data), dtype="<f8").reshape(blob["shape"])
                if not np.isfinite(array).all():
                    raise CheckpointError("nonfinite checkpoint array")
                arrays[blob["name"]] = array
        if f.read(1):
            raise CheckpointError("checkpoint changed while reading")
    return doc, arrays, size


def inspect_checkpoint(path, *, max_bytes: int = 256 << 20) -> dict:
    """Verify header, schema, lengths and every checksum; returns the manifest.
    Nothing is loaded into any batch and no device is touched."""
    return _read(path, max_bytes, retain=False)[0]


# -- save / load ---------------------------------------------------------------


def _warm_state(batch, index: int):
    lib = batch._lib
    st = _native.descriptor(_native.HfWarmState)
    _native.check(lib.generativeqc_batch_get_hf_warm_state(batch._handle, index, ct.byref(st)), "warm state query failed")
    if not st.present:
        return None
    density = np.empty(st.density_count)
    coords = np.empty(st.coordinate_count)
    st.density, st.coordinates = _native.pointer(density), _native.pointer(coords)
    _native.check(lib.generativeqc_batch_get_hf_warm_state(batch._handle, index, ct.byref(st)), "warm state copy failed")
    return density, coords.reshape(-1, 3), st


def build_checkpoint(calculator, systems, charges, multiplicities, statuses, seeds, *, max_bytes: int = 256 << 20) -> bytes:
    """The bytes of a checkpoint file.  seeds[i] is None or a dict with the
    retained density (n x n restricted, 2 x n x n alpha/beta for UHF), the
    coordinates it converged at
    (natom x 3, Bohr) and its energy, energy_change, density_rms, iterations."""
    from .basis import Atom

    if len(systems) > MAX_ITEMS:
        raise CheckpointError("checkpoint item count exceeds schema limit")
    items, arrays, blobs = [], [], []
    offset = 0
    for i, atoms in enumerate(systems):
        q, m = charges[i], multiplicities[i]
        _, meta = resolved_model(calculator, atoms, q, m)
        item = {
            "index": i,
            "atomic_numbers": [a.atomic_number for a in atoms],
            "request_identity": meta["model_identity"],
            "basis_metadata": meta,
            "model": None,
            "provider": None,
            "controls": controls(calculator),
            # The ABI backend id this engine answers to (GENERATIVEQC_BACKEND_CUDA).
            "backend": "cuda",
            "status": int(statuses[i]),
            "seed": None,
        }
        seed = seeds[i]
        if seed is not None:
            spins = 2 if calculator.method == "uhf" else 1
            density = np.asarray(seed["density"], dtype=np.float64)
            coords = np.asarray(seed["coordinates"], dtype=np.float64).reshape(-1, 3)
            n = density.shape[-1]
            source_atoms = tuple(Atom(a.atomic_number, tuple(map(float, x))) for a, x in zip(atoms, coords))
            source_model, _ = resolved_model(calculator, source_atoms, q, m)
            item["model"], item["provider"] = source_model, PROVIDER
            item["seed"] = {
                "density": f"density_{i}",
                "coordinates": f"coordinates_{i}",
                "nbf": int(n),
                "energy": float(seed["energy"]),
                "energy_change": float(seed["energy_change"]),
                "density_rms": float(seed["density_rms"]),
                "iterations": int(seed["iterations"]),
                "converged": True,
            }
            for name, array in (("density", density.reshape(spins, n, n)), ("coordinates", coords)):
                a = np.ascontiguousarray(array, dtype="<f8")
                if not np.isfinite(a).all():
                    raise CheckpointError("nonfinite retained state")
                raw = a.tobytes()
                blobs.append({"name": f"{name}_{i}", "shape": list(a.shape), "offset": offset, "bytes": len(raw), "sha256": hashlib.sha256(raw).hexdigest()})
                arrays.append(raw)
                offset += len(raw)
        items.append(item)
    doc = {"schema": SCHEMA, "schema_version": 1, "array_format": ARRAY_FORMAT, "items": items, "blobs": blobs}
    raw = _manifest_bytes(doc)
    if len(raw) > MAX_MANIFEST or HEADER.size + len(raw) + offset > max_bytes:
        raise MemoryError("checkpoint manifest/file exceeds byte limit")
    parse_manifest(raw, offset, max_bytes)
    return HEADER.pack(MAGIC, len(raw), hashlib.sha256(raw).digest()) + raw + b"".join(arrays)


def write_checkpoint(path, data: bytes, *, max_bytes: int = 256 << 20) -> None:
    """Write to a temporary file beside ``path``, fsync, verify, then rename."""
    path = Path(path)
    tmp = None
    try:
        with tempfile.NamedTemporaryFile(dir=path.parent, prefix=f".{path.name}.", suffix=".tmp", delete=False) as f:
            tmp = Path(f.name)
            f.write(data)
            f.flush()
            os.fsync(f.fileno())
        inspect_checkpoint(tmp, max_bytes=max_bytes)
        os.replace(tmp, path)
        tmp = None
    finally:
        if tmp is not None:
            tmp.unlink(missing_ok=True)


def save_checkpoint(batch, path, *, max_bytes: int = 256 << 20) -> dict:
    """Persist a prepared batch's retained seeds (atomic replace of ``path``)."""
    start = time.perf_counter()
    batch._ensure_open()
    seeds = []
    for i in range(len(batch._systems)):
        state = _warm_state(batch, i)
        if state is None:
            seeds.append(None)
            continue
        density, coords, st = state
        spins = 2 if batch._calculator.method == "uhf" else 1
        n = math.isqrt(density.size // spins)
        seeds.append(
            {
                "density": density.reshape(spins, n, n) if spins == 2 else density.reshape(n, n),
                "coordinates": coords,
                "energy": st.energy,
                "energy_change": st.energy_change,
                "density_rms": st.density_rms,
                "iterations": st.iterations,
            }
        )
    data = build_checkpoint(
        batch._calculator, batch._systems, batch._charges, batch._multiplicities, batch._last_status, seeds, max_bytes=max_bytes
    )
    write_checkpoint(path, data, max_bytes=max_bytes)
    length = HEADER.unpack(data[: HEADER.size])[1]
    return {
        "schema_version": 1,
        "bytes": len(data),
        "write_seconds": time.perf_counter() - start,
        "manifest_sha256": hashlib.sha256(data[HEADER.size : HEADER.size + length]).hexdigest(),
        "saved_indices": [i for i, sd in enumerate(seeds) if sd is not None],
    }


def load_checkpoint(batch, path, *, allow_warm: bool = False, strict: bool = True, max_bytes: int = 256 << 20) -> dict:
    """Validate the whole file, classify every slot against the target batch,
    then import the accepted seeds in one native call."""
    from .basis import Atom

    start = time.perf_counter()
    batch._ensure_open()
    doc, arrays, size = _read(path, max_bytes)
    if len(doc["items"]) != len(batch._systems):
        raise CheckpointError("checkpoint item count/order differs from target batch")
    calc = batch._calculator
    reports, states, keep = [], [], []
    for i, item in enumerate(doc["items"]):
        level, reason = "incompatible", "missing reusable state"
        seed = item["seed"]
        target_atoms = batch._systems[i]
        st = _native.descriptor(_native.HfWarmState)
        if seed is not None:
            coords = arrays[seed["coordinates"]]
            if geometry_hash(item["atomic_numbers"], coords) != item["model"]["geometry_hash"]:
                raise CheckpointError("source coordinate/model identity mismatch")
            target, _ = resolved_model(calc, target_atoms, batch._charges[i], batch._multiplicities[i])
            src = {k: v for k, v in item["model"].items() if k not in ("identity", "geometry_hash")}
            tgt = {k: v for k, v in target.items() if k not in ("identity", "geometry_hash")}
            if [a.atomic_number for a in target_atoms] != item["atomic_numbers"]:
                reason = "ordered nuclei differ"
            elif src != tgt:
                basis_differs = src["basis_hash"] != tgt["basis_hash"] or src["representation"] != tgt["representation"]
                level = "transport_required" if basis_differs else "incompatible"
                reason = "basis/AO transport required" if basis_differs else "method/core/spin/provider identity differs"
            elif target["geometry_hash"] != item["model"]["geometry_hash"] or item["controls"] != controls(calc):
                level, reason = "warm_start_compatible", "geometry or numerical controls differ"
            else:
                level, reason = "exact_restart", "same scientific identity, geometry and controls"
        accepted = level == "exact_restart" or (allow_warm and level == "warm_start_compatible")
        if accepted:
            density = np.ascontiguousarray(arrays[seed["density"]].reshape(-1), dtype=np.float64)
            coords = np.ascontiguousarray(arrays[seed["coordinates"]].reshape(-1), dtype=np.float64)
            keep.extend((density, coords))
            st.density, st.coordinates = _native.pointer(density), _native.pointer(coords)
            st.density_count, st.coordinate_count = density.size, coords.size
            st.energy, st.energy_change, st.density_rms = seed["energy"], seed["energy_change"], seed["density_rms"]
            st.iterations, st.present = seed["iterations"], 1
        states.append(st)
        reports.append(
            {
                "index": i,
                "compatibility": level,
                "reason": reason,
                "restored_fields": ["density"] if accepted else [],
                "rejected_fields": [] if accepted or seed is None else ["density"],
                "source_status": item["status"],
                "source_model_identity": item["model"]["identity"] if seed else None,
            }
        )
    rejected = [r for r in reports if not r["restored_fields"]]
    if strict and rejected:
        raise CheckpointError("checkpoint restore rejected: " + "; ".join(f"{r['index']}: {r['reason']}" for r in rejected))
    accepted = {r["index"] for r in reports if r["restored_fields"]}
    if accepted:
        arr = (_native.HfWarmState * len(states))(*states)
        status = batch._lib.generativeqc_batch_restore_hf_warm_states(batch._handle, arr, len(states))
        if status != 0:
            raise CheckpointError(f"native checkpoint validation failed (status {status})")
    batch._restart_indices |= accepted
    return {
        "schema_version": 1,
        "bytes": size,
        "manifest_identity": canonical_hash(doc),
        "read_seconds": time.perf_counter() - start,
        "items": reports,
        "target_verification": "pending_execute",
    }


__all__ = ["CheckpointError", "build_checkpoint", "inspect_checkpoint", "load_checkpoint", "save_checkpoint", "write_checkpoint"]
