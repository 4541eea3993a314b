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
from dataclasses import dataclass
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


# -- manifest schema -------------------------------------------------------------
#
# The manifest is checked in two layers.  First every record is conformed to a
# field table: the record must carry exactly the listed fields and each field
# must pass its predicate (each table entry names what is wrong when it does
# not).  Then the rules that tie fields together run per item and over the blob
# table.  Nothing is imported from a file that fails either layer.


def _sha_hex(v) -> bool:
    return isinstance(v, str) and len(v) == 64 and set(v) <= set("0123456789abcdef")


def _count(lo, hi):
    return lambda v: type(v) is int and lo <= v <= hi


def _real(nonnegative=False):
    return lambda v: type(v) in (int, float) and math.isfinite(v) and (v >= 0 or not nonnegative)


_ANY = lambda v: True  # noqa: E731
_I32 = 2**31 - 1

_RECORDS = {
    "manifest": {
        "schema": (lambda v: v == SCHEMA, "unsupported checkpoint schema"),
        "schema_version": (lambda v: type(v) is int and v == 1, "unsupported checkpoint schema version"),
        "array_format": (lambda v: v == ARRAY_FORMAT, "unsupported array format (dtype, endianness or order)"),
        "items": (lambda v: isinstance(v, list) and 1 <= len(v) <= MAX_ITEMS, "item count outside 1..10000"),
        "blobs": (lambda v: isinstance(v, list), "blob table is not a list"),
    },
    "item": {
        "index": (lambda v: type(v) is int, "item index is not an integer"),
        "atomic_numbers": (lambda v: isinstance(v, list) and len(v) > 0 and all(_count(1, 118)(z) for z in v),
                           "ordered nuclei must be atomic numbers 1..118"),
        "request_identity": (_sha_hex, "request identity is not a SHA-256 digest"),
        "basis_metadata": (lambda v: isinstance(v, dict), "basis metadata is not an object"),
        "model": (_ANY, ""),
        "provider": (_ANY, ""),
        "controls": (_ANY, ""),
        "backend": (lambda v: v in ("cpu", "cuda"), "unknown source backend"),
        "status": (_count(-1, _I32), "item status out of range"),
        "seed": (_ANY, ""),
    },
    "controls": {
        "max_iterations": (_count(0, _I32), "max_iterations out of range"),
        "diis_history": (_count(0, _I32), "diis_history out of range"),
        "energy_tolerance": (_real(True), "energy_tolerance must be finite and non-negative"),
        "density_tolerance": (_real(True), "density_tolerance must be finite and non-negative"),
        "screening_tolerance": (_real(True), "screening_tolerance must be finite and non-negative"),
        "target_accuracy": (_ANY, ""),
        "precision_policy": (_ANY, ""),
        "runtime_policy": (_ANY, ""),
    },
    "model": {key: (_ANY, "") for key in MODEL_KEYS},
    "seed": {
        "density": (lambda v: isinstance(v, str), "density blob name is not a string"),
        "coordinates": (lambda v: isinstance(v, str), "coordinates blob name is not a string"),
        "nbf": (_count(1, _I32), "AO dimension out of range"),
        "energy": (_real(), "source energy must be finite"),
        "energy_change": (_real(), "source energy change must be finite"),
        "density_rms": (_real(True), "source density residual must be finite and non-negative"),
        "iterations": (_count(0, _I32), "source iteration count out of range"),
        "converged": (lambda v: v is True, "schema 1 stores converged seeds only; this seed is not converged"),
    },
    "blob": {
        "name": (lambda v: isinstance(v, str), "blob name is not a string"),
        "shape": (lambda v: isinstance(v, list) and all(type(d) is int and d >= 0 for d in v), "blob dimensions are not non-negative integers"),
        "offset": (lambda v: type(v) is int, "blob offset is not an integer"),
        "bytes": (lambda v: type(v) is int, "blob byte count is not an integer"),
        "sha256": (_sha_hex, "blob checksum is not a SHA-256 digest"),
    },
}


def _conform(kind: str, record, where: str) -> None:
    table = _RECORDS[kind]
    if not isinstance(record, dict):
        raise CheckpointError(f"{where}: {kind} record is not an object")
    if record.keys() != table.keys():
        extra = sorted(set(record) - set(table))
        absent = sorted(set(table) - set(record))
        raise CheckpointError(f"{where}: unknown or missing {kind} fields (unknown {extra}, missing {absent})")
    for key, (ok, problem) in table.items():
        if not ok(record[key]):
            raise CheckpointError(f"{where}: {problem}")


def _strict_object(pairs):
    seen = {}
    for key, value in pairs:
        if key in seen:
            raise CheckpointError(f"duplicate manifest key {key!r}")
        seen[key] = value
    return seen


def _no_constants(token):
    raise CheckpointError(f"nonfinite JSON constant {token} in the manifest")


def _seeded_item_rules(item, where, max_bytes):
    """Cross-field rules of an item that carries a seed; returns the blob
    shapes the seed needs, by name."""
    model, seed, index = item["model"], item["seed"], item["index"]
    _conform("model", model, where)
    if canonical_hash({k: v for k, v in model.items() if k != "identity"}) != model["identity"]:
        raise CheckpointError(f"{where}: model identity checksum mismatch")
    if item["basis_metadata"]["orbital"]["mathematical_identity"] != model["basis_hash"]:
        raise CheckpointError(f"{where}: orbital basis identity differs from the model's basis hash")
    if sum(item["atomic_numbers"]) - model["charge"] != model["electron_count"]:
        raise CheckpointError(f"{where}: electron count impossible for these nuclei and charge")
    if item["provider"] != PROVIDER:
        raise CheckpointError(f"{where}: unknown scientific provider {item['provider']!r}")
    _conform("seed", seed, where)
    n = seed["nbf"]
    if n * n * 8 > max_bytes:
        raise CheckpointError(f"{where}: AO dimension exceeds the byte limit")
    spins = 2 if model["method"] == "uhf" else 1
    wanted = {f"density_{index}": [spins, n, n], f"coordinates_{index}": [len(item["atomic_numbers"]), 3]}
    if (seed["density"], seed["coordinates"]) != tuple(wanted):
        raise CheckpointError(f"{where}: seed blob names must be density_{index} and coordinates_{index}")
    return wanted


def _blob_layout(blobs, wanted, max_bytes) -> int:
    """Check the blob table against the shapes the seeds need; blobs follow
    each other without gaps from payload offset 0.  Returns the payload length."""
    end = 0
    named = set()
    for k, blob in enumerate(blobs):
        where = f"blob {k}"
        _conform("blob", blob, where)
        name = blob["name"]
        if name in named or name not in wanted:
            raise CheckpointError(f"{where}: unexpected or duplicate blob {name!r}")
        if blob["shape"] != wanted[name]:
            raise CheckpointError(f"{where}: blob dimensions {blob['shape']} differ from the seed's {wanted[name]}")
        nbytes = 8 * math.prod(blob["shape"])
        if nbytes > max_bytes or blob["bytes"] != nbytes:
            raise CheckpointError(f"{where}: blob byte count does not match its dimensions")
        if blob["offset"] != end:
            raise CheckpointError(f"{where}: blob offset {blob['offset']} is not the end of the previous blob ({end})")
        named.add(name)
        end += nbytes
    if named != set(wanted):
        raise CheckpointError(f"missing blobs: {sorted(set(wanted) - named)}")
    return end


def parse_manifest(raw: bytes, payload_bytes: int, max_bytes: int) -> dict:
    """Decode and validate a manifest for a payload of ``payload_bytes``."""
    try:
        doc = json.loads(raw, object_pairs_hook=_strict_object, parse_constant=_no_constants)
    except CheckpointError:
        raise
    except (UnicodeError, ValueError, RecursionError) as error:
        raise CheckpointError(f"manifest is not valid JSON: {error}") from error
    _conform("manifest", doc, "manifest")
    if len(doc["blobs"]) > 2 * len(doc["items"]):
        raise CheckpointError("manifest: more blobs than the items' seeds can own")
    wanted = {}
    for position, item in enumerate(doc["items"]):
        where = f"item {position}"
        _conform("item", item, where)
        if item["index"] != position:
            raise CheckpointError(f"{where}: item ordering broken (index {item['index']})")
        if len(item["atomic_numbers"]) > max_bytes // 24:
            raise CheckpointError(f"{where}: too many nuclei for the byte limit")
        if item["basis_metadata"].get("model_identity") != item["request_identity"]:
            raise CheckpointError(f"{where}: prepared identity differs from the basis metadata's model identity")
        _conform("controls", item["controls"], where)
        if item["seed"] is None:
            if item["model"] is not None or item["provider"] is not None:
                raise CheckpointError(f"{where}: no-state item carries a model or provider")
            continue
        wanted.update(_seeded_item_rules(item, where, max_bytes))
    declared = _blob_layout(doc["blobs"], wanted, max_bytes)
    if payload_bytes < declared:
        raise CheckpointError(f"payload truncated: {payload_bytes} bytes of {declared}")
    if payload_bytes > declared:
        raise CheckpointError(f"trailing bytes after the last blob: {payload_bytes - declared}")
    return doc


# -- reading ---------------------------------------------------------------------


class _CheckpointFile:
    """Random-access view of one checkpoint file.  Blobs are read at their
    absolute offsets (header + manifest + blob offset) with positional reads,
    so verification does not depend on read order, and the file's size and
    modification time are compared at the end to catch a concurrent writer."""

    def __init__(self, path, max_bytes: int):
        if type(max_bytes) is not int or max_bytes < HEADER.size:
            raise CheckpointError("byte limit must be an integer of at least the header size")
        self._fd = os.open(path, os.O_RDONLY)
        try:
            stat = os.fstat(self._fd)
            self.size, self._stamp = stat.st_size, (stat.st_size, stat.st_mtime_ns)
            if self.size > max_bytes:
                raise MemoryError(f"checkpoint of {self.size} bytes exceeds max_bytes={max_bytes}")
            magic, length, digest = HEADER.unpack(self._exact(0, HEADER.size, "header"))
            if magic != MAGIC or length > min(MAX_MANIFEST, self.size - HEADER.size):
                raise CheckpointError("not a checkpoint header (magic or manifest length)")
            raw = self._exact(HEADER.size, length, "manifest")
            if hashlib.sha256(raw).digest() != digest:
                raise CheckpointError("manifest checksum mismatch")
            self.payload_start = HEADER.size + length
            self.manifest = parse_manifest(raw, self.size - self.payload_start, max_bytes)
        except BaseException:
            os.close(self._fd)
            raise

    def _exact(self, offset: int, length: int, what: str) -> bytes:
        data = os.pread(self._fd, length, offset)
        if len(data) != length:
            raise CheckpointError(f"truncated checkpoint {what}")
        return data

    def blob(self, blob: dict, keep: bool):
        """Verify one blob's checksum; with keep, return it as an array."""
        digest = hashlib.sha256()
        pieces = [] if keep else None
        base = self.payload_start + blob["offset"]
        for at in range(0, blob["bytes"], 1 << 22):
            piece = self._exact(base + at, min(1 << 22, blob["bytes"] - at), "blob")
            digest.update(piece)
            if keep:
                pieces.append(piece)
        if digest.hexdigest() != blob["sha256"]:
            raise CheckpointError(f"blob checksum mismatch for {blob['name']}")
        if not keep:
            return None
        array = np.frombuffer(b"".join(pieces), dtype="<f8").reshape(blob["shape"])
        if not np.all(np.isfinite(array)):
            raise CheckpointError(f"nonfinite values in {blob['name']}")
        return array

    def close(self) -> None:
        stat = os.fstat(self._fd)
        os.close(self._fd)
        if (stat.st_size, stat.st_mtime_ns) != self._stamp:
            raise CheckpointError("checkpoint file changed while it was read")


def _open_verified(path, max_bytes: int, keep_arrays: bool):
    """(manifest, {blob name: array}, file size) after every check passed."""
    view = _CheckpointFile(path, max_bytes)
    try:
        arrays = {b["name"]: view.blob(b, keep_arrays) for b in view.manifest["blobs"]}
    except BaseException:
        os.close(view._fd)
        raise
    view.close()
    return view.manifest, (arrays if keep_arrays else {}), view.size


def inspect_checkpoint(path, *, max_bytes: int = 256 << 20) -> dict:
    """Verify header, schema, lengths and every checksum; returns the manifest.
    Nothing is loaded into any batch and no device is touched."""
    return _open_verified(path, max_bytes, keep_arrays=False)[0]


# -- save / load ---------------------------------------------------------------


def _manifest_bytes(doc) -> bytes:
    """Canonical manifest encoding: sorted keys, compact separators, no NaN."""
    return json.dumps(doc, sort_keys=True, separators=(",", ":"), allow_nan=False).encode("utf-8")


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


@dataclass
class _SlotDecision:
    """What loading does with one slot of the target batch."""

    index: int
    level: str
    reason: str
    source_status: int
    source_identity: str | None
    seed: dict | None = None  # the seed record when the slot held one
    density: np.ndarray | None = None  # set when the seed is accepted
    coordinates: np.ndarray | None = None

    @property
    def accepted(self) -> bool:
        return self.density is not None

    def report(self) -> dict:
        return {
            "index": self.index,
            "compatibility": self.level,
            "reason": self.reason,
            "restored_fields": ["density"] if self.accepted else [],
            "rejected_fields": ["density"] if self.seed is not None and not self.accepted else [],
            "source_status": self.source_status,
            "source_model_identity": self.source_identity,
        }


def _without(model: dict, *keys) -> dict:
    return {k: v for k, v in model.items() if k not in keys}


def _level_of(item, target_atoms, target_model, target_controls) -> tuple[str, str]:
    """Compatibility level of a seeded item against its target slot: the first
    mismatch found, from the strongest (different chemistry) to the weakest
    (different geometry or numerical controls), decides."""
    source = item["model"]
    same_identity = _without(source, "identity", "geometry_hash") == _without(target_model, "identity", "geometry_hash")
    basis_moved = source["basis_hash"] != target_model["basis_hash"] or source["representation"] != target_model["representation"]
    checks = (
        ([a.atomic_number for a in target_atoms] == item["atomic_numbers"], "incompatible", "ordered nuclei differ"),
        (same_identity or not basis_moved, "transport_required", "basis/AO transport required"),
        (same_identity, "incompatible", "method/core/spin/provider identity differs"),
        (target_model["geometry_hash"] == source["geometry_hash"] and item["controls"] == target_controls,
         "warm_start_compatible", "geometry or numerical controls differ"),
    )
    for holds, level, reason in checks:
        if not holds:
            return level, reason
    return "exact_restart", "same scientific identity, geometry and controls"


def _install(batch, decisions) -> None:
    """Hand every accepted seed to the native batch in one call; the native
    side validates them all before it replaces any."""
    states = []
    for d in decisions:
        st = _native.descriptor(_native.HfWarmState)
        if d.accepted:
            st.density, st.coordinates = _native.pointer(d.density), _native.pointer(d.coordinates)
            st.density_count, st.coordinate_count = d.density.size, d.coordinates.size
            st.energy, st.energy_change = d.seed["energy"], d.seed["energy_change"]
            st.density_rms, st.iterations, st.present = d.seed["density_rms"], d.seed["iterations"], 1
        states.append(st)
    table = (_native.HfWarmState * len(states))(*states)
    status = batch._lib.generativeqc_batch_restore_hf_warm_states(batch._handle, table, len(states))
    if status != 0:
        raise CheckpointError(f"native checkpoint validation failed (status {status})")


def load_checkpoint(batch, path, *, allow_warm: bool = False, strict: bool = True, max_bytes: int = 256 << 20) -> dict:
    """Validate the whole file, classify every slot against the target batch,
    then import the accepted seeds in one native call."""
    began = time.perf_counter()
    batch._ensure_open()
    doc, arrays, size = _open_verified(path, max_bytes, keep_arrays=True)
    targets = batch._systems
    if len(doc["items"]) != len(targets):
        raise CheckpointError(f"checkpoint holds {len(doc['items'])} slots, the target batch {len(targets)}")
    calc = batch._calculator
    wanted_controls = controls(calc)
    decisions = []
    for item, atoms, charge, mult in zip(doc["items"], targets, batch._charges, batch._multiplicities):
        seed = item["seed"]
        d = _SlotDecision(item["index"], "incompatible", "missing reusable state", item["status"],
                          item["model"]["identity"] if seed is not None else None, seed)
        if seed is not None:
            coords = arrays[seed["coordinates"]]
            if geometry_hash(item["atomic_numbers"], coords) != item["model"]["geometry_hash"]:
                raise CheckpointError(f"item {d.index}: stored coordinates do not match the model's geometry hash")
            target_model, _ = resolved_model(calc, atoms, charge, mult)
            d.level, d.reason = _level_of(item, atoms, target_model, wanted_controls)
            if d.level == "exact_restart" or (allow_warm and d.level == "warm_start_compatible"):
                d.density = np.ascontiguousarray(arrays[seed["density"]], dtype=np.float64).reshape(-1)
                d.coordinates = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1)
        decisions.append(d)
    refused = [d for d in decisions if not d.accepted]
    if strict and refused:
        raise CheckpointError("checkpoint restore rejected: " + "; ".join(f"{d.index}: {d.reason}" for d in refused))
    restored = {d.index for d in decisions if d.accepted}
    if restored:
        _install(batch, decisions)
    batch._restart_indices |= restored
    return {
        "schema_version": 1,
        "bytes": size,
        "manifest_identity": canonical_hash(doc),
        "read_seconds": time.perf_counter() - began,
        "items": [d.report() for d in decisions],
        "target_verification": "pending_execute",
    }


__all__ = ["CheckpointError", "build_checkpoint", "inspect_checkpoint", "load_checkpoint", "save_checkpoint", "write_checkpoint"]
