"""CPU checks of the C ABI library and the host logic (no device calls)."""

from __future__ import annotations

import ctypes as ct
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT
from vibeqc_amd import _native
from vibeqc_amd.basis import AOBasis, load_basis, shells_for
from vibeqc_amd.geometry import alkane, lowdin_occupied, water_cluster


def _declared_exports():
    names = []
    for header in (ROOT / "include").rglob("*.h"):
        text = header.read_text()
        names += re.findall(r"GENERATIVEQC_API\s+[\w\s\*]+?\b((?:generativeqc|vibeqc)_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_loads_and_exports_every_declared_symbol():
    lib = ct.CDLL(str(_native.library_path()))
    declared = _declared_exports()
    assert len(declared) >= 19
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # The Python binding binds exactly the declared surface.
    assert set(_native.EXPORTS) == set(declared)


def test_abi_constants_match_reference_values():
    lib = _native.load()
    assert lib.generativeqc_get_abi_version() == 0
    assert lib.generativeqc_status_message(4) == b"not converged"
    assert b"gfx950" in lib.generativeqc_get_source_identity()


def test_descriptor_layouts():
    # Field layouts match the reference's fock.h / generativeqc.h structs.
    assert ct.sizeof(_native.Atom) == 32
    assert ct.sizeof(_native.FockTermC) == 40
    assert ct.sizeof(_native.FockResultC) == 8 + 16 + 6 * 8 + 24
    assert ct.sizeof(_native.ScfControlsC) == 32


def test_cpu_backend_is_refused_without_device_work():
    lib = _native.load()
    desc = _native.descriptor(_native.ContextDescriptor, device_id=0, backend=_native.BACKEND_CPU_REFERENCE)
    handle = ct.c_void_p()
    assert lib.generativeqc_context_create(ct.byref(desc), ct.byref(handle)) == 3  # NOT_IMPLEMENTED
    assert not handle
    bad = _native.ContextDescriptor(struct_size=4, abi_version=0, device_id=0, backend=1)
    assert lib.generativeqc_context_create(ct.byref(bad), ct.byref(handle)) == 2  # ABI_MISMATCH
    assert lib.generativeqc_context_create(None, ct.byref(handle)) == 1


def test_basis_expansion_matches_reference_counts():
    atoms = water_cluster(4, 6.0, seed=1)
    orb = AOBasis.build(atoms, "cc-pvdz")
    aux = AOBasis.build(atoms, "cc-pvdz-jkfit")
    assert orb.nao == 96 and aux.nao == 464
    # O cc-pVDZ has a general contraction: its 8-primitive s block expands to two shells.
    o = load_basis("cc-pvdz")[8]
    assert len(o[0]["coefficients"]) == 2
    assert sum(1 for s in shells_for("cc-pvdz", atoms[:1]) if s.l == 0) == 3


def test_water_cluster_construction():
    atoms = water_cluster(16, 3.0 * 16 ** (1 / 3), seed=5)
    xyz = np.array([a.position for a in atoms]) * 0.529177210903
    o = xyz[0::3]
    for i in range(16):
        assert abs(np.linalg.norm(xyz[3 * i + 1] - o[i]) - 0.9572) < 1e-12
        assert abs(np.linalg.norm(xyz[3 * i + 2] - o[i]) - 0.9572) < 1e-12
        v1, v2 = xyz[3 * i + 1] - o[i], xyz[3 * i + 2] - o[i]
        ang = np.degrees(np.arccos(v1 @ v2 / (np.linalg.norm(v1) * np.linalg.norm(v2))))
        assert abs(ang - 104.52) < 1e-9
    d = np.linalg.norm(o[:, None] - o[None], axis=-1) + np.eye(16) * 10
    assert d.min() >= 2.6 - 1e-12


def test_alkane_construction():
    atoms = alkane(20)
    z = [a.atomic_number for a in atoms]
    assert z.count(6) == 20 and z.count(1) == 42
    orb = AOBasis.build(atoms, "def2-svp")
    assert orb.nao == 490
    xyz = np.array([a.position for a in atoms]) * 0.529177210903
    c = xyz[:20]
    assert np.allclose(np.linalg.norm(np.diff(c, axis=0), axis=1), 1.54)
    h = xyz[20:]
    dch = np.linalg.norm(h[:, None] - c[None], axis=-1).min(axis=1)
    assert np.allclose(dch, 1.09)


def test_lowdin_occupied_is_orthonormal():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(10, 10))
    S = a @ a.T + 10 * np.eye(10)
    C = lowdin_occupied(S, 4, seed=3)
    assert np.allclose(C.T @ S @ C, np.eye(4), atol=1e-12)
