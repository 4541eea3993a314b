"""Seeded synthetic molecules for the benchmark workloads.

* Water clusters (H2O)n: rigid monomers (O-H 0.9572 A, H-O-H 104.52 deg),
  uniformly random orientations (normalised Gaussian quaternion), O centres
  uniform in a cube, rejection-sampled until every O-O distance >= 2.6 A;
  optional N(0, sigma) per-atom displacements.
* All-trans n-alkanes CnH(2n+2): zig-zag carbon chain (C-C 1.54 A, C-H 1.09 A,
  tetrahedral angles).
* Lowdin-orthonormalised random occupied factors C <- C (C^T S C)^(-1/2).
Coordinates are returned in Bohr (1 A = 1/0.529177210903 Bohr).
"""

from __future__ import annotations

import math

import numpy as np

from .basis import BOHR_PER_ANGSTROM, Atom

OH = 0.9572
HOH = math.radians(104.52)
TETRAHEDRAL = math.acos(-1.0 / 3.0)


def _rotation(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ]
    )


def water_cluster(n: int, box_angstrom: float, seed: int, displacement_angstrom: float = 0.0):
    """(H2O)n as a tuple of Atoms (Bohr); O, H, H per monomer."""
    rng = np.random.default_rng(seed)
    monomer = np.array(
        [[0.0, 0.0, 0.0], [OH, 0.0, 0.0], [OH * math.cos(HOH), OH * math.sin(HOH), 0.0]]
    )
    centres = []
    attempts = 0
    while len(centres) < n:
        attempts += 1
        if attempts > 1_000_000:
            raise RuntimeError("could not place water molecules with O-O >= 2.6 A")
        c = rng.uniform(0.0, box_angstrom, size=3)
        if all(np.linalg.norm(c - o) >= 2.6 for o in centres):
            centres.append(c)
    xyz = []
    z = []
    for c in centres:
        rot = _rotation(rng.normal(size=4))
        for k, zz in enumerate((8, 1, 1)):
            xyz.append(c + rot @ monomer[k])
            z.append(zz)
    xyz = np.array(xyz)
    if displacement_angstrom > 0.0:
        xyz = xyz + rng.normal(scale=displacement_angstrom, size=xyz.shape)
    xyz *= BOHR_PER_ANGSTROM
    return tuple(Atom(zz, tuple(map(float, p))) for zz, p in zip(z, xyz))


def ragged_water_batch(count: int = 512, seed: int = 2):
    """The ragged benchmark batch: (H2O)n with n ~ U{1..16} drawn from `seed`,
    member k built by water_cluster(n, 3 n^(1/3) A, seed=1000 + k).
    Returns (systems, sizes)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 17, size=count)
    systems = [water_cluster(int(n), 3.0 * float(n) ** (1.0 / 3.0), seed=1000 + k) for k, n in enumerate(sizes)]
    return systems, sizes


def alkane(n_carbon: int):
    """All-trans CnH(2n+2) zig-zag in the xy plane (Bohr)."""
    cc, ch = 1.54, 1.09
    half = TETRAHEDRAL / 2.0
    dx = cc * math.sin(half)
    dy = cc * math.cos(half)
    carbons = [np.array([i * dx, (i % 2) * dy, 0.0]) for i in range(n_carbon)]
    atoms = [(6, c) for c in carbons]
    for i, c in enumerate(carbons):
        # In-plane bisector points away from the chain; the two H lie above/below the plane.
        outward = np.array([0.0, -1.0 if i % 2 == 0 else 1.0, 0.0])
        hs = []
        for sign in (1.0, -1.0):
            v = outward * math.cos(TETRAHEDRAL / 2.0) + np.array([0.0, 0.0, sign * math.sin(TETRAHEDRAL / 2.0)])
            hs.append(c + ch * v / np.linalg.norm(v))
        if i == 0 or i == n_carbon - 1:
            axis = np.array([-1.0 if i == 0 else 1.0, 0.0, 0.0])
            tip = axis * math.sin(half) + (-outward) * math.cos(half)
            hs.append(c + ch * tip / np.linalg.norm(tip))
        atoms.extend((1, h) for h in hs)
    return tuple(Atom(z, tuple(map(float, p * BOHR_PER_ANGSTROM))) for z, p in atoms)


def lowdin_occupied(S: np.ndarray, nocc: int, seed: int) -> np.ndarray:
    """Random N(0,1) nbf x nocc factor, orthonormalised in the S metric."""
    rng = np.random.default_rng(seed)
    C = rng.normal(size=(S.shape[0], nocc))
    M = C.T @ S @ C
    w, v = np.linalg.eigh(M)
    return C @ (v / np.sqrt(w)) @ v.T
