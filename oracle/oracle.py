"""FP64 CPU oracle for the density-fitted J/K Fock build and the RHF SCF.

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP product path is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to check or to
time a CPU baseline.  The product package (``vibeqc_amd``) never imports it.

Integrals come from ``df_oracle.c`` (plain C, McMurchie-Davidson, built into
``liboracle_df.so`` by ``oracle/Makefile``); the DF algebra and the SCF are
restated in numpy.  Every function cites the reference file:line it follows.

Pinning (each function is checked by tests/test_oracle_*.py against):
  * normalise / spherical transforms / integrals: the libcint known answers in
    tests/native/test_density_fitting.cpp:468-499 and the PySCF fixtures
    tests/reference_data/posthf/*.npz (raw_three_center, metric, df_S, df_h);
    closed forms for (s|s) and (ss|s) with F0 = 1/2 sqrt(pi/T) erf(sqrt T);
    the overlap diagonal equal to one;
  * metric_inverse_sqrt: X M X = I and the rank-deficient {2,2,2,2} case
    (test_density_fitting.cpp:574-597);
  * jk / fock: a brute-force four-index RI reconstruction
    (test_density_fitting.cpp:47-76) and the PySCF df_F - df_h = J - K/2 at
    the fixture density;
  * rhf_scf: the PySCF DF-RHF energies in posthf/*.json and the GPU4PySCF
    cc-pVDZ/cc-pVDZ-JKFIT energies in
    benchmarks/results/issue409-packed-values/references/*-reference.json;
  * df_rhf_gradient: central finite differences of df_rhf energies, the
    translational invariance sum_A dE/dR_A = 0, and the GPU4PySCF forces in
    the same issue409 reference JSONs.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle_df.so"
_lib = None

BOHR_PER_ANGSTROM = 1.0 / 0.529177210903


class _CBasis(ctypes.Structure):
    _fields_ = [
        ("centers", ctypes.POINTER(ctypes.c_double)),
        ("l", ctypes.POINTER(ctypes.c_int)),
        ("prim_offset", ctypes.POINTER(ctypes.c_int)),
        ("prim_count", ctypes.POINTER(ctypes.c_int)),
        ("exps", ctypes.POINTER(ctypes.c_double)),
        ("coefs", ctypes.POINTER(ctypes.c_double)),
        ("nshell", ctypes.c_int),
    ]


def library():
    """Load liboracle_df.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        lib = ctypes.CDLL(str(_LIB_PATH))
        b = ctypes.POINTER(_CBasis)
        d = ctypes.POINTER(ctypes.c_double)
        lib.oracle_three_center.argtypes = [b, b, d]
        lib.oracle_metric.argtypes = [b, d]
        lib.oracle_one_electron.argtypes = [b, ctypes.c_int, d, d, d, d, d]
        lib.oracle_boys.argtypes = [ctypes.c_int, ctypes.c_double, d]
        lib.oracle_boys.restype = None
        i = ctypes.POINTER(ctypes.c_int)
        lib.oracle_grad_three_center.argtypes = [b, b, i, i, d, ctypes.c_int, d]
        lib.oracle_grad_metric.argtypes = [b, i, d, ctypes.c_int, d]
        lib.oracle_grad_one_electron.argtypes = [b, ctypes.c_int, d, d, i, d, d, d]
        _lib = lib
    return _lib


# ----------------------------------------------------------------------------
# Basis handling (src/molecule/basis.cpp; python/generativeqc/basis.py)
# ----------------------------------------------------------------------------


@dataclass(frozen=True)
class Shell:
    atom_index: int
    l: int
    exponents: tuple
    coefficients: tuple


def ncart(l: int) -> int:
    return (l + 1) * (l + 2) // 2


def cart_components(l: int):
    """CCA order: lx descending, then lz ascending (basis.cpp:33-45)."""
    out = []
    for lx in range(l, -1, -1):
        for lz in range(0, l - lx + 1):
            out.append((lx, l - lx - lz, lz))
    return out


def _odd_df(k: int) -> float:
    v = 1.0
    for f in range(1, 2 * k, 2):
        v *= f
    return v


def normalise_shell(shell: Shell):
    """Fold (2a/pi)^{3/4} (4a)^{l/2} into each coefficient and renormalise the
    contraction (basis.cpp:13-22, 254-282)."""
    l = shell.l
    e = np.asarray(shell.exponents, dtype=np.float64)
    c = np.asarray(shell.coefficients, dtype=np.float64)
    ratio = 2.0 * np.sqrt(np.outer(e, e)) / np.add.outer(e, e)
    norm2 = float(c @ (ratio ** (l + 1.5)) @ c)
    if not norm2 > 0.0:
        raise ValueError("contracted shell has an invalid normalisation")
    radial = (2.0 * e / math.pi) ** 0.75 * (4.0 * e) ** (0.5 * l)
    return e, c * radial / math.sqrt(norm2)


def spherical_rows(l: int) -> np.ndarray:
    """Rows = real-spherical AOs (libcint m = -l..l order), columns = individually
    normalised Cartesian AOs in CCA order (basis.cpp:50-175)."""
    cart = cart_components(l)
    idx = {c: i for i, c in enumerate(cart)}
    out = np.zeros((2 * l + 1, len(cart)))
    if l < 2:
        return np.eye(len(cart))
    if l == 2:
        r3 = math.sqrt(3.0) / 2.0
        out[0, idx[(1, 1, 0)]] = 1.0
        out[1, idx[(0, 1, 1)]] = 1.0
        out[2, idx[(2, 0, 0)]] = -0.5
        out[2, idx[(0, 2, 0)]] = -0.5
        out[2, idx[(0, 0, 2)]] = 1.0
        out[3, idx[(1, 0, 1)]] = 1.0
        out[4, idx[(2, 0, 0)]] = r3
        out[4, idx[(0, 2, 0)]] = -r3
        return out
    if l == 3:
        a = 3.0 / math.sqrt(8.0)
        b = math.sqrt(5.0 / 8.0)
        c = math.sqrt(3.0 / 40.0)
        d = math.sqrt(3.0 / 8.0)
        e = math.sqrt(6.0 / 5.0)
        f = 3.0 / math.sqrt(20.0)
        g = math.sqrt(3.0) / 2.0
        rows = [
            {(2, 1, 0): a, (0, 3, 0): -b},
            {(1, 1, 1): 1.0},
            {(2, 1, 0): -c, (0, 3, 0): -d, (0, 1, 2): e},
            {(2, 0, 1): -f, (0, 2, 1): -f, (0, 0, 3): 1.0},
            {(3, 0, 0): -d, (1, 2, 0): -c, (1, 0, 2): e},
            {(2, 0, 1): g, (0, 2, 1): -g},
            {(3, 0, 0): b, (1, 2, 0): -a},
        ]
        for m, row in enumerate(rows):
            for comp, v in row.items():
                out[m, idx[comp]] = v
        return out
    # l >= 4: real regular solid harmonics without Condon-Shortley phase,
    # normalised in the Cartesian Gaussian metric (basis.cpp:97-152).
    fact = math.factorial
    for row, sm in enumerate(range(-l, l + 1)):
        m = abs(sm)
        poly = np.zeros(len(cart))
        for k in range(0, (l - m) // 2 + 1):
            leg = (-1.0) ** k * fact(2 * l - 2 * k) / (2.0**l * fact(k) * fact(l - k) * fact(l - m - 2 * k))
            for y in range(0, m + 1):
                if y % 2 != (1 if sm < 0 else 0):
                    continue
                phase = -1.0 if (y // 2) % 2 else 1.0
                cterm = phase * fact(m) / (fact(y) * fact(m - y))
                for rx in range(0, k + 1):
                    for ry in range(0, k - rx + 1):
                        rz = k - rx - ry
                        comp = (m - y + 2 * rx, y + 2 * ry, l - m - 2 * k + 2 * rz)
                        poly[idx[comp]] += leg * cterm * fact(k) / (fact(rx) * fact(ry) * fact(rz))
        norm = 0.0
        for i, ci in enumerate(cart):
            for j, cj in enumerate(cart):
                mom = 1.0
                for ax in range(3):
                    pw = ci[ax] + cj[ax]
                    mom *= 0.0 if pw % 2 else _odd_df(pw // 2)
                norm += poly[i] * poly[j] * mom
        for i, ci in enumerate(cart):
            if poly[i] != 0.0:
                cn = 1.0 / math.sqrt(_odd_df(ci[0]) * _odd_df(ci[1]) * _odd_df(ci[2]))
                out[row, i] = poly[i] / (math.sqrt(norm) * cn)
    return out


def transform_matrix(shells, spherical: bool) -> np.ndarray:
    """Public-AO x Cartesian-AO block-diagonal transform."""
    nc = sum(ncart(s.l) for s in shells)
    rows = []
    col = 0
    for s in shells:
        t = spherical_rows(s.l) if spherical else np.eye(ncart(s.l))
        blk = np.zeros((t.shape[0], nc))
        blk[:, col : col + ncart(s.l)] = t
        rows.append(blk)
        col += ncart(s.l)
    return np.vstack(rows)


class _Packed:
    """Flat C-layout view of a shell list for df_oracle.c."""

    def __init__(self, atoms_xyz: np.ndarray, shells):
        self.centers = np.ascontiguousarray([atoms_xyz[s.atom_index] for s in shells], dtype=np.float64).reshape(-1)
        self.l = np.array([s.l for s in shells], dtype=np.int32)
        exps, coefs, off, cnt = [], [], [], []
        for s in shells:
            e, c = normalise_shell(s)
            off.append(len(exps))
            cnt.append(len(e))
            exps.extend(e.tolist())
            coefs.extend(c.tolist())
        self.off = np.array(off, dtype=np.int32)
        self.cnt = np.array(cnt, dtype=np.int32)
        self.exps = np.array(exps, dtype=np.float64)
        self.coefs = np.array(coefs, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        self.c = _CBasis(
            self.centers.ctypes.data_as(dp),
            self.l.ctypes.data_as(ip),
            self.off.ctypes.data_as(ip),
            self.cnt.ctypes.data_as(ip),
            self.exps.ctypes.data_as(dp),
            self.coefs.ctypes.data_as(dp),
            len(shells),
        )


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ----------------------------------------------------------------------------
# Integrals (src/integrals/s_integrals.cpp)
# ----------------------------------------------------------------------------


def boys(mmax: int, t: float) -> np.ndarray:
    f = np.zeros(mmax + 1)
    library().oracle_boys(mmax, float(t), _dptr(f))
    return f


def one_electron(atoms_z, atoms_xyz, shells, spherical: bool):
    """S, T, V in public AO order (s_integrals.cpp:1865-2060)."""
    xyz = np.ascontiguousarray(atoms_xyz, dtype=np.float64)
    pk = _Packed(xyz, shells)
    nc = sum(ncart(s.l) for s in shells)
    S = np.zeros((nc, nc))
    T = np.zeros((nc, nc))
    V = np.zeros((nc, nc))
    z = np.ascontiguousarray(atoms_z, dtype=np.float64)
    rc = library().oracle_one_electron(ctypes.byref(pk.c), len(z), _dptr(z), _dptr(xyz), _dptr(S), _dptr(T), _dptr(V))
    if rc:
        raise ValueError("angular momentum above the oracle limit")
    U = transform_matrix(shells, spherical)
    return U @ S @ U.T, U @ T @ U.T, U @ V @ U.T


def three_center(atoms_xyz, shells, spherical: bool, aux_shells, aux_spherical: bool) -> np.ndarray:
    """Raw (mu nu|P) as an (N, N, naux) array (s_integrals.cpp:1201-1490)."""
    xyz = np.ascontiguousarray(atoms_xyz, dtype=np.float64)
    po = _Packed(xyz, shells)
    pa = _Packed(xyz, aux_shells)
    nc = sum(ncart(s.l) for s in shells)
    na = sum(ncart(s.l) for s in aux_shells)
    out = np.zeros((nc, nc, na))
    if library().oracle_three_center(ctypes.byref(po.c), ctypes.byref(pa.c), _dptr(out)):
        raise ValueError("angular momentum above the oracle limit")
    U = transform_matrix(shells, spherical)
    W = transform_matrix(aux_shells, aux_spherical)
    out = np.einsum("ai,ijk->ajk", U, out, optimize=True)
    out = np.einsum("bj,ajk->abk", U, out, optimize=True)
    return np.einsum("pk,abk->abp", W, out, optimize=True)


def metric(atoms_xyz, aux_shells, aux_spherical: bool) -> np.ndarray:
    """(P|Q) in public auxiliary order."""
    xyz = np.ascontiguousarray(atoms_xyz, dtype=np.float64)
    pa = _Packed(xyz, aux_shells)
    na = sum(ncart(s.l) for s in aux_shells)
    out = np.zeros((na, na))
    if library().oracle_metric(ctypes.byref(pa.c), _dptr(out)):
        raise ValueError("angular momentum above the oracle limit")
    W = transform_matrix(aux_shells, aux_spherical)
    return W @ out @ W.T


def nuclear_repulsion(atoms_z, atoms_xyz) -> float:
    z = np.asarray(atoms_z, dtype=np.float64)
    x = np.asarray(atoms_xyz, dtype=np.float64)
    e = 0.0
    for i in range(len(z)):
        for j in range(i):
            e += z[i] * z[j] / np.linalg.norm(x[i] - x[j])
    return float(e)


# ----------------------------------------------------------------------------
# DF algebra (src/integrals/density_fitting_metric.cpp, src/scf/density_fitting.cpp)
# ----------------------------------------------------------------------------


def metric_inverse_sqrt(M: np.ndarray, relative_threshold: float = 1e-10):
    """Symmetrise, eigh, keep lambda > thr * lambda_max, X = V diag(lambda^-1/2) V^T.
    Returns (X, effective_rank, condition_number)
    (density_fitting_metric.cpp:101-152)."""
    Ms = 0.5 * (M + M.T)
    w, v = np.linalg.eigh(Ms)
    largest = w[-1]
    if not largest > 0.0:
        raise ValueError("Coulomb metric has no positive eigenspace")
    keep = w > relative_threshold * largest
    if not keep.any():
        raise ValueError("metric threshold removed every auxiliary direction")
    vk = v[:, keep]
    X = (vk / np.sqrt(w[keep])) @ vk.T
    return X, int(keep.sum()), float(largest / w[keep].min())


def df_tensor(A: np.ndarray, X: np.ndarray) -> np.ndarray:
    """B = A X over the auxiliary index (density_fitting.cpp:711-764)."""
    n = A.shape[0]
    return (A.reshape(n * n, -1) @ X).reshape(n, n, -1)


def jk(B: np.ndarray, D: np.ndarray):
    """Raw J and K: rho_Q = sum B_{mn,Q} D_mn, J = B rho; K_mn = sum_Q (B_Q D B_Q)_mn
    (density_fitting.cpp:101-224)."""
    n = B.shape[0]
    Bf = B.reshape(n * n, -1)
    rho = D.reshape(-1) @ Bf
    J = (Bf @ rho).reshape(n, n)
    T = np.einsum("mlq,ls->msq", B, D, optimize=True)  # (B_Q D)_{m s}
    K = np.einsum("msq,snq->mn", T, B, optimize=True)
    return J, K


def jk_occupied(B: np.ndarray, C: np.ndarray):
    """J/K from occupied factors: D = C C^T, U_{mu,Q,i} = sum_nu B C, K = U U^T
    (df_occupied_exchange.cpp:236-300)."""
    n = B.shape[0]
    naux = B.shape[2]
    D = C @ C.T
    Bf = B.reshape(n * n, naux)
    rho = D.reshape(-1) @ Bf
    J = (Bf @ rho).reshape(n, n)
    U = np.einsum("mnq,ni->mqi", B, C, optimize=True).reshape(n, -1)
    return J, U @ U.T


def brute_force_jk(B: np.ndarray, D: np.ndarray):
    """Four-index RI reconstruction used as an independent J/K check
    (tests/native/test_density_fitting.cpp:47-76)."""
    eri = np.einsum("ijq,klq->ijkl", B, B)
    J = np.einsum("ijkl,kl->ij", eri, D)
    K = np.einsum("ikjl,kl->ij", eri, D)
    return J, K


def fock(H, J, K, cj=1.0, ck=-0.5):
    """F = H + cJ J + cK K (fock_build.cpp:429-458)."""
    return H + cj * J + ck * K


def two_electron_energy(D, J, K, cj=1.0, ck=-0.5) -> float:
    """E2 = 1/2 sum D (cJ J + cK K) (fock_build.cpp:460-491)."""
    return float(0.5 * np.sum(D * (cj * J)) + 0.5 * np.sum(D * (ck * K)))


# ----------------------------------------------------------------------------
# SCF (src/scf/solver/mean_field_driver.cpp:233-345, diis.cpp:17-46,
#      solver/diis_coefficients.hpp, solver/self_consistent.hpp:59-96)
# ----------------------------------------------------------------------------


def symmetric_orthogonalizer(S: np.ndarray) -> np.ndarray:
    """X = S^{-1/2}; rejects eigenvalues below 1e-10 (reference/linalg.cpp:123-142)."""
    w, v = np.linalg.eigh(S)
    if w.min() < 1e-10:
        raise ValueError("overlap matrix is singular or severely linearly dependent")
    return (v / np.sqrt(w)) @ v.T


def generalized_eigen(F: np.ndarray, X: np.ndarray):
    """Solve X^T F X C' = C' e, C = X C' (reference/linalg.cpp:144-151)."""
    e, c = np.linalg.eigh(X.T @ F @ X)
    return e, X @ c


class Diis:
    """Pulay DIIS on FDS - SDF residuals, no metric scaling, no retirement
    (diis.cpp:17-46; diis_coefficients.hpp)."""

    def __init__(self, capacity: int = 8):
        self.capacity = capacity
        self.focks: list = []
        self.errors: list = []

    def update(self, F, r):
        if self.capacity < 2:
            return F
        self.focks.append(F.copy())
        self.errors.append(r.copy())
        if len(self.focks) > self.capacity:
            self.focks.pop(0)
            self.errors.pop(0)
        m = len(self.focks)
        if m < 2:
            return F
        G = np.array([[np.sum(a * b) for b in self.errors] for a in self.errors])
        A = np.zeros((m + 1, m + 1))
        A[:m, :m] = G
        A[:m, m] = -1.0
        A[m, :m] = -1.0
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        try:
            c = np.linalg.solve(A, rhs)
        except np.linalg.LinAlgError:
            return F
        if not np.all(np.isfinite(c)):
            return F
        return sum(ci * Fi for ci, Fi in zip(c[:m], self.focks))


@dataclass
class ScfResult:
    energy: float
    density: np.ndarray
    iterations: int
    converged: bool
    energy_change: float
    density_rms: float
    orbital_energies: np.ndarray
    coefficients: np.ndarray
    fock_builds: int


def rhf_scf(
    S,
    H,
    enuc: float,
    nocc: int,
    jk_builder,
    max_iterations: int = 100,
    diis_history: int = 8,
    energy_tolerance: float = 1e-10,
    density_tolerance: float = 1e-8,
    initial_density=None,
) -> ScfResult:
    """Host RHF loop: core-Hamiltonian guess, DIIS(FDS-SDF), converged when
    iteration > 1 and |dE| < e_tol and RMS(D_new - D) < d_tol; then a final
    rebuild/diagonalise/rebuild (mean_field_driver.cpp:181-199, 233-345)."""
    n = S.shape[0]
    X = symmetric_orthogonalizer(S)
    if initial_density is None:
        _, c0 = generalized_eigen(H, X)
        D = 2.0 * c0[:, :nocc] @ c0[:, :nocc].T
    else:
        D = np.array(initial_density, dtype=np.float64)
    diis = Diis(diis_history)
    prev = math.inf
    builds = 0
    converged = False
    e = de = rms = math.nan
    for it in range(1, max_iterations + 1):
        J, K = jk_builder(D)
        builds += 1
        F = fock(H, J, K)
        e = float(0.5 * np.sum(D * (H + F)) + enuc)
        r = F @ D @ S - S @ D @ F
        Fe = diis.update(F, r)
        eps, C = generalized_eigen(Fe, X)
        Dn = 2.0 * C[:, :nocc] @ C[:, :nocc].T
        rms = float(np.sqrt(np.mean((Dn - D) ** 2)))
        de = abs(e - prev) if math.isfinite(prev) else math.inf
        D = Dn
        if it > 1 and de < energy_tolerance and rms < density_tolerance:
            converged = True
            break
        prev = e
    if not converged:
        return ScfResult(e, D, it, False, de, rms, np.array([]), np.array([]), builds)
    J, K = jk_builder(D)
    F = fock(H, J, K)
    eps, C = generalized_eigen(F, X)
    D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    J, K = jk_builder(D)
    F = fock(H, J, K)
    builds += 2
    energy = float(0.5 * np.sum(D * (H + F)) + enuc)
    return ScfResult(energy, D, it, True, de, rms, eps, C, builds)


# ----------------------------------------------------------------------------
# Convenience: a whole DF-RHF problem from atoms + basis shells
# ----------------------------------------------------------------------------


@dataclass
class DfProblem:
    S: np.ndarray
    H: np.ndarray
    A: np.ndarray
    M: np.ndarray
    X: np.ndarray
    B: np.ndarray
    enuc: float
    nocc: int
    rank: int
    condition: float


def build_problem(atoms_z, atoms_xyz, shells, spherical, aux_shells, aux_spherical, charge=0, threshold=1e-10):
    S, T, V = one_electron(atoms_z, atoms_xyz, shells, spherical)
    A = three_center(atoms_xyz, shells, spherical, aux_shells, aux_spherical)
    M = metric(atoms_xyz, aux_shells, aux_spherical)
    X, rank, cond = metric_inverse_sqrt(M, threshold)
    B = df_tensor(A, X)
    nelec = int(sum(atoms_z)) - charge
    return DfProblem(S, T + V, A, M, X, B, nuclear_repulsion(atoms_z, atoms_xyz), nelec // 2, rank, cond)


def df_rhf(problem: DfProblem, **kw) -> ScfResult:
    return rhf_scf(problem.S, problem.H, problem.enuc, problem.nocc, lambda D: jk(problem.B, D), **kw)


# ----------------------------------------------------------------------------
# Unrestricted HF (density_fitting.cpp:780-801; rhf.cpp:451-463;
# mean_field_driver.cpp:201-227, 346-488; reference/mean_field.cpp:26-35;
# initial_guess/density.cpp:602-623)
# ----------------------------------------------------------------------------


def spin_occupations(electrons: int, multiplicity: int):
    """(n_alpha, n_beta) from the electron count and multiplicity."""
    excess = multiplicity - 1
    if excess < 0 or excess > electrons or (electrons + excess) % 2:
        raise ValueError("electron count and multiplicity do not define integral UHF occupations")
    return (electrons + excess) // 2, (electrons - excess) // 2


def uhf_jk(B: np.ndarray, Da: np.ndarray, Db: np.ndarray):
    """Raw J of Da + Db and K of each spin density (density_fitting.cpp:780-801)."""
    J, Ka = jk(B, Da + Db)
    _, Ka = jk(B, Da)
    _, Kb = jk(B, Db)
    return J, Ka, Kb


def uhf_energy(Da, Db, H, Fa, Fb) -> float:
    """1/2 sum_s D_s (H + F_s), without E_nuc (reference/mean_field.cpp:26-35)."""
    return float(0.5 * np.sum(Da * (H + Fa)) + 0.5 * np.sum(Db * (H + Fb)))


@dataclass
class UhfResult:
    energy: float
    density: np.ndarray  # (2, n, n): alpha, beta
    iterations: int
    converged: bool
    energy_change: float
    density_rms: float
    orbital_energies: np.ndarray  # (2, n)
    coefficients: np.ndarray  # (2, n, n)
    fock_builds: int


def uhf_scf(
    S,
    H,
    enuc: float,
    nalpha: int,
    nbeta: int,
    jk_builder,
    max_iterations: int = 100,
    diis_history: int = 8,
    energy_tolerance: float = 1e-10,
    density_tolerance: float = 1e-8,
    initial_density=None,
) -> UhfResult:
    """Host UHF loop: core guess with the beta HOMO/LUMO pair rotated by 45
    degrees when the spins differ, one DIIS over stacked (F_a, F_b) and
    (F_a D_a S - S D_a F_a, F_b D_b S - S D_b F_b), the density change as the
    RMS over both blocks, then rebuild / diagonalise / rebuild.
    jk_builder(Da, Db) returns (J, Ka, Kb)."""
    n = S.shape[0]
    X = symmetric_orthogonalizer(S)
    occ = (nalpha, nbeta)
    if initial_density is None:
        _, ca = generalized_eigen(H, X)
        cb = ca.copy()
        if nalpha != nbeta and 0 < nbeta < n:
            c = s = 0.7071067811865476
            i, j = nbeta - 1, nbeta
            a, b = cb[:, i].copy(), cb[:, j].copy()
            cb[:, i], cb[:, j] = c * a + s * b, -s * a + c * b
        D = np.stack([ca[:, :nalpha] @ ca[:, :nalpha].T, cb[:, :nbeta] @ cb[:, :nbeta].T])
    else:
        D = np.array(initial_density, dtype=np.float64).reshape(2, n, n)
    diis = Diis(diis_history)
    prev = math.inf
    builds = 0
    converged = False
    e = de = rms = math.nan

    def focks(Dm):
        J, Ka, Kb = jk_builder(Dm[0], Dm[1])
        return np.stack([H + J - Ka, H + J - Kb])

    def densities(Fm):
        eps, C = [], []
        for s in range(2):
            es, cs = generalized_eigen(Fm[s], X)
            eps.append(es)
            C.append(cs)
        return np.stack([C[s][:, : occ[s]] @ C[s][:, : occ[s]].T for s in range(2)]), np.stack(eps), np.stack(C)

    for it in range(1, max_iterations + 1):
        F = focks(D)
        builds += 2
        e = uhf_energy(D[0], D[1], H, F[0], F[1]) + enuc
        r = np.stack([F[s] @ D[s] @ S - S @ D[s] @ F[s] for s in range(2)])
        Fe = diis.update(F, r)
        Dn, _, _ = densities(Fe)
        rms = float(np.sqrt(np.mean((Dn - D) ** 2)))
        de = abs(e - prev) if math.isfinite(prev) else math.inf
        D = Dn
        if it > 1 and de < energy_tolerance and rms < density_tolerance:
            converged = True
            break
        prev = e
    if not converged:
        return UhfResult(e, D, it, False, de, rms, np.array([]), np.array([]), builds)
    D, eps, C = densities(focks(D))
    F = focks(D)
    builds += 4
    energy = uhf_energy(D[0], D[1], H, F[0], F[1]) + enuc
    return UhfResult(energy, D, it, True, de, rms, eps, C, builds)


def df_uhf(problem: DfProblem, nalpha: int, nbeta: int, **kw) -> UhfResult:
    return uhf_scf(problem.S, problem.H, problem.enuc, nalpha, nbeta, lambda a, b: uhf_jk(problem.B, a, b), **kw)


# ----------------------------------------------------------------------------
# Nuclear gradient of the DF-RHF energy (density_fitting.cpp:898-917, 464-595;
# density_fitting_metric.cpp:50-99; hf_gradient.cpp:8-18; mean_field.cpp:18-24)
# ----------------------------------------------------------------------------


def _atom_of(shells):
    return np.array([s.atom_index for s in shells], dtype=np.int32)


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def grad_three_center(atoms_xyz, natom, shells, spherical, aux_shells, aux_spherical, w) -> np.ndarray:
    """sum_{mu nu P} w[mu, nu, P] d(mu nu|P)/dR, (natom, 3); w in public AOs."""
    xyz = np.ascontiguousarray(atoms_xyz, dtype=np.float64)
    po, pa = _Packed(xyz, shells), _Packed(xyz, aux_shells)
    U = transform_matrix(shells, spherical)
    Wx = transform_matrix(aux_shells, aux_spherical)
    wc = np.einsum("mnp,ma->anp", w, U, optimize=True)
    wc = np.einsum("anp,nb->abp", wc, U, optimize=True)
    wc = np.ascontiguousarray(np.einsum("abp,pk->abk", wc, Wx, optimize=True))
    g = np.zeros((natom, 3))
    ao, aa = _atom_of(shells), _atom_of(aux_shells)
    if library().oracle_grad_three_center(ctypes.byref(po.c), ctypes.byref(pa.c), _iptr(ao), _iptr(aa), _dptr(wc), natom, _dptr(g)):
        raise ValueError("angular momentum above the oracle limit")
    return g


def grad_metric(atoms_xyz, natom, aux_shells, aux_spherical, w) -> np.ndarray:
    """sum_{PQ} w[P, Q] d(P|Q)/dR, (natom, 3); w in public auxiliary order."""
    xyz = np.ascontiguousarray(atoms_xyz, dtype=np.float64)
    pa = _Packed(xyz, aux_shells)
    Wx = transform_matrix(aux_shells, aux_spherical)
    wc = np.ascontiguousarray(Wx.T @ w @ Wx)
    g = np.zeros((natom, 3))
    aa = _atom_of(aux_shells)
    if library().oracle_grad_metric(ctypes.byref(pa.c), _iptr(aa), _dptr(wc), natom, _dptr(g)):
        raise ValueError("angular momentum above the oracle limit")
    return g


def grad_one_electron(atoms_z, atoms_xyz, shells, spherical, D, W) -> np.ndarray:
    """sum D d(T + V)/dR - sum W dS/dR, (natom, 3) (hf_gradient.cpp:8-18)."""
    xyz = np.ascontiguousarray(atoms_xyz, dtype=np.float64)
    z = np.ascontiguousarray(atoms_z, dtype=np.float64)
    pk = _Packed(xyz, shells)
    U = transform_matrix(shells, spherical)
    dc = np.ascontiguousarray(U.T @ D @ U)
    wc = np.ascontiguousarray(U.T @ W @ U)
    g = np.zeros((len(z), 3))
    ao = _atom_of(shells)
    if library().oracle_grad_one_electron(ctypes.byref(pk.c), len(z), _dptr(z), _dptr(xyz), _iptr(ao), _dptr(dc), _dptr(wc), _dptr(g)):
        raise ValueError("angular momentum above the oracle limit")
    return g


def nuclear_repulsion_gradient(atoms_z, atoms_xyz) -> np.ndarray:
    z = np.asarray(atoms_z, dtype=np.float64)
    x = np.asarray(atoms_xyz, dtype=np.float64)
    d = x[:, None, :] - x[None, :, :]
    r = np.linalg.norm(d, axis=-1)
    np.fill_diagonal(r, np.inf)
    return -np.einsum("i,j,ijx->ix", z, z, d / r[..., None] ** 3)


def df_weights(A, M, D, C_occ):
    """Response weights of the DF two-electron energy for a full-rank metric
    (density_fitting.cpp:464-595): dE2 = sum W_A dA + sum W_M dM with
      c = M^-1 rho, rho_P = sum D A_P,   T_Pij = sum C_mi C_nj A_mnP,
      Y = M^-1 T,
      W_A = D c_P - 2 sum_ij C_mi C_nj Y_Pij,
      W_M = -1/2 c c^T + sum_ij Y_Pij Y_Qij
    (E_J = 1/2 rho^T M^-1 rho and E_K = -sum_ij T_ij^T M^-1 T_ij for
    D = 2 C_occ C_occ^T)."""
    Minv = np.linalg.inv(M)
    rho = np.einsum("mn,mnp->p", D, A, optimize=True)
    c = Minv @ rho
    T = np.einsum("mi,nj,mnp->pij", C_occ, C_occ, A, optimize=True)
    Y = np.einsum("pq,qij->pij", Minv, T, optimize=True)
    WA = np.einsum("mn,p->mnp", D, c) - 2.0 * np.einsum("mi,nj,pij->mnp", C_occ, C_occ, Y, optimize=True)
    WM = -0.5 * np.outer(c, c) + np.einsum("pij,qij->pq", Y, Y, optimize=True)
    return WA, WM


def df_weights_dense(A, M, D, cj=1.0, ck=-0.5):
    """Response weights of E2 = 1/2 sum D (cJ J + cK K) at a fixed symmetric
    (possibly indefinite) density, without a factor of D:
      At_P = sum_Q (M^-1)_PQ A_Q,   c = M^-1 sum D A,
      W_A[P] = cJ D c_P + cK D At_P D,
      W_M[P, Q] = -cJ/2 c_P c_Q - cK/2 tr(D At_P D At_Q)."""
    Minv = np.linalg.inv(M)
    At = np.einsum("mnq,pq->mnp", A, Minv, optimize=True)
    c = np.einsum("mn,mnp->p", D, At, optimize=True)
    DAD = np.einsum("ml,lsp,sn->mnp", D, At, D, optimize=True)
    WA = cj * np.einsum("mn,p->mnp", D, c) + ck * DAD
    WM = -0.5 * cj * np.outer(c, c) - 0.5 * ck * np.einsum("mnp,mnq->pq", DAD, At, optimize=True)
    return WA, WM


def df_two_electron_gradient(atoms_xyz, natom, shells, spherical, aux_shells, aux_spherical, A, M, D, cj=1.0, ck=-0.5):
    """Fixed-density dE2/dR (natom, 3) of 1/2 sum D (cJ J + cK K): the
    gradient the plan's evaluate(derivative=True) returns (fock.h:73-80)."""
    WA, WM = df_weights_dense(A, M, D, cj, ck)
    return grad_three_center(atoms_xyz, natom, shells, spherical, aux_shells, aux_spherical, WA) + grad_metric(
        atoms_xyz, natom, aux_shells, aux_spherical, WM
    )


def df_rhf_gradient(atoms_z, atoms_xyz, shells, spherical, aux_shells, aux_spherical, problem: DfProblem, scf: ScfResult):
    """dE/dR (natom, 3) of a converged DF-RHF energy; forces are its negative
    (hf_gradient.cpp:8-18 with the DF two-electron response of
    density_fitting.cpp:898-917)."""
    if problem.rank != problem.M.shape[0]:
        raise ValueError("the gradient oracle assumes a full-rank metric")
    natom = len(atoms_z)
    Co = scf.coefficients[:, : problem.nocc]
    eps = scf.orbital_energies[: problem.nocc]
    D = 2.0 * Co @ Co.T
    W = 2.0 * (Co * eps) @ Co.T
    WA, WM = df_weights(problem.A, problem.M, D, Co)
    g = nuclear_repulsion_gradient(atoms_z, atoms_xyz)
    g += grad_one_electron(atoms_z, atoms_xyz, shells, spherical, D, W)
    g += grad_three_center(atoms_xyz, natom, shells, spherical, aux_shells, aux_spherical, WA)
    g += grad_metric(atoms_xyz, natom, aux_shells, aux_spherical, WM)
    return g


def df_uhf_gradient(atoms_z, atoms_xyz, shells, spherical, aux_shells, aux_spherical, problem: DfProblem, scf: UhfResult, occ):
    """dE/dR (natom, 3) of a converged DF-UHF energy (analytic_uhf_forces:
    one-electron terms of D_a + D_b and of E = sum_s C_s diag(e_s) C_s^T; the
    DF response of E2 = 1/2 (D_a + D_b) J - 1/2 sum_s D_s K_s)."""
    if problem.rank != problem.M.shape[0]:
        raise ValueError("the gradient oracle assumes a full-rank metric")
    natom = len(atoms_z)
    D = scf.density
    W = sum((scf.coefficients[s][:, : occ[s]] * scf.orbital_energies[s][: occ[s]]) @ scf.coefficients[s][:, : occ[s]].T for s in range(2))
    WA, WM = df_weights_dense(problem.A, problem.M, D[0] + D[1], 1.0, 0.0)
    for s in range(2):
        wa, wm = df_weights_dense(problem.A, problem.M, D[s], 0.0, -1.0)
        WA = WA + wa
        WM = WM + wm
    g = nuclear_repulsion_gradient(atoms_z, atoms_xyz)
    g += grad_one_electron(atoms_z, atoms_xyz, shells, spherical, D[0] + D[1], W)
    g += grad_three_center(atoms_xyz, natom, shells, spherical, aux_shells, aux_spherical, WA)
    g += grad_metric(atoms_xyz, natom, aux_shells, aux_spherical, WM)
    return g
