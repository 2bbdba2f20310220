"""Bogoliubov (HFB) states U, V.

Layout everywhere: real n x n arrays, ``U[i, k]`` with i the single-particle
index (basis order of ``sccm_amd.basis``) and k the quasiparticle index; the
C-ABI takes them column-major (Fortran order), as the reference's module
arrays ``HFB%U0(NLEV,NLEV)`` are.

* ``read_wf`` reads the reference's ``.wf`` files
  (``Codes/PNAMP/PNAMP_Kernel3D.f90:90-97``: header ``ide(n) ide(p)``, then
  n*n lines ``U(i,k) V(i,k)`` with i outer, k inner).
* ``synthetic_state`` makes the seeded states of the benchmark workloads:
  W = expm([[X, Y], [Y, X]]) with X, Y real antisymmetric and
  species-block-diagonal, applied to a Slater reference with an even number
  of occupied m-states per species.  W is orthogonal with the Bogoliubov block
  structure, so U^T U + V^T V = 1 and V U^{-1} is antisymmetric.  X (entries
  N(0, 0.3)) rotates the single-particle basis and keeps the particle number;
  Y (pairing, entries N(0, 0.05) by default) mixes particle numbers.  With Y
  as large as X the quasiparticle vacuum is half filled (<Z> ~ 10 for a
  Z = 2 pf state) and its Z = 2 component is ~1e-9; with 0.05 the states
  carry <Z> ~ 2.6, <N> ~ 4.5 for Z, N = 2, 4, like a constrained HFB state.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.linalg import expm

from .basis import Basis


def read_wf(path: str, n: int) -> tuple[np.ndarray, np.ndarray, tuple[int, int]]:
    with open(path) as f:
        head = f.readline().split()
        data = np.loadtxt(f)
    if data.shape != (n * n, 2):
        raise ValueError(f"{path}: expected {n*n} (U,V) lines, got {data.shape}")
    U = data[:, 0].reshape(n, n)
    V = data[:, 1].reshape(n, n)
    return U, V, (int(head[0]), int(head[1]))


def slater_occupations(b: Basis, zval: int, nval: int) -> np.ndarray:
    """Deterministic even occupation per species: time-reversed pairs
    (m, -m), lowest |m| first, orbits in .val order."""
    if zval % 2 or nval % 2:
        raise ValueError("valence numbers must be even")
    occ = np.zeros(b.n)
    for species, count in ((-1, zval), (1, nval)):
        idx = [k for k in range(b.n) if b.t[k] == species]
        order = sorted(idx, key=lambda k: (int(b.orbit[k]), abs(int(b.twom[k])), -int(b.twom[k])))
        if count > len(order):
            raise ValueError("too many particles for the space")
        for k in order[:count]:
            occ[k] = 1.0
    return occ


def _skew_block(b: Basis, rng: np.random.Generator, sigma: float) -> np.ndarray:
    n = b.n
    X = np.zeros((n, n))
    for species in (-1, 1):
        idx = np.nonzero(b.t == species)[0]
        m = idx.size
        a = rng.normal(0.0, sigma, size=(m, m))
        a = np.triu(a, 1)
        X[np.ix_(idx, idx)] = a - a.T
    return X


def ry_pi(b: Basis) -> np.ndarray:
    """Single-particle matrix of exp(-i pi J_y) in the m basis: inside each
    orbit m -> -m with the phase (-1)^(j - m) (a signed permutation; the
    phase convention is immaterial where it enters twice, as in P X P^T)."""
    n = b.n
    P = np.zeros((n, n))
    for k in range(n):
        for l in range(n):
            if b.t[l] == b.t[k] and b.orbit[l] == b.orbit[k] and b.twom[l] == -b.twom[k]:
                P[l, k] = -1.0 if ((int(b.twoj[k]) - int(b.twom[k])) // 2) % 2 else 1.0
    return P


def d2_symmetrize(b: Basis, X: np.ndarray, Y: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Projects a Bogoliubov generator onto its D2-invariant part: invariant
    under exp(-i pi J_z) (X_ij only for m_i - m_j even, Y_ij only for
    m_i + m_j even) and under exp(-i pi J_y) (X = P X P^T, Y = P Y P^T).
    With a real generator and a Slater reference of time-reversed pairs
    the state is real and D2 symmetric: the symmetries the reference's 3D
    mirror fold of the projection sums relies on
    (Codes/PNAMP/PNAMP_Kernel3D.f90:415-526), and the reason its GCM keeps
    only even K >= 0 (Codes/GCM/GCM_HWG.f90:166-170)."""
    P = ry_pi(b)
    X = 0.5 * (X + P @ X @ P.T)
    Y = 0.5 * (Y + P @ Y @ P.T)
    tm = b.twom.astype(int)
    X = X * ((((tm[:, None] - tm[None, :]) // 2) % 2) == 0)
    Y = Y * ((((tm[:, None] + tm[None, :]) // 2) % 2) == 0)
    return X, Y


def bogoliubov_from_generator(b: Basis, X: np.ndarray, Y: np.ndarray,
                              occ: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    n = b.n
    G = np.block([[X, Y], [Y, X]])
    W = expm(G)
    A, B = W[:n, :n], W[n:, :n]
    U0 = np.diag(1.0 - occ)
    V0 = np.diag(occ)
    U = A @ U0 + B @ V0
    V = B @ U0 + A @ V0
    return U, V


def synthetic_state(b: Basis, zval: int, nval: int, seed: int, sigma: float = 0.3,
                    sigma_pair: float = 0.05, beta2: float = 0.0, gamma_deg: float = 0.0, offset_seed: int | None = None,
                    offset_scale: float = 1.0, noise: float = 0.0, d2: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """One seeded state.  With ``offset_seed`` set, the generator is
    G(seed) + beta2 (cos(gamma) Gc + sin(gamma) Gs) * offset_scale, where Gc, Gs
    are fixed seeded generators: the GCM configurations of one workload share
    the base generator and differ by a smooth constraint offset."""
    rng = np.random.default_rng(seed)
    X = _skew_block(b, rng, sigma)
    Y = _skew_block(b, rng, sigma_pair)
    if offset_seed is not None:
        ro = np.random.default_rng(offset_seed)
        Xc, Yc = _skew_block(b, ro, sigma), _skew_block(b, ro, sigma_pair)
        Xs, Ys = _skew_block(b, ro, sigma), _skew_block(b, ro, sigma_pair)
        g = math.radians(gamma_deg)
        X = X + offset_scale * beta2 * (math.cos(g) * Xc + math.sin(g) * Xs)
        Y = Y + offset_scale * beta2 * (math.cos(g) * Yc + math.sin(g) * Ys)
    if noise > 0.0:
        rn = np.random.default_rng(seed + 7919)
        X = X + _skew_block(b, rn, noise)
        Y = Y + _skew_block(b, rn, noise * sigma_pair / sigma)
    if d2:
        X, Y = d2_symmetrize(b, X, Y)
    occ = slater_occupations(b, zval, nval)
    return bogoliubov_from_generator(b, X, Y, occ)


def check_state(U: np.ndarray, V: np.ndarray) -> tuple[float, float]:
    """(max |U^T U + V^T V - 1|, max |Z + Z^T|) with Z = V U^{-1}."""
    n = U.shape[0]
    unit = np.abs(U.T @ U + V.T @ V - np.eye(n)).max()
    Z = V @ np.linalg.inv(U)
    return float(unit), float(np.abs(Z + Z.T).max())
