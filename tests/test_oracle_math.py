"""Closed-form and textbook checks of the oracle's building blocks."""
import itertools

import numpy as np
import pytest

from sccm_amd.basis import make_basis
from sccm_amd.interaction import Interaction, cg
from sccm_amd.mesh import make_mesh


def _dummy_inter(b):
    z = np.zeros(0, dtype=np.int32)
    return Interaction(z, z, z, z, np.zeros(0), np.zeros((b.n, b.n)), 0.0)


def test_wigner_d_closed_forms(oracle_mod):
    d = oracle_mod.wigner_d
    for beta in (0.0, 0.3, 1.7, 3.0):
        c, s = np.cos(beta / 2), np.sin(beta / 2)
        assert abs(d(1, 1, 1, beta) - c) < 1e-15
        assert abs(d(1, 1, -1, beta) + s) < 1e-15
        assert abs(d(1, -1, 1, beta) - s) < 1e-15
        assert abs(d(2, 2, 2, beta) - (1 + np.cos(beta)) / 2) < 1e-15
        assert abs(d(2, 2, 0, beta) + np.sin(beta) / np.sqrt(2)) < 1e-15
        assert abs(d(2, 0, 0, beta) - np.cos(beta)) < 1e-15
        # d^J_00 = Legendre P_J(cos beta)
        for J in range(9):
            P = np.polynomial.legendre.Legendre.basis(J)(np.cos(beta))
            assert abs(d(2 * J, 0, 0, beta) - P) < 1e-12


def test_wigner_d_matches_exponential_of_jy(oracle_mod):
    """d^j(beta) = exp(-i beta J_y) in the |j m> basis, m descending."""
    from scipy.linalg import expm
    for tj in (1, 3, 5, 9, 16):
        j = tj / 2
        ms = np.arange(j, -j - 1, -1)
        Jp = np.zeros((len(ms), len(ms)))
        for c, m in enumerate(ms):
            if m < j:
                Jp[c - 1, c] = np.sqrt(j * (j + 1) - m * (m + 1))
        Jy = (Jp - Jp.T) / 2j
        for beta in (0.4, 2.2):
            D = expm(-1j * beta * Jy)
            mine = np.array([[oracle_mod.wigner_d(tj, int(2 * a), int(2 * b), beta) for b in ms] for a in ms])
            assert np.abs(D - mine).max() < 1e-12


def test_rotation_unitary_and_composition(oracle_mod):
    b = make_basis("f5pg9")
    o = oracle_mod.Oracle(b, _dummy_inter(b), make_mesh(1, 1, 1, 1, 0), 0, 0)
    R1 = o.rotation(0.3, 1.1, 2.0)
    assert np.abs(R1.conj().T @ R1 - np.eye(b.n)).max() < 1e-13
    # rotations about z compose additively; about y too
    A = o.rotation(0.0, 0.4, 0.0) @ o.rotation(0.0, 0.9, 0.0)
    assert np.abs(A - o.rotation(0.0, 1.3, 0.0)).max() < 1e-13
    Z = o.rotation(0.2, 0.0, 0.0) @ o.rotation(0.5, 0.0, 0.0)
    assert np.abs(Z - o.rotation(0.7, 0.0, 0.0)).max() < 1e-13
    # R(a,b,g) = Rz(a) Ry(b) Rz(g)
    C = o.rotation(0.3, 0.0, 0.0) @ o.rotation(0.0, 1.1, 0.0) @ o.rotation(2.0, 0.0, 0.0)
    assert np.abs(C - R1).max() < 1e-13


def _pf_brute(A):
    n = A.shape[0]
    if n == 0:
        return 1.0
    s = 0.0
    for j in range(1, n):
        idx = [k for k in range(n) if k not in (0, j)]
        s += (-1) ** (j - 1) * A[0, j] * _pf_brute(A[np.ix_(idx, idx)])
    return s


@pytest.mark.parametrize("N", [2, 4, 6, 8])
def test_pfaffian_bruteforce_and_det(oracle_mod, N):
    rng = np.random.default_rng(N)
    X = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    A = X - X.T
    pf = oracle_mod.pfaffian(A)
    assert abs(pf - _pf_brute(A)) < 1e-12 * max(1, abs(pf))
    assert abs(pf ** 2 - np.linalg.det(A)) < 1e-11 * abs(pf) ** 2


def test_pfaffian_large_det_and_blocks(oracle_mod):
    rng = np.random.default_rng(3)
    N = 40
    X = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    A = (X - X.T) / np.sqrt(N)
    pf = oracle_mod.pfaffian(A)
    sign, logdet = np.linalg.slogdet(A)
    assert abs(pf ** 2 / (sign * np.exp(logdet)) - 1) < 1e-10
    # block-diagonal skew: Pf = prod of 2x2 upper elements
    vals = rng.normal(size=N // 2) + 1j * rng.normal(size=N // 2)
    B = np.zeros((N, N), dtype=complex)
    for k, v in enumerate(vals):
        B[2 * k, 2 * k + 1], B[2 * k + 1, 2 * k] = v, -v
    assert abs(oracle_mod.pfaffian(B) - np.prod(vals)) < 1e-12 * abs(np.prod(vals))
    # Pf(A^T) = (-1)^{N/2} Pf(A) and Pf(P A P^T) = det(P) Pf(A)
    assert abs(oracle_mod.pfaffian(A.T) - (-1) ** (N // 2) * pf) < 1e-10 * abs(pf)
    perm = rng.permutation(N)
    P = np.eye(N)[perm]
    assert abs(oracle_mod.pfaffian(P @ A @ P.T) - np.linalg.det(P) * pf) < 1e-10 * abs(pf)


def test_zbareover_closed_form(oracle_mod):
    from sccm_amd.states import synthetic_state
    b = make_basis("sd")
    U, V = synthetic_state(b, 2, 2, seed=9)
    Zc, det = oracle_mod.zbareover(U, V)
    Z = V @ np.linalg.inv(U)
    assert np.abs(Zc - Z).max() < 1e-12
    assert abs(det - np.linalg.det(np.eye(b.n) + Z.T @ Z) ** -0.25) < 1e-13


def test_hndiag_textbook(oracle_mod):
    rng = np.random.default_rng(7)
    dim = 12
    G = rng.normal(size=(dim, 4))
    N = G @ G.T + 1e-9 * np.eye(dim)       # rank-deficient norm: cutoff matters
    Hs = rng.normal(size=(dim, dim))
    H = Hs + Hs.T
    r = oracle_mod.hndiag(N, H, 1e-6)
    w, vec = np.linalg.eigh(N)
    w, vec = w[::-1], vec[:, ::-1]
    keep = w / w[0] >= 1e-6
    assert r["m"] == keep.sum() == 4
    Rr = vec[:, keep] / np.sqrt(w[keep])
    E = np.linalg.eigvalsh(Rr.T @ H @ Rr)
    assert np.abs(np.sort(r["E"]) - E).max() < 1e-10
    # F solves H F = E N F in the kept space
    for k in range(r["m"]):
        f = r["F"][:, k]
        res = H @ f - r["E"][k] * (N @ f)
        assert np.abs(vec[:, keep].T @ res).max() < 1e-8


def test_cg_textbook():
    # <1/2 1/2, 1/2 -1/2 | 1 0> = 1/sqrt2 ; <1/2 1/2, 1/2 -1/2 | 0 0> = 1/sqrt2
    assert abs(cg(1, 1, 1, -1, 2, 0) - 2 ** -0.5) < 1e-15
    assert abs(cg(1, 1, 1, -1, 0, 0) - 2 ** -0.5) < 1e-15
    assert abs(cg(1, -1, 1, 1, 0, 0) + 2 ** -0.5) < 1e-15
    # orthonormality for j1 = 5/2, j2 = 3/2
    for J2, M2 in itertools.product(range(2, 9, 2), range(-8, 9, 2)):
        s = sum(cg(5, m1, 3, M2 - m1, J2, M2) ** 2 for m1 in range(-5, 6, 2))
        if abs(M2) <= J2:
            assert abs(s - 1) < 1e-13
